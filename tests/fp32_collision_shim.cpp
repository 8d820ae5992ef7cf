// Host build of the FP32 deviation-form collision and of the FP64 collision
// from the library's arithmetic header, for tests/test_fp32_collision.py.
#include "../fortran_loadfile_mitiwing_amd/csrc/lbmib_exact.h"

using namespace lbmib;

extern "C" {

// n nodes of deviations g (9 per node, FP32): the FP32 collision into g32 and
// the FP64 collision of f = W + g, returned as deviations, into g64.
void collide_both(int n, const double* s, const double* W, const float* g, float* g32, double* g64)
{
    const DevConst K = make_devconst(s, W);
    Rates R;
    for (int k = 0; k < 9; ++k) R.s[k] = s[k];
    for (int c = 0; c < n; ++c) {
        float a[9];
        double f[9];
        for (int k = 0; k < 9; ++k) {
            a[k] = g[9 * c + k];
            f[k] = W[k] + (double)g[9 * c + k];
        }
        float drho, jx, jy;
        collide_dev_f32(a, K, drho, jx, jy);
        double rho, u, v;
        moments(f, rho, u, v);
        collide_node(f, rho, u, v, R, false, 0.0, 0.0);
        for (int k = 0; k < 9; ++k) {
            g32[9 * c + k] = a[k];
            g64[9 * c + k] = f[k] - W[k];
        }
    }
}

}
