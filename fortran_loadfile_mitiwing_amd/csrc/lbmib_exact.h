// lbmib_exact.h -- FP64 building blocks shared by the HIP kernels and the
// host side of liblbmib.  Every expression is evaluated in the order the
// reference writes it; the library is compiled with -ffp-contract=off so no
// multiply-add is fused, which makes results reproducible bit for bit on the
// host and on gfx950.  Terms the reference multiplies by an exact zero are
// omitted (x + (+-0) == x in IEEE arithmetic).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define LBM_HD __host__ __device__ __forceinline__
#else
#define LBM_HD inline
#endif

namespace lbmib {

constexpr int Q = 9;
// D2Q9 lattice vectors (D2Q9model.f90:5-22), 0-based k
constexpr int EX[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int EY[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
// pai (main.f90:9)
constexpr double PAI = 3.1415926535897932384626433832795;

// Inverse of the MRT transform as the reference computes it: Gauss-Jordan
// with partial pivoting (nizhen, solution.f90:1462-1517) applied to matm
// (solution.f90:17-25).  These are that routine's outputs; lbmib_create
// recomputes them on the host and refuses to run if any bit differs.  Entries
// not listed are exactly zero (one is -0.0).
struct MinvTab {
    double c00, c01a, c02a;           // row 0: cols 0,1,2
    double c01b, c02b, c03, c04, c07; // rows 1-4 share |values|
    double c01c, c02c, c03c, c04c, c08;
};
constexpr double MI_0 = 0x1.c71c71c71c71cp-4;   //  0.111111111111111105
constexpr double MI_1a = -0x1.c71c71c71c71ep-4; // -0.111111111111111133
constexpr double MI_2a = 0x1.c71c71c71c71cp-4;  //  0.111111111111111105
constexpr double MI_1b = -0x1.c71c71c71c71cp-6; // -0.0277777777777777762
constexpr double MI_2b = -0x1.c71c71c71c71cp-5; // -0.0555555555555555525
constexpr double MI_3 = 0x1.5555555555555p-3;   //  0.166666666666666657
constexpr double MI_4 = -0x1.5555555555556p-3;  // -0.166666666666666685
constexpr double MI_7 = 0.25;
constexpr double MI_1c = 0x1.c71c71c71c71cp-5;  //  0.0555555555555555525
constexpr double MI_2c = 0x1.c71c71c71c71cp-6;  //  0.0277777777777777762
constexpr double MI_4c = 0x1.5555555555555p-4;  //  0.0833333333333333287
constexpr double MI_8 = 0.25;

// Full table (row-major) of the values above, for the host-side check.
LBM_HD void minv_table(double* m)
{
    for (int i = 0; i < 81; ++i) m[i] = 0.0;
    m[0 * 9 + 0] = MI_0; m[0 * 9 + 1] = MI_1a; m[0 * 9 + 2] = MI_2a;
    for (int r = 1; r <= 4; ++r) { m[r * 9 + 0] = MI_0; m[r * 9 + 1] = MI_1b; m[r * 9 + 2] = MI_2b; }
    m[1 * 9 + 3] = MI_3; m[1 * 9 + 4] = MI_4; m[1 * 9 + 7] = MI_7;
    m[2 * 9 + 5] = MI_3; m[2 * 9 + 6] = MI_4; m[2 * 9 + 7] = -MI_7;
    m[3 * 9 + 3] = -MI_3; m[3 * 9 + 4] = -MI_4; m[3 * 9 + 7] = MI_7;
    m[4 * 9 + 5] = -MI_3; m[4 * 9 + 6] = -MI_4; m[4 * 9 + 7] = -MI_7;
    for (int r = 5; r <= 8; ++r) { m[r * 9 + 0] = MI_0; m[r * 9 + 1] = MI_1c; m[r * 9 + 2] = MI_2c; }
    m[5 * 9 + 3] = MI_3; m[5 * 9 + 4] = MI_4c; m[5 * 9 + 5] = MI_3; m[5 * 9 + 6] = MI_4c; m[5 * 9 + 8] = MI_8;
    m[6 * 9 + 3] = -MI_3; m[6 * 9 + 4] = -MI_4c; m[6 * 9 + 5] = MI_3; m[6 * 9 + 6] = MI_4c; m[6 * 9 + 8] = -MI_8;
    m[7 * 9 + 3] = -MI_3; m[7 * 9 + 4] = -MI_4c; m[7 * 9 + 5] = -MI_3; m[7 * 9 + 6] = -MI_4c; m[7 * 9 + 8] = MI_8;
    m[8 * 9 + 3] = MI_3; m[8 * 9 + 4] = MI_4c; m[8 * 9 + 5] = -MI_3; m[8 * 9 + 6] = -MI_4c; m[8 * 9 + 8] = -MI_8;
    m[8 * 9 + 7] = -0.0;
}

// m = M f with the sequential dot products of matmulvoc
// (boundary_condition.f90:1747-1753); M's entries are 0, +-1, +-2, +-4 so
// every product is exact and zero terms drop out.
template <typename T>
LBM_HD void mrt_forward(const T* f, T* m)
{
    m[0] = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    m[1] = -4 * f[0] - f[1] - f[2] - f[3] - f[4] + 2 * f[5] + 2 * f[6] + 2 * f[7] + 2 * f[8];
    m[2] = 4 * f[0] - 2 * f[1] - 2 * f[2] - 2 * f[3] - 2 * f[4] + f[5] + f[6] + f[7] + f[8];
    m[3] = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
    m[4] = -2 * f[1] + 2 * f[3] + f[5] - f[6] - f[7] + f[8];
    m[5] = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
    m[6] = -2 * f[2] + 2 * f[4] + f[5] + f[6] - f[7] - f[8];
    m[7] = f[1] - f[2] + f[3] - f[4];
    m[8] = f[5] - f[6] + f[7] - f[8];
}

// f = Minv m, sequential over the non-zero columns of each row.
LBM_HD void mrt_inverse(const double* m, double* f)
{
    const double a0 = MI_0 * m[0];
    const double r1 = MI_1b * m[1], r2 = MI_2b * m[2];
    const double p3 = MI_3 * m[3], p4 = MI_4 * m[4], p5 = MI_3 * m[5], p6 = MI_4 * m[6], p7 = MI_7 * m[7];
    const double s1 = MI_1c * m[1], s2 = MI_2c * m[2];
    const double q4 = MI_4c * m[4], q6 = MI_4c * m[6], p8 = MI_8 * m[8];
    f[0] = a0 + MI_1a * m[1] + MI_2a * m[2];
    f[1] = a0 + r1 + r2 + p3 + p4 + p7;
    f[2] = a0 + r1 + r2 + p5 + p6 - p7;
    f[3] = a0 + r1 + r2 - p3 - p4 + p7;
    f[4] = a0 + r1 + r2 - p5 - p6 - p7;
    f[5] = a0 + s1 + s2 + p3 + q4 + p5 + q6 + p8;
    f[6] = a0 + s1 + s2 - p3 - q4 + p5 + q6 - p8;
    f[7] = a0 + s1 + s2 - p3 - q4 - p5 - q6 + p8;
    f[8] = a0 + s1 + s2 + p3 + q4 - p5 - q6 - p8;
}

// Moments of the (pre-collision) populations as physical_variables sums them
// (solution.f90:511-537): rho, then u = tempu/rho, v = tempv/rho.
LBM_HD void moments(const double* f, double& rho, double& u, double& v)
{
    double tr = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    double tu = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
    double tv = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
    rho = tr;
    u = tu / tr;
    v = tv / tr;
}

struct Rates {
    double s[9];
};

// MRT collision of one node (solution.f90:598-638).  f: the node's
// post-stream populations in, post-collision out.  forced: add the Guo force
// moments of (gfx, gfy).
LBM_HD void collide_node(double* f, double rt, double ut, double vt, const Rates& R,
                         bool forced, double gfx, double gfy)
{
    double m[9], t[9];
    mrt_forward(f, m);
    const double* s = R.s;
    double a = rt * ut * rt * ut;
    double b = rt * vt * rt * vt;
    t[0] = 0.0;
    t[1] = -2.0 * rt + 3.0 * (a + b);
    t[2] = rt - 3.0 * (a + b);
    t[3] = 0.0;
    t[4] = -rt * ut;
    t[5] = 0.0;
    t[6] = -rt * vt;
    t[7] = (a - b);
    t[8] = (rt * ut * rt * vt);
    if (forced) {
        double fp[9];
        fp[0] = 0.0;
        fp[1] = 6.0 * (1.0 - 0.5 * s[1]) * (ut * gfx + vt * gfy);
        fp[2] = -6.0 * (1.0 - 0.5 * s[2]) * (ut * gfx + vt * gfy);
        fp[3] = gfx;
        fp[4] = -(1.0 - 0.5 * s[4]) * gfx;
        fp[5] = gfy;
        fp[6] = -(1.0 - 0.5 * s[6]) * gfy;
        fp[7] = 2.0 * (1.0 - 0.5 * s[7]) * (ut * gfx - vt * gfy);
        fp[8] = (1.0 - 0.5 * s[8]) * (ut * gfy + vt * gfx);
#pragma unroll
        for (int k = 0; k < 9; ++k) t[k] = m[k] + (t[k] - m[k]) * s[k] + fp[k];
    } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) t[k] = m[k] + (t[k] - m[k]) * s[k];
    }
    mrt_inverse(t, f);
}

// Compressible moment equilibrium of the NEE edges (boundary_condition.f90:444-455).
LBM_HD void meq_nee(double rh, double uu, double vv, double* t)
{
    double jx = rh * uu;
    double jy = rh * vv;
    t[0] = rh;
    t[1] = -2.0 * rh + 3.0 * (jx * jx + jy * jy) / rh;
    t[2] = rh - 3.0 * (jx * jx + jy * jy) / rh;
    t[3] = jx;
    t[4] = -jx;
    t[5] = jy;
    t[6] = -jy;
    t[7] = (jx * jx - jy * jy) / rh;
    t[8] = (jx * jy) / rh;
}

// Non-equilibrium extrapolation in moment space (boundary_condition.f90:444-474):
// out = Minv [ meq(edge) + (1-s) (M f0nb - meq(nb)) ].
LBM_HD void nee_node(double rhe, double ue, double ve, double rhn, double un, double vn,
                     const double* f0nb, const Rates& R, double* out)
{
    double t1[9], t2[9], t3[9];
    meq_nee(rhe, ue, ve, t1);
    meq_nee(rhn, un, vn, t2);
    mrt_forward(f0nb, t3);
#pragma unroll
    for (int k = 0; k < 9; ++k) t3[k] = t1[k] + (1.0 - R.s[k]) * (t3[k] - t2[k]);
    mrt_inverse(t3, out);
}

// ---- FP32 deviation form ----
// Populations stored as g = f - W with W the reference's REAL(4) weights
// (the rest state, rho0 = sum W).  The collision is linear in f, so in moment
// space it relaxes m(g) = M g towards m_eq(rho, j) - M W:
//   k=1: -2 drho + 3 j^2 + c1,  k=2: drho - 3 j^2 + c2  (c1 = -2 rho0 - (M W)_1,
//   c2 = rho0 - (M W)_2, round-off sized), k=4: -jx, k=6: -jy, k=7: jx^2 - jy^2,
//   k=8: jx jy, k=0: -rho0 and k=3,5: 0 (the reference writes zero equilibria
//   there and relaxes them with s = 0), with drho = sum g and j = sum e g.
struct DevConst {
    float s[9];
    float c0, c1, c2;
    float oms[9];  // 1 - s[k]
    float s0c0;    // s[0] * c0
};

// The constants of the FP32 form for relaxation rates s and rest state W
// (rho0 = sum W, summed in index order).
LBM_HD DevConst make_devconst(const double* s, const double* W)
{
    DevConst K;
    double rho0 = 0.0;
    for (int k = 0; k < 9; ++k) {
        rho0 += W[k];
        K.s[k] = (float)s[k];
    }
    const double e_w = -4 * W[0] - W[1] - W[2] - W[3] - W[4] + 2 * (W[5] + W[6] + W[7] + W[8]);
    const double eps_w = 4 * W[0] - 2 * (W[1] + W[2] + W[3] + W[4]) + W[5] + W[6] + W[7] + W[8];
    K.c0 = (float)(-rho0);
    K.c1 = (float)(-2.0 * rho0 - e_w);
    K.c2 = (float)(rho0 - eps_w);
    for (int k = 0; k < 9; ++k) K.oms[k] = 1.0f - K.s[k];
    K.s0c0 = K.s[0] * K.c0;
    return K;
}

// Unforced collision of one node in deviation form; returns drho, jx, jy.
// The FP32 form has no operation order to reproduce, so the transforms take
// the fewest operations: M g from the sums and differences of opposite
// populations, each moment relaxed as m (1 - s) + s m_eq, and the exact
// rational inverse M^T diag(1/9, 1/36, 1/36, 1/6, 1/12, 1/6, 1/12, 1/4, 1/4)
// from shared partial sums, products and sums fused where they meet (77
// operations a node instead of the 111 of the row-by-row forms).
LBM_HD void collide_dev_f32(float* g, const DevConst& K, float& drho, float& jx, float& jy)
{
#pragma clang fp contract(fast)
    // m = M g
    const float a = g[1] + g[3], b = g[2] + g[4], c = g[5] + g[7], d = g[6] + g[8];
    const float e = g[1] - g[3], h = g[2] - g[4], p = g[5] - g[7], q = g[6] - g[8];
    const float sa = a + b, sd = c + d, pmq = p - q, ppq = p + q;
    const float m0 = g[0] + sa + sd;
    const float m1 = -4.0f * g[0] + (2.0f * sd - sa);
    const float m2 = 4.0f * g[0] + (sd - 2.0f * sa);
    const float m3 = e + pmq;
    const float m4 = -2.0f * e + pmq;
    const float m5 = h + ppq;
    const float m6 = -2.0f * h + ppq;
    const float m7 = a - b;
    const float m8 = c - d;
    drho = m0;
    jx = m3;
    jy = m5;
    // relaxation towards m_eq(drho, j) - M W
    const float jxx = jx * jx, jyy = jy * jy, j2 = jxx + jyy;
    const float q1 = -2.0f * drho + (3.0f * j2 + K.c1);
    const float q2 = drho + (-3.0f * j2 + K.c2);
    const float t0 = m0 * K.oms[0] + K.s0c0;
    const float t1 = m1 * K.oms[1] + K.s[1] * q1;
    const float t2 = m2 * K.oms[2] + K.s[2] * q2;
    const float t3 = m3 * K.oms[3];
    const float t4 = m4 * K.oms[4] - K.s[4] * jx;
    const float t5 = m5 * K.oms[5];
    const float t6 = m6 * K.oms[6] - K.s[6] * jy;
    const float t7 = m7 * K.oms[7] + K.s[7] * (jxx - jyy);
    const float t8 = m8 * K.oms[8] + K.s[8] * (jx * jy);
    // g = M^-1 t
    const float a0 = t0 * (1.0f / 9.0f);
    const float A = a0 - t1 * (1.0f / 36.0f) - t2 * (1.0f / 18.0f);
    const float B = a0 + t1 * (1.0f / 18.0f) + t2 * (1.0f / 36.0f);
    const float Ap = A + t7 * 0.25f, Am = A - t7 * 0.25f;
    const float Bp = B + t8 * 0.25f, Bm = B - t8 * 0.25f;
    const float dx = t3 - t4, dy = t5 - t6;
    const float U = t3 * (1.0f / 6.0f) + t4 * (1.0f / 12.0f);
    const float V = t5 * (1.0f / 6.0f) + t6 * (1.0f / 12.0f);
    const float uv = U + V, vu = V - U;
    g[0] = a0 + (t2 - t1) * (1.0f / 9.0f);
    g[1] = Ap + dx * (1.0f / 6.0f);
    g[3] = Ap - dx * (1.0f / 6.0f);
    g[2] = Am + dy * (1.0f / 6.0f);
    g[4] = Am - dy * (1.0f / 6.0f);
    g[5] = Bp + uv;
    g[7] = Bp - uv;
    g[6] = Bm + vu;
    g[8] = Bm - vu;
}

// ---- delta kernel ----

// asin by fdlibm's e_asin method (Sun Microsystems, 1993): x + x^3 R(x^2)
// on |x| < 0.5, the pi/2 - 2 asin(sqrt((1-x)/2)) reduction with a split
// square root above.  Plain IEEE operations only.
LBM_HD double asin_fd(double x)
{
    const double pio2_hi = 1.57079632679489655800e+00;
    const double pio2_lo = 6.12323399573676603587e-17;
    const double pio4_hi = 7.85398163397448278999e-01;
    const double pS0 = 1.66666666666666657415e-01, pS1 = -3.25565818622400915405e-01,
                 pS2 = 2.01212532134862925881e-01, pS3 = -4.00555345006794114027e-02,
                 pS4 = 7.91534994289814532176e-04, pS5 = 3.47933107596021167570e-05;
    const double qS1 = -2.40339491173441421878e+00, qS2 = 2.02094576023350569471e+00,
                 qS3 = -6.88283971605453293030e-01, qS4 = 7.70381505559019352791e-02;
    double ax = x < 0 ? -x : x;
    if (!(ax <= 1.0)) return __builtin_nan("");
    if (ax == 1.0) return x > 0 ? pio2_hi + pio2_lo : -(pio2_hi + pio2_lo);
    if (ax < 0.5) {
        if (ax < 7.450580596923828125e-9) return x;
        double t = x * x;
        double p = t * (pS0 + t * (pS1 + t * (pS2 + t * (pS3 + t * (pS4 + t * pS5)))));
        double q = 1.0 + t * (qS1 + t * (qS2 + t * (qS3 + t * qS4)));
        double w = p / q;
        return x + x * w;
    }
    double w = 1.0 - ax;
    double t = w * 0.5;
    double p = t * (pS0 + t * (pS1 + t * (pS2 + t * (pS3 + t * (pS4 + t * pS5)))));
    double q = 1.0 + t * (qS1 + t * (qS2 + t * (qS3 + t * qS4)));
    double s = __builtin_sqrt(t);
    double r;
    if (ax >= 0.975) {
        w = p / q;
        r = pio2_hi - (2.0 * (s + s * w) - pio2_lo);
    } else {
        uint64_t bits = __builtin_bit_cast(uint64_t, s) & 0xFFFFFFFF00000000ULL;
        double wh = __builtin_bit_cast(double, bits);
        double c = (t - wh * wh) / (s + wh);
        double rr = p / q;
        double pp = 2.0 * s * rr - (pio2_lo - 2.0 * c);
        double qq = pio4_hi - 2.0 * wh;
        r = pio4_hi - (pp - qq);
    }
    return x > 0 ? r : -r;
}

// One-axis delta kernel for disr = |x - X|/scale >= 0
// (boundary_condition.f90:1563-1600).
LBM_HD double delta1(double disr, int smooth)
{
    const double SQ2 = 1.4142135623730951; // sqrt(2.0d0)
    if (smooth == 0) {
        if (disr < 1.0)
            return 0.125 * (3.0 - 2.0 * disr + __builtin_sqrt(1.0 + 4.0 * disr - 4.0 * disr * disr));
        else if (disr < 2.0)
            return 0.125 * (5.0 - 2.0 * disr - __builtin_sqrt(-7.0 + 12.0 * disr - 4.0 * disr * disr));
        return 0.0;
    }
    if (disr <= 0.5)
        return 0.375 + 0.03125 * PAI - 0.25 * disr * disr;
    else if (disr <= 1.5)
        return 0.25 + 0.125 * (1.0 - disr) * __builtin_sqrt(-2.0 + 8.0 * disr - 4.0 * disr * disr)
               - 0.125 * asin_fd(SQ2 * (disr - 1.0));
    else if (disr <= 2.5)
        return 1.0625 - 0.015625 * PAI - 0.75 * disr + 0.125 * disr * disr
               + 0.0625 * (disr - 2.0) * __builtin_sqrt(-14.0 + 16.0 * disr - 4.0 * disr * disr)
               + 0.0625 * asin_fd(SQ2 * (disr - 2.0));
    return 0.0;
}

}  // namespace lbmib
