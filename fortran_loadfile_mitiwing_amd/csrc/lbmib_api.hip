// lbmib_api.hip -- host side of liblbmib: the C ABI declared in include/lbmib.h.
//
// Owns the device state of a case, builds the per-step kinematics from the
// motion table exactly as the reference does (mode.f90:199-283), and drives
// the kernels of lbmib_kernels.hip on two HIP streams per handle:
//   s0: lattice kernel, outer-boundary kernel, IB-node collide
//   s1: immersed-boundary chain (runs beside the lattice kernel)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lbmib.h"
#include "lbmib_kernels.h"

namespace lbmib {
// kernels (lbmib_kernels.hip)
__global__ void k_collide_f64(const double*, double*, int, int, long long, Rates, double*);
__global__ void k_edges_f64(DevCase, int, double, double*);
__global__ void k_markers(DevCase, int);
__global__ void k_clear(DevCase);
__global__ void k_claim(DevCase);
__global__ void k_count(DevCase);
__global__ void k_fill(DevCase);
__global__ void k_nodes(DevCase, int);
__global__ void k_dvals(DevCase);
__global__ void k_assemble(DevCase);
__global__ void k_chol(DevCase);
__global__ void k_spread(DevCase);
__global__ void k_band_f64(DevCase, int, double*);
__global__ void k_forces(DevCase, int);
__global__ void k_residual(const double*, const double*, int, int, long long, double*);
__global__ void k_nan_check(const double*, long long, int*);
}  // namespace lbmib

using namespace lbmib;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(LBMIB_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));       \
    } while (0)

// Gauss-Jordan inverse of the MRT transform as nizhen computes it
// (solution.f90:1462-1517), row-major.
void nizhen9(const double* AA, double* b1)
{
    const int Nm = 9;
    double a[81];
    std::memcpy(a, AA, sizeof a);
    for (int i = 0; i < 81; ++i) b1[i] = 0.0;
    for (int k = 0; k < Nm; ++k) b1[k * Nm + k] = 1.0;
    for (int k = 0; k < Nm - 1; ++k) {
        int ik = k;
        double temp = std::fabs(a[ik * Nm + k]);
        for (int i = k; i < Nm; ++i)
            if (std::fabs(a[i * Nm + k]) > std::fabs(temp)) { temp = std::fabs(a[i * Nm + k]); ik = i; }
        for (int j = 0; j < Nm; ++j) {
            std::swap(a[ik * Nm + j], a[k * Nm + j]);
            std::swap(b1[ik * Nm + j], b1[k * Nm + j]);
        }
        for (int i = 0; i < Nm; ++i) {
            double mik = a[i * Nm + k] / a[k * Nm + k];
            temp = a[k * Nm + k];
            for (int j = k; j < Nm; ++j)
                if (i != k) a[i * Nm + j] = a[i * Nm + j] - mik * a[k * Nm + j];
            for (int j = k; j < Nm; ++j) a[k * Nm + j] = a[k * Nm + j] / temp;
            for (int j = 0; j < Nm; ++j)
                if (i != k) b1[i * Nm + j] = b1[i * Nm + j] - mik * b1[k * Nm + j];
            for (int j = 0; j < Nm; ++j) b1[k * Nm + j] = b1[k * Nm + j] / temp;
        }
    }
    double temp = a[80];
    for (int i = 0; i < Nm - 1; ++i) {
        double mik = a[i * Nm + 8] / a[80];
        a[i * Nm + 8] = 0.0;
        for (int j = 0; j < Nm; ++j) b1[i * Nm + j] = b1[i * Nm + j] - mik * b1[8 * Nm + j];
    }
    a[80] = 1.0;
    for (int j = 0; j < Nm; ++j) b1[8 * Nm + j] = b1[8 * Nm + j] / temp;
}

const double MATM[81] = {1, 1, 1, 1, 1, 1, 1, 1, 1,  -4, -1, -1, -1, -1, 2, 2, 2, 2,
                         4, -2, -2, -2, -2, 1, 1, 1, 1, 0, 1, 0, -1, 0, 1, -1, -1, 1,
                         0, -2, 0, 2, 0, 1, -1, -1, 1, 0, 0, 1, 0, -1, 1, 1, -1, -1,
                         0, 0, -2, 0, 2, 1, 1, -1, -1, 0, 1, -1, 1, -1, 0, 0, 0, 0,
                         0, 0, 0, 0, 0, 1, -1, 1, -1};

// equilibrium with the REAL(4) weights (D2Q9model.f90:24-46)
double feq_ref(double rh, double u1, double v1, int k)
{
    const double w = k == 0 ? (double)(4.0f / 9.0f) : (k <= 4 ? (double)(1.0f / 9.0f) : (double)(1.0f / 36.0f));
    double eu = (double)EX[k] * u1 + (double)EY[k] * v1;
    double utwo = u1 * u1 + v1 * v1;
    return w * rh * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * utwo);
}

// ---- motion tables (mode.f90:176-269) ----
struct MotionTable {
    std::vector<double> t;  // rows x ncols
    int rows = 0, ncols = 0, index = 1;
    double at(int r, int c) const { return t[(size_t)(r - 1) * ncols + (c - 1)]; }  // 1-based
};

void interpolate_row(const MotionTable& m, int i1, int i2, double target, double* row)
{
    double t1 = m.at(i1, 1), t2 = m.at(i2, 1);
    if (std::fabs(t2 - t1) <= 1.0e-14) {
        for (int c = 1; c <= m.ncols; ++c) row[c - 1] = m.at(i1, c);
        row[0] = target;
        return;
    }
    double weight = (target - t1) / (t2 - t1);
    for (int c = 1; c <= m.ncols; ++c) row[c - 1] = (1.0 - weight) * m.at(i1, c) + weight * m.at(i2, c);
    row[0] = target;
}

void lookup_row(MotionTable& m, double target, double* row)
{
    const int n = m.rows;
    auto copy = [&](int r) {
        for (int c = 1; c <= m.ncols; ++c) row[c - 1] = m.at(r, c);
        row[0] = target;
    };
    if (target <= m.at(1, 1)) { m.index = 1; copy(1); return; }
    if (target >= m.at(n, 1)) { m.index = std::max(1, n - 1); copy(n); return; }
    int idx = std::max(1, std::min(m.index, n - 1));
    while (idx > 1 && target < m.at(idx, 1)) idx--;
    while (idx < n - 1 && target > m.at(idx + 1, 1)) idx++;
    m.index = idx;
    if (std::fabs(target - m.at(idx, 1)) <= 1.0e-12) { copy(idx); return; }
    if (std::fabs(target - m.at(idx + 1, 1)) <= 1.0e-12) { copy(idx + 1); return; }
    interpolate_row(m, idx, idx + 1, target, row);
}

void zigzag(int nw, int spw, std::vector<int>& perm)
{
    perm.clear();
    for (int w = 0; w < nw; ++w) {
        int lo = 0, hi = spw - 1, t = 0;
        while (lo <= hi) {
            perm.push_back(w * spw + ((t & 1) == 0 ? lo++ : hi--));
            t++;
        }
    }
}

}  // namespace

struct lbmib_handle {
    lbmib_params p;
    DevCase D;
    hipStream_t s0 = nullptr, s1 = nullptr;
    hipEvent_t ev_ib = nullptr, ev_step = nullptr;
    std::vector<double> xb0, yb0, arc_l;  // polygon after the (xc-xp, yc-yp) shift
    std::vector<int> perm, iperm;
    double cs = 0.0;
    // kinematics
    bool have_motion = false;
    std::vector<double> step_rows;  // per step: the motion row used (ncols)
    int ncols = 0, max_steps = 0;
    double motion_end_time = 0.0;
    int steps_done = 0;
    int last_flag_step[2] = {0, -1};  // step whose macros each slot holds
    // device allocations to free
    std::vector<void*> allocs;
    // profiling
    bool profiling = false;
    std::vector<hipEvent_t> evpool;
    double tms[6] = {0, 0, 0, 0, 0, 0};
    int dev = 0;
};

namespace {

template <typename T>
int dalloc(lbmib_handle* h, T** ptr, size_t count)
{
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sizeof(T) * std::max<size_t>(count, 1));
    if (e != hipSuccess)
        return fail(LBMIB_E_HIP, "hipMalloc of %zu bytes failed: %s", sizeof(T) * count, hipGetErrorString(e));
    h->allocs.push_back(p);
    *ptr = (T*)p;
    return 0;
}

int fill_i32(int* d, int v, size_t n, hipStream_t s)
{
    std::vector<int> hv(n, v);
    HIP_TRY(hipMemcpyAsync(d, hv.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

// Largest zig-zag distance between two segments whose +-2.5 lattice-unit
// boxes can share a node under any rigid motion: such segments are within
// 5*sqrt(2) lattice units of each other in the body frame.
int geometric_bandwidth(const lbmib_handle* h)
{
    const int nw = h->p.n_wings, ppw = h->p.pts_per_wing, spw = ppw - 1, ns = nw * spw;
    std::vector<double> xm(ns), ym(ns);
    for (int w = 0; w < nw; ++w)
        for (int i = 0; i < spw; ++i) {
            int p = w * ppw + i;
            xm[w * spw + i] = 0.5 * (h->xb0[p + 1] + h->xb0[p]) / h->p.scale;
            ym[w * spw + i] = 0.5 * (h->yb0[p + 1] + h->yb0[p]) / h->p.scale;
        }
    const double r2 = 50.0 * (1.0 + 1e-9) + 1e-9;  // (5 sqrt 2)^2
    int bw = 0;
    for (int a = 0; a < ns; ++a)
        for (int b = 0; b < a; ++b) {
            if (a / spw != b / spw) continue;  // other wings handled separately
            double dx = xm[a] - xm[b], dy = ym[a] - ym[b];
            if (dx * dx + dy * dy <= r2) bw = std::max(bw, std::abs(h->iperm[a] - h->iperm[b]));
        }
    return bw;
}

}  // namespace

extern "C" {

const char* lbmib_last_error(void) { return g_err.c_str(); }

const char* lbmib_version(void) { return "lbmib 0.1 gfx950"; }

int lbmib_destroy(lbmib_handle* h)
{
    if (!h) return 0;
    (void)hipSetDevice(h->dev);
    if (h->s0) (void)hipStreamSynchronize(h->s0);
    if (h->s1) (void)hipStreamSynchronize(h->s1);
    for (void* p : h->allocs) (void)hipFree(p);
    for (hipEvent_t e : h->evpool) (void)hipEventDestroy(e);
    if (h->ev_ib) (void)hipEventDestroy(h->ev_ib);
    if (h->ev_step) (void)hipEventDestroy(h->ev_step);
    if (h->s0) (void)hipStreamDestroy(h->s0);
    if (h->s1) (void)hipStreamDestroy(h->s1);
    delete h;
    return 0;
}

int lbmib_create(const lbmib_params* p, const double* xb0, const double* yb0, lbmib_handle** out)
{
    if (!p || !xb0 || !yb0 || !out) return fail(LBMIB_E_ARG, "null argument");
    *out = nullptr;
    if (p->nx < 8 || p->ny < 8) return fail(LBMIB_E_ARG, "lattice must be at least 8x8 (got %dx%d)", p->nx, p->ny);
    if (p->n_wings < 1 || p->n_wings > 8) return fail(LBMIB_E_ARG, "wing_count %d out of range 1..8", p->n_wings);
    if (p->pts_per_wing < 4) return fail(LBMIB_E_ARG, "need at least 4 boundary points per wing");
    if (!(p->scale > 0.0)) return fail(LBMIB_E_ARG, "scale must be positive");
    if (p->precision != LBMIB_PREC_FP64) return fail(LBMIB_E_ARG, "precision %d not available in this build", p->precision);
    if (p->smooth != 0 && p->smooth != 1) return fail(LBMIB_E_ARG, "smooth must be 0 or 1");

    // the kernels carry nizhen's inverse as constants; check them bit for bit
    double minv_host[81], minv_dev[81];
    nizhen9(MATM, minv_host);
    minv_table(minv_dev);
    for (int i = 0; i < 81; ++i)
        if (std::memcmp(&minv_host[i], &minv_dev[i], sizeof(double)) != 0)
            return fail(LBMIB_E_STATE, "MRT inverse constant %d differs from the host Gauss-Jordan result", i);

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(LBMIB_E_HIP, "no HIP device available");
    if (p->device < 0 || p->device >= ndev) return fail(LBMIB_E_ARG, "device %d out of range (%d devices)", p->device, ndev);

    lbmib_handle* h = new lbmib_handle();
    h->p = *p;
    h->dev = p->device;
    int rc = 0;
#define TRY(x)                         \
    do {                               \
        rc = (x);                      \
        if (rc) { lbmib_destroy(h); return rc; } \
    } while (0)
#define HTRY(x)                                                                                 \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) {                                                                 \
            lbmib_destroy(h);                                                                   \
            return fail(LBMIB_E_HIP, "%s failed: %s", #x, hipGetErrorString(e_));               \
        }                                                                                       \
    } while (0)
    HTRY(hipSetDevice(h->dev));
    HTRY(hipStreamCreateWithFlags(&h->s0, hipStreamNonBlocking));
    HTRY(hipStreamCreateWithFlags(&h->s1, hipStreamNonBlocking));
    HTRY(hipEventCreateWithFlags(&h->ev_ib, hipEventDisableTiming));
    HTRY(hipEventCreateWithFlags(&h->ev_step, hipEventDisableTiming));

    const int nx = p->nx, ny = p->ny, nw = p->n_wings, ppw = p->pts_per_wing;
    const long long N = (long long)nx * ny;
    DevCase& D = h->D;
    std::memset(&D, 0, sizeof D);
    D.nx = nx; D.ny = ny; D.N = N;
    D.ox = p->ox; D.oy = p->oy; D.scale = p->scale; D.U = p->U; D.xc = p->xc; D.yc = p->yc;
    for (int k = 0; k < 9; ++k) D.R.s[k] = p->s[k];
    D.smooth = p->smooth;
    D.nw = nw; D.ppw = ppw; D.spw = ppw - 1; D.nseg = nw * (ppw - 1);
    D.cap = D.nseg * BOXMAX;
    h->cs = (double)std::sqrt(3.0f) / 3.0;  // REAL(4) sqrt (boundary_condition.f90:304)

    // polygon after read_boundary's shift (airfoil.f90:235-241)
    int npts = nw * ppw;
    h->xb0.resize(npts); h->yb0.resize(npts); h->arc_l.resize(D.nseg);
    for (int w = 0; w < nw; ++w)
        for (int i = 0; i < ppw - 1; ++i) {
            double dx = xb0[w * ppw + i + 1] - xb0[w * ppw + i];
            double dy = yb0[w * ppw + i + 1] - yb0[w * ppw + i];
            h->arc_l[w * (ppw - 1) + i] = std::sqrt(dx * dx + dy * dy) / p->scale;
        }
    for (int i = 0; i < npts; ++i) {
        h->xb0[i] = xb0[i] + (p->xc - p->xp);
        h->yb0[i] = yb0[i] + (p->yc - p->yp);
    }
    zigzag(nw, D.spw, h->perm);
    h->iperm.assign(D.nseg, 0);
    for (int q = 0; q < D.nseg; ++q) h->iperm[h->perm[q]] = q;

    // lattice
    TRY(dalloc(h, &D.f[0], 9 * (size_t)N));
    TRY(dalloc(h, &D.f[1], 9 * (size_t)N));
    TRY(dalloc(h, &D.f0L, 9 * (size_t)ny));
    TRY(dalloc(h, &D.f0R, 9 * (size_t)ny));
    TRY(dalloc(h, &D.macro[0], 3 * (size_t)N));
    TRY(dalloc(h, &D.macro[1], 3 * (size_t)N));
    // init (solution.f90:34-59): f = f0 = feq(1, U, 0); rho = 1, u = U, v = 0
    {
        double feq0[9];
        for (int k = 0; k < 9; ++k) feq0[k] = feq_ref(1.0, p->U, 0.0, k);
        std::vector<double> plane((size_t)N);
        for (int k = 0; k < 9; ++k) {
            std::fill(plane.begin(), plane.end(), feq0[k]);
            HTRY(hipMemcpy(D.f[0] + (size_t)k * N, plane.data(), sizeof(double) * N, hipMemcpyHostToDevice));
            HTRY(hipMemcpy(D.f[1] + (size_t)k * N, plane.data(), sizeof(double) * N, hipMemcpyHostToDevice));
        }
        std::vector<double> col(9 * (size_t)ny);
        for (int k = 0; k < 9; ++k)
            for (int j = 0; j < ny; ++j) col[(size_t)k * ny + j] = feq0[k];
        HTRY(hipMemcpy(D.f0L, col.data(), sizeof(double) * col.size(), hipMemcpyHostToDevice));
        HTRY(hipMemcpy(D.f0R, col.data(), sizeof(double) * col.size(), hipMemcpyHostToDevice));
        std::fill(plane.begin(), plane.end(), 1.0);
        HTRY(hipMemcpy(D.macro[0], plane.data(), sizeof(double) * N, hipMemcpyHostToDevice));
        std::fill(plane.begin(), plane.end(), p->U);
        HTRY(hipMemcpy(D.macro[0] + N, plane.data(), sizeof(double) * N, hipMemcpyHostToDevice));
        std::fill(plane.begin(), plane.end(), 0.0);
        HTRY(hipMemcpy(D.macro[0] + 2 * N, plane.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    }
    // immersed boundary
    double *dxb0, *dyb0;
    TRY(dalloc(h, &dxb0, npts));
    TRY(dalloc(h, &dyb0, npts));
    HTRY(hipMemcpy(dxb0, h->xb0.data(), sizeof(double) * npts, hipMemcpyHostToDevice));
    HTRY(hipMemcpy(dyb0, h->yb0.data(), sizeof(double) * npts, hipMemcpyHostToDevice));
    D.xb0 = dxb0; D.yb0 = dyb0;
    const int ns = D.nseg, cap = D.cap;
    TRY(dalloc(h, &D.xb2, ns)); TRY(dalloc(h, &D.yb2, ns));
    TRY(dalloc(h, &D.ub, ns)); TRY(dalloc(h, &D.vb, ns));
    TRY(dalloc(h, &D.box, 4 * (size_t)ns));
    TRY(dalloc(h, &D.claim, (size_t)N));
    TRY(dalloc(h, &D.nodeidx, (size_t)N));
    TRY(dalloc(h, &D.segcnt, ns));
    TRY(dalloc(h, &D.adj, 3 * (size_t)cap));
    TRY(dalloc(h, &D.nadj, 1));
    TRY(dalloc(h, &D.ustar, cap)); TRY(dalloc(h, &D.vstar, cap)); TRY(dalloc(h, &D.rhoadj, cap));
    TRY(dalloc(h, &D.dcnt, ns));
    TRY(dalloc(h, &D.didx, (size_t)ns * BOXMAX));
    TRY(dalloc(h, &D.dval, (size_t)ns * BOXMAX));
    TRY(dalloc(h, &D.nmkcnt, cap));
    TRY(dalloc(h, &D.nmk, (size_t)cap * NCAP));
    TRY(dalloc(h, &D.nmkv, (size_t)cap * NCAP));
    int *dperm, *diperm;
    TRY(dalloc(h, &dperm, ns)); TRY(dalloc(h, &diperm, ns));
    HTRY(hipMemcpy(dperm, h->perm.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
    HTRY(hipMemcpy(diperm, h->iperm.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
    D.perm = dperm; D.iperm = diperm;
    TRY(dalloc(h, &D.rhs, 2 * (size_t)ns));
    TRY(dalloc(h, &D.y, 2 * (size_t)ns));
    TRY(dalloc(h, &D.glx, cap)); TRY(dalloc(h, &D.gly, cap));
    TRY(dalloc(h, &D.err, 1));
    TRY(dalloc(h, &D.red, 2 * 4096));
    TRY(fill_i32(D.claim, INT_MAX, (size_t)N, h->s0));
    TRY(fill_i32(D.nodeidx, -1, (size_t)N, h->s0));
    TRY(fill_i32(D.nadj, 0, 1, h->s0));
    TRY(fill_i32(D.err, 0, 1, h->s0));
    D.bw = geometric_bandwidth(h);
    if (D.bw >= ns) D.bw = ns - 1;
    HTRY(hipEventRecord(h->ev_step, h->s0));
    HTRY(hipStreamSynchronize(h->s0));
    *out = h;
    return 0;
#undef TRY
#undef HTRY
}

int lbmib_set_motion_table(lbmib_handle* h, const double* rows, int32_t nrows, int32_t ncols)
{
    if (!h || !rows) return fail(LBMIB_E_ARG, "null argument");
    const int nw = h->p.n_wings;
    if (ncols != 1 + 9 * nw) return fail(LBMIB_E_ARG, "motion table needs %d columns, got %d", 1 + 9 * nw, ncols);
    if (nrows < 1) return fail(LBMIB_E_ARG, "motion table is empty");
    for (int r = 1; r < nrows; ++r)
        if (rows[(size_t)r * ncols] < rows[(size_t)(r - 1) * ncols])
            return fail(LBMIB_E_ARG, "motion table time column must be non-decreasing (row %d)", r + 1);
    if (h->steps_done > 0) return fail(LBMIB_E_STATE, "motion table must be set before the first step");
    HIP_TRY(hipSetDevice(h->dev));
    MotionTable m;
    m.t.assign(rows, rows + (size_t)nrows * ncols);
    m.rows = nrows; m.ncols = ncols;
    const double dt = h->p.scale * h->p.U;  // motion_dt = scale*UU (main.f90:255)
    h->motion_end_time = m.at(nrows, 1);
    const int count = (int)((h->motion_end_time + 1.0e-12) / dt) + 1;  // mode.f90:258
    std::vector<double> table((size_t)count * ncols);
    for (int i = 1; i <= count; ++i) lookup_row(m, (double)(i - 1) * dt, &table[(size_t)(i - 1) * ncols]);
    // reference loop bound (main.f90:352-353)
    int maxn = 0;
    while ((double)maxn * dt <= h->motion_end_time + 1.0e-12) maxn++;
    h->max_steps = maxn;
    h->ncols = ncols;
    // per step n: mode's row (mode.f90:280-282)
    h->step_rows.assign((size_t)maxn * ncols, 0.0);
    std::vector<double> kin((size_t)maxn * nw * KIN, 0.0);
    for (int n = 1; n <= maxn; ++n) {
        double time = (double)(n - 1) * dt;
        long idx = std::lround(time / dt) + 1;
        idx = std::max(1L, std::min(idx, (long)count));
        const double* row = &table[(size_t)(idx - 1) * ncols];
        std::memcpy(&h->step_rows[(size_t)(n - 1) * ncols], row, sizeof(double) * ncols);
        for (int w = 0; w < nw; ++w) {
            const double* c = row + 1 + 9 * w;
            double* k = &kin[((size_t)(n - 1) * nw + w) * KIN];
            k[0] = std::sin(c[0]);
            k[1] = std::cos(c[0]);
            k[2] = c[1];  // omega
            k[3] = c[3];  // sa
            k[4] = c[4];  // sau
            k[5] = c[6];  // sb
            k[6] = c[7];  // sbv
        }
    }
    // wings that may approach each other couple across their blocks
    if (nw > 1) {
        double margin = 6.0 * h->p.scale;
        bool close = false;
        for (int n = 0; n < maxn && !close; ++n) {
            double lo[8][2], hi[8][2];
            for (int w = 0; w < nw; ++w) {
                const double* k = &kin[((size_t)n * nw + w) * KIN];
                lo[w][0] = lo[w][1] = 1e300; hi[w][0] = hi[w][1] = -1e300;
                for (int i = 0; i < h->p.pts_per_wing; ++i) {
                    double x = h->xb0[w * h->p.pts_per_wing + i], y = h->yb0[w * h->p.pts_per_wing + i];
                    double X = k[1] * x + k[0] * y + k[3], Y = -k[0] * x + k[1] * y + k[5];
                    lo[w][0] = std::min(lo[w][0], X); hi[w][0] = std::max(hi[w][0], X);
                    lo[w][1] = std::min(lo[w][1], Y); hi[w][1] = std::max(hi[w][1], Y);
                }
            }
            for (int a = 0; a < nw && !close; ++a)
                for (int b = 0; b < a && !close; ++b)
                    if (lo[a][0] - margin <= hi[b][0] && lo[b][0] - margin <= hi[a][0] &&
                        lo[a][1] - margin <= hi[b][1] && lo[b][1] - margin <= hi[a][1])
                        close = true;
        }
        if (close) h->D.bw = h->D.nseg - 1;
    }
    DevCase& D = h->D;
    double* dkin;
    if (dalloc(h, &dkin, kin.size())) return LBMIB_E_HIP;
    HIP_TRY(hipMemcpy(dkin, kin.data(), sizeof(double) * kin.size(), hipMemcpyHostToDevice));
    D.kin = dkin;
    D.steps_cap = maxn;
    if (dalloc(h, &D.hist, (size_t)maxn * 3 * nw)) return LBMIB_E_HIP;
    if (dalloc(h, &D.G, (size_t)D.nseg * (D.bw + 1))) return LBMIB_E_HIP;
    h->have_motion = true;
    return 0;
}

static int launch_step(lbmib_handle* h, int n, bool flag, hipEvent_t* ev)
{
    DevCase& D = h->D;
    const int cur = n & 1;
    const int nx = D.nx, ny = D.ny;
    const int step = n - 1;
    double* macro = flag ? D.macro[cur] : nullptr;
    const int ns = D.nseg, cap = D.cap;
    // IB chain on s1, after the previous step completed on s0
    HIP_TRY(hipStreamWaitEvent(h->s1, h->ev_step, 0));
    if (ev) HIP_TRY(hipEventRecord(ev[2], h->s1));
    hipLaunchKernelGGL(k_markers, dim3((ns + 255) / 256), dim3(256), 0, h->s1, D, step);
    hipLaunchKernelGGL(k_clear, dim3((cap + 255) / 256), dim3(256), 0, h->s1, D);
    hipLaunchKernelGGL(k_claim, dim3((ns * BOXMAX + 255) / 256), dim3(256), 0, h->s1, D);
    hipLaunchKernelGGL(k_count, dim3((ns + 255) / 256), dim3(256), 0, h->s1, D);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(1024), 0, h->s1, D);
    hipLaunchKernelGGL(k_nodes, dim3((cap + 255) / 256), dim3(256), 0, h->s1, D, cur);
    hipLaunchKernelGGL(k_dvals, dim3((ns + 63) / 64), dim3(64), 0, h->s1, D);
    long long nb = (long long)ns * (D.bw + 1);
    hipLaunchKernelGGL(k_assemble, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, h->s1, D);
    hipLaunchKernelGGL(k_chol, dim3(1), dim3(1024), 0, h->s1, D);
    hipLaunchKernelGGL(k_spread, dim3((cap + 255) / 256), dim3(256), 0, h->s1, D);
    if (ev) HIP_TRY(hipEventRecord(ev[3], h->s1));
    HIP_TRY(hipEventRecord(h->ev_ib, h->s1));
    hipLaunchKernelGGL(k_forces, dim3(1), dim3(64), 0, h->s1, D, step);
    // lattice on s0
    if (ev) HIP_TRY(hipEventRecord(ev[0], h->s0));
    hipLaunchKernelGGL(k_collide_f64, dim3((nx - 2 + 255) / 256, ny - 2), dim3(256), 0, h->s0,
                       (const double*)D.f[cur ^ 1], D.f[cur], nx, ny, D.N, D.R, macro);
    if (ev) HIP_TRY(hipEventRecord(ev[1], h->s0));
    hipLaunchKernelGGL(k_edges_f64, dim3((ny - 2 + nx + 255) / 256), dim3(256), 0, h->s0, D, cur, h->cs, macro);
    if (ev) HIP_TRY(hipEventRecord(ev[6], h->s0));
    HIP_TRY(hipStreamWaitEvent(h->s0, h->ev_ib, 0));
    if (ev) HIP_TRY(hipEventRecord(ev[4], h->s0));
    hipLaunchKernelGGL(k_band_f64, dim3((cap + 255) / 256), dim3(256), 0, h->s0, D, cur, macro);
    if (ev) HIP_TRY(hipEventRecord(ev[5], h->s0));
    HIP_TRY(hipEventRecord(h->ev_step, h->s0));
    HIP_TRY(hipGetLastError());
    if (flag) {
        h->last_flag_step[cur] = n;
    }
    return 0;
}

int lbmib_step(lbmib_handle* h, int32_t n_steps)
{
    if (!h) return fail(LBMIB_E_ARG, "null handle");
    if (!h->have_motion) return fail(LBMIB_E_STATE, "lbmib_set_motion_table must be called before lbmib_step");
    if (n_steps < 0) return fail(LBMIB_E_ARG, "n_steps must be >= 0");
    if (h->steps_done + n_steps > h->max_steps)
        return fail(LBMIB_E_ARG, "step %d beyond the motion table (last step %d)", h->steps_done + n_steps, h->max_steps);
    HIP_TRY(hipSetDevice(h->dev));
    const int E = 7;
    if (h->profiling && (int)h->evpool.size() < E * n_steps) {
        while ((int)h->evpool.size() < E * n_steps) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            h->evpool.push_back(e);
        }
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (h->profiling) {
        if (h->evpool.size() < (size_t)E * n_steps + 2) {
            for (int i = 0; i < 2; ++i) {
                hipEvent_t e;
                HIP_TRY(hipEventCreate(&e));
                h->evpool.push_back(e);
            }
        }
        t0 = h->evpool[(size_t)E * n_steps];
        t1 = h->evpool[(size_t)E * n_steps + 1];
        HIP_TRY(hipEventRecord(t0, h->s0));
    }
    for (int s = 0; s < n_steps; ++s) {
        int n = h->steps_done + 1;
        bool flag = s >= n_steps - 2;
        int rc = launch_step(h, n, flag, h->profiling ? &h->evpool[(size_t)E * s] : nullptr);
        if (rc) return rc;
        h->steps_done = n;
    }
    if (h->profiling) {
        HIP_TRY(hipStreamWaitEvent(h->s0, h->ev_step, 0));
        HIP_TRY(hipEventRecord(t1, h->s0));
    }
    HIP_TRY(hipStreamSynchronize(h->s0));
    HIP_TRY(hipStreamSynchronize(h->s1));
    int err = 0;
    HIP_TRY(hipMemcpy(&err, h->D.err, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
        if (err & ERR_GEOMETRY) return fail(LBMIB_E_GEOMETRY, "an immersed-boundary node left the interior margin (4..nx-3, 4..ny-3)");
        if (err & ERR_CAPACITY) return fail(LBMIB_E_CAPACITY, "an immersed-boundary list overflowed");
        if (err & ERR_SOLVE) return fail(LBMIB_E_SOLVE, "immersed-boundary system not positive definite");
    }
    if (h->profiling && n_steps > 0) {
        float ms;
        HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
        h->tms[0] += ms;
        for (int s = 0; s < n_steps; ++s) {
            hipEvent_t* ev = &h->evpool[(size_t)E * s];
            HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1])); h->tms[1] += ms;
            HIP_TRY(hipEventElapsedTime(&ms, ev[2], ev[3])); h->tms[2] += ms;
            HIP_TRY(hipEventElapsedTime(&ms, ev[4], ev[5])); h->tms[3] += ms;
            HIP_TRY(hipEventElapsedTime(&ms, ev[1], ev[6])); h->tms[4] += ms;
        }
        h->tms[5] += n_steps;
    }
    if (h->p.nan_guard) {
        int* flag;
        if (dalloc(h, &flag, 1)) return LBMIB_E_HIP;
        HIP_TRY(hipMemset(flag, 0, sizeof(int)));
        hipLaunchKernelGGL(k_nan_check, dim3(1024), dim3(256), 0, h->s0,
                           (const double*)h->D.f[h->steps_done & 1], 9 * h->D.N, flag);
        int f = 0;
        HIP_TRY(hipMemcpy(&f, flag, sizeof(int), hipMemcpyDeviceToHost));
        h->allocs.pop_back();
        (void)hipFree(flag);
        if (f) return fail(LBMIB_E_NAN, "NaN in the populations after step %d", h->steps_done);
    }
    return 0;
}

int lbmib_steps_done(const lbmib_handle* h, int32_t* n)
{
    if (!h || !n) return fail(LBMIB_E_ARG, "null argument");
    *n = h->steps_done;
    return 0;
}

int lbmib_max_steps(const lbmib_handle* h, int32_t* n)
{
    if (!h || !n) return fail(LBMIB_E_ARG, "null argument");
    if (!h->have_motion) return fail(LBMIB_E_STATE, "no motion table");
    *n = h->max_steps;
    return 0;
}

int lbmib_get_history(lbmib_handle* h, double* rows, int32_t max_rows, int32_t* n_rows)
{
    if (!h) return fail(LBMIB_E_ARG, "null handle");
    if (n_rows) *n_rows = h->steps_done;
    if (!rows || max_rows <= 0 || h->steps_done == 0) return 0;
    HIP_TRY(hipSetDevice(h->dev));
    const int nw = h->p.n_wings, cols = 4 + 12 * nw;
    int n = std::min(max_rows, h->steps_done);
    std::vector<double> coef((size_t)n * 3 * nw);
    HIP_TRY(hipMemcpy(coef.data(), h->D.hist, sizeof(double) * coef.size(), hipMemcpyDeviceToHost));
    const double dt = h->p.scale * h->p.U;
    for (int r = 0; r < n; ++r) {
        double* row = rows + (size_t)r * cols;
        const double* c = &coef[(size_t)r * 3 * nw];
        const double* mrow = &h->step_rows[(size_t)r * h->ncols];
        int col = 0;
        row[col++] = (double)r * dt;  // time=(cycle1-1)*(scale*UU) (solution.f90:289)
        double fx = 0.0, fy = 0.0, mt = 0.0;
        for (int w = 0; w < nw; ++w) { row[col++] = c[3 * w]; row[col++] = c[3 * w + 1]; row[col++] = c[3 * w + 2]; }
        for (int w = 0; w < nw; ++w) { fx = fx + c[3 * w]; fy = fy + c[3 * w + 1]; }
        for (int w = 0; w < nw; ++w) mt = mt + c[3 * w + 2];
        row[col++] = fx; row[col++] = fy; row[col++] = mt;
        for (int w = 0; w < nw; ++w) {
            const double* m = mrow + 1 + 9 * w;  // angle, omega, omgac, sa, sau, saac, sb, sbv, sbac
            row[col++] = m[3]; row[col++] = m[4]; row[col++] = m[5];
            row[col++] = m[6]; row[col++] = m[7]; row[col++] = m[8];
            row[col++] = m[0]; row[col++] = m[1]; row[col++] = m[2];
        }
    }
    return 0;
}

int lbmib_get_forces(lbmib_handle* h, double* wing_force, double* wing_moment)
{
    if (!h) return fail(LBMIB_E_ARG, "null handle");
    if (h->steps_done == 0) return fail(LBMIB_E_STATE, "no step taken yet");
    HIP_TRY(hipSetDevice(h->dev));
    const int nw = h->p.n_wings;
    std::vector<double> c(3 * nw);
    HIP_TRY(hipMemcpy(c.data(), h->D.hist + (size_t)(h->steps_done - 1) * 3 * nw, sizeof(double) * 3 * nw,
                      hipMemcpyDeviceToHost));
    for (int w = 0; w < nw; ++w) {
        if (wing_force) { wing_force[2 * w] = c[3 * w]; wing_force[2 * w + 1] = c[3 * w + 1]; }
        if (wing_moment) wing_moment[w] = c[3 * w + 2];
    }
    return 0;
}

int lbmib_download(lbmib_handle* h, double* rho, double* u, double* v, double* f)
{
    if (!h) return fail(LBMIB_E_ARG, "null handle");
    HIP_TRY(hipSetDevice(h->dev));
    HIP_TRY(hipDeviceSynchronize());
    const long long N = h->D.N;
    const int slot = h->steps_done & 1;
    if (h->last_flag_step[slot] != h->steps_done)
        return fail(LBMIB_E_STATE, "macroscopic fields of step %d were not recorded", h->steps_done);
    const double* m = h->D.macro[slot];
    if (rho) HIP_TRY(hipMemcpy(rho, m, sizeof(double) * N, hipMemcpyDeviceToHost));
    if (u) HIP_TRY(hipMemcpy(u, m + N, sizeof(double) * N, hipMemcpyDeviceToHost));
    if (v) HIP_TRY(hipMemcpy(v, m + 2 * N, sizeof(double) * N, hipMemcpyDeviceToHost));
    if (f) HIP_TRY(hipMemcpy(f, h->D.f[slot], sizeof(double) * 9 * N, hipMemcpyDeviceToHost));
    return 0;
}

int lbmib_upload_state(lbmib_handle* h, const double* f)
{
    if (!h || !f) return fail(LBMIB_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(h->dev));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(h->D.f[h->steps_done & 1], f, sizeof(double) * 9 * h->D.N, hipMemcpyHostToDevice));
    return 0;
}

int lbmib_residual(lbmib_handle* h, double* err)
{
    if (!h || !err) return fail(LBMIB_E_ARG, "null argument");
    if (h->steps_done < 1) return fail(LBMIB_E_STATE, "no step taken yet");
    const int n = h->steps_done, sc = n & 1, sp = sc ^ 1;
    if (h->last_flag_step[sc] != n || h->last_flag_step[sp] != n - 1)
        return fail(LBMIB_E_STATE, "velocity fields of steps %d and %d were not recorded", n - 1, n);
    HIP_TRY(hipSetDevice(h->dev));
    const int blocks = 1024;
    hipLaunchKernelGGL(k_residual, dim3(blocks), dim3(256), 0, h->s0, (const double*)h->D.macro[sc],
                       (const double*)h->D.macro[sp], h->D.nx, h->D.ny, h->D.N, h->D.red);
    std::vector<double> part(2 * blocks);
    HIP_TRY(hipMemcpyAsync(part.data(), h->D.red, sizeof(double) * part.size(), hipMemcpyDeviceToHost, h->s0));
    HIP_TRY(hipStreamSynchronize(h->s0));
    double t1 = 0.0, t2 = 0.0;
    for (int b = 0; b < blocks; ++b) { t1 += part[2 * b]; t2 += part[2 * b + 1]; }
    *err = std::sqrt(t1) / (std::sqrt(t2) + 1e-30);
    return 0;
}

int lbmib_ib_nodes(lbmib_handle* h, int32_t* n, int32_t* pi, int32_t* pj, int32_t* wing, double* glx, double* gly)
{
    if (!h || !n) return fail(LBMIB_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(h->dev));
    HIP_TRY(hipDeviceSynchronize());
    int na = 0;
    HIP_TRY(hipMemcpy(&na, h->D.nadj, sizeof(int), hipMemcpyDeviceToHost));
    *n = na;
    if (na == 0) return 0;
    if (pi || pj || wing) {
        std::vector<int> adj(3 * (size_t)na);
        HIP_TRY(hipMemcpy(adj.data(), h->D.adj, sizeof(int) * adj.size(), hipMemcpyDeviceToHost));
        for (int a = 0; a < na; ++a) {
            if (pi) pi[a] = adj[3 * a];
            if (pj) pj[a] = adj[3 * a + 1];
            if (wing) wing[a] = adj[3 * a + 2];
        }
    }
    if (glx) HIP_TRY(hipMemcpy(glx, h->D.glx, sizeof(double) * na, hipMemcpyDeviceToHost));
    if (gly) HIP_TRY(hipMemcpy(gly, h->D.gly, sizeof(double) * na, hipMemcpyDeviceToHost));
    return 0;
}

int lbmib_set_profiling(lbmib_handle* h, int32_t on)
{
    if (!h) return fail(LBMIB_E_ARG, "null handle");
    h->profiling = on != 0;
    for (double& t : h->tms) t = 0.0;
    return 0;
}

int lbmib_get_timings(lbmib_handle* h, double* ms6)
{
    if (!h || !ms6) return fail(LBMIB_E_ARG, "null argument");
    for (int i = 0; i < 6; ++i) ms6[i] = h->tms[i];
    return 0;
}

}  // extern "C"
