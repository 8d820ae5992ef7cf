// lbmib_api.hip -- host side of liblbmib: the C ABI declared in include/lbmib.h.
//
// Owns the device state of a case and builds the per-step kinematics from
// the motion table exactly as the reference does (mode.f90:199-283).  Steps
// run on the case's own engine (lbmib_engine.hip), or on a batch's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "lbmib_host.h"

using namespace lbmib;

namespace lbmib {

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

}  // namespace lbmib

using namespace lbmib;

namespace {

constexpr int GEO_BATCH = LBMIB_GEOMETRY_BATCH;  // slots = 2 * batch

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

namespace {

template <typename T>
int dalloc(lbmib_handle* h, T** ptr, size_t count)
{
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sizeof(T) * std::max<size_t>(count, 1));
    if (e != hipSuccess)
        return fail(LBMIB_E_HIP, "hipMalloc of %zu bytes failed: %s", sizeof(T) * count, hipGetErrorString(e));
    h->allocs.push_back(p);
    h->device_bytes += (long long)(sizeof(T) * std::max<size_t>(count, 1));
    *ptr = (T*)p;
    return 0;
}

// Device fills on the null stream, which every later engine call and copy
// of the handle follows.
template <typename T>
int fill_dev(T* d, T v, size_t n)
{
    if (n == 0) return 0;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill<T>, dim3(blocks), dim3(256), 0, nullptr, d, (long long)n, v);
    HIP_TRY(hipGetLastError());
    return 0;
}

int fill_i32(int* d, int v, size_t n) { return fill_dev<int>(d, v, n); }

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

static int scratch_on(int dev, double** out);  // staging buffer of lbmib_surface_pressure_noalloc

extern "C" {

const char* lbmib_last_error(void) { return g_err.c_str(); }

const char* lbmib_version(void) { return "lbmib 0.1 gfx950"; }

int lbmib_destroy(lbmib_handle* h)
{
    if (!h) return 0;
    (void)hipSetDevice(h->dev);
    if (h->batch) (void)lbmib_batch_remove(h->batch, h);
    engine_destroy(h->own);
    h->own = nullptr;
    (void)hipDeviceSynchronize();
    for (void* p : h->allocs) (void)hipFree(p);
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
    if (p->precision != LBMIB_PREC_FP64 && p->precision != LBMIB_PREC_FP32DEV)
        return fail(LBMIB_E_ARG, "precision %d unknown (LBMIB_PREC_FP64 or LBMIB_PREC_FP32DEV)", p->precision);
    if (p->smooth != 0 && p->smooth != 1) return fail(LBMIB_E_ARG, "smooth must be 0 or 1");
    if (p->ib_solver < LBMIB_SOLVER_AUTO || p->ib_solver > LBMIB_SOLVER_SEPARATOR)
        return fail(LBMIB_E_ARG, "ib_solver %d unknown", p->ib_solver);
    if ((long long)p->n_wings * (p->pts_per_wing - 1) > MAXSEG)
        return fail(LBMIB_E_ARG, "at most %d marker segments supported", MAXSEG);

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
#define TRY(x)                                   \
    do {                                         \
        rc = (x);                                \
        if (rc) { lbmib_destroy(h); return rc; } \
    } while (0)
#define HTRY(x)                                                                   \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            lbmib_destroy(h);                                                     \
            return fail(LBMIB_E_HIP, "%s failed: %s", #x, hipGetErrorString(e_)); \
        }                                                                         \
    } while (0)
    HTRY(hipSetDevice(h->dev));
    {
        double* sbuf = nullptr;  // the device's staging buffer, recorded once per device
        TRY(scratch_on(h->dev, &sbuf));
    }
    TRY(engine_create(h->dev, &h->own));
    h->own->members.push_back(h);
    h->in_xb0.assign(xb0, xb0 + (size_t)p->n_wings * p->pts_per_wing);
    h->in_yb0.assign(yb0, yb0 + (size_t)p->n_wings * p->pts_per_wing);
    const int nx = p->nx, ny = p->ny, nw = p->n_wings, ppw = p->pts_per_wing;
    const long long N = (long long)nx * ny;
    DevCase& D = h->D;
    std::memset(&D, 0, sizeof D);
    D.nx = nx; D.ny = ny; D.N = N;
    D.ox = p->ox; D.oy = p->oy; D.scale = p->scale; D.U = p->U; D.xc = p->xc; D.yc = p->yc;
    for (int k = 0; k < 9; ++k) D.R.s[k] = p->s[k];
    // weights and the FP32 deviation-form constants (rest moments of W)
    D.prec = p->precision;
    D.rho0 = 0.0;
    for (int k = 0; k < 9; ++k) {
        D.W[k] = feq_ref(1.0, 0.0, 0.0, k);
        D.rho0 += D.W[k];
    }
    D.K = make_devconst(p->s, D.W);
    D.smooth = p->smooth;
    D.nw = nw; D.ppw = ppw; D.spw = ppw - 1; D.nseg = nw * (ppw - 1);
    D.batch = GEO_BATCH;
    D.nslot = 2 * GEO_BATCH;
    D.cs = (double)std::sqrt(3.0f) / 3.0;  // REAL(4) sqrt (boundary_condition.f90:304)

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
    // IB node capacity: a node belongs to the set when it lies in the
    // +-2.5 lattice-unit box (at most 6 x 6 nodes) of a marker; sweeping that
    // box along a contour of length P covers at most 6 (|cos a| + |sin a|) <=
    // 8.49 nodes per lattice unit of arc, so 10 P + 256 per wing bounds the
    // set; a set that still overflows reports LBMIB_E_CAPACITY.
    {
        double cap = 0.0;
        for (int w = 0; w < nw; ++w) {
            double P = 0.0;
            for (int i = 0; i < ppw - 1; ++i) P += h->arc_l[w * (ppw - 1) + i];
            cap += 10.0 * P + 256.0;
        }
        D.cap = (int)std::min((double)D.nseg * BOXMAX, std::ceil(cap));
    }
    zigzag(nw, D.spw, h->perm);
    h->iperm.assign(D.nseg, 0);
    for (int q = 0; q < D.nseg; ++q) h->iperm[h->perm[q]] = q;

    // lattice
    if (D.prec == LBMIB_PREC_FP64) {
        TRY(dalloc(h, &D.f[0], 9 * (size_t)N));
        TRY(dalloc(h, &D.f[1], 9 * (size_t)N));
    } else {
        TRY(dalloc(h, &D.g[0], 9 * (size_t)N));
        TRY(dalloc(h, &D.g[1], 9 * (size_t)N));
    }
    TRY(dalloc(h, &D.f0L, 9 * (size_t)ny));
    TRY(dalloc(h, &D.f0R, 9 * (size_t)ny));
    {
        double* body = nullptr;  // body offset terms, zero until the dynamics move the body
        TRY(dalloc(h, &body, 32));
        HIP_TRY(hipMemset(body, 0, sizeof(double) * 32));
        D.body = body;
    }
    TRY(dalloc(h, &D.macro[0], 3 * (size_t)N));
    TRY(dalloc(h, &D.macro[1], 3 * (size_t)N));
    // init (solution.f90:34-59): f = f0 = feq(1, U, 0); rho = 1, u = U, v = 0;
    // filled on the device, plane by plane
    {
        double feq0[9];
        for (int k = 0; k < 9; ++k) feq0[k] = feq_ref(1.0, p->U, 0.0, k);
        for (int k = 0; k < 9; ++k) {
            if (D.prec == LBMIB_PREC_FP64) {
                TRY(fill_dev<double>(D.f[0] + (size_t)k * N, feq0[k], (size_t)N));
                TRY(fill_dev<double>(D.f[1] + (size_t)k * N, feq0[k], (size_t)N));
            } else {
                TRY(fill_dev<float>(D.g[0] + (size_t)k * N, (float)(feq0[k] - D.W[k]), (size_t)N));
                TRY(fill_dev<float>(D.g[1] + (size_t)k * N, (float)(feq0[k] - D.W[k]), (size_t)N));
            }
            TRY(fill_dev<double>(D.f0L + (size_t)k * ny, feq0[k], (size_t)ny));
            TRY(fill_dev<double>(D.f0R + (size_t)k * ny, feq0[k], (size_t)ny));
        }
        TRY(fill_dev<double>(D.macro[0], 1.0, (size_t)N));
        TRY(fill_dev<double>(D.macro[0] + N, p->U, (size_t)N));
        TRY(fill_dev<double>(D.macro[0] + 2 * N, 0.0, (size_t)N));
    }
    // markers
    double *dxb0, *dyb0;
    TRY(dalloc(h, &dxb0, npts));
    TRY(dalloc(h, &dyb0, npts));
    HTRY(hipMemcpy(dxb0, h->xb0.data(), sizeof(double) * npts, hipMemcpyHostToDevice));
    HTRY(hipMemcpy(dyb0, h->yb0.data(), sizeof(double) * npts, hipMemcpyHostToDevice));
    D.xb0 = dxb0; D.yb0 = dyb0;
    const int ns = D.nseg;
    int *dperm, *diperm;
    TRY(dalloc(h, &dperm, ns)); TRY(dalloc(h, &diperm, ns));
    HTRY(hipMemcpy(dperm, h->perm.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
    HTRY(hipMemcpy(diperm, h->iperm.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
    D.perm = dperm; D.iperm = diperm;
    TRY(dalloc(h, &D.rhs, 2 * (size_t)ns));
    TRY(dalloc(h, &D.y, 2 * (size_t)ns));
    TRY(dalloc(h, &D.err, 2));  // [0] ErrBits, [1] NaN-guard flag
    TRY(dalloc(h, &D.red, 2 * 4096));
    TRY(fill_i32(D.err, 0, 2));
    D.bw = geometric_bandwidth(h);
    if (D.bw >= ns) D.bw = ns - 1;
    // the fills ran on the null stream; the engine's streams do not wait for it
    HTRY(hipStreamSynchronize(nullptr));
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
    if (h->have_motion) return fail(LBMIB_E_STATE, "the motion table can be set once per handle");
    HIP_TRY(hipSetDevice(h->dev));
    h->motion_in.assign(rows, rows + (size_t)nrows * ncols);
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
    DevCase& D = h->D;
    // lattice window holding every IB node of every step, and whether wings
    // ever come within coupling distance of each other
    {
        const int ppw = h->p.pts_per_wing;
        double bx0[8], bx1[8], by0[8], by1[8];
        for (int w = 0; w < nw; ++w) {
            bx0[w] = by0[w] = 1e300; bx1[w] = by1[w] = -1e300;
            for (int i = 0; i < ppw; ++i) {
                bx0[w] = std::min(bx0[w], h->xb0[w * ppw + i]); bx1[w] = std::max(bx1[w], h->xb0[w * ppw + i]);
                by0[w] = std::min(by0[w], h->yb0[w * ppw + i]); by1[w] = std::max(by1[w], h->yb0[w * ppw + i]);
            }
        }
        const double margin = 6.0 * h->p.scale;
        double X0 = 1e300, X1 = -1e300, Y0 = 1e300, Y1 = -1e300;
        bool close = false;
        for (int n = 0; n < maxn; ++n) {
            double lo[8][2], hi[8][2];
            for (int w = 0; w < nw; ++w) {
                const double* k = &kin[((size_t)n * nw + w) * KIN];
                lo[w][0] = lo[w][1] = 1e300; hi[w][0] = hi[w][1] = -1e300;
                const double cx[4] = {bx0[w], bx1[w], bx0[w], bx1[w]}, cy[4] = {by0[w], by0[w], by1[w], by1[w]};
                for (int q = 0; q < 4; ++q) {
                    double X = k[1] * cx[q] + k[0] * cy[q] + k[3] + h->p.xc;
                    double Y = -k[0] * cx[q] + k[1] * cy[q] + k[5] + h->p.yc;
                    lo[w][0] = std::min(lo[w][0], X); hi[w][0] = std::max(hi[w][0], X);
                    lo[w][1] = std::min(lo[w][1], Y); hi[w][1] = std::max(hi[w][1], Y);
                }
                X0 = std::min(X0, lo[w][0]); X1 = std::max(X1, hi[w][0]);
                Y0 = std::min(Y0, lo[w][1]); Y1 = std::max(Y1, hi[w][1]);
            }
            for (int a = 0; a < nw && !close; ++a)
                for (int b = 0; b < a && !close; ++b)
                    if (lo[a][0] - margin <= hi[b][0] && lo[b][0] - margin <= hi[a][0] &&
                        lo[a][1] - margin <= hi[b][1] && lo[b][1] - margin <= hi[a][1])
                        close = true;
        }
        if (close) D.bw = D.nseg - 1;
        const double sc = h->p.scale;
        // with the dynamics the body drifts up to one node from its prescribed
        // path before the domain follows it
        const int mx = h->dyn.enabled ? 9 : 6;
        int i0 = (int)std::floor((X0 - h->p.ox) / sc) - mx, i1 = (int)std::ceil((X1 - h->p.ox) / sc) + mx;
        int j0 = (int)std::floor((Y0 - h->p.oy) / sc) - 6, j1 = (int)std::ceil((Y1 - h->p.oy) / sc) + 6;
        i0 = std::max(0, i0); j0 = std::max(0, j0);
        i1 = std::min(D.nx - 1, i1); j1 = std::min(D.ny - 1, j1);
        if (i1 < i0 || j1 < j0) return fail(LBMIB_E_GEOMETRY, "the body never overlaps the lattice");
        D.gi0 = i0; D.gj0 = j0; D.gw = i1 - i0 + 1; D.gh = j1 - j0 + 1;
    }
    if (D.bw > MAXBW)
        return fail(LBMIB_E_ARG, "immersed-boundary bandwidth %d exceeds the banded solver's limit of %d "
                    "(wings too close together, or markers much denser than the lattice)", D.bw, MAXBW);
    D.ws = 64;
    D.npad = (D.nseg + 63) / 64 * 64;
    // separator form of the IB solve: rows mT .. mT+bw-1 split the band into
    // two halves solved as concurrent chains (a different round-off order
    // than the oracle's single chain, so the FP64 drop-in uses it only when
    // asked for)
    {
        const int sol = h->p.ib_solver;
        const bool want = sol == LBMIB_SOLVER_SEPARATOR || (sol == LBMIB_SOLVER_AUTO && D.prec == LBMIB_PREC_FP32DEV);
        const int b = D.bw, mT = (D.nseg - b) / 2, mB = D.nseg - b - mT;
        D.compact = b < CW ? 1 : 0;
        const int np2 = (std::max(mT, mB) + 127) / 128 * 128;
        const bool fits = b >= 1 && b <= SEPMAXBW && sep_len1(b) <= SEP_L1 && sep_len2(b) <= SEP_L2 && mT >= b &&
                          mB >= b && mT >= 64;
        // AUTO falls back to the single chain where the separator form does not apply
        D.dissect = want && D.compact && fits && solve2_unit(np2) > 0 ? 1 : 0;
        D.mT = D.dissect ? mT : 0;
        D.mB = D.dissect ? mB : 0;
        D.np2 = D.dissect ? np2 : 0;
        if (sol == LBMIB_SOLVER_SEPARATOR && !D.dissect) {
            if (D.compact && fits)
                return fail(LBMIB_E_ARG, "too many segments for the separator solve's LDS (%d segments)", D.nseg);
            return fail(LBMIB_E_ARG, "the separator solve needs at least 2*64 + bandwidth segments and bandwidth "
                        "<= %d (have %d segments, bandwidth %d)", SEPMAXBW, D.nseg, b);
        }
    }
    double* dkin;
    if (dalloc(h, &dkin, kin.size())) return LBMIB_E_HIP;
    HIP_TRY(hipMemcpy(dkin, kin.data(), sizeof(double) * kin.size(), hipMemcpyHostToDevice));
    D.kin = dkin;
    D.steps_cap = maxn;
    if (dalloc(h, &D.hist, (size_t)maxn * 3 * nw)) return LBMIB_E_HIP;
    if (h->dyn.enabled) {
        // each step's geometry follows the body state after the step before
        if (dalloc(h, &D.raw, (size_t)maxn * 3 * nw)) return LBMIB_E_HIP;
        D.batch = 1;
        h->own->per_step = true;
    }
    // slot pools
    {
        const long long ns = D.nseg, cap = D.cap, W = D.bw + 1, grid = (long long)D.gw * D.gh;
        auto al = [](long long x) { return (x + 63) / 64 * 64; };
        SlotInts& io = D.io;
        long long o = 0;
        io.claim = o; o += al(grid);
        io.nodeidx = o; o += al(grid);
        io.box = o; o += al(4 * ns);
        io.draw_idx = o; o += al(ns * BOXMAX);
        io.nraw_s = o; o += al(cap * NCAP);
        io.adj = o; o += al(3 * cap);
        io.nadj = o; o += al(1);
        io.dcnt = o; o += al(ns);
        io.didx = o; o += al(ns * BOXMAX);
        io.nmkcnt = o; o += al(cap);
        io.nmk = o; o += al(cap * NCAP);
        io.total = o;
        SlotDoubles& dd = D.dof;
        o = 0;
        dd.xb2 = o; o += al(ns);
        dd.yb2 = o; o += al(ns);
        dd.ub = o; o += al(ns);
        dd.vb = o; o += al(ns);
        dd.phi = o; o += al(12 * ns);
        dd.draw_val = o; o += al(ns * BOXMAX);
        dd.nraw_w = o; o += al(cap * NCAP);
        dd.dval = o; o += al(ns * BOXMAX);
        dd.nmkv = o; o += al(cap * NCAP);
        dd.G = o; o += al(ns * W);
        dd.coef = o; o += al(D.compact ? (long long)(2 * D.npad + 256) * CW : (long long)(2 * D.npad + 64) * 2 * D.ws);
        dd.dinv = o; o += al(D.npad);
        {
            const long long b = D.dissect ? D.bw : 0, np2 = D.np2;
            dd.coef2 = o; o += al(D.dissect ? 2 * (2 * np2 + 256) * CW : 0);
            dd.dinv2 = o; o += al(2 * np2);
            dd.ltail2 = o; o += al(2 * b * (b + 1));
            dd.zsep = o; o += al(5 * b * b);
        }
        dd.rhoadj = o; o += al(cap);
        dd.glx = o; o += al(cap);
        dd.gly = o; o += al(cap);
        dd.ust = o; o += al(cap);
        dd.vst = o; o += al(cap);
        dd.total = o;
        if (dalloc(h, &D.ipool, (size_t)D.nslot * io.total)) return LBMIB_E_HIP;
        if (dalloc(h, &D.dpool, (size_t)D.nslot * dd.total)) return LBMIB_E_HIP;
        // the solve coefficients of the padding rows stay zero
        HIP_TRY(hipMemset(D.dpool, 0, sizeof(double) * (size_t)D.nslot * dd.total));
        for (int g = 0; g < D.nslot; ++g) {
            Slot S = slot_of(D, g);
            if (fill_i32(S.claim, INT_MAX, (size_t)grid)) return LBMIB_E_HIP;
            if (fill_i32(S.nodeidx, -1, (size_t)grid)) return LBMIB_E_HIP;
            if (fill_i32(S.nadj, 0, 1)) return LBMIB_E_HIP;
        }
        // the fills ran on the null stream; the engine's streams do not wait for it
        HIP_TRY(hipStreamSynchronize(nullptr));
    }
    h->have_motion = true;
    return 0;
}

}  // extern "C"

namespace lbmib {

int ensure_mid_plane(lbmib_handle* h)
{
    DevCase& D = h->D;
    const size_t n = 9 * (size_t)D.N;
    if (D.prec == LBMIB_PREC_FP64) {
        if (!D.f[2]) return dalloc(h, &D.f[2], n);
    } else {
        if (!D.g[2]) return dalloc(h, &D.g[2], n);
    }
    return 0;
}

int nan_guard_check(lbmib_handle* h)
{
    if (!h->p.nan_guard || h->steps_done == 0) return 0;
    int* flag = h->D.err + 1;
    if (fill_i32(flag, 0, 1)) return LBMIB_E_HIP;
    hipStream_t st = nullptr;  // every engine call has synchronised its own streams
    if (h->D.prec == LBMIB_PREC_FP64)
        hipLaunchKernelGGL(k_nan_check<double>, dim3(1024), dim3(256), 0, st,
                           (const double*)h->D.f[state_plane(h)], 9 * h->D.N, flag);
    else
        hipLaunchKernelGGL(k_nan_check<float>, dim3(1024), dim3(256), 0, st,
                           (const float*)h->D.g[state_plane(h)], 9 * h->D.N, flag);
    int f = 0;
    HIP_TRY(hipMemcpyAsync(&f, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (f) return fail(LBMIB_E_NAN, "NaN in the populations after step %d", h->steps_done);
    return 0;
}

}  // namespace lbmib

extern "C" {

// ---- rigid-body dynamics (the host side of the reference's own host code) ----

// update_body_extra_from_rel (mode.f90:285-307): running sums over wings
static void body_extra(lbmib_handle* h)
{
    double av = 0.0, as = 0.0, ap = 0.0;
    for (int w = 0; w < 8; ++w) h->ext[0][w] = h->ext[1][w] = h->ext[2][w] = 0.0;
    for (int w = 0; w < h->p.n_wings; ++w) {
        av = av + h->rel[2][w];
        as = as + h->rel[4][w];
        ap = ap + h->rel[6][w];
        h->ext[0][w] = av;
        h->ext[1][w] = as;
        h->ext[2][w] = ap;
    }
}

// rk4 (boundary_condition.f90:3168-3202): the acceleration is held over the
// four stages; the pitch equation reads alphad and the last moment
static void body_rk4(const lbmib_handle* h, double s1, double& s, double& s2, double v1, double& v, double a1,
                     double jd1, double& jd, double jsd1, double& jsd, double p1, double p2)
{
    const double alphad = h->dyn.alphad, mom = h->mom;
    double k11 = v1, k21 = a1;
    double k12 = v1 + 0.5 * k21, k22 = a1;
    double k13 = v1 + 0.5 * k22, k23 = a1;
    double k14 = v1 + k23, k24 = a1;
    s = s1 + (k11 + 2.0 * k12 + 2.0 * k13 + k14) / 6.0;
    v = v1 + (k21 + 2.0 * k22 + 2.0 * k23 + k24) / 6.0;
    s2 = s2 + (k11 + 2.0 * k12 + 2.0 * k13 + k14) / 6.0;
    k11 = jsd1;
    k21 = p1 * (jd1 + alphad) + p2 * mom;
    k12 = jsd1 + 0.5 * k21;
    k22 = p1 * (jd1 + alphad + 0.5 * k11) + p2 * mom;
    k13 = jsd1 + 0.5 * k22;
    k23 = p1 * (jd1 + alphad + 0.5 * k12) + p2 * mom;
    k14 = jsd1 + k23;
    k24 = p1 * (jd1 + alphad + k13) + p2 * mom;
    jd = jd1 + (k11 + 2.0 * k12 + 2.0 * k13 + k14) / 6.0;
    jsd = jsd1 + (k21 + 2.0 * k22 + 2.0 * k23 + k24) / 6.0;
}

// the dynamics block of immersed_boundary_Wu (boundary_condition.f90:1676-1719)
// on the raw per-wing forces of the step
static void body_update(lbmib_handle* h, const double* raw)
{
    const double sc = h->p.scale, UU = h->p.U;
    const lbmib_dynamics& d = h->dyn;
    double (*r)[8] = h->rel;
    double force1 = 0.0;
    for (int w = 0; w < h->p.n_wings; ++w) force1 = force1 + raw[3 * w];
    h->mom = raw[2];
    r[1][0] = r[0][0];
    if (!d.closed_boundary)
        r[0][0] = force1 / (d.solid_den / (sc * sc));
    else
        r[0][0] = force1 / (d.solid_den * d.area_airfoil / (sc * sc));
    r[3][0] = r[2][0];
    r[5][0] = r[4][0];
    r[7][0] = r[8][0];
    r[9][0] = r[10][0];
    if (!d.rk4) {
        r[2][0] = r[0][0] + r[2][0];
        r[4][0] = r[4][0] + r[3][0];
        r[6][0] = r[6][0] + r[3][0];
        r[8][0] = r[7][0] + r[9][0];
        r[10][0] = r[9][0] + d.par1 * UU * UU * sc * sc * (r[8][0] + d.alphad) + d.par2 * (sc * sc * sc * sc) * h->mom;
    } else {
        body_rk4(h, r[5][0], r[4][0], r[6][0], r[3][0], r[2][0], r[0][0], r[7][0], r[8][0], r[9][0], r[10][0],
                 d.par1 * UU * UU * sc * sc, d.par2 * (sc * sc * sc * sc));
    }
    body_extra(h);
}

// the body offset terms the geometry and force kernels add (airfoil.f90:262,
// boundary_condition.f90:1548, 1672)
static int upload_body(lbmib_handle* h)
{
    double b[32] = {};
    for (int w = 0; w < h->p.n_wings; ++w) {
        b[4 * w] = h->ext[1][w] / (1.0 / h->p.scale);
        b[4 * w + 1] = (double)h->movenum * h->p.scale;
        b[4 * w + 2] = h->ext[0][w];
        b[4 * w + 3] = h->ext[2][w];
    }
    HIP_TRY(hipMemcpy(const_cast<double*>(h->D.body), b, sizeof b, hipMemcpyHostToDevice));
    std::memcpy(h->body_terms, b, sizeof b);
    return 0;
}

// move_mesh on the device state after the last step: populations, the
// refilled edge column's stale f0, and the recorded rho/u/v of the step
// and the step before
static int move_domain(lbmib_handle* h, int dr)
{
    DevCase& D = h->D;
    const long long N = D.N;
    const int nx = D.nx, ny = D.ny, slot = state_plane(h), mslot = h->steps_done & 1;
    RingW w;
    for (int k = 0; k < 9; ++k) w.w[k] = D.W[k];
    const dim3 gp((unsigned)((9 * N + 255) / 256)), gr((unsigned)((9 * ny + 255) / 256)), b(256);
    double* f0col = dr == 1 ? D.f0L : D.f0R;
    if (D.prec == LBMIB_PREC_FP64) {
        hipLaunchKernelGGL(k_move_ring<double>, gr, b, 0, nullptr, (const double*)D.f[slot], f0col, nx, ny, N, dr, w);
        hipLaunchKernelGGL(k_move_pops<double>, gp, b, 0, nullptr, (const double*)D.f[slot], D.f[slot ^ 1], nx, N, dr);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(D.f[slot], D.f[slot ^ 1], sizeof(double) * 9 * N, hipMemcpyDeviceToDevice, nullptr));
    } else {
        hipLaunchKernelGGL(k_move_ring<float>, gr, b, 0, nullptr, (const float*)D.g[slot], f0col, nx, ny, N, dr, w);
        hipLaunchKernelGGL(k_move_pops<float>, gp, b, 0, nullptr, (const float*)D.g[slot], D.g[slot ^ 1], nx, N, dr);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(D.g[slot], D.g[slot ^ 1], sizeof(float) * 9 * N, hipMemcpyDeviceToDevice, nullptr));
    }
    if (h->last_flag_step[mslot] == h->steps_done) {
        const dim3 gm((unsigned)((3 * N + 255) / 256));
        hipLaunchKernelGGL(k_move_macro, gm, b, 0, nullptr, D.macro[mslot], D.macro[mslot ^ 1], nx, N, dr, 0);
        hipLaunchKernelGGL(k_move_macro, gm, b, 0, nullptr, D.macro[mslot], D.macro[mslot ^ 1], nx, N, dr, 1);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipDeviceSynchronize());
    return 0;
}

// One step with the dynamics on: the body offsets of the state so far, the
// step (its geometry built first), the body update from the step's raw
// forces, the airfoil_motion row, and the domain shift when the body has
// moved a node (main.f90:370-388).
static int dynamics_step(lbmib_handle* h)
{
    const int nw = h->p.n_wings;
    if (int rc = upload_body(h)) return rc;
    h->geo_fresh = true;
    if (int rc = engine_run(h->own, 1)) return rc;
    if (int rc = nan_guard_check(h)) return rc;
    const int n = h->steps_done;
    double raw[24];
    HIP_TRY(hipMemcpy(raw, h->D.raw + (size_t)(n - 1) * 3 * nw, sizeof(double) * 3 * nw, hipMemcpyDeviceToHost));
    body_update(h, raw);
    const double UU = h->p.U, sc = h->p.scale;
    h->mrows.resize((size_t)n * (1 + 3 * nw));
    double* m = &h->mrows[(size_t)(n - 1) * (1 + 3 * nw)];
    int mc = 0;
    m[mc++] = (double)(n - 1) * (sc * UU);
    for (int w = 0; w < nw; ++w) {
        m[mc++] = 0.0;
        m[mc++] = h->ext[0][w] / UU;
        m[mc++] = h->ext[1][w] / (1.0 / sc);
    }
    double& pos = h->rel[6][0];
    if (pos >= 1.0 || pos <= -1.0) {
        const int dr = pos >= 1.0 ? 1 : -1;
        if (int rc = move_domain(h, dr)) return rc;
        pos = dr == 1 ? pos - 1.0 : pos + 1.0;
        body_extra(h);
        h->movenum = h->movenum + dr;
    }
    return 0;
}

int lbmib_set_dynamics(lbmib_handle* h, const lbmib_dynamics* d)
{
    if (!h || !d) return fail(LBMIB_E_ARG, "null argument");
    if (h->have_motion) return fail(LBMIB_E_STATE, "set the dynamics before the motion table");
    if (h->batch) return fail(LBMIB_E_STATE, "the case belongs to a batch");
    if (d->enabled != 0 && d->enabled != 1) return fail(LBMIB_E_ARG, "enabled must be 0 or 1");
    if ((d->rk4 != 0 && d->rk4 != 1) || (d->closed_boundary != 0 && d->closed_boundary != 1) || d->reserved != 0)
        return fail(LBMIB_E_ARG, "rk4 and closed_boundary must be 0 or 1, reserved 0");
    if (d->enabled && !(d->solid_den > 0.0)) return fail(LBMIB_E_ARG, "solid_den must be positive");
    if (d->enabled && d->closed_boundary && !(d->area_airfoil > 0.0))
        return fail(LBMIB_E_ARG, "area_airfoil must be positive for a closed boundary");
    h->dyn = *d;
    return 0;
}

int lbmib_get_body_state(const lbmib_handle* h, double* state, int32_t n)
{
    if (!h || !state) return fail(LBMIB_E_ARG, "null argument");
    const int nw = h->p.n_wings;
    if (n < 11 * nw + 1) return fail(LBMIB_E_ARG, "room for %d values needed", 11 * nw + 1);
    int o = 0;
    for (int a = 0; a < 11; ++a)
        for (int w = 0; w < nw; ++w) state[o++] = h->rel[a][w];
    state[o] = (double)h->movenum;
    return 0;
}

int lbmib_set_body_state(lbmib_handle* h, const double* state, int32_t n)
{
    if (!h || !state) return fail(LBMIB_E_ARG, "null argument");
    const int nw = h->p.n_wings;
    if (n != 11 * nw + 1) return fail(LBMIB_E_ARG, "%d values expected", 11 * nw + 1);
    const double mv = state[11 * nw];
    if (!(mv == std::floor(mv)) || std::fabs(mv) > 1e9) return fail(LBMIB_E_ARG, "movenum must be an integer");
    int o = 0;
    for (int a = 0; a < 11; ++a)
        for (int w = 0; w < nw; ++w) h->rel[a][w] = state[o++];
    h->movenum = (int)mv;
    body_extra(h);
    return 0;
}

int lbmib_get_edge_f0(lbmib_handle* h, double* first, double* last)
{
    if (!h) return fail(LBMIB_E_ARG, "null handle");
    HIP_TRY(hipSetDevice(h->dev));
    HIP_TRY(hipDeviceSynchronize());
    const size_t n = 9 * (size_t)h->D.ny;
    if (first) HIP_TRY(hipMemcpy(first, h->D.f0L, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (last) HIP_TRY(hipMemcpy(last, h->D.f0R, sizeof(double) * n, hipMemcpyDeviceToHost));
    return 0;
}

int lbmib_set_edge_f0(lbmib_handle* h, const double* first, const double* last)
{
    if (!h) return fail(LBMIB_E_ARG, "null handle");
    HIP_TRY(hipSetDevice(h->dev));
    HIP_TRY(hipDeviceSynchronize());
    const size_t n = 9 * (size_t)h->D.ny;
    if (first) HIP_TRY(hipMemcpy(h->D.f0L, first, sizeof(double) * n, hipMemcpyHostToDevice));
    if (last) HIP_TRY(hipMemcpy(h->D.f0R, last, sizeof(double) * n, hipMemcpyHostToDevice));
    return 0;
}

int lbmib_get_motion_rows(const lbmib_handle* h, double* rows, int32_t max_rows, int32_t* n_rows)
{
    if (!h) return fail(LBMIB_E_ARG, "null handle");
    if (n_rows) *n_rows = h->steps_done;
    if (!rows || max_rows <= 0) return 0;
    const int nw = h->p.n_wings, cols = 1 + 3 * nw;
    const int n = std::min(max_rows, h->steps_done);
    const double dt = h->p.scale * h->p.U;
    for (int r = 0; r < n; ++r) {
        double* row = rows + (size_t)r * cols;
        if (h->dyn.enabled && (size_t)(r + 1) * cols <= h->mrows.size()) {
            std::memcpy(row, &h->mrows[(size_t)r * cols], sizeof(double) * cols);
        } else {
            row[0] = (double)r * dt;
            for (int c = 1; c < cols; ++c) row[c] = 0.0;
        }
    }
    return 0;
}

int lbmib_step(lbmib_handle* h, int32_t n_steps)
{
    if (!h) return fail(LBMIB_E_ARG, "null handle");
    if (!h->have_motion) return fail(LBMIB_E_STATE, "lbmib_set_motion_table must be called before lbmib_step");
    if (h->batch) return fail(LBMIB_E_STATE, "the case belongs to a batch: step it with lbmib_batch_step");
    if (n_steps < 0) return fail(LBMIB_E_ARG, "n_steps must be >= 0");
    if (h->steps_done + n_steps > h->max_steps)
        return fail(LBMIB_E_ARG, "step %d beyond the motion table (last step %d)", h->steps_done + n_steps, h->max_steps);
    if (h->dyn.enabled) {
        for (int s = 0; s < n_steps; ++s)
            if (int rc = dynamics_step(h)) return rc;
        return 0;
    }
    int rc = engine_run(h->own, n_steps);
    if (rc) return rc;
    return nan_guard_check(h);
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
    const int slot = h->steps_done & 1, plane = state_plane(h);
    if (h->last_flag_step[slot] != h->steps_done)
        return fail(LBMIB_E_STATE, "macroscopic fields of step %d were not recorded", h->steps_done);
    const double* m = h->D.macro[slot];
    if (rho) HIP_TRY(hipMemcpy(rho, m, sizeof(double) * N, hipMemcpyDeviceToHost));
    if (u) HIP_TRY(hipMemcpy(u, m + N, sizeof(double) * N, hipMemcpyDeviceToHost));
    if (v) HIP_TRY(hipMemcpy(v, m + 2 * N, sizeof(double) * N, hipMemcpyDeviceToHost));
    if (f && h->D.prec == LBMIB_PREC_FP64)
        HIP_TRY(hipMemcpy(f, h->D.f[plane], sizeof(double) * 9 * N, hipMemcpyDeviceToHost));
    if (f && h->D.prec != LBMIB_PREC_FP64) {
        std::vector<float> g((size_t)N);
        for (int k = 0; k < 9; ++k) {
            HIP_TRY(hipMemcpy(g.data(), h->D.g[plane] + (size_t)k * N, sizeof(float) * N, hipMemcpyDeviceToHost));
            for (long long c = 0; c < N; ++c) f[(size_t)k * N + c] = h->D.W[k] + (double)g[c];
        }
    }
    return 0;
}

int lbmib_upload_state(lbmib_handle* h, const double* f)
{
    if (!h || !f) return fail(LBMIB_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(h->dev));
    HIP_TRY(hipDeviceSynchronize());
    const long long N = h->D.N;
    const int plane = state_plane(h);
    if (h->D.prec == LBMIB_PREC_FP64) {
        HIP_TRY(hipMemcpy(h->D.f[plane], f, sizeof(double) * 9 * N, hipMemcpyHostToDevice));
    } else {
        std::vector<float> g((size_t)N);
        for (int k = 0; k < 9; ++k) {
            for (long long c = 0; c < N; ++c) g[c] = (float)(f[(size_t)k * N + c] - h->D.W[k]);
            HIP_TRY(hipMemcpy(h->D.g[plane] + (size_t)k * N, g.data(), sizeof(float) * N, hipMemcpyHostToDevice));
        }
    }
    return 0;
}

int lbmib_set_step(lbmib_handle* h, int32_t n)
{
    if (!h) return fail(LBMIB_E_ARG, "null handle");
    if (!h->have_motion) return fail(LBMIB_E_STATE, "lbmib_set_motion_table must come first");
    if (h->batch) return fail(LBMIB_E_STATE, "the case belongs to a batch");
    if (n < 0 || n > h->max_steps) return fail(LBMIB_E_ARG, "step %d outside 0..%d", n, h->max_steps);
    h->steps_done = n;
    h->geo_fresh = true;
    h->last_flag_step[0] = h->last_flag_step[1] = -1;
    const size_t mcols = 1 + 3 * (size_t)h->p.n_wings;
    if (h->mrows.size() > (size_t)n * mcols) h->mrows.resize((size_t)n * mcols);
    // a resumed run starts vor_mom_rate's history afresh, as a new process does
    if (h->vort_hist) {
        HIP_TRY(hipSetDevice(h->dev));
        HIP_TRY(hipMemset(h->vort_hist, 0, sizeof(double) * 3 * h->D.N));
    }
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
    hipStream_t st = nullptr;  // every engine call has synchronised its own streams
    hipLaunchKernelGGL(k_residual, dim3(blocks), dim3(256), 0, st, (const double*)h->D.macro[sc],
                       (const double*)h->D.macro[sp], h->D.nx, h->D.ny, h->D.N, h->D.red);
    std::vector<double> part(2 * blocks);
    HIP_TRY(hipMemcpyAsync(part.data(), h->D.red, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    double t1 = 0.0, t2 = 0.0;
    for (int b = 0; b < blocks; ++b) { t1 += part[2 * b]; t2 += part[2 * b + 1]; }
    *err = std::sqrt(t1) / (std::sqrt(t2) + 1e-30);
    return 0;
}

// Device memory of the post-processing calls, allocated on first use and
// kept until lbmib_destroy.
static int diag_alloc(lbmib_handle* h, double** p, size_t bytes)
{
    if (*p) return 0;
    void* q = nullptr;
    HIP_TRY(hipMalloc(&q, bytes));
    h->allocs.push_back(q);
    h->device_bytes += (long long)bytes;
    *p = static_cast<double*>(q);
    return 0;
}

int lbmib_vorticity(lbmib_handle* h, int32_t rate, double* vor, double* dvordt)
{
    if (!h) return fail(LBMIB_E_ARG, "null handle");
    if (dvordt && !rate) return fail(LBMIB_E_ARG, "dvordt is given by the rate update only (rate != 0)");
    HIP_TRY(hipSetDevice(h->dev));
    HIP_TRY(hipDeviceSynchronize());
    const long long N = h->D.N;
    const int slot = h->steps_done & 1;
    if (h->last_flag_step[slot] != h->steps_done)
        return fail(LBMIB_E_STATE, "velocity field of step %d was not recorded", h->steps_done);
    if (int rc = diag_alloc(h, &h->vort, sizeof(double) * N)) return rc;
    if (rate && !h->vort_hist) {
        if (int rc = diag_alloc(h, &h->vort_hist, sizeof(double) * 3 * N)) return rc;
        HIP_TRY(hipMemset(h->vort_hist, 0, sizeof(double) * 3 * N));
    }
    double* hv = rate ? h->vort_hist : nullptr;
    // with profiling on, the kernel is timed by events on its own stream
    lbmib::Engine* E = h->own;
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (E && E->prof_every > 0) {
        HIP_TRY(hipEventCreate(&ev[0]));
        HIP_TRY(hipEventCreate(&ev[1]));
        HIP_TRY(hipEventRecord(ev[0], nullptr));
    }
    hipLaunchKernelGGL(k_vorticity, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, nullptr,
                       (const double*)h->D.macro[slot], h->D.nx, h->D.ny, N, h->D.U, h->D.scale, h->vort, hv,
                       hv ? hv + N : nullptr, hv ? hv + 2 * N : nullptr);
    HIP_TRY(hipGetLastError());
    if (ev[0]) {
        float ms = 0.0f;
        HIP_TRY(hipEventRecord(ev[1], nullptr));
        HIP_TRY(hipEventSynchronize(ev[1]));
        HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        E->tms[9] += ms;
        E->tms[10] += 1.0;
    }
    if (vor) HIP_TRY(hipMemcpy(vor, h->vort, sizeof(double) * N, hipMemcpyDeviceToHost));
    if (dvordt) HIP_TRY(hipMemcpy(dvordt, h->vort_hist + 2 * N, sizeof(double) * N, hipMemcpyDeviceToHost));
    if (!vor && !dvordt) HIP_TRY(hipDeviceSynchronize());
    return 0;
}

int lbmib_surface_pressure(lbmib_handle* h, double* p, int32_t n)
{
    if (!h || !p) return fail(LBMIB_E_ARG, "null argument");
    if (h->steps_done < 1 || h->geo_fresh) return fail(LBMIB_E_STATE, "no step taken yet");
    if (h->p.smooth != 1)
        return fail(LBMIB_E_ARG, "cal_sur_cp interpolates with the smoothed kernel (smooth = 1), the case uses smooth = 0");
    const int ns = h->D.nseg;
    if (n < ns) return fail(LBMIB_E_ARG, "room for %d markers needed (got %d)", ns, n);
    HIP_TRY(hipSetDevice(h->dev));
    HIP_TRY(hipDeviceSynchronize());
    if (int rc = diag_alloc(h, &h->surf, sizeof(double) * ns)) return rc;
    const Slot S = slot_of(h->D, (h->steps_done - 1) % h->D.nslot);
    hipLaunchKernelGGL(k_surface_pressure, dim3((unsigned)((ns + 63) / 64)), dim3(64), 0, nullptr,
                       (const int*)S.dcnt, (const int*)S.didx, (const double*)S.dval, (const double*)S.rhoadj, ns,
                       h->surf);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(p, h->surf, sizeof(double) * ns, hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"

// Staging buffer of lbmib_surface_pressure_noalloc: one value per marker
// segment, LBMIB_SCRATCH_SEGMENTS of them (16 KiB), in the code object's
// global memory, so each device holds its own copy from the library's first
// use on it until the process ends.  Host code reaches a device's copy only
// through the address hipGetSymbolAddress returns on that device (the
// array's name is a host-side stand-in); lbmib_create records it.
static __device__ __align__(256) double scratch[LBMIB_SCRATCH_SEGMENTS];
static std::mutex g_scratch_mu;
static constexpr int SCRATCH_DEVS = 64;
static double* g_scratch_ptr[SCRATCH_DEVS] = {};

// The device copy of the staging buffer on `dev`, looked up once per device
// with that device current.
static int scratch_on(int dev, double** out)
{
    *out = nullptr;
    if (dev < 0 || dev >= SCRATCH_DEVS) return fail(LBMIB_E_ARG, "device %d beyond the scratch table", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        HIP_TRY(hipSetDevice(dev));
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(LBMIB_E_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

extern "C" {

int lbmib_surface_pressure_noalloc(lbmib_handle* h, double* p, int32_t n)
{
    if (!h || !p) return fail(LBMIB_E_ARG, "null argument");
    const int ns = h->D.nseg;
    if (ns > LBMIB_SCRATCH_SEGMENTS)
        return fail(LBMIB_E_ARG, "the case has %d marker segments; the allocation-free call stages at most %d", ns,
                    LBMIB_SCRATCH_SEGMENTS);
    if (h->steps_done < 1 || h->geo_fresh) return fail(LBMIB_E_STATE, "no step taken yet");
    if (h->p.smooth != 1)
        return fail(LBMIB_E_ARG, "cal_sur_cp interpolates with the smoothed kernel (smooth = 1), the case uses smooth = 0");
    if (n < ns) return fail(LBMIB_E_ARG, "room for %d markers needed (got %d)", ns, n);
    double* buf = nullptr;
    if (int rc = scratch_on(h->dev, &buf)) return rc;
    HIP_TRY(hipSetDevice(h->dev));
    HIP_TRY(hipDeviceSynchronize());
    // every one of the ns values is written by the kernel: nothing to clear
    const Slot S = slot_of(h->D, (h->steps_done - 1) % h->D.nslot);
    hipLaunchKernelGGL(k_surface_pressure, dim3((unsigned)((ns + 63) / 64)), dim3(64), 0, nullptr,
                       (const int*)S.dcnt, (const int*)S.didx, (const double*)S.dval, (const double*)S.rhoadj, ns,
                       buf);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(p, buf, sizeof(double) * ns, hipMemcpyDeviceToHost));
    return 0;
}

int lbmib_ib_nodes(lbmib_handle* h, int32_t* n, int32_t* pi, int32_t* pj, int32_t* wing, double* glx, double* gly)
{
    if (!h || !n) return fail(LBMIB_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(h->dev));
    HIP_TRY(hipDeviceSynchronize());
    *n = 0;
    if (h->steps_done == 0) return 0;
    const Slot S = slot_of(h->D, (h->steps_done - 1) % h->D.nslot);
    int na = 0;
    HIP_TRY(hipMemcpy(&na, S.nadj, sizeof(int), hipMemcpyDeviceToHost));
    *n = na;
    if (na == 0) return 0;
    if (pi || pj || wing) {
        std::vector<int> adj(3 * (size_t)na);
        HIP_TRY(hipMemcpy(adj.data(), S.adj, sizeof(int) * adj.size(), hipMemcpyDeviceToHost));
        for (int a = 0; a < na; ++a) {
            if (pi) pi[a] = adj[3 * a];
            if (pj) pj[a] = adj[3 * a + 1];
            if (wing) wing[a] = adj[3 * a + 2];
        }
    }
    if (glx) HIP_TRY(hipMemcpy(glx, S.glx, sizeof(double) * na, hipMemcpyDeviceToHost));
    if (gly) HIP_TRY(hipMemcpy(gly, S.gly, sizeof(double) * na, hipMemcpyDeviceToHost));
    return 0;
}

int lbmib_lattice_window(const lbmib_handle* h, int32_t* i0, int32_t* j0, int32_t* w, int32_t* hgt)
{
    if (!h || !i0 || !j0 || !w || !hgt) return fail(LBMIB_E_ARG, "null argument");
    if (!h->have_motion) return fail(LBMIB_E_STATE, "no motion table");
    *i0 = h->D.gi0;
    *j0 = h->D.gj0;
    *w = h->D.gw;
    *hgt = h->D.gh;
    return 0;
}

int lbmib_markers(const lbmib_handle* h, double* xb1, double* yb1)
{
    if (!h || !xb1 || !yb1) return fail(LBMIB_E_ARG, "null argument");
    if (!h->have_motion) return fail(LBMIB_E_STATE, "lbmib_set_motion_table must come first");
    // move_boundary's rigid motion of the shifted polygon (airfoil.f90:256-263)
    // with the kinematics row of the last step taken (of step 1 before any)
    const int n = std::max(1, h->steps_done);
    const double* row = &h->step_rows[(size_t)(n - 1) * h->ncols];
    const int nw = h->p.n_wings, ppw = h->p.pts_per_wing;
    for (int w = 0; w < nw; ++w) {
        const double* c = row + 1 + 9 * w;
        const double sn = std::sin(c[0]), ca = std::cos(c[0]), sa = c[3], sb = c[6];
        // the body offsets of the step (the dynamics' state before its update)
        const double es = h->body_terms[4 * w], mn = h->body_terms[4 * w + 1];
        for (int i = 0; i < ppw; ++i) {
            const double x = h->xb0[w * ppw + i], y = h->yb0[w * ppw + i];
            xb1[w * ppw + i] = ca * x + sn * y + sa + h->p.xc + es - mn;
            yb1[w * ppw + i] = -sn * x + ca * y + sb + h->p.yc;
        }
    }
    return 0;
}

int lbmib_device_bytes(const lbmib_handle* h, int64_t* bytes)
{
    if (!h || !bytes) return fail(LBMIB_E_ARG, "null argument");
    *bytes = h->device_bytes;
    return 0;
}

int lbmib_mem_info(int32_t device, int64_t* free_bytes, int64_t* total_bytes)
{
    if (!free_bytes || !total_bytes) return fail(LBMIB_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(device));
    size_t fr = 0, to = 0;
    HIP_TRY(hipMemGetInfo(&fr, &to));
    *free_bytes = (int64_t)fr;
    *total_bytes = (int64_t)to;
    return 0;
}

int lbmib_set_profiling(lbmib_handle* h, int32_t every)
{
    if (!h) return fail(LBMIB_E_ARG, "null handle");
    if (every < 0) return fail(LBMIB_E_ARG, "profiling interval must be >= 0");
    h->own->prof_every = every;
    for (double& t : h->own->tms) t = 0.0;
    return 0;
}

int lbmib_get_timings(lbmib_handle* h, double* t, int32_t n)
{
    if (!h || !t) return fail(LBMIB_E_ARG, "null argument");
    for (int i = 0; i < n && i < LBMIB_NTIMINGS; ++i) t[i] = h->own->tms[i];
    for (int i = LBMIB_NTIMINGS; i < n; ++i) t[i] = 0.0;
    return 0;
}

}  // extern "C"
