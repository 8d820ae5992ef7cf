// lbmib_engine.hip -- the stepping engine of liblbmib and the batched form.
//
// An Engine advances a set of cases (one for a single handle, many for a
// batch) on one device.  Every kernel of a step covers all cases in one
// launch: the lattice update walks a tile table spanning every case, the
// IB kernels take one grid row (or workgroup) per case.  Each case carries
// its own step clock, so cases of one window may stand at different steps
// (a case admitted later, or one that finished early and idles).
//
// Per step (window offset off), as for one case:
//   s3  right-hand sides of the IB systems (after step off-1's lattice and
//       IB-node collides, and the geometry of the batch)
//   s1  triangular solves, then spread + forced collide of the IB nodes
//   s0  lattice update of every other node (after step off-1's IB-node collides)
// and at every geometry-batch boundary, on s2: forces of the batch just
// finished, then the geometry of the batch after the one starting (and, for
// cases taking their first step, of batch 0 first).
//
// In graph mode windows of 2*batch steps that start at a batch boundary are
// captured once as a HIP graph (its last node advances every case's clock)
// and replayed.
#include "lbmib_host.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <map>
#include <mutex>

namespace lbmib {

__global__ void k_gather_err(const DevCase* __restrict__ C, int nc, int* __restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc) out[c] = C[c].err[0];
}

namespace {

// Stream sets are shared: one set of NSTREAM streams per (device, reserved
// CU count), made on first use and kept until the process ends.  A stream
// made with a CU mask takes a hardware queue of its own, so engines that
// made and destroyed their own sets (a process stepping many single cases,
// the autotuner's trials) would create and release dozens of such queues;
// with shared sets a process holds at most one set per CU split it uses.
// Engines of one process run one at a time on a set: every engine call ends
// with its streams synchronised, so no wait of one engine is pending when
// another engine's work follows on the same streams.
struct StreamSet {
    int dev = -1, cus = -1;
    hipStream_t s[Engine::NSTREAM] = {};
};
std::mutex g_sets_mu;
std::vector<StreamSet*> g_sets;

void destroy_streams(Engine* E)
{
    for (hipStream_t& st : E->s) {
        if (st) (void)hipStreamSynchronize(st);
        st = nullptr;
    }
    E->streams_cus = -1;
}

// The triangular solves run as one workgroup per case whose critical
// wavefront must not queue behind the lattice kernel's memory traffic or
// wait for a CU that the lattice kernel keeps full, so stream s1 gets
// ib_cus CUs that no other stream uses.  The mask bits are spread over the
// accelerator's 8 XCDs and a queue whose mask leaves an XCD without CUs is
// not usable, so only whole multiples of 8 are taken; 0 turns the split off.
int make_streams(Engine* E, int ib_cus)
{
    destroy_streams(E);
    ib_cus = std::max(0, ib_cus) / 8 * 8;
    const int ncu = E->ncu;
    const bool split = ib_cus > 0 && ib_cus < ncu / 2 && ncu % 8 == 0;
    const int key = split ? ib_cus : 0;
    std::lock_guard<std::mutex> lock(g_sets_mu);
    StreamSet* set = nullptr;
    for (StreamSet* q : g_sets)
        if (q->dev == E->dev && q->cus == key) set = q;
    if (!set) {
        set = new StreamSet();
        set->dev = E->dev;
        set->cus = key;
        hipError_t e = hipSuccess;
        if (split) {
            const int nwords = (ncu + 31) / 32;
            std::vector<uint32_t> lat(nwords, 0u), ib(nwords, 0u);
            for (int c = 0; c < ncu; ++c) {
                if (c >= ncu - ib_cus) ib[c / 32] |= 1u << (c % 32);
                else lat[c / 32] |= 1u << (c % 32);
            }
            for (int i = 0; i < Engine::NSTREAM && e == hipSuccess; ++i)
                e = hipExtStreamCreateWithCUMask(&set->s[i], nwords, i == 1 ? ib.data() : lat.data());
        } else {
            for (int i = 0; i < Engine::NSTREAM && e == hipSuccess; ++i)
                e = hipStreamCreateWithFlags(&set->s[i], hipStreamNonBlocking);
        }
        if (e != hipSuccess) {
            for (hipStream_t st : set->s)
                if (st) (void)hipStreamDestroy(st);
            delete set;
            return fail(LBMIB_E_HIP, "stream creation failed: %s", hipGetErrorString(e));
        }
        g_sets.push_back(set);
    }
    for (int i = 0; i < Engine::NSTREAM; ++i) E->s[i] = set->s[i];
    E->streams_cus = std::max(0, E->shape.ib_cus) / 8 * 8;
    return 0;
}

template <typename T>
int dev_reserve(T** p, size_t& cap, size_t need)
{
    if (need <= cap && *p) return 0;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    size_t n = std::max<size_t>(need, 1);
    HIP_TRY(hipMalloc((void**)p, sizeof(T) * n));
    cap = n;
    return 0;
}

// Tile table of the lattice kernel.  Case-major order takes each case's
// tiles in turn.  XCD order deals the tiles so that the workgroups the
// hardware places on XCD x (launch index v with v % 8 == x) take one
// contiguous band of rows of every case: the rows a pull reads above and
// below a tile then sit in the same XCD's L2.
void build_tiles(const Engine* E, std::vector<int2>& t)
{
    const int npt = E->shape.npt;
    t.clear();
    if (E->shape.order == 0) {
        for (size_t c = 0; c < E->members.size(); ++c) {
            const DevCase& D = E->members[c]->D;
            const int nt = lattice_tiles(D.nx, D.ny, npt);
            for (int b = 0; b < nt; ++b) t.push_back(make_int2((int)c, b));
        }
        return;
    }
    std::vector<int2> L[8];
    int rot = 0;
    for (size_t c = 0; c < E->members.size(); ++c) {
        const DevCase& D = E->members[c]->D;
        const int nt = lattice_tiles(D.nx, D.ny, npt), rb = lattice_ring_blocks(D.nx, D.ny);
        for (int b = 0; b < rb; ++b) L[(rot + b) % 8].push_back(make_int2((int)c, b));
        rot = (rot + rb) % 8;
        const long long ni = nt - rb;
        for (int x = 0; x < 8; ++x) {
            const int lo = rb + (int)(ni * x / 8), hi = rb + (int)(ni * (x + 1) / 8);
            for (int b = lo; b < hi; ++b) L[(x + rot) % 8].push_back(make_int2((int)c, b));
        }
        rot = (rot + 1) % 8;
    }
    size_t q = 0;
    bool any = true;
    while (any) {
        any = false;
        for (int x = 0; x < 8; ++x)
            if (q < L[x].size()) { t.push_back(L[x][q]); any = true; }
        ++q;
    }
}

struct Dims {
    int nc = 0, maxns = 0, maxcap = 0, maxW = 0, maxnpad = 0, prec = 0;
    int chain = 0, sep = 0, maxnp2 = 0;  // members solved as one chain / in separator form
    int plain = 0;                       // chain members with per-lane (not compact) coefficients
};

// Staging units (steps per LDS ring unit) of the two solve kernels: the
// largest whose rings fit beside the values, capped by shape.unit (0: no
// cap).  The unit never changes results.
int chain_unit(const Engine* E, const Dims& d)
{
    const int u = solve_unit(d.maxnpad);
    return E->shape.unit && E->shape.unit < u ? 32 : u;
}
int sep_unit(const Engine* E, const Dims& d)
{
    const int u = solve2_unit(d.maxnp2);
    return E->shape.unit && E->shape.unit < u ? 16 : u;
}

const void* solve2_kernel(int U)
{
    return U == 32 ? (const void*)k_crit_solve2<32> : (const void*)k_crit_solve2<16>;
}

const void* solve_kernel(int U, bool cmp)
{
    if (U == 64) return cmp ? (const void*)k_crit_solve<64, true> : (const void*)k_crit_solve<64, false>;
    return cmp ? (const void*)k_crit_solve<32, true> : (const void*)k_crit_solve<32, false>;
}

Dims dims_of(const Engine* E)
{
    Dims d;
    d.nc = (int)E->members.size();
    for (lbmib_handle* h : E->members) {
        d.maxns = std::max(d.maxns, h->D.nseg);
        d.maxcap = std::max(d.maxcap, h->D.cap);
        d.maxW = std::max(d.maxW, h->D.bw + 1);
        d.maxnpad = std::max(d.maxnpad, h->D.npad);
        if (h->D.dissect) {
            d.sep++;
            d.maxnp2 = std::max(d.maxnp2, h->D.np2);
        } else {
            d.chain++;
            if (!h->D.compact) d.plain++;
        }
    }
    d.prec = E->members.empty() ? 0 : E->members[0]->D.prec;
    return d;
}

// Whether an engine takes two-step passes: prescribed motion (the dynamics
// step one at a time), members of one precision (every member a pass) whose
// lattice windows lie at least 4 nodes away from every edge (the window's
// single steps then never touch the outer ring), and lattices narrower than
// 2^16 columns (the far tiles pack two columns in one int).
bool fuse_ok(const Engine* E)
{
    if (!E->shape.fuse || E->shape.graph || E->per_step || E->members.empty()) return false;
    for (const lbmib_handle* h : E->members) {
        const DevCase& D = h->D;
        if (D.nx >= (1 << 16)) return false;
        if (!(D.gi0 >= 4 && D.gj0 >= 4 && D.gi0 + D.gw <= D.nx - 4 && D.gj0 + D.gh <= D.ny - 4)) return false;
    }
    return true;
}

// Workgroup tables of the two-step passes over every member.  Far tiles:
// strips of at most FAR_OUT columns (lane L of a strip at column a-1+L) and
// segments of about 5 rows.  Many short segments beat one workgroup per CU
// slot with tall ones (4096x2048, FP64: 0.279 ms per pass at 410 segments,
// 0.291 ms at 256, 0.49 ms at 28 with an older kernel): the pass is bound by
// the rows the workgroups have in flight, not by the halo rows that short
// segments read again, and many workgroups balance the load at the end.
// LBMIB_FAR_SEGS overrides the segment count of every member.  The window,
// ring and window + ring (+ layer) tables take window_blocks, ring_blocks1
// and window_blocks + ring_blocks1 + layer_blocks workgroups per member.
int build_pass_tables(Engine* E)
{
    std::vector<int> key;
    const char* es = getenv("LBMIB_FAR_SEGS");
    key.push_back(es ? atoi(es) : -1);
    for (const lbmib_handle* h : E->members) {
        const DevCase& D = h->D;
        for (int v : {D.nx, D.ny, D.gi0, D.gj0, D.gw, D.gh}) key.push_back(v);
    }
    if (key == E->pass_key && E->dfar) return 0;
    std::vector<int4> far;
    std::vector<int2> win, ring, wr;
    for (size_t c = 0; c < E->members.size(); ++c) {
        const DevCase& D = E->members[c]->D;
        const int nx = D.nx, ny = D.ny;
        // FP64: strips of at most FAR_OUT columns; FP32: of at most FAR_OUT2
        // columns starting at even columns (two columns per lane)
        const bool f64 = D.prec == LBMIB_PREC_FP64;
        const int units = f64 ? nx : (nx + 1) / 2, per = f64 ? FAR_OUT : FAR_OUT2 / 2, scale = f64 ? 1 : 2;
        const int strips = (units + per - 1) / per;
        int segs = std::max(1, ny / 5);
        if (es) segs = std::max(1, std::min(ny / 4, atoi(es)));
        for (int sg = 0; sg < segs; ++sg) {
            const int j0 = (int)((long long)sg * ny / segs), j1 = (int)((long long)(sg + 1) * ny / segs) - 1;
            for (int st = 0; st < strips; ++st) {
                const int a = scale * (int)((long long)st * units / strips);
                const int b = std::min(nx, scale * (int)((long long)(st + 1) * units / strips));
                far.push_back(make_int4((int)c, a | (b << 16), j0, j1));
            }
        }
        const int wb = window_blocks(D.gw, D.gh), rb = ring_blocks1(nx, ny), lb = layer_blocks(nx, ny);
        for (int b = 0; b < wb; ++b) win.push_back(make_int2((int)c, b));
        for (int b = 0; b < rb; ++b) ring.push_back(make_int2((int)c, b));
        for (int b = 0; b < wb + rb + lb; ++b) wr.push_back(make_int2((int)c, b));
    }
    std::vector<int2> all;
    all.insert(all.end(), win.begin(), win.end());
    all.insert(all.end(), ring.begin(), ring.end());
    all.insert(all.end(), wr.begin(), wr.end());
    int rc = dev_reserve(&E->dfar, E->far_cap, far.size());
    if (!rc) rc = dev_reserve(&E->dwin, E->win_cap, all.size());
    if (rc) return rc;
    HIP_TRY(hipMemcpy(E->dfar, far.data(), sizeof(int4) * far.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(E->dwin, all.data(), sizeof(int2) * all.size(), hipMemcpyHostToDevice));
    E->nfar = (int)far.size();
    E->nwin = (int)win.size();
    E->nring = (int)ring.size();
    E->nwr = (int)wr.size();
    E->pass_key.swap(key);
    return 0;
}

// The case table goes to the device only when it changed since the last
// window (byte copies, so that padding compares equal too).
int upload_cases(Engine* E)
{
    const size_t nc = E->members.size();
    std::vector<unsigned char> tab(sizeof(DevCase) * nc);
    for (size_t c = 0; c < nc; ++c) std::memcpy(&tab[c * sizeof(DevCase)], &E->members[c]->D, sizeof(DevCase));
    if (tab != E->tab_bytes) {
        HIP_TRY(hipMemcpy(E->dcases, tab.data(), tab.size(), hipMemcpyHostToDevice));
        E->tab_bytes.swap(tab);
    }
    return 0;
}

// Before a window of two-step passes every member that steps in it holds its
// state in the plane of the first such member (the planes named by a launch
// are then the same for all members): a member whose state lies in the other
// plane swaps its two state planes and its parity.  Returns that plane in
// *plane (0 if no member steps) and whether any member swapped.
bool normalize_planes(Engine* E, const std::vector<int>& len_end, int* plane)
{
    bool any = false;
    int target = -1;
    for (size_t c = 0; c < E->members.size(); ++c) {
        lbmib_handle* h = E->members[c];
        if (len_end[c] <= h->steps_done) continue;
        if (target < 0) target = state_plane(h);
        if (state_plane(h) == target) continue;
        DevCase& D = h->D;
        std::swap(D.f[0], D.f[1]);
        std::swap(D.g[0], D.g[1]);
        D.pofs ^= 1;
        any = true;
    }
    *plane = target < 0 ? 0 : target;
    return any;
}

// hipFuncSetAttribute(MaxDynamicSharedMemorySize) only when kernel f on
// device dev needs more than any earlier call set (the limit is a maximum,
// so the largest serves every engine).
int lds_limit(int dev, const void* f, size_t bytes)
{
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, size_t> set;
    std::lock_guard<std::mutex> lock(mu);
    size_t& have = set[{dev, f}];
    if (bytes <= have) return 0;
    HIP_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    have = bytes;
    return 0;
}

int engine_prepare(Engine* E)
{
    HIP_TRY(hipSetDevice(E->dev));
    const int want = std::max(0, E->shape.ib_cus) / 8 * 8;
    if (E->streams_cus != want || !E->s[0]) {
        int rc = make_streams(E, want);
        if (rc) return rc;
        E->graph_valid = false;
    }
    const size_t nc = E->members.size();
    if (nc > E->case_cap || !E->dcases) {
        // tables sized for the largest membership seen; a move invalidates the graph
        size_t cc = 0, kc = 0, ec = 0;
        if (E->dcases) (void)hipFree(E->dcases);
        if (E->dclock) (void)hipFree(E->dclock);
        if (E->derr) (void)hipFree(E->derr);
        E->dcases = nullptr; E->dclock = nullptr; E->derr = nullptr;
        E->case_cap = 0;
        int rc = dev_reserve(&E->dcases, cc, nc);
        if (!rc) rc = dev_reserve(&E->dclock, kc, nc);
        if (!rc) rc = dev_reserve(&E->derr, ec, nc);
        if (rc) return rc;
        E->case_cap = nc;
        E->graph_valid = false;
        E->tab_bytes.clear();
    }
    int rc = 0;
    if (fuse_ok(E)) {
        // every member's middle plane, and the workgroup tables
        for (lbmib_handle* h : E->members)
            if ((rc = ensure_mid_plane(h))) return rc;
        if ((rc = build_pass_tables(E))) return rc;
    }
    if ((rc = upload_cases(E))) return rc;
    if (nc > E->pin_cap) {
        if (E->hclk_pin) (void)hipHostFree(E->hclk_pin);
        if (E->herr_pin) (void)hipHostFree(E->herr_pin);
        E->hclk_pin = nullptr;
        E->herr_pin = nullptr;
        E->pin_cap = 0;
        HIP_TRY(hipHostMalloc((void**)&E->hclk_pin, sizeof(CaseClock) * nc, hipHostMallocDefault));
        HIP_TRY(hipHostMalloc((void**)&E->herr_pin, sizeof(int) * nc, hipHostMallocDefault));
        E->pin_cap = nc;
    }
    if (!E->tiles_valid) {
        std::vector<int2> t;
        build_tiles(E, t);
        rc = dev_reserve(&E->dtiles, E->tile_cap, t.size());
        if (rc) return rc;
        HIP_TRY(hipMemcpy(E->dtiles, t.data(), sizeof(int2) * t.size(), hipMemcpyHostToDevice));
        E->ntiles = (int)t.size();
        E->tiles_valid = true;
        E->graph_valid = false;
    }
    // the kernels' dynamic LDS limits: each is a maximum, kept per device and
    // kernel at the largest any engine of the process has needed
    const Dims d = dims_of(E);
    if (int rc2 = lds_limit(E->dev, (const void*)k_far2<double>, far2_smem_bytes(true))) return rc2;
    if (int rc2 = lds_limit(E->dev, (const void*)k_far2<float>, far2_smem_bytes(false))) return rc2;
    if (int rc2 = lds_limit(E->dev, (const void*)k_geo_factor, factor_smem_bytes())) return rc2;
    for (bool cmp : {false, true})
        if (int rc2 = lds_limit(E->dev, solve_kernel(chain_unit(E, d), cmp), solve_smem_bytes(d.maxnpad))) return rc2;
    if (d.sep) {
        if (int rc2 = lds_limit(E->dev, (const void*)k_geo_factor2, factor_smem_bytes())) return rc2;
        if (int rc2 = lds_limit(E->dev, solve2_kernel(sep_unit(E, d)), solve2_smem_bytes(d.maxnp2))) return rc2;
    }
    return 0;
}

// Forces of window offsets a..b (one workgroup per step and case) on s2
// (or, for the window's last steps, s1), after the IB-node collide of engine
// step gs_b (offset b).
int enqueue_forces(Engine* E, const Dims& d, int a, int b, long long gs_b, bool tail = false)
{
    if (b < a) return 0;
    // with per-step geometry the forces gate the next step, and a window's
    // last forces end it: they go to the solves' stream, whose reserved CUs
    // are idle once the last IB nodes collide (on the shared CUs they would
    // wait for the far pass to free room, and on s2 behind geometry built
    // ahead)
    hipStream_t st = E->per_step || tail ? E->s[1] : E->s[2];
    HIP_TRY(hipStreamWaitEvent(st, E->ev_ib[gs_b % Engine::RING], 0));
    hipLaunchKernelGGL(k_forces, dim3(b - a + 1, d.nc), dim3(256), 0, st, (const DevCase*)E->dcases,
                       (const CaseClock*)E->dclock, a);
    return 0;
}

int launch_geometry(Engine* E, const Dims& d, int off, int mode)
{
    hipStream_t st = E->s[2];
    const DevCase* C = E->dcases;
    const CaseClock* K = E->dclock;
    const int B = E->B, nc = d.nc, ns = d.maxns;
    hipLaunchKernelGGL(k_geo_clear, dim3(16, B, nc), dim3(256), 0, st, C, K, off, mode);
    hipLaunchKernelGGL(k_geo_markers, dim3((ns + 255) / 256, B, nc), dim3(256), 0, st, C, K, off, mode);
    hipLaunchKernelGGL(k_geo_fill, dim3(B, nc), dim3(1024), 0, st, C, K, off, mode);
    hipLaunchKernelGGL(k_geo_phi, dim3((ns * 12 + 255) / 256, B, nc), dim3(256), 0, st, C, K, off, mode);
    hipLaunchKernelGGL(k_geo_pairs, dim3((ns * BOXMAX + 255) / 256, B, nc), dim3(256), 0, st, C, K, off, mode);
    hipLaunchKernelGGL(k_geo_dsort, dim3((ns * BOXMAX + 255) / 256, B, nc), dim3(256), 0, st, C, K, off, mode);
    hipLaunchKernelGGL(k_geo_nsort, dim3(64, B, nc), dim3(256), 0, st, C, K, off, mode);
    const long long nbw = (long long)ns * d.maxW;
    hipLaunchKernelGGL(k_geo_assemble, dim3((unsigned)((nbw + 255) / 256), B, nc), dim3(256), 0, st, C, K, off, mode);
    return 0;
}

int launch_factor(Engine* E, const Dims& d, int off, int mode, hipEvent_t* gev)
{
    hipStream_t st = E->s[2];
    const DevCase* C = E->dcases;
    const CaseClock* K = E->dclock;
    if (gev) HIP_TRY(hipEventRecord(gev[1], st));
    if (d.chain) hipLaunchKernelGGL(k_geo_factor, dim3(E->B, d.nc), dim3(256), factor_smem_bytes(), st, C, K, off, mode);
    if (d.sep) {
        hipLaunchKernelGGL(k_geo_factor2, dim3(E->B, d.nc, 2), dim3(256), factor_smem_bytes(), st, C, K, off, mode);
        hipLaunchKernelGGL(k_geo_schur, dim3(E->B, d.nc), dim3(256), 0, st, C, K, off, mode);
    }
    if (gev) HIP_TRY(hipEventRecord(gev[2], st));
    return 0;
}

// Host check of the planes a launch names: each index one of the case's
// allocated planes (0, 1, or the middle plane 2).
bool planes_ok(const DevCase& D, std::initializer_list<int> planes)
{
    for (int p : planes) {
        if (p < 0 || p > 2) return false;
        if (D.prec == LBMIB_PREC_FP64 ? D.f[p] == nullptr : D.g[p] == nullptr) return false;
    }
    return true;
}

// Right-hand sides of the IB systems of the step at window offset off on
// stream st: one case takes the single pass (the shorter chain of dependent
// loads), several cases u* once per IB node and then the per-marker sums.
int launch_rhs(Engine* E, const Dims& d, hipStream_t st, int off, int psel)
{
    const DevCase* C = E->dcases;
    const CaseClock* K = E->dclock;
    const bool f64 = d.prec == LBMIB_PREC_FP64;
    if (d.nc == 1) {
        hipLaunchKernelGGL(f64 ? k_crit_rhs1<double> : k_crit_rhs1<float>, dim3((d.maxns + 6) / 7, d.nc), dim3(256),
                           0, st, C, K, off, psel);
    } else {
        hipLaunchKernelGGL(f64 ? k_crit_ustar<double> : k_crit_ustar<float>, dim3((d.maxcap + 255) / 256, d.nc),
                           dim3(256), 0, st, C, K, off, psel);
        hipLaunchKernelGGL(k_crit_rhs, dim3((d.maxns + 255) / 256, d.nc), dim3(256), 0, st, C, K, off);
    }
    return 0;
}

// Triangular solves of the step at window offset off on stream st: chain
// members of each coefficient layout and separator-form members (a kernel
// returns at once for the members of the other kinds).
int launch_solve(Engine* E, const Dims& d, hipStream_t st, int off)
{
    const DevCase* C = E->dcases;
    const CaseClock* K = E->dclock;
    const size_t smc = solve_smem_bytes(d.maxnpad);
    const bool u64 = chain_unit(E, d) == 64;
    if (d.chain > d.plain) {
        if (u64) hipLaunchKernelGGL((k_crit_solve<64, true>), dim3(d.nc), dim3(solve_threads()), smc, st, C, K, off);
        else hipLaunchKernelGGL((k_crit_solve<32, true>), dim3(d.nc), dim3(solve_threads()), smc, st, C, K, off);
    }
    if (d.plain) {
        if (u64) hipLaunchKernelGGL((k_crit_solve<64, false>), dim3(d.nc), dim3(solve_threads()), smc, st, C, K, off);
        else hipLaunchKernelGGL((k_crit_solve<32, false>), dim3(d.nc), dim3(solve_threads()), smc, st, C, K, off);
    }
    if (d.sep) {
        const size_t sm = solve2_smem_bytes(d.maxnp2);
        if (sep_unit(E, d) == 32)
            hipLaunchKernelGGL(k_crit_solve2<32>, dim3(d.nc), dim3(solve2_threads()), sm, st, C, K, off);
        else
            hipLaunchKernelGGL(k_crit_solve2<16>, dim3(d.nc), dim3(solve2_threads()), sm, st, C, K, off);
    }
    return 0;
}

// Step at window offset off.  ev (when this step is sampled): [0]/[1]
// around the lattice update, [2] before the right-hand sides, [3] after the
// solves, [4] after the IB-node collides.
// Flow control: the host stays at most LAG steps ahead of the lattice
// stream, so an event of the step ring (RING = 2 LAG slots) is never
// recorded again while a wait on its previous record may still be pending,
// and the work in flight stays bounded however long the window.
constexpr int LAG = Engine::RING / 2;

int launch_step(Engine* E, const Dims& d, int off, bool after_prev, hipEvent_t wfirst, hipEvent_t wnext,
                hipEvent_t wnodes, hipEvent_t* ev, int psel)
{
    const int R = Engine::RING;
    const long long gs = E->gstep;
    if (psel >= 0 && !planes_ok(E->members[0]->D, {psel & 3, (psel >> 2) & 3}))
        return fail(LBMIB_E_STATE, "step planes %d -> %d not allocated", psel & 3, (psel >> 2) & 3);
    if (off >= LAG) HIP_TRY(hipEventSynchronize(E->ev_step[(gs - LAG) % R]));
    const DevCase* C = E->dcases;
    const CaseClock* K = E->dclock;
    hipStream_t s0 = E->s[0], s1 = E->s[1], s3 = E->s[3];
    // right-hand sides on the shared CUs (s3), or (shape.rhs) in the solves'
    // stream, which saves the hand-over between the two streams
    const bool rr = E->shape.rhs != 0;
    hipStream_t sr = rr ? s1 : s3;
    if (after_prev) {
        HIP_TRY(hipStreamWaitEvent(sr, E->ev_step[(gs - 1) % R], 0));
        if (!rr || E->shape.band) HIP_TRY(hipStreamWaitEvent(sr, E->ev_ib[(gs - 1) % R], 0));
    }
    if (wfirst) HIP_TRY(hipStreamWaitEvent(sr, wfirst, 0));
    if (wnext) HIP_TRY(hipStreamWaitEvent(sr, wnext, 0));
    if (ev) HIP_TRY(hipEventRecord(ev[2], sr));
    const bool f64 = d.prec == LBMIB_PREC_FP64;
    if (int rc = launch_rhs(E, d, sr, off, psel)) return rc;
    if (!rr) {
        HIP_TRY(hipEventRecord(E->ev_rhs[gs % R], s3));
        HIP_TRY(hipStreamWaitEvent(s1, E->ev_rhs[gs % R], 0));
    }
    if (int rc = launch_solve(E, d, s1, off)) return rc;
    if (ev) HIP_TRY(hipEventRecord(ev[3], s1));
    // the forced collide of the IB nodes follows the solves on their reserved
    // CUs, or (shape.band) on the shared CUs of the right-hand-side stream
    hipStream_t sb = s1;
    if (E->shape.band) {
        HIP_TRY(hipEventRecord(E->ev_solve[gs % R], s1));
        HIP_TRY(hipStreamWaitEvent(s3, E->ev_solve[gs % R], 0));
        sb = s3;
    }
    const int band_x = std::max(1, (3 * d.maxns + 255) / 256);
    hipLaunchKernelGGL(f64 ? k_band<double> : k_band<float>, dim3(band_x, d.nc), dim3(256), 0, sb, C, K, off, psel);
    if (ev) HIP_TRY(hipEventRecord(ev[4], sb));
    HIP_TRY(hipEventRecord(E->ev_ib[gs % R], sb));
    if (after_prev) HIP_TRY(hipStreamWaitEvent(s0, E->ev_ib[(gs - 1) % R], 0));
    if (wfirst) HIP_TRY(hipStreamWaitEvent(s0, wnodes ? wnodes : wfirst, 0));
    if (wnext) HIP_TRY(hipStreamWaitEvent(s0, wnext, 0));
    if (ev) HIP_TRY(hipEventRecord(ev[0], s0));
    const int npt = E->shape.npt == 2 ? 2 : 1;
    int grid = E->ntiles;
    if (E->shape.persist > 0) grid = std::min(grid, std::max(8, E->shape.persist * E->ncu / 8 * 8));
    const int2* tiles = E->dtiles;
    const int nt = E->ntiles;
    if (f64) {
        if (npt == 2) hipLaunchKernelGGL((k_lattice<double, 2>), dim3(grid), dim3(256), 0, s0, C, K, tiles, nt, off, psel);
        else hipLaunchKernelGGL((k_lattice<double, 1>), dim3(grid), dim3(256), 0, s0, C, K, tiles, nt, off, psel);
    } else {
        if (npt == 2) hipLaunchKernelGGL((k_lattice<float, 2>), dim3(grid), dim3(256), 0, s0, C, K, tiles, nt, off, psel);
        else hipLaunchKernelGGL((k_lattice<float, 1>), dim3(grid), dim3(256), 0, s0, C, K, tiles, nt, off, psel);
    }
    if (ev) HIP_TRY(hipEventRecord(ev[1], s0));
    HIP_TRY(hipEventRecord(E->ev_step[gs % R], s0));
    E->gstep++;
    return 0;
}

// Steps n, n+1 at window offsets off, off+1 of an engine of one case as a
// two-step pass (k_far2 on s0) over the far field beside the window's single
// steps (k_window on s3) and the IB chain of both steps (s1).  Planes: the
// state of step n-1 in X, of step n+1 written to Y = X ^ 1, step n in the
// middle plane 2 where the window's second step and the IB chain of step
// n+1 read it.  Order (every wait an event of the step before or of this
// pass):
//   IB(n):   after the window/lattice and the IB collide of step n-1
//   W(n):    after the update and IB collide of step n-1 (plane X complete)
//   F:       after the window's and the ring's update of step n-1 (same
//            stream).  F reads X outside W and in W's outer two layers and
//            writes Y and the middle plane outside W; the IB kernels of
//            step n-1 read and write only nodes at least 2 nodes inside W
//            (the geometry flags an IB box closer than 3 nodes to W's
//            edge), so F does not wait for the IB collide of step n-1
//   IB(n+1): after W(n) (and IB(n) in its own stream)
//   W(n+1):  after F in F's stream (its layer around W), W(n) and the IB
//            collide of step n
// The outer ring's step n (k_ring, X -> middle) follows W(n) in W's stream,
// beside F: it reads only the ring's neighbourhood of X and writes only the
// ring's nodes of the middle plane, which F neither reads nor writes (F
// leaves out the ring and the second step of the layer next to it).  Its
// step n+1 and the layer's (middle -> Y) ride along with W(n+1)
// (k_window_ring), after F, which puts step n of the nodes inwards of the
// ring into the middle plane.
// ev as in launch_step for step n, plus [5]/[6] around W(n); [0]/[1] around F.
int launch_pair(Engine* E, const Dims& d, int off, bool after_prev, hipEvent_t wfirst, hipEvent_t wnext,
                hipEvent_t wnodes, hipEvent_t* ev, int X)
{
    const int R = Engine::RING;
    const long long gs = E->gstep;
    if (off >= LAG) HIP_TRY(hipEventSynchronize(E->ev_step[(gs - LAG) % R]));
    if (off + 1 >= LAG) HIP_TRY(hipEventSynchronize(E->ev_step[(gs + 1 - LAG) % R]));
    const DevCase* C = E->dcases;
    const CaseClock* K = E->dclock;
    hipStream_t s0 = E->s[0], s1 = E->s[1], s3 = E->s[3];
    const bool f64 = d.prec == LBMIB_PREC_FP64;
    const int Y = X ^ 1, M = 2;
    const int pa = make_psel(X, M), pb = make_psel(M, Y);
    for (const lbmib_handle* h : E->members)
        if (!planes_ok(h->D, {X, Y, M}))
            return fail(LBMIB_E_STATE, "two-step pass: planes %d, %d, %d not all allocated", X, Y, M);
    // the far field, both steps, enqueued first: it waits for nothing but
    // the window + ring update of step n-1 before it in its stream
    if (ev) HIP_TRY(hipEventRecord(ev[0], s0));
    // one workgroup per far tile, or the persistent grid of the launch shape
    int fgrid = E->nfar;
    if (E->shape.persist > 0) fgrid = std::min(fgrid, std::max(8, E->shape.persist * E->ncu / 8 * 8));
    const dim3 fg((unsigned)fgrid);
    const size_t fsm = far2_smem_bytes(f64);
    const int pf = X | (Y << 2) | (M << 4);  // k_far2's planes: src, dst, middle
    if (f64) hipLaunchKernelGGL(k_far2<double>, fg, dim3(FAR_LANES), fsm, s0, C, K, E->dfar, E->nfar, off, pf);
    else hipLaunchKernelGGL(k_far2<float>, fg, dim3(FAR_LANES), fsm, s0, C, K, E->dfar, E->nfar, off, pf);
    if (ev) HIP_TRY(hipEventRecord(ev[1], s0));
    // IB chain of step n on the solves' stream
    if (after_prev) HIP_TRY(hipStreamWaitEvent(s1, E->ev_step[(gs - 1) % R], 0));
    if (wfirst) HIP_TRY(hipStreamWaitEvent(s1, wfirst, 0));
    if (wnext) HIP_TRY(hipStreamWaitEvent(s1, wnext, 0));
    if (ev) HIP_TRY(hipEventRecord(ev[2], s1));
    if (int rc = launch_rhs(E, d, s1, off, pa)) return rc;
    if (int rc = launch_solve(E, d, s1, off)) return rc;
    if (ev) HIP_TRY(hipEventRecord(ev[3], s1));
    const int band_x = std::max(1, (3 * d.maxns + 255) / 256);
    hipLaunchKernelGGL(f64 ? k_band<double> : k_band<float>, dim3(band_x, d.nc), dim3(256), 0, s1, C, K, off, pa);
    if (ev) HIP_TRY(hipEventRecord(ev[4], s1));
    HIP_TRY(hipEventRecord(E->ev_ib[gs % R], s1));
    const int2* wtab = E->dwin;
    const int2* rtab = E->dwin + E->nwin;
    const int2* wrtab = rtab + E->nring;
    // window step n on the shared CUs
    const dim3 wg(E->nwin);
    if (after_prev) {
        HIP_TRY(hipStreamWaitEvent(s3, E->ev_step[(gs - 1) % R], 0));
        HIP_TRY(hipStreamWaitEvent(s3, E->ev_ib[(gs - 1) % R], 0));
    }
    if (wfirst) HIP_TRY(hipStreamWaitEvent(s3, wnodes ? wnodes : wfirst, 0));
    if (wnext) HIP_TRY(hipStreamWaitEvent(s3, wnext, 0));
    if (ev) HIP_TRY(hipEventRecord(ev[5], s3));
    hipLaunchKernelGGL(f64 ? k_window<double> : k_window<float>, wg, dim3(256), 0, s3, C, K, wtab, off, pa);
    if (ev) HIP_TRY(hipEventRecord(ev[6], s3));
    HIP_TRY(hipEventRecord(E->ev_win[gs % R], s3));  // the IB chain of step n+1 needs only W(n)
    // the ring's step n into the middle plane, from the ring's neighbourhood
    // of plane X (written by the window + ring update, or the lattice update,
    // of step n-1 and the far pass before it, all before ev_step of step n-1)
    hipLaunchKernelGGL(f64 ? k_ring<double> : k_ring<float>, dim3(E->nring), dim3(256), 0, s3, C, K, rtab, off, pa);
    // step n is complete once its IB collide is too: the window + ring
    // update after F then waits for one event (each wait is a packet that
    // F's stream processes, even when long satisfied)
    HIP_TRY(hipStreamWaitEvent(s3, E->ev_ib[gs % R], 0));
    HIP_TRY(hipEventRecord(E->ev_step[gs % R], s3));
    // IB chain of step n+1
    HIP_TRY(hipStreamWaitEvent(s1, E->ev_win[gs % R], 0));
    if (int rc = launch_rhs(E, d, s1, off + 1, pb)) return rc;
    if (int rc = launch_solve(E, d, s1, off + 1)) return rc;
    hipLaunchKernelGGL(f64 ? k_band<double> : k_band<float>, dim3(band_x, d.nc), dim3(256), 0, s1, C, K, off + 1, pb);
    HIP_TRY(hipEventRecord(E->ev_ib[(gs + 1) % R], s1));
    // window step n+1 right after the far pass in its stream (no hand-over on
    // the path from one pass to the next); W(n) and the IB collide of step n
    // have long finished by then.  The ring's step n+1 rides along.
    HIP_TRY(hipStreamWaitEvent(s0, E->ev_step[gs % R], 0));  // W(n), the ring's step n, the IB collide of step n
    hipLaunchKernelGGL(f64 ? k_window_ring<double> : k_window_ring<float>, dim3(E->nwr), dim3(256), 0, s0, C, K, wrtab,
                       off + 1, pb, make_psel(M, Y));
    HIP_TRY(hipEventRecord(E->ev_step[(gs + 1) % R], s0));
    E->gstep += 2;
    return 0;
}

// Enqueue `len` steps from window offset 0; phase = position of offset 0
// inside its geometry batch, fresh = some case takes its first step at
// offset 0.  With prof, samples every prof_every-th step.
//
// An engine that takes two-step passes (par given) names the planes of every
// step itself: every member that steps starts the window with its state in
// plane par (normalize_planes), the step at offset off reads plane
// (off + par) & 1, and each pass flips par (passes only at offsets below
// pair_limit, while every member that steps in the window still steps).  Passes start at any position; a
// pass whose second step opens a geometry batch takes that batch's waits at
// its start (the geometry of the batch it enters was built a batch earlier),
// and the boundary's own work (the forces of the batch just finished, the
// geometry of the batch after next into the finished batch's slots) follows
// the pass, waiting for both of its steps.
//
// A boundary fewer than GEO_DEFER steps before the window's end leaves the
// next batch's geometry (about 1.3 ms of latency-bound kernels) to the
// start of the next window (due: some member's next batch is not built),
// so that the window does not end waiting for geometry its steps do not
// use.  *next_off receives the offset of the last boundary that built the
// next batch (-1: none).
constexpr int GEO_DEFER = 12;
int launch_window(Engine* E, const Dims& d, int len, int phase, bool fresh, bool prof, int* par, int pair_limit,
                  bool due = false, int* next_off = nullptr)
{
    const int B = E->B, BR = Engine::BRING;
    int fupto = 0;
    if (next_off) *next_off = -1;
    hipEvent_t prev_next = nullptr;
    const int every = prof ? E->prof_every : 0;
    const int Eper = Engine::EPER;
    if (every > 0) E->samp_kind.assign((size_t)(len + every - 1) / every, 2);
    // Work of the geometry-batch boundary at window offset off, after every
    // step enqueued so far: the forces of offsets fupto..off-1, then on s2
    // (after the last enqueued step, whose batch's slots the next geometry
    // reuses) the geometry of the batch after the one starting, and for a
    // fresh case at offset 0 that of its first batch before.  Returns the
    // events the step at off waits for (wfirst/wnodes: the first batch;
    // wnext: the batch starting, built at the boundary before).
    auto boundary = [&](int off, hipEvent_t& wfirst, hipEvent_t& wnext, hipEvent_t& wnodes) -> int {
        if (off > fupto) {
            int rc = enqueue_forces(E, d, fupto, off - 1, E->gstep - 1);
            if (rc) return rc;
            fupto = off;
        }
        hipStream_t st = E->s[2];
        if (off > 0) {
            HIP_TRY(hipStreamWaitEvent(st, E->ev_step[(E->gstep - 1) % Engine::RING], 0));
            HIP_TRY(hipStreamWaitEvent(st, E->ev_ib[(E->gstep - 1) % Engine::RING], 0));
        }
        const long long bi = E->gbound++;
        hipEvent_t* gev = nullptr;
        if (every > 0) {
            while (E->gevpool.size() < (size_t)3 * (E->gev_used + 1)) {
                hipEvent_t e;
                HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));  // timing events
                E->gevpool.push_back(e);
            }
            gev = &E->gevpool[(size_t)3 * E->gev_used++];
            HIP_TRY(hipEventRecord(gev[0], st));
        }
        int rc;
        if (fresh && off == 0) {
            // the lattice update needs only the node sets; the IB
            // solves wait for the factorisation as well
            if ((rc = launch_geometry(E, d, off, GEO_FIRST))) return rc;
            HIP_TRY(hipEventRecord(E->ev_nodes[bi % BR], st));
            wnodes = E->ev_nodes[bi % BR];
            if ((rc = launch_factor(E, d, off, GEO_FIRST, nullptr))) return rc;
            HIP_TRY(hipEventRecord(E->ev_first[bi % BR], st));
            wfirst = E->ev_first[bi % BR];
        }
        const bool defer = len - off < GEO_DEFER && !(off == 0 && (fresh || due));
        if (!E->per_step && !defer) {
            if ((rc = launch_geometry(E, d, off, GEO_NEXT))) return rc;
            if ((rc = launch_factor(E, d, off, GEO_NEXT, gev))) return rc;
            HIP_TRY(hipEventRecord(E->ev_next[bi % BR], st));
            wnext = prev_next;
            prev_next = E->ev_next[bi % BR];
            if (next_off) *next_off = off;
        } else if (!E->per_step) {
            wnext = prev_next;
            prev_next = nullptr;
            if (gev) {
                HIP_TRY(hipEventRecord(gev[1], st));
                HIP_TRY(hipEventRecord(gev[2], st));
            }
        } else if (gev) {
            HIP_TRY(hipEventRecord(gev[1], st));
            HIP_TRY(hipEventRecord(gev[2], st));
        }
        return 0;
    };
    auto is_boundary = [&](int off) { return (phase + off) % B == 0 || ((fresh || due) && off == 0); };
    for (int off = 0; off < len; ++off) {
        hipEvent_t wfirst = nullptr, wnext = nullptr, wnodes = nullptr;
        if (is_boundary(off)) {
            if (int rc = boundary(off, wfirst, wnext, wnodes)) return rc;
        }
        const bool pair = par && off + 1 < len && off + 1 < pair_limit;
        // a pass entering a new batch at its second step waits at its start
        // for that batch's geometry, built at the boundary before (at most one
        // boundary falls inside a pass: batches are longer than two steps)
        const bool straddle = pair && is_boundary(off + 1);
        if (straddle && !wnext) wnext = prev_next;
        hipEvent_t* ev = nullptr;
        int q = -1;
        if (every > 0 && off % every == 0) q = off / every;
        else if (every > 0 && pair && (off + 1) % every == 0) q = (off + 1) / every;
        if (q >= 0) {
            const size_t need = (size_t)Eper * (q + 1);
            while (E->evpool.size() < need + 2) {
                hipEvent_t e;
                HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));  // timing events
                E->evpool.push_back(e);
            }
            ev = &E->evpool[2 + (size_t)Eper * q];
            E->samp_kind[q] = pair ? 1 : 0;
        }
        int rc;
        if (pair) {
            const int X = (off + *par) & 1;
            rc = launch_pair(E, d, off, off > 0, wfirst, wnext, wnodes, ev, X);
            *par ^= 1;
            ++off;
            if (!rc && straddle) {
                hipEvent_t f2 = nullptr, n2 = nullptr, w2 = nullptr;
                rc = boundary(off, f2, n2, w2);
            }
        } else {
            const int psel = par ? make_psel((off + *par) & 1, (off + 1 + *par) & 1) : -1;
            rc = launch_step(E, d, off, off > 0, wfirst, wnext, wnodes, ev, psel);
        }
        if (rc) return rc;
    }
    if (len > fupto) return enqueue_forces(E, d, fupto, len - 1, E->gstep - 1, true);
    return 0;
}

// After enqueue_error_gather, s0 has waited for every other stream's work of
// the window, so synchronising s0 alone finishes the window (each
// synchronisation of an idle stream still costs about 15 us).
int sync_joined(Engine* E)
{
    HIP_TRY(hipStreamSynchronize(E->s[0]));
    return 0;
}

int upload_clocks(Engine* E, const std::vector<int>& len_end, const std::vector<int>& final_end,
                  const std::vector<int>& due)
{
    const size_t nc = E->members.size();
    E->hclock.resize(nc);
    for (size_t c = 0; c < nc; ++c) {
        const lbmib_handle* h = E->members[c];
        CaseClock& k = E->hclock[c];
        k.n0 = h->steps_done + 1;
        k.n_end = len_end[c];
        k.flag_from = final_end[c] - 1;
        k.fresh = h->geo_fresh ? 1 : 0;
        k.next_due = due[c];
    }
    // asynchronous from pinned memory on s0; the other streams wait for it
    // (every engine call ends with its streams synchronised, so the pinned
    // copy is free again)
    std::memcpy(E->hclk_pin, E->hclock.data(), sizeof(CaseClock) * nc);
    HIP_TRY(hipMemcpyAsync(E->dclock, E->hclk_pin, sizeof(CaseClock) * nc, hipMemcpyHostToDevice, E->s[0]));
    HIP_TRY(hipEventRecord(E->ev_clk, E->s[0]));
    for (int i = 1; i < Engine::NSTREAM; ++i) HIP_TRY(hipStreamWaitEvent(E->s[i], E->ev_clk, 0));
    return 0;
}

// At the end of a window's work: the members' error flags gathered on s0
// after every stream's work, and copied to pinned memory, so that the
// window's closing synchronisation also brings them to the host.
int enqueue_error_gather(Engine* E, const Dims& d)
{
    for (int i = 1; i < Engine::NSTREAM; ++i) {
        HIP_TRY(hipEventRecord(E->ev_end[i], E->s[i]));
        HIP_TRY(hipStreamWaitEvent(E->s[0], E->ev_end[i], 0));
    }
    hipLaunchKernelGGL(k_gather_err, dim3((d.nc + 255) / 256), dim3(256), 0, E->s[0], (const DevCase*)E->dcases, d.nc,
                       E->derr);
    HIP_TRY(hipMemcpyAsync(E->herr_pin, E->derr, sizeof(int) * d.nc, hipMemcpyDeviceToHost, E->s[0]));
    return 0;
}

// The gathered flags, after the window's synchronisation.
int check_errors(Engine* E, const Dims& d)
{
    for (int c = 0; c < d.nc; ++c) {
        const int v = E->herr_pin[c];
        if (!v) continue;
        if (v & ERR_GEOMETRY)
            return fail(LBMIB_E_GEOMETRY, "case %d: an immersed-boundary node left the interior margin (4..nx-3, "
                        "4..ny-3) or came within 3 nodes of the lattice window's edge", c);
        if (v & ERR_CAPACITY) return fail(LBMIB_E_CAPACITY, "case %d: an immersed-boundary list overflowed", c);
        if (v & ERR_SOLVE) return fail(LBMIB_E_SOLVE, "case %d: immersed-boundary system not positive definite", c);
    }
    return 0;
}

// Members' step counters and recorded-macro slots after a segment whose
// cases ran steps n0 .. len_end.
void advance_members(Engine* E, const std::vector<int>& len_end, const std::vector<int>& final_end)
{
    for (size_t c = 0; c < E->members.size(); ++c) {
        lbmib_handle* h = E->members[c];
        const int n0 = h->steps_done + 1, n1 = len_end[c];
        for (int n = std::max(n0, final_end[c] - 1); n <= n1; ++n) h->last_flag_step[n & 1] = n;
        if (n1 >= n0) h->geo_fresh = false;
        h->steps_done = std::max(h->steps_done, n1);
    }
}

// Capture a window of 2*B steps from a batch boundary, no case fresh.
int capture_graph(Engine* E, const Dims& d)
{
    if (E->gexec) {
        (void)hipGraphExecDestroy(E->gexec);
        E->gexec = nullptr;
    }
    const int len = 2 * E->B;
    hipStream_t s0 = E->s[0];
    HIP_TRY(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
    auto abort_capture = [&](int rc) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s0, &g);
        if (g) (void)hipGraphDestroy(g);
        return rc;
    };
    if (hipEventRecord(E->ev_fork, s0) != hipSuccess) return abort_capture(fail(LBMIB_E_HIP, "graph fork record failed"));
    for (int i = 1; i < Engine::NSTREAM; ++i)
        if (hipStreamWaitEvent(E->s[i], E->ev_fork, 0) != hipSuccess)
            return abort_capture(fail(LBMIB_E_HIP, "graph fork wait failed"));
    int rc = launch_window(E, d, len, 0, false, false, nullptr, len);
    if (rc) return abort_capture(rc);
    for (int i = 1; i < Engine::NSTREAM; ++i) {
        if (hipEventRecord(E->ev_join[i], E->s[i]) != hipSuccess || hipStreamWaitEvent(s0, E->ev_join[i], 0) != hipSuccess)
            return abort_capture(fail(LBMIB_E_HIP, "graph join failed"));
    }
    hipLaunchKernelGGL(k_clock_advance, dim3((d.nc + 255) / 256), dim3(256), 0, s0, E->dclock, d.nc, len);
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamEndCapture(s0, &graph));
    hipError_t e = hipGraphInstantiateWithFlags(&E->gexec, graph, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(LBMIB_E_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    E->graph_len = len;
    E->graph_valid = true;
    return 0;
}

}  // namespace

int shape_from_env(Shape& s)
{
    if (const char* e = getenv("LBMIB_IB_CUS")) s.ib_cus = std::max(0, atoi(e)) / 8 * 8;
    return 0;
}

int engine_create(int dev, Engine** out)
{
    *out = nullptr;
    HIP_TRY(hipSetDevice(dev));
    Engine* E = new Engine();
    E->dev = dev;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) {
        delete E;
        return fail(LBMIB_E_HIP, "hipGetDeviceProperties failed: %s", hipGetErrorString(e));
    }
    E->ncu = prop.multiProcessorCount;
    shape_from_env(E->shape);
    // The engine's events order queues of this one device (and pace the
    // host, which reads no memory through them), so they skip the
    // system-scope fence, a cache write-back and invalidation per record
    // that cost 2-4 us per hand-over in the step.
    auto mk = [&](hipEvent_t* ev) {
        return hipEventCreateWithFlags(ev, hipEventDisableTiming | hipEventDisableSystemFence);
    };
    bool ok = true;
    for (int i = 0; i < Engine::RING; ++i)
        for (hipEvent_t* ev : {&E->ev_ib[i], &E->ev_step[i], &E->ev_rhs[i], &E->ev_solve[i], &E->ev_win[i]})
            ok = ok && mk(ev) == hipSuccess;
    for (int i = 0; i < Engine::BRING; ++i)
        for (hipEvent_t* ev : {&E->ev_first[i], &E->ev_next[i], &E->ev_nodes[i]}) ok = ok && mk(ev) == hipSuccess;
    ok = ok && mk(&E->ev_fork) == hipSuccess;
    ok = ok && mk(&E->ev_clk) == hipSuccess;
    for (int i = 0; i < Engine::NSTREAM; ++i) ok = ok && mk(&E->ev_end[i]) == hipSuccess;
    for (int i = 0; i < Engine::NSTREAM; ++i) ok = ok && mk(&E->ev_join[i]) == hipSuccess;
    if (!ok) {
        engine_destroy(E);
        return fail(LBMIB_E_HIP, "event creation failed");
    }
    // streams (with their CU masks, one hardware queue each) are made on the
    // first run: a case that only ever steps inside a batch needs none
    *out = E;
    return 0;
}

void engine_destroy(Engine* E)
{
    if (!E) return;
    (void)hipSetDevice(E->dev);
    destroy_streams(E);
    if (E->gexec) (void)hipGraphExecDestroy(E->gexec);
    for (void* p : {(void*)E->dcases, (void*)E->dclock, (void*)E->dtiles, (void*)E->derr, (void*)E->dfar,
                    (void*)E->dwin})
        if (p) (void)hipFree(p);
    for (hipEvent_t e : E->evpool) (void)hipEventDestroy(e);
    for (hipEvent_t e : E->gevpool) (void)hipEventDestroy(e);
    for (int i = 0; i < Engine::RING; ++i)
        for (hipEvent_t e : {E->ev_ib[i], E->ev_step[i], E->ev_rhs[i], E->ev_solve[i], E->ev_win[i]})
            if (e) (void)hipEventDestroy(e);
    for (int i = 0; i < Engine::BRING; ++i)
        for (hipEvent_t e : {E->ev_first[i], E->ev_next[i], E->ev_nodes[i]})
            if (e) (void)hipEventDestroy(e);
    if (E->ev_fork) (void)hipEventDestroy(E->ev_fork);
    if (E->ev_clk) (void)hipEventDestroy(E->ev_clk);
    for (hipEvent_t e : E->ev_end)
        if (e) (void)hipEventDestroy(e);
    if (E->hclk_pin) (void)hipHostFree(E->hclk_pin);
    if (E->herr_pin) (void)hipHostFree(E->herr_pin);
    for (hipEvent_t e : E->ev_join)
        if (e) (void)hipEventDestroy(e);
    delete E;
}

int engine_set_shape(Engine* E, const Shape& s)
{
    if (s.npt != 1 && s.npt != 2) return fail(LBMIB_E_ARG, "nodes_per_thread must be 1 or 2");
    if (s.order != 0 && s.order != 1) return fail(LBMIB_E_ARG, "tile_order must be 0 or 1");
    if (s.persist < 0 || s.persist > 64) return fail(LBMIB_E_ARG, "persistent must be 0..64");
    if (s.ib_cus < 0 || s.ib_cus % 8) return fail(LBMIB_E_ARG, "ib_cus must be a non-negative multiple of 8");
    if (s.graph != 0 && s.graph != 1) return fail(LBMIB_E_ARG, "graph must be 0 or 1");
    if (s.band != 0 && s.band != 1) return fail(LBMIB_E_ARG, "ib_collide_shared must be 0 or 1");
    if (s.rhs != 0 && s.rhs != 1) return fail(LBMIB_E_ARG, "ib_rhs_reserved must be 0 or 1");
    if (s.unit != 0 && s.unit != 16 && s.unit != 32 && s.unit != 64)
        return fail(LBMIB_E_ARG, "solve_unit must be 0, 16, 32 or 64");
    if (s.fuse != 0 && s.fuse != 1) return fail(LBMIB_E_ARG, "fuse_steps must be 0 or 1");
    // Graph replay is switched off: its first run on the GPU (a three-case
    // batch, replaying captured 16-step windows on CU-masked streams) ended
    // the host process with a segmentation fault whose cause is not known.
    if (s.graph) return fail(LBMIB_E_ARG, "graph replay is switched off in this build (capture of the step crashed "
                             "the host process; cause not identified)");
    if (s.npt != E->shape.npt || s.order != E->shape.order) E->tiles_valid = false;
    E->shape = s;
    E->graph_valid = false;
    return 0;
}

int engine_run(Engine* E, int k)
{
    if (E->members.empty() || k <= 0) return 0;
    int rc = engine_prepare(E);
    if (rc) return rc;
    const Dims d = dims_of(E);
    const size_t nc = E->members.size();
    std::vector<int> final_end(nc), len_end(nc);
    for (size_t c = 0; c < nc; ++c) {
        const lbmib_handle* h = E->members[c];
        final_end[c] = std::min(h->steps_done + k, h->max_steps);
    }
    const int B = E->B;
    const int every = E->prof_every;
    const int Eper = Engine::EPER;
    const bool fuse = fuse_ok(E);
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (every > 0) {
        while (E->evpool.size() < 2) {
            hipEvent_t e;
            HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));  // timing events
            E->evpool.push_back(e);
        }
        t0 = E->evpool[0];
        t1 = E->evpool[1];
        HIP_TRY(hipEventRecord(t0, E->s[0]));
        E->gev_used = 0;
    }
    int steps_total = 0, sampled = 0;
    std::vector<int> due(nc), cur0(nc);
    for (;;) {
        // geometry a window needs before its steps: a member whose current
        // batch was never built (its boundary built nothing, last window)
        // starts fresh; one whose next batch is missing builds it at offset 0
        bool any_due = false;
        for (size_t c = 0; c < nc; ++c) {
            lbmib_handle* h = E->members[c];
            due[c] = 0;
            cur0[c] = h->steps_done / B;
            if (E->per_step || h->geo_fresh || final_end[c] <= h->steps_done) continue;
            if (h->geo_built < cur0[c]) h->geo_fresh = true;
            else if (h->geo_built < cur0[c] + 1) due[c] = 1;
            any_due = any_due || due[c];
        }
        int kwin = 0;
        bool fresh = false;
        for (size_t c = 0; c < nc; ++c) {
            kwin = std::max(kwin, final_end[c] - E->members[c]->steps_done);
            fresh = fresh || E->members[c]->geo_fresh;
        }
        if (kwin <= 0) break;
        // the batch phase is that of the members that still step (members
        // whose motion table has ended idle and may stand anywhere)
        int phase = -1;
        for (size_t c = 0; c < nc; ++c) {
            if (final_end[c] <= E->members[c]->steps_done) continue;
            const int ph = E->members[c]->steps_done % B;
            if (phase < 0) phase = ph;
            else if (ph != phase)
                return fail(LBMIB_E_STATE, "members stand at different positions inside a geometry batch (%d, %d)",
                            phase, ph);
        }
        const bool graph = E->shape.graph && phase == 0 && !fresh && kwin >= 2 * B;
        int len = kwin;
        if (E->shape.graph && !graph) {
            if (phase != 0) len = std::min(kwin, B - phase);
            else if (fresh) len = std::min(kwin, B);
        }
        if (graph) len = kwin / (2 * B) * (2 * B);
        for (size_t c = 0; c < nc; ++c) len_end[c] = std::min(E->members[c]->steps_done + len, final_end[c]);
        // passes only while every member that steps in the window steps
        int pair_limit = len, plane0 = 0;
        if (fuse) {
            for (size_t c = 0; c < nc; ++c) {
                const int mc = len_end[c] - E->members[c]->steps_done;
                if (mc > 0) pair_limit = std::min(pair_limit, mc);
            }
            if (normalize_planes(E, len_end, &plane0) && (rc = upload_cases(E))) return rc;
        }
        if ((rc = upload_clocks(E, len_end, final_end, due))) return rc;
        std::vector<bool> was_fresh(nc);
        for (size_t c = 0; c < nc; ++c) was_fresh[c] = E->members[c]->geo_fresh;
        int next_off = -1;
        if (graph) {
            if (!E->graph_valid) {
                // capture uses the clocks only through device pointers
                if ((rc = capture_graph(E, d))) return rc;
            }
            for (int r = 0; r < len / E->graph_len; ++r) HIP_TRY(hipGraphLaunch(E->gexec, E->s[0]));
            if ((rc = enqueue_error_gather(E, d))) return rc;
        } else {
            if (every > 0) E->gev_used = 0;
            int par = plane0;
            if ((rc = launch_window(E, d, len, phase, fresh, every > 0, fuse ? &par : nullptr, pair_limit, any_due,
                                    &next_off)))
                return rc;
            if ((rc = enqueue_error_gather(E, d))) return rc;
            if (fuse && par != plane0)
                for (size_t c = 0; c < nc; ++c)
                    if (len_end[c] > E->members[c]->steps_done) E->members[c]->D.pofs ^= 1;
            if (every > 0) {
                const int nsamp = (len + every - 1) / every;
                if ((rc = sync_joined(E))) return rc;
                for (int q = 0; q < nsamp; ++q) {
                    const int kind = E->samp_kind[q];
                    if (kind == 2) continue;
                    hipEvent_t* ev = &E->evpool[2 + (size_t)Eper * q];
                    float ms;
                    HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
                    if (kind == 0) {
                        E->tms[3] += ms;
                        E->tms[15] += 1.0;
                    } else {
                        E->tms[11] += ms;
                        E->tms[12] += 1.0;
                        HIP_TRY(hipEventElapsedTime(&ms, ev[5], ev[6]));
                        E->tms[13] += ms;
                        E->tms[14] += 1.0;
                    }
                    HIP_TRY(hipEventElapsedTime(&ms, ev[2], ev[3])); E->tms[4] += ms;
                    HIP_TRY(hipEventElapsedTime(&ms, ev[3], ev[4])); E->tms[5] += ms;
                    ++sampled;
                }
                for (int q = 0; q < E->gev_used; ++q) {
                    hipEvent_t* ev = &E->gevpool[(size_t)3 * q];
                    float ms;
                    HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1])); E->tms[6] += ms;
                    HIP_TRY(hipEventElapsedTime(&ms, ev[1], ev[2])); E->tms[7] += ms;
                }
                E->tms[8] += E->gev_used;
            }
        }
        if ((rc = sync_joined(E))) return rc;
        if ((rc = check_errors(E, d))) return rc;
        // the batches whose geometry now exists: the first one of a fresh
        // member, the one after each boundary that built it
        for (size_t c = 0; c < nc; ++c) {
            lbmib_handle* h = E->members[c];
            if (len_end[c] <= h->steps_done || E->per_step) continue;
            if (was_fresh[c]) h->geo_built = cur0[c];
            if (next_off >= 0 && h->steps_done + 1 + next_off <= len_end[c])
                h->geo_built = std::max(h->geo_built, (h->steps_done + next_off) / B + 1);
        }
        advance_members(E, len_end, final_end);
        steps_total += len;
    }
    if (every > 0 && steps_total > 0) {
        HIP_TRY(hipEventRecord(t1, E->s[0]));
        HIP_TRY(hipEventSynchronize(t1));
        float ms;
        HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
        E->tms[0] += ms;
        E->tms[1] += steps_total;
        E->tms[2] += sampled;
    }
    return 0;
}

}  // namespace lbmib

using namespace lbmib;

struct lbmib_batch {
    Engine* E = nullptr;
    int dev = 0;
};

namespace {

Shape shape_in(const lbmib_launch_shape* s)
{
    Shape o;
    o.npt = s->nodes_per_thread;
    o.order = s->tile_order;
    o.persist = s->persistent;
    o.ib_cus = s->ib_cus;
    o.graph = s->graph;
    o.band = s->ib_collide_shared;
    o.rhs = s->ib_rhs_reserved;
    o.unit = s->solve_unit;
    o.fuse = s->fuse_steps;
    return o;
}

void shape_out(const Shape& s, lbmib_launch_shape* o)
{
    std::memset(o, 0, sizeof *o);
    o->nodes_per_thread = s.npt;
    o->tile_order = s.order;
    o->persistent = s.persist;
    o->ib_cus = s.ib_cus;
    o->graph = s.graph;
    o->ib_collide_shared = s.band;
    o->ib_rhs_reserved = s.rhs;
    o->solve_unit = s.unit;
    o->fuse_steps = s.fuse;
}

}  // namespace

extern "C" {

int lbmib_set_shape(lbmib_handle* h, const lbmib_launch_shape* s)
{
    if (!h || !s) return fail(LBMIB_E_ARG, "null argument");
    return engine_set_shape(h->own, shape_in(s));
}

int lbmib_get_shape(const lbmib_handle* h, lbmib_launch_shape* s)
{
    if (!h || !s) return fail(LBMIB_E_ARG, "null argument");
    shape_out(h->own->shape, s);
    return 0;
}

int lbmib_batch_create(int32_t device, lbmib_batch** out)
{
    if (!out) return fail(LBMIB_E_ARG, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(LBMIB_E_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(LBMIB_E_ARG, "device %d out of range (%d devices)", device, ndev);
    lbmib_batch* b = new lbmib_batch();
    b->dev = device;
    int rc = engine_create(device, &b->E);
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return 0;
}

int lbmib_batch_add(lbmib_batch* b, lbmib_handle* h)
{
    if (!b || !h) return fail(LBMIB_E_ARG, "null argument");
    if (h->batch) return fail(LBMIB_E_STATE, "the case already belongs to a batch");
    if (!h->have_motion) return fail(LBMIB_E_STATE, "set the case's motion table before adding it to a batch");
    if (h->dyn.enabled) return fail(LBMIB_E_STATE, "a case with body dynamics takes its steps alone");
    if (h->dev != b->dev) return fail(LBMIB_E_ARG, "the case lives on device %d, the batch on %d", h->dev, b->dev);
    Engine* E = b->E;
    if (!E->members.empty()) {
        if (E->members[0]->D.prec != h->D.prec)
            return fail(LBMIB_E_ARG, "all cases of a batch must use the same precision");
        // the position inside the geometry batch is set by the members that
        // still have steps to take; finished members idle wherever they stand
        for (const lbmib_handle* m : E->members) {
            if (m->steps_done >= m->max_steps) continue;
            if (m->steps_done % E->B != h->steps_done % E->B)
                return fail(LBMIB_E_STATE, "the case stands at step %d, the batch at %d (mod %d): cases join at the "
                            "same position inside a geometry batch", h->steps_done, m->steps_done, E->B);
            break;
        }
    }
    if (h->D.batch != E->B) return fail(LBMIB_E_STATE, "geometry batch length differs");
    E->members.push_back(h);
    h->batch = b;
    E->tiles_valid = false;
    E->graph_valid = false;
    return 0;
}

int lbmib_batch_remove(lbmib_batch* b, lbmib_handle* h)
{
    if (!b || !h) return fail(LBMIB_E_ARG, "null argument");
    if (h->batch != b) return fail(LBMIB_E_STATE, "the case is not in this batch");
    Engine* E = b->E;
    for (hipStream_t st : E->s)
        if (st) HIP_TRY(hipStreamSynchronize(st));
    E->members.erase(std::remove(E->members.begin(), E->members.end(), h), E->members.end());
    h->batch = nullptr;
    E->tiles_valid = false;
    E->graph_valid = false;
    return 0;
}

int lbmib_batch_size(const lbmib_batch* b, int32_t* n)
{
    if (!b || !n) return fail(LBMIB_E_ARG, "null argument");
    *n = (int32_t)b->E->members.size();
    return 0;
}

int lbmib_batch_step(lbmib_batch* b, int32_t n_steps)
{
    if (!b) return fail(LBMIB_E_ARG, "null batch");
    if (n_steps < 0) return fail(LBMIB_E_ARG, "n_steps must be >= 0");
    int rc = engine_run(b->E, n_steps);
    if (rc) return rc;
    for (lbmib_handle* h : b->E->members)
        if ((rc = nan_guard_check(h))) return rc;
    return 0;
}

int lbmib_batch_set_shape(lbmib_batch* b, const lbmib_launch_shape* s)
{
    if (!b || !s) return fail(LBMIB_E_ARG, "null argument");
    return engine_set_shape(b->E, shape_in(s));
}

int lbmib_batch_get_shape(const lbmib_batch* b, lbmib_launch_shape* s)
{
    if (!b || !s) return fail(LBMIB_E_ARG, "null argument");
    shape_out(b->E->shape, s);
    return 0;
}

int lbmib_batch_set_profiling(lbmib_batch* b, int32_t every)
{
    if (!b) return fail(LBMIB_E_ARG, "null batch");
    if (every < 0) return fail(LBMIB_E_ARG, "profiling interval must be >= 0");
    b->E->prof_every = every;
    for (double& t : b->E->tms) t = 0.0;
    return 0;
}

int lbmib_batch_get_timings(const lbmib_batch* b, double* t, int32_t n)
{
    if (!b || !t) return fail(LBMIB_E_ARG, "null argument");
    for (int i = 0; i < n && i < LBMIB_NTIMINGS; ++i) t[i] = b->E->tms[i];
    for (int i = LBMIB_NTIMINGS; i < n; ++i) t[i] = 0.0;
    return 0;
}

// Times candidate launch shapes on copies of the members (same lattice,
// markers and motion; their own state), coordinate by coordinate: lattice
// tile shape and order with the persistent-grid choice, then the CUs
// reserved for the solves, then graph replay.  Each trial runs `trial_steps`
// timed steps after one untimed geometry batch.  The best shape is applied
// to the batch; the report lists every trial as JSON.  (Graph replay, while
// switched off, is not tried.)
int lbmib_batch_autotune(lbmib_batch* b, int32_t trial_steps, lbmib_launch_shape* best, char* report,
                         int32_t report_len)
{
    if (!b) return fail(LBMIB_E_ARG, "null batch");
    if (trial_steps < 1) return fail(LBMIB_E_ARG, "trial_steps must be >= 1");
    Engine* E = b->E;
    if (E->members.empty()) return fail(LBMIB_E_STATE, "the batch is empty");
    std::vector<lbmib_handle*> clones;
    auto cleanup = [&]() {
        for (lbmib_handle* c : clones) lbmib_destroy(c);
    };
    for (lbmib_handle* m : E->members) {
        lbmib_handle* c = nullptr;
        int rc = lbmib_create(&m->p, m->in_xb0.data(), m->in_yb0.data(), &c);
        if (!rc) {
            clones.push_back(c);
            rc = lbmib_set_motion_table(c, m->motion_in.data(), (int32_t)(m->motion_in.size() / m->ncols), m->ncols);
        }
        if (rc) {
            cleanup();
            return rc;
        }
    }
    Engine* T = nullptr;
    int rc = engine_create(b->dev, &T);
    if (rc) {
        cleanup();
        return rc;
    }
    T->members = clones;
    int room = INT32_MAX;
    for (lbmib_handle* c : clones) room = std::min(room, c->max_steps);
    std::string rep = "{\"trials\":[";
    bool firstrec = true;
    // the first geometry batch of the copies, outside every trial
    rc = engine_run(T, T->B);
    const int timed = (trial_steps + T->B - 1) / T->B * T->B;
    auto trial = [&](const Shape& s, double& ms) -> int {
        // warm-up of one geometry batch, then the timed batches
        if (clones[0]->steps_done + T->B + timed > room) return 1;  // out of motion table
        int r = engine_set_shape(T, s);
        if (r) return r;
        if ((r = engine_run(T, T->B))) return r;
        auto t0 = std::chrono::steady_clock::now();
        if ((r = engine_run(T, timed))) return r;
        auto t1 = std::chrono::steady_clock::now();
        ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / timed;
        char buf[256];
        snprintf(buf, sizeof buf, "%s{\"nodes_per_thread\":%d,\"tile_order\":%d,\"persistent\":%d,\"ib_cus\":%d,"
                 "\"ib_collide_shared\":%d,\"ib_rhs_reserved\":%d,\"solve_unit\":%d,\"graph\":%d,"
                 "\"ms_per_step\":%.5f}",
                 firstrec ? "" : ",", s.npt, s.order, s.persist, s.ib_cus, s.band, s.rhs, s.unit, s.graph, ms);
        rep += buf;
        firstrec = false;
        return 0;
    };
    Shape cur = E->shape;
    double cur_ms = 1e300;
    auto consider = [&](const Shape& s) -> int {
        double ms = 0.0;
        int r = trial(s, ms);
        if (r < 0) return r;
        if (r == 0 && ms < cur_ms) {
            cur_ms = ms;
            cur = s;
        }
        return 0;
    };
    if (!rc) {
        Shape s0 = cur;
        s0.graph = 0;
        for (int npt : {1, 2})
            for (int order : {0, 1})
                for (int persist : {0, 2, 4, 8}) {
                    Shape s = s0;
                    s.npt = npt; s.order = order; s.persist = persist;
                    if ((rc = consider(s))) break;
                }
        if (!rc) {
            const Shape base = cur;
            for (int cus : {0, 8, 16, 32})
                for (int band : {0, 1})
                    for (int rhs : {0, 1}) {
                        if (rc || (cus == base.ib_cus && band == base.band && rhs == base.rhs)) continue;
                        Shape s = base;
                        s.ib_cus = cus;
                        s.band = band;
                        s.rhs = rhs;
                        rc = consider(s);
                    }
        }
        // graph replay stays out of the trials while it is switched off
    }
    engine_destroy(T);
    cleanup();
    if (rc) return rc;
    if (cur_ms >= 1e300) return fail(LBMIB_E_STATE, "the members' motion tables are too short for a trial");
    rc = engine_set_shape(E, cur);
    if (rc) return rc;
    char buf[256];
    snprintf(buf, sizeof buf, "],\"best\":{\"nodes_per_thread\":%d,\"tile_order\":%d,\"persistent\":%d,\"ib_cus\":%d,"
             "\"ib_collide_shared\":%d,\"ib_rhs_reserved\":%d,\"solve_unit\":%d,\"graph\":%d,"
             "\"ms_per_step\":%.5f}}", cur.npt, cur.order, cur.persist, cur.ib_cus, cur.band, cur.rhs, cur.unit,
             cur.graph, cur_ms);
    rep += buf;
    if (best) shape_out(cur, best);
    if (report && report_len > 0) {
        std::strncpy(report, rep.c_str(), (size_t)report_len - 1);
        report[report_len - 1] = '\0';
        if ((int)rep.size() >= report_len) return fail(LBMIB_E_ARG, "report buffer too small (%zu bytes needed)", rep.size() + 1);
    }
    return 0;
}

int lbmib_batch_destroy(lbmib_batch* b)
{
    if (!b) return 0;
    for (lbmib_handle* h : b->E->members) h->batch = nullptr;
    b->E->members.clear();
    engine_destroy(b->E);
    delete b;
    return 0;
}

}  // extern "C"
