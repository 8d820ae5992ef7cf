// lbmib_kernels.hip -- gfx950 kernels of the timestep: the FP64 drop-in and
// the FP32 deviation form (populations stored as f - W in FP32).
//
// One timestep (reference evolution, solution.f90:269-410) is
//   lattice:  k_lattice<T>   fused pull-stream + moments + MRT collide of every
//                            interior node, prev -> cur (HBM-bound: 144 B/node in
//                            FP64, 72 B/node in FP32), and the characteristic/NEE
//                            outer boundary ring (FP64 arithmetic in both forms)
//   IB:       geometry of the next step (node set, kernel weights, L D L^T
//             factor of D D^T) on its own stream; per step u* -> D u* ->
//             triangular solves -> spread + forced collide of the IB nodes;
//             per-wing forces in fixed order
// Nothing here uses floating-point atomics; every reduction has a fixed order.
#include "lbmib_kernels.h"

#include <climits>
#include <type_traits>

namespace lbmib {

// ---------------------------------------------------------------------------
// lattice
// ---------------------------------------------------------------------------

// Population planes are global memory: pointers read from the case table
// are generic, so they are cast to the global address space to keep every
// population access a global (not flat) memory instruction.
#define GAS __attribute__((address_space(1)))

// Workgroup barrier for LDS traffic only: waits for this wave's LDS
// operations, not for its outstanding global stores.
__device__ __forceinline__ void lds_barrier()
{
    __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Stored population -> FP64 population: FP64 storage holds f itself, the
// FP32 deviation form holds g = f - W.
__device__ __forceinline__ double lift(double v, double) { return v; }
__device__ __forceinline__ double lift(float v, double w) { return w + (double)v; }
__device__ __forceinline__ void lower(double f, double, GAS double* dst) { *dst = f; }
__device__ __forceinline__ void lower(double f, double w, GAS float* dst) { *dst = (float)(f - w); }

template <typename T> __device__ __forceinline__ GAS T* pops(const DevCase& D, int i);
template <> __device__ __forceinline__ GAS double* pops<double>(const DevCase& D, int i) { return (GAS double*)D.f[i]; }
template <> __device__ __forceinline__ GAS float* pops<float>(const DevCase& D, int i) { return (GAS float*)D.g[i]; }

// Pull the nine populations of interior node (i,j) from prev
// (streaming, solution.f90:438-454: f0(i,j,k) = f(i-ex,j-ey,k)), in FP64.
template <typename T>
__device__ __forceinline__ void pull9(const GAS T* __restrict__ src, long long N, int nx, long long c,
                                      double* f, const double* W)
{
    f[0] = lift(src[c], W[0]);
    f[1] = lift(src[1 * N + c - 1], W[1]);
    f[2] = lift(src[2 * N + c - nx], W[2]);
    f[3] = lift(src[3 * N + c + 1], W[3]);
    f[4] = lift(src[4 * N + c + nx], W[4]);
    f[5] = lift(src[5 * N + c - nx - 1], W[5]);
    f[6] = lift(src[6 * N + c - nx + 1], W[6]);
    f[7] = lift(src[7 * N + c + nx + 1], W[7]);
    f[8] = lift(src[8 * N + c + nx - 1], W[8]);
}

// Moments of a boundary-ring node from its stale f0 (never streamed,
// solution.f90:516-520).
__device__ __forceinline__ void ring_moments(const GAS double* f0col, int ny, int j, double& r,
                                             double& u, double& v)
{
    double f[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) f[k] = f0col[k * ny + j];
    moments(f, r, u, v);
}

// Left inflow characteristic update at row j (boundary_condition.f90:418-440):
// inputs the ring moments of column 1 at j (tr,tu,tv) and the interior
// moments of columns 2,3.
__device__ __forceinline__ void left_char(double cs, double tr, double tu, double tv, double r2,
                                          double u2, double v2, double r3, double u3, double v3,
                                          double& ro, double& uo, double& vo)
{
    double rhox = (-3.0 * tr + 4.0 * r2 - r3) * 0.5;
    double ux = (-3.0 * tu + 4.0 * u2 - u3) * 0.5;
    double vx = (-3.0 * tv + 4.0 * v2 - v3) * 0.5;
    double lx1 = (tu - cs) * (cs * cs * rhox - cs * tr * ux);
    double lx2 = (tu < 0.0) ? tu * vx : 0.0;
    double lx3 = 0.0;
    ro = tr - (lx1 + lx3) * 1.5;
    uo = tu + (lx1 - lx3) / (2.0 * cs * ro);
    vo = tv - lx2;
}

// Right outflow characteristic update (boundary_condition.f90:500-521).
__device__ __forceinline__ void right_char(double cs, double tr, double tu, double tv, double r2,
                                           double u2, double v2, double r3, double u3, double v3,
                                           double& ro, double& uo, double& vo)
{
    double rhox = (3.0 * tr - 4.0 * r2 + r3) * 0.5;
    double ux = (3.0 * tu - 4.0 * u2 + u3) * 0.5;
    double vx = (3.0 * tv - 4.0 * v2 + v3) * 0.5;
    double lx1 = 0.0;
    double lx2 = (tu >= 0.0) ? tu * vx : 0.0;
    double lx3 = (tu + cs) * (cs * cs * rhox + cs * tr * ux);
    ro = tr - (lx1 + lx3) * 1.5;
    uo = tu + (lx1 - lx3) / (2.0 * cs * ro);
    vo = tv - lx2;
}

// The steps below read the previous step's populations through a pull
// accessor: pull(i, j, f) gives the nine populations streamed into node
// (i, j), f[k] = prev(i - e_kx, j - e_ky, k), in FP64.  GlobalPull reads a
// population plane; the two-step pass reads its LDS ring instead.  Both
// feed the same arithmetic.
template <typename T>
struct GlobalPull {
    const GAS T* src;
    long long N;
    int nx;
    const double* W;
    __device__ __forceinline__ void operator()(int i, int j, double* f) const
    {
        pull9(src, N, nx, (long long)j * nx + i, f, W);
    }
};

// Edge state of column `col` (0: left, 1: right) at interior row j: the
// characteristic-updated rho/u/v and the neighbour's f0 and moments.
template <typename PULL>
__device__ __forceinline__ void edge_column_state(const DevCase& D, const PULL& pull, double cs, int col, int j,
                                                  double& ro, double& uo, double& vo, double* f0n, double& rn,
                                                  double& un, double& vn)
{
    const int nx = D.nx;
    const GAS double* f0c = (const GAS double*)(col == 0 ? D.f0L : D.f0R);
    double tr, tu, tv;
    ring_moments(f0c, D.ny, j, tr, tu, tv);
    int i2 = col == 0 ? 1 : nx - 2;
    int i3 = col == 0 ? 2 : nx - 3;
    double f3[9], r3, u3, v3;
    pull(i2, j, f0n);
    moments(f0n, rn, un, vn);
    pull(i3, j, f3);
    moments(f3, r3, u3, v3);
    if (col == 0)
        left_char(cs, tr, tu, tv, rn, un, vn, r3, u3, v3, ro, uo, vo);
    else
        right_char(cs, tr, tu, tv, rn, un, vn, r3, u3, v3, ro, uo, vo);
}

// boundary_condition1 (boundary_condition.f90:288-858) with no periodic
// edges at ring node (i, j): the left and right columns at rows 1..ny-2 take
// the characteristic state and the NEE from the next column; the top and
// bottom rows take rho of the next row at rest (the reference's top and
// bottom characteristic updates and corner states are overwritten by rho =
// rho(.,Ny-1), u = v = 0 before use, 658-661 and 785-788, so only that
// effective state is computed), the corners reading the left/right results
// at rows 1 and ny-2, recomputed here.  Reads only the previous populations
// and the stale ring f0.
// The top or bottom wall node away from the corners, from f0n, the pull of
// its neighbour node in the next row in (rho of that node, at rest).
__device__ __forceinline__ void wall_node(const DevCase& D, const double* f0n, double* out, double& ro, double& uo,
                                          double& vo)
{
    double rn, un, vn;
    moments(f0n, rn, un, vn);
    ro = rn;
    uo = 0.0;
    vo = 0.0;
    nee_node(ro, uo, vo, rn, un, vn, f0n, D.R, out);
}

template <typename PULL>
__device__ __forceinline__ void ring_node(const DevCase& D, const PULL& pull, int i, int j, double* out, double& ro,
                                          double& uo, double& vo)
{
    const int nx = D.nx, ny = D.ny;
    if (j >= 1 && j <= ny - 2) {
        double f0n[9], rn, un, vn;
        edge_column_state(D, pull, D.cs, i == 0 ? 0 : 1, j, ro, uo, vo, f0n, rn, un, vn);
        nee_node(ro, uo, vo, rn, un, vn, f0n, D.R, out);
        return;
    }
    // top wall (row ny-1, neighbour ny-2) or bottom wall (row 0, neighbour 1)
    const int jn = j == 0 ? 1 : ny - 2;
    double f0n[9];
    if (i == 0 || i == nx - 1) {
        const int col = i == 0 ? 0 : 1;
        double rn, un, vn, fdummy[9], r2, u2, v2;
        edge_column_state(D, pull, D.cs, col, jn, rn, un, vn, fdummy, r2, u2, v2);
        const GAS double* f0c = (const GAS double*)(col == 0 ? D.f0L : D.f0R);
#pragma unroll
        for (int k = 0; k < 9; ++k) f0n[k] = f0c[k * ny + jn];
        ro = rn;
        uo = 0.0;
        vo = 0.0;
        nee_node(ro, uo, vo, rn, un, vn, f0n, D.R, out);
    } else {
        pull(i, jn, f0n);
        wall_node(D, f0n, out, ro, uo, vo);
    }
}

// Ring thread t of a lattice update: threads 0..ny-3 take the left and right
// columns at rows 1..ny-2, threads ny-2..ny-2+nx-1 the top and bottom rows at
// column t-(ny-2).  Runs beside the interior update.
template <typename T>
__device__ __forceinline__ void ring_thread(const DevCase& D, int src, int dst, double* __restrict__ macro, int t)
{
    const int nx = D.nx, ny = D.ny;
    const long long N = D.N;
    const GlobalPull<T> pull{pops<T>(D, src), N, nx, D.W};
    GAS T* out_p = pops<T>(D, dst);
    int ii[2], jj[2];
    if (t < ny - 2) {
        ii[0] = 0; ii[1] = nx - 1; jj[0] = jj[1] = t + 1;
    } else {
        t -= ny - 2;
        if (t >= nx) return;
        ii[0] = ii[1] = t; jj[0] = ny - 1; jj[1] = 0;
    }
    for (int q = 0; q < 2; ++q) {
        double out[9], ro, uo, vo;
        ring_node(D, pull, ii[q], jj[q], out, ro, uo, vo);
        long long c = (long long)jj[q] * nx + ii[q];
#pragma unroll
        for (int k = 0; k < 9; ++k) lower(out[k], D.W[k], out_p + k * N + c);
        if (macro) { macro[c] = ro; macro[N + c] = uo; macro[2 * N + c] = vo; }
    }
}

// The moments a node's update leaves behind, in the form its arithmetic
// has them: rho, u, v in FP64; in the FP32 deviation form the density
// deviation and the momentum, from which rho, u, v follow only when the step
// is recorded.
struct NodeMom64 { double rho, u, v; };
struct NodeMom32 { float drho, jx, jy; };
template <typename T> struct NodeMomOf { typedef NodeMom64 type; };
template <> struct NodeMomOf<float> { typedef NodeMom32 type; };

// Fused streaming + physical_variables + colission (solution.f90:428-643) of
// one interior node without forcing, from its pulled populations p (the
// stored type: FP64 populations, or FP32 deviations g = f - W) into out.
// gforce is zero away from the IB nodes, so the force moments vanish
// exactly.
__device__ __forceinline__ void interior_step(const DevCase& D, const double* p, double* out, NodeMom64& m)
{
#pragma unroll
    for (int k = 0; k < 9; ++k) out[k] = p[k];
    moments(out, m.rho, m.u, m.v);
    collide_node(out, m.rho, m.u, m.v, D.R, false, 0.0, 0.0);
}
__device__ __forceinline__ void interior_step(const DevCase& D, const float* p, float* out, NodeMom32& m)
{
#pragma unroll
    for (int k = 0; k < 9; ++k) out[k] = p[k];
    collide_dev_f32(out, D.K, m.drho, m.jx, m.jy);
}

// rho/u/v of a recorded step into the macroscopic planes at node c.
template <typename P>
__device__ __forceinline__ void record_macro(const DevCase&, const NodeMom64& m, P* mac, long long N, long long c)
{
    mac[c] = m.rho;
    mac[N + c] = m.u;
    mac[2 * N + c] = m.v;
}
template <typename P>
__device__ __forceinline__ void record_macro(const DevCase& D, const NodeMom32& m, P* mac, long long N, long long c)
{
    const double rho = D.rho0 + (double)m.drho;
    mac[c] = rho;
    mac[N + c] = (double)m.jx / rho;
    mac[2 * N + c] = (double)m.jy / rho;
}

// The nine stored populations streamed into node c = j*nx + i of plane src.
template <typename T>
__device__ __forceinline__ void pull_raw(const GAS T* __restrict__ src, long long N, int nx, long long c, T* g)
{
    g[0] = src[c];
    g[1] = src[1 * N + c - 1];
    g[2] = src[2 * N + c - nx];
    g[3] = src[3 * N + c + 1];
    g[4] = src[4 * N + c + nx];
    g[5] = src[5 * N + c - nx - 1];
    g[6] = src[6 * N + c - nx + 1];
    g[7] = src[7 * N + c + nx + 1];
    g[8] = src[8 * N + c + nx - 1];
}

// Is (i, j) one of step slot g's IB nodes?  (Only nodes inside the lattice
// window of the run can be.)
__device__ __forceinline__ bool is_ib_node(const DevCase& D, int g, int i, int j)
{
    const unsigned wi = (unsigned)(i - D.gi0), wj = (unsigned)(j - D.gj0);
    if (wi < (unsigned)D.gw && wj < (unsigned)D.gh) {
        const GAS int* nodeidx = (const GAS int*)(D.ipool + (long long)g * D.io.total + D.io.nodeidx);
        return nodeidx[(long long)wj * D.gw + wi] >= 0;
    }
    return false;
}

// One interior node (i, j) of the lattice update src -> dst.  The IB nodes
// of the step (slot g) are skipped: k_band collides them with their
// forcing, concurrently.  Optionally records rho/u/v (download / residual
// support).
template <typename T>
__device__ __forceinline__ void lattice_node(const DevCase& D, int g, int src, int dst, double* __restrict__ macro,
                                             int i, int j)
{
    const int nx = D.nx;
    const long long N = D.N;
    if (i < 1 || i > nx - 2) return;
    if (is_ib_node(D, g, i, j)) return;
    const long long c = (long long)j * nx + i;
    T p[9], out[9];
    pull_raw(pops<T>(D, src), N, nx, c, p);
    typename NodeMomOf<T>::type m;
    interior_step(D, p, out, m);
    if (macro) record_macro(D, m, macro, N, c);
    GAS T* __restrict__ o = pops<T>(D, dst);
#pragma unroll
    for (int k = 0; k < 9; ++k) __builtin_nontemporal_store(out[k], o + k * N + c);
}

// One lattice update of every case of the window: the interior nodes and
// the outer boundary ring (boundary_condition1).  Work is cut into tiles;
// tile (c, b) of case c is ring block b (256 ring threads) for
// b < lattice_ring_blocks, otherwise the 256 * NPT consecutive nodes
// lattice_base(nx) + q * 256 * NPT .. of the node order c = j * nx + i, with
// q = b - ring blocks.  Tiles run on over row ends, so no lane idles at a
// row's short last tile, and they start at multiples of 64 nodes, so every
// wave writes whole cache lines; lanes on the ring columns and rows skip.
// Thread t of a tile takes nodes t, t + 256, ..  A grid smaller than the
// tile count makes the kernel persistent: block v takes tiles v,
// v + gridDim.x, ..  The order of the
// tile table (case-major, or dealt to the accelerator's XCDs) is the
// launch shape's choice; the result does not depend on it.
template <typename T, int NPT>
__global__ __launch_bounds__(256) void k_lattice(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                                 const int2* __restrict__ tiles, int ntiles, int off, int psel)
{
    for (int v = blockIdx.x; v < ntiles; v += gridDim.x) {
        // the tile's case and block are uniform over the workgroup: keep them
        // (and so every field of the case) in scalar registers
        const int2 tl = tiles[v];
        const int cix = __builtin_amdgcn_readfirstlane(tl.x);
        const int blk = __builtin_amdgcn_readfirstlane(tl.y);
        const DevCase& D = C[cix];
        const CaseClock kc = K[cix];
        const int n = kc.n0 + off;
        if (n > kc.n_end) continue;
        int src, dst;
        step_planes(D, n, psel, src, dst);
        double* macro = n >= kc.flag_from ? D.macro[n & 1] : nullptr;
        const int rb = lattice_ring_blocks(D.nx, D.ny);
        if (blk < rb) {
            ring_thread<T>(D, src, dst, macro, blk * 256 + threadIdx.x);
            continue;
        }
        const int q = blk - rb;
        const int g = (n - 1) % D.nslot;
        const int nx = D.nx;
        const int c0 = lattice_base(nx) + q * 256 * NPT;  // wave-uniform
        const int j0 = c0 / nx;
#pragma unroll
        for (int r = 0; r < NPT; ++r) {
            int i = c0 - j0 * nx + r * 256 + (int)threadIdx.x, j = j0;
            while (i >= nx) {  // at most (256 * NPT) / nx + 1 rows on
                i -= nx;
                ++j;
            }
            if (j >= 1 && j <= D.ny - 2) lattice_node<T>(D, g, src, dst, macro, i, j);
        }
    }
}
template __global__ void k_lattice<double, 1>(const DevCase*, const CaseClock*, const int2*, int, int, int);
template __global__ void k_lattice<double, 2>(const DevCase*, const CaseClock*, const int2*, int, int, int);
template __global__ void k_lattice<float, 1>(const DevCase*, const CaseClock*, const int2*, int, int, int);
template __global__ void k_lattice<float, 2>(const DevCase*, const CaseClock*, const int2*, int, int, int);

// One step of the nodes of the lattice window W (the rectangle gi0.. x
// gj0.. holding every IB node of the run) but its IB nodes, for an engine of
// one case whose far field takes two steps per pass (k_far2).  Thread t of
// the grid takes node t of W in row order.
template <typename T>
__device__ __forceinline__ void window_thread(const DevCase& D, const CaseClock& kc, int off, int psel, long long t)
{
    const int n = kc.n0 + off;
    if (n > kc.n_end) return;
    if (t >= (long long)D.gw * D.gh) return;
    int src, dst;
    step_planes(D, n, psel, src, dst);
    double* macro = n >= kc.flag_from ? D.macro[n & 1] : nullptr;
    const int i = D.gi0 + (int)(t % D.gw);
    const int j = D.gj0 + (int)(t / D.gw);
    lattice_node<T>(D, (n - 1) % D.nslot, src, dst, macro, i, j);
}

// Step n of one node of the outer ring, src -> dst, thread t: threads
// 0..ny-3 the left column at rows 1..ny-2, the next ny-2 the right column,
// then nx the top row and nx the bottom row.  Records rho/u/v when step n is
// recorded.
template <typename T>
__device__ __forceinline__ void ring_thread1(const DevCase& D, const CaseClock& kc, int off, int psel, int t)
{
    const int n = kc.n0 + off;
    if (n > kc.n_end) return;
    const int nx = D.nx, ny = D.ny;
    int i, j;
    if (t < 2 * (ny - 2)) {
        const bool right = t >= ny - 2;
        i = right ? nx - 1 : 0;
        j = 1 + (right ? t - (ny - 2) : t);
    } else {
        t -= 2 * (ny - 2);
        if (t >= 2 * nx) return;
        const bool bottom = t >= nx;
        i = bottom ? t - nx : t;
        j = bottom ? 0 : ny - 1;
    }
    int src, dst;
    step_planes(D, n, psel, src, dst);
    double* macro = n >= kc.flag_from ? D.macro[n & 1] : nullptr;
    const long long N = D.N;
    const GlobalPull<T> pull{pops<T>(D, src), N, nx, D.W};
    double out[9], ro, uo, vo;
    ring_node(D, pull, i, j, out, ro, uo, vo);
    const long long c = (long long)j * nx + i;
    GAS T* out_p = pops<T>(D, dst);
#pragma unroll
    for (int k = 0; k < 9; ++k) lower(out[k], D.W[k], out_p + k * N + c);
    if (macro) { macro[c] = ro; macro[N + c] = uo; macro[2 * N + c] = vo; }
}

// Step n of one node of the layer next to the outer ring (columns 1 and
// nx-2 at rows 1..ny-2, rows 1 and ny-2 at columns 2..nx-3), src -> dst,
// thread t: threads 0..ny-3 the left column, the next ny-2 the right one,
// then nx-4 the top row and nx-4 the bottom row.  The ordinary interior
// update (the far pass leaves these nodes' second step out: their pulls
// reach the ring's step n, which k_ring forms beside that pass).
template <typename T>
__device__ __forceinline__ void layer_thread(const DevCase& D, const CaseClock& kc, int off, int psel, int t)
{
    const int n = kc.n0 + off;
    if (n > kc.n_end) return;
    const int nx = D.nx, ny = D.ny;
    int i, j;
    if (t < 2 * (ny - 2)) {
        const bool right = t >= ny - 2;
        i = right ? nx - 2 : 1;
        j = 1 + (right ? t - (ny - 2) : t);
    } else {
        t -= 2 * (ny - 2);
        if (t >= 2 * (nx - 4)) return;
        const bool bottom = t >= nx - 4;
        i = 2 + (bottom ? t - (nx - 4) : t);
        j = bottom ? 1 : ny - 2;
    }
    int src, dst;
    step_planes(D, n, psel, src, dst);
    double* macro = n >= kc.flag_from ? D.macro[n & 1] : nullptr;
    lattice_node<T>(D, (n - 1) % D.nslot, src, dst, macro, i, j);
}

// The case of a table entry, made wave-uniform so that every field of the
// case stays in scalar registers.
__device__ __forceinline__ int table_case(int c) { return __builtin_amdgcn_readfirstlane(c); }

template <typename T>
__global__ __launch_bounds__(256) void k_window(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                                const int2* __restrict__ tab, int off, int psel)
{
    const int2 e = tab[blockIdx.x];
    const int c = table_case(e.x);
    const CaseClock kc = K[c];
    window_thread<T>(C[c], kc, off, psel, (long long)e.y * 256 + threadIdx.x);
}
template __global__ void k_window<double>(const DevCase*, const CaseClock*, const int2*, int, int);
template __global__ void k_window<float>(const DevCase*, const CaseClock*, const int2*, int, int);

// Step n of the outer ring alone, src -> dst (the middle plane of a two-step
// pass, which then reads the ring's step n there instead of recomputing it
// in phase 1 one row at a time); one node per thread.
template <typename T>
__global__ __launch_bounds__(256) void k_ring(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                             const int2* __restrict__ tab, int off, int psel)
{
    const int2 e = tab[blockIdx.x];
    const int c = table_case(e.x);
    const CaseClock kc = K[c];
    ring_thread1<T>(C[c], kc, off, psel, e.y * 256 + (int)threadIdx.x);
}
template __global__ void k_ring<double>(const DevCase*, const CaseClock*, const int2*, int, int);
template __global__ void k_ring<float>(const DevCase*, const CaseClock*, const int2*, int, int);

// The tail of a two-step pass: step n+1 (offset off) of the window (planes
// psel, a case's workgroups 0..window_blocks-1), of the outer ring and of the
// layer next to it (planes ring_psel: the middle plane to the pass's output
// plane; the case's next ring_blocks1 workgroups, then layer_blocks), in one
// launch.
// They touch disjoint nodes: the window step reads the middle plane inside W
// and writes W; the ring and layer steps read the ring's neighbourhood of
// the middle plane (k_ring put the ring's step n there, the far pass that of
// the nodes inwards of it) and write the ring and the layer.
template <typename T>
__global__ __launch_bounds__(256) void k_window_ring(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                                     const int2* __restrict__ tab, int off, int psel, int ring_psel)
{
    const int2 e = tab[blockIdx.x];
    const int c = table_case(e.x);
    const DevCase& D = C[c];
    const CaseClock kc = K[c];
    const int wb = window_blocks(D.gw, D.gh), rb = ring_blocks1(D.nx, D.ny), b = e.y;
    if (b < wb) window_thread<T>(D, kc, off, psel, (long long)b * 256 + threadIdx.x);
    else if (b < wb + rb) ring_thread1<T>(D, kc, off, ring_psel, (b - wb) * 256 + (int)threadIdx.x);
    else layer_thread<T>(D, kc, off, ring_psel, (b - wb - rb) * 256 + (int)threadIdx.x);
}
template __global__ void k_window_ring<double>(const DevCase*, const CaseClock*, const int2*, int, int, int);
template __global__ void k_window_ring<float>(const DevCase*, const CaseClock*, const int2*, int, int, int);

// ---------------------------------------------------------------------------
// Two steps of the far field in one pass
//
// An engine of one case may take its steps in pairs (n, n+1): the nodes of
// the window W (every IB node of the run lies at least 3 nodes inside it)
// advance one step at a time with the IB chain between the steps (k_window,
// k_band), the outer ring takes its steps in kernels of its own (k_ring
// beside the pass, step n; k_window_ring after it, step n+1, together with
// the second step of the layer next to the ring), and every other node --
// the far field F -- takes both steps in one pass here, reading each
// population of step n-1 once and writing each of step n+1 once: 144 (FP64)
// or 72 (FP32) bytes per node for two steps.
//
// A workgroup of FAR_LANES lanes owns a strip of output columns [a, b)
// (b - a <= FAR_OUT) and a segment of output rows [j0, j1]; lane L stands at
// column i = a - 1 + L, and the lanes of columns a-1..b take part.  It sweeps
// the rows upwards.  At row r it forms step n at (i, r) -- phase 1: pull from
// plane src, moments, collide, at every node but the ring's -- then, after
// one workgroup barrier, step n+1 at (i, r-1) (phase 2) for the lanes of the
// strip's columns but the ring and the layer next to it (whose pulls would
// reach the ring's step n), pulling step n from the neighbouring lanes.
//
// Step n of a row is kept only as far as phase 2 pulls it across lanes: the
// populations k = 1, 3 (pulled from the same row), 5, 6 (pulled by the row
// above) and 7, 8 (by the row below) go to LDS rings of 3, 4 and 2 rows;
// the lane's own k = 0, 2, 4 stay in its registers (rows r-1, r-2 and r).
// Phase 2 of row q reads k = 7, 8 of row q+1, 1, 3 of row q and 5, 6 of row
// q-1; a wave can be at most one row ahead of another (it waits at the next
// barrier), so the row written next never lands on a slot still being read.
// 36 KB of LDS in FP64, 18 KB in FP32.  With the ring's updates (which pull
// up to three columns or two rows inwards, and fill registers) kept out of
// this kernel, it runs four workgroups per CU in FP64 and eight in FP32.
//
// Every node keeps the single-step arithmetic (the same functions, the same
// operations in the same order), so two steps here equal two single steps
// bit for bit.  Phase 1 skips W's nodes but its outer layer (phase 2 pulls
// only that layer of W); it writes step n of F's layer around W (the
// window's second step pulls it) and of the ring's neighbourhood (columns
// 1..3 and nx-4..nx-2, rows 1, 2 and ny-3, ny-2: the ring's update pulls
// them) into the middle plane, and both phases
// record rho/u/v of their step when it is recorded.  Each of these writes
// comes from the workgroup whose output range holds the node.
//
// psel packs the planes: src = psel & 3, dst = (psel >> 2) & 3, middle =
// (psel >> 4) & 3.
constexpr int FAR_R13 = 3, FAR_R56 = 4, FAR_R78 = 2;
constexpr int FAR_O56 = FAR_R13 * 2 * FAR_LANES;         // element offsets of the rings
constexpr int FAR_O78 = FAR_O56 + FAR_R56 * 2 * FAR_LANES;
constexpr int FAR_NELEM = FAR_O78 + FAR_R78 * 2 * FAR_LANES;
size_t far2_smem_bytes(bool f64) { return (size_t)FAR_NELEM * (f64 ? 8 : 4); }

template <typename T>
__device__ __forceinline__ void far2_body(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                          const int4 tl, int off, int psel)
{
    extern __shared__ __align__(16) unsigned char far_smem[];
    T* lds = reinterpret_cast<T*>(far_smem);
    const int cs = table_case(tl.x);
    const DevCase& D = C[cs];
    const CaseClock kc = K[cs];
    const int n = kc.n0 + off;  // the pair's first step
    if (n + 1 > kc.n_end) return;
    const int src = psel & 3, dst = (psel >> 2) & 3, mid = (psel >> 4) & 3;
    const int nx = D.nx, ny = D.ny;
    const long long N = D.N;
    const int a = tl.y & 0xffff, b = tl.y >> 16;
    const int j0 = tl.z, j1 = tl.w;
    const int L = (int)threadIdx.x;
    const int c0 = a - 1;
    const int i = c0 + L;
    const int wi0 = D.gi0, wi1 = D.gi0 + D.gw - 1, wj0 = D.gj0, wj1 = D.gj0 + D.gh - 1;
    const bool col_in = i >= 0 && i < nx && i <= b;
    // the strip's columns but the outer ones
    const bool own_col = i >= a && i < b && i >= 1 && i <= nx - 2;
    // ... and of those the ones whose step n+1 this pass forms (not the
    // layer next to the ring, which k_window_ring advances)
    const bool own2 = own_col && i >= 2 && i <= nx - 3;
    // columns whose step n the outer columns' update pulls
    const bool ring_nb = (i >= 1 && i <= 3) || (i >= nx - 4 && i <= nx - 2);
    double* mac1 = n >= kc.flag_from ? D.macro[n & 1] : nullptr;
    double* mac2 = n + 1 >= kc.flag_from ? D.macro[(n + 1) & 1] : nullptr;
    GAS T* __restrict__ out2 = pops<T>(D, dst);
    GAS T* __restrict__ outm = pops<T>(D, mid);
    // the lane's own k = 0, 2, 4 of step n: k = 0 and 2 of the row before
    // (h0, h2a), k = 2 of the row before that (h2b), and of the current row
    T h0 = 0, h2a = 0, h2b = 0, c0v = 0, c2v = 0, c4v = 0;

    // phase 2 at output row q for this lane
    auto phase2 = [&](int q) {
        const int s13 = q % FAR_R13, s56 = (q - 1) & (FAR_R56 - 1), s78 = (q + 1) & (FAR_R78 - 1);
        T p[9], out[9];
        p[0] = h0;
        p[1] = lds[(s13 * 2 + 0) * FAR_LANES + L - 1];
        p[2] = h2b;
        p[3] = lds[(s13 * 2 + 1) * FAR_LANES + L + 1];
        p[4] = c4v;
        p[5] = lds[FAR_O56 + (s56 * 2 + 0) * FAR_LANES + L - 1];
        p[6] = lds[FAR_O56 + (s56 * 2 + 1) * FAR_LANES + L + 1];
        p[7] = lds[FAR_O78 + (s78 * 2 + 0) * FAR_LANES + L + 1];
        p[8] = lds[FAR_O78 + (s78 * 2 + 1) * FAR_LANES + L - 1];
        const long long c = (long long)q * nx + i;
        typename NodeMomOf<T>::type m;
        interior_step(D, p, out, m);
#pragma unroll
        for (int k = 0; k < 9; ++k) __builtin_nontemporal_store(out[k], out2 + k * N + c);
        if (mac2) record_macro(D, m, mac2, N, c);
    };

    // Phase 1 of row r at this lane: 0 nothing, 1 an interior node (the ring
    // is not needed: phase 2 never pulls it).  The loads are issued one row
    // ahead, so they overlap the barrier and phase 2 of the row before.
    auto kind1 = [&](int r) -> int {
        if (r < 1 || r > ny - 2 || !col_in || i < 1 || i > nx - 2) return 0;
        const bool in_w = i >= wi0 && i <= wi1 && r >= wj0 && r <= wj1;
        const bool w_edge = in_w && (i == wi0 || i == wi1 || r == wj0 || r == wj1);
        return in_w && !w_edge ? 0 : 1;
    };
    const GAS T* __restrict__ srcp = pops<T>(D, src);
    T pre[9];
    auto load1 = [&](int kind, int r) {
        if (kind == 1) pull_raw(srcp, N, nx, (long long)r * nx + i, pre);
    };
    // phase 2 of row q (2 .. ny-3) needs phase 1 of rows q-1..q+1 (the rows
    // past those are kind 0; the row loop keeps its bounds, which hold the
    // register count down)
    const int rend = min(j1 + 1, ny - 1);
    int kpre = kind1(j0 - 1);
    load1(kpre, j0 - 1);
    for (int r = j0 - 1; r <= rend; ++r) {
        const int kr = kpre;
        T p[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) p[k] = pre[k];
        kpre = r + 1 <= rend ? kind1(r + 1) : 0;
        load1(kpre, r + 1);
        // phase 1: step n at (i, r)
        if (kr != 0) {
            const bool in_w = i >= wi0 && i <= wi1 && r >= wj0 && r <= wj1;
            const long long c = (long long)r * nx + i;
            T st[9];
            typename NodeMomOf<T>::type m;
            interior_step(D, p, st, m);
            const int s13 = r % FAR_R13, s56 = r & (FAR_R56 - 1), s78 = r & (FAR_R78 - 1);
            lds[(s13 * 2 + 0) * FAR_LANES + L] = st[1];
            lds[(s13 * 2 + 1) * FAR_LANES + L] = st[3];
            lds[FAR_O56 + (s56 * 2 + 0) * FAR_LANES + L] = st[5];
            lds[FAR_O56 + (s56 * 2 + 1) * FAR_LANES + L] = st[6];
            lds[FAR_O78 + (s78 * 2 + 0) * FAR_LANES + L] = st[7];
            lds[FAR_O78 + (s78 * 2 + 1) * FAR_LANES + L] = st[8];
            c0v = st[0];
            c2v = st[2];
            c4v = st[4];
            const bool own = own_col && r >= j0 && r <= j1 && !in_w;
            if (own && (ring_nb || r <= 2 || r >= ny - 3 ||
                        (i >= wi0 - 1 && i <= wi1 + 1 && r >= wj0 - 1 && r <= wj1 + 1))) {
#pragma unroll
                for (int k = 0; k < 9; ++k) outm[k * N + c] = st[k];  // step n for the window and the ring
            }
            if (own && mac1) record_macro(D, m, mac1, N, c);
        }
        lds_barrier();  // the next row's pulls stay in flight across it
        // phase 2: step n+1 at row r-1 (the ring and the layer next to it
        // excluded)
        const int q = r - 1;
        if (own2 && q >= max(j0, 2) && q <= min(j1, ny - 3) && !(i >= wi0 && i <= wi1 && q >= wj0 && q <= wj1))
            phase2(q);
        h2b = h2a;
        h2a = c2v;
        h0 = c0v;
    }
}

// The FP32 far pass with two adjacent columns per lane.  A 4-byte
// population per lane and memory instruction moves half the bytes of an FP64
// one per instruction, and an x-shifted pull of 256 bytes crosses one cache
// line more in five than one of 512 bytes in nine: nine FP32 planes pulled
// and nine written stream at 3.92 TB/s one node per lane and at 4.73 TB/s two
// nodes per lane (FP64 one per lane: 4.85 TB/s; a probe of 8192x4096 nodes).
// Lane L stands at columns a-2+2L and a-1+2L (a even, so the pair is 8-byte
// aligned when nx is even; the x-shifted pulls are 4-byte aligned pairs); a
// strip holds at most FAR_OUT2 output columns.  Rows, phases and every
// node's arithmetic are those of far2_body.  Of the populations phase 2 pulls
// across columns, those a lane's other node supplies stay in its registers
// (kept for the rows they are pulled from); only those of the lane's edge
// nodes go to the LDS rings, one float per lane and ring row as in far2_body
// (18 KB).  A pair whose nodes differ in kind (the strip's edges, the outer
// ring, the window's edge) loads and stores node by node.
__device__ __forceinline__ void far2_pairs(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                           const int4 tl, int off, int psel)
{
    typedef float f2u __attribute__((ext_vector_type(2), aligned(4)));
    constexpr int LW = FAR_LANES;  // floats per ring row
    extern __shared__ __align__(16) unsigned char far_smem[];
    float* lds = reinterpret_cast<float*>(far_smem);
    const int cs = table_case(tl.x);
    const DevCase& D = C[cs];
    const CaseClock kc = K[cs];
    const int n = kc.n0 + off;  // the pair's first step
    if (n + 1 > kc.n_end) return;
    const int src = psel & 3, dst = (psel >> 2) & 3, mid = (psel >> 4) & 3;
    const int nx = D.nx, ny = D.ny;
    const long long N = D.N;
    const int a = tl.y & 0xffff, b = tl.y >> 16;
    const int j0 = tl.z, j1 = tl.w;
    const int L = (int)threadIdx.x;
    const int ci = a - 2 + 2 * L;  // the lane's first column
    const int wi0 = D.gi0, wi1 = D.gi0 + D.gw - 1, wj0 = D.gj0, wj1 = D.gj0 + D.gh - 1;
    bool col_in[2], own_col[2], ring_nb[2];
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        const int i = ci + x;
        col_in[x] = i >= 0 && i < nx && i >= a - 1 && i <= b;
        own_col[x] = i >= a && i < b && i >= 1 && i <= nx - 2;
        ring_nb[x] = (i >= 1 && i <= 3) || (i >= nx - 4 && i <= nx - 2);
    }
    GAS double* mac1 = n >= kc.flag_from ? (GAS double*)D.macro[n & 1] : nullptr;
    GAS double* mac2 = n + 1 >= kc.flag_from ? (GAS double*)D.macro[(n + 1) & 1] : nullptr;
    GAS float* __restrict__ out2 = pops<float>(D, dst);
    GAS float* __restrict__ outm = pops<float>(D, mid);
    const GAS float* __restrict__ srcp = pops<float>(D, src);
    // each node's own k = 0, 2, 4 of step n (as in far2_body), and what one
    // node of the lane pulls from the other: node 1's k = 1 and 5 come from
    // node 0 (rows q and q-1: p1, p5a/p5b), node 0's k = 3 and 6 from node 1
    // (p3, p6a/p6b); k = 7 and 8 come from the row phase 1 just formed
    float h0[2] = {0, 0}, h2a[2] = {0, 0}, h2b[2] = {0, 0}, c0v[2] = {0, 0}, c2v[2] = {0, 0}, c4v[2] = {0, 0};
    float p1 = 0, p3 = 0, p5a = 0, p5b = 0, p6a = 0, p6b = 0, c1 = 0, c3 = 0, c5 = 0, c6 = 0, c7 = 0, c8 = 0;

    auto phase2 = [&](int q) {
        const int s13 = q % FAR_R13, s56 = (q - 1) & (FAR_R56 - 1), s78 = (q + 1) & (FAR_R78 - 1);
        const float* r1 = lds + (s13 * 2 + 0) * LW;
        const float* r3 = lds + (s13 * 2 + 1) * LW;
        const float* r5 = lds + FAR_O56 + (s56 * 2 + 0) * LW;
        const float* r6 = lds + FAR_O56 + (s56 * 2 + 1) * LW;
        const float* r7 = lds + FAR_O78 + (s78 * 2 + 0) * LW;
        const float* r8 = lds + FAR_O78 + (s78 * 2 + 1) * LW;
        float out[2][9];
        NodeMom32 m[2];
        {
            float p[9];  // node 0 at column ci
            p[0] = h0[0];
            p[1] = r1[L - 1];
            p[2] = h2b[0];
            p[3] = p3;
            p[4] = c4v[0];
            p[5] = r5[L - 1];
            p[6] = p6b;
            p[7] = c7;
            p[8] = r8[L - 1];
            interior_step(D, p, out[0], m[0]);
        }
        {
            float p[9];  // node 1 at column ci + 1
            p[0] = h0[1];
            p[1] = p1;
            p[2] = h2b[1];
            p[3] = r3[L + 1];
            p[4] = c4v[1];
            p[5] = p5b;
            p[6] = r6[L + 1];
            p[7] = r7[L + 1];
            p[8] = c8;
            interior_step(D, p, out[1], m[1]);
        }
        const long long c = (long long)q * nx + ci;
        const bool in_wq = q >= wj0 && q <= wj1;
        // not the layer next to the ring (k_window_ring advances it)
        const bool ok0 = own_col[0] && ci >= 2 && ci <= nx - 3 && !(in_wq && ci >= wi0 && ci <= wi1);
        const bool ok1 = own_col[1] && ci + 1 >= 2 && ci + 1 <= nx - 3 && !(in_wq && ci + 1 >= wi0 && ci + 1 <= wi1);
        if (ok0 && ok1) {
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                f2u v = {out[0][k], out[1][k]};
                __builtin_nontemporal_store(v, (GAS f2u*)(out2 + k * N + c));
            }
        } else {
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                if (ok0) __builtin_nontemporal_store(out[0][k], out2 + k * N + c);
                if (ok1) __builtin_nontemporal_store(out[1][k], out2 + k * N + c + 1);
            }
        }
        if (mac2) {
            if (ok0) record_macro(D, m[0], mac2, N, c);
            if (ok1) record_macro(D, m[1], mac2, N, c + 1);
        }
    };

    // phase 1 at node x of row r: 0 nothing, 1 an interior node, 2 a ring
    // node, which phase 2 never pulls: it is carried through phase 1 with
    // whatever its registers hold, no load (that form keeps the kernel at
    // five workgroups per CU; zeroing the ring's values costs registers)
    auto kind1 = [&](int x, int r) -> int {
        const int i = ci + x;
        if (r < 0 || r >= ny || !col_in[x]) return 0;
        const bool in_w = i >= wi0 && i <= wi1 && r >= wj0 && r <= wj1;
        const bool w_edge = in_w && (i == wi0 || i == wi1 || r == wj0 || r == wj1);
        if (in_w && !w_edge) return 0;
        return (i == 0 || i == nx - 1 || r == 0 || r == ny - 1) ? 2 : 1;
    };
    float pre[2][9] = {};
    auto load1 = [&](int k0, int k1, int r) {
        const long long c = (long long)r * nx + ci;
        if (k0 == 1 && k1 == 1) {
            const long long off[9] = {0, -1, -(long long)nx, 1, nx, -(long long)nx - 1, -(long long)nx + 1,
                                      (long long)nx + 1, (long long)nx - 1};
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const f2u v = *(const GAS f2u*)(srcp + k * N + c + off[k]);
                pre[0][k] = v.x;
                pre[1][k] = v.y;
            }
        } else {
            if (k0 == 1) pull_raw(srcp, N, nx, c, pre[0]);
            if (k1 == 1) pull_raw(srcp, N, nx, c + 1, pre[1]);
        }
    };
    const int rend = min(j1 + 1, ny - 1);
    int kp0 = kind1(0, j0 - 1), kp1 = kind1(1, j0 - 1);
    load1(kp0, kp1, j0 - 1);
    for (int r = j0 - 1; r <= rend; ++r) {
        const int kr[2] = {kp0, kp1};
        float p[2][9];
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int k = 0; k < 9; ++k) p[x][k] = pre[x][k];
        kp0 = r + 1 <= rend ? kind1(0, r + 1) : 0;
        kp1 = r + 1 <= rend ? kind1(1, r + 1) : 0;
        load1(kp0, kp1, r + 1);
        // phase 1: step n at (ci, r) and (ci + 1, r)
        const int s13 = r % FAR_R13, s56 = r & (FAR_R56 - 1), s78 = r & (FAR_R78 - 1);
        float st[2][9];
        NodeMom32 m[2] = {};
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            if (kr[x] == 2) {
#pragma unroll
                for (int k = 0; k < 9; ++k) st[x][k] = p[x][k];
            } else if (kr[x] == 1) {
                interior_step(D, p[x], st[x], m[x]);
            } else {
#pragma unroll
                for (int k = 0; k < 9; ++k) st[x][k] = 0.0f;
            }
        }
        if (kr[0] != 0 || kr[1] != 0) {
            // the lane's edge nodes to the neighbouring lanes: node 1's k = 1,
            // 5, 8 (pulled by the next lane's node 0), node 0's k = 3, 6, 7
            // (pulled by the previous lane's node 1)
            lds[(s13 * 2 + 0) * LW + L] = st[1][1];
            lds[(s13 * 2 + 1) * LW + L] = st[0][3];
            lds[FAR_O56 + (s56 * 2 + 0) * LW + L] = st[1][5];
            lds[FAR_O56 + (s56 * 2 + 1) * LW + L] = st[0][6];
            lds[FAR_O78 + (s78 * 2 + 0) * LW + L] = st[0][7];
            lds[FAR_O78 + (s78 * 2 + 1) * LW + L] = st[1][8];
            c1 = st[0][1];
            c3 = st[1][3];
            c5 = st[0][5];
            c6 = st[1][6];
            c7 = st[1][7];
            c8 = st[0][8];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                c0v[x] = st[x][0];
                c2v[x] = st[x][2];
                c4v[x] = st[x][4];
                const int i = ci + x;
                const bool in_w = i >= wi0 && i <= wi1 && r >= wj0 && r <= wj1;
                const bool own = own_col[x] && r >= j0 && r <= j1 && !in_w && kr[x] == 1;
                const long long c = (long long)r * nx + i;
                if (own && (ring_nb[x] || r <= 2 || r >= ny - 3 ||
                            (i >= wi0 - 1 && i <= wi1 + 1 && r >= wj0 - 1 && r <= wj1 + 1))) {
#pragma unroll
                    for (int k = 0; k < 9; ++k) outm[k * N + c] = st[x][k];  // step n for the window and the ring
                }
                if (own && mac1) record_macro(D, m[x], mac1, N, c);
            }
        }
        lds_barrier();  // the next row's pulls stay in flight across it
        // phase 2: step n+1 at row r-1 (the ring and the layer next to it
        // excluded)
        const int q = r - 1;
        if ((own_col[0] || own_col[1]) && q >= max(j0, 2) && q <= min(j1, ny - 3)) phase2(q);
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            h2b[x] = h2a[x];
            h2a[x] = c2v[x];
            h0[x] = c0v[x];
        }
        p1 = c1;
        p3 = c3;
        p5b = p5a;
        p5a = c5;
        p6b = p6a;
        p6a = c6;
    }
}

// One far tile per workgroup, or (a grid smaller than the tile table: the
// launch shape's persistent grid) tiles v, v + gridDim.x, .. in turn, with a
// barrier between two tiles so that no wave writes the LDS rings of the next
// tile while another still reads those of the last.
template <typename T>
__global__ __launch_bounds__(FAR_LANES) void k_far2(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                                    const int4* __restrict__ tiles, int ntiles, int off, int psel)
{
    for (int v = blockIdx.x; v < ntiles; v += gridDim.x) {
        if (v != (int)blockIdx.x) lds_barrier();
        const int4 tl = tiles[v];
        if constexpr (sizeof(T) == 4) far2_pairs(C, K, tl, off, psel);
        else far2_body<T>(C, K, tl, off, psel);
    }
}
template __global__ void k_far2<double>(const DevCase*, const CaseClock*, const int4*, int, int, int);
template __global__ void k_far2<float>(const DevCase*, const CaseClock*, const int4*, int, int, int);

int lattice_tiles(int nx, int ny, int npt)
{
    const long long span = (long long)(ny - 1) * nx - lattice_base(nx);  // through the last interior row
    return lattice_ring_blocks(nx, ny) + (int)((span + 256 * npt - 1) / (256 * npt));
}

// Fills n values of p with v (grid-stride; set-up of planes and tables).
template <typename T>
__global__ __launch_bounds__(256) void k_fill(T* __restrict__ p, long long n, T v)
{
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        p[t] = v;
}
template __global__ void k_fill<double>(double*, long long, double);
template __global__ void k_fill<float>(float*, long long, float);
template __global__ void k_fill<int>(int*, long long, int);

// Box search of move_boundary (airfoil.f90:290-321) for one axis: the 1-based
// index range of nodes with |(o + (i-1)*sc - X)/sc| <= 2.5.
__device__ __forceinline__ void axis_box(double X, double o, double sc, int n, int& lo, int& hi)
{
    int i1 = (int)((X - o) / sc + 1.0);
    for (int i = i1; i >= 1; --i) {
        double temp = o + (double)(i - 1) * sc;
        if (fabs((temp - X) / sc) > 2.5) { i1 = i + 1; break; }
    }
    int i2 = (int)((X - o) / sc + 1.0);
    for (int i = i2; i <= n; ++i) {
        double temp = o + (double)(i - 1) * sc;
        if (fabs((temp - X) / sc) > 2.5) { i2 = i - 1; break; }
    }
    lo = i1;
    hi = i2;
}

// ---------------------------------------------------------------------------
// immersed boundary
//
// Geometry of a batch of steps (one slot per step; grid z, or y for the
// one-workgroup-per-slot kernels, is the case), built ahead of time on its
// own stream because it depends only on the prescribed kinematics:
//   k_geo_clear -> k_geo_markers -> k_geo_fill -> k_geo_phi -> k_geo_pairs ->
//   k_geo_dsort -> k_geo_nsort -> k_geo_assemble -> k_geo_factor
// Per-step solve, on the critical path beside the lattice kernel:
//   k_crit_ustar -> k_crit_rhs -> k_crit_solve -> k_band (beside the lattice kernel)
// Per-wing forces of a batch of steps: k_forces.
// ---------------------------------------------------------------------------


__device__ __forceinline__ bool box_cell(const int* b, int q, int& i, int& j)
{
    int nyb = b[3] - b[2] + 1;
    int n = (b[1] - b[0] + 1) * nyb;
    if (nyb <= 0 || q >= n) return false;
    i = b[0] + q / nyb;  // i-major, j-minor: the reference's loop order (airfoil.f90:322-323)
    j = b[2] + q % nyb;
    return true;
}

// The geometry batch that a launch at window offset `off` builds for a case:
// GEO_FIRST builds the batch holding the step a fresh case takes at offset 0,
// GEO_NEXT the batch after the one holding the step at this offset (at
// batch starts, or at offset 0 for a fresh case).  Block s of the
// batch works on step n0b + s (1-based); returns false when there is none.
__device__ __forceinline__ bool geo_step(const DevCase& D, const CaseClock& k, int off, int mode, int s,
                                         int& step0)
{
    const int n = k.n0 + off;
    // a fresh case (first step of a run or of a restart) builds at offset 0
    // whatever its position inside the batch, and so does a case whose next
    // batch was left unbuilt at the end of its last window; the others at
    // batch starts
    const bool first = k.fresh && off == 0;
    const bool start = (k.fresh || k.next_due) && off == 0;
    if (n > k.n_end || ((n - 1) % D.batch != 0 && !start)) return false;
    int j = (n - 1) / D.batch;
    if (mode == GEO_FIRST) {
        if (!first) return false;
    } else {
        j += 1;
    }
    const int n0b = j * D.batch + 1;
    const int nb = min(D.batch, D.steps_cap - n0b + 1);
    if (s >= nb) return false;
    step0 = n0b - 1 + s;  // 0-based step index
    return true;
}

// window index of 1-based lattice node (i, j)
__device__ __forceinline__ long long wcell(const DevCase& D, int i, int j)
{
    return (long long)(j - 1 - D.gj0) * D.gw + (i - 1 - D.gi0);
}

// Clear the slot's previous node set (airfoil.f90:276-283).
__global__ __launch_bounds__(256) void k_geo_clear(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                                   int off, int mode)
{
    const DevCase& D = C[blockIdx.z];
    int step;
    if (!geo_step(D, K[blockIdx.z], off, mode, blockIdx.y, step)) return;
    const Slot S = slot_of(D, step % D.nslot);
    const int na = *S.nadj;
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < na; a += gridDim.x * blockDim.x) {
        long long c = (long long)(S.adj[3 * a + 1] - D.gj0) * D.gw + (S.adj[3 * a] - D.gi0);
        S.claim[c] = INT_MAX;
        S.nodeidx[c] = -1;
    }
}

// Markers, marker velocities and boxes (move_boundary airfoil.f90:256-321,
// boundary_condition.f90:1542-1553), then the first claim of every box node:
// the smallest segment index wins, which is the segment that claims it first
// in the reference's serial loop.  sin/cos of the pitch angle come
// precomputed from the host's libm.
__global__ __launch_bounds__(256) void k_geo_markers(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                                     int off, int mode)
{
    const DevCase& D = C[blockIdx.z];
    int step;
    if (!geo_step(D, K[blockIdx.z], off, mode, blockIdx.y, step)) return;
    const Slot S = slot_of(D, step % D.nslot);
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.nseg) return;
    int w = s / D.spw;
    int p = w * D.ppw + (s - w * D.spw);
    const double* kn = D.kin + ((long long)step * D.nw + w) * KIN;
    double sn = kn[0], ca = kn[1], om = kn[2], sa = kn[3], sau = kn[4], sb = kn[5], sbv = kn[6];
    double xa = D.xb0[p], ya = D.yb0[p], xb = D.xb0[p + 1], yb = D.yb0[p + 1];
    // body offsets of the dynamics (zero without): body_extra_s/(1/scale) and
    // the domain shift movenum*mesh_max_scale (airfoil.f90:261-262)
    const double es = D.body[4 * w], mn = D.body[4 * w + 1];
    double x1a = ca * xa + sn * ya + sa + D.xc + es - mn;
    double x1b = ca * xb + sn * yb + sa + D.xc + es - mn;
    double y1a = -sn * xa + ca * ya + sb + D.yc;
    double y1b = -sn * xb + ca * yb + sb + D.yc;
    double X = 0.5 * (x1b + x1a);
    double Y = 0.5 * (y1b + y1a);
    S.xb2[s] = X;
    S.yb2[s] = Y;
    double xs = xb + xa, ys = yb + ya;
    double U = D.U;
    S.ub[s] = U * (-om * sn * 0.5 * xs + om * ca * 0.5 * ys + sau) + D.body[4 * w + 2];
    S.vb[s] = U * (-om * ca * 0.5 * xs - om * sn * 0.5 * ys + sbv);
    int i1, i2, j1, j2;
    axis_box(X, D.ox, D.scale, D.nx, i1, i2);
    axis_box(Y, D.oy, D.scale, D.ny, j1, j2);
    // the box (0-based nodes i1-1 .. i2-1) lies at least 3 nodes inside the
    // lattice window, so the IB kernels (the box's nodes and their pull
    // neighbours) never touch the window's outer two layers, which the far
    // pass reads
    bool ok = i1 >= 4 && i2 <= D.nx - 3 && j1 >= 4 && j2 <= D.ny - 3 && i2 >= i1 && j2 >= j1 &&
              (i2 - i1) < 6 && (j2 - j1) < 6 && i1 - 1 >= D.gi0 + 3 && i2 - 1 < D.gi0 + D.gw - 3 &&
              j1 - 1 >= D.gj0 + 3 && j2 - 1 < D.gj0 + D.gh - 3;
    if (!ok) {
        atomicOr(D.err, ERR_GEOMETRY);
        i1 = D.gi0 + 2; i2 = D.gi0 + 1; j1 = D.gj0 + 2; j2 = D.gj0 + 1;  // empty box
    }
    int* b = S.box + 4 * s;
    b[0] = i1; b[1] = i2; b[2] = j1; b[3] = j2;
    for (int i = i1; i <= i2; ++i)
        for (int j = j1; j <= j2; ++j) atomicMin(&S.claim[wcell(D, i, j)], s);
}

// Ordered compaction: the node list in the reference's order (segment-major,
// then i, then j; airfoil.f90:322-331), the node -> list index grid, the node
// count, and cleared marker-list counters.  One workgroup per slot.
__global__ __launch_bounds__(1024) void k_geo_fill(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                                  int off, int mode)
{
    const DevCase& D = C[blockIdx.y];
    int step;
    if (!geo_step(D, K[blockIdx.y], off, mode, blockIdx.x, step)) return;
    const Slot S = slot_of(D, step % D.nslot);
    __shared__ int part[16];  // wavefront totals (1024 threads)
    __shared__ int s_total;
    const int ns = D.nseg, T = blockDim.x, tid = threadIdx.x;
    const int chunk = (ns + T - 1) / T;
    const int s0 = tid * chunk, s1 = min(ns, s0 + chunk);
    int sum = 0;
    for (int s = s0; s < s1; ++s) {
        const int* b = S.box + 4 * s;
        int cl[BOXMAX];
#pragma unroll
        for (int q = 0; q < BOXMAX; ++q) {
            int i, j;
            cl[q] = box_cell(b, q, i, j) ? S.claim[wcell(D, i, j)] : -1;
        }
#pragma unroll
        for (int q = 0; q < BOXMAX; ++q) sum += cl[q] == s;
    }
    // inclusive scan of the counts: within each wavefront by shuffles, then
    // over the wavefront totals (one barrier instead of two per doubling)
    const int lane = tid & 63, wv = tid >> 6;
    int v = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    if (lane == 63) part[wv] = v;
    __syncthreads();
    for (int q = 0; q < wv; ++q) v += part[q];
    int base = v - sum;
    if (tid == T - 1) s_total = v;
    for (int s = s0; s < s1; ++s) {
        const int* b = S.box + 4 * s;
        int w = s / D.spw;
        int cl[BOXMAX];
#pragma unroll
        for (int q = 0; q < BOXMAX; ++q) {
            int i, j;
            cl[q] = box_cell(b, q, i, j) ? S.claim[wcell(D, i, j)] : -1;
        }
#pragma unroll
        for (int q = 0; q < BOXMAX; ++q) {
            int i, j;
            if (cl[q] == s && box_cell(b, q, i, j)) {
                if (base < D.cap) {
                    S.adj[3 * base] = i - 1;
                    S.adj[3 * base + 1] = j - 1;
                    S.adj[3 * base + 2] = w;
                    S.nodeidx[wcell(D, i, j)] = base;
                }
                base++;
            }
        }
    }
    __syncthreads();
    const int total = s_total;
    if (tid == 0) {
        if (total > D.cap) atomicOr(D.err, ERR_CAPACITY);
        *S.nadj = min(total, D.cap);
    }
    for (int a = tid; a < min(total, D.cap); a += T) S.nmkcnt[a] = 0;
}

// One-axis kernel values of every box column (q < 6) and row (q >= 6)
// (boundary_condition.f90:1581-1600).
__global__ __launch_bounds__(256) void k_geo_phi(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                              int off, int mode)
{
    const DevCase& D = C[blockIdx.z];
    int step;
    if (!geo_step(D, K[blockIdx.z], off, mode, blockIdx.y, step)) return;
    const Slot S = slot_of(D, step % D.nslot);
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int s = t / 12, q = t % 12;
    if (s >= D.nseg) return;
    const int* b = S.box + 4 * s;
    const double sc = D.scale;
    double v = 0.0;
    if (q < 6) {
        if (q <= b[1] - b[0]) {
            double x0 = D.ox + sc * (double)(b[0] + q - 1);
            v = delta1(fabs((x0 - S.xb2[s]) / sc), D.smooth);
        }
    } else if (q - 6 <= b[3] - b[2]) {
        double y0 = D.oy + sc * (double)(b[2] + q - 6 - 1);
        v = delta1(fabs((y0 - S.yb2[s]) / sc), D.smooth);
    }
    S.phi[t] = v;
}

// Entry xx1(s, node) = kernel1*kernel2 (boundary_condition.f90:1602) of every
// box node, in box order, and registration of s in the node's marker list.
__global__ __launch_bounds__(256) void k_geo_pairs(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                              int off, int mode)
{
    const DevCase& D = C[blockIdx.z];
    int step;
    if (!geo_step(D, K[blockIdx.z], off, mode, blockIdx.y, step)) return;
    const Slot S = slot_of(D, step % D.nslot);
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int s = t / BOXMAX, q = t % BOXMAX;
    if (s >= D.nseg) return;
    const int* b = S.box + 4 * s;
    int i, j;
    if (!box_cell(b, q, i, j)) return;
    int nyb = b[3] - b[2] + 1;
    double w = S.phi[s * 12 + q / nyb] * S.phi[s * 12 + 6 + q % nyb];
    int node = S.nodeidx[wcell(D, i, j)];
    S.draw_idx[t] = node;
    S.draw_val[t] = w;
    if (node < 0) { atomicOr(D.err, ERR_CAPACITY); return; }
    int slot = atomicAdd(&S.nmkcnt[node], 1);
    if (slot < NCAP) {
        S.nraw_s[(long long)node * NCAP + slot] = s;
        S.nraw_w[(long long)node * NCAP + slot] = w;
    } else {
        atomicOr(D.err, ERR_CAPACITY);
    }
}

// A row of xx1 in ascending node order: each entry moves to its rank.
__global__ __launch_bounds__(256) void k_geo_dsort(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                              int off, int mode)
{
    const DevCase& D = C[blockIdx.z];
    int step;
    if (!geo_step(D, K[blockIdx.z], off, mode, blockIdx.y, step)) return;
    const Slot S = slot_of(D, step % D.nslot);
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int s = t / BOXMAX, q = t % BOXMAX;
    if (s >= D.nseg) return;
    const int* b = S.box + 4 * s;
    int cnt = max(0, (b[1] - b[0] + 1) * (b[3] - b[2] + 1));
    if (q == 0) S.dcnt[s] = cnt;
    if (q >= cnt) return;
    const int* ix = S.draw_idx + s * BOXMAX;
    int me = ix[q], rank = 0;
    for (int r = 0; r < cnt; ++r) rank += ix[r] < me;
    S.didx[s * BOXMAX + rank] = me < 0 ? 0 : me;
    S.dval[s * BOXMAX + rank] = S.draw_val[t];
}

// Each node's marker list in ascending segment order (the order of the
// spread sum); grid-stride over (node, entry).
__global__ __launch_bounds__(256) void k_geo_nsort(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                              int off, int mode)
{
    const DevCase& D = C[blockIdx.z];
    int step;
    if (!geo_step(D, K[blockIdx.z], off, mode, blockIdx.y, step)) return;
    const Slot S = slot_of(D, step % D.nslot);
    const long long total = (long long)(*S.nadj) * NCAP;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        int a = (int)(t / NCAP), q = (int)(t % NCAP);
        int cnt = min(S.nmkcnt[a], NCAP);
        if (q >= cnt) continue;
        const int* ks = S.nraw_s + (long long)a * NCAP;
        int me = ks[q], rank = 0;
        for (int r = 0; r < cnt; ++r) rank += ks[r] < me;
        S.nmk[(long long)a * NCAP + rank] = me;
        S.nmkv[(long long)a * NCAP + rank] = S.nraw_w[t];
    }
}

// Band of G = D D^T in zig-zag order: G(p, p-d) is the sum over the nodes
// shared by the boxes of segments perm[p] and perm[p-d], ascending list index.
__global__ __launch_bounds__(256) void k_geo_assemble(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                              int off, int mode)
{
    const DevCase& D = C[blockIdx.z];
    int step;
    if (!geo_step(D, K[blockIdx.z], off, mode, blockIdx.y, step)) return;
    const Slot S = slot_of(D, step % D.nslot);
    const int W = D.bw + 1;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)D.nseg * W) return;
    int p = (int)(t / W), d = (int)(t % W), q = p - d;
    double sum = 0.0;
    if (q >= 0) {
        int k = D.perm[p], l = D.perm[q];
        const int* bk = S.box + 4 * k;
        const int* bl = S.box + 4 * l;
        if (!(bk[1] < bl[0] || bl[1] < bk[0] || bk[3] < bl[2] || bl[3] < bk[2])) {
            const int* ik = S.didx + k * BOXMAX;
            const int* il = S.didx + l * BOXMAX;
            const double* vk = S.dval + k * BOXMAX;
            const double* vl = S.dval + l * BOXMAX;
            int nk = S.dcnt[k], nl = S.dcnt[l], a = 0, c = 0;
            while (a < nk && c < nl) {
                int xa = ik[a], xc = il[c];
                if (xa < xc) a++;
                else if (xa > xc) c++;
                else { sum = sum + vk[a] * vl[c]; a++; c++; }
            }
        }
    }
    S.G[t] = sum;
}

// Value of v in lane `lane` of the wavefront (wave-uniform result).
__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)b, lane);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

constexpr int WIN = 64;  // LDS window slots of the banded factorisation (rows k..k+bw, bw <= WIN-2)
constexpr int CH = 32;   // rows / columns staged per chunk
constexpr int PAIRMAX = MAXBW * (MAXBW + 1) / 2;

// Banded L D L^T of an n x n SPD band of lower bandwidth bw, right-looking
// over a sliding LDS window of rows k..k+bw; ga(r, d) reads entry
// a(r, r-d).  Step k takes the pivot d_k = a(k,k), forms the column
// l(i,k) = t_i * (1/d_k) from its unscaled entries t_i and updates the
// trailing band a(i,j) -= l(i,k) * t_j.  Each entry receives its updates in
// increasing k, so the result equals the dense right-looking algorithm bit
// for bit.  Rows entering the window are staged through LDS one chunk ahead;
// results leave through plain stores that are never waited on inside the
// loop.  Besides 1/d_k the kernel writes the coefficients of step k of both
// triangular solves in the per-lane order of k_crit_solve (see there), the
// forward step at timeline position k, the backward one at 2*np-1-k:
//   forward  [h] = {l(k+e1,k), l(k+e2,k)}, e1 = h-q (h > q), e2 = ws+h-q (h <= q)
//   backward [h] = {l(k,k-e1), l(k,k-e2)}, e1 = q-h (h < q), e2 = ws+q-h (h >= q)
// with q = k mod ws, zero outside the band.  In the compact layout
// (compact, bw < CW) each timeline row holds CW values instead, entry e the
// coefficient of the row e away from the pivot: forward l(k+e,k), backward
// l(k,k-e), entry 0 and entries past the band zero.  With ltail, row k's
// final band (d_k, l(k,k-1), .., l(k,k-bw)) of the last bw rows is kept as
// well.  One workgroup of 256 threads.
template <typename GA>
__device__ __forceinline__ void band_factor(const GA& ga, int n, int bw, int ws, int np, double* __restrict__ coef,
                                            double* __restrict__ dinv, double* __restrict__ ltail, int* err,
                                            bool compact)
{
    extern __shared__ double smem[];
    const int tid = threadIdx.x, T = blockDim.x;
    const int W = bw + 1;
    double* win = smem;
    double* colv = win + WIN * WIN;
    double* colt = colv + WIN;
    double* stage = colt + WIN;  // [2][CH][WIN]
    double* piv = stage + 2 * CH * WIN;  // [2] d_k and [2] 1/d_k, alternating steps
    int* pairs = (int*)(piv + 4);
    for (int a = tid; a < bw; a += T)
        for (int b = 0; b <= a; ++b) pairs[a * (a + 1) / 2 + b] = (a << 8) | b;
    for (int e = tid; e < min(W, n) * W; e += T) {
        int r = e / W, d = e % W;
        win[(r % WIN) * WIN + d] = ga(r, d);
    }
    // chunk c: rows W + c*CH .. W + c*CH + CH - 1, entering at steps c*CH .. c*CH + CH - 1
    constexpr int PER = (CH * WIN + 255) / 256;
    double pre[PER];
    auto load_chunk = [&](int c) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            int e = tid + q * T;
            int r = W + c * CH + e / W, d = e % W;
            pre[q] = (e < CH * W && r < n) ? ga(r, d) : 0.0;
        }
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            int e = tid + q * T;
            if (e < CH * W) stage[buf * CH * WIN + (e / W) * WIN + e % W] = pre[q];
        }
    };
    load_chunk(0);
    store_chunk(0);
    load_chunk(1);
    __syncthreads();
    // each thread's update pairs (a, b), fixed for the whole factorisation
    constexpr int PP = (PAIRMAX + 255) / 256;
    int myab[PP];
#pragma unroll
    for (int r = 0; r < PP; ++r) {
        const int t = tid + r * T;
        myab[r] = t < bw * (bw + 1) / 2 ? pairs[t] : 0;
    }
    // the pivot of step k and its reciprocal are taken at the end of step
    // k-1 by the thread that gives the pivot its last update
    if (tid == 0 && n > 0) {
        piv[0] = win[0];
        piv[2] = 1.0 / piv[0];
    }
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        const int c = k / CH;
        if (k > 0 && k % CH == 0) {
            store_chunk(c & 1);  // chunk c, prefetched during chunk c-1
            load_chunk(c + 1);
        }
        const int m = min(k + bw, n - 1) - k;
        const int sk = (k % WIN) * WIN;
        const double akk = piv[k & 1];
        const double rk = piv[2 + (k & 1)];
        if (tid < m) {
            int i = k + 1 + tid;
            int e = (i % WIN) * WIN + (i - k);
            double t = win[e];
            double l = t * rk;
            win[e] = l;
            colv[tid] = l;
            colt[tid] = t;
        }
        lds_barrier();
        if (tid == 0 && !(akk > 0.0)) atomicOr(err, ERR_SOLVE);
        const int npairs = m * (m + 1) / 2;
#pragma unroll
        for (int r = 0; r < PP; ++r) {
            const int t = tid + r * T;
            if (t >= npairs) break;
            const int ab = myab[r];
            const int a = ab >> 8, b = ab & 255;
            const int i = k + 1 + a, j = k + 1 + b;
            const int e = (i % WIN) * WIN + (i - j);
            const double v = win[e] - colv[a] * colt[b];
            win[e] = v;
            if (t == 0) {  // a(k+1, k+1) is final: the next pivot
                piv[(k + 1) & 1] = v;
                piv[2 + ((k + 1) & 1)] = 1.0 / v;
            }
        }
        // solve coefficients of step k: column k (final now) and row k (final
        // since step k-1)
        if (compact) {
            // row k (forward) and row 2*np-1-k (backward) of CW values: entry
            // e is the coefficient of the row e away from the pivot
            if (tid < CW) {
                const int e = tid;
                coef[(long long)k * CW + e] = (e >= 1 && e <= m) ? colv[e - 1] : 0.0;
            } else if (tid < 2 * CW) {
                const int e = tid - CW;
                coef[(long long)(2 * np - 1 - k) * CW + e] = (e >= 1 && e <= min(bw, k)) ? win[sk + e] : 0.0;
            }
            if (tid == 0) dinv[k] = rk;
            if (ltail && k >= n - bw)
                for (int d = tid; d < W; d += T)
                    ltail[(long long)(k - (n - bw)) * W + d] = d == 0 ? akk : (d <= k ? win[sk + d] : 0.0);
        } else {
            const int q = k & (ws - 1);
            const int hh = (tid >> 1) & (ws - 1), j = tid & 1;
            if (tid < 2 * ws) {
                const int e = j == 0 ? hh - q : ws + hh - q;
                const bool ok = (j == 0 ? hh > q : hh <= q) && e <= m;
                coef[(long long)k * 2 * ws + tid] = ok ? colv[e - 1] : 0.0;
            } else if (tid < 4 * ws) {
                const int e = j == 0 ? q - hh : ws + q - hh;
                const bool ok = (j == 0 ? hh < q : hh >= q) && e <= min(bw, k);
                coef[(long long)(2 * np - 1 - k) * 2 * ws + tid - 2 * ws] = ok ? win[sk + e] : 0.0;
            }
            if (tid == 0) dinv[k] = rk;
            if (ltail && k >= n - bw)
                for (int d = tid; d < W; d += T)
                    ltail[(long long)(k - (n - bw)) * W + d] = d == 0 ? akk : (d <= k ? win[sk + d] : 0.0);
        }
        // row k+W enters the window (its slot differs from every row in use)
        const int rn = k + W;
        if (rn < n)
            for (int d = tid; d < W; d += T)
                win[(rn % WIN) * WIN + d] = stage[(c & 1) * CH * WIN + (k - c * CH) * WIN + d];
        lds_barrier();
    }
}

// Solve coefficients of step k (column k and row k final, both in the LDS
// window win of band_factor_blocked) in band_factor's layouts; u is the
// entry (0 .. 4*ws-1, or 0 .. 2*CW-1 compact).
__device__ __forceinline__ void factor_coefs(const double* __restrict__ win, int k, int n, int bw, int ws, int np,
                                             GAS double* __restrict__ coef, bool compact, int u)
{
    const int m = min(k + bw, n - 1) - k;
    const int sk = (k % WIN) * WIN;
    auto fwd = [&](int e) { return win[((k + e) % WIN) * WIN + e]; };  // l(k+e, k), 1 <= e <= m
    if (compact) {
        if (u < CW) {
            const int e = u;
            coef[(long long)k * CW + e] = (e >= 1 && e <= m) ? fwd(e) : 0.0;
        } else if (u < 2 * CW) {
            const int e = u - CW;
            coef[(long long)(2 * np - 1 - k) * CW + e] = (e >= 1 && e <= min(bw, k)) ? win[sk + e] : 0.0;
        }
    } else {
        const int q = k & (ws - 1);
        const int hh = (u >> 1) & (ws - 1), j = u & 1;
        if (u < 2 * ws) {
            const int e = j == 0 ? hh - q : ws + hh - q;
            const bool ok = (j == 0 ? hh > q : hh <= q) && e <= m;
            coef[(long long)k * 2 * ws + u] = ok ? fwd(e) : 0.0;
        } else if (u < 4 * ws) {
            const int e = j == 0 ? q - hh : ws + q - hh;
            const bool ok = (j == 0 ? hh < q : hh >= q) && e <= min(bw, k);
            coef[(long long)(2 * np - 1 - k) * 2 * ws + u - 2 * ws] = ok ? win[sk + e] : 0.0;
        }
    }
}

// Blocked form of band_factor (bw <= WIN - 2*FNB): the same
// right-looking L D L^T with every entry receiving the same updates in
// increasing k, in blocks of FNB columns k0 .. k0+FNB-1.
//  * Panel (wavefront 0): lane a holds row k0+a (rows k0 .. k0+FNB-1+bw)
//    with the block's columns in registers (register c = column k0+c) and
//    factors the block column by column: the pivot and the unscaled entries
//    of other rows come from their lanes by readlane, the updates among the
//    block's own columns stay in registers.  It leaves the block's l and t
//    in LDS (two buffers, by block parity), 1/d_k in dinv and, in the
//    compact layout, the forward coefficients of the block's columns.
//  * Trailing update of block b, split by column: T1 (the FNB columns of
//    the next panel) by every wavefront right after the panel, T2 (the
//    columns after those) by wavefronts 1-3 beside the next panel, together
//    with block b's backward coefficients.  Each pair (a, c) (row
//    k0+FNB+a >= column k0+FNB+c) applies the block's FNB updates in
//    increasing k.  An entry of column j receives T2(b-1) before T1(b)
//    (a barrier apart), and the panel of block b after T2(b-2) and T1(b-1).
//  * Rows enter the window one group of FNB rows per block (wavefront 3):
//    group g (rows g*FNB+bw .. g*FNB+FNB-1+bw, first read by block g) is
//    loaded into registers during block g-3 and written to the window
//    after the panel of block g-1, into slots of rows retired (backward
//    coefficients written) before.
// Each wavefront's memory counters hold one kind of traffic: the panel's
// LDS waits never wait for global loads or stores, and the group loads are
// unconditional, so wavefront 3 waits only for the groups it writes.  Two
// barriers per block instead of two per column, and the panel's critical
// chain (the pivot's reciprocal and the next pivot's update) shares its
// phase only with T2 and the coefficients.
constexpr int FNB = 8;
constexpr int FGQ = (FNB * (WIN - 2 * FNB + 1) + 63) / 64;
constexpr int FBW = WIN - 2 * FNB;                                   // largest bw of the blocked form
constexpr int FPP1 = (FNB * FBW + 255) / 256;                        // T1 pairs per thread (256 threads)
constexpr int FPP2 = ((FBW - FNB) * (FBW - FNB + 1) / 2 + 191) / 192;  // T2 pairs per thread (192 threads)
template <typename GA>
__device__ __forceinline__ void band_factor_blocked(const GA& ga, int n, int bw, int ws, int np,
                                                    GAS double* __restrict__ coef, GAS double* __restrict__ dinv,
                                                    GAS double* __restrict__ ltail, int* err, bool compact)
{
    extern __shared__ double smem[];
    const int tid = threadIdx.x, T = blockDim.x;
    const int W = bw + 1;
    double* win = smem;                 // [WIN][WIN]: row r in slot r % WIN, entry d = a(r, r-d)
    double* lbuf = win + WIN * WIN;     // [2][l, t][FNB][WIN]: l(k0+a, k0+c) and t(k0+a, k0+c) at c*WIN + a
    int* pairs1 = (int*)(lbuf + 4 * FNB * WIN);
    int* pairs2 = pairs1 + FNB * FBW;
    // T1 pairs (c < FNB) ordered by c then a, T2 pairs (c >= FNB) likewise
    const int c1 = min(FNB, bw);
    const int np1 = c1 * bw - c1 * (c1 - 1) / 2;
    const int np2 = bw > FNB ? (bw - FNB) * (bw - FNB + 1) / 2 : 0;
    for (int a = tid; a < bw; a += T)
        for (int c = 0; c <= a; ++c) {
            if (c < FNB) pairs1[c * bw - c * (c - 1) / 2 + a - c] = (a << 8) | c;
            else pairs2[(c - FNB) * bw - (c * (c - 1) / 2 - FNB * (FNB - 1) / 2) + a - c] = (a << 8) | c;
        }
    for (int e = tid; e < min(FNB + bw, n) * W; e += T) {
        const int r = e / W, d = e % W;
        win[(r % WIN) * WIN + d] = ga(r, d);
    }
    const int lt = tid - 192;  // lane of wavefront 3
    auto load_group = [&](int g, double (&p)[FGQ]) {
#pragma unroll
        for (int q = 0; q < FGQ; ++q) {
            const int e = lt + q * 64;
            const int r = g * FNB + bw + e / W, d = e % W;
            const bool ok = e < FNB * W && r < n;
            p[q] = ga(ok ? r : 0, ok ? d : 0);  // every lane loads (exact wait counts); store_group skips !ok
        }
    };
    auto store_group = [&](int g, const double (&p)[FGQ]) {
#pragma unroll
        for (int q = 0; q < FGQ; ++q) {
            const int e = lt + q * 64;
            const int r = g * FNB + bw + e / W, d = e % W;
            if (e < FNB * W && r < n) win[(r % WIN) * WIN + d] = p[q];
        }
    };
    double pa[FGQ], pb[FGQ];
    if (lt >= 0) {
        load_group(1, pa);
        load_group(2, pb);
    }
    __syncthreads();
    int my1[FPP1], my2[FPP2];
#pragma unroll
    for (int r = 0; r < FPP1; ++r) {
        const int t = tid + r * T;
        my1[r] = t < np1 ? pairs1[t] : 0;
    }
#pragma unroll
    for (int r = 0; r < FPP2; ++r) {
        const int t = tid - 64 + r * 192;
        my2[r] = (tid >= 64 && t < np2) ? pairs2[t] : 0;
    }
    const int ncoef = compact ? 2 * CW : 4 * ws;
    // pair pr = (a, c) of the trailing update of the block at k0
    auto trail = [&](int k0, int nb, const double* lc, int pr) {
        const int a = pr >> 8, c = pr & 255;
        const int i = k0 + FNB + a, j = k0 + FNB + c;
        if (i >= n) return;
        const double* tc = lc + FNB * WIN;
        const int e = (i % WIN) * WIN + (i - j);
        double v = win[e];
#pragma unroll
        for (int q = 0; q < FNB; ++q)
            if (q < nb && i - k0 - q <= bw) v = v - lc[q * WIN + (i - k0)] * tc[q * WIN + (j - k0)];
        win[e] = v;
    };
    // backward (and, in the ws layout, forward) coefficients of the block
    // at k0, from the window; u0 .. ncoef-1 in steps of du
    auto coefs = [&](int k0, int nb, int u0, int du) {
        if (compact) {
            // entry e of row k: l(k, k-e), 1 <= e <= min(bw, k)
            for (int u = u0; u < FNB * CW; u += du) {
                const int c = u / CW, e = u % CW, k = k0 + c;
                if (c < nb)
                    coef[(long long)(2 * np - 1 - k) * CW + e] =
                        (e >= 1 && e <= min(bw, k)) ? win[(k % WIN) * WIN + e] : 0.0;
            }
        } else {
            for (int c = 0; c < nb; ++c)
                for (int u = u0; u < ncoef; u += du) factor_coefs(win, k0 + c, n, bw, ws, np, coef, compact, u);
        }
        // the final band (d_k, l(k,k-1), .., l(k,k-bw)) of the last bw rows
        if (ltail && k0 + nb > n - bw)
            for (int u = u0; u < FNB * W; u += du) {
                const int c = u / W, d = u % W, k = k0 + c;
                if (c < nb && k >= n - bw) {
                    const int sk = (k % WIN) * WIN;
                    ltail[(long long)(k - (n - bw)) * W + d] = d == 0 ? win[sk] : (d <= k ? win[sk + d] : 0.0);
                }
            }
    };
    auto block = [&](int b, double (&p)[FGQ]) {
        const int k0 = b * FNB, nb = min(FNB, n - k0);
        double* lc = lbuf + (b & 1) * 2 * FNB * WIN;
        if (tid < 64) {
            const int a = tid, i = k0 + a;
            const bool live = a < FNB + bw && i < n;
            const int si = (i % WIN) * WIN;
            double P[FNB], Tt[FNB], R[FNB];
#pragma unroll
            for (int c = 0; c < FNB; ++c) {
                const int dd = a - c;
                P[c] = (live && dd >= 0 && dd <= bw) ? win[si + dd] : 0.0;
                Tt[c] = 0.0;
                R[c] = 0.0;
            }
            bool bad = false;
            // column k0+c of the panel: selects, no branches, so that the
            // updates of later columns fill the latency of the next pivot's
            // reciprocal
            auto column = [&](int c) {
                const double d = readlane_f64(P[c], c);
                const double r = 1.0 / d;
                R[c] = r;
                bad = bad || !(d > 0.0);
                const bool below = live && a - c >= 1 && a - c <= bw;
                const double t = P[c], l = t * r;
                P[c] = below ? l : t;
                Tt[c] = below ? t : 0.0;
#pragma unroll
                for (int c2 = c + 1; c2 < FNB; ++c2) {
                    const double tj = readlane_f64(t, c2);
                    const double u = P[c2] - l * tj;
                    P[c2] = (below && a >= c2 && c2 - c <= bw) ? u : P[c2];
                }
            };
            if (nb == FNB) {
#pragma unroll
                for (int c = 0; c < FNB; ++c) column(c);
            } else {
#pragma unroll
                for (int c = 0; c < FNB; ++c)
                    if (c < nb) column(c);
            }
            if (a < nb) {
                double ra = R[0];
#pragma unroll
                for (int c = 1; c < FNB; ++c) ra = a == c ? R[c] : ra;
                dinv[k0 + a] = ra;
            }
            if (bad && a == 0) atomicOr(err, ERR_SOLVE);
#pragma unroll
            for (int c = 0; c < FNB; ++c) {
                const int dd = a - c;
                if (live && dd >= 0 && dd <= bw) win[si + dd] = P[c];
                lc[c * WIN + a] = P[c];
                lc[FNB * WIN + c * WIN + a] = Tt[c];
                // compact forward coefficients of column k0+c from the lanes
                // holding it: entry e = a-c is l(k0+c+e, k0+c), 1 <= e <= m
                if (compact && c < nb && dd >= 0 && dd < CW) {
                    const int k = k0 + c, m = min(k + bw, n - 1) - k;
                    coef[(long long)k * CW + dd] = (dd >= 1 && dd <= m) ? P[c] : 0.0;
                }
            }
        } else if (b > 0) {
            // beside the panel: T2 of the block before (a full block) and its
            // coefficients
            const int kp = k0 - FNB;
            const double* lp = lbuf + ((b - 1) & 1) * 2 * FNB * WIN;
#pragma unroll
            for (int r = 0; r < FPP2; ++r) {
                const int t = tid - 64 + r * 192;
                if (t >= np2) break;
                trail(kp, FNB, lp, my2[r]);
            }
            coefs(kp, FNB, tid - 64, 192);
        }
        lds_barrier();
#pragma unroll
        for (int r = 0; r < FPP1; ++r) {
            const int t = tid + r * T;
            if (t >= np1) break;
            trail(k0, nb, lc, my1[r]);
        }
        if (lt >= 0) {
            // group b+1 is first read by the panel of block b+1
            store_group(b + 1, p);
            load_group(b + 3, p);
        }
        lds_barrier();
    };
    const int nblk = (n + FNB - 1) / FNB;
    for (int b = 0; b < nblk; b += 2) {
        block(b, pa);
        if (b + 1 < nblk) block(b + 1, pb);
    }
    // the last block's coefficients (its T2 reaches no row below n)
    if (nblk > 0) coefs((nblk - 1) * FNB, n - (nblk - 1) * FNB, tid, T);
}

// L D L^T of the whole zig-zag band G (cases solved as one chain).
__global__ __launch_bounds__(256) void k_geo_factor(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                                    int off, int mode)
{
    const DevCase& D = C[blockIdx.y];
    if (D.dissect) return;
    int step;
    if (!geo_step(D, K[blockIdx.y], off, mode, blockIdx.x, step)) return;
    const Slot S = slot_of(D, step % D.nslot);
    const double* G = S.G;
    const int W = D.bw + 1;
    auto ga = [G, W](int r, int d) { return G[(long long)r * W + d]; };
    if (D.bw <= WIN - 2 * FNB) {
        // global (not flat) accesses: flat ones would count in the LDS wait counter too
        const GAS double* Gg = (const GAS double*)G;
        auto gg = [Gg, W](int r, int d) { return Gg[(long long)r * W + d]; };
        band_factor_blocked(gg, D.nseg, D.bw, D.ws, D.npad, (GAS double*)S.coef, (GAS double*)S.dinv, nullptr,
                            D.err, D.compact != 0);
    } else {
        band_factor(ga, D.nseg, D.bw, D.ws, D.npad, S.coef, S.dinv, nullptr, D.err, D.compact != 0);
    }
}

// Separator form of the solve.  With bw the band's lower bandwidth, rows
// mT .. mT+bw-1 (the separator S) split G into a top half T (rows 0..mT-1)
// and a bottom half B (rows mT+bw .. nseg-1) that share no entry.  B is
// taken in reverse row order, so that both halves meet S with their last bw
// rows.  Each half is factored here (blockIdx.z: 0 = T, 1 = B reversed) as
// its own band; k_geo_schur then forms the separator's Schur complement.
__global__ __launch_bounds__(256) void k_geo_factor2(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                                     int off, int mode)
{
    const DevCase& D = C[blockIdx.y];
    if (!D.dissect) return;
    int step;
    if (!geo_step(D, K[blockIdx.y], off, mode, blockIdx.x, step)) return;
    const Slot S = slot_of(D, step % D.nslot);
    const double* G = S.G;
    const int W = D.bw + 1, ns = D.nseg, half = blockIdx.z;
    double* coef = S.coef2 + (long long)half * (2 * D.np2 + 256) * CW;
    double* dinv = S.dinv2 + (long long)half * D.np2;
    double* lt = S.ltail2 + (long long)half * D.bw * W;
    // the blocked form (global, not flat, accesses) where the band allows it
    const GAS double* Gg = (const GAS double*)G;
    const bool blk = D.bw <= FBW;
    if (half == 0) {
        auto ga = [Gg, W](int r, int d) { return Gg[(long long)r * W + d]; };
        if (blk) band_factor_blocked(ga, D.mT, D.bw, D.ws, D.np2, (GAS double*)coef, (GAS double*)dinv,
                                     (GAS double*)lt, D.err, true);
        else band_factor(ga, D.mT, D.bw, D.ws, D.np2, coef, dinv, lt, D.err, true);
    } else {
        // reversed row r is row ns-1-r; its entry a'(r, r-d) = a(ns-1-r+d, ns-1-r)
        auto ga = [Gg, W, ns](int r, int d) { return d <= r ? Gg[(long long)(ns - 1 - r + d) * W + d] : 0.0; };
        if (blk) band_factor_blocked(ga, D.mB, D.bw, D.ws, D.np2, (GAS double*)coef, (GAS double*)dinv,
                                     (GAS double*)lt, D.err, true);
        else band_factor(ga, D.mB, D.bw, D.ws, D.np2, coef, dinv, lt, D.err, true);
    }
}

// Schur complement of the separator, S' = A_S - sum over the halves of
// Y^T D^-1 Y with Y = L^-1 C (L, D the half's factor, C its coupling to S,
// nonzero only in the half's last bw rows, so Y is the bw x bw block
// Lt^-1 Ct of the factor's last bw rows), S'^-1 by Gauss-Jordan (S' is SPD),
// and the separator matrices of the solve: with v = [f_S; w_T; w_B] (w the
// halves' forward results D^-1 L^-1 f over their last bw rows),
//   x_S = M v,                 M = [S'^-1, -S'^-1 Y_T^T, -S'^-1 Y_B^T]
//   [w_T; w_B] -= N x_S,       N = [D_T Y_T; D_B Y_B]  (D_* the tail's 1/d)
// One workgroup of 256 threads per slot.  Every entry is formed by the
// same operations in the same order whichever thread takes it.
__global__ __launch_bounds__(256) void k_geo_schur(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                                   int off, int mode)
{
    const DevCase& D = C[blockIdx.y];
    if (!D.dissect) return;
    int step;
    if (!geo_step(D, K[blockIdx.y], off, mode, blockIdx.x, step)) return;
    const Slot S = slot_of(D, step % D.nslot);
    constexpr int MB = SEPMAXBW, T = 256;
    __shared__ double lt[2][MB][MB + 1], Y[2][MB][MB + 1], Sm[MB][MB + 1], Iv[MB][MB + 1];
    __shared__ double fcol[MB], dts[2][MB];
    const int b = D.bw, W = b + 1, mT = D.mT, mB = D.mB, np2 = D.np2, b3 = 3 * b;
    const int tid = threadIdx.x;
    const double* G = S.G;
    for (int e = tid; e < 2 * b * W; e += T) {
        const int h = e / (b * W), r = e % (b * W);
        lt[h][r / W][r % W] = S.ltail2[e];
    }
    // the tails' 1/d: half 0 rows mT-b.., half 1 rows mB-b.. of its own factor
    for (int e = tid; e < 2 * b; e += T) {
        const int h = e / b, i = e % b;
        dts[h][i] = h == 0 ? S.dinv2[mT - b + i] : S.dinv2[np2 + mB - b + i];
    }
    __syncthreads();
    if (tid < 2 * b) {
        // column s of half h: Y = Lt^-1 Ct by forward substitution
        const int h = tid / b, s = tid % b;
        for (int i = 0; i < b; ++i) {
            // coupling of the half's row (last bw rows, i) to separator row s
            double cv;
            if (h == 0) {
                const int d = b + s - i;  // a(mT+s, mT-b+i)
                cv = d <= b ? G[(long long)(mT + s) * W + d] : 0.0;
            } else {
                const int d = 2 * b - 1 - i - s;  // a(mT+2b-1-i, mT+s)
                cv = d <= b ? G[(long long)(mT + 2 * b - 1 - i) * W + d] : 0.0;
            }
            for (int j = 0; j < i; ++j) cv = cv - lt[h][i][i - j] * Y[h][j][s];
            Y[h][i][s] = cv;
        }
    }
    __syncthreads();
    for (int e = tid; e < b * b; e += T) {
        const int s = e / b, t = e % b;
        const int hi = max(s, t), lo = min(s, t);
        double v = G[(long long)(mT + hi) * W + (hi - lo)];
        for (int i = 0; i < b; ++i) v = v - Y[0][i][s] * dts[0][i] * Y[0][i][t];
        for (int i = 0; i < b; ++i) v = v - Y[1][i][s] * dts[1][i] * Y[1][i][t];
        Sm[s][t] = v;
        Iv[s][t] = s == t ? 1.0 : 0.0;
    }
    // S'^-1 by Gauss-Jordan without pivoting (S' is SPD)
    for (int k = 0; k < b; ++k) {
        __syncthreads();
        const double piv = Sm[k][k];
        if (tid == 0 && !(piv > 0.0)) atomicOr(D.err, ERR_SOLVE);
        const double p = 1.0 / piv;
        __syncthreads();
        if (tid < b) {
            Sm[k][tid] = Sm[k][tid] * p;
            Iv[k][tid] = Iv[k][tid] * p;
        } else if (tid >= 64 && tid < 64 + b && tid - 64 != k) {
            fcol[tid - 64] = Sm[tid - 64][k];  // column k below and above row k
        }
        __syncthreads();
        for (int e = tid; e < 2 * b * b; e += T) {
            const int m = e / (b * b), r = e % (b * b), i = r / b, s = r % b;
            if (i == k) continue;
            if (m == 0) Sm[i][s] = Sm[i][s] - fcol[i] * Sm[k][s];
            else Iv[i][s] = Iv[i][s] - fcol[i] * Iv[k][s];
        }
    }
    __syncthreads();
    double* Z = S.zsep;
    for (int e = tid; e < b * b3; e += T) {
        const int s = e / b3, c = e % b3;
        double v;
        if (c < b) {
            v = Iv[s][c];
        } else {
            const int h = c < 2 * b ? 0 : 1, i = c - (h + 1) * b;
            v = 0.0;
            for (int t = 0; t < b; ++t) v = v + Iv[s][t] * Y[h][i][t];
            v = -v;
        }
        Z[e] = v;
    }
    for (int e = tid; e < 2 * b * b; e += T) {
        const int h = e / (b * b), r = e % (b * b), s = r / b, t = r % b;
        Z[(long long)b * b3 + e] = dts[h][s] * Y[h][s][t];
    }
}

// u* at the step's IB nodes: the moments of the nodes' post-stream
// populations (physical_variables, solution.f90:504-545), once per node, so
// that the interpolation below reads each node's value instead of pulling
// the node again for every marker box that holds it (about 11 boxes per
// node at the standard marker spacing).
template <typename T>
__global__ __launch_bounds__(256) void k_crit_ustar(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                                    int off, int psel)
{
    __builtin_amdgcn_s_setprio(3);  // critical path: ahead of co-resident lattice waves
    const DevCase& D = C[blockIdx.y];
    const int nstep = K[blockIdx.y].n0 + off;
    if (nstep > K[blockIdx.y].n_end || (int)blockIdx.x * 256 >= D.cap) return;
    const Slot S = slot_of(D, (nstep - 1) % D.nslot);
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (a >= *S.nadj) return;
    const long long c = (long long)S.adj[3 * a + 1] * D.nx + S.adj[3 * a];
    double f[9], r, u, v;
    int src, dst;
    step_planes(D, nstep, psel, src, dst);
    pull9(pops<T>(D, src), D.N, D.nx, c, f, D.W);
    moments(f, r, u, v);
    S.ust[a] = u;
    S.vst[a] = v;
}
template __global__ void k_crit_ustar<double>(const DevCase*, const CaseClock*, int, int);
template __global__ void k_crit_ustar<float>(const DevCase*, const CaseClock*, int, int);

// Interpolation D u* (the DGEMVs, boundary_condition.f90:1620-1621) and the
// right-hand side U_B - D u* (1624-1629) in zig-zag order: one thread per
// marker segment sums xx1(s,node) * u*(node) over its box's nodes in
// ascending node order, each product rounded before it is added.  This
// two-kernel form does a third of the memory work of k_crit_rhs1 and is
// the batched form's; a single case takes k_crit_rhs1, whose one pass has
// the shorter dependency chain.
__global__ __launch_bounds__(256) void k_crit_rhs(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                                  int off)
{
    __builtin_amdgcn_s_setprio(3);
    const DevCase& D = C[blockIdx.y];
    const int nstep = K[blockIdx.y].n0 + off;
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (nstep > K[blockIdx.y].n_end || s >= D.nseg) return;
    const Slot S = slot_of(D, (nstep - 1) % D.nslot);
    const int n = S.dcnt[s];
    const int* __restrict__ ix = S.didx + s * BOXMAX;
    const double* __restrict__ w = S.dval + s * BOXMAX;
    double su = 0.0, sv = 0.0;
    for (int q = 0; q < n; ++q) {
        const int node = ix[q];
        const double pu = w[q] * S.ust[node];
        const double pv = w[q] * S.vst[node];
        su = su + pu;
        sv = sv + pv;
    }
    const int p = D.iperm[s];
    D.rhs[p] = S.ub[s] - su;
    D.rhs[D.nseg + p] = S.vb[s] - sv;
}

// The same right-hand side in one pass, for an engine of one case, whose IB
// chain is latency-bound: each thread forms one product xx1(s,node) *
// u*(node) with u* taken from the moments of the node's post-stream
// populations, pulled by the thread itself; the segment's products are then
// summed in ascending node order.  Bit-identical to k_crit_ustar +
// k_crit_rhs (the same products, summed in the same order).
constexpr int RHS_SEG = 7;  // segments per 256-thread block (7 * 36 = 252)
template <typename T>
__global__ __launch_bounds__(256) void k_crit_rhs1(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                                                  int off, int psel)
{
    __builtin_amdgcn_s_setprio(3);  // critical path: ahead of co-resident lattice waves
    const DevCase& D = C[blockIdx.y];
    const int nstep = K[blockIdx.y].n0 + off;
    if (nstep > K[blockIdx.y].n_end || (int)blockIdx.x * RHS_SEG >= D.nseg) return;
    int src, dst;
    step_planes(D, nstep, psel, src, dst);
    const Slot S = slot_of(D, (nstep - 1) % D.nslot);
    __shared__ double pu[RHS_SEG][BOXMAX], pv[RHS_SEG][BOXMAX];
    const int sl = threadIdx.x / BOXMAX, q = threadIdx.x % BOXMAX;
    const int s = blockIdx.x * RHS_SEG + sl;
    const bool live = sl < RHS_SEG && s < D.nseg;
    const int n = live ? S.dcnt[s] : 0;
    if (live && q < n) {
        int node = S.didx[s * BOXMAX + q];
        double w = S.dval[s * BOXMAX + q];
        long long c = (long long)S.adj[3 * node + 1] * D.nx + S.adj[3 * node];
        double f[9], r, u, v;
        pull9(pops<T>(D, src), D.N, D.nx, c, f, D.W);
        moments(f, r, u, v);
        pu[sl][q] = w * u;
        pv[sl][q] = w * v;
    }
    __syncthreads();
    if (live && q == 0) {
        double su = 0.0, sv = 0.0;
        for (int a = 0; a < n; ++a) { su = su + pu[sl][a]; sv = sv + pv[sl][a]; }
        int p = D.iperm[s];
        D.rhs[p] = S.ub[s] - su;
        D.rhs[D.nseg + p] = S.vb[s] - sv;
    }
}
template __global__ void k_crit_rhs1<double>(const DevCase*, const CaseClock*, int, int);
template __global__ void k_crit_rhs1<float>(const DevCase*, const CaseClock*, int, int);

// Forward L z = b, scaling x = D^-1 z and backward L^T y = x for both
// right-hand sides.  Rows are dealt to the lanes of a wavefront in chunks of
// 64: lane h keeps the right-hand side(s) of row c*64 + h of the active chunk
// c ("cur") and of row (c+1)*64 + h (forward) / (c-1)*64 + h (backward) of
// its neighbour ("nxt"), which covers every row within bw <= 64 of the pivot.
// Step k (q = k mod 64) takes the finished value of row k from lane q by
// readlane and every lane applies
//   cur -= c1 * z_k,   nxt -= c2 * z_k
// with this lane's two coefficients of step k (zero outside the band).  A zero coefficient leaves a finite value unchanged, so every
// row receives exactly the oracle's updates in increasing (forward) /
// decreasing (backward) k, and the step has no lane-dependent control flow.
//
// The 2*npad steps form one timeline (forward steps, then backward steps);
// their coefficients lie in that order in coef, 1 KB per step as per-lane
// pairs or 256 B per step in the compact layout.  In k_crit_solve wavefronts
// 0 and 1 run the steps of right-hand sides 0 and 1 (the two chains share the
// coefficients and nothing else, and one wave issues one FP64 operation at a
// time, so two waves halve the instructions each chain issues per step).  A
// chain wave touches only LDS and registers: a global load in its
// instruction stream would wait for the CU's vector-memory queue, which the
// lattice kernel keeps full, and stall the whole dependency chain.  The
// SV_NPROD producer wavefronts stream units of U steps from HBM into a
// double-buffered LDS ring of per-lane pairs, loading each unit two units
// before it is used; one workgroup barrier per unit hands the buffers over.
constexpr int SV_NPROD = 8;  // producer wavefronts
typedef double sv_d2 __attribute__((ext_vector_type(2)));  // register-friendly pair
constexpr int SV_P = 4;      // steps of coefficients read ahead from LDS, divides U

// FMA: each update is one fused multiply-add (one rounding) instead of a
// product and a difference; the separator form, whose operation order differs
// from the oracle's anyway, takes it for the shorter dependency chain.
template <bool FWD, int U, bool FMA = false, int NR = 2>
__device__ __forceinline__ void sv_unit(const double2* __restrict__ rb, int lane, int i0, double (&cur)[NR],
                                        double (&nxt)[NR])
{
    double2 cf[SV_P];
#pragma unroll
    for (int i = 0; i < SV_P; ++i) cf[i] = rb[i * 64 + lane];
#pragma unroll
    for (int i = 0; i < U; ++i) {
        const double2 c = cf[i % SV_P];
        if (i + SV_P < U) cf[i % SV_P] = rb[(i + SV_P) * 64 + lane];
        const int q = FWD ? i0 + i : 63 - i0 - i;
        double z[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) z[r] = readlane_f64(cur[r], q);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if constexpr (FMA) {
                cur[r] = __builtin_fma(-c.x, z[r], cur[r]);
                nxt[r] = __builtin_fma(-c.y, z[r], nxt[r]);
            } else {
                cur[r] = cur[r] - c.x * z[r];
                nxt[r] = nxt[r] - c.y * z[r];
            }
        }
        // keep this step's updates in this step
        if constexpr (NR == 2) asm volatile("" : "+v"(cur[0]), "+v"(cur[1]), "+v"(nxt[0]), "+v"(nxt[1]));
        else asm volatile("" : "+v"(cur[0]), "+v"(nxt[0]));
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Per-lane coefficient pair (c1, c2) of timeline step t for this lane, from
// the compact row of step t (band_factor's compact layout, np the chain's
// padded length): the inverse of that layout, so the pair equals the one the
// per-lane layout stores.
__device__ __forceinline__ int compact_entry(int t, int np, int lane)
{
    const bool fwd = t < np;
    const int q = fwd ? (t & 63) : 63 - (t & 63);
    const int e = (fwd ? lane - q : q - lane) & 63;
    return e < CW ? e : 0;  // entry 0 is always zero
}
__device__ __forceinline__ sv_d2 compact_pair(double v, int t, int np, int lane)
{
    const bool fwd = t < np;
    const int q = fwd ? (t & 63) : 63 - (t & 63);
    const bool act = fwd ? lane > q : lane < q;  // the row lies in the active chunk
    sv_d2 r;
    r.x = act ? v : 0.0;
    r.y = act ? 0.0 : v;
    return r;
}

// CMP: the coefficients are in the compact layout (DevCase::compact); the
// producers fetch those CW values per step (a quarter of the per-lane
// layout's bytes) and expand them into the per-lane LDS ring, so the chain's
// instruction stream is the same for both layouts.
template <int U, bool CMP>
__global__ __launch_bounds__(64 * (2 + SV_NPROD)) void k_crit_solve(const DevCase* __restrict__ C,
                                                                     const CaseClock* __restrict__ K, int off)
{
    constexpr int UPC = 64 / U;              // units per lane chunk
    constexpr int SPW = U / SV_NPROD;        // steps per producer and unit
    __shared__ double2 ring[2 * U * 64];     // two units
    extern __shared__ double zb[];           // [2][npad] values, then [npad] 1/d
    const DevCase& D = C[blockIdx.x];
    if (D.dissect || (D.compact != 0) != CMP) return;
    const int nstep = K[blockIdx.x].n0 + off;
    if (nstep > K[blockIdx.x].n_end) return;
    const Slot S = slot_of(D, (nstep - 1) % D.nslot);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int n = D.nseg, np = D.npad, nch = np / 64, nun = 2 * nch * UPC;
    double* dv = zb + 2 * np;
    for (int e = tid; e < 2 * np; e += blockDim.x) {
        const int r = e >= np, p = e - r * np;
        zb[e] = p < n ? D.rhs[r * n + p] : 0.0;
    }
    for (int p = tid; p < np; p += blockDim.x) dv[p] = S.dinv[p];
    if (wave > 1) {
        // ---- producer: unit u, steps u*U + w*SPW + s ----
        const int w = wave - 2;
        sv_d2* rg = (sv_d2*)ring;
        typedef typename std::conditional<CMP, double, sv_d2>::type Val;
        Val A[SPW], B[SPW];
        // units past the end reload the last one and are never read
        auto load = [&](Val (&R)[SPW], int u) {
            const long long u_ = min(u, nun - 1);
#pragma unroll
            for (int q = 0; q < SPW; ++q) {
                const int t = (int)(u_ * U) + w * SPW + q;
                if constexpr (CMP)
                    R[q] = S.coef[(long long)t * CW + compact_entry(t, np, lane)];
                else
                    R[q] = ((const sv_d2*)S.coef)[(long long)t * 64 + lane];
            }
        };
        auto store = [&](const Val (&R)[SPW], int u) {
#pragma unroll
            for (int q = 0; q < SPW; ++q) {
                const int t = u * U + w * SPW + q;
                sv_d2 v;
                if constexpr (CMP) v = compact_pair(R[q], t, np, lane);
                else v = R[q];
                rg[(u & 1) * U * 64 + (w * SPW + q) * 64 + lane] = v;
            }
        };
        load(A, 0);
        load(B, 1);
        store(A, 0);
        load(A, 2);
        __syncthreads();
        for (int u = 0; u < nun; u += 2) {
            store(B, u + 1);
            load(B, u + 3);
            __syncthreads();
            store(A, u + 2);
            load(A, u + 4);
            __syncthreads();
        }
    } else {
        // ---- the chain of right-hand side r = wave ----
        const int r = wave;
        double* zr = zb + r * np;
        __syncthreads();
        double cur[1], nxt[1];
        cur[0] = zr[lane];
        nxt[0] = nch > 1 ? zr[64 + lane] : 0.0;
        // forward, then x = z * (1/d)
        for (int c = 0; c < nch; ++c) {
#pragma unroll
            for (int j = 0; j < UPC; ++j) {
                const int u = c * UPC + j;
                sv_unit<true, U, false, 1>(ring + (u & 1) * U * 64, lane, j * U, cur, nxt);
                if (j == UPC - 1) {
                    const int row = c * 64 + lane;
                    zr[row] = cur[0] * dv[row];
                    cur[0] = nxt[0];
                    nxt[0] = c + 2 < nch ? zr[row + 128] : 0.0;
                }
                __syncthreads();
            }
        }
        // backward; lane h starts on rows (nch-1)*64 + h and (nch-2)*64 + h,
        // which it finished itself
        cur[0] = zr[(nch - 1) * 64 + lane];
        nxt[0] = nch > 1 ? zr[(nch - 2) * 64 + lane] : 0.0;
        for (int c = nch - 1; c >= 0; --c) {
#pragma unroll
            for (int j = 0; j < UPC; ++j) {
                const int u = (2 * nch - 1 - c) * UPC + j;
                sv_unit<false, U, false, 1>(ring + (u & 1) * U * 64, lane, j * U, cur, nxt);
                if (j == UPC - 1) {
                    const int row = c * 64 + lane;
                    zr[row] = cur[0];
                    cur[0] = nxt[0];
                    nxt[0] = c >= 2 ? zr[row - 128] : 0.0;
                }
                __syncthreads();
            }
        }
    }
    __syncthreads();
    for (int p = tid; p < n; p += blockDim.x) {
        const int s = D.perm[p];
        D.y[s] = zb[p];
        D.y[n + s] = zb[np + p];
    }
}
template __global__ void k_crit_solve<64, false>(const DevCase*, const CaseClock*, int);
template __global__ void k_crit_solve<32, false>(const DevCase*, const CaseClock*, int);
template __global__ void k_crit_solve<64, true>(const DevCase*, const CaseClock*, int);
template __global__ void k_crit_solve<32, true>(const DevCase*, const CaseClock*, int);

// Separator form of the solve (DevCase::dissect; factors from k_geo_factor2
// and k_geo_schur; compact coefficients).  Wavefronts 0 and 2 run the top
// half's chains (right-hand sides 0 and 1), wavefronts 1 and 3 the reversed
// bottom half's, each as in k_crit_solve
// (per-lane LDS rings that the producer wavefronts fill from the compact
// rows) but with fused multiply-adds, so the critical path is one half's
// length at one operation per update.  Between the forward and
// backward sweeps the separator step (matrices from k_geo_schur) forms
// x_S = M v from v = [f_S; both halves' forward results over their last bw
// rows], then the halves' tail corrections N x_S.  The 512 producer threads,
// idle there, hold their pieces of M and N in registers from the start and
// write partial sums into the free half of a coefficient ring; the chain
// threads add them in fixed order.  The backward sweeps then give x_T and x_B.
template <int U>
__global__ __launch_bounds__(64 * (4 + SV_NPROD)) void k_crit_solve2(const DevCase* __restrict__ C,
                                                                      const CaseClock* __restrict__ K, int off)
{
    constexpr int UPC = 64 / U;              // units per lane chunk
    constexpr int SPW = U / SV_NPROD;        // steps per producer, half and unit
    __shared__ double2 ring[2][2 * U * 64];  // per half, two units
    extern __shared__ double zb[];           // [2 halves][2 rhs][np2], [2][np2] 1/d, f_S, x_S
    const DevCase& D = C[blockIdx.x];
    if (!D.dissect) return;
    const int nstep = K[blockIdx.x].n0 + off;
    if (nstep > K[blockIdx.x].n_end) return;
    const Slot S = slot_of(D, (nstep - 1) % D.nslot);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int ns = D.nseg, b = D.bw, mT = D.mT, mB = D.mB, np = D.np2, nch = np / 64;
    const int nun = 2 * nch * UPC, F = nch * UPC;
    double* dv = zb + 4 * np;     // [2][np]
    double* fs = zb + 6 * np;     // [2 rhs][64] separator right-hand sides
    double* xs = fs + 128;        // [2 rhs][64] x_S
    for (int e = tid; e < 4 * np; e += blockDim.x) {
        const int h = e / (2 * np), r = (e / np) & 1, p = e % np;
        double v = 0.0;
        if (h == 0 && p < mT) v = D.rhs[r * ns + p];
        if (h == 1 && p < mB) v = D.rhs[r * ns + ns - 1 - p];
        zb[e] = v;
    }
    for (int e = tid; e < 2 * np; e += blockDim.x) dv[e] = S.dinv2[e];
    for (int e = tid; e < 2 * b; e += blockDim.x) fs[(e / b) * 64 + e % b] = D.rhs[(e / b) * ns + mT + e % b];
    // partial sums of the separator step go to the half of ring 0 that
    // neither the last forward unit nor the first backward unit uses
    double* part = (double*)(ring[0] + ((F + 1) & 1) * U * 64);
    const long long cstride = (long long)(2 * np + 256) * CW;
    const int b3 = 3 * b, n1 = sep_seg1(b), L1 = sep_len1(b), n2 = sep_seg2(b), L2 = sep_len2(b);
    if (wave >= 4) {
        const int w = wave - 4, pt = tid - 256;
        // pieces of M (row o1, terms q1*L1 ..) and N (row o2, terms q2*L2 ..)
        const int o1 = pt / n1, q1 = pt % n1, o2 = pt / n2, q2 = pt % n2;
        const bool a1 = o1 < b, a2 = o2 < 2 * b;
        double m1[SEP_L1], m2[SEP_L2];
#pragma unroll
        for (int j = 0; j < SEP_L1; ++j) {
            const int c = q1 * L1 + j;
            m1[j] = (a1 && j < L1 && c < b3) ? S.zsep[(long long)o1 * b3 + c] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < SEP_L2; ++j) {
            const int t = q2 * L2 + j;
            m2[j] = (a2 && j < L2 && t < b) ? S.zsep[(long long)b * b3 + (long long)o2 * b + t] : 0.0;
        }
        auto mid = [&]() {
            if (a1) {
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    double acc = 0.0;
#pragma unroll
                    for (int j = 0; j < SEP_L1; ++j) {
                        const int c = q1 * L1 + j;
                        if (j < L1 && c < b3) {
                            const double v = c < b ? fs[r * 64 + c]
                                           : c < 2 * b ? zb[r * np + mT - b + (c - b)]
                                                       : zb[2 * np + r * np + mB - b + (c - 2 * b)];
                            acc = acc + m1[j] * v;
                        }
                    }
                    part[(r * b + o1) * n1 + q1] = acc;
                }
            }
            __syncthreads();
            __syncthreads();
            if (a2) {
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    double acc = 0.0;
#pragma unroll
                    for (int j = 0; j < SEP_L2; ++j) {
                        const int t = q2 * L2 + j;
                        if (j < L2 && t < b) acc = acc + m2[j] * xs[r * 64 + t];
                    }
                    part[1024 + (r * 2 * b + o2) * n2 + q2] = acc;
                }
            }
            __syncthreads();
            __syncthreads();
        };
        // ---- producer: unit u, steps u*U + w*SPW + q of both halves ----
        const double* cf[2] = {S.coef2, S.coef2 + cstride};
        // unit v sits in register set v % NB from its load, NB units ahead
        // of its use, until it is stored to the ring just before barrier v
        constexpr int NB = 4;
        double R[NB][2][SPW];
        auto load = [&](double (&R)[2][SPW], int u) {
            const long long u_ = min(u, nun - 1);  // units past the end reload the last one, never read
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int q = 0; q < SPW; ++q) {
                    const int t = (int)(u_ * U) + w * SPW + q;
                    R[h][q] = cf[h][(long long)t * CW + compact_entry(t, np, lane)];
                }
        };
        auto store = [&](const double (&R)[2][SPW], int u) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                sv_d2* rg = (sv_d2*)ring[h];
#pragma unroll
                for (int q = 0; q < SPW; ++q) {
                    const int t = u * U + w * SPW + q;
                    rg[(u & 1) * U * 64 + (w * SPW + q) * 64 + lane] = compact_pair(R[h][q], t, np, lane);
                }
            }
        };
#pragma unroll
        for (int k = 0; k < NB; ++k) load(R[k], k);
        store(R[0], 0);
        load(R[0], NB);
        __syncthreads();
        // barriers 1 .. nun, as the chains take them (unit nun is a copy of
        // the last one and never read)
        for (int u = 1; u <= nun; u += NB) {
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int v = u + k;
                if (v <= nun) {
                    store(R[(1 + k) % NB], v);
                    load(R[(1 + k) % NB], v + NB);
                    __syncthreads();
                    if (v == F) mid();
                }
            }
        }
    } else {
        // ---- the four chains; wave w runs half w & 1, right-hand side w >> 1 ----
        const int h = wave & 1;
        double* z = zb + 2 * h * np + (wave >> 1) * np;
        const double* di = dv + h * np;
        const double2* rb = ring[h];
        __syncthreads();
        double cur[1], nxt[1];
        cur[0] = z[lane];
        nxt[0] = nch > 1 ? z[64 + lane] : 0.0;
        for (int c = 0; c < nch; ++c) {
#pragma unroll
            for (int j = 0; j < UPC; ++j) {
                const int u = c * UPC + j;
                sv_unit<true, U, true, 1>(rb + (u & 1) * U * 64, lane, j * U, cur, nxt);
                if (j == UPC - 1) {
                    const int row = c * 64 + lane;
                    z[row] = cur[0] * di[row];
                    cur[0] = nxt[0];
                    nxt[0] = c + 2 < nch ? z[row + 128] : 0.0;
                }
                __syncthreads();
            }
        }
        // ---- separator step, chain side (four barriers shared with the producers) ----
        __syncthreads();
        if (tid < 2 * b) {
            const int r = tid / b, o = tid % b;
            double sum = 0.0;
            for (int q = 0; q < n1; ++q) sum = sum + part[(r * b + o) * n1 + q];
            xs[r * 64 + o] = sum;
        }
        __syncthreads();
        __syncthreads();
        if (tid < 4 * b) {
            const int r = tid / (2 * b), o = tid % (2 * b);
            double sum = 0.0;
            for (int q = 0; q < n2; ++q) sum = sum + part[1024 + (r * 2 * b + o) * n2 + q];
            if (o < b) zb[r * np + mT - b + o] -= sum;
            else zb[2 * np + r * np + mB - b + (o - b)] -= sum;
        }
        __syncthreads();
        // backward; lane l starts on rows (nch-1)*64 + l and (nch-2)*64 + l
        cur[0] = z[(nch - 1) * 64 + lane];
        nxt[0] = nch > 1 ? z[(nch - 2) * 64 + lane] : 0.0;
        for (int c = nch - 1; c >= 0; --c) {
#pragma unroll
            for (int j = 0; j < UPC; ++j) {
                const int u = (2 * nch - 1 - c) * UPC + j;
                sv_unit<false, U, true, 1>(rb + (u & 1) * U * 64, lane, j * U, cur, nxt);
                if (j == UPC - 1) {
                    const int row = c * 64 + lane;
                    z[row] = cur[0];
                    cur[0] = nxt[0];
                    nxt[0] = c >= 2 ? z[row - 128] : 0.0;
                }
                __syncthreads();
            }
        }
    }
    __syncthreads();
    for (int p = tid; p < ns; p += blockDim.x) {
        double v0, v1;
        if (p < mT) {
            v0 = zb[p];
            v1 = zb[np + p];
        } else if (p < mT + b) {
            v0 = xs[p - mT];
            v1 = xs[64 + p - mT];
        } else {
            const int q = ns - 1 - p;
            v0 = zb[2 * np + q];
            v1 = zb[3 * np + q];
        }
        const int sg = D.perm[p];
        D.y[sg] = v0;
        D.y[ns + sg] = v1;
    }
}
template __global__ void k_crit_solve2<32>(const DevCase*, const CaseClock*, int);
template __global__ void k_crit_solve2<16>(const DevCase*, const CaseClock*, int);

size_t solve2_smem_bytes(int np2) { return sizeof(double) * (6 * (size_t)np2 + 256); }  // values, 1/d, f_S, x_S
// the largest staging unit whose two per-lane rings fit beside the values
int solve2_unit(int np2)
{
    for (int U : {32, 16})
        if (solve2_smem_bytes(np2) + 2 * 2 * (size_t)U * 64 * 16 <= 160 * 1024) return U;
    return 0;
}
int solve2_threads() { return 64 * (4 + SV_NPROD); }

// The 64-step staging unit when its two LDS buffers (128 KB) fit beside the
// right-hand sides, the 32-step one otherwise.
int solve_unit(int npad) { return 3 * 8 * (size_t)npad + 2 * 64 * 64 * 16 <= 160 * 1024 ? 64 : 32; }
int solve_threads() { return 64 * (2 + SV_NPROD); }

size_t factor_smem_bytes() { return sizeof(double) * ((size_t)WIN * WIN + 2 * WIN + 2 * CH * WIN + 4) + sizeof(int) * PAIRMAX; }
size_t solve_smem_bytes(int npad) { return sizeof(double) * 3 * (size_t)npad; }

// Spread delta_u = sum_k D(k,a) y(k) in ascending k (the fixed-order gather in
// place of the spread DGEMV, boundary_condition.f90:1639-1640), velocity
// correction and body force (1647-1657), and the forced collision
// (solution.f90:598-638) of the IB nodes, which k_lattice leaves out;
// the two run concurrently.
template <typename T>
__device__ __forceinline__ void band_node(const DevCase& D, const Slot& S, int a, int src, int dst,
                                          double* __restrict__ macro)
{
    const int cnt = min(S.nmkcnt[a], NCAP);
    const int* km = S.nmk + (long long)a * NCAP;
    const double* wm = S.nmkv + (long long)a * NCAP;
    const int ns = D.nseg;
    const int* ad = S.adj + 3 * a;
    long long c = (long long)ad[1] * D.nx + ad[0];
    double f[9], r, us, vs;
    pull9(pops<T>(D, src), D.N, D.nx, c, f, D.W);
    double gx = 0.0, gy = 0.0;
    for (int q = 0; q < cnt; ++q) {
        int s = km[q];
        double w = wm[q];
        gx = gx + w * D.y[s];
        gy = gy + w * D.y[ns + s];
    }
    S.glx[a] = gx;
    S.gly[a] = gy;
    moments(f, r, us, vs);
    S.rhoadj[a] = r;
    double u = us + gx;
    double v = vs + gy;
    double gfx = 2.0 * r * gx;
    double gfy = 2.0 * r * gy;
    bool forced = (gfx != 0.0) || (gfy != 0.0);
    collide_node(f, r, u, v, D.R, forced, gfx, gfy);
    GAS T* o = pops<T>(D, dst);
#pragma unroll
    for (int k = 0; k < 9; ++k) lower(f[k], D.W[k], o + k * D.N + c);
    if (macro) { macro[c] = r; macro[D.N + c] = u; macro[2 * D.N + c] = v; }
}

template <typename T>
__global__ __launch_bounds__(256) void k_band(const DevCase* __restrict__ C, const CaseClock* __restrict__ K, int off,
                                              int psel)
{
    __builtin_amdgcn_s_setprio(3);  // critical path: ahead of co-resident lattice waves
    const DevCase& D = C[blockIdx.y];
    const CaseClock kc = K[blockIdx.y];
    const int n = kc.n0 + off;
    if (n > kc.n_end) return;
    int src, dst;
    step_planes(D, n, psel, src, dst);
    double* macro = n >= kc.flag_from ? D.macro[n & 1] : nullptr;
    const Slot S = slot_of(D, (n - 1) % D.nslot);
    const int na = *S.nadj;
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < na; a += gridDim.x * blockDim.x)
        band_node<T>(D, S, a, src, dst, macro);
}
template __global__ void k_band<double>(const DevCase*, const CaseClock*, int, int);
template __global__ void k_band<float>(const DevCase*, const CaseClock*, int, int);

// Per-wing force and moment, accumulated in node order
// (boundary_condition.f90:1662-1675), and their normalisation
// (solution.f90:327-330).  One workgroup per step of a batch.  The nodes'
// terms are formed in chunks of FCH nodes into one of two LDS buffers by
// wavefront 3 while wavefronts 0, 1 and 2 advance the running sums F_x, F_y
// and M node by node over the other buffer: each serial chain has a wave of
// its own (a wave issues one FP64 add at a time), and the sums overlap the
// loads.  The nodes of a wing are contiguous in the list (segment-major
// order).  Block (x, c) takes step n0 + off + x of case c.
constexpr int FCH = 1024;
__global__ __launch_bounds__(256) void k_forces(const DevCase* __restrict__ C, const CaseClock* __restrict__ K, int off)
{
    const DevCase& D = C[blockIdx.y];
    const int n = K[blockIdx.y].n0 + off + (int)blockIdx.x;
    if (n > K[blockIdx.y].n_end) return;
    const int step = n - 1;  // 0-based
    const Slot S = slot_of(D, step % D.nslot);
    __shared__ double terms[2][FCH][4];
    __shared__ int wing[2][FCH];
    __shared__ double sab[8][3];  // body offsets sa, sb and body_extra_pos of each wing at this step
    const int tid = threadIdx.x;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // 0, 1, 2: F_x, F_y, M; 3: forms terms
    const int na = *S.nadj;
    const double sc = D.scale;
    if (tid < D.nw) {
        const double* kn = D.kin + ((long long)step * D.nw + tid) * KIN;
        sab[tid][0] = kn[3];
        sab[tid][1] = kn[5];
        sab[tid][2] = D.body[4 * tid + 3];
    }
    __syncthreads();
    // terms of chunk ch into buffer ch & 1, by threads t0 .. t0 + nt - 1
    auto form = [&](int ch, int t0, int nt) {
        const int base = ch * FCH, cnt = min(FCH, na - base), b = ch & 1;
        constexpr int UN = 4;
        for (int q0 = tid - t0; q0 < cnt; q0 += UN * nt) {
            int ax[UN], ay[UN], aw[UN];
            double rho[UN], gx[UN], gy[UN];
#pragma unroll
            for (int u = 0; u < UN; ++u) {
                const int q = min(q0 + u * nt, cnt - 1), a = base + q;
                ax[u] = S.adj[3 * a];
                ay[u] = S.adj[3 * a + 1];
                aw[u] = S.adj[3 * a + 2];
                rho[u] = S.rhoadj[a];
                gx[u] = S.glx[a];
                gy[u] = S.gly[a];
            }
#pragma unroll
            for (int u = 0; u < UN; ++u) {
                const int q = q0 + u * nt;
                if (q >= cnt) break;
                const int w = aw[u];
                const double t1 = 2.0 * rho[u] * gx[u];
                const double t2 = 2.0 * rho[u] * gy[u];
                const double x0 = (D.ox + sc * (double)ax[u] - (D.xc + sab[w][0])) / sc - sab[w][2];
                const double y0 = (D.oy + sc * (double)ay[u] - (D.yc + sab[w][1])) / sc;
                terms[b][q][0] = t1;
                terms[b][q][1] = t2;
                terms[b][q][2] = x0 * t2;
                terms[b][q][3] = y0 * t1;
                wing[b][q] = w;
            }
        }
    };
    double A = 0.0;  // this wave's running sum
    int curw = 0;
    double* hist = D.hist + (long long)step * 3 * D.nw;
    const double U = D.U;
    double* raw = D.raw ? D.raw + (long long)step * 3 * D.nw : nullptr;
    auto flush = [&](int w) {
        if ((tid & 63) != 0) return;
        if (raw) raw[3 * w + wv] = A;
        if (wv == 2) hist[3 * w + 2] = -2.0 * A / (U * U * (1.0 / (sc * sc)));
        else hist[3 * w + wv] = 2.0 * A / (U * U * (1.0 / sc));
    };
    const int nch = (na + FCH - 1) / FCH;
    if (nch > 0) form(0, 0, 256);
    __syncthreads();
    for (int ch = 0; ch < nch; ++ch) {
        if (wv == 3) {
            if (ch + 1 < nch) form(ch + 1, 192, 64);
        } else {
            // the wave sums its quantity over the chunk in ascending node
            // order, every lane the same values: each run of one wing's nodes
            // (wing ids never decrease along the list) is found with a
            // ballot, then summed with no branch per node, the terms of PF
            // nodes read ahead
            const int cnt = min(FCH, na - ch * FCH), b = ch & 1;
            int q = 0;
            while (q < cnt) {
                const int w = __builtin_amdgcn_readfirstlane(wing[b][q]);
                if (w != curw) {
                    flush(curw);
                    for (int x = curw + 1; x < w; ++x) { A = 0.0; flush(x); }
                    curw = w;
                    A = 0.0;
                }
                int e = cnt;
                for (int base = q + 1; base < cnt; base += 64) {
                    const int idx = base + (tid & 63);
                    const unsigned long long other = __ballot(idx < cnt && wing[b][idx] != w);
                    if (other) {
                        e = base + __builtin_ctzll(other);
                        break;
                    }
                }
                constexpr int PF = 8;
                if (wv == 2) {
                    for (; q + PF <= e; q += PF) {
                        double t2[PF], t3[PF];
#pragma unroll
                        for (int u = 0; u < PF; ++u) {
                            t2[u] = terms[b][q + u][2];
                            t3[u] = terms[b][q + u][3];
                        }
#pragma unroll
                        for (int u = 0; u < PF; ++u) A = A - t2[u] + t3[u];
                    }
                    for (; q < e; ++q) A = A - terms[b][q][2] + terms[b][q][3];
                } else {
                    for (; q + PF <= e; q += PF) {
                        double t[PF];
#pragma unroll
                        for (int u = 0; u < PF; ++u) t[u] = terms[b][q + u][wv];
#pragma unroll
                        for (int u = 0; u < PF; ++u) A = A - t[u];
                    }
                    for (; q < e; ++q) A = A - terms[b][q][wv];
                }
            }
        }
        __syncthreads();
    }
    if (wv < 3) {
        flush(curw);
        for (int x = curw + 1; x < D.nw; ++x) { A = 0.0; flush(x); }
    }
}

// Moves every case's window forward by `steps` (the last node of a captured
// window graph, so that its next replay takes the following steps); n_end
// and flag_from are absolute steps and stay.
__global__ void k_clock_advance(CaseClock* __restrict__ K, int ncases, int steps)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncases) K[c].n0 += steps;
}

// erranalysis partial sums over the interior (solution.f90:102-109): per
// block sum of |u-u0|^2 and |u|^2 (fixed block-tree order).
__global__ __launch_bounds__(256) void k_residual(const double* __restrict__ mc,
                                                  const double* __restrict__ mp, int nx, int ny,
                                                  long long N, double* __restrict__ out)
{
    __shared__ double s1[256], s2[256];
    long long n_int = (long long)(nx - 2) * (ny - 2);
    double a1 = 0.0, a2 = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n_int;
         t += (long long)gridDim.x * blockDim.x) {
        int i = (int)(t % (nx - 2)) + 1;
        int j = (int)(t / (nx - 2)) + 1;
        long long c = (long long)j * nx + i;
        double u = mc[N + c], v = mc[2 * N + c];
        double du = u - mp[N + c], dv = v - mp[2 * N + c];
        a1 = a1 + du * du + dv * dv;
        a2 = a2 + u * u + v * v;
    }
    s1[threadIdx.x] = a1;
    s2[threadIdx.x] = a2;
    __syncthreads();
    for (unsigned off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            s1[threadIdx.x] += s1[threadIdx.x + off];
            s2[threadIdx.x] += s2[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = s1[0];
        out[2 * blockIdx.x + 1] = s2[0];
    }
}

// The vorticity column of output (solution.f90:181-210): second-order
// one-sided differences on the first and last column (row), central ones
// elsewhere, in units of U/scale.  With hv set, the same pass is one call of
// vor_mom_rate (solution.f90:1880-1897) on the interior nodes: hv and hv1 hold
// the vorticity of the last two calls and shift by one, dvdt gets the
// second-order backward difference.  (vor_mom_rate's vor2 is written and
// never read, so it is not kept.)  One thread per node in x-fastest order;
// the j-1 and j+1 rows of u come through L2, so u and v are read from HBM
// about once.
__global__ __launch_bounds__(256) void k_vorticity(const double* __restrict__ m, int nx, int ny, long long N,
                                                   double UU, double scale, double* __restrict__ vor,
                                                   double* __restrict__ hv, double* __restrict__ hv1,
                                                   double* __restrict__ dvdt)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= N) return;
    const int j = (int)(c / nx), i = (int)(c - (long long)j * nx);
    const double* u = m + N;
    const double* v = m + 2 * N;
    double dvx, duy;
    if (i == nx - 1)
        dvx = (1.5 * v[c] - 2.0 * v[c - 1] + 0.5 * v[c - 2]) / (UU * scale);
    else if (i == 0)
        dvx = (-1.5 * v[c] + 2.0 * v[c + 1] - 0.5 * v[c + 2]) / (UU * scale);
    else
        dvx = (v[c + 1] - v[c - 1]) / (UU * 2.0 * scale);
    if (j == ny - 1)
        duy = (1.5 * u[c] - 2.0 * u[c - nx] + 0.5 * u[c - 2 * nx]) / (UU * scale);
    else if (j == 0)
        duy = (-1.5 * u[c] + 2.0 * u[c + nx] - 0.5 * u[c + 2 * nx]) / (UU * scale);
    else
        duy = (u[c + nx] - u[c - nx]) / (UU * 2.0 * scale);
    const double w = dvx - duy;
    vor[c] = w;
    if (hv && i > 0 && i < nx - 1 && j > 0 && j < ny - 1) {
        const double tem1 = hv1[c], tem2 = hv[c];
        hv1[c] = tem2;
        hv[c] = w;
        dvdt[c] = (1.5 * w - 2.0 * tem2 + 0.5 * tem1) / (UU * scale);
    }
}

// cal_sur_cp (solution.f90:1912-2118): the pressure rho/3 of the IB nodes
// interpolated onto each marker midpoint with the step's kernel weights
// phi_x phi_y, summed in ascending node order as matmulvocxs does
// (boundary_condition.f90:1887-1900).  One lane per marker.
__global__ __launch_bounds__(64) void k_surface_pressure(const int* __restrict__ dcnt, const int* __restrict__ didx,
                                                         const double* __restrict__ dval,
                                                         const double* __restrict__ rhoadj, int nseg,
                                                         double* __restrict__ out)
{
    const int k = blockIdx.x * 64 + threadIdx.x;
    if (k >= nseg) return;
    const int n = min(dcnt[k], BOXMAX);
    double bb = 0.0;
    for (int p = 0; p < n; ++p)
        bb = bb + dval[k * BOXMAX + p] * (rhoadj[didx[k * BOXMAX + p]] / 3.0);
    out[k] = bb;
}

// move_mesh (move_mesh.f90:16-58 for dr = 1, 172-214 for dr = -1): the
// populations of every node take those of the node dr columns further, the
// column without a source keeps its own.  Out of place; the caller copies
// back.
template <typename T>
__global__ __launch_bounds__(256) void k_move_pops(const T* __restrict__ src, T* __restrict__ dst, int nx, long long N,
                                                   int dr)
{
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= 9 * N) return;
    const int i = (int)((t % N) % nx);
    const bool keep = dr == 1 ? i == nx - 1 : i == 0;
    dst[t] = src[keep ? t : t + dr];
}
template __global__ void k_move_pops<double>(const double*, double*, int, long long, int);
template __global__ void k_move_pops<float>(const float*, float*, int, long long, int);

// The stale f0 of the edge column that the shift refills (f0(1,:) from
// f(2,:) for dr = 1, f0(Nx,:) from f(Nx-1,:) for dr = -1), as full FP64
// populations; the other edge column keeps its stale f0.
template <typename T>
__global__ __launch_bounds__(256) void k_move_ring(const T* __restrict__ src, double* __restrict__ f0col, int nx,
                                                   int ny, long long N, int dr, RingW w)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= 9 * ny) return;
    const int k = t / ny, j = t - k * ny;
    const int i = dr == 1 ? 1 : nx - 2;
    f0col[(long long)k * ny + j] = lift(src[(long long)k * N + (long long)j * nx + i], w.w[k]);
}
template __global__ void k_move_ring<double>(const double*, double*, int, int, long long, int, RingW);
template __global__ void k_move_ring<float>(const float*, double*, int, int, long long, int, RingW);

// The recorded rho/u/v of the step (cur) and of the step before (prev): in
// pass 0 prev takes cur shifted (u0/v0 of move_mesh), in pass 1 cur takes
// prev back; the column without a source keeps its values in both.
__global__ __launch_bounds__(256) void k_move_macro(double* __restrict__ cur, double* __restrict__ prev, int nx,
                                                    long long N, int dr, int pass)
{
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= 3 * N) return;
    const int i = (int)((t % N) % nx);
    if (dr == 1 ? i == nx - 1 : i == 0) return;
    if (pass == 0)
        prev[t] = cur[t + dr];
    else
        cur[t] = prev[t];
}

// NaN guard over the populations.
template <typename T>
__global__ __launch_bounds__(256) void k_nan_check(const T* __restrict__ f, long long n, int* __restrict__ flag)
{
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        if (f[t] != f[t]) { *flag = 1; return; }
}
template __global__ void k_nan_check<double>(const double*, long long, int*);
template __global__ void k_nan_check<float>(const float*, long long, int*);

}  // namespace lbmib
