// lbmib_kernels.hip -- gfx950 kernels of the FP64 drop-in timestep.
//
// One timestep (reference evolution, solution.f90:269-410) is
//   lattice:  k_collide_f64   fused pull-stream + moments + MRT collide of every
//                             interior node, prev -> cur (HBM-bound, 144 B/node)
//             k_edges_f64     characteristic/NEE outer boundary, prev -> cur ring
//   IB chain: k_markers -> k_clear -> k_claim -> k_count -> k_fill -> k_nodes ->
//             k_dvals -> k_assemble -> k_chol -> k_spread   (one small stream)
//             k_band_f64      forced collide of the IB nodes, overwrites cur there
//             k_forces        per-wing force / moment, fixed order
// Nothing here uses floating-point atomics; every reduction has a fixed order.
#include "lbmib_kernels.h"

#include <climits>

namespace lbmib {

// ---------------------------------------------------------------------------
// lattice
// ---------------------------------------------------------------------------

// Pull the nine populations of interior node (i,j) from prev
// (streaming, solution.f90:438-454: f0(i,j,k) = f(i-ex,j-ey,k)).
__device__ __forceinline__ void pull9(const double* __restrict__ src, long long N, int nx,
                                      long long c, double* f)
{
    f[0] = src[c];
    f[1] = src[1 * N + c - 1];
    f[2] = src[2 * N + c - nx];
    f[3] = src[3 * N + c + 1];
    f[4] = src[4 * N + c + nx];
    f[5] = src[5 * N + c - nx - 1];
    f[6] = src[6 * N + c - nx + 1];
    f[7] = src[7 * N + c + nx + 1];
    f[8] = src[8 * N + c + nx - 1];
}

// Fused streaming + physical_variables + colission for all interior nodes
// (solution.f90:428-643).  The IB nodes are recomputed with forcing by
// k_band_f64 afterwards; everywhere else gforce is zero and the force moments
// vanish exactly.  Optionally records rho/u/v (download / residual support).
__global__ __launch_bounds__(256) void k_collide_f64(const double* __restrict__ src,
                                                     double* __restrict__ dst, int nx, int ny,
                                                     long long N, Rates R,
                                                     double* __restrict__ macro)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y + 1;
    if (i > nx - 2) return;
    long long c = (long long)j * nx + i;
    double f[9];
    pull9(src, N, nx, c, f);
    double rho, u, v;
    moments(f, rho, u, v);
    if (macro) {
        macro[c] = rho;
        macro[N + c] = u;
        macro[2 * N + c] = v;
    }
    collide_node(f, rho, u, v, R, false, 0.0, 0.0);
#pragma unroll
    for (int k = 0; k < 9; ++k) dst[k * N + c] = f[k];
}

// Moments of a boundary-ring node from its stale f0 (never streamed,
// solution.f90:516-520).
__device__ __forceinline__ void ring_moments(const double* f0col, int ny, int j, double& r,
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

// Edge state of column `col` (0: left, 1: right) at interior row j: the
// characteristic-updated rho/u/v and the neighbour's f0 and moments.
__device__ void edge_column_state(const DevCase& D, const double* src, double cs, int col, int j,
                                  double& ro, double& uo, double& vo, double* f0n, double& rn,
                                  double& un, double& vn)
{
    const int nx = D.nx;
    const long long N = D.N;
    const double* f0c = col == 0 ? D.f0L : D.f0R;
    double tr, tu, tv;
    ring_moments(f0c, D.ny, j, tr, tu, tv);
    int i2 = col == 0 ? 1 : nx - 2;
    int i3 = col == 0 ? 2 : nx - 3;
    double f3[9], r3, u3, v3;
    pull9(src, N, nx, (long long)j * nx + i2, f0n);
    moments(f0n, rn, un, vn);
    pull9(src, N, nx, (long long)j * nx + i3, f3);
    moments(f3, r3, u3, v3);
    if (col == 0)
        left_char(cs, tr, tu, tv, rn, un, vn, r3, u3, v3, ro, uo, vo);
    else
        right_char(cs, tr, tu, tv, rn, un, vn, r3, u3, v3, ro, uo, vo);
}

// boundary_condition1 (boundary_condition.f90:288-858) with no periodic
// edges.  Threads 0..ny-3 take the left and right columns at rows 1..ny-2;
// threads ny-2..ny-2+nx-1 take the top and bottom rows at column x.  The top
// and bottom characteristic updates and corner states of the reference are
// overwritten by rho = rho(.,Ny-1), u = v = 0 before use (658-661, 785-788),
// so only that effective state is computed.  Corner columns of the walls read
// the left/right results at rows 1 and ny-2, recomputed here.
__global__ __launch_bounds__(256) void k_edges_f64(DevCase D, int cur, double cs,
                                                   double* __restrict__ macro)
{
    const int nx = D.nx, ny = D.ny;
    const long long N = D.N;
    const double* src = D.f[cur ^ 1];
    double* dst = D.f[cur];
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ny - 2) {
        int j = t + 1;
        for (int col = 0; col < 2; ++col) {
            double ro, uo, vo, f0n[9], rn, un, vn, out[9];
            edge_column_state(D, src, cs, col, j, ro, uo, vo, f0n, rn, un, vn);
            nee_node(ro, uo, vo, rn, un, vn, f0n, D.R, out);
            int i = col == 0 ? 0 : nx - 1;
            long long c = (long long)j * nx + i;
#pragma unroll
            for (int k = 0; k < 9; ++k) dst[k * N + c] = out[k];
            if (macro) { macro[c] = ro; macro[N + c] = uo; macro[2 * N + c] = vo; }
        }
        return;
    }
    t -= ny - 2;
    if (t >= nx) return;
    int x = t;
    for (int wall = 0; wall < 2; ++wall) {
        // wall 0: top (row ny-1, neighbour ny-2); wall 1: bottom (row 0, neighbour 1)
        int jn = wall == 0 ? ny - 2 : 1;
        int je = wall == 0 ? ny - 1 : 0;
        double rn, un, vn, f0n[9];
        if (x == 0 || x == nx - 1) {
            int col = x == 0 ? 0 : 1;
            double fdummy[9], r2, u2, v2;
            edge_column_state(D, src, cs, col, jn, rn, un, vn, fdummy, r2, u2, v2);
            const double* f0c = col == 0 ? D.f0L : D.f0R;
#pragma unroll
            for (int k = 0; k < 9; ++k) f0n[k] = f0c[k * ny + jn];
        } else {
            pull9(src, N, nx, (long long)jn * nx + x, f0n);
            moments(f0n, rn, un, vn);
        }
        double re = rn, ue = 0.0, ve = 0.0, out[9];
        nee_node(re, ue, ve, rn, un, vn, f0n, D.R, out);
        long long c = (long long)je * nx + x;
#pragma unroll
        for (int k = 0; k < 9; ++k) dst[k * N + c] = out[k];
        if (macro) { macro[c] = re; macro[N + c] = ue; macro[2 * N + c] = ve; }
    }
}

// ---------------------------------------------------------------------------
// immersed boundary
// ---------------------------------------------------------------------------

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

// Markers, marker velocities and boxes of step `step` (move_boundary
// airfoil.f90:256-270 and boundary_condition.f90:1542-1553).  sin/cos of the
// pitch angle come precomputed from the host's libm.
__global__ __launch_bounds__(256) void k_markers(DevCase D, int step)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.nseg) return;
    int w = s / D.spw;
    int p = w * D.ppw + (s - w * D.spw);
    const double* kn = D.kin + ((long long)step * D.nw + w) * KIN;
    double sn = kn[0], ca = kn[1], om = kn[2], sa = kn[3], sau = kn[4], sb = kn[5], sbv = kn[6];
    double xa = D.xb0[p], ya = D.yb0[p], xb = D.xb0[p + 1], yb = D.yb0[p + 1];
    double x1a = ca * xa + sn * ya + sa + D.xc;
    double x1b = ca * xb + sn * yb + sa + D.xc;
    double y1a = -sn * xa + ca * ya + sb + D.yc;
    double y1b = -sn * xb + ca * yb + sb + D.yc;
    double X = 0.5 * (x1b + x1a);
    double Y = 0.5 * (y1b + y1a);
    D.xb2[s] = X;
    D.yb2[s] = Y;
    double xs = xb + xa, ys = yb + ya;
    double U = D.U;
    D.ub[s] = U * (-om * sn * 0.5 * xs + om * ca * 0.5 * ys + sau);
    D.vb[s] = U * (-om * ca * 0.5 * xs - om * sn * 0.5 * ys + sbv);
    int i1, i2, j1, j2;
    axis_box(X, D.ox, D.scale, D.nx, i1, i2);
    axis_box(Y, D.oy, D.scale, D.ny, j1, j2);
    bool ok = i1 >= 4 && i2 <= D.nx - 3 && j1 >= 4 && j2 <= D.ny - 3 && i2 >= i1 && j2 >= j1 &&
              (i2 - i1) < 6 && (j2 - j1) < 6;
    if (!ok) {
        atomicOr(D.err, ERR_GEOMETRY);
        i1 = 4; i2 = 3; j1 = 4; j2 = 3;  // empty box keeps every later access in bounds
    }
    int* b = D.box + 4 * s;
    b[0] = i1; b[1] = i2; b[2] = j1; b[3] = j2;
}

// Clear the previous node set (airfoil.f90:276-283).
__global__ __launch_bounds__(256) void k_clear(DevCase D)
{
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= *D.nadj) return;
    long long c = (long long)D.adj[3 * a + 1] * D.nx + D.adj[3 * a];
    D.claim[c] = INT_MAX;
    D.nodeidx[c] = -1;
}

__device__ __forceinline__ bool box_cell(const int* b, int q, int& i, int& j)
{
    int nyb = b[3] - b[2] + 1;
    int n = (b[1] - b[0] + 1) * nyb;
    if (nyb <= 0 || q >= n) return false;
    i = b[0] + q / nyb;  // i-major, j-minor: the reference's loop order (airfoil.f90:322-323)
    j = b[2] + q % nyb;
    return true;
}

// First claim of every box node: the smallest segment index wins, which is
// the segment that claims it first in the reference's serial loop.
__global__ __launch_bounds__(256) void k_claim(DevCase D)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int s = t / BOXMAX, q = t % BOXMAX;
    if (s >= D.nseg) return;
    int i, j;
    if (!box_cell(D.box + 4 * s, q, i, j)) return;
    atomicMin(&D.claim[(long long)(j - 1) * D.nx + (i - 1)], s);
}

__global__ __launch_bounds__(256) void k_count(DevCase D)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.nseg) return;
    const int* b = D.box + 4 * s;
    int cnt = 0, i, j;
    for (int q = 0; q < BOXMAX; ++q) {
        if (!box_cell(b, q, i, j)) break;
        if (D.claim[(long long)(j - 1) * D.nx + (i - 1)] == s) cnt++;
    }
    D.segcnt[s] = cnt;
}

// Ordered compaction: the node list in the reference's order (segment-major,
// then i, then j), the node -> list index grid and the node count.
__global__ __launch_bounds__(1024) void k_fill(DevCase D)
{
    __shared__ int part[1024];
    const int ns = D.nseg;
    const int T = blockDim.x;
    int chunk = (ns + T - 1) / T;
    int s0 = threadIdx.x * chunk, s1 = min(ns, s0 + chunk);
    int sum = 0;
    for (int s = s0; s < s1; ++s) sum += D.segcnt[s];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < T; off <<= 1) {
        int v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int base = part[threadIdx.x] - sum;
    if (threadIdx.x == T - 1) {
        if (part[T - 1] > D.cap) atomicOr(D.err, ERR_CAPACITY);
        *D.nadj = min(part[T - 1], D.cap);
    }
    for (int s = s0; s < s1; ++s) {
        const int* b = D.box + 4 * s;
        int w = s / D.spw, i, j;
        for (int q = 0; q < BOXMAX; ++q) {
            if (!box_cell(b, q, i, j)) break;
            long long c = (long long)(j - 1) * D.nx + (i - 1);
            if (D.claim[c] == s) {
                if (base < D.cap) {
                    D.adj[3 * base] = i - 1;
                    D.adj[3 * base + 1] = j - 1;
                    D.adj[3 * base + 2] = w;
                    D.nodeidx[c] = base;
                }
                base++;
            }
        }
    }
}

// u*, v*, rho at the IB nodes (physical_variables at those nodes).
__global__ __launch_bounds__(256) void k_nodes(DevCase D, int cur)
{
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= *D.nadj) return;
    const double* src = D.f[cur ^ 1];
    long long c = (long long)D.adj[3 * a + 1] * D.nx + D.adj[3 * a];
    double f[9], r, u, v;
    pull9(src, D.N, D.nx, c, f);
    moments(f, r, u, v);
    D.rhoadj[a] = r;
    D.ustar[a] = u;
    D.vstar[a] = v;
    D.nmkcnt[a] = 0;
}

// Per segment: kernel weights of its box nodes sorted by list index
// (xx1 row, boundary_condition.f90:1558-1606), the interpolation D u*
// (DGEMV, 1620-1621) in that order, the right-hand side U_B - D u* (1624-1629)
// in zig-zag order, and registration in the nodes' marker lists.
__global__ __launch_bounds__(64) void k_dvals(DevCase D)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.nseg) return;
    const int* b = D.box + 4 * s;
    const double sc = D.scale;
    double px[6], py[6];
    int nxb = b[1] - b[0] + 1, nyb = b[3] - b[2] + 1;
    double X = D.xb2[s], Y = D.yb2[s];
    for (int q = 0; q < nxb; ++q) {
        double x0 = D.ox + sc * (double)(b[0] + q - 1);
        px[q] = delta1(fabs((x0 - X) / sc), D.smooth);
    }
    for (int q = 0; q < nyb; ++q) {
        double y0 = D.oy + sc * (double)(b[2] + q - 1);
        py[q] = delta1(fabs((y0 - Y) / sc), D.smooth);
    }
    int idx[BOXMAX];
    double val[BOXMAX];
    int n = 0;
    for (int qi = 0; qi < nxb; ++qi)
        for (int qj = 0; qj < nyb; ++qj) {
            long long c = (long long)(b[2] + qj - 1) * D.nx + (b[0] + qi - 1);
            int node = D.nodeidx[c];
            double w = px[qi] * py[qj];
            // insertion by list index
            int m = n++;
            while (m > 0 && idx[m - 1] > node) { idx[m] = idx[m - 1]; val[m] = val[m - 1]; --m; }
            idx[m] = node;
            val[m] = w;
            if (node >= 0) {
                int slot = atomicAdd(&D.nmkcnt[node], 1);
                if (slot < NCAP) {
                    D.nmk[(long long)node * NCAP + slot] = s;
                    D.nmkv[(long long)node * NCAP + slot] = w;
                } else {
                    atomicOr(D.err, ERR_CAPACITY);
                }
            } else {
                atomicOr(D.err, ERR_CAPACITY);
            }
        }
    double su = 0.0, sv = 0.0;
    for (int a = 0; a < n; ++a) {
        int node = idx[a] < 0 ? 0 : idx[a];
        su = su + val[a] * D.ustar[node];
        sv = sv + val[a] * D.vstar[node];
        D.didx[s * BOXMAX + a] = idx[a];
        D.dval[s * BOXMAX + a] = val[a];
    }
    D.dcnt[s] = n;
    int p = D.iperm[s];
    D.rhs[p] = D.ub[s] - su;
    D.rhs[D.nseg + p] = D.vb[s] - sv;
}

// Band of G = D D^T in zig-zag order: G(p,q) = sum over the nodes shared by
// the boxes of segments perm[p], perm[q], ascending list index.
__global__ __launch_bounds__(256) void k_assemble(DevCase D)
{
    const int W = D.bw + 1;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)D.nseg * W) return;
    int p = (int)(t / W), d = (int)(t % W), q = p - d;
    double sum = 0.0;
    if (q >= 0) {
        int k = D.perm[p], l = D.perm[q];
        const int* bk = D.box + 4 * k;
        const int* bl = D.box + 4 * l;
        if (!(bk[1] < bl[0] || bl[1] < bk[0] || bk[3] < bl[2] || bl[3] < bk[2])) {
            const int* ik = D.didx + k * BOXMAX;
            const int* il = D.didx + l * BOXMAX;
            const double* vk = D.dval + k * BOXMAX;
            const double* vl = D.dval + l * BOXMAX;
            int nk = D.dcnt[k], nl = D.dcnt[l], a = 0, c = 0;
            while (a < nk && c < nl) {
                int xa = ik[a], xc = il[c];
                if (xa < xc) a++;
                else if (xa > xc) c++;
                else { sum = sum + vk[a] * vl[c]; a++; c++; }
            }
        }
    }
    D.G[t] = sum;
}

// Banded Cholesky of G and the two triangular solves for both right-hand
// sides, right-looking: element (i,j) receives its updates in increasing k.
// One workgroup; the band lives in global memory (L2-resident).
__global__ __launch_bounds__(1024) void k_chol(DevCase D)
{
    const int n = D.nseg, bw = D.bw, W = bw + 1;
    double* G = D.G;
    double* x = D.rhs;
    const int tid = threadIdx.x, T = blockDim.x;
    for (int k = 0; k < n; ++k) {
        int m = min(k + bw, n - 1) - k;
        double akk = G[(long long)k * W];
        double lkk = sqrt(akk);
        __syncthreads();
        if (tid == 0) {
            G[(long long)k * W] = lkk;
            if (!(akk > 0.0)) atomicOr(D.err, ERR_SOLVE);
        }
        for (int t = tid; t < m; t += T) {
            int i = k + 1 + t;
            long long e = (long long)i * W + (i - k);
            G[e] = G[e] / lkk;
        }
        __syncthreads();
        for (int t = tid; t < m * m; t += T) {
            int a = t / m, b = t % m;
            if (b > a) continue;
            int i = k + 1 + a, j = k + 1 + b;
            long long e = (long long)i * W + (i - j);
            G[e] = G[e] - G[(long long)i * W + (i - k)] * G[(long long)j * W + (j - k)];
        }
        __syncthreads();
    }
    // forward L z = b
    for (int k = 0; k < n; ++k) {
        int m = min(k + bw, n - 1) - k;
        if (tid < 2) x[tid * n + k] = x[tid * n + k] / G[(long long)k * W];
        __syncthreads();
        for (int t = tid; t < 2 * m; t += T) {
            int r = t & 1, i = k + 1 + (t >> 1);
            x[r * n + i] = x[r * n + i] - G[(long long)i * W + (i - k)] * x[r * n + k];
        }
        __syncthreads();
    }
    // backward L^T x = z
    for (int k = n - 1; k >= 0; --k) {
        int m = k - max(0, k - bw);
        if (tid < 2) x[tid * n + k] = x[tid * n + k] / G[(long long)k * W];
        __syncthreads();
        for (int t = tid; t < 2 * m; t += T) {
            int r = t & 1, i = k - 1 - (t >> 1);
            x[r * n + i] = x[r * n + i] - G[(long long)k * W + (k - i)] * x[r * n + k];
        }
        __syncthreads();
    }
    for (int p = tid; p < n; p += T) {
        int s = D.perm[p];
        D.y[s] = x[p];
        D.y[n + s] = x[n + p];
    }
}

// Spread delta_u(i) = sum_k D(k,i) y(k), k ascending (the fixed-order gather
// replacing the spread DGEMV, boundary_condition.f90:1639-1640).
__global__ __launch_bounds__(256) void k_spread(DevCase D)
{
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= *D.nadj) return;
    int cnt = min(D.nmkcnt[a], NCAP);
    int ks[NCAP];
    double ws[NCAP];
    for (int q = 0; q < cnt; ++q) {
        int s = D.nmk[(long long)a * NCAP + q];
        double w = D.nmkv[(long long)a * NCAP + q];
        int m = q;
        while (m > 0 && ks[m - 1] > s) { ks[m] = ks[m - 1]; ws[m] = ws[m - 1]; --m; }
        ks[m] = s;
        ws[m] = w;
    }
    const int n = D.nseg;
    double sx = 0.0, sy = 0.0;
    for (int q = 0; q < cnt; ++q) {
        sx = sx + ws[q] * D.y[ks[q]];
        sy = sy + ws[q] * D.y[n + ks[q]];
    }
    D.glx[a] = sx;
    D.gly[a] = sy;
}

// Velocity correction, body force and forced collision at the IB nodes
// (boundary_condition.f90:1647-1657, solution.f90:598-638).  Runs after
// k_collide_f64 and overwrites its unforced result at these nodes.
__global__ __launch_bounds__(256) void k_band_f64(DevCase D, int cur, double* __restrict__ macro)
{
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= *D.nadj) return;
    const double* src = D.f[cur ^ 1];
    double* dst = D.f[cur];
    long long c = (long long)D.adj[3 * a + 1] * D.nx + D.adj[3 * a];
    double f[9], r, us, vs;
    pull9(src, D.N, D.nx, c, f);
    moments(f, r, us, vs);
    double u = us + D.glx[a];
    double v = vs + D.gly[a];
    double gfx = 2.0 * r * D.glx[a];
    double gfy = 2.0 * r * D.gly[a];
    bool forced = (gfx != 0.0) || (gfy != 0.0);
    collide_node(f, r, u, v, D.R, forced, gfx, gfy);
#pragma unroll
    for (int k = 0; k < 9; ++k) dst[k * D.N + c] = f[k];
    if (macro) { macro[D.N + c] = u; macro[2 * D.N + c] = v; }
}

// Per-wing force and moment, sequential in node order
// (boundary_condition.f90:1662-1675), and their normalisation
// (solution.f90:327-330).
__global__ void k_forces(DevCase D, int step)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double F[8][2], M[8];
    for (int w = 0; w < 8; ++w) { F[w][0] = 0.0; F[w][1] = 0.0; M[w] = 0.0; }
    const int n = *D.nadj;
    const double sc = D.scale;
    for (int a = 0; a < n; ++a) {
        int w = D.adj[3 * a + 2];
        const double* kn = D.kin + ((long long)step * D.nw + w) * KIN;
        double sa = kn[3], sb = kn[5];
        double t1 = 2.0 * D.rhoadj[a] * D.glx[a];
        double t2 = 2.0 * D.rhoadj[a] * D.gly[a];
        F[w][0] = F[w][0] - t1;
        F[w][1] = F[w][1] - t2;
        double x0 = (D.ox + sc * (double)D.adj[3 * a] - (D.xc + sa)) / sc;
        double y0 = (D.oy + sc * (double)D.adj[3 * a + 1] - (D.yc + sb)) / sc;
        M[w] = M[w] - x0 * t2 + y0 * t1;
    }
    const double U = D.U;
    double* h = D.hist + (long long)step * 3 * D.nw;
    for (int w = 0; w < D.nw; ++w) {
        h[3 * w + 0] = 2.0 * F[w][0] / (U * U * (1.0 / sc));
        h[3 * w + 1] = 2.0 * F[w][1] / (U * U * (1.0 / sc));
        h[3 * w + 2] = -2.0 * M[w] / (U * U * (1.0 / (sc * sc)));
    }
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
    for (int off = 128; off > 0; off >>= 1) {
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

// NaN guard over the populations.
__global__ __launch_bounds__(256) void k_nan_check(const double* __restrict__ f, long long n,
                                                   int* __restrict__ flag)
{
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        if (f[t] != f[t]) { *flag = 1; return; }
}

}  // namespace lbmib
