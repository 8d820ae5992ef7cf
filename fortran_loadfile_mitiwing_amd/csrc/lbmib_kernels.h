// lbmib_kernels.h -- device-side data of one lbmib case and the kernel
// declarations shared by the HIP translation units of liblbmib.
#pragma once

#include <hip/hip_runtime.h>

#include "lbmib_exact.h"

namespace lbmib {

constexpr int BOXMAX = 36;  // nodes in one marker's +-2.5 lattice-unit box (<= 6x6)
constexpr int NCAP = 64;    // markers whose boxes may contain one IB node
constexpr int KIN = 8;      // per wing per step: sin, cos, omega, sa, sau, sb, sbv, pad
constexpr int MAXBW = 62;   // largest bandwidth of the banded IB solver
constexpr int MAXSEG = 4096;
// separator form of the IB solve: largest bandwidth, and the slices of the
// separator matrices M (bw x 3bw) and N (2bw x bw) that each of the 512
// producer threads holds in registers: row pieces of SEP_L1 and SEP_L2 terms
constexpr int SEPMAXBW = 31;
// compact solve coefficients: CW values per timeline step (bandwidth < CW)
constexpr int CW = 32;
constexpr int SEP_L1 = 8;
constexpr int SEP_L2 = 8;
__host__ __device__ inline int sep_seg1(int bw) { return 512 / bw; }
__host__ __device__ inline int sep_len1(int bw) { return (3 * bw + sep_seg1(bw) - 1) / sep_seg1(bw); }
__host__ __device__ inline int sep_seg2(int bw) { return 512 / (2 * bw); }
__host__ __device__ inline int sep_len2(int bw) { return (bw + sep_seg2(bw) - 1) / sep_seg2(bw); }

enum ErrBits { ERR_GEOMETRY = 1, ERR_CAPACITY = 2, ERR_SOLVE = 4 };

// Offsets (in elements) of the per-step immersed-boundary arrays inside one
// slot of the int and double pools.
struct SlotInts {
    long long claim;    // [gw*gh] first claiming segment of each window node, INT_MAX when free
    long long nodeidx;  // [gw*gh] IB list index of each window node, or -1
    long long box;      // [nseg][4] i1,i2,j1,j2 (1-based)
    long long draw_idx; // [nseg][BOXMAX] box nodes in box order
    long long nraw_s;   // [cap][NCAP] markers of each node, arrival order
    long long adj;      // [cap][3] node i, j (0-based), wing
    long long nadj;     // [1] node count
    long long dcnt;     // [nseg] nodes in the segment's box
    long long didx;     // [nseg][BOXMAX] their list indices, ascending
    long long nmkcnt;   // [cap] markers whose box holds the node
    long long nmk;      // [cap][NCAP] those markers, ascending
    long long total;
};
struct SlotDoubles {
    long long xb2, yb2; // [nseg] marker midpoints
    long long ub, vb;   // [nseg] marker velocities
    long long phi;      // [nseg][12] one-axis kernel values of the box columns / rows
    long long draw_val; // [nseg][BOXMAX]
    long long nraw_w;   // [cap][NCAP]
    long long dval;     // [nseg][BOXMAX] kernel weights phi_x*phi_y, ascending node order
    long long nmkv;     // [cap][NCAP]
    long long G;        // [nseg][bw+1] band of D D^T, G[p][d] = A(p, p-d)
    long long coef;     // [2*npad+64][ws][2] per-lane coefficients of the solve timeline, or in the
                        // compact layout [2*npad+256][CW]
    long long dinv;     // [npad] 1/d_k of G = L D L^T (zero in the padding)
    // separator form of the solve (DevCase::dissect): two half systems, each
    // factored as above, and the dense Schur complement of the separator
    long long coef2;    // [2][2*np2+256][CW] solve timelines of the two halves (compact layout)
    long long dinv2;    // [2][np2]
    long long ltail2;   // [2][bw][bw+1] last bw rows of each half's factor, l(i, i-d)
    long long zsep;     // separator step: M [bw][3bw] (x_S = M v), then N [2bw][bw] (tail corrections = N x_S)
    long long rhoadj;   // [cap] rho at the nodes
    long long glx, gly; // [cap] velocity corrections
    long long ust, vst; // [cap] u*, v* at the nodes (moments of the post-stream populations)
    long long total;
};

// All device pointers and scalars of one case.  Kernels read it from a
// device table of the cases they work on (one for a single handle).
// The immersed-boundary geometry of step n (node set, kernel weights, factor
// of D D^T) lives in slot (n-1) % nslot.  Geometry depends only on the
// prescribed kinematics, so it is built in batches of steps ahead of time.
struct DevCase {
    int nx, ny;
    long long N;            // nx*ny
    double ox, oy, scale, U, xc, yc;
    Rates R;
    int smooth;
    int nw, ppw, spw, nseg; // wings, points per wing, segments per wing, all segments
    int cap;                // IB node capacity
    int bw;                 // lower bandwidth of the zig-zag ordered IB matrix
    int ws;                 // rows per lane group of the triangular solves (32 or 64)
    int npad;               // nseg rounded up to a multiple of 64
    int dissect;            // 1: solve through the separator form (two concurrent half chains)
    int compact;            // 1: solve coefficients in the compact layout (bw < CW)
    int mT, mB, np2;        // separator form: rows of the top half, of the bottom half (taken in
                            // reverse), and the halves' padded length
    int steps_cap;          // rows of kin / hist
    int nslot;
    int batch;              // steps per geometry batch (nslot = 2 * batch)
    int gi0, gj0, gw, gh;   // lattice window (0-based origin, size) holding every IB node
    double cs;              // sound speed of the characteristic boundary, REAL(4) sqrt(3)/3
    // lattice (population-planar, x fastest): FP64 populations f, or in the
    // FP32 deviation form g = f - W
    int prec;               // LBMIB_PREC_FP64 / LBMIB_PREC_FP32DEV
    double* f[3];           // [0], [1]: the state planes; [2]: the middle step of two-step passes (or null)
    float* g[3];
    int pofs;               // plane parity: the state after step m lies in plane (m + pofs) & 1
    double W[9];            // the reference's REAL(4) weights, as doubles
    double rho0;            // sum of W
    DevConst K;             // FP32 collision constants
    double* f0L;            // stale boundary-ring f0 of column 1 and column nx: [9][ny]
    double* f0R;
    double* macro[2];       // rho, u, v planes (3*N) of the last two flagged steps
    // marker input
    const double* xb0;
    const double* yb0;
    const double* kin;      // [steps_cap][nw][KIN]
    const int* perm;        // [nseg] zig-zag position -> segment
    const int* iperm;       // [nseg] segment -> zig-zag position
    // per-step slots
    int* ipool;             // nslot * SlotInts::total
    double* dpool;          // nslot * SlotDoubles::total
    SlotInts io;
    SlotDoubles dof;
    // per-step solve (one step at a time)
    double* rhs;            // [2][nseg] U_B - D u*, zig-zag order
    double* y;              // [2][nseg] (D D^T)^-1 rhs, segment order
    double* hist;           // [steps_cap][3*nw] C_x, C_y, C_M per wing
    double* raw;            // [steps_cap][nw][3] raw force x, y and moment per wing (dynamics only), or null
    const double* body;     // [nw][4] body offset terms of the next step: body_extra_s/(1/scale),
                            // movenum*scale, body_extra_v, body_extra_pos (zero with the dynamics off)
    int* err;               // [1] ErrBits
    double* red;            // reduction scratch
};

// Steps of one case inside one launch window (one lbmib_step or batch call):
// a kernel launched at window offset `off` works on step n = n0 + off of the
// case, skips the case when n > n_end, and records rho/u/v when
// n >= flag_from.  Cases of one window may stand at different steps, but all
// at the same position inside a geometry batch.
struct CaseClock {
    int n0, n_end, flag_from;
    int fresh;     // 1: no geometry built for this case's steps yet (new run or restart)
    int next_due;  // 1: the batch after the one holding step n0 is not built yet: build it at offset 0
};

// Geometry launch modes: GEO_FIRST builds the batch of a fresh case's first
// step; GEO_NEXT builds the batch after the current one.
enum GeoMode { GEO_FIRST = 0, GEO_NEXT = 1 };

// One slot's arrays.
struct Slot {
    int *claim, *nodeidx, *box, *draw_idx, *nraw_s, *adj, *nadj, *dcnt, *didx, *nmkcnt, *nmk;
    double *xb2, *yb2, *ub, *vb, *phi, *draw_val, *nraw_w, *dval, *nmkv, *G, *coef, *dinv, *rhoadj, *glx, *gly;
    double *coef2, *dinv2, *ltail2, *zsep, *ust, *vst;
};

__host__ __device__ inline Slot slot_of(const DevCase& D, int g)
{
    Slot S;
    int* ib = D.ipool + (long long)g * D.io.total;
    double* db = D.dpool + (long long)g * D.dof.total;
    S.claim = ib + D.io.claim; S.nodeidx = ib + D.io.nodeidx; S.box = ib + D.io.box;
    S.draw_idx = ib + D.io.draw_idx; S.nraw_s = ib + D.io.nraw_s; S.adj = ib + D.io.adj;
    S.nadj = ib + D.io.nadj; S.dcnt = ib + D.io.dcnt; S.didx = ib + D.io.didx;
    S.nmkcnt = ib + D.io.nmkcnt; S.nmk = ib + D.io.nmk;
    S.xb2 = db + D.dof.xb2; S.yb2 = db + D.dof.yb2; S.ub = db + D.dof.ub; S.vb = db + D.dof.vb;
    S.phi = db + D.dof.phi; S.draw_val = db + D.dof.draw_val; S.nraw_w = db + D.dof.nraw_w;
    S.dval = db + D.dof.dval; S.nmkv = db + D.dof.nmkv; S.G = db + D.dof.G; S.coef = db + D.dof.coef;
    S.dinv = db + D.dof.dinv; S.rhoadj = db + D.dof.rhoadj; S.glx = db + D.dof.glx; S.gly = db + D.dof.gly;
    S.coef2 = db + D.dof.coef2; S.dinv2 = db + D.dof.dinv2; S.ltail2 = db + D.dof.ltail2;
    S.zsep = db + D.dof.zsep;
    S.ust = db + D.dof.ust; S.vst = db + D.dof.vst;
    return S;
}

// Planes a step reads and writes.  psel < 0: the case's own parity (the
// state of step m in plane (m + pofs) & 1); otherwise the launch names them,
// src = psel & 3 and dst = psel >> 2 (an engine of one case that runs two-step
// passes through the middle plane 2).
__host__ __device__ inline void step_planes(const DevCase& D, int n, int psel, int& src, int& dst)
{
    if (psel < 0) {
        dst = (n + D.pofs) & 1;
        src = dst ^ 1;
    } else {
        src = psel & 3;
        dst = (psel >> 2) & 3;
    }
}
__host__ __device__ inline int make_psel(int src, int dst) { return src | (dst << 2); }

// ---- kernels (lbmib_kernels.hip); C: case table, K: clocks of the window ----
template <typename T, int NPT>
__global__ void k_lattice(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                          const int2* __restrict__ tiles, int ntiles, int off, int psel);
// two-step far-field passes and the windows' single steps.  Each launch
// covers every case of the engine through a table of workgroups: far tiles
// {case, a | b << 16, j0, j1} (output columns [a, b), rows j0..j1), and
// {case, block} for the window, ring and window + ring launches (in the
// last, blocks below the case's window block count step the window, the
// rest the ring).
constexpr int FAR_LANES = 256;             // phase-1 columns per strip
constexpr int FAR_OUT = FAR_LANES - 2;     // output columns per strip (FP64: one column per lane)
constexpr int FAR_OUT2 = 2 * FAR_LANES - 4;  // FP32: two columns per lane, strips start at even columns
template <typename T>
__global__ void k_far2(const DevCase* __restrict__ C, const CaseClock* __restrict__ K, const int4* __restrict__ tiles,
                       int ntiles, int off, int psel);
template <typename T>
__global__ void k_window(const DevCase* __restrict__ C, const CaseClock* __restrict__ K, const int2* __restrict__ tab,
                         int off, int psel);
template <typename T>
__global__ void k_ring(const DevCase* __restrict__ C, const CaseClock* __restrict__ K, const int2* __restrict__ tab,
                       int off, int psel);
template <typename T>
__global__ void k_window_ring(const DevCase* __restrict__ C, const CaseClock* __restrict__ K,
                              const int2* __restrict__ tab, int off, int psel, int ring_psel);
__host__ __device__ inline int window_blocks(int gw, int gh) { return (int)(((long long)gw * gh + 255) / 256); }
size_t far2_smem_bytes(bool f64);
__host__ __device__ inline int lattice_ring_blocks(int nx, int ny) { return (ny - 2 + nx + 255) / 256; }
// workgroups of k_ring: one ring node per thread
__host__ __device__ inline int ring_blocks1(int nx, int ny) { return (2 * (ny - 2) + 2 * nx + 255) / 256; }
// workgroups of 256 threads over the layer next to the outer ring
__host__ __device__ inline int layer_blocks(int nx, int ny) { return (2 * (ny - 2) + 2 * (nx - 4) + 255) / 256; }
// first node of the interior tiles: row 1 rounded down to a multiple of 64 nodes
__host__ __device__ inline int lattice_base(int nx) { return nx / 64 * 64; }
int lattice_tiles(int nx, int ny, int npt);  // ring blocks + interior tiles of one case
template <typename T> __global__ void k_fill(T* __restrict__ p, long long n, T v);
__global__ void k_geo_clear(const DevCase* __restrict__, const CaseClock* __restrict__, int, int);
__global__ void k_geo_markers(const DevCase* __restrict__, const CaseClock* __restrict__, int, int);
__global__ void k_geo_fill(const DevCase* __restrict__, const CaseClock* __restrict__, int, int);
__global__ void k_geo_phi(const DevCase* __restrict__, const CaseClock* __restrict__, int, int);
__global__ void k_geo_pairs(const DevCase* __restrict__, const CaseClock* __restrict__, int, int);
__global__ void k_geo_dsort(const DevCase* __restrict__, const CaseClock* __restrict__, int, int);
__global__ void k_geo_nsort(const DevCase* __restrict__, const CaseClock* __restrict__, int, int);
__global__ void k_geo_assemble(const DevCase* __restrict__, const CaseClock* __restrict__, int, int);
__global__ void k_geo_factor(const DevCase* __restrict__, const CaseClock* __restrict__, int, int);
__global__ void k_geo_factor2(const DevCase* __restrict__, const CaseClock* __restrict__, int, int);
__global__ void k_geo_schur(const DevCase* __restrict__, const CaseClock* __restrict__, int, int);
template <typename T>
__global__ void k_crit_ustar(const DevCase* __restrict__, const CaseClock* __restrict__, int, int);
__global__ void k_crit_rhs(const DevCase* __restrict__, const CaseClock* __restrict__, int);
template <typename T>
__global__ void k_crit_rhs1(const DevCase* __restrict__, const CaseClock* __restrict__, int, int);
template <int U, bool CMP>
__global__ void k_crit_solve(const DevCase* __restrict__, const CaseClock* __restrict__, int);
int solve_unit(int npad);
int solve_threads();
template <int U>
__global__ void k_crit_solve2(const DevCase* __restrict__, const CaseClock* __restrict__, int);
int solve2_unit(int np2);  // 0 when the values leave no room for the rings
int solve2_threads();
size_t solve2_smem_bytes(int np2);
template <typename T>
__global__ void k_band(const DevCase* __restrict__, const CaseClock* __restrict__, int, int);
__global__ void k_forces(const DevCase* __restrict__, const CaseClock* __restrict__, int);
__global__ void k_clock_advance(CaseClock* __restrict__, int ncases, int steps);
size_t factor_smem_bytes();
size_t solve_smem_bytes(int npad);
__global__ void k_residual(const double*, const double*, int, int, long long, double*);
struct RingW { double w[9]; };  // the weights, by value
template <typename T>
__global__ void k_move_pops(const T*, T*, int, long long, int);
template <typename T>
__global__ void k_move_ring(const T*, double*, int, int, long long, int, RingW);
__global__ void k_move_macro(double*, double*, int, long long, int, int);
__global__ void k_vorticity(const double*, int, int, long long, double, double, double*, double*, double*, double*);
__global__ void k_surface_pressure(const int*, const int*, const double*, const double*, int, double*);
template <typename T> __global__ void k_nan_check(const T*, long long, int*);

}  // namespace lbmib
