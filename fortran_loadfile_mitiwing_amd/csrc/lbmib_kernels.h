// lbmib_kernels.h -- device-side data of one lbmib case and the kernel
// declarations shared by the HIP translation units of liblbmib.
#pragma once

#include <hip/hip_runtime.h>

#include "lbmib_exact.h"

namespace lbmib {

constexpr int BOXMAX = 36;  // nodes in one marker's +-2.5 lattice-unit box (<= 6x6)
constexpr int NCAP = 64;    // markers whose boxes may contain one IB node
constexpr int KIN = 8;      // per wing per step: sin, cos, omega, sa, sau, sb, sbv, pad

enum ErrBits { ERR_GEOMETRY = 1, ERR_CAPACITY = 2, ERR_SOLVE = 4 };

// All device pointers and scalars of one case.  Passed by value to kernels.
struct DevCase {
    int nx, ny;
    long long N;            // nx*ny
    double ox, oy, scale, U, xc, yc;
    Rates R;
    int smooth;
    int nw, ppw, spw, nseg; // wings, points per wing, segments per wing, all segments
    int cap;                // IB node capacity
    int bw;                 // lower bandwidth of the zig-zag ordered IB matrix
    int steps_cap;          // rows of kin / hist
    // lattice (population-planar, x fastest)
    double* f[2];
    double* f0L;            // stale boundary-ring f0 of column 1 and column nx: [9][ny]
    double* f0R;
    double* macro[2];       // rho, u, v planes (3*N) of the last two flagged steps
    // immersed boundary
    const double* xb0;
    const double* yb0;
    const double* kin;      // [steps_cap][nw][KIN]
    double *xb2, *yb2, *ub, *vb;
    int* box;               // [nseg][4] i1,i2,j1,j2 (1-based)
    int* claim;             // [N] first claiming segment, INT_MAX when free
    int* nodeidx;           // [N] IB list index or -1
    int* segcnt;            // [nseg] nodes first claimed by each segment
    int* adj;               // [cap][3] node i, j (0-based), wing
    int* nadj;              // [1] node count of the current set
    double *ustar, *vstar, *rhoadj;
    int* dcnt;              // [nseg] nodes in the segment's box
    int* didx;              // [nseg][BOXMAX] their list indices, ascending
    double* dval;           // [nseg][BOXMAX] kernel weights phi_x*phi_y
    int* nmkcnt;            // [cap] markers whose box holds the node
    int* nmk;               // [cap][NCAP] those markers (sorted before use)
    double* nmkv;           // [cap][NCAP] their weights
    const int* perm;        // [nseg] zig-zag position -> segment
    const int* iperm;       // [nseg] segment -> zig-zag position
    double* rhs;            // [2][nseg] U_B - D u*, zig-zag order
    double* G;              // [nseg][bw+1] band of D D^T in zig-zag order, G[p][d] = A(p,p-d)
    double* y;              // [2][nseg] (D D^T)^-1 rhs, segment order
    double *glx, *gly;      // [cap] velocity corrections
    double* hist;           // [steps_cap][3*nw] C_x, C_y, C_M per wing
    int* err;               // [1] ErrBits
    double* red;            // reduction scratch
};

}  // namespace lbmib
