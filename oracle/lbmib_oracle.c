/*
 * lbmib_oracle.c -- CPU restatement of the reference solver's per-timestep hot
 * path.  TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
 * product library.
 *
 * Conventions
 *   - Arrays are Fortran column-major with 0-based indices here: node (i,j)
 *     (Fortran (i+1,j+1)) is cell c = j*nx + i; populations are planar,
 *     f[k*nx*ny + c] is Fortran f(i+1,j+1,k+1).
 *   - Every floating-point expression is written in the order the Fortran
 *     source writes it and evaluated left to right.  Build with
 *     -ffp-contract=off -fno-fast-math.
 *   - Terms that the reference multiplies by an exact zero (e.g. e(k,1)=0 in
 *     the moment sums, zero entries of M, structurally zero entries of the IB
 *     matrices) are skipped: x + (+-0) == x in IEEE arithmetic, so skipping is
 *     bit-identical.  The "dense" IB mode keeps every term literally and the
 *     tests check that both modes agree bit for bit.
 *
 * Parity notes (round-off-level choices where the reference calls closed
 * libraries whose operation order is not visible):
 *   - DGEMM/DGEMV (boundary_condition.f90:1619-1621,1639-1640): sequential
 *     dot products in index order.
 *   - DGETRF/DGETRS (boundary_condition.f90:1632-1638): solver=1 restates them
 *     as unblocked right-looking LU with partial pivoting (first maximal
 *     pivot), division by the pivot, column-oriented triangular solves.
 *   - solver=0 (default) solves the mathematically identical SPD system
 *     (D D^T) y = b, since A = D D^T S and delta_u = D^T S A^-1 b = D^T y
 *     (S cancels), by square-root-free Cholesky (L D L^T) in a zig-zag
 *     marker order (LE outwards along both surfaces), pivots applied
 *     through their reciprocals.  The tests pin it against solver=1 and
 *     numpy.
 *   - asin (delta kernel) is evaluated with fdlibm's published e_asin method
 *     (rational approximation, Sun Microsystems 1993), implemented with plain
 *     IEEE operations so that the GPU evaluates the identical sequence.
 *   - sin/cos of the pitch angle use the host libm; the product computes the
 *     same per-step values with the same host libm before upload.
 */
#include "lbmib_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define Q 9

/* D2Q9 lattice vectors e(k,1:2) (D2Q9model.f90:5-22), k 0-based */
static const int EX[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int EY[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

/* MRT transform matm (solution.f90:17-25), row-major */
static const double MATM[Q][Q] = {
    {1, 1, 1, 1, 1, 1, 1, 1, 1},
    {-4, -1, -1, -1, -1, 2, 2, 2, 2},
    {4, -2, -2, -2, -2, 1, 1, 1, 1},
    {0, 1, 0, -1, 0, 1, -1, -1, 1},
    {0, -2, 0, 2, 0, 1, -1, -1, 1},
    {0, 0, 1, 0, -1, 1, 1, -1, -1},
    {0, 0, -2, 0, 2, 1, 1, -1, -1},
    {0, 1, -1, 1, -1, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 1, -1, 1, -1}};

/* pai (main.f90:9) */
static const double PAI = 3.1415926535897932384626433832795;

/* ------------------------------------------------------------------ */
/* constants                                                           */
/* ------------------------------------------------------------------ */

void orc_weights(double w[9])
{
    /* w(k)=4.0/9.0 ... are default-REAL (single precision) expressions
     * assigned to REAL(8) (D2Q9model.f90:24-32). */
    w[0] = (double)(4.0f / 9.0f);
    for (int k = 1; k <= 4; ++k) w[k] = (double)(1.0f / 9.0f);
    for (int k = 5; k <= 8; ++k) w[k] = (double)(1.0f / 36.0f);
}

double orc_cs(void)
{
    /* cs=sqrt(3.0)/3.0d0: single-precision sqrt (boundary_condition.f90:304) */
    return (double)sqrtf(3.0f) / 3.0;
}

/* nizhen (solution.f90:1462-1517): Gauss-Jordan with partial pivoting.
 * a, b row-major Nm x Nm; b returns the inverse. */
static void nizhen(const double* AA, double* b1, int Nm)
{
    double a[Q * Q];
    memcpy(a, AA, sizeof(double) * Nm * Nm);
#define A_(i, j) a[(i) * Nm + (j)]
#define B_(i, j) b1[(i) * Nm + (j)]
    for (int i = 0; i < Nm * Nm; ++i) b1[i] = 0.0;
    for (int k = 0; k < Nm; ++k) B_(k, k) = 1.0;
    for (int k = 0; k < Nm - 1; ++k) {
        int ik = k;
        double temp = fabs(A_(ik, k));
        for (int i = k; i < Nm; ++i) {
            if (fabs(A_(i, k)) > fabs(temp)) {
                temp = fabs(A_(i, k));
                ik = i;
            }
        }
        for (int j = 0; j < Nm; ++j) {
            temp = A_(ik, j);
            A_(ik, j) = A_(k, j);
            A_(k, j) = temp;
            temp = B_(ik, j);
            B_(ik, j) = B_(k, j);
            B_(k, j) = temp;
        }
        for (int i = 0; i < Nm; ++i) {
            double mik = A_(i, k) / A_(k, k);
            temp = A_(k, k);
            for (int j = k; j < Nm; ++j)
                if (i != k) A_(i, j) = A_(i, j) - mik * A_(k, j);
            for (int j = k; j < Nm; ++j) A_(k, j) = A_(k, j) / temp;
            for (int j = 0; j < Nm; ++j)
                if (i != k) B_(i, j) = B_(i, j) - mik * B_(k, j);
            for (int j = 0; j < Nm; ++j) B_(k, j) = B_(k, j) / temp;
        }
    }
    {
        double temp = A_(Nm - 1, Nm - 1);
        for (int i = 0; i < Nm - 1; ++i) {
            double mik = A_(i, Nm - 1) / A_(Nm - 1, Nm - 1);
            A_(i, Nm - 1) = 0.0;
            for (int j = 0; j < Nm; ++j) B_(i, j) = B_(i, j) - mik * B_(Nm - 1, j);
        }
        A_(Nm - 1, Nm - 1) = 1.0;
        for (int j = 0; j < Nm; ++j) B_(Nm - 1, j) = B_(Nm - 1, j) / temp;
    }
#undef A_
#undef B_
}

void orc_minv(double minv[81]) { nizhen(&MATM[0][0], minv, Q); }

/* matmulvoc (boundary_condition.f90:1738-1755): c = a*b, sequential j. */
static inline void matmulvoc(double* c, const double* a /* row-major */, const double* b)
{
    for (int i = 0; i < Q; ++i) {
        double sum = 0.0;
        for (int j = 0; j < Q; ++j) {
            double aij = a[i * Q + j];
            if (aij != 0.0) sum = sum + aij * b[j]; /* exact-zero terms skipped */
        }
        c[i] = sum;
    }
}

/* D2Q9feq (D2Q9model.f90:35-46), k 0-based */
double orc_feq(double rh, double u1, double v1, int k)
{
    double w[Q];
    orc_weights(w);
    double eu = (double)EX[k] * u1 + (double)EY[k] * v1;
    double utwo = u1 * u1 + v1 * v1;
    return w[k] * rh * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * utwo);
}

/* ------------------------------------------------------------------ */
/* delta kernel                                                        */
/* ------------------------------------------------------------------ */

/* asin by fdlibm's e_asin method (x + x^3 R(x^2) on |x|<0.5; the
 * pi/2 - 2 asin(sqrt((1-x)/2)) reduction with a split sqrt above). */
static const double AS_pio2_hi = 1.57079632679489655800e+00;
static const double AS_pio2_lo = 6.12323399573676603587e-17;
static const double AS_pio4_hi = 7.85398163397448278999e-01;
static const double AS_pS0 = 1.66666666666666657415e-01;
static const double AS_pS1 = -3.25565818622400915405e-01;
static const double AS_pS2 = 2.01212532134862925881e-01;
static const double AS_pS3 = -4.00555345006794114027e-02;
static const double AS_pS4 = 7.91534994289814532176e-04;
static const double AS_pS5 = 3.47933107596021167570e-05;
static const double AS_qS1 = -2.40339491173441421878e+00;
static const double AS_qS2 = 2.02094576023350569471e+00;
static const double AS_qS3 = -6.88283971605453293030e-01;
static const double AS_qS4 = 7.70381505559019352791e-02;

double orc_asin(double x)
{
    double ax = fabs(x);
    if (!(ax <= 1.0)) return NAN;
    if (ax == 1.0) return x > 0 ? AS_pio2_hi + AS_pio2_lo : -(AS_pio2_hi + AS_pio2_lo);
    if (ax < 0.5) {
        if (ax < 7.450580596923828125e-9) return x; /* 2^-27 */
        double t = x * x;
        double p = t * (AS_pS0 + t * (AS_pS1 + t * (AS_pS2 + t * (AS_pS3 + t * (AS_pS4 + t * AS_pS5)))));
        double q = 1.0 + t * (AS_qS1 + t * (AS_qS2 + t * (AS_qS3 + t * AS_qS4)));
        double w = p / q;
        return x + x * w;
    }
    double w = 1.0 - ax;
    double t = w * 0.5;
    double p = t * (AS_pS0 + t * (AS_pS1 + t * (AS_pS2 + t * (AS_pS3 + t * (AS_pS4 + t * AS_pS5)))));
    double q = 1.0 + t * (AS_qS1 + t * (AS_qS2 + t * (AS_qS3 + t * AS_qS4)));
    double s = sqrt(t);
    double r;
    if (ax >= 0.975) {
        w = p / q;
        r = AS_pio2_hi - (2.0 * (s + s * w) - AS_pio2_lo);
    } else {
        uint64_t bits;
        memcpy(&bits, &s, 8);
        bits &= 0xFFFFFFFF00000000ULL; /* high word of s */
        double wh;
        memcpy(&wh, &bits, 8);
        double c = (t - wh * wh) / (s + wh);
        double rr = p / q;
        double pp = 2.0 * s * rr - (AS_pio2_lo - 2.0 * c);
        double qq = AS_pio4_hi - 2.0 * wh;
        r = AS_pio4_hi - (pp - qq);
    }
    return x > 0 ? r : -r;
}

/* One-axis kernel value for disr = |x0-xb2|/scale >= 0
 * (boundary_condition.f90:1563-1600). */
double orc_delta(double disr, int smooth)
{
    if (smooth == 0) {
        if (fabs(disr) < 1.0)
            return 0.125 * (3.0 - 2.0 * fabs(disr) + sqrt(1.0 + 4.0 * fabs(disr) - 4.0 * disr * disr));
        else if ((fabs(disr) < 2.0) && (fabs(disr) >= 1.0))
            return 0.125 * (5.0 - 2.0 * fabs(disr) - sqrt(-7.0 + 12.0 * fabs(disr) - 4.0 * disr * disr));
        return 0.0;
    }
    if (fabs(disr) <= 0.5)
        return 0.375 + 0.03125 * PAI - 0.25 * disr * disr;
    else if ((fabs(disr) <= 1.5) && (fabs(disr) >= 0.5))
        return 0.25 + 0.125 * (1.0 - disr) * sqrt(-2.0 + 8.0 * disr - 4.0 * disr * disr)
               - 0.125 * orc_asin(sqrt(2.0) * (disr - 1.0));
    else if ((fabs(disr) <= 2.5) && (fabs(disr) >= 1.5))
        return 1.0625 - 0.015625 * PAI - 0.75 * disr + 0.125 * disr * disr
               + 0.0625 * (disr - 2.0) * sqrt(-14.0 + 16.0 * disr - 4.0 * disr * disr)
               + 0.0625 * orc_asin(sqrt(2.0) * (disr - 2.0));
    return 0.0;
}

/* ------------------------------------------------------------------ */
/* kinematics tables (mode.f90)                                        */
/* ------------------------------------------------------------------ */

typedef struct motion_tab {
    const double* t; /* rows x ncols */
    int rows, ncols;
    int index;       /* motion_index, 1-based */
} motion_tab;

#define MT(m, r, c) ((m)->t[(size_t)((r) - 1) * (m)->ncols + (c) - 1]) /* 1-based */

/* interpolate_motion_row (mode.f90:176-197) */
static void interpolate_row(const motion_tab* m, int i1, int i2, double target, double* row)
{
    double t1 = MT(m, i1, 1), t2 = MT(m, i2, 1);
    if (fabs(t2 - t1) <= 1.0e-14) {
        for (int c = 1; c <= m->ncols; ++c) row[c - 1] = MT(m, i1, c);
        row[0] = target;
        return;
    }
    double weight = (target - t1) / (t2 - t1);
    for (int c = 1; c <= m->ncols; ++c)
        row[c - 1] = (1.0 - weight) * MT(m, i1, c) + weight * MT(m, i2, c);
    row[0] = target;
}

/* lookup_motion_row (mode.f90:199-243) */
static void lookup_row(motion_tab* m, double target, double* row)
{
    int n = m->rows;
    if (target <= MT(m, 1, 1)) {
        m->index = 1;
        for (int c = 1; c <= m->ncols; ++c) row[c - 1] = MT(m, 1, c);
        row[0] = target;
        return;
    }
    if (target >= MT(m, n, 1)) {
        m->index = n - 1 > 1 ? n - 1 : 1;
        for (int c = 1; c <= m->ncols; ++c) row[c - 1] = MT(m, n, c);
        row[0] = target;
        return;
    }
    int idx = m->index;
    if (idx > n - 1) idx = n - 1;
    if (idx < 1) idx = 1;
    while (idx > 1 && target < MT(m, idx, 1)) idx--;
    while (idx < n - 1 && target > MT(m, idx + 1, 1)) idx++;
    m->index = idx;
    if (fabs(target - MT(m, idx, 1)) <= 1.0e-12) {
        for (int c = 1; c <= m->ncols; ++c) row[c - 1] = MT(m, idx, c);
        row[0] = target;
        return;
    }
    if (fabs(target - MT(m, idx + 1, 1)) <= 1.0e-12) {
        for (int c = 1; c <= m->ncols; ++c) row[c - 1] = MT(m, idx + 1, c);
        row[0] = target;
        return;
    }
    interpolate_row(m, idx, idx + 1, target, row);
}

static int step_count_for(double end_time, double dt)
{
    /* motion_step_count=int((motion_end_time+1.0d-12)/motion_dt)+1 (mode.f90:258) */
    return (int)((end_time + 1.0e-12) / dt) + 1;
}

int orc_motion_step_table(const double* motion, int rows, int ncols, double dt,
                          double* out, int max_steps)
{
    if (rows < 1 || dt <= 0.0) return -1;
    motion_tab m = {motion, rows, ncols, 1};
    double end_time = motion[(size_t)(rows - 1) * ncols];
    int count = step_count_for(end_time, dt);
    if (out == NULL) return count;
    if (count > max_steps) return -2;
    double row[1 + 9 * 8];
    for (int i = 1; i <= count; ++i) {
        double target = (double)(i - 1) * dt;
        lookup_row(&m, target, row);
        memcpy(out + (size_t)(i - 1) * ncols, row, sizeof(double) * ncols);
    }
    return count;
}

/* ------------------------------------------------------------------ */
/* case state                                                          */
/* ------------------------------------------------------------------ */

#define MAXW 8

struct orc_case {
    orc_params p;
    int nx, ny;
    size_t N;
    double w[Q], minv[Q * Q], cs;
    double *f, *f0, *rho, *u, *v, *u0, *v0, *pre, *gforce;
    int* body;
    int* nodeidx; /* IB list index per cell or -1 (sparse bookkeeping) */
    /* markers */
    int npts, nseg, segs_per_wing;
    double *xb0, *yb0, *xb1, *yb1, *xb2, *yb2, *arc_l, *ub, *vb;
    /* IB node set */
    int nadj, cap;
    int *adj_i, *adj_j, *adj_w;
    int *box; /* per segment i1,i2,j1,j2 (1-based) */
    double *glx, *gly, *rho_adj;
    /* kinematics */
    double* step_table;
    int step_count, ncols;
    double motion_end_time;
    double angle[MAXW], omega[MAXW], omgac[MAXW], sa[MAXW], sau[MAXW], saac[MAXW], sb[MAXW],
        sbv[MAXW], sbac[MAXW];
    double wing_force[MAXW][2], wing_moment[MAXW];
    /* rigid-body dynamics (main.f90:60-66; index 0 is wing 1) and the
     * domain shift count movenum (main.f90:121) */
    orc_dynamics dyn;
    double rel_a[MAXW], rel_a_prev[MAXW], rel_v[MAXW], rel_v_prev[MAXW], rel_s[MAXW], rel_s_prev[MAXW],
        rel_pos[MAXW], rel_theta[MAXW], rel_theta_prev[MAXW], rel_tt[MAXW], rel_tt_prev[MAXW];
    double extra_v[MAXW], extra_s[MAXW], extra_pos[MAXW];
    double mom; /* com_date's mom: the raw moment of wing 1 in the last IB call */
    int movenum;
    /* history: force rows and airfoil_motion rows */
    double* hist;
    double* mrows;
    int hist_cols, hist_rows, hist_cap;
    int n; /* steps done */
    double err;
};

static void set_err(char* err, const char* msg)
{
    if (err) snprintf(err, 256, "%s", msg);
}

static int omp_threads(const orc_case* c)
{
    return c->p.n_threads > 0 ? c->p.n_threads : 1;
}

void orc_destroy(orc_case* c)
{
    if (!c) return;
    free(c->f); free(c->f0); free(c->rho); free(c->u); free(c->v); free(c->u0); free(c->v0);
    free(c->pre); free(c->gforce); free(c->body); free(c->nodeidx);
    free(c->xb0); free(c->yb0); free(c->xb1); free(c->yb1); free(c->xb2); free(c->yb2);
    free(c->arc_l); free(c->ub); free(c->vb);
    free(c->adj_i); free(c->adj_j); free(c->adj_w); free(c->box);
    free(c->glx); free(c->gly); free(c->rho_adj);
    free(c->step_table); free(c->hist); free(c->mrows);
    free(c);
}

orc_case* orc_create(const orc_params* p, const double* xb0, const double* yb0,
                     const double* motion, int rows, int ncols, char* err)
{
    if (p->nx < 8 || p->ny < 8) { set_err(err, "lattice too small"); return NULL; }
    if (p->n_wings < 1 || p->n_wings > MAXW) { set_err(err, "wing_count out of range"); return NULL; }
    if (p->pts_per_wing < 3) { set_err(err, "too few boundary points"); return NULL; }
    if (ncols != 1 + 9 * p->n_wings) { set_err(err, "motion table column count must be 1+9*n_wings"); return NULL; }
    if (rows < 1) { set_err(err, "motion table is empty"); return NULL; }
    for (int r = 1; r < rows; ++r)
        if (motion[(size_t)r * ncols] < motion[(size_t)(r - 1) * ncols]) {
            set_err(err, "motion table time column must be non-decreasing");
            return NULL;
        }
    if (!(p->scale > 0.0)) { set_err(err, "scale must be positive"); return NULL; }

    orc_case* c = (orc_case*)calloc(1, sizeof(orc_case));
    c->p = *p;
    c->nx = p->nx;
    c->ny = p->ny;
    c->N = (size_t)p->nx * p->ny;
    size_t N = c->N;
    orc_weights(c->w);
    orc_minv(c->minv);
    c->cs = orc_cs();
    c->f = (double*)malloc(sizeof(double) * Q * N);
    c->f0 = (double*)malloc(sizeof(double) * Q * N);
    c->rho = (double*)malloc(sizeof(double) * N);
    c->u = (double*)malloc(sizeof(double) * N);
    c->v = (double*)malloc(sizeof(double) * N);
    c->u0 = (double*)malloc(sizeof(double) * N);
    c->v0 = (double*)malloc(sizeof(double) * N);
    c->pre = (double*)malloc(sizeof(double) * N);
    c->gforce = (double*)calloc(2 * N, sizeof(double));
    c->body = (int*)calloc(N, sizeof(int));
    c->nodeidx = (int*)malloc(sizeof(int) * N);
    for (size_t i = 0; i < N; ++i) c->nodeidx[i] = -1;

    /* init (solution.f90:34-59): u=u0=UU, v=v0=0, rho=1, pre=1/3, f=f0=feq */
    double feq0[Q];
    for (int k = 0; k < Q; ++k) feq0[k] = orc_feq(1.0, p->U, 0.0, k);
    for (size_t i = 0; i < N; ++i) {
        c->u[i] = p->U; c->u0[i] = p->U; c->v[i] = 0.0; c->v0[i] = 0.0;
        c->rho[i] = 1.0; c->pre[i] = 1.0 / 3.0;
    }
    for (int k = 0; k < Q; ++k)
        for (size_t i = 0; i < N; ++i) { c->f[k * N + i] = feq0[k]; c->f0[k * N + i] = feq0[k]; }

    /* read_boundary (airfoil.f90:216-241) */
    int nw = p->n_wings, ppw = p->pts_per_wing;
    c->npts = nw * ppw;
    c->segs_per_wing = ppw - 1;
    c->nseg = nw * (ppw - 1);
    c->xb0 = (double*)malloc(sizeof(double) * c->npts);
    c->yb0 = (double*)malloc(sizeof(double) * c->npts);
    c->xb1 = (double*)malloc(sizeof(double) * c->npts);
    c->yb1 = (double*)malloc(sizeof(double) * c->npts);
    c->xb2 = (double*)malloc(sizeof(double) * c->nseg);
    c->yb2 = (double*)malloc(sizeof(double) * c->nseg);
    c->arc_l = (double*)malloc(sizeof(double) * c->nseg);
    c->ub = (double*)malloc(sizeof(double) * c->nseg);
    c->vb = (double*)malloc(sizeof(double) * c->nseg);
    c->box = (int*)malloc(sizeof(int) * 4 * c->nseg);
    memcpy(c->xb0, xb0, sizeof(double) * c->npts);
    memcpy(c->yb0, yb0, sizeof(double) * c->npts);
    for (int wg = 0; wg < nw; ++wg) {
        int ps = wg * ppw, ss = wg * (ppw - 1);
        for (int i0 = 0; i0 < ppw - 1; ++i0) {
            double dx = c->xb0[ps + i0 + 1] - c->xb0[ps + i0];
            double dy = c->yb0[ps + i0 + 1] - c->yb0[ps + i0];
            c->arc_l[ss + i0] = sqrt(dx * dx + dy * dy) / p->scale;
        }
    }
    for (int i = 0; i < c->npts; ++i) {
        c->xb0[i] = c->xb0[i] + (p->xc - p->xp);
        c->yb0[i] = c->yb0[i] + (p->yc - p->yp);
    }
    c->cap = 64 * c->nseg + 64;
    c->adj_i = (int*)malloc(sizeof(int) * c->cap);
    c->adj_j = (int*)malloc(sizeof(int) * c->cap);
    c->adj_w = (int*)malloc(sizeof(int) * c->cap);
    c->glx = (double*)malloc(sizeof(double) * c->cap);
    c->gly = (double*)malloc(sizeof(double) * c->cap);
    c->rho_adj = (double*)malloc(sizeof(double) * c->cap);
    c->nadj = 0;

    /* load_motion_table + build_motion_step_table(scale*UU) (main.f90:254-255) */
    c->ncols = ncols;
    c->motion_end_time = motion[(size_t)(rows - 1) * ncols];
    double dt = p->scale * p->U;
    int count = orc_motion_step_table(motion, rows, ncols, dt, NULL, 0);
    c->step_table = (double*)malloc(sizeof(double) * (size_t)count * ncols);
    orc_motion_step_table(motion, rows, ncols, dt, c->step_table, count);
    c->step_count = count;

    c->hist_cols = 4 + 12 * nw;
    c->hist_cap = 1024;
    c->hist = (double*)malloc(sizeof(double) * c->hist_cap * c->hist_cols);
    c->mrows = (double*)malloc(sizeof(double) * c->hist_cap * (1 + 3 * c->p.n_wings));
    c->n = 0;
    c->err = 0.0;
    return c;
}

int orc_steps_done(const orc_case* c) { return c->n; }

int orc_max_steps(const orc_case* c)
{
    double dt = c->p.scale * c->p.U;
    int n = 1;
    while ((double)(n - 1) * dt <= c->motion_end_time + 1.0e-12) n++;
    return n - 1;
}

/* mode (mode.f90:271-283) -> apply_motion_row (mode.f90:144-174) */
static void mode(orc_case* c, double time)
{
    double dt = c->p.scale * c->p.U;
    long idx = lround(time / dt) + 1;
    if (idx < 1) idx = 1;
    if (idx > c->step_count) idx = c->step_count;
    const double* row = c->step_table + (size_t)(idx - 1) * c->ncols;
    for (int wg = 0; wg < c->p.n_wings; ++wg) {
        int col = 1 + wg * 9;
        c->angle[wg] = row[col + 0];
        c->omega[wg] = row[col + 1];
        c->omgac[wg] = row[col + 2];
        c->sa[wg] = row[col + 3];
        c->sau[wg] = row[col + 4];
        c->saac[wg] = row[col + 5];
        c->sb[wg] = row[col + 6];
        c->sbv[wg] = row[col + 7];
        c->sbac[wg] = row[col + 8];
    }
}

/* the box search of move_boundary (airfoil.f90:290-321), 1-based results */
static void marker_box(const orc_case* c, double xb2, double yb2, int* bi1, int* bi2, int* bj1, int* bj2)
{
    const double ox = c->p.ox, oy = c->p.oy, sc = c->p.scale;
    int nx = c->nx, ny = c->ny;
    int i1 = (int)((xb2 - ox) / sc + 1.0);
    for (int i = i1; i >= 1; --i) {
        double temp = ox + (double)(i - 1) * sc;
        if (fabs((temp - xb2) / sc) > 2.5) { i1 = i + 1; break; }
    }
    int i2 = (int)((xb2 - ox) / sc + 1.0);
    for (int i = i2; i <= nx; ++i) {
        double temp = ox + (double)(i - 1) * sc;
        if (fabs((temp - xb2) / sc) > 2.5) { i2 = i - 1; break; }
    }
    int j2 = (int)((yb2 - oy) / sc + 1.0);
    for (int i = j2; i <= ny; ++i) {
        double temp = oy + (double)(i - 1) * sc;
        if (fabs((temp - yb2) / sc) > 2.5) { j2 = i - 1; break; }
    }
    int j1 = (int)((yb2 - oy) / sc + 1.0);
    for (int i = j1; i >= 1; --i) {
        double temp = oy + (double)(i - 1) * sc;
        if (fabs((temp - yb2) / sc) > 2.5) { j1 = i + 1; break; }
    }
    *bi1 = i1; *bi2 = i2; *bj1 = j1; *bj2 = j2;
}

/* IB node set build (airfoil.f90:286-335) on cleared body flags.  Returns
 * node count or <0 if a node falls outside the interior margin
 * 4..Nx-3 x 4..Ny-3 (1-based) that keeps IB nodes off the outer-boundary
 * stencils, or the list overflows. */
static int build_set(const orc_case* c, const double* xb2, const double* yb2, int* body,
                     int* ai, int* aj, int* aw, int cap, int* box)
{
    int n = 0;
    for (int s = 0; s < c->nseg; ++s) {
        int wing = s / c->segs_per_wing; /* 0-based */
        int i1, i2, j1, j2;
        marker_box(c, xb2[s], yb2[s], &i1, &i2, &j1, &j2);
        if (box) { box[4 * s] = i1; box[4 * s + 1] = i2; box[4 * s + 2] = j1; box[4 * s + 3] = j2; }
        if (i1 < 4 || i2 > c->nx - 3 || j1 < 4 || j2 > c->ny - 3 || i2 < i1 || j2 < j1) return -3;
        for (int i = i1; i <= i2; ++i)
            for (int j = j1; j <= j2; ++j) {
                size_t cc = (size_t)(j - 1) * c->nx + (i - 1);
                if (body[cc] == 0) {
                    if (n >= cap) return -4;
                    body[cc] = 1;
                    ai[n] = i - 1; aj[n] = j - 1; aw[n] = wing;
                    n++;
                }
            }
    }
    return n;
}

int orc_build_ib_set(const orc_case* c, const double* xb2, const double* yb2,
                     int* pi, int* pj, int* wing, int max_nodes)
{
    int* body = (int*)calloc(c->N, sizeof(int));
    int n = build_set(c, xb2, yb2, body, pi, pj, wing, max_nodes, NULL);
    free(body);
    return n;
}

/* move_boundary (airfoil.f90:244-338) */
static int move_boundary(orc_case* c)
{
    int nw = c->p.n_wings, ppw = c->p.pts_per_wing;
    for (int wg = 0; wg < nw; ++wg) {
        int ps = wg * ppw, ss = wg * (ppw - 1);
        double ca = cos(c->angle[wg]), sn = sin(c->angle[wg]);
        for (int i0 = ps; i0 < ps + ppw; ++i0) {
            /* body_extra_s and movenum are zero with the dynamics off (main.f90:103,319);
             * the shift per domain move is mesh_max_scale, the lattice spacing (ini_mesh.f90:21) */
            c->xb1[i0] = ca * c->xb0[i0] + sn * c->yb0[i0] + c->sa[wg] + c->p.xc +
                         c->extra_s[wg] / (1.0 / c->p.scale) - (double)c->movenum * c->p.scale;
            c->yb1[i0] = -sn * c->xb0[i0] + ca * c->yb0[i0] + c->sb[wg] + c->p.yc;
        }
        for (int i0 = ss; i0 < ss + ppw - 1; ++i0) {
            c->xb2[i0] = 0.5 * (c->xb1[ps + i0 - ss + 1] + c->xb1[ps + i0 - ss]);
            c->yb2[i0] = 0.5 * (c->yb1[ps + i0 - ss + 1] + c->yb1[ps + i0 - ss]);
        }
    }
    /* clear the previous node set (airfoil.f90:276-283) */
    for (int a = 0; a < c->nadj; ++a) {
        size_t cc = (size_t)c->adj_j[a] * c->nx + c->adj_i[a];
        c->body[cc] = 0;
        c->gforce[cc] = 0.0;
        c->gforce[c->N + cc] = 0.0;
        c->nodeidx[cc] = -1;
    }
    int n = build_set(c, c->xb2, c->yb2, c->body, c->adj_i, c->adj_j, c->adj_w, c->cap, c->box);
    if (n < 0) return n;
    c->nadj = n;
    for (int a = 0; a < n; ++a) c->nodeidx[(size_t)c->adj_j[a] * c->nx + c->adj_i[a]] = a;
    return 0;
}

/* streaming (solution.f90:428-454): interior pull */
static void streaming(orc_case* c)
{
    int nx = c->nx, ny = c->ny;
    size_t N = c->N;
#pragma omp parallel for num_threads(omp_threads(c)) schedule(static)
    for (int j = 1; j < ny - 1; ++j)
        for (int k = 0; k < Q; ++k) {
            const double* fs = c->f + k * N + (size_t)(j - EY[k]) * nx - EX[k];
            double* fd = c->f0 + k * N + (size_t)j * nx;
            for (int i = 1; i < nx - 1; ++i) fd[i] = fs[i];
        }
}

/* physical_variables (solution.f90:471-542) */
static void physical_variables(orc_case* c)
{
    int nx = c->nx, ny = c->ny;
    size_t N = c->N;
#pragma omp parallel for num_threads(omp_threads(c)) schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t cc = (size_t)j * nx + i;
            c->u0[cc] = c->u[cc];
            c->v0[cc] = c->v[cc];
            int ring = (i == 0) || (i == nx - 1) || (j == 0) || (j == ny - 1);
            double fk[Q];
            for (int k = 0; k < Q; ++k) {
                fk[k] = c->f0[k * N + cc];
                if (!ring) c->f[k * N + cc] = fk[k];
            }
            /* temprho/tempu/tempv sums in k order; e(k,.)=0 terms are exact no-ops */
            double tr = 0.0;
            for (int k = 0; k < Q; ++k) tr = tr + fk[k];
            double tu = fk[1] - fk[3] + fk[5] - fk[6] - fk[7] + fk[8];
            double tv = fk[2] - fk[4] + fk[5] + fk[6] - fk[7] - fk[8];
            c->rho[cc] = tr;
            c->pre[cc] = tr / 3.0;
            c->u[cc] = tu / tr;
            c->v[cc] = tv / tr;
        }
}

/* One node of colission (solution.f90:573-643): m = M f, relaxation towards
 * the incompressible moment equilibrium with jx = rho*u, Guo forcing in
 * moment space when a force acts, f = M^-1 m*. */
static void collide_node(const double* f, double rt, double ut, double vt, const double* s,
                         double gfx, double gfy, const double* minv, double* out)
{
    double temp[Q], temp1[Q], fp[Q];
    for (int k = 0; k < Q; ++k) temp[k] = f[k];
    matmulvoc(temp1, &MATM[0][0], temp);
    temp[0] = 0.0;
    temp[1] = -2.0 * rt + 3.0 * (rt * ut * rt * ut + rt * vt * rt * vt);
    temp[2] = rt - 3.0 * (rt * ut * rt * ut + rt * vt * rt * vt);
    temp[3] = 0.0;
    temp[4] = -rt * ut;
    temp[5] = 0.0;
    temp[6] = -rt * vt;
    temp[7] = (rt * ut * rt * ut - rt * vt * rt * vt);
    temp[8] = (rt * ut * rt * vt);
    int forced = (gfx != 0.0) || (gfy != 0.0);
    if (forced) {
        fp[0] = 0.0;
        fp[1] = 6.0 * (1.0 - 0.5 * s[1]) * (ut * gfx + vt * gfy);
        fp[2] = -6.0 * (1.0 - 0.5 * s[2]) * (ut * gfx + vt * gfy);
        fp[3] = gfx;
        fp[4] = -(1.0 - 0.5 * s[4]) * gfx;
        fp[5] = gfy;
        fp[6] = -(1.0 - 0.5 * s[6]) * gfy;
        fp[7] = 2.0 * (1.0 - 0.5 * s[7]) * (ut * gfx - vt * gfy);
        fp[8] = (1.0 - 0.5 * s[8]) * (ut * gfy + vt * gfx);
        for (int k = 0; k < Q; ++k) temp[k] = temp1[k] + (temp[k] - temp1[k]) * s[k] + fp[k];
    } else {
        /* fp == +-0: adding it is an exact no-op */
        for (int k = 0; k < Q; ++k) temp[k] = temp1[k] + (temp[k] - temp1[k]) * s[k];
    }
    matmulvoc(out, minv, temp);
}

void orc_collide_node(const double f[9], double rho, double u, double v, const double s[9],
                      double gfx, double gfy, double out[9])
{
    double minv[Q * Q];
    nizhen(&MATM[0][0], minv, Q);
    collide_node(f, rho, u, v, s, gfx, gfy, minv, out);
}

/* colission (solution.f90:573-643) on the interior; gforce is zero away from
 * the IB nodes so the force moments vanish there exactly. */
static void colission(orc_case* c)
{
    int nx = c->nx, ny = c->ny;
    size_t N = c->N;
    const double* s = c->p.s;
    const double* minv = c->minv;
#pragma omp parallel for num_threads(omp_threads(c)) schedule(static)
    for (int j = 1; j < ny - 1; ++j)
        for (int i = 1; i < nx - 1; ++i) {
            size_t cc = (size_t)j * nx + i;
            double fin[Q], fout[Q];
            for (int k = 0; k < Q; ++k) fin[k] = c->f[k * N + cc];
            collide_node(fin, c->rho[cc], c->u[cc], c->v[cc], s, c->gforce[cc], c->gforce[N + cc], minv, fout);
            for (int k = 0; k < Q; ++k) c->f[k * N + cc] = fout[k];
        }
}

/* ---------------- IB linear algebra ---------------- */

/* zig-zag order per wing: LE segment first, then alternating ends */
static void zigzag_perm(int nw, int spw, int* perm)
{
    int p = 0;
    for (int wg = 0; wg < nw; ++wg) {
        int lo = 0, hi = spw - 1, t = 0;
        while (lo <= hi) {
            if ((t & 1) == 0) perm[p++] = wg * spw + lo++;
            else perm[p++] = wg * spw + hi--;
            t++;
        }
    }
}

/* Square-root-free Cholesky (L D L^T) of the symmetric matrix a (n x n
 * row-major, lower triangle used), right-looking, element (i,j) updated in
 * increasing k: with t_i = a(i,k) before scaling and l_i = t_i * (1/d_k),
 * a(i,j) -= l_i * t_j.  dinv[k] keeps 1/d_k for the solves; L (unit lower)
 * replaces the strict lower triangle.  bw bounds the lower bandwidth of the
 * profile (n-1 = dense).  Returns 0 or -5 (pivot not positive). */
static int ldlt(double* a, int n, int bw, double* dinv)
{
    double* t = (double*)malloc(sizeof(double) * (size_t)(bw + 1));
    for (int k = 0; k < n; ++k) {
        double d = a[(size_t)k * n + k];
        if (!(d > 0.0)) { free(t); return -5; }
        double r = 1.0 / d;
        dinv[k] = r;
        int iend = k + bw < n - 1 ? k + bw : n - 1;
        for (int i = k + 1; i <= iend; ++i) {
            t[i - k - 1] = a[(size_t)i * n + k];
            a[(size_t)i * n + k] = t[i - k - 1] * r;
        }
        for (int j = k + 1; j <= iend; ++j) {
            double tj = t[j - k - 1];
            for (int i = j; i <= iend; ++i) {
                double lik = a[(size_t)i * n + k];
                a[(size_t)i * n + j] = a[(size_t)i * n + j] - lik * tj;
            }
        }
    }
    free(t);
    return 0;
}

/* forward L z = b, scaling x = D^-1 z, backward L^T y = x, column oriented
 * (nrhs columns of length n stored consecutively) */
static void ldlt_solve(const double* a, const double* dinv, int n, int bw, double* b, int nrhs)
{
    for (int r = 0; r < nrhs; ++r) {
        double* x = b + (size_t)r * n;
        for (int k = 0; k < n; ++k) {
            int iend = k + bw < n - 1 ? k + bw : n - 1;
            for (int i = k + 1; i <= iend; ++i) x[i] = x[i] - a[(size_t)i * n + k] * x[k];
        }
        for (int k = 0; k < n; ++k) x[k] = x[k] * dinv[k];
        for (int k = n - 1; k >= 0; --k) {
            int ibeg = k - bw > 0 ? k - bw : 0;
            for (int i = ibeg; i < k; ++i) x[i] = x[i] - a[(size_t)k * n + i] * x[k];
        }
    }
}

/* getrf/getrs restated: unblocked right-looking LU with partial pivoting on
 * a (n x n row-major), then two triangular solves per RHS. */
static int lu_solve(double* a, int n, double* b, int nrhs)
{
    int* ipiv = (int*)malloc(sizeof(int) * n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        double amax = fabs(a[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(a[(size_t)i * n + k]) > amax) { amax = fabs(a[(size_t)i * n + k]); p = i; }
        ipiv[k] = p;
        if (a[(size_t)p * n + k] == 0.0) { free(ipiv); return -5; }
        if (p != k)
            for (int j = 0; j < n; ++j) {
                double t = a[(size_t)k * n + j];
                a[(size_t)k * n + j] = a[(size_t)p * n + j];
                a[(size_t)p * n + j] = t;
            }
        double piv = a[(size_t)k * n + k];
        for (int i = k + 1; i < n; ++i) a[(size_t)i * n + k] = a[(size_t)i * n + k] / piv;
        for (int i = k + 1; i < n; ++i) {
            double lik = a[(size_t)i * n + k];
            if (lik == 0.0) continue;
            for (int j = k + 1; j < n; ++j) a[(size_t)i * n + j] = a[(size_t)i * n + j] - lik * a[(size_t)k * n + j];
        }
    }
    for (int r = 0; r < nrhs; ++r) {
        double* x = b + (size_t)r * n;
        for (int k = 0; k < n; ++k)
            if (ipiv[k] != k) { double t = x[k]; x[k] = x[ipiv[k]]; x[ipiv[k]] = t; }
        for (int k = 0; k < n; ++k)
            for (int i = k + 1; i < n; ++i) x[i] = x[i] - a[(size_t)i * n + k] * x[k];
        for (int k = n - 1; k >= 0; --k) {
            x[k] = x[k] / a[(size_t)k * n + k];
            for (int i = 0; i < k; ++i) x[i] = x[i] - a[(size_t)i * n + k] * x[k];
        }
    }
    free(ipiv);
    return 0;
}

int orc_solve(const double* G, const double* S, const double* b, int n,
              const int* perm, int solver, double* y)
{
    size_t nn = (size_t)n * n;
    double* a = (double*)malloc(sizeof(double) * nn);
    double* x = (double*)malloc(sizeof(double) * 2 * n);
    int rc;
    if (solver == 0) {
        for (int p = 0; p < n; ++p)
            for (int q = 0; q < n; ++q) a[(size_t)p * n + q] = G[(size_t)perm[p] * n + perm[q]];
        for (int r = 0; r < 2; ++r)
            for (int p = 0; p < n; ++p) x[r * n + p] = b[r * n + perm[p]];
        double* dinv = (double*)malloc(sizeof(double) * n);
        rc = ldlt(a, n, n - 1, dinv);
        if (rc == 0) {
            ldlt_solve(a, dinv, n, n - 1, x, 2);
            for (int r = 0; r < 2; ++r)
                for (int p = 0; p < n; ++p) y[r * n + perm[p]] = x[r * n + p];
        }
        free(dinv);
    } else {
        for (int p = 0; p < n; ++p)
            for (int q = 0; q < n; ++q) a[(size_t)p * n + q] = G[(size_t)p * n + q] * S[q];
        memcpy(x, b, sizeof(double) * 2 * n);
        rc = lu_solve(a, n, x, 2);
        if (rc == 0)
            for (int r = 0; r < 2; ++r)
                for (int p = 0; p < n; ++p) y[r * n + p] = x[r * n + p] * S[p];
    }
    free(a);
    free(x);
    return rc;
}

/* ---------------- rigid-body dynamics ---------------- */

/* update_body_extra_from_rel (mode.f90:285-307): running sums over wings */
static void update_body_extra(orc_case* c)
{
    double av = 0.0, as = 0.0, ap = 0.0;
    for (int w = 0; w < MAXW; ++w) { c->extra_v[w] = 0.0; c->extra_s[w] = 0.0; c->extra_pos[w] = 0.0; }
    for (int w = 0; w < c->p.n_wings; ++w) {
        av = av + c->rel_v[w];
        as = as + c->rel_s[w];
        ap = ap + c->rel_pos[w];
        c->extra_v[w] = av;
        c->extra_s[w] = as;
        c->extra_pos[w] = ap;
    }
}

/* rk4 (boundary_condition.f90:3168-3202): four stages with the acceleration
 * held over the step; the pitch equation reads alphad and the last moment */
static void rk4(const orc_case* c, double s1, double* s, double* s2, double v1, double* v, double a1, double jd1,
                double* jd, double jsd1, double* jsd, double p1, double p2)
{
    const double alphad = c->dyn.alphad, mom = c->mom;
    double k11 = v1, k21 = a1;
    double k12 = v1 + 0.5 * k21, k22 = a1;
    double k13 = v1 + 0.5 * k22, k23 = a1;
    double k14 = v1 + k23, k24 = a1;
    *s = s1 + (k11 + 2.0 * k12 + 2.0 * k13 + k14) / 6.0;
    *v = v1 + (k21 + 2.0 * k22 + 2.0 * k23 + k24) / 6.0;
    *s2 = *s2 + (k11 + 2.0 * k12 + 2.0 * k13 + k14) / 6.0;
    k11 = jsd1;
    k21 = p1 * (jd1 + alphad) + p2 * mom;
    k12 = jsd1 + 0.5 * k21;
    k22 = p1 * (jd1 + alphad + 0.5 * k11) + p2 * mom;
    k13 = jsd1 + 0.5 * k22;
    k23 = p1 * (jd1 + alphad + 0.5 * k12) + p2 * mom;
    k14 = jsd1 + k23;
    k24 = p1 * (jd1 + alphad + k13) + p2 * mom;
    *jd = jd1 + (k11 + 2.0 * k12 + 2.0 * k13 + k14) / 6.0;
    *jsd = jsd1 + (k21 + 2.0 * k22 + 2.0 * k23 + k24) / 6.0;
}

/* the dynamics block of immersed_boundary_Wu (boundary_condition.f90:1676-1719,
 * 1731): wing 1 accelerates by the total raw x force over its mass */
static void body_update(orc_case* c)
{
    const double sc = c->p.scale, UU = c->p.U;
    const orc_dynamics* d = &c->dyn;
    double force1 = 0.0;
    for (int w = 0; w < c->p.n_wings; ++w) force1 = force1 + c->wing_force[w][0];
    c->rel_a_prev[0] = c->rel_a[0];
    if (!d->closed_boundary)
        c->rel_a[0] = force1 / (d->solid_den / (sc * sc));
    else
        c->rel_a[0] = force1 / (d->solid_den * d->area_airfoil / (sc * sc));
    c->rel_v_prev[0] = c->rel_v[0];
    c->rel_s_prev[0] = c->rel_s[0];
    c->rel_theta_prev[0] = c->rel_theta[0];
    c->rel_tt_prev[0] = c->rel_tt[0];
    if (!d->rk4) {
        c->rel_v[0] = c->rel_a[0] + c->rel_v[0];
        c->rel_s[0] = c->rel_s[0] + c->rel_v_prev[0];
        c->rel_pos[0] = c->rel_pos[0] + c->rel_v_prev[0];
        c->rel_theta[0] = c->rel_theta_prev[0] + c->rel_tt_prev[0];
        c->rel_tt[0] = c->rel_tt_prev[0] + d->par1 * UU * UU * sc * sc * (c->rel_theta[0] + d->alphad) +
                       d->par2 * (sc * sc * sc * sc) * c->mom;
    } else {
        rk4(c, c->rel_s_prev[0], &c->rel_s[0], &c->rel_pos[0], c->rel_v_prev[0], &c->rel_v[0], c->rel_a[0],
            c->rel_theta_prev[0], &c->rel_theta[0], c->rel_tt_prev[0], &c->rel_tt[0], d->par1 * UU * UU * sc * sc,
            d->par2 * (sc * sc * sc * sc));
    }
    update_body_extra(c);
}

/* move_mesh (move_mesh.f90:1-118 for dr = 1, 186-214 for dr = -1; one mesh,
 * one process, no periodic edges): every array shifts by one column towards
 * -dr; the column that has no source keeps its values.  u0/v0 take the
 * shifted velocity and pre the shifted density (not rho/3), as there. */
static void move_mesh(orc_case* c, int dr)
{
    const int nx = c->nx, ny = c->ny;
    const size_t N = c->N;
    const int k0 = dr == 1 ? 0 : 1, k1 = dr == 1 ? nx - 2 : nx - 1;
    for (int j = 0; j < ny; ++j)
        for (int k = k0; k <= k1; ++k) {
            const size_t d = (size_t)j * nx + k, s = d + dr;
            for (int q = 0; q < Q; ++q) c->f0[q * N + d] = c->f[q * N + s];
            c->u0[d] = c->u[s];
            c->v0[d] = c->v[s];
            c->pre[d] = c->rho[s];
        }
    for (int j = 0; j < ny; ++j)
        for (int k = k0; k <= k1; ++k) {
            const size_t d = (size_t)j * nx + k;
            for (int q = 0; q < Q; ++q) c->f[q * N + d] = c->f0[q * N + d];
            c->u[d] = c->u0[d];
            c->v[d] = c->v0[d];
            c->rho[d] = c->pre[d];
        }
}

void orc_move_mesh(orc_case* c, int dr) { move_mesh(c, dr >= 0 ? 1 : -1); }

int orc_set_dynamics(orc_case* c, const orc_dynamics* d)
{
    if (c->n != 0 || !d) return -1;
    if (d->enabled && !(d->solid_den > 0.0)) return -1;
    c->dyn = *d;
    return 0;
}

int orc_body_state(const orc_case* c, double* out)
{
    const int nw = c->p.n_wings;
    const double* arr[11] = {c->rel_a, c->rel_a_prev, c->rel_v, c->rel_v_prev, c->rel_s, c->rel_s_prev,
                             c->rel_pos, c->rel_theta_prev, c->rel_theta, c->rel_tt_prev, c->rel_tt};
    int o = 0;
    for (int a = 0; a < 11; ++a)
        for (int w = 0; w < nw; ++w) out[o++] = arr[a][w];
    out[o++] = (double)c->movenum;
    return o;
}

int orc_motion_rows(const orc_case* c, double* rows, int max_rows)
{
    const int cols = 1 + 3 * c->p.n_wings;
    const int n = c->hist_rows < max_rows ? c->hist_rows : max_rows;
    if (rows) memcpy(rows, c->mrows, sizeof(double) * (size_t)n * cols);
    return c->hist_rows;
}

/* per-marker sparse row of D = xx1: box nodes sorted by IB list index */
typedef struct drow {
    int cnt;
    int idx[64];
    double val[64];
} drow;

static void sort_drow(drow* r)
{
    for (int a = 1; a < r->cnt; ++a) {
        int ki = r->idx[a];
        double kv = r->val[a];
        int b = a - 1;
        while (b >= 0 && r->idx[b] > ki) { r->idx[b + 1] = r->idx[b]; r->val[b + 1] = r->val[b]; b--; }
        r->idx[b + 1] = ki;
        r->val[b + 1] = kv;
    }
}

/* immersed_boundary_Wu (boundary_condition.f90:1512-1692) */
static int immersed_boundary(orc_case* c)
{
    const int nw = c->p.n_wings, ppw = c->p.pts_per_wing, spw = c->segs_per_wing;
    const int ns = c->nseg, na = c->nadj;
    const double ox = c->p.ox, oy = c->p.oy, sc = c->p.scale, U = c->p.U;
    const int smooth = c->p.smooth;
    size_t N = c->N;

    /* marker velocities (boundary_condition.f90:1542-1553) */
    for (int wg = 0; wg < nw; ++wg) {
        int ss = wg * spw, ps = wg * ppw;
        double sn = sin(c->angle[wg]), ca = cos(c->angle[wg]), om = c->omega[wg];
        for (int i = ss; i < ss + spw; ++i) {
            double xs = c->xb0[ps + i - ss + 1] + c->xb0[ps + i - ss];
            double ys = c->yb0[ps + i - ss + 1] + c->yb0[ps + i - ss];
            c->ub[i] = U * (-om * sn * 0.5 * xs + om * ca * 0.5 * ys + c->sau[wg]) + c->extra_v[wg];
            c->vb[i] = U * (-om * ca * 0.5 * xs - om * sn * 0.5 * ys + c->sbv[wg]);
        }
    }

    double* ustar = (double*)malloc(sizeof(double) * na);
    double* vstar = (double*)malloc(sizeof(double) * na);
    for (int k = 0; k < na; ++k) {
        size_t cc = (size_t)c->adj_j[k] * c->nx + c->adj_i[k];
        ustar[k] = c->u[cc];
        vstar[k] = c->v[cc];
    }
    double* bu = (double*)malloc(sizeof(double) * ns);
    double* bv = (double*)malloc(sizeof(double) * ns);
    double* G = NULL;   /* dense n x n */
    double* xx1 = NULL; /* dense ns x na (dense mode) */
    drow* rows = NULL;  /* sparse mode */
    int rc = 0;

    if (c->p.dense_ib) {
        xx1 = (double*)malloc(sizeof(double) * (size_t)ns * na);
        for (int k = 0; k < ns; ++k)
            for (int i = 0; i < na; ++i) {
                double x0 = ox + sc * (double)(c->adj_i[i]);
                double y0 = oy + sc * (double)(c->adj_j[i]);
                double k1 = orc_delta(fabs((x0 - c->xb2[k]) / sc), smooth);
                double k2 = orc_delta(fabs((y0 - c->yb2[k]) / sc), smooth);
                xx1[(size_t)k * na + i] = k1 * k2;
            }
        for (int k = 0; k < ns; ++k) {
            double su = 0.0, sv = 0.0;
            for (int i = 0; i < na; ++i) {
                su = su + xx1[(size_t)k * na + i] * ustar[i];
                sv = sv + xx1[(size_t)k * na + i] * vstar[i];
            }
            bu[k] = su; bv[k] = sv;
        }
    } else {
        rows = (drow*)malloc(sizeof(drow) * ns);
        for (int k = 0; k < ns; ++k) {
            const int* bx = c->box + 4 * k;
            drow* r = rows + k;
            r->cnt = 0;
            for (int i = bx[0]; i <= bx[1]; ++i) {
                double x0 = ox + sc * (double)(i - 1);
                double k1 = orc_delta(fabs((x0 - c->xb2[k]) / sc), smooth);
                for (int j = bx[2]; j <= bx[3]; ++j) {
                    double y0 = oy + sc * (double)(j - 1);
                    double k2 = orc_delta(fabs((y0 - c->yb2[k]) / sc), smooth);
                    r->idx[r->cnt] = c->nodeidx[(size_t)(j - 1) * c->nx + (i - 1)];
                    r->val[r->cnt] = k1 * k2;
                    r->cnt++;
                }
            }
            sort_drow(r);
            double su = 0.0, sv = 0.0;
            for (int a = 0; a < r->cnt; ++a) {
                su = su + r->val[a] * ustar[r->idx[a]];
                sv = sv + r->val[a] * vstar[r->idx[a]];
            }
            bu[k] = su; bv[k] = sv;
        }
    }

    /* comj = ub - bu (boundary_condition.f90:1624-1629) */
    double* rhs = (double*)malloc(sizeof(double) * 2 * ns);
    for (int k = 0; k < ns; ++k) { rhs[k] = c->ub[k] - bu[k]; rhs[ns + k] = c->vb[k] - bv[k]; }

    /* assemble and solve */
    G = (double*)calloc((size_t)ns * ns, sizeof(double));
    double* y = (double*)malloc(sizeof(double) * 2 * ns);
    if (c->p.solver == 1) {
        /* as = xx1 * xx2, xx2(i,l) = kernel1*kernel2*arc_l(l) (literal) */
        if (c->p.dense_ib) {
            for (int k = 0; k < ns; ++k)
                for (int l = 0; l < ns; ++l) {
                    double sum = 0.0;
                    for (int i = 0; i < na; ++i)
                        sum = sum + xx1[(size_t)k * na + i] * (xx1[(size_t)l * na + i] * c->arc_l[l]);
                    G[(size_t)k * ns + l] = sum;
                }
        } else {
            for (int k = 0; k < ns; ++k)
                for (int l = 0; l < ns; ++l) {
                    const drow *rk = rows + k, *rl = rows + l;
                    double sum = 0.0;
                    int a = 0, b = 0;
                    while (a < rk->cnt && b < rl->cnt) {
                        if (rk->idx[a] < rl->idx[b]) a++;
                        else if (rk->idx[a] > rl->idx[b]) b++;
                        else { sum = sum + rk->val[a] * (rl->val[b] * c->arc_l[l]); a++; b++; }
                    }
                    G[(size_t)k * ns + l] = sum;
                }
        }
        memcpy(y, rhs, sizeof(double) * 2 * ns);
        rc = lu_solve(G, ns, y, 2); /* y now holds c = comj solution */
    } else {
        int* perm = (int*)malloc(sizeof(int) * ns);
        zigzag_perm(nw, spw, perm);
        /* G_perm(p,q) = sum_i D(perm p, i) D(perm q, i) in ascending node order */
        if (c->p.dense_ib) {
            for (int p = 0; p < ns; ++p)
                for (int q = 0; q <= p; ++q) {
                    int k = perm[p], l = perm[q];
                    double sum = 0.0;
                    for (int i = 0; i < na; ++i) sum = sum + xx1[(size_t)k * na + i] * xx1[(size_t)l * na + i];
                    G[(size_t)p * ns + q] = sum;
                }
        } else {
            for (int p = 0; p < ns; ++p)
                for (int q = 0; q <= p; ++q) {
                    int k = perm[p], l = perm[q];
                    const int* bk = c->box + 4 * k;
                    const int* bl = c->box + 4 * l;
                    if (bk[1] < bl[0] || bl[1] < bk[0] || bk[3] < bl[2] || bl[3] < bk[2]) continue;
                    const drow *rk = rows + k, *rl = rows + l;
                    double sum = 0.0;
                    int a = 0, b = 0;
                    while (a < rk->cnt && b < rl->cnt) {
                        if (rk->idx[a] < rl->idx[b]) a++;
                        else if (rk->idx[a] > rl->idx[b]) b++;
                        else { sum = sum + rk->val[a] * rl->val[b]; a++; b++; }
                    }
                    G[(size_t)p * ns + q] = sum;
                }
        }
        /* lower bandwidth of the profile */
        int bw = ns - 1;
        if (!c->p.dense_ib) {
            bw = 0;
            for (int p = 0; p < ns; ++p)
                for (int q = 0; q < p; ++q)
                    if (G[(size_t)p * ns + q] != 0.0) { if (p - q > bw) bw = p - q; break; }
        }
        double* x = (double*)malloc(sizeof(double) * 2 * ns);
        for (int r = 0; r < 2; ++r)
            for (int p = 0; p < ns; ++p) x[r * ns + p] = rhs[r * ns + perm[p]];
        double* dinv = (double*)malloc(sizeof(double) * ns);
        rc = ldlt(G, ns, bw, dinv);
        if (rc == 0) {
            ldlt_solve(G, dinv, ns, bw, x, 2);
            for (int r = 0; r < 2; ++r)
                for (int p = 0; p < ns; ++p) y[r * ns + perm[p]] = x[r * ns + p];
        }
        free(dinv);
        free(x);
        free(perm);
    }

    if (rc == 0) {
        /* spread: glx(i) = sum_k xx2(i,k) comj(k) (solver 1) or sum_k D(k,i) y(k)
         * (solver 0), k ascending (boundary_condition.f90:1639-1640) */
        for (int i = 0; i < na; ++i) { c->glx[i] = 0.0; c->gly[i] = 0.0; }
        if (c->p.dense_ib) {
            for (int i = 0; i < na; ++i) {
                double sx = 0.0, sy = 0.0;
                for (int k = 0; k < ns; ++k) {
                    double d = xx1[(size_t)k * na + i];
                    if (c->p.solver == 1) d = d * c->arc_l[k];
                    sx = sx + d * y[k];
                    sy = sy + d * y[ns + k];
                }
                c->glx[i] = sx; c->gly[i] = sy;
            }
        } else {
            for (int k = 0; k < ns; ++k) {
                const drow* r = rows + k;
                for (int a = 0; a < r->cnt; ++a) {
                    int i = r->idx[a];
                    double d = r->val[a];
                    if (c->p.solver == 1) d = d * c->arc_l[k];
                    c->glx[i] = c->glx[i] + d * y[k];
                    c->gly[i] = c->gly[i] + d * y[ns + k];
                }
            }
        }
        /* velocity correction and body force (boundary_condition.f90:1647-1657) */
        for (int k = 0; k < na; ++k) {
            size_t cc = (size_t)c->adj_j[k] * c->nx + c->adj_i[k];
            c->u[cc] = ustar[k] + c->glx[k];
            c->v[cc] = vstar[k] + c->gly[k];
            double t1 = 2.0 * c->rho[cc] * c->glx[k];
            double t2 = 2.0 * c->rho[cc] * c->gly[k];
            c->rho_adj[k] = c->rho[cc];
            c->gforce[cc] = t1;
            c->gforce[N + cc] = t2;
        }
        /* per-wing force and moment, sequential in node order (1662-1675) */
        for (int wg = 0; wg < MAXW; ++wg) { c->wing_force[wg][0] = 0.0; c->wing_force[wg][1] = 0.0; c->wing_moment[wg] = 0.0; }
        for (int k = 0; k < na; ++k) {
            int wg = c->adj_w[k];
            double t1 = 2.0 * c->rho_adj[k] * c->glx[k];
            double t2 = 2.0 * c->rho_adj[k] * c->gly[k];
            c->wing_force[wg][0] = c->wing_force[wg][0] - t1;
            c->wing_force[wg][1] = c->wing_force[wg][1] - t2;
            double x0 = (ox + sc * (double)(c->adj_i[k]) - (c->p.xc + c->sa[wg])) / sc - c->extra_pos[wg];
            double y0 = (oy + sc * (double)(c->adj_j[k]) - (c->p.yc + c->sb[wg])) / sc;
            c->wing_moment[wg] = c->wing_moment[wg] - x0 * t2 + y0 * t1;
        }
        c->mom = c->wing_moment[0];
        if (c->dyn.enabled) body_update(c);
    }
    free(ustar); free(vstar); free(bu); free(bv); free(rhs); free(G); free(y);
    free(xx1); free(rows);
    return rc;
}

/* compressible moment-space equilibrium used by the NEE edges
 * (boundary_condition.f90:444-455) */
static inline void meq_nee(double rh, double uu, double vv, double* t)
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

/* f_edge = M^-1 [ meq(edge) + (1-s) (M f0_nb - meq(nb)) ] */
static inline void nee_s(const double* minv, const double* s, double rhe, double ue, double ve, double rhn,
                         double un, double vn, const double* f0nb, double* out)
{
    double t1[Q], t2[Q], t3[Q];
    meq_nee(rhe, ue, ve, t1);
    meq_nee(rhn, un, vn, t2);
    matmulvoc(t3, &MATM[0][0], f0nb);
    for (int k = 0; k < Q; ++k) t3[k] = t1[k] + (1.0 - s[k]) * (t3[k] - t2[k]);
    matmulvoc(out, minv, t3);
}
static inline void nee(const orc_case* c, double rhe, double ue, double ve,
                       double rhn, double un, double vn, const double* f0nb, double* out)
{
    nee_s(c->minv, c->p.s, rhe, ue, ve, rhn, un, vn, f0nb, out);
}

/* characteristic update of an inflow (side 0, boundary_condition.f90:418-440)
 * or outflow (side 1, 500-521) edge node from its own state (tr, tu, tv) and
 * the next two nodes inwards */
static inline void char_update(int side, double cs, double tr, double tu, double tv, double r2, double u2,
                               double v2, double r3, double u3, double v3, double* ro, double* uo, double* vo)
{
    double lx1, lx2, lx3;
    if (side == 0) {
        double rhox = (-3.0 * tr + 4.0 * r2 - r3) * 0.5;
        double ux = (-3.0 * tu + 4.0 * u2 - u3) * 0.5;
        double vx = (-3.0 * tv + 4.0 * v2 - v3) * 0.5;
        lx1 = (tu - cs) * (cs * cs * rhox - cs * tr * ux);
        lx2 = (tu < 0.0) ? tu * vx : 0.0;
        lx3 = 0.0;
    } else {
        double rhox = (3.0 * tr - 4.0 * r2 + r3) * 0.5;
        double ux = (3.0 * tu - 4.0 * u2 + u3) * 0.5;
        double vx = (3.0 * tv - 4.0 * v2 + v3) * 0.5;
        lx1 = 0.0;
        lx2 = (tu >= 0.0) ? tu * vx : 0.0;
        lx3 = (tu + cs) * (cs * cs * rhox + cs * tr * ux);
    }
    *ro = tr - (lx1 + lx3) * 1.5;
    *uo = tu + (lx1 - lx3) / (2.0 * cs * *ro);
    *vo = tv - lx2;
}

void orc_nee_node(const double s[9], double rhe, double ue, double ve, double rhn, double un, double vn,
                  const double f0n[9], double out[9])
{
    double minv[Q * Q];
    orc_minv(minv);
    nee_s(minv, s, rhe, ue, ve, rhn, un, vn, f0n, out);
}

void orc_char_update(int side, double cs, const double in[9], double out[3])
{
    char_update(side, cs, in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8], &out[0], &out[1], &out[2]);
}

/* boundary_condition1 (boundary_condition.f90:288-858) with periodical_bc=0,
 * ref_freestream=1 and no neighbouring blocks.  The top/bottom characteristic
 * updates and the corner states (303-401, 599-656, 728-783) are overwritten
 * by the no-slip assignment (658-661, 785-788) before use, so only their
 * effective result is restated. */
static void boundary_condition1(orc_case* c)
{
    int nx = c->nx, ny = c->ny;
    size_t N = c->N;
    double cs = c->cs;
    double* rho = c->rho;
    double* u = c->u;
    double* v = c->v;
#define IDX(i, j) ((size_t)(j) * nx + (i))
    double* tr = (double*)malloc(sizeof(double) * ny);
    double* tu = (double*)malloc(sizeof(double) * ny);
    double* tv = (double*)malloc(sizeof(double) * ny);
    double f0n[Q], out[Q];

    /* left inflow (413-483) */
    for (int j = 0; j < ny; ++j) { tr[j] = rho[IDX(0, j)]; tu[j] = u[IDX(0, j)]; tv[j] = v[IDX(0, j)]; }
    for (int j = 1; j < ny - 1; ++j)
        char_update(0, cs, tr[j], tu[j], tv[j], rho[IDX(1, j)], u[IDX(1, j)], v[IDX(1, j)], rho[IDX(2, j)],
                    u[IDX(2, j)], v[IDX(2, j)], &rho[IDX(0, j)], &u[IDX(0, j)], &v[IDX(0, j)]);
    for (int j = 1; j < ny - 1; ++j) {
        for (int k = 0; k < Q; ++k) f0n[k] = c->f0[k * N + IDX(1, j)];
        nee(c, rho[IDX(0, j)], u[IDX(0, j)], v[IDX(0, j)], rho[IDX(1, j)], u[IDX(1, j)], v[IDX(1, j)], f0n, out);
        for (int k = 0; k < Q; ++k) c->f[k * N + IDX(0, j)] = out[k];
    }
    /* right outflow (487-569) */
    for (int j = 0; j < ny; ++j) { tr[j] = rho[IDX(nx - 1, j)]; tu[j] = u[IDX(nx - 1, j)]; tv[j] = v[IDX(nx - 1, j)]; }
    for (int j = 1; j < ny - 1; ++j)
        char_update(1, cs, tr[j], tu[j], tv[j], rho[IDX(nx - 2, j)], u[IDX(nx - 2, j)], v[IDX(nx - 2, j)],
                    rho[IDX(nx - 3, j)], u[IDX(nx - 3, j)], v[IDX(nx - 3, j)], &rho[IDX(nx - 1, j)],
                    &u[IDX(nx - 1, j)], &v[IDX(nx - 1, j)]);
    for (int j = 1; j < ny - 1; ++j) {
        for (int k = 0; k < Q; ++k) f0n[k] = c->f0[k * N + IDX(nx - 2, j)];
        nee(c, rho[IDX(nx - 1, j)], u[IDX(nx - 1, j)], v[IDX(nx - 1, j)], rho[IDX(nx - 2, j)], u[IDX(nx - 2, j)],
            v[IDX(nx - 2, j)], f0n, out);
        for (int k = 0; k < Q; ++k) c->f[k * N + IDX(nx - 1, j)] = out[k];
    }
    /* top wall (598-726): effective rho=rho(.,Ny-1), u=v=0, NEE from row Ny-1 */
    for (int i = 0; i < nx; ++i) {
        rho[IDX(i, ny - 1)] = rho[IDX(i, ny - 2)];
        u[IDX(i, ny - 1)] = 0.0;
        v[IDX(i, ny - 1)] = 0.0;
        for (int k = 0; k < Q; ++k) f0n[k] = c->f0[k * N + IDX(i, ny - 2)];
        nee(c, rho[IDX(i, ny - 1)], u[IDX(i, ny - 1)], v[IDX(i, ny - 1)], rho[IDX(i, ny - 2)], u[IDX(i, ny - 2)],
            v[IDX(i, ny - 2)], f0n, out);
        for (int k = 0; k < Q; ++k) c->f[k * N + IDX(i, ny - 1)] = out[k];
    }
    /* bottom wall (727-853) */
    for (int i = 0; i < nx; ++i) {
        rho[IDX(i, 0)] = rho[IDX(i, 1)];
        u[IDX(i, 0)] = 0.0;
        v[IDX(i, 0)] = 0.0;
        for (int k = 0; k < Q; ++k) f0n[k] = c->f0[k * N + IDX(i, 1)];
        nee(c, rho[IDX(i, 0)], u[IDX(i, 0)], v[IDX(i, 0)], rho[IDX(i, 1)], u[IDX(i, 1)], v[IDX(i, 1)], f0n, out);
        for (int k = 0; k < Q; ++k) c->f[k * N + IDX(i, 0)] = out[k];
    }
#undef IDX
    free(tr); free(tu); free(tv);
}

/* erranalysis (solution.f90:87-123), sequential i-outer j-inner */
static void erranalysis(orc_case* c)
{
    int nx = c->nx, ny = c->ny;
    double t1 = 0.0, t2 = 0.0;
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j) {
            size_t cc = (size_t)j * nx + i;
            double du = c->u[cc] - c->u0[cc], dv = c->v[cc] - c->v0[cc];
            t1 = t1 + du * du + dv * dv;
            t2 = t2 + c->u[cc] * c->u[cc] + c->v[cc] * c->v[cc];
        }
    c->err = sqrt(t1) / (sqrt(t2) + 1e-30);
}

/* evolution (solution.f90:269-410) */
static int evolution(orc_case* c, int cycle1)
{
    double sc = c->p.scale, U = c->p.U;
    int nw = c->p.n_wings;
    double time = (double)(cycle1 - 1) * (sc * U);
    mode(c, time);
    int rc = move_boundary(c);
    if (rc < 0) return rc;
    streaming(c);
    physical_variables(c);
    rc = immersed_boundary(c);
    if (rc < 0) return rc;
    colission(c);
    /* coefficients (solution.f90:327-376) */
    double force[2] = {0.0, 0.0}, total_moment = 0.0;
    for (int wg = 0; wg < nw; ++wg) {
        c->wing_force[wg][0] = 2.0 * c->wing_force[wg][0] / (U * U * (1.0 / sc));
        c->wing_force[wg][1] = 2.0 * c->wing_force[wg][1] / (U * U * (1.0 / sc));
        c->wing_moment[wg] = -2.0 * c->wing_moment[wg] / (U * U * (1.0 / (sc * sc)));
    }
    for (int wg = 0; wg < nw; ++wg) { force[0] = force[0] + c->wing_force[wg][0]; force[1] = force[1] + c->wing_force[wg][1]; }
    for (int wg = 0; wg < nw; ++wg) total_moment = total_moment + c->wing_moment[wg];
    if (c->hist_rows >= c->hist_cap) {
        c->hist_cap *= 2;
        c->hist = (double*)realloc(c->hist, sizeof(double) * c->hist_cap * c->hist_cols);
        c->mrows = (double*)realloc(c->mrows, sizeof(double) * c->hist_cap * (1 + 3 * nw));
    }
    /* airfoil_motion row (solution.f90:378-388): after the IB's dynamics update */
    {
        double* m = c->mrows + (size_t)c->hist_rows * (1 + 3 * nw);
        int mc = 0;
        m[mc++] = time;
        for (int wg = 0; wg < nw; ++wg) {
            m[mc++] = 0.0;
            m[mc++] = c->extra_v[wg] / U;
            m[mc++] = c->extra_s[wg] / (1.0 / sc);
        }
    }
    double* row = c->hist + (size_t)c->hist_rows * c->hist_cols;
    int col = 0;
    row[col++] = time;
    for (int wg = 0; wg < nw; ++wg) { row[col++] = c->wing_force[wg][0]; row[col++] = c->wing_force[wg][1]; row[col++] = c->wing_moment[wg]; }
    row[col++] = force[0]; row[col++] = force[1]; row[col++] = total_moment;
    for (int wg = 0; wg < nw; ++wg) {
        row[col++] = c->sa[wg]; row[col++] = c->sau[wg]; row[col++] = c->saac[wg];
        row[col++] = c->sb[wg]; row[col++] = c->sbv[wg]; row[col++] = c->sbac[wg];
        row[col++] = c->angle[wg]; row[col++] = c->omega[wg]; row[col++] = c->omgac[wg];
    }
    c->hist_rows++;
    boundary_condition1(c);
    return 0;
}

int orc_step(orc_case* c, int nsteps)
{
    for (int s = 0; s < nsteps; ++s) {
        int n = c->n + 1;
        if (n >= 2) erranalysis(c);
        int rc = evolution(c, n);
        if (rc < 0) return rc;
        /* the domain follows the body one node at a time (main.f90:370-388) */
        if (c->rel_pos[0] >= 1.0) {
            move_mesh(c, 1);
            c->rel_pos[0] = c->rel_pos[0] - 1.0;
            update_body_extra(c);
            c->movenum = c->movenum + 1;
        } else if (c->rel_pos[0] <= -1.0) {
            move_mesh(c, -1);
            c->rel_pos[0] = c->rel_pos[0] + 1.0;
            update_body_extra(c);
            c->movenum = c->movenum - 1;
        }
        c->n = n;
    }
    return 0;
}

void orc_get_fields(const orc_case* c, double* rho, double* u, double* v, double* f, double* f0)
{
    size_t N = c->N;
    if (rho) memcpy(rho, c->rho, sizeof(double) * N);
    if (u) memcpy(u, c->u, sizeof(double) * N);
    if (v) memcpy(v, c->v, sizeof(double) * N);
    if (f) memcpy(f, c->f, sizeof(double) * Q * N);
    if (f0) memcpy(f0, c->f0, sizeof(double) * Q * N);
}

void orc_set_f(orc_case* c, const double* f) { memcpy(c->f, f, sizeof(double) * Q * c->N); }

int orc_history_cols(const orc_case* c) { return c->hist_cols; }

int orc_get_history(const orc_case* c, double* rows, int max_rows)
{
    int n = c->hist_rows < max_rows ? c->hist_rows : max_rows;
    if (rows) memcpy(rows, c->hist, sizeof(double) * (size_t)n * c->hist_cols);
    return c->hist_rows;
}

double orc_err(const orc_case* c) { return c->err; }

int orc_ib_count(const orc_case* c) { return c->nadj; }

void orc_ib_nodes(const orc_case* c, int* pi, int* pj, int* wing, double* glx, double* gly, double* rho_adj)
{
    for (int k = 0; k < c->nadj; ++k) {
        if (pi) pi[k] = c->adj_i[k];
        if (pj) pj[k] = c->adj_j[k];
        if (wing) wing[k] = c->adj_w[k];
        if (glx) glx[k] = c->glx[k];
        if (gly) gly[k] = c->gly[k];
        if (rho_adj) rho_adj[k] = c->rho_adj[k];
    }
}

int orc_n_segments(const orc_case* c) { return c->nseg; }

void orc_markers(const orc_case* c, double* xb2, double* yb2, double* ub, double* vb, double* arc_l)
{
    if (xb2) memcpy(xb2, c->xb2, sizeof(double) * c->nseg);
    if (yb2) memcpy(yb2, c->yb2, sizeof(double) * c->nseg);
    if (ub) memcpy(ub, c->ub, sizeof(double) * c->nseg);
    if (vb) memcpy(vb, c->vb, sizeof(double) * c->nseg);
    if (arc_l) memcpy(arc_l, c->arc_l, sizeof(double) * c->nseg);
}

/* ------------------------------------------------------------------ */
/* post-processing                                                     */
/* ------------------------------------------------------------------ */

/* The Vorticity column of output (solution.f90:181-210): dv/dx - du/dy in
 * units of U/scale, second-order one-sided differences on the first and last
 * column (row), central differences elsewhere. */
void orc_vorticity(int nx, int ny, double UU, double scale, const double* u, const double* v, double* vor)
{
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const size_t c = (size_t)j * nx + i;
            double dvx, duy;
            if (i == nx - 1)
                dvx = (1.5 * v[c] - 2.0 * v[c - 1] + 0.5 * v[c - 2]) / (UU * scale);
            else if (i == 0)
                dvx = (-1.5 * v[c] + 2.0 * v[c + 1] - 0.5 * v[c + 2]) / (UU * scale);
            else
                dvx = (v[c + 1] - v[c - 1]) / (UU * 2.0 * scale);
            if (j == ny - 1)
                duy = (1.5 * u[c] - 2.0 * u[c - nx] + 0.5 * u[c - 2 * (size_t)nx]) / (UU * scale);
            else if (j == 0)
                duy = (-1.5 * u[c] + 2.0 * u[c + nx] - 0.5 * u[c + 2 * (size_t)nx]) / (UU * scale);
            else
                duy = (u[c + nx] - u[c - nx]) / (UU * 2.0 * scale);
            vor[c] = dvx - duy;
        }
}

/* One call of vor_mom_rate (solution.f90:1880-1897) on the interior nodes:
 * vor and vor1 hold the vorticity of the last two calls (shifted here), the
 * rate is the second-order backward difference in units of U/scale per step.
 * Edge nodes are left untouched. */
void orc_vorticity_rate(int nx, int ny, double UU, double scale, const double* u, const double* v,
                        double* vor, double* vor1, double* dvordt)
{
    for (int j = 1; j < ny - 1; ++j)
        for (int i = 1; i < nx - 1; ++i) {
            const size_t c = (size_t)j * nx + i;
            const double tem1 = vor1[c];
            const double tem2 = vor[c];
            vor1[c] = tem2;
            const double dvx = (v[c + 1] - v[c - 1]) / (UU * 2.0 * scale);
            const double duy = (u[c + nx] - u[c - nx]) / (UU * 2.0 * scale);
            const double tem = dvx - duy;
            vor[c] = tem;
            dvordt[c] = (1.5 * tem - 2.0 * tem2 + 0.5 * tem1) / (UU * scale);
        }
}

/* cal_sur_cp (solution.f90:1912-2118, single process): the pressure at each
 * marker midpoint, sum over the IB nodes in list order of
 * phi(x_i - xb2) phi(y_i - yb2) pre_i with the smoothed kernel (smooth = 1),
 * nodes outside the kernel's support skipped (matmulvocxs over the sparse
 * rows, boundary_condition.f90:1869-1900).  pre = NULL takes the case's own
 * pressure field rho/3; any other field of nx*ny values may be given.  One
 * value per marker of the live IB (nseg). */
int orc_surface_pressure(const orc_case* c, const double* pre, double* cp)
{
    if (!pre) pre = c->pre;
    const double sc = c->p.scale;
    for (int k = 0; k < c->nseg; ++k) {
        double bb = 0.0;
        for (int a = 0; a < c->nadj; ++a) {
            const double x0 = c->p.ox + sc * (double)c->adj_i[a];
            const double y0 = c->p.oy + sc * (double)c->adj_j[a];
            const double dx = fabs((x0 - c->xb2[k]) / sc);
            const double dy = fabs((y0 - c->yb2[k]) / sc);
            if (dx > 2.5 || dy > 2.5) continue;
            const double w = orc_delta(dx, 1) * orc_delta(dy, 1);
            bb = bb + w * pre[(size_t)c->adj_j[a] * c->nx + c->adj_i[a]];
        }
        cp[k] = bb;
    }
    return c->nseg;
}
