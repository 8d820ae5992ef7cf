/*
 * lbmib_oracle.h -- CPU restatement of the reference solver's per-timestep
 * hot path (D2Q9 MRT lattice Boltzmann + implicit velocity-correction
 * immersed boundary + characteristic/NEE outer boundaries + prescribed
 * marker kinematics).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (liblbmib.so) never links, loads or calls anything in oracle/.
 *
 * Every function cites the reference file:line it restates.  Arithmetic is
 * plain IEEE double, evaluated left to right exactly as the Fortran source is
 * written, compiled with -ffp-contract=off (no FMA contraction) so that the
 * GPU path, built under the same rule, can be compared bit for bit.
 *
 * How each part is pinned (tests/test_oracle.py, tests/test_cases.py):
 *   weights, cs, M^-1, asin, both delta kernels, the collision (mass, force,
 *   He-Luo equilibrium fixed point, Guo forcing moments), feq, the motion
 *   table, the IB solvers (numpy, LU vs L D L^T, elimination order), the IB
 *   node set (brute force), marker positions and velocities (closed-form
 *   rotation, finite differences), the IB correction (D (u* + du) = U_B),
 *   dense vs sparse assembly (bit-identical), symmetric-airfoil zero lift,
 *   mirrored-plunge lift, the force-row layout and erranalysis;
 *   boundary_condition1's characteristic update against the textbook LODI
 *   equations (within 1% of the update), and to round-off (1e-13 of the
 *   update) against the same equations with the reference's two readings
 *   written in -- the velocity step divided by the already updated density,
 *   1/(2 c^2) as the literal 1.5 -- each of which the bar tells from its
 *   alternative; its NEE nodes against the second-order equilibrium and the
 *   per-moment (1 - s) scaling; the stale ring f0 against the closed-form
 *   initial equilibrium D2Q9feq(1, U, 0) with REAL(4) weights, bit for bit
 *   after stepping; every wall node (corners included) and the edge columns
 *   as the NEE of a state recomputed from those closed forms
 *   (tests/test_oracle_boundary.py); the IB system's conditioning and the
 *   spread among exact LU/L D L^T solvers (tests/test_ib_conditioning.py);
 *   post-processing, body dynamics and move_mesh (tests/test_oracle_post.py,
 *   test_oracle_dynamics.py).  The oracle is built with -fno-builtin-sin
 *   -fno-builtin-cos: its sin and cos of the pitch angle stay two libm
 *   calls, as the library's host code makes them, instead of one sincos,
 *   which differs from them in the last bit (tests/test_oracle.py).
 * PARITY UNPINNED:
 *   - the round-off order of the reference's cuBLAS DGEMM/DGEMV and cuSOLVER
 *     DGETRF/DGETRS calls (NVHPC 23.11 / CUDA 12.3, closed source): solver=1
 *     restates them as textbook LAPACK; the GPU path and solver=0 use a banded
 *     L D L^T of the same SPD system.  At cond(D D^T) ~ 1e8 (128 markers) to
 *     2.8e9 (1024 markers) two valid LU codes already differ by about as much
 *     as L D L^T and the textbook LU do, so this order cannot be pinned.
 *   The two readings of the characteristic update above come from the
 *   reference's statements (boundary_condition.f90:413-569); the reference
 *   itself does not build here (it needs mpi, cublas and cusolverDn
 *   modules), so they rest on reading it.
 */
#ifndef LBMIB_ORACLE_H
#define LBMIB_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int nx, ny;            /* lattice nodes, mesh%Nx, mesh%Ny (ini_mesh.f90:22-23) */
    double ox, oy;         /* physical origin of node (1,1) (ini_mesh.f90:36-37) */
    double scale;          /* lattice spacing in chord units, mesh%scale (ini_mesh.f90:21) */
    double U;              /* inflow lattice velocity UU (main.f90:11-12) */
    double xc, yc;         /* body reference point (main.f90:85) */
    double xp, yp;         /* rotation axis offset applied to the polygon (main.f90:86, airfoil.f90:240-241) */
    double s[9];           /* MRT relaxation rates mesh%s (ini_mesh.f90:59-67) */
    int smooth;            /* 1: smoothed 3-point kernel, 0: Peskin 4-point (boundary_condition.f90:1530,1563-1600) */
    int n_wings;           /* wing_count (main.f90:18) */
    int pts_per_wing;      /* boundary_points_per_wing, closed polygon (main.f90:21) */
    int solver;            /* 0: L D L^T of D D^T in zig-zag order; 1: dense LU with partial pivoting of D D^T S */
    int dense_ib;          /* 1: literal dense Nseg x Nadj matrices; 0: box-sparse evaluation (bit-identical) */
    int n_threads;         /* OpenMP threads for the lattice passes (results do not depend on it) */
} orc_params;

typedef struct orc_case orc_case;

/* Rigid-body dynamics of wing 1 driven by the IB force (solve_dynamic_equations,
 * main.f90:103; boundary_condition.f90:1694-1719, 3168-3202) and the domain
 * shift that follows the body (main.f90:370-388, move_mesh.f90). */
typedef struct orc_dynamics {
    int enabled;           /* solve_dynamic_equations */
    int rk4;               /* if_rk4 (boundary_condition.f90:1527-1529): 1 rk4, 0 explicit update */
    int closed_boundary;   /* main.f90:101: mass solid_den*area_airfoil/scale^2 (1) or solid_den/scale^2 (0) */
    int reserved;
    double solid_den;      /* main.f90:123 */
    double area_airfoil;   /* main.f90:125 */
    double par1, par2;     /* main.f90:128-129 */
    double alphad;         /* main.f90:76, radians */
} orc_dynamics;

/* init (solution.f90:1-72) + read_boundary's post-processing (airfoil.f90:235-241)
 * + build_motion_step_table (mode.f90:245-269).
 * xb0/yb0: n_wings*pts_per_wing polygon points in chord units, body frame, as
 * read with '(2f14.8)' (airfoil.f90:229).  motion: rows x ncols row-major table
 * in the motion_params.dat column order (mode.f90:161-172), ncols = 1+9*n_wings.
 * Returns NULL and fills err (>=256 bytes) on invalid input. */
orc_case* orc_create(const orc_params* p, const double* xb0, const double* yb0,
                     const double* motion, int rows, int ncols, char* err);
void orc_destroy(orc_case* c);

/* Run nsteps timesteps of the main loop body (main.f90:352-359): erranalysis
 * when n >= 2, then evolution(n).  Returns 0, or <0 on a solver failure. */
int orc_step(orc_case* c, int nsteps);

/* Steps taken so far, motion_end_time, and the reference's loop bound
 * (the last n with (n-1)*scale*U <= end_time + 1e-12, main.f90:352-353). */
int orc_steps_done(const orc_case* c);
int orc_max_steps(const orc_case* c);

/* Fields (Fortran column-major, x fastest): rho,u,v: nx*ny; f,f0: 9*nx*ny
 * population-planar.  Any pointer may be NULL. */
void orc_get_fields(const orc_case* c, double* rho, double* u, double* v,
                    double* f, double* f0);
/* Overwrite the populations f (9*nx*ny) -- test hook (restart upload). */
void orc_set_f(orc_case* c, const double* f);

/* Force-row history (solution.f90:349-376): row n holds time, per wing
 * (Cx,Cy,CM), totals (Cx,Cy,CM), per wing 9 kinematics = 4+12*nw columns. */
int orc_history_cols(const orc_case* c);
int orc_get_history(const orc_case* c, double* rows, int max_rows);

/* erranalysis value err (solution.f90:74-127) as of the last step. */
double orc_err(const orc_case* c);

/* IB diagnostics of the last step: node count, node list (0-based i,j,wing),
 * glx/gly (velocity corrections), rho at nodes, marker midpoints. */
int orc_ib_count(const orc_case* c);
void orc_ib_nodes(const orc_case* c, int* pi, int* pj, int* wing,
                  double* glx, double* gly, double* rho_adj);
int orc_n_segments(const orc_case* c);
void orc_markers(const orc_case* c, double* xb2, double* yb2, double* ub, double* vb,
                 double* arc_l);
/* Rebuild the IB node set from given midpoints (test hook on move_boundary). */
int orc_build_ib_set(const orc_case* c, const double* xb2, const double* yb2,
                     int* pi, int* pj, int* wing, int max_nodes);

/* Dynamics: set before the first step; state per wing in the restart file's
 * order (solution.f90:1650-1652): body_rel_a, _a_prev, _v, _v_prev, _s,
 * _s_prev, _pos, _theta_prev, _theta, _temp_theta_prev, _temp_theta (each
 * n_wings values), then movenum; motion rows of airfoil_motion.dat
 * (solution.f90:378-388), 1 + 3*n_wings per step. */
int orc_set_dynamics(orc_case* c, const orc_dynamics* d);
int orc_body_state(const orc_case* c, double* out);
int orc_motion_rows(const orc_case* c, double* rows, int max_rows);
/* Test hooks on boundary_condition1's pieces: one NEE edge node
 * (boundary_condition.f90:444-483), f = M^-1 [meq(edge) + (1-s)(M f0n -
 * meq(nb))], and the characteristic update of an inflow (side 0, 418-440) or
 * outflow (side 1, 500-521) edge node; in = {rho, u, v} of the edge node, of
 * the next node inwards and of the one after, out = the updated rho, u, v. */
void orc_nee_node(const double s[9], double rhe, double ue, double ve, double rhn, double un, double vn,
                  const double f0n[9], double out[9]);
void orc_char_update(int side, double cs, const double in[9], double out[3]);
/* Test hook: one move_mesh of the current state (dr = 1 or -1). */
void orc_move_mesh(orc_case* c, int dr);

/* Post-processing: output's vorticity column (solution.f90:181-210), one
 * vor_mom_rate call (solution.f90:1841-1900) and cal_sur_cp
 * (solution.f90:1912-2118). */
void orc_vorticity(int nx, int ny, double UU, double scale, const double* u, const double* v, double* vor);
void orc_vorticity_rate(int nx, int ny, double UU, double scale, const double* u, const double* v,
                        double* vor, double* vor1, double* dvordt);
int orc_surface_pressure(const orc_case* c, const double* pre, double* cp);

/* ---- component-level functions (unit tests) ---- */
void orc_weights(double w[9]);                    /* D2Q9model.f90:24-32 (REAL(4) weights) */
void orc_minv(double minv[81]);                   /* nizhen on matm (solution.f90:17-27,1462-1517), row-major */
double orc_asin(double x);                        /* asin used by the delta kernel */
double orc_delta(double r, int smooth);           /* one-axis kernel (boundary_condition.f90:1563-1600) */
double orc_cs(void);                              /* sqrt(3.0)/3.0d0 with REAL(4) sqrt (boundary_condition.f90:304) */
double orc_feq(double rh, double u1, double v1, int k);  /* D2Q9feq (D2Q9model.f90:35-46), k 0-based */
/* one node of colission (solution.f90:573-643): f, rho, u, v, rates, Guo force -> f */
void orc_collide_node(const double f[9], double rho, double u, double v, const double s[9],
                      double gfx, double gfy, double out[9]);
/* Motion step table (mode.f90:245-269) on a dt grid: out has nsteps x ncols. */
int orc_motion_step_table(const double* motion, int rows, int ncols, double dt,
                          double* out, int max_steps);
/* Solve the IB system for a given dense symmetric G (n x n row-major), rhs b
 * (n x 2, column k at b+k*n), scaling S (arc lengths), with solver 0 (SPD
 * L D L^T in the given order perm) or 1 (LU of G*S with partial pivoting).
 * Result y satisfies G y = b (solver 0) or the c = S^-1 y equivalent is
 * returned already multiplied by S (solver 1), i.e. both return y = G^-1 b. */
int orc_solve(const double* G, const double* S, const double* b, int n,
              const int* perm, int solver, double* y);

#ifdef __cplusplus
}
#endif
#endif
