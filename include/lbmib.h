/*
 * lbmib.h -- C ABI of liblbmib.so, the MI355X (gfx950) implementation of the
 * reference solver's per-timestep hot path: D2Q9 MRT collide-and-stream,
 * implicit velocity-correction immersed boundary (interpolate / solve /
 * spread / force) and the prescribed marker kinematics, with the
 * characteristic/NEE outer boundaries.
 *
 * The reference keeps this path behind Fortran subroutines on module state;
 * each entry point below names the reference interface it replaces.  A
 * Fortran caller binds these with iso_c_binding (module lbmib in
 * fortran_loadfile_mitiwing_amd/fortran/lbmib_mod.f90).
 *
 * Conventions
 *   - All entry points return 0 on success and a negative LBMIB_E_* code on
 *     failure; lbmib_last_error() returns the message of the calling thread's
 *     last failure.  Nothing in the library calls exit/abort/stop.
 *   - Arrays crossing the boundary are plain host pointers in Fortran
 *     column-major order: node (i,j) of an nx x ny lattice is element
 *     (j*nx + i); populations are planar, f[k*nx*ny + j*nx + i] is the
 *     reference's f(i+1,j+1,k+1).
 *   - A handle owns its device memory (allocated once in lbmib_create /
 *     lbmib_set_motion_table, nothing per step) and, from its first step, four
 *     HIP streams.  A handle is not thread-safe; distinct handles may be used
 *     concurrently.
 *   - Calls are synchronous: host arrays are only touched during the call.
 */
#ifndef LBMIB_H
#define LBMIB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBMIB_OK 0
#define LBMIB_E_ARG -1       /* invalid argument */
#define LBMIB_E_HIP -2       /* HIP runtime failure (no device, out of memory, launch) */
#define LBMIB_E_GEOMETRY -3  /* an immersed-boundary node left the interior margin */
#define LBMIB_E_CAPACITY -4  /* a fixed-capacity IB list overflowed */
#define LBMIB_E_SOLVE -5     /* IB system not positive definite */
#define LBMIB_E_STATE -6     /* call out of order (e.g. step before a motion table) */
#define LBMIB_E_NAN -7       /* NaN guard tripped */

#define LBMIB_PREC_FP64 0    /* plain FP64 drop-in path */
#define LBMIB_PREC_FP32DEV 1 /* FP32 populations stored as deviations from w_i*rho0 */

/* Steps per geometry batch: the immersed-boundary geometry (markers, node
 * set, interpolation weights, band of D D^T and its factors) of this many
 * steps is built in one launch, a batch ahead of the steps that use it. */
#define LBMIB_GEOMETRY_BATCH 16

/* The immersed-boundary system G y = b (G = D D^T, banded in zig-zag marker
 * order) is solved with a banded L D L^T.  CHAIN runs its triangular sweeps
 * as one sequential chain, in the oracle's operation order (bit-exact);
 * SEPARATOR splits the band at a separator of bandwidth rows into two halves
 * whose chains run concurrently, with the separator's small Schur complement
 * in between (half the chain length, a different round-off order).  AUTO
 * takes CHAIN for LBMIB_PREC_FP64 and SEPARATOR for LBMIB_PREC_FP32DEV (when
 * the system has at least 2*64 + bandwidth rows). */
#define LBMIB_SOLVER_AUTO 0
#define LBMIB_SOLVER_CHAIN 1
#define LBMIB_SOLVER_SEPARATOR 2

typedef struct lbmib_params {
    int32_t nx, ny;        /* lattice nodes, mesh%Nx/mesh%Ny (ini_mesh.f90:22-23) */
    double ox, oy;         /* physical position of node (1,1) (ini_mesh.f90:36-37) */
    double scale;          /* lattice spacing in chords, mesh%scale (ini_mesh.f90:21) */
    double U;              /* inflow lattice velocity UU (main.f90:12) */
    double xc, yc;         /* body reference point (main.f90:85) */
    double xp, yp;         /* rotation-axis offset of the polygon (main.f90:86) */
    double s[9];           /* MRT relaxation rates mesh%s (ini_mesh.f90:59-67) */
    int32_t smooth;        /* 1: smoothed 3-point kernel (reference default), 0: Peskin 4-point
                              (boundary_condition.f90:1530) */
    int32_t precision;     /* LBMIB_PREC_* */
    int32_t n_wings;       /* wing_count, 1..8 (main.f90:18) */
    int32_t pts_per_wing;  /* closed polygon points per wing (main.f90:21) */
    int32_t device;        /* HIP device ordinal */
    int32_t nan_guard;     /* 1: check the populations for NaN after every lbmib_step call */
    int32_t ib_solver;     /* LBMIB_SOLVER_* */
    int32_t reserved[5];
} lbmib_params;

typedef struct lbmib_handle lbmib_handle;

/* Replaces ini_mesh + init + read_boundary's post-processing
 * (ini_mesh.f90:1-80, solution.f90:1-72, airfoil.f90:207-242): allocates the
 * HBM-resident lattice, sets f = feq(rho0=1, u=U, v=0) with the reference's
 * REAL(4) weights, and takes the marker polygons.  xb0/yb0 hold
 * n_wings*pts_per_wing points in chord units (body frame, as read with
 * '(2f14.8)' from boundary.dat / boundary<k>.dat). */
int lbmib_create(const lbmib_params* p, const double* xb0, const double* yb0,
                 lbmib_handle** out);

/* Replaces load_motion_table + build_motion_step_table(scale*UU)
 * (mode.f90:68-142, 245-269): rows x ncols row-major, ncols = 1+9*n_wings,
 * columns time, then per wing angle, omega, alpha-dot, sa, sau, saac, sb, sbv,
 * sbac (mode.f90:161-172).  Must be called before lbmib_step. */
int lbmib_set_motion_table(lbmib_handle* h, const double* rows, int32_t nrows, int32_t ncols);

/* Replaces n_steps iterations of the main-loop body (main.f90:352-359):
 * mode -> move_boundary -> streaming -> physical_variables ->
 * immersed_boundary_Wu -> colission -> coefficients -> boundary_condition1
 * (solution.f90:269-410).  Steps continue from the last one taken. */
int lbmib_step(lbmib_handle* h, int32_t n_steps);

/* Steps taken; and the reference loop bound for the loaded motion table
 * (last n with (n-1)*scale*U <= motion_end_time + 1e-12, main.f90:352-353). */
int lbmib_steps_done(const lbmib_handle* h, int32_t* n);
int lbmib_max_steps(const lbmib_handle* h, int32_t* n);

/* Normalised per-wing coefficients of the last step (solution.f90:327-330):
 * wing_force[2*w+0] = C_x, wing_force[2*w+1] = C_y, wing_moment[w] = C_M. */
int lbmib_get_forces(lbmib_handle* h, double* wing_force, double* wing_moment);

/* Force-row history, one row per step taken, 4+12*n_wings columns in the
 * force.dat order (solution.f90:349-376): time, per wing (C_x, C_y, C_M),
 * totals (C_x, C_y, C_M), per wing (sa, sau, saac, sb, sbv, sbac, angle,
 * omega, alpha-dot).  Copies min(max_rows, steps) rows; *n_rows = steps. */
int lbmib_get_history(lbmib_handle* h, double* rows, int32_t max_rows, int32_t* n_rows);

/* Replaces the `update host` + output/saveflowfield reads (main.f90:408-415,
 * solution.f90:129-238, 1604-1707): rho/u/v (nx*ny each) as the reference
 * holds them after the last step, and the populations f (9*nx*ny).  Any
 * pointer may be NULL. */
int lbmib_download(lbmib_handle* h, double* rho, double* u, double* v, double* f);

/* Replaces readflowfield's population restore (solution.f90:1709-1839):
 * uploads f (9*nx*ny); the next step streams from it. */
int lbmib_upload_state(lbmib_handle* h, const double* f);

/* Replaces readflowfield's step counter (cycle1, solution.f90:1750): the
 * next lbmib_step takes step n + 1, with that step's kinematics and
 * geometry (0 <= n <= lbmib_max_steps).  Call it before
 * lbmib_upload_state to restart a run from a saved state.  Fields and the
 * residual of earlier steps are no longer available; history rows up to n
 * stay as they were.  Not for a case that belongs to a batch. */
int lbmib_set_step(lbmib_handle* h, int32_t n);

/* Replaces erranalysis (solution.f90:74-127): ||u-u0|| / (||u|| + 1e-30)
 * over the interior between the last two steps -- the value the reference
 * prints at the start of the following step. */
int lbmib_residual(lbmib_handle* h, double* err);

/* IB node set of the last step (move_boundary's adjoint_p / adjoint_wing_id,
 * airfoil.f90:286-335) and its velocity corrections glx/gly
 * (boundary_condition.f90:1639-1650).  Arrays sized by *n from a first call
 * with NULL pointers. */
int lbmib_ib_nodes(lbmib_handle* h, int32_t* n, int32_t* pi, int32_t* pj, int32_t* wing,
                   double* glx, double* gly);

/* The lattice window of the run: the rectangle of nodes (0-based origin i0,
 * j0; width w, height hgt) that holds every IB node of every step of the
 * motion table (move_boundary's node sets, airfoil.f90:276-335, with a margin
 * of 6 nodes).  Its nodes step one step at a time in two-step passes; the
 * rest, the far field, two at a time.  Needs the motion table. */
int lbmib_lattice_window(const lbmib_handle* h, int32_t* i0, int32_t* j0, int32_t* w, int32_t* hgt);

/* Marker polygon of every wing at the last step taken, n_wings *
 * pts_per_wing points (move_boundary's xb1/yb1, airfoil.f90:256-263; the
 * points of step 1 before any step), for the airfoil zone of the flow-field
 * output (solution.f90:219-233). */
int lbmib_markers(const lbmib_handle* h, double* xb1, double* yb1);

/* Rigid-body dynamics of wing 1, driven by the IB force (replaces
 * solve_dynamic_equations and its parameters, main.f90:76,101-129; the
 * dynamics block of immersed_boundary_Wu, boundary_condition.f90:1676-1731;
 * rk4, 3168-3202; update_body_extra_from_rel, mode.f90:285-307), and the
 * domain shift that follows the body one node at a time (main.f90:370-388,
 * move_mesh.f90).  The body offsets enter the markers (airfoil.f90:261-262),
 * their velocities (boundary_condition.f90:1546-1548) and the moment arm
 * (1672).  Set before lbmib_set_motion_table.  A case with the dynamics on
 * takes its steps one at a time (each step's geometry follows the force of
 * the step before) and cannot join a batch. */
typedef struct lbmib_dynamics {
    int32_t enabled;          /* solve_dynamic_equations */
    int32_t rk4;              /* if_rk4: 1 the four-stage update, 0 the explicit one */
    int32_t closed_boundary;  /* mass solid_den*area_airfoil/scale^2 (1) or solid_den/scale^2 (0) */
    int32_t reserved;         /* 0 */
    double solid_den;         /* > 0 */
    double area_airfoil;
    double par1, par2;        /* pitch equation coefficients */
    double alphad;            /* radians */
} lbmib_dynamics;
int lbmib_set_dynamics(lbmib_handle* h, const lbmib_dynamics* d);

/* The body state as the restart file holds it (saveflowfield,
 * solution.f90:1650-1653): body_rel_a, _a_prev, _v, _v_prev, _s, _s_prev,
 * _pos, _theta_prev, _theta, _temp_theta_prev, _temp_theta (n_wings values
 * each), then movenum: 11*n_wings + 1 values.  Set it after lbmib_set_step
 * to resume a run (readflowfield, 1750-1760). */
int lbmib_get_body_state(const lbmib_handle* h, double* state, int32_t n);
int lbmib_set_body_state(lbmib_handle* h, const double* state, int32_t n);

/* The stale f0 of the first and last lattice columns (f0(1,:,:) and
 * f0(Nx,:,:) of the restart file's edge records, solution.f90:1663-1668,
 * 1750-1756), each ny*9 values with j fastest.  They hold the initial
 * equilibrium until a domain shift refills one of them.  Either pointer may
 * be NULL. */
int lbmib_get_edge_f0(lbmib_handle* h, double* first, double* last);
int lbmib_set_edge_f0(lbmib_handle* h, const double* first, const double* last);

/* airfoil_motion.dat rows (solution.f90:378-388), one per step taken:
 * time, then per wing 0, body_extra_v/UU, body_extra_s*scale (1 + 3*n_wings
 * columns; zero offsets with the dynamics off). */
int lbmib_get_motion_rows(const lbmib_handle* h, double* rows, int32_t max_rows, int32_t* n_rows);

/* Replaces the vorticity of output (solution.f90:181-210) and vor_mom_rate
 * (solution.f90:1841-1900, main.f90:391), computed on the device from the
 * last step's velocity field.  vor (nx*ny) gets dv/dx - du/dy in units of
 * U/scale, one-sided on the edge columns and rows.  With rate != 0 the call
 * is also one vor_mom_rate call: the vorticity history of the last two calls
 * (zero before the first, and again after lbmib_set_step) shifts by one and
 * dvordt (nx*ny, may be NULL) gets the second-order backward difference on
 * the interior, zero on the edges; call it once per step for the reference's
 * per-step rate.  vor may be NULL (the result stays on the device).  The
 * first call allocates the planes (one, four with rate). */
int lbmib_vorticity(lbmib_handle* h, int32_t rate, double* vor, double* dvordt);

/* Replaces cal_sur_cp (solution.f90:1912-2118): the pressure rho/3 of the
 * last step's IB nodes interpolated onto each marker midpoint with the
 * smoothed kernel, one value per marker segment (n >= the case's segment
 * count, n_wings * (pts_per_wing - 1)).  Needs smooth = 1. */
int lbmib_surface_pressure(lbmib_handle* h, double* p, int32_t n);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): lbmib_surface_pressure with the same arguments and
 * results, staged through a fixed per-device buffer of the library instead
 * of a per-handle one allocated on first use; it calls no allocating
 * routine.  At most LBMIB_SCRATCH_SEGMENTS marker segments (LBMIB_E_ARG
 * beyond).  The buffer is shared by every call on one device: calls that
 * overlap on one device race. */
#define LBMIB_SCRATCH_SEGMENTS 2048
int lbmib_surface_pressure_noalloc(lbmib_handle* h, double* p, int32_t n);

/* Device timing of the hot path (the reference's cpu_time table,
 * main.f90:338-369, 420-433), measured with HIP events on the streams the
 * kernels run on.  lbmib_set_profiling(h, every) resets the counters and,
 * for every > 0, records the phase events of every every-th step and of each
 * geometry batch (0 turns profiling off; events cost a few microseconds per
 * sampled step).  lbmib_get_timings copies min(n, LBMIB_NTIMINGS) values:
 *   [0] ms of all steps taken while profiling   [1] those steps
 *   [2] sampled steps                            [3] ms in the lattice update
 *       (fused stream + collide + outer boundary, all nodes but the IB nodes)
 *   [4] ms from the IB right-hand side to the end of the triangular solves
 *   [5] ms in the forced collide of the IB nodes ([3]-[5] summed over the
 *       sampled steps)
 *   [6] ms of geometry batches before the factorisation, [7] ms of their
 *       factorisations, [8] batches timed (geometry overlaps the steps)
 *   [9] ms in the vorticity kernel of lbmib_vorticity, [10] its calls timed
 *       (a handle's own counters; zero for a batch)
 *   [11] ms in two-step far-field passes, [12] passes timed
 *   [13] ms in the lattice window's single steps of those passes, [14] timed
 *   [15] single-step lattice updates timed ([3] is summed over them; [4] and
 *        [5] over all sampled steps, [2]). */
#define LBMIB_NTIMINGS 16
int lbmib_set_profiling(lbmib_handle* h, int32_t every);
int lbmib_get_timings(lbmib_handle* h, double* t, int32_t n);

/* Device memory the handle holds (allocated by lbmib_create and
 * lbmib_set_motion_table, plus the planes of the post-processing calls once
 * used), for admission by HBM footprint; and the device's free and total
 * memory.  No reference counterpart. */
int lbmib_device_bytes(const lbmib_handle* h, int64_t* bytes);
int lbmib_mem_info(int32_t device, int64_t* free_bytes, int64_t* total_bytes);

/* Releases the handle and its device memory (removing it from its batch). */
int lbmib_destroy(lbmib_handle* h);

/* ---- launch shape and the batched form ----
 *
 * The reference runs one case per process and has no counterpart of the
 * entry points below; they serve parameter sweeps (many kinematics cases of
 * uneven lattice sizes advanced together on one GPU) and a serving loop that
 * admits and retires cases while others run.  A case keeps its own handle:
 * its inputs, state, history, download and residual work as above.  While it
 * belongs to a batch it is stepped only by lbmib_batch_step (lbmib_step
 * returns LBMIB_E_STATE).  Every kernel of a batch step covers all member
 * cases in one launch, each at its own step; the per-case arithmetic is the
 * single-case path's, so a case's results are identical alone and in a
 * batch. */

/* How the steps are launched (never changes results). */
typedef struct lbmib_launch_shape {
    int32_t nodes_per_thread; /* lattice tiles of 256 x nodes_per_thread nodes of one row: 1 or 2 */
    int32_t tile_order;       /* 0: case-major tile table; 1: tiles dealt to the 8 XCDs in row bands */
    int32_t persistent;       /* 0: one workgroup per tile; k > 0: k workgroups per CU looping over tiles */
    int32_t ib_cus;           /* CUs reserved for the triangular solves (multiple of 8; 0: none) */
    int32_t graph;            /* 1: windows of 16 steps captured once as a HIP graph and replayed */
    int32_t ib_collide_shared; /* 1: the forced collide of the IB nodes runs on the shared CUs, not the
                                  solves' reserved ones */
    int32_t ib_rhs_reserved;  /* 1: the right-hand sides of the IB systems run in the solves' stream (on
                                 their reserved CUs), not beside the lattice update */
    int32_t solve_unit;       /* cap on the steps the solves stage per LDS ring unit: 0 (none: the
                                 largest that fits), 16, 32 or 64 */
    int32_t fuse_steps;       /* 1: an engine of one case (a handle stepped alone, or a batch of one)
                                 with prescribed motion takes its steps in pairs: the nodes outside
                                 the body's lattice window take both steps in one pass, reading and
                                 writing each population once; 0: one step at a time */
} lbmib_launch_shape;

/* Launch shape of a single handle's steps (default 1, 0, 0, 16, 0, 0, 1, 0, 1;
 * the environment variable LBMIB_IB_CUS overrides ib_cus). */
int lbmib_set_shape(lbmib_handle* h, const lbmib_launch_shape* s);
int lbmib_get_shape(const lbmib_handle* h, lbmib_launch_shape* s);

typedef struct lbmib_batch lbmib_batch;

/* An empty batch on a device (its own HIP streams). */
int lbmib_batch_create(int32_t device, lbmib_batch** out);

/* Admits a case (motion table set, same device and precision as the other
 * members, and at the same step modulo LBMIB_GEOMETRY_BATCH as them, e.g.
 * a fresh case beside members at multiples of LBMIB_GEOMETRY_BATCH steps). */
int lbmib_batch_add(lbmib_batch* b, lbmib_handle* h);

/* Retires a case from the batch (it may then step alone again). */
int lbmib_batch_remove(lbmib_batch* b, lbmib_handle* h);

int lbmib_batch_size(const lbmib_batch* b, int32_t* n);

/* Advances every member min(n_steps, its remaining motion-table steps)
 * steps; synchronous.  Members that run out of steps idle. */
int lbmib_batch_step(lbmib_batch* b, int32_t n_steps);

int lbmib_batch_set_shape(lbmib_batch* b, const lbmib_launch_shape* s);
int lbmib_batch_get_shape(const lbmib_batch* b, lbmib_launch_shape* s);

/* Phase timings of the batch's steps, as lbmib_set_profiling /
 * lbmib_get_timings for a handle (graph replays are not sampled). */
int lbmib_batch_set_profiling(lbmib_batch* b, int32_t every);
int lbmib_batch_get_timings(const lbmib_batch* b, double* t, int32_t n);

/* Autotuner: times launch shapes on copies of the members (same lattices,
 * markers and motion tables, own state; the members are not advanced), one
 * coordinate at a time -- lattice tile shape, order and persistent grid, then
 * the CUs reserved for the solves with the placement of the IB-node collide
 * -- over trial_steps
 * steps each, applies the fastest to the batch and returns it in *best.
 * report (may be NULL) receives every trial as JSON. */
int lbmib_batch_autotune(lbmib_batch* b, int32_t trial_steps, lbmib_launch_shape* best, char* report,
                         int32_t report_len);

/* Releases the batch; its members remain valid handles. */
int lbmib_batch_destroy(lbmib_batch* b);

/* Message of the calling thread's last failure ("" if none). */
const char* lbmib_last_error(void);

/* Library version string. */
const char* lbmib_version(void);

#ifdef __cplusplus
}
#endif
#endif
