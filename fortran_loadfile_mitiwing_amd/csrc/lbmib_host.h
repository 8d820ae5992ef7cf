// lbmib_host.h -- host-side state of liblbmib shared by its translation
// units: the case handle, and the engine that drives the kernels of one or
// more cases on one device.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/lbmib.h"
#include "lbmib_kernels.h"

namespace lbmib {

int fail(int code, const char* fmt, ...);

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::lbmib::fail(LBMIB_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

struct Engine;

}  // namespace lbmib

struct lbmib_handle {
    lbmib_params p;
    lbmib::DevCase D;
    lbmib::Engine* own = nullptr;    // engine that steps this case alone
    lbmib_batch* batch = nullptr;    // batch the case belongs to, if any
    std::vector<double> in_xb0, in_yb0;   // polygon as given to lbmib_create
    std::vector<double> motion_in;        // motion table as given (rows x ncols)
    std::vector<double> xb0, yb0, arc_l;  // polygon after the (xc-xp, yc-yp) shift
    std::vector<int> perm, iperm;
    // kinematics
    bool have_motion = false;
    std::vector<double> step_rows;  // per step: the motion row used (ncols)
    int ncols = 0, max_steps = 0;
    double motion_end_time = 0.0;
    int steps_done = 0;
    bool geo_fresh = true;   // the next run builds the geometry of its first step's batch first
    int geo_built = -1;      // (valid while !geo_fresh) highest geometry batch built or enqueued
    int last_flag_step[2] = {0, -1};  // step whose macros each slot holds
    // rigid-body dynamics: parameters, state (wing-indexed, as the
    // reference's body_rel_* / body_extra_* arrays), domain shift count,
    // airfoil_motion rows of the steps taken
    lbmib_dynamics dyn{};
    double rel[11][8] = {};           // a, a_prev, v, v_prev, s, s_prev, pos, theta_prev, theta, tt_prev, tt
    double ext[3][8] = {};            // body_extra_v, _s, _pos
    double mom = 0.0;                 // raw moment of wing 1 in the last step
    int movenum = 0;
    double body_terms[32] = {};       // the offset terms the last step's kernels used
    std::vector<double> mrows;
    double* vort = nullptr;           // post-processing: vorticity plane,
    double* vort_hist = nullptr;      //   vor_mom_rate's vor, vor1 and dvordt planes,
    double* surf = nullptr;           //   surface pressure per marker
    std::vector<void*> allocs;        // device allocations to free
    long long device_bytes = 0;       // their total size
    int dev = 0;
};

namespace lbmib {

// How a window's kernels are launched; see lbmib_launch_shape in lbmib.h.
struct Shape {
    int npt = 1;
    int order = 0;
    int persist = 0;
    int ib_cus = 16;
    int graph = 0;
    int band = 0;
    int rhs = 1;
    int unit = 0;
    int fuse = 1;  // two-step passes over the far field
};

// Drives the kernels of a set of cases on one device.  Cases advance in
// windows (one call); within a window every kernel covers all cases in one
// launch, each case at its own step (CaseClock).  Streams, all at the
// default priority:
//   s0  lattice update (every node but the step's IB nodes, and the ring)
//   s1  triangular solves (reserved CUs), then the forced collide of the IB
//       nodes (or on s3, shape.band); with shape.rhs the right-hand sides too
//   s2  geometry of later batches, forces of earlier ones
//   s3  right-hand sides of the IB systems
struct Engine {
    int dev = 0;
    int ncu = 0;
    int B = LBMIB_GEOMETRY_BATCH;    // steps per geometry batch
    bool per_step = false;           // geometry of each step only, built after the step before (body dynamics)
    Shape shape;
    static constexpr int NSTREAM = 4;
    hipStream_t s[NSTREAM] = {};
    int streams_cus = -1;            // ib_cus the streams were made with
    static constexpr int RING = 64;  // per-step events of engine step q live in slot q % RING
    hipEvent_t ev_ib[RING] = {}, ev_step[RING] = {}, ev_rhs[RING] = {}, ev_solve[RING] = {};
    hipEvent_t ev_win[RING] = {};  // two-step passes: the window's first step done
    static constexpr int BRING = 8;  // per-boundary events
    hipEvent_t ev_first[BRING] = {}, ev_next[BRING] = {}, ev_fork = nullptr, ev_join[NSTREAM] = {};
    hipEvent_t ev_nodes[BRING] = {};  // a fresh batch's node sets built (its factorisation may still run)
    long long gstep = 0, gbound = 0;
    std::vector<lbmib_handle*> members;
    // device tables
    DevCase* dcases = nullptr;
    CaseClock* dclock = nullptr;
    int2* dtiles = nullptr;
    int* derr = nullptr;             // gathered error flags of the members
    // pinned host copies of the clock table and the gathered error flags
    // (asynchronous copies at a window's start and end)
    CaseClock* hclk_pin = nullptr;
    int* herr_pin = nullptr;
    size_t pin_cap = 0;
    hipEvent_t ev_clk = nullptr, ev_end[4] = {};
    size_t case_cap = 0, tile_cap = 0, err_cap = 0;
    int ntiles = 0;
    bool tiles_valid = false;
    std::vector<CaseClock> hclock;
    std::vector<unsigned char> tab_bytes;  // the case table as last uploaded
    // captured window (graph mode)
    hipGraphExec_t gexec = nullptr;
    int graph_len = 0;
    bool graph_valid = false;
    // two-step passes: workgroup tables of the far pass and of the window,
    // ring and window + ring launches over every member (see k_far2), and
    // the layout they were built for
    int4* dfar = nullptr;
    int2* dwin = nullptr;
    size_t far_cap = 0, win_cap = 0;
    int nfar = 0, nwin = 0, nring = 0, nwr = 0;  // dwin holds the window, ring and window + ring tables in turn
    std::vector<int> pass_key;
    // profiling: EPER events per sampled step or pass (see launch_pair), the
    // kind of each sample (0 single step, 1 pass, 2 unused)
    static constexpr int EPER = 7;
    std::vector<char> samp_kind;
    int prof_every = 0;
    std::vector<hipEvent_t> evpool, gevpool;
    int gev_used = 0;
    double tms[LBMIB_NTIMINGS] = {};
};

int engine_create(int dev, Engine** out);
void engine_destroy(Engine* E);
// every member advances min(n_steps, its remaining steps); synchronous
int engine_run(Engine* E, int n_steps);
int engine_set_shape(Engine* E, const Shape& s);
int shape_from_env(Shape& s);
// NaN guard of one case after a call (no-op unless params.nan_guard)
int nan_guard_check(lbmib_handle* h);
// The middle plane of two-step passes, allocated on first use.
int ensure_mid_plane(lbmib_handle* h);
// Plane holding the state after the case's last step.
inline int state_plane(const lbmib_handle* h) { return (h->steps_done + h->D.pofs) & 1; }

}  // namespace lbmib
