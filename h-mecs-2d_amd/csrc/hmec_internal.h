// Internal declarations of libhmec_hip: context, device state, kernel launch
// wrappers and the multifrontal solver plan.
#pragma once
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "hmec.h"

#define HMEC_HIP(call)                                                            \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            hmec_set_last_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return HMEC_E_HIP;                                                    \
        }                                                                         \
    } while (0)

#define HMEC_TRY(call)              \
    do {                            \
        int rc_ = (call);           \
        if (rc_ != HMEC_OK) return rc_; \
    } while (0)

// A linear solve whose final ||b - A x||inf / ||b||inf exceeds this bound is
// reported as HMEC_E_NOCONV (the tolerance the parity and precision targets
// use); below it a refinement that missed its tighter target is accepted.
constexpr double kAcceptResidual = 1e-10;

// Length of the per-step series of the snapshot files (constants.hpp:14
// num_timesteps, the reference's maximum number of timesteps).
constexpr int64_t kNumTimesteps = 100000;
// partial extrema per workgroup of the series reduction, kept after the series
constexpr int kSeriesBlocks = 512;

void hmec_set_last_error(const std::string &s);
double hmec_seeded_normal(uint64_t seed, int64_t idx, unsigned stream);

// Number of live scatter accumulators per grid and the marker property
// streams the scatter reads (derived once per marker set).
enum MarkerProp {
    MP_ETAMM = 0,   // effective viscosity (clamped)
    MP_RHOMM,       // total density
    MP_KKKMM,       // permeability (porosity-scaled, clamped)
    MP_TMARK,       // rock type
    MP_INVG,        // 1 / shear modulus
    MP_ETAMM0,      // matrix viscosity (clamped)
    MP_COHT, MP_FRIT, MP_COHC, MP_FRIC, MP_DILC,
    MP_PORM,        // porosity
    MP_ETASMM0,     // solid viscosity (bulk viscosity numerator)
    MP_INVETAD,     // 1 / Darcy viscosity = kkk / etaf
    MP_RHOFM,       // fluid density
    MP_NPROP
};

// Dense-front multifrontal plan (symbolic analysis, host side), see
// solver_symbolic.cpp.
struct FrontPlan {
    int dpn = 4;                        // unknowns per node of the factored system
    // per front
    std::vector<int> np, nu;            // pivot / update node counts
    std::vector<int64_t> off;           // offset of the n x n front in the arena
    std::vector<int64_t> voff;          // offset of the front in the per-front vector arena
    std::vector<int> level;             // 0 = leaf
    std::vector<int> parent;            // -1 = root
    std::vector<std::vector<int>> children;
    std::vector<std::vector<int>> nodes;  // local node list: pivots then update nodes
    std::vector<int64_t> ea_off;        // offset into ea_map of this front's (child) map
    // global
    int nfronts = 0, nlevels = 0, max_front = 0;
    int64_t arena = 0;                  // doubles in the front arena
    int64_t varena = 0;                 // doubles in the vector arena
    double flops = 0;                   // dense flops of one factorization
    int64_t factor_nnz = 0;             // entries of the factor blocks
    std::vector<std::vector<int>> levels;   // fronts per level
    // maps
    std::vector<int64_t> nz_dst;        // CSR entry -> arena offset
    // the CSR entries of each front column (rows ascending): column c of
    // front f has entries [col_ptr[col_base[f] + c], col_ptr[col_base[f] + c + 1])
    std::vector<int64_t> col_base;      // nfronts + 1
    std::vector<int32_t> col_ptr;       // sum of front orders + 1
    std::vector<int32_t> col_row;       // local row of the entry
    std::vector<int32_t> col_nz;        // CSR entry
    std::vector<int32_t> ea_map;        // child update node -> parent local node
    std::vector<int32_t> dof_local;     // unused placeholder
    std::vector<int32_t> node_front;    // front that eliminates each node
    std::vector<int32_t> node_local;    // local node index inside that front
};

// dpn: unknowns per node of the system the plan factors (6: the coupled
// system, 4: the Darcy-condensed system of CondPlan)
int hmec_plan_build(int Nx1, int Ny1, const std::vector<int32_t> &rowp,
                    const std::vector<int32_t> &coli, FrontPlan &plan, int dpn = 4);
// position of variable v inside a node block of a front (dpn 4 or 6)
int hmec_var_pos(int dpn, int v);

// Darcy condensation (solver_condense.cpp).  The Darcy unknowns (vxD, vyD)
// of the coupled system are eliminated exactly: every Darcy row holds its
// diagonal plus either up to three non-Darcy entries (its node's velocity and
// two Pf) or one other Darcy unknown (boundary rows), so x_D = g + E x_R
// with E = -A_DD^-1 A_DR of at most three entries per row, and the factored
// system is S = A_RR + A_RD E on the 4 remaining unknowns per node (index
// node*4 + v, v = 0 Pt, 1 vx, 2 vy, 3 Pf).  S keeps the node graph of the
// coupled system (axis neighbours and one anti-diagonal pair).
struct CondPlan {
    int64_t N = 0, NS = 0, ND = 0;
    std::vector<int32_t> srowp, scoli;     // S pattern (condensed numbering)
    std::vector<int32_t> ssrc;             // per S entry: A entry of A_rc, -1: none
    std::vector<int32_t> s_of_full;        // N: condensed index, -1 for Darcy unknowns
    std::vector<int32_t> d_of_full;        // N: Darcy index, -1 otherwise
    std::vector<int32_t> full_of_s;        // NS: full index
    // per Darcy unknown d
    std::vector<int32_t> dglob;            // full index
    std::vector<int32_t> ddiag;            // A entry of its diagonal
    std::vector<int32_t> dnb, dnbdiag, dnbe;   // neighbour Darcy unknown (full index), its
                                               // diagonal entry, the coupling entry (-1: none)
    std::vector<int32_t> des, dec;         // 3 per d: A entries of the interior row's
                                           // non-Darcy part (own row or the neighbour's)
                                           // and their condensed columns (-1: none)
    // per condensed row: Darcy terms A_{r,d}
    std::vector<int32_t> rptr, raslot, rd;
    // per condensed row: S fill terms A_{r,d} E_{d,k} -> S entry
    std::vector<int32_t> tptr, taslot, tdk, tsslot;   // tdk = 3 d + k
};
int hmec_condense_build(int64_t nnodes, const std::vector<int32_t> &rowp,
                        const std::vector<int32_t> &coli, CondPlan &cp);

// Device-side task descriptors for the batched dense kernels.
struct FrontTask {
    int64_t off;    // arena offset of the front
    int n;          // front order
    int p;          // pivots
    int64_t doff;   // offset of the inverted diagonal blocks (2 x NB x NB per NB pivots)
    int64_t poff;   // offset of the front's partial sums in the solve scratch (per level)
    int64_t coff;   // offset of the front's pivot-block copy in the inversion scratch (per level)
};

// One workgroup's share of a level's front pull: rows [r0, r0 + nr) x
// columns [c0, c0 + nc) (at most kPullTile^2 entries; whole columns, so a
// tile is one contiguous piece of the front), with what the pull needs of the
// front and of its children's contribution blocks, and the tile's original
// entries ent[e0, e0 + ne).
struct PullTile {
    int64_t off;         // arena offset of the front
    int64_t coff[2];     // arena offsets of the children (slot order)
    int64_t e0;          // first original entry of the tile
    int32_t imoff[2];    // per child: parent local node -> child update node (-1: none)
    int n;               // front order
    int cn[2], cp[2];    // children's order and pivots
    int nch;             // children (0..2)
    int r0, c0, nr, nc, ne;
};
constexpr int kPullTile = 64;
// An original entry of a pull tile: its place in the tile (row | column << 12),
// the CSR entry of the condensed system, and that entry's global row and
// column (equilibration scales)
struct alignas(16) ColEnt {
    int32_t rc, nz, grow, gcol;
};

struct DevPlan {
    int nlevels = 0;
    // per level task arrays on the device
    std::vector<FrontTask *> d_tasks;     // fronts of the level
    std::vector<std::vector<FrontTask>> h_tasks;
    std::vector<PullTile *> d_pull_t;     // per level: the pull's tiles
    std::vector<int> n_pull_t;
    std::vector<int> maxp, maxn;          // per level maxima
    std::vector<int64_t> lvl_begin, lvl_end;  // arena range touched by the level
    int32_t *d_ea_map = nullptr;
    int32_t *d_imap = nullptr;            // PullTile::imoff maps
    ColEnt *d_pull_ent = nullptr;         // the pull tiles' original entries
    int32_t *d_pull_rc = nullptr;         // their places in the tiles (ColEnt::rc)
    int32_t *d_csc_ptr = nullptr;         // column pointers of the transposed pattern
    int32_t *d_csc_idx = nullptr;         // CSR entry index of each CSC entry
    int32_t *d_csc_row = nullptr;         // row of each CSC entry
    // solve: per front pivot/update dof -> global dof
    int32_t *d_front_dofs = nullptr;      // per front: n global dof ids (pivots, then update)
    int64_t *d_front_dof_off = nullptr;
    std::vector<int64_t> h_front_dof_off;
};

// Named device timer: event pairs recorded around launches on the context's
// stream, resolved lazily (no synchronization inside the timed region).
struct Timer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    std::vector<double> pending_work;
    std::vector<int> pending_count;   // launches covered by each interval
    std::vector<hipEvent_t> pool;
    hipEvent_t open_ev = nullptr;
    double total_ms = 0;
    double bytes = 0;     // algorithmic work (bytes or flops) of resolved samples
    int n = 0;
    bool open = false;
    std::string name;     // the name hmec_kernel_time and hmec_timing_select use
    int sel_gen = -1;     // selection generation sel was computed for
    bool sel = true;      // recorded under the current selection
    int64_t calls = 0;    // begin calls (sampling by the context's stride)
};

struct hmec_ctx_impl;
struct hmec_batcher;

// Device buffers owned by a context.
// device copies of the condensation structures and the condensed system
struct CondDev {
    int32_t *srowp = nullptr, *scoli = nullptr, *ssrc = nullptr;
    double *sval = nullptr;                  // S values (formed each factorization)
    int32_t *s_of_full = nullptr, *d_of_full = nullptr, *full_of_s = nullptr;
    int32_t *dglob = nullptr, *ddiag = nullptr, *dnb = nullptr, *dnbdiag = nullptr,
            *dnbe = nullptr, *des = nullptr, *dec = nullptr;
    int32_t *rptr = nullptr, *raslot = nullptr, *rd = nullptr;
    int32_t *tptr = nullptr, *taslot = nullptr, *tdk = nullptr, *tsslot = nullptr;
    double *e = nullptr;                     // 3 x ND: E of the current matrix
    double *dcoef = nullptr;                 // 3 x ND: a_dd, a_nn, a_dn of the current matrix
    double *rcoef = nullptr;                 // A_RD entries of the right-hand-side terms
    double *gv = nullptr;                    // ND: A_DD^-1 v_D of the current vector
    double *vs = nullptr;                    // NS: condensed right-hand side / solution
    double *rs = nullptr, *cs = nullptr;     // NS: equilibration of the active S factor
    int64_t snnz = 0;
};

struct DeviceState {
    double *fields = nullptr;        // HMEC_NFIELDS x stag_size
    double *mk = nullptr;            // 19 x M markers
    double *prop = nullptr;          // MP_NPROP x M derived properties
    int32_t *cell = nullptr;         // 4 x M anchor cells
    int32_t *perm = nullptr;         // 4 x M markers sorted by cell
    int32_t *cell_start = nullptr;   // 4 x (ncell + 1)
    double *sorted = nullptr;        // 31 x M: per grid, its scatter inputs in cell order
    void *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int32_t *keys_tmp = nullptr;     // 2 x M scratch for sorting
    double *coords = nullptr;        // x, y, xvx, yvx, xvy, yvy, xp, yp
    double *scal = nullptr;          // device scalars (ptscale, reductions)
    double *h_scal = nullptr;        // pinned host mirror
    // linear system
    int32_t *rowp = nullptr, *coli = nullptr;
    double *val = nullptr, *rhs = nullptr, *sol = nullptr;
    int32_t *trip_slot = nullptr;    // CSR slot of each triplet (emission order)
    int32_t *node_trip = nullptr;    // first triplet of each node (N/6 + 1)
    uint8_t *sec_off = nullptr;      // per node: first triplet of each row section (8 bytes)
    // solver
    double *arena = nullptr;         // FP64 fronts
    float *arena32 = nullptr;        // FP32 fronts
    double *vec = nullptr;           // per-front vector arena (solve)
    double *rscale = nullptr, *cscale = nullptr;
    double *work = nullptr;          // N-vector scratch (residual, correction)
    double *work2 = nullptr;
    int *counters = nullptr;         // perturbed pivots etc.
    double *dsylsq = nullptr;        // niterglobal + 1 iteration errors of the device controller
    double *series = nullptr;        // 6 x kNumTimesteps per-step series (timesumcur, dtcur,
                                     // maxvxsmod, minvxsmod, maxvysmod, minvysmod), then
                                     // 4 x kSeriesBlocks partial extrema
};

struct hmec_ctx {
    hmec_params P;
    int Nx, Ny, Nx1, Ny1;
    int64_t nb, ns;                  // basic / staggered node counts
    int64_t N, nnz, ntrip;
    int64_t M = 0;
    int precision = HMEC_PREC_FP64;
    hipStream_t stream = nullptr;
    hipEvent_t ev_scal = nullptr;    // after the device scalars' copy to the host
    int device = 0;
    DeviceState d;
    // host copies of the pattern
    std::vector<int32_t> h_rowp, h_coli;
    // scalar state of the time loop
    int timestep = 1, ynlast = 0, iterstep = 0, yndtdecrease = 1;
    double timesum = 0, dt = 0, dt00 = 0, dtx = 0, dty = 0, dtlapusta = 0, Vmax = 0, maxvxy = 0;
    double ptscale = 0;
    int64_t asm_lds_entries = 0;     // largest CSR range of an assembly workgroup
    bool rscale_ready = false;       // d.rscale holds the current values' row scales
    std::vector<double> DSYLSQ;
    bool markers_dirty = true;       // cell keys / sort need recomputing
    bool have_markers = false;
    // solver
    bool analyzed = false;
    FrontPlan plan;
    DevPlan dp;
    CondPlan cp;                     // Darcy condensation (host structures)
    CondDev dc;                      // ... and their device copies
    int factor_prec = -1;            // precision of the factor in the arena, -1: none
    bool factor_valid = false;
    int last_ir = 0, last_fallback = 0, last_perturbed = 0;
    double last_res = 0;
    int stale_solves = 0;            // FP32: solves since the last factorization
    double factor_dt = -1;           // dt of the system the factor was computed for
    // FP32 GMRES initial guesses (hmec_driver.cpp): per key (dt within 1 %,
    // Picard iteration it, -1 for any iteration) the last solution and the
    // one kept at the key in an earlier step, with the model times of their
    // steps; the dt of the last solve
    struct XGuess {
        double *x = nullptr, *x2 = nullptr, *x3 = nullptr;   // newest, and two earlier steps'
        double dt = 0, t = 0, t2 = 0, t3 = 0;
        bool has2 = false, has3 = false;
        int it = -1;
        bool quad = false;           // extrapolate quadratically (else linearly) next time
        int64_t used = 0;
    };
    static constexpr int NGUESS = 6;
    XGuess xguess[NGUESS];
    int64_t xguess_clock = 0;
    int xguess_used = -1;            // entry the current solve's guess came from
    double last_solve_dt = 0;
    int ir_max_stale = 12;           // FP32 policy: GMRES budget with a reused factor
    int ir_max_fresh = 40;           //              ... with a fresh factor
    double ir_tol = 1e-12;           //              target relative inf-norm residual
    int fp64_max_ir = 3;             // FP64 factor: refinement sweeps after the first solve
    double fp64_tol = 1e-16;         //              target (unscaled and row-equilibrated):
                                     //              rounding level, i.e. refine until the
                                     //              equilibrated residual stops halving
    hmec_batcher *batcher = nullptr; // server member: factorizations go through the batcher
    int batched_facs = 0;            // factorizations done inside a batch of > 1 members
    // statistics of the current step
    hmec_step_info acc;
    // timing
    bool timing = false;
    bool phase_sync = false;         // wait for the device at step phase boundaries
    Timer t_scatter, t_assemble, t_factor, t_solve;
    std::map<std::string, Timer> kt;    // per kernel class (schur, fwd, bwd, ...)
    std::vector<std::string> timing_sel;   // timers recorded (empty: all)
    int timing_gen = 0;
    int timing_stride = 1;                 // record every stride-th interval of each timer
};

// ---- kernel launch wrappers (kernels_phys.hip) ----
int k_marker_prep(hmec_ctx *c);          // derived props + anchor cells + sort
int k_scatter(hmec_ctx *c);
int k_step_begin_fields(hmec_ctx *c);
// dev: dt from the device controller, launches gated on its exit flag
int k_vep_pre(hmec_ctx *c, double dt, bool dev = false);
int k_scale(hmec_ctx *c, double dt, double *ptscale, bool dev = false);
int k_assemble(hmec_ctx *c, double dt, bool dev = false);   // ptscale from the device scalars
// device-side Picard / dt controller (csrc/controller.h)
int k_controller_begin(hmec_ctx *c);
int k_controller_commit(hmec_ctx *c);
bool k_controller_read(hmec_ctx *c);
int k_controller_stage(hmec_ctx *c, const hmec_iter_scalars &sc, int iterstep, bool *stop);
int k_update_launch(hmec_ctx *c, double dt, int iterstep, bool gated = false);
int k_exit_release(hmec_ctx *c);
int k_scalars_enqueue(hmec_ctx *c);
int k_scalars_wait(hmec_ctx *c, hmec_iter_scalars *out);
int hmec_update_dev(hmec_ctx *c, double dt, hmec_iter_scalars *out);
int k_series_step(hmec_ctx *c, double timesum, double dt);
constexpr int SC_PTSCALE_SLOT = 0;   // ptscale's index in the device scalars
// the device controller's exit flag: 0 go on, 1 the Picard loop exits, 2 the
// linear solve left a decision to the host (the update and the next
// iteration's prologue enqueued behind it do nothing)
constexpr int SC_EXIT_SLOT = 18;
constexpr double SC_EXIT_HOLD = 2.0;
// outcome of a solve's first GMRES step when it was decided on the device
// (GmState in solver_numeric.hip)
constexpr int SC_GMSTATE_SLOT = 19;
constexpr int HMEC_FP32_SLOTS = 3;   // FP32 factors kept by default (keyed by dt)
int k_update(hmec_ctx *c, double dt, hmec_iter_scalars *out);
int k_iter_commit(hmec_ctx *c, int ynpl, int converged);
int k_final_stress(hmec_ctx *c, double dt);
// series of the fault-row output streams: field, subtracted field (-1 none)
// and grid of each
struct FaultRows {
    static constexpr int NSERIES = 10;
    int f[NSERIES], sub[NSERIES], rows[NSERIES], cols[NSERIES];
};
// row line_fault of every series into dev_out[k * Nx1 + j]
int k_fault_rows_gather(hmec_ctx *c, const FaultRows &a, double *dev_out);
int k_history(hmec_ctx *c);
int k_advect(hmec_ctx *c, double dt);
int k_init_state(hmec_ctx *c);
int k_copy_fields(hmec_ctx *c, const int *dst, const int *src, int count);
int k_copy_field(hmec_ctx *c, int dst, int src);
// symbolic CSR pattern of the reference's triplet emission (host)
int hmec_build_pattern(hmec_ctx *c);

// ---- solver (solver_numeric.hip) ----
int sv_setup_device(hmec_ctx *c);
int sv_free_device(hmec_ctx *c);
int sv_factorize(hmec_ctx *c, int precision);
// HMEC_E_PIVOT when the last factorization met an unrepairable pivot (syncs)
int factor_check(hmec_ctx *c);
int sv_factorize_batch(hmec_ctx *const *cs, int n, int precision);
int sv_solve(hmec_ctx *c, int max_ir, double tol, hmec_solve_stats *st);
int sv_batch_factorize(hmec_ctx *const *cs, int n, int precision);
int sv_gmres(hmec_ctx *c, int max_it, double tol, hmec_solve_stats *st, int resume = 0);
// a solve's first GMRES step with its decision on the device (nothing waits);
// the state lands in the device scalars' SC_GMSTATE_SLOT
int sv_gmres_first(hmec_ctx *c, double tol);
int sv_gmres_first_result(hmec_ctx *c, int state, hmec_solve_stats *st);
int sv_gmres_first_state(hmec_ctx *c, int *state);
constexpr int kGmLeadOk = 1, kGmContinue = 3, kGmHold = 4, kGmDone = 5, kGmRestart = 6;
int sv_lincomb(hmec_ctx *c, double *y, double a, const double *u, double b, const double *v);
int sv_lincomb3(hmec_ctx *c, double *y, double a, const double *u, double b, const double *v,
                double d, const double *w);
void sv_gmres_free(hmec_ctx *c);
// ---- serving (hmec_serve.cpp) ----
int batcher_factorize(hmec_batcher *b, hmec_ctx *c, int precision);
// ---- sweep form of the factor (solver_sweep.hip) ----
int sv_sweep_build_batch(hmec_ctx *const *cs, int n, int precision);
bool sv_sweep_ready(hmec_ctx *c, int precision);
void sv_sweep_invalidate(hmec_ctx *c, int precision);
int sv_slots_set(hmec_ctx *c, int n);
int sv_slots_count(hmec_ctx *c);
int sv_slot_nearest(hmec_ctx *c, double dt, double *dist);
int sv_slot_victim(hmec_ctx *c, int keep);
bool sv_slot_idle(hmec_ctx *c, int k, int64_t n);
void sv_slot_target(hmec_ctx *c, int k);
int sv_slot_activate(hmec_ctx *c, int k);
void sv_active_scales(hmec_ctx *c, const double **rs, const double **cs);
void sv_sweep_rhs_target(hmec_ctx *c, int precision, double **zf, const int32_t **fpos);
void sv_sweep_rhs_given(hmec_ctx *c);
int sv_sweep_solve(hmec_ctx *c, double *z);
void sv_sweep_free(hmec_ctx *c);

// timing helpers
void tm_begin(hmec_ctx *c, Timer &t);
void tm_end(hmec_ctx *c, Timer &t, double bytes, int launches = 1);
void tm_resolve(Timer &t);
void tm_free(Timer &t);
inline void kt_begin(hmec_ctx *c, const char *name) {
    if (c->timing) {
        Timer &t = c->kt[name];
        if (t.name.empty()) t.name = name;
        tm_begin(c, t);
    }
}
inline void kt_end(hmec_ctx *c, const char *name, double work, int launches = 1) {
    if (c->timing) tm_end(c, c->kt[name], work, launches);
}

inline double *F(hmec_ctx *c, int f) { return c->d.fields + (int64_t)f * c->ns; }
