// C++ host driver: one timestep of h_mec_rsf_v9.cpp:376-1726 around the HIP
// stages, with the reference's Picard / timestep controller (:1310-1456)
// evaluated on the device (csrc/controller.h), and the serving entry points
// that advance several contexts concurrently.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "hmec_internal.h"

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Numeric factorization, through the server's batcher for server members
int factor(hmec_ctx *c, int precision) {
    return c->batcher ? batcher_factorize(c->batcher, c, precision) : sv_factorize(c, precision);
}

// FP32 GMRES starts from the context's last solution, or from a kept one
// that is usually closer.  Kept per key (dt within 1 %, Picard iteration):
// the newest solution and the one of an earlier step, so the start can be
// extrapolated linearly in model time (the interseismic state moves
// steadily from step to step):
// - the first solve of a step, or the first at a changed dt: the key (dt,
//   any iteration), i.e. the last solutions computed at that dt, whose
//   newest is the previous step's final one at the reduced dt (the
//   controller alternates dt between an increased value on a step's first
//   Picard iteration and the reduced one after it);
// - from the third Picard iteration on: the key (dt, that iteration), the
//   same iteration of earlier steps (the iterations of a step repeat their
//   pattern from one step to the next).
double dt_dist(double a, double b) { return (a > 0 && b > 0) ? std::fabs(std::log(a / b)) : INFINITY; }

// the entry kept for (dt within 1 %, it), -1 if none
int guess_find(hmec_ctx *c, int it) {
    for (int k = 0; k < hmec_ctx::NGUESS; k++)
        if (c->xguess[k].x && c->xguess[k].it == it && dt_dist(c->xguess[k].dt, c->dt) < 0.01) return k;
    return -1;
}

constexpr double kQuadSpacing = 0.05;   // most relative change of the step between kept solutions

// x0 from entry k: its newest solution, extrapolated to the current step's
// time through the earlier ones when there are (and the current step is no
// farther from the newest than twice the newest's distance from the earlier
// one): linearly through one, or quadratically (Lagrange weights in model
// time) through two when the three are nearly evenly spaced and the entry
// has switched to it.  An entry switches order whenever a solve started
// from it needs more than one GMRES step, so each keeps the order that
// predicts its solutions (early in a run linear mostly, on the smooth
// interseismic trajectory later quadratic for the entries linear misses).
int guess_load(hmec_ctx *c, int k) {
    hmec_ctx::XGuess &g = c->xguess[k];
    g.used = ++c->xguess_clock;
    c->xguess_used = k;
    if (g.has2 && g.t > g.t2 && c->timesum > g.t) {
        const double T = c->timesum;
        const double r = (T - g.t) / (g.t - g.t2);
        // quadratic only over nearly even steps: uneven spacing (dt still
        // growing) makes the weights amplify the kept solutions' errors
        const double h1 = g.t - g.t2, h0 = g.t2 - g.t3;
        if (r <= 2.0 && g.has3 && g.quad && h0 > 0 && std::fabs(h1 / h0 - 1.0) <= kQuadSpacing) {
            const double w1 = (T - g.t2) * (T - g.t3) / ((g.t - g.t2) * (g.t - g.t3));
            const double w2 = (T - g.t) * (T - g.t3) / ((g.t2 - g.t) * (g.t2 - g.t3));
            const double w3 = (T - g.t) * (T - g.t2) / ((g.t3 - g.t) * (g.t3 - g.t2));
            return sv_lincomb3(c, c->d.sol, w1, g.x, w2, g.x2, w3, g.x3);
        }
        if (r <= 2.0) return sv_lincomb(c, c->d.sol, 1.0 + r, g.x, -r, g.x2);
    }
    HMEC_HIP(hipMemcpyAsync(c->d.sol, g.x, sizeof(double) * c->N, hipMemcpyDeviceToDevice, c->stream));
    return HMEC_OK;
}

int guess_start(hmec_ctx *c) {
    c->xguess_used = -1;
    if (c->iterstep >= 2) {
        const int k = guess_find(c, c->iterstep);
        if (k >= 0) return guess_load(c, k);
    }
    if (c->iterstep == 0 || c->last_solve_dt != c->dt) {
        const int k = guess_find(c, -1);
        if (k >= 0) return guess_load(c, k);
    }
    if (c->last_solve_dt == c->dt) return HMEC_OK;
    // no entry at this dt: the kept solution of the nearest dt, when that
    // is nearer than the previous solve's
    int best = -1;
    double bd = dt_dist(c->dt, c->last_solve_dt);
    for (int k = 0; k < hmec_ctx::NGUESS; k++) {
        const double d = (c->xguess[k].x && c->xguess[k].it < 0) ? dt_dist(c->dt, c->xguess[k].dt) : INFINITY;
        if (d < bd) {
            bd = d;
            best = k;
        }
    }
    if (best < 0) return HMEC_OK;
    c->xguess[best].used = ++c->xguess_clock;
    HMEC_HIP(hipMemcpyAsync(c->d.sol, c->xguess[best].x, sizeof(double) * c->N, hipMemcpyDeviceToDevice,
                            c->stream));
    return HMEC_OK;
}

// keep the solution for (dt, it): its entry (the newest solution of an
// earlier step becomes the entry's earlier one; within a step the newest is
// replaced), else an empty or the least recently used entry
int guess_store(hmec_ctx *c, int it) {
    int k = guess_find(c, it);
    if (k < 0) {
        k = 0;
        for (int q = 0; q < hmec_ctx::NGUESS; q++) {
            if (!c->xguess[q].x) {
                k = q;
                break;
            }
            if (c->xguess[q].used < c->xguess[k].used) k = q;
        }
        hmec_ctx::XGuess &g = c->xguess[k];
        if (!g.x) HMEC_HIP(hipMalloc(&g.x, sizeof(double) * c->N));
        g.has2 = false;
        g.has3 = false;
        g.t = c->timesum;
    }
    hmec_ctx::XGuess &g = c->xguess[k];
    if (g.t != c->timesum) {
        // x3 <- x2 <- x <- the new solution (buffers rotated)
        if (g.has2) {
            if (!g.x3) HMEC_HIP(hipMalloc(&g.x3, sizeof(double) * c->N));
            std::swap(g.x2, g.x3);
            g.t3 = g.t2;
            g.has3 = true;
        }
        if (!g.x2) HMEC_HIP(hipMalloc(&g.x2, sizeof(double) * c->N));
        std::swap(g.x, g.x2);
        g.t2 = g.t;
        g.has2 = true;
    }
    g.dt = c->dt;
    g.t = c->timesum;
    g.it = it;
    g.used = ++c->xguess_clock;
    HMEC_HIP(hipMemcpyAsync(g.x, c->d.sol, sizeof(double) * c->N, hipMemcpyDeviceToDevice, c->stream));
    return HMEC_OK;
}

int guess_keep(hmec_ctx *c) {
    HMEC_TRY(guess_store(c, -1));
    if (c->iterstep >= 2) HMEC_TRY(guess_store(c, c->iterstep));
    return HMEC_OK;
}

// Linear solve of the freshly assembled system under the context's precision
// policy.  FP64: numeric refactorization every Picard iteration, like
// PardisoLU (h_mec_rsf_v9.cpp:1050-1052).  FP32: the equilibrated FP32 factor
// is reused while FP64 iterative refinement converges; a stalled refinement
// refactors in FP32, and a second stall falls back to an FP64 factor.
//
// The FP32 solve is split so that hmec_step can enqueue its first GMRES
// step with the decision on the device (sv_gmres_first) and the rest of the
// Picard iteration behind it: fp32_begin chooses or refreshes the factor and
// the initial guess, fp32_finish handles a stalled solve and keeps the
// solution.
struct Fp32Solve {
    double dist = 0;
    int k = -1;
    bool fresh = false;
    int max_it = 0;
};

int fp32_refactor(hmec_ctx *c, hmec_step_info &acc, Fp32Solve &fs, int slot) {
    sv_slot_target(c, slot);
    tm_begin(c, c->t_factor);
    HMEC_TRY(factor(c, HMEC_PREC_FP32));
    tm_end(c, c->t_factor, c->plan.flops);
    acc.factorizations++;
    c->factor_dt = c->dt;
    fs.fresh = true;
    fs.max_it = c->ir_max_fresh;
    return HMEC_OK;
}

// FP32: GMRES preconditioned by a cached FP32 factor.  The cache keeps a few
// factors keyed by the dt they were computed at (the controller alternates
// between an increased dt on a step's first Picard iteration and the reduced
// one after it), so each solve starts from the factor of the nearest dt.
int fp32_begin(hmec_ctx *c, hmec_step_info &acc, Fp32Solve &fs) {
    fs = Fp32Solve();
    fs.k = sv_slot_nearest(c, c->dt, &fs.dist);
    fs.max_it = c->ir_max_stale;
    // a dt more than 10 % from every kept factor, with a free slot (or one
    // whose factor no solve has used lately): factor for it right away
    // rather than spend the GMRES budget on the factor of another dt
    int idle = -1;
    if (fs.k >= 0 && fs.dist > 0.1) {
        const int v = sv_slot_victim(c, fs.k);
        if (v != fs.k && sv_slot_idle(c, v, 16)) idle = v;
    }
    if (fs.k < 0) HMEC_TRY(fp32_refactor(c, acc, fs, sv_slot_victim(c, -1)));
    else if (idle >= 0) HMEC_TRY(fp32_refactor(c, acc, fs, idle));
    else HMEC_TRY(sv_slot_activate(c, fs.k));
    return guess_start(c);
}

// rc of the first GMRES attempt: if GMRES needs more than its budget the
// system has moved: refactor in FP32, into the same slot when dt is close to
// the slot's, else into an empty or the least recently used slot; a fresh
// factor that still fails falls back to an FP64 factor.
int fp32_finish(hmec_ctx *c, hmec_step_info &acc, Fp32Solve &fs, int rc, hmec_solve_stats &st) {
    acc.ir_iterations += st.ir_iterations;
    // a stalled refinement, or an FP32 factor with an unrepairable pivot:
    // a fresh FP32 factor, then an FP64 one
    if ((rc == HMEC_E_NOCONV || rc == HMEC_E_PIVOT) && !fs.fresh) {
        const int slot = fs.dist <= 0.05 ? fs.k : sv_slot_victim(c, fs.k);
        HMEC_TRY(fp32_refactor(c, acc, fs, slot));
        tm_begin(c, c->t_solve);
        rc = sv_gmres(c, c->ir_max_fresh, c->ir_tol, &st);
        tm_end(c, c->t_solve, 0);
        acc.ir_iterations += st.ir_iterations;
    }
    if (rc == HMEC_E_NOCONV || rc == HMEC_E_PIVOT) {
        acc.fallbacks++;
        HMEC_TRY(factor(c, HMEC_PREC_FP64));
        acc.factorizations++;
        rc = sv_solve(c, c->fp64_max_ir, c->fp64_tol, &st);
        if (rc == HMEC_E_NOCONV && st.rel_residual <= kAcceptResidual) rc = HMEC_OK;
    }
    if (rc != HMEC_OK) return rc;
    HMEC_TRY(guess_keep(c));
    c->last_solve_dt = c->dt;
    return HMEC_OK;
}

// after an FP32 solve: a guess that needed more than one GMRES step
// switches its entry's extrapolation order
void guess_outcome(hmec_ctx *c, const hmec_solve_stats &st) {
    if (c->xguess_used >= 0 && st.ir_iterations > 1)
        c->xguess[c->xguess_used].quad = !c->xguess[c->xguess_used].quad;
    c->xguess_used = -1;
}

void solve_stats(hmec_step_info &acc, const hmec_solve_stats &st) {
    acc.solves++;
    acc.perturbed_pivots += st.perturbed_pivots;
    acc.max_rel_residual = std::max(acc.max_rel_residual, st.rel_residual);
}

int linear_solve(hmec_ctx *c, hmec_step_info &acc) {
    hmec_solve_stats st;
    std::memset(&st, 0, sizeof(st));
    if (c->precision == HMEC_PREC_FP64) {
        tm_begin(c, c->t_factor);
        HMEC_TRY(factor(c, HMEC_PREC_FP64));
        tm_end(c, c->t_factor, c->plan.flops);
        acc.factorizations++;
        tm_begin(c, c->t_solve);
        int rc = sv_solve(c, c->fp64_max_ir, c->fp64_tol, &st);
        tm_end(c, c->t_solve, 0);
        if (rc == HMEC_E_NOCONV && st.rel_residual <= kAcceptResidual) rc = HMEC_OK;
        if (rc != HMEC_OK) return rc;
        acc.ir_iterations += st.ir_iterations;
    } else {
        Fp32Solve fs;
        HMEC_TRY(fp32_begin(c, acc, fs));
        tm_begin(c, c->t_solve);
        const int rc = sv_gmres(c, fs.max_it, c->ir_tol, &st);
        tm_end(c, c->t_solve, 0);
        HMEC_TRY(fp32_finish(c, acc, fs, rc, st));
        guess_outcome(c, st);
    }
    solve_stats(acc, st);
    return HMEC_OK;
}

}  // namespace

// One complete timestep, h_mec_rsf_v9.cpp:376-1726 (outputs excluded).
extern "C" int hmec_step(hmec_ctx *c, hmec_step_info *info) {
    if (!c || !c->have_markers) return HMEC_E_ARG;
    hipSetDevice(c->device);
    HMEC_TRY(hmec_analyze(c));
    const hmec_params &P = c->P;
    hmec_step_info &acc = c->acc;
    std::memset(&acc, 0, sizeof(acc));
    const bool sync_timing = c->phase_sync;
    auto sync = [&]() {
        if (sync_timing) hipStreamSynchronize(c->stream);
    };
    const double t0 = now_ms();
    HMEC_TRY(hmec_scatter(c));
    sync();
    const double t1 = now_ms();
    acc.ms_scatter = t1 - t0;
    HMEC_TRY(hmec_step_begin(c));
    HMEC_TRY(k_controller_begin(c));
    // Picard iterations (:610-1457).  The controller runs on the device after
    // the update's reductions, the iteration epilogue (:1440-1448) reads its
    // decision there, and the next iteration's prologue (VEP pre-update,
    // scaling, assembly) is enqueued right behind them with dt from the
    // device and gated on the exit flag: the host reads the decision (one
    // wait per iteration) while the device already assembles the next system.
    auto prologue = [&](bool dev) -> int {
        HMEC_TRY(k_vep_pre(c, c->dt, dev));
        HMEC_TRY(k_scale(c, c->dt, nullptr, dev));
        tm_begin(c, c->t_assemble);
        HMEC_TRY(k_assemble(c, c->dt, dev));
        tm_end(c, c->t_assemble, 8.0 * (c->nnz + c->N + 26.0 * c->ns) + 4.0 * (c->nnz + c->ntrip));
        return HMEC_OK;
    };
    double a = now_ms();
    HMEC_TRY(prologue(false));
    int capped = 0;
    // the iteration's tail: reload, update and controller, the epilogue with
    // its decision, the scalars to the host, and the next iteration's
    // prologue (gated on the exit flag)
    auto tail = [&](bool gated) -> int {
        HMEC_TRY(k_update_launch(c, c->dt, c->iterstep, gated));
        HMEC_TRY(k_controller_commit(c));
        HMEC_TRY(k_scalars_enqueue(c));
        if (c->iterstep + 1 < P.niterglobal) HMEC_TRY(prologue(true));
        return HMEC_OK;
    };
    for (c->iterstep = 0; c->iterstep < P.niterglobal; c->iterstep++) {
        sync();
        const double b = now_ms();
        // FP32: the solve's first GMRES step is decided on the device and the
        // tail is enqueued behind it, gated on that decision; the host waits
        // once, and only a solve that needs more than one step (or a
        // factor, or a fallback) continues on the host before the tail runs
        Fp32Solve fs;
        bool first = false;
        if (c->precision == HMEC_PREC_FP32 && !sync_timing) {
            HMEC_TRY(fp32_begin(c, acc, fs));
            if (fs.max_it >= 1) {
                tm_begin(c, c->t_solve);
                HMEC_TRY(sv_gmres_first(c, c->ir_tol));
                tm_end(c, c->t_solve, 0);
                first = true;
            } else {
                hmec_solve_stats st;
                std::memset(&st, 0, sizeof(st));
                const int rc = sv_gmres(c, fs.max_it, c->ir_tol, &st);
                HMEC_TRY(fp32_finish(c, acc, fs, rc, st));
                guess_outcome(c, st);
                solve_stats(acc, st);
            }
        } else {
            HMEC_TRY(linear_solve(c, acc));
        }
        sync();
        const double cc = now_ms();
        if (first) HMEC_TRY(guess_keep(c));   // kept again if the host finishes the solve
        HMEC_TRY(tail(first));
        hmec_iter_scalars sc;
        if (first) {
            // the first step's outcome arrives before the (gated) tail has run
            int state = 0;
            HMEC_TRY(sv_gmres_first_state(c, &state));
            hmec_solve_stats st;
            std::memset(&st, 0, sizeof(st));
            if (state == kGmLeadOk || state == kGmDone) {
                HMEC_TRY(k_scalars_wait(c, &sc));
                HMEC_TRY(sv_gmres_first_result(c, state, &st));
                acc.ir_iterations += st.ir_iterations;
                c->last_solve_dt = c->dt;
            } else {
                // the update and the prologue hold: finish the solve here,
                // release the flag and run the tail again
                tm_begin(c, c->t_solve);
                const int rc = sv_gmres(c, fs.max_it, c->ir_tol, &st, state);
                tm_end(c, c->t_solve, 0);
                HMEC_TRY(fp32_finish(c, acc, fs, rc, st));
                HMEC_TRY(k_exit_release(c));
                HMEC_TRY(tail(false));
                HMEC_TRY(k_scalars_wait(c, &sc));
            }
            guess_outcome(c, st);
            solve_stats(acc, st);
        } else {
            HMEC_TRY(k_scalars_wait(c, &sc));
        }
        c->Vmax = sc.Vmax;
        c->ptscale = c->d.h_scal[SC_PTSCALE_SLOT];
        c->dtlapusta = sc.dtlapusta;
        capped += sc.bisect_capped;
        const bool stop = k_controller_read(c);   // dt, dtx, dty, maxvxy, ynlast (:1455)
        const double d = now_ms();
        acc.ms_assemble += b - a;
        acc.ms_factor += cc - b;
        acc.ms_update += d - cc;
        a = d;
        if (stop) break;
    }
    // a loop that runs out leaves iterstep == niterglobal, as in the reference
    // (:1434-1456): the EVO_data.txt column and the step's iteration count
    // then read niterglobal + 1
    const double e = now_ms();
    // :1464-1472
    if (c->dt00 > c->dt) c->yndtdecrease = 1;
    c->dt = std::min(std::min(c->dtx, c->dty), std::min(c->dt, c->dtlapusta));
    HMEC_TRY(hmec_step_end(c, c->dt));     // (also timesum += dt, :1661)
    HMEC_HIP(hipStreamSynchronize(c->stream));
    const double f = now_ms();
    acc.ms_update += f - e;
    acc.ms_total = f - t0;
    acc.bisect_capped = capped;
    hmec_get_state(c, &acc);
    acc.timestep = c->timestep;
    if (info) *info = acc;
    c->timestep++;
    return HMEC_OK;
}

// Serving: advance n contexts by one step each.  Every context owns a HIP
// stream; one host thread per context drives its Picard loop so the members'
// kernels overlap on the device.
extern "C" int hmec_batch_step(hmec_ctx *const *cs, int n, hmec_step_info *infos) {
    if (!cs || n <= 0) return HMEC_E_ARG;
    std::vector<int> rc(n, HMEC_OK);
    std::vector<std::thread> th;
    th.reserve(n);
    for (int k = 0; k < n; k++)
        th.emplace_back([&, k]() { rc[k] = hmec_step(cs[k], infos ? &infos[k] : nullptr); });
    for (auto &t : th) t.join();
    for (int k = 0; k < n; k++)
        if (rc[k] != HMEC_OK) return rc[k];
    return HMEC_OK;
}

extern "C" int hmec_batch_factorize(hmec_ctx *const *cs, int n, int precision) {
    if (!cs || n <= 0) return HMEC_E_ARG;
    for (int k = 0; k < n; k++) {
        if (cs[k]->Nx != cs[0]->Nx || cs[k]->Ny != cs[0]->Ny) return HMEC_E_ARG;
        HMEC_TRY(hmec_analyze(cs[k]));
    }
    HMEC_TRY(sv_batch_factorize(cs, n, precision));
    int rc = HMEC_OK;
    for (int k = 0; k < n; k++) {
        const int r = factor_check(cs[k]);
        if (rc == HMEC_OK) rc = r;
    }
    return rc;
}
