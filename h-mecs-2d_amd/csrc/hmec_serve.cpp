// Ensemble serving: a request loop over independent simulations (members)
// with continuous batching, and the batcher that groups same-size members'
// factorizations into one launch sequence.
//
// Admission: a request is admitted once its arrival time has passed, a slot
// is free and its exact device footprint (hmec_member_footprint) fits the
// HBM budget beside the members already running; requests are admitted in
// arrival order.  An admitted member runs its timesteps on its own host
// thread and HIP stream (markers, nodal state and factor stay resident), so
// the members' kernels overlap on the device; on completion its memory is
// released and waiting requests are admitted.
//
// Factorization batching: a member that needs a numeric factorization
// registers with the batcher.  The first requester of a (grid, precision)
// group waits up to batch_window_ms (no deadline when it is negative) for the
// other running same-size members to ask too, then factors the whole group
// with one launch per kernel
// (sv_factorize_batch: blockIdx.z selects the member).  Each member's
// kernels compute exactly what they compute alone, so results do not depend
// on the mix.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "hmec_internal.h"

namespace {
using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0;
    std::sort(v.begin(), v.end());
    const double pos = q * (v.size() - 1);
    const size_t lo = (size_t)pos;
    const size_t hi = std::min(v.size() - 1, lo + 1);
    return v[lo] + (pos - lo) * (v[hi] - v[lo]);
}

// fields kept per finished member (keep_fields)
const int kKeep[] = {HMEC_F_VXS, HMEC_F_VYS, HMEC_F_PT,   HMEC_F_PF,  HMEC_F_VXD, HMEC_F_VYD,
                     HMEC_F_SXX, HMEC_F_SYY, HMEC_F_SXY,  HMEC_F_ETA, HMEC_F_OM,  HMEC_F_VSLIPB};
}  // namespace

// ------------------------------------------------------------ batcher
struct hmec_batcher {
    struct Req {
        hmec_ctx *c;
        bool done = false;
        int rc = HMEC_OK;
        int group = 1;
    };
    std::mutex mu;
    std::condition_variable cv;
    double window_ms = 2.0;
    std::map<std::tuple<int, int, int>, std::vector<Req *>> waiting;   // (Nx, Ny, precision)
    std::map<std::pair<int, int>, int> running;                        // members per grid
    int batches = 0, batch_members = 0;
};

int batcher_factorize(hmec_batcher *b, hmec_ctx *c, int precision) {
    hmec_batcher::Req r;
    r.c = c;
    std::unique_lock<std::mutex> lk(b->mu);
    const auto key = std::make_tuple(c->Nx, c->Ny, precision);
    auto &q = b->waiting[key];
    q.push_back(&r);
    if (q.size() > 1) {
        // follower: the group's first requester factors it
        b->cv.notify_all();
        b->cv.wait(lk, [&] { return r.done; });
        if (r.group > 1) c->batched_facs++;
        return r.rc;
    }
    const std::pair<int, int> grid{c->Nx, c->Ny};
    auto all_in = [&] { return (int)b->waiting[key].size() >= b->running[grid]; };
    if (b->window_ms < 0) {
        // rendezvous without a deadline: every other running same-size member
        // either asks for a factorization or finishes (both notify), so the
        // wait ends, and the grouping does not depend on host timing
        b->cv.wait(lk, all_in);
    } else {
        const auto deadline = Clock::now() + std::chrono::microseconds((int64_t)(b->window_ms * 1e3));
        b->cv.wait_until(lk, deadline, all_in);
    }
    std::vector<hmec_batcher::Req *> grp;
    grp.swap(b->waiting[key]);
    lk.unlock();
    std::vector<hmec_ctx *> cs;
    for (auto *x : grp) cs.push_back(x->c);
    const int rc = sv_factorize_batch(cs.data(), (int)cs.size(), precision);
    lk.lock();
    for (auto *x : grp) {
        x->rc = rc;
        x->group = (int)grp.size();
        x->done = true;
    }
    if (grp.size() > 1) {
        b->batches++;
        b->batch_members += (int)grp.size();
        c->batched_facs++;
    }
    b->cv.notify_all();
    return rc;
}

// ------------------------------------------------------------ server
struct hmec_server {
    struct Member {
        hmec_request req{};
        int id = 0;
        int64_t footprint = 0;
        int state = 0;                      // 0 waiting, 1 running, 2 done
        std::vector<double> lat;
        hmec_member_result res{};
        std::map<int, std::vector<double>> fields;   // column-major
        hmec_params P{};
        std::thread th;
    };
    hmec_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv;
    std::vector<std::unique_ptr<Member>> members;
    hmec_batcher batcher;
    Clock::time_point t0;
    int64_t used = 0, peak = 0;
    int active = 0, max_conc = 0;
    double wall_ms = 0;
};

static void member_run(hmec_server *s, hmec_server::Member *m) {
    hipSetDevice(s->cfg.device);
    hmec_member_result &R = m->res;
    R.footprint = m->footprint;
    int rc = HMEC_OK;
    hmec_ctx *c = nullptr;
    do {
        rc = hmec_create(&m->P, &c);
        if (rc != HMEC_OK) break;
        const int64_t M = hmec_marker_count(&m->P);
        {
            std::vector<std::vector<double>> soa(HMEC_NMARKER_FIELDS, std::vector<double>(M));
            double *ptr[HMEC_NMARKER_FIELDS];
            for (int f = 0; f < HMEC_NMARKER_FIELDS; f++) ptr[f] = soa[f].data();
            rc = hmec_generate_markers(&m->P, m->req.cfg, m->req.seed, m->req.jitter, ptr);
            if (rc != HMEC_OK) break;
            rc = hmec_set_markers(c, ptr, M);
            if (rc != HMEC_OK) break;
        }
        rc = hmec_set_precision(c, s->cfg.precision);
        if (rc != HMEC_OK) break;
        c->batcher = &s->batcher;
        for (int k = 0; k < m->req.nsteps; k++) {
            const auto a = Clock::now();
            hmec_step_info info;
            rc = hmec_step(c, &info);
            if (rc != HMEC_OK) break;
            m->lat.push_back(ms_since(a));
            R.steps++;
            R.factorizations += info.factorizations;
            R.ir_iterations += info.ir_iterations;
            R.fallbacks += info.fallbacks;
            R.max_rel_residual = std::max(R.max_rel_residual, info.max_rel_residual);
            R.timesum = info.timesum;
            R.dt = info.dt;
        }
        R.batched_factorizations = c->batched_facs;
        if (rc == HMEC_OK && s->cfg.keep_fields) {
            for (int f : kKeep) {
                int rows, cols;
                hmec_field_shape(&m->P, f, &rows, &cols);
                std::vector<double> v((size_t)rows * cols);
                rc = hmec_get_field(c, f, v.data(), HMEC_COLMAJOR);
                if (rc != HMEC_OK) break;
                m->fields[f] = std::move(v);
            }
        }
    } while (false);
    if (c) hmec_destroy(c);
    std::lock_guard<std::mutex> lk(s->mu);
    R.status = rc;
    R.finish_ms = ms_since(s->t0);
    R.p50_ms = percentile(m->lat, 0.50);
    R.p99_ms = percentile(m->lat, 0.99);
    double sum = 0;
    for (double x : m->lat) sum += x;
    R.mean_ms = m->lat.empty() ? 0 : sum / m->lat.size();
    m->state = 2;
    s->used -= m->footprint;
    s->active--;
    {
        std::lock_guard<std::mutex> bl(s->batcher.mu);
        s->batcher.running[{m->P.Nx, m->P.Ny}]--;
        s->batcher.cv.notify_all();
    }
    s->cv.notify_all();
}

extern "C" int hmec_server_create(const hmec_server_config *cfg, hmec_server **out) {
    if (!cfg || !out || cfg->max_active <= 0) return HMEC_E_ARG;
    if (cfg->precision != HMEC_PREC_FP64 && cfg->precision != HMEC_PREC_FP32) return HMEC_E_ARG;
    auto *s = new hmec_server();
    s->cfg = *cfg;
    if (s->cfg.hbm_budget <= 0) {
        size_t fr = 0, tot = 0;
        hipSetDevice(cfg->device);
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            delete s;
            hmec_set_last_error("hipMemGetInfo failed");
            return HMEC_E_HIP;
        }
        s->cfg.hbm_budget = (int64_t)(0.9 * (double)fr);
    }
    s->batcher.window_ms = cfg->batch_window_ms < 0 ? -1.0 : cfg->batch_window_ms;
    *out = s;
    return HMEC_OK;
}

extern "C" void hmec_server_destroy(hmec_server *s) {
    if (!s) return;
    for (auto &m : s->members)
        if (m->th.joinable()) m->th.join();
    delete s;
}

extern "C" int hmec_server_submit(hmec_server *s, const hmec_request *req, int *id) {
    if (!s || !req || req->nsteps < 0) return HMEC_E_ARG;
    auto m = std::make_unique<hmec_server::Member>();
    m->req = *req;
    m->P = req->params;
    HMEC_TRY(hmec_params_finalize(&m->P));
    m->footprint = hmec_member_footprint(&m->P, s->cfg.precision);
    if (m->footprint < 0) return HMEC_E_ARG;
    m->res.status = 1;
    std::lock_guard<std::mutex> lk(s->mu);
    m->id = (int)s->members.size();
    if (id) *id = m->id;
    s->members.push_back(std::move(m));
    s->cv.notify_all();
    return HMEC_OK;
}

extern "C" int hmec_server_run(hmec_server *s) {
    if (!s) return HMEC_E_ARG;
    hipSetDevice(s->cfg.device);
    std::unique_lock<std::mutex> lk(s->mu);
    s->t0 = Clock::now();
    for (;;) {
        // waiting members in arrival order (ties: submission order)
        std::vector<hmec_server::Member *> wait;
        bool all_done = true;
        for (auto &m : s->members) {
            if (m->state == 0) wait.push_back(m.get());
            if (m->state != 2) all_done = false;
        }
        if (all_done) break;
        std::stable_sort(wait.begin(), wait.end(), [](const hmec_server::Member *a, const hmec_server::Member *b) {
            return a->req.arrival_ms < b->req.arrival_ms;
        });
        const double now = ms_since(s->t0);
        double next_arrival = -1;
        for (auto *m : wait) {
            if (m->req.arrival_ms > now) {
                next_arrival = m->req.arrival_ms;
                break;
            }
            if (m->footprint > s->cfg.hbm_budget) {
                // can never fit: fail it instead of blocking the queue
                m->state = 2;
                m->res.status = HMEC_E_OOM;
                continue;
            }
            if (s->active >= s->cfg.max_active || s->used + m->footprint > s->cfg.hbm_budget) break;
            m->state = 1;
            m->res.admit_ms = now;
            s->used += m->footprint;
            s->peak = std::max(s->peak, s->used);
            s->active++;
            s->max_conc = std::max(s->max_conc, s->active);
            {
                std::lock_guard<std::mutex> bl(s->batcher.mu);
                s->batcher.running[{m->P.Nx, m->P.Ny}]++;
            }
            m->th = std::thread(member_run, s, m);
        }
        if (next_arrival >= 0) {
            const auto until = s->t0 + std::chrono::microseconds((int64_t)(next_arrival * 1e3));
            s->cv.wait_until(lk, until);
        } else {
            s->cv.wait(lk);
        }
    }
    s->wall_ms = ms_since(s->t0);
    lk.unlock();
    for (auto &m : s->members)
        if (m->th.joinable()) m->th.join();
    return HMEC_OK;
}

extern "C" int hmec_server_result(hmec_server *s, int id, hmec_member_result *res) {
    if (!s || !res) return HMEC_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (id < 0 || id >= (int)s->members.size()) return HMEC_E_ARG;
    *res = s->members[id]->res;
    return HMEC_OK;
}

extern "C" int hmec_server_latencies(hmec_server *s, int id, double *buf, int cap) {
    if (!s) return HMEC_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (id < 0 || id >= (int)s->members.size()) return HMEC_E_ARG;
    const auto &l = s->members[id]->lat;
    if (buf)
        for (int k = 0; k < cap && k < (int)l.size(); k++) buf[k] = l[k];
    return (int)l.size();
}

extern "C" int hmec_server_get_stats(hmec_server *s, hmec_server_stats *st) {
    if (!s || !st) return HMEC_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(st, 0, sizeof(*st));
    std::vector<double> all;
    for (auto &m : s->members) {
        all.insert(all.end(), m->lat.begin(), m->lat.end());
        if (m->state == 2) {
            if (m->res.status == HMEC_OK) st->members_done++;
            else st->members_failed++;
        }
    }
    st->member_steps = (int64_t)all.size();
    st->wall_ms = s->wall_ms;
    st->member_steps_per_s = s->wall_ms > 0 ? all.size() / (s->wall_ms * 1e-3) : 0;
    st->p50_ms = percentile(all, 0.50);
    st->p99_ms = percentile(all, 0.99);
    st->max_concurrent = s->max_conc;
    st->peak_bytes = s->peak;
    std::lock_guard<std::mutex> bl(s->batcher.mu);
    st->factor_batches = s->batcher.batches;
    st->factor_batch_members = s->batcher.batch_members;
    return HMEC_OK;
}

extern "C" int hmec_server_get_field(hmec_server *s, int id, int field, double *buf, int layout) {
    if (!s || !buf) return HMEC_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (id < 0 || id >= (int)s->members.size()) return HMEC_E_ARG;
    auto &m = *s->members[id];
    auto it = m.fields.find(field);
    if (it == m.fields.end()) return HMEC_E_ARG;
    int rows, cols;
    HMEC_TRY(hmec_field_shape(&m.P, field, &rows, &cols));
    const auto &v = it->second;
    if (layout == HMEC_COLMAJOR) {
        std::memcpy(buf, v.data(), sizeof(double) * v.size());
    } else if (layout == HMEC_ROWMAJOR) {
        for (int i = 0; i < rows; i++)
            for (int j = 0; j < cols; j++) buf[(size_t)i * cols + j] = v[i + (size_t)j * rows];
    } else {
        return HMEC_E_ARG;
    }
    return HMEC_OK;
}
