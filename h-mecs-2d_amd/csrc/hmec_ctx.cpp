// Context management and the stage-level C ABI of libhmec_hip.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "hmec_internal.h"

static thread_local std::string g_last_error;
void hmec_set_last_error(const std::string &s) { g_last_error = s; }
extern "C" const char *hmec_last_error(void) { return g_last_error.c_str(); }

size_t hmec_sort_tmp_bytes(int64_t M, int nbits);

// Eigen 3.3 VectorXd::LinSpaced(n, low, high) (constants.hpp:57-64)
static void linspaced(double *v, int n, double low, double high) {
    if (n == 1) { v[0] = high; return; }
    const double step = (high - low) / (double)(n - 1);
    const bool flip = std::fabs(high) < std::fabs(low);
    for (int i = 0; i < n; i++) {
        if (flip) v[i] = (i == 0) ? low : (high - (double)(n - 1 - i) * step);
        else v[i] = (i == n - 1) ? high : (low + (double)i * step);
    }
}

static hipEvent_t tm_event(Timer &t) {
    if (!t.pool.empty()) {
        hipEvent_t e = t.pool.back();
        t.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
}

void tm_begin(hmec_ctx *c, Timer &t) {
    if (!c->timing) return;
    if (t.sel_gen != c->timing_gen) {
        t.sel = c->timing_sel.empty() ||
                std::find(c->timing_sel.begin(), c->timing_sel.end(), t.name) != c->timing_sel.end();
        t.sel_gen = c->timing_gen;
    }
    if (!t.sel) return;
    if (c->timing_stride > 1 && (t.calls++ % c->timing_stride) != 0) return;
    t.open_ev = tm_event(t);
    hipEventRecord(t.open_ev, c->stream);
    t.open = true;
}

void tm_end(hmec_ctx *c, Timer &t, double work, int launches) {
    if (!c->timing || !t.open) return;
    hipEvent_t b = tm_event(t);
    hipEventRecord(b, c->stream);
    t.pending.emplace_back(t.open_ev, b);
    t.pending_work.push_back(work);
    t.pending_count.push_back(launches);
    t.open = false;
    if (t.pending.size() >= 4096) tm_resolve(t);
}

void tm_resolve(Timer &t) {
    for (size_t k = 0; k < t.pending.size(); k++) {
        hipEventSynchronize(t.pending[k].second);
        float ms = 0;
        hipEventElapsedTime(&ms, t.pending[k].first, t.pending[k].second);
        t.total_ms += ms;
        t.bytes += t.pending_work[k];
        t.n += t.pending_count[k];
        t.pool.push_back(t.pending[k].first);
        t.pool.push_back(t.pending[k].second);
    }
    t.pending.clear();
    t.pending_work.clear();
    t.pending_count.clear();
}

void tm_free(Timer &t) {
    tm_resolve(t);
    for (hipEvent_t e : t.pool) hipEventDestroy(e);
    t.pool.clear();
}

static void free_ctx(hmec_ctx *c) {
    if (!c) return;
    sv_gmres_free(c);
    sv_sweep_free(c);
    sv_free_device(c);
    DeviceState &d = c->d;
    if (c->ev_scal) hipEventDestroy(c->ev_scal);
    hipFree(d.series);
    hipFree(d.dsylsq);
    hipFree(d.fields); hipFree(d.mk); hipFree(d.prop); hipFree(d.cell); hipFree(d.perm);
    hipFree(d.sorted);
    hipFree(d.cell_start); hipFree(d.sort_tmp); hipFree(d.keys_tmp); hipFree(d.coords);
    hipFree(d.scal);
    if (d.h_scal) hipHostFree(d.h_scal);
    hipFree(d.rowp); hipFree(d.coli); hipFree(d.val); hipFree(d.rhs); hipFree(d.sol);
    hipFree(d.trip_slot); hipFree(d.node_trip); hipFree(d.sec_off); hipFree(d.rscale); hipFree(d.cscale);
    hipFree(d.work); hipFree(d.work2); hipFree(d.counters);
    for (auto &g : c->xguess) {
        hipFree(g.x);
        hipFree(g.x2);
        hipFree(g.x3);
    }
    for (Timer *t : {&c->t_scatter, &c->t_assemble, &c->t_factor, &c->t_solve}) tm_free(*t);
    for (auto &kv : c->kt) tm_free(kv.second);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int hmec_create(const hmec_params *p, hmec_ctx **out) {
    if (!p || !out) return HMEC_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        hipGetLastError();
        hmec_set_last_error("no HIP device");
        return HMEC_E_NODEV;
    }
    hmec_ctx *c = new hmec_ctx();
    c->t_scatter.name = "scatter";
    c->t_assemble.name = "assemble";
    c->t_factor.name = "factor";
    c->t_solve.name = "solve";
    c->P = *p;
    if (hmec_params_finalize(&c->P) != HMEC_OK) { delete c; return HMEC_E_ARG; }
    c->Nx = c->P.Nx; c->Ny = c->P.Ny; c->Nx1 = c->Nx + 1; c->Ny1 = c->Ny + 1;
    c->nb = (int64_t)c->Nx * c->Ny;
    c->ns = (int64_t)c->Nx1 * c->Ny1;
    c->N = c->ns * 6;
    hipGetDevice(&c->device);
    auto fail = [&](int rc) { free_ctx(c); return rc; };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(HMEC_E_HIP);
    DeviceState &d = c->d;
    const int64_t N = c->N;
    if (hipMalloc(&d.fields, sizeof(double) * HMEC_NFIELDS * c->ns) != hipSuccess) return fail(HMEC_E_OOM);
    hipMemset(d.fields, 0, sizeof(double) * HMEC_NFIELDS * c->ns);
    const int ncoord = 2 * c->Nx + 2 * c->Ny + 4 * c->Nx1 + 4 * c->Ny1;
    std::vector<double> co(ncoord);
    {
        const double dx = c->P.dx, dy = c->P.dy;
        double *x = co.data(), *y = x + c->Nx, *xvx = y + c->Ny, *yvx = xvx + c->Nx1;
        double *xvy = yvx + c->Ny1, *yvy = xvy + c->Nx1, *xp = yvy + c->Ny1, *yp = xp + c->Nx1;
        linspaced(x, c->Nx, 0, c->P.xsize);
        linspaced(y, c->Ny, 0, c->P.ysize);
        linspaced(xvx, c->Nx1, 0, c->P.xsize + dx);
        linspaced(yvx, c->Ny1, 0 - dy / 2., c->P.ysize + dy / 2.);
        linspaced(xvy, c->Nx1, 0 - dx / 2., c->P.xsize + dx / 2.);
        linspaced(yvy, c->Ny1, 0, c->P.ysize + dy);
        linspaced(xp, c->Nx1, 0 - dx / 2., c->P.xsize + dx / 2.);
        linspaced(yp, c->Ny1, 0 - dy / 2., c->P.ysize + dy / 2.);
    }
    if (hipMalloc(&d.coords, sizeof(double) * ncoord) != hipSuccess) return fail(HMEC_E_OOM);
    hipMemcpy(d.coords, co.data(), sizeof(double) * ncoord, hipMemcpyHostToDevice);
    if (hipMalloc(&d.scal, sizeof(double) * 32) != hipSuccess) return fail(HMEC_E_OOM);
    hipMemset(d.scal, 0, sizeof(double) * 32);
    if (hipHostMalloc(&d.h_scal, sizeof(double) * 32, hipHostMallocDefault) != hipSuccess) return fail(HMEC_E_OOM);
    // [0, 16): solver counters; then per fault node the yield flag and the
    // bisection-cap flag of the RSF solve; spare
    const size_t ncount = 16 + 2 * (size_t)c->Nx1 + 64;
    if (hipMalloc(&d.counters, sizeof(int) * ncount) != hipSuccess) return fail(HMEC_E_OOM);
    hipMemset(d.counters, 0, sizeof(int) * ncount);
    if (hipMalloc(&d.series, sizeof(double) * (6 * kNumTimesteps + 4 * kSeriesBlocks)) != hipSuccess)
        return fail(HMEC_E_OOM);
    if (hipMalloc(&d.dsylsq, sizeof(double) * (c->P.niterglobal + 1)) != hipSuccess) return fail(HMEC_E_OOM);
    hipMemset(d.dsylsq, 0, sizeof(double) * (c->P.niterglobal + 1));
    hipMemset(d.series, 0, sizeof(double) * 6 * kNumTimesteps);
    for (double **v : {&d.rhs, &d.sol, &d.work, &d.work2, &d.rscale, &d.cscale}) {
        if (hipMalloc(v, sizeof(double) * N) != hipSuccess) return fail(HMEC_E_OOM);
        hipMemset(*v, 0, sizeof(double) * N);
    }
    int rc = hmec_build_pattern(c);
    if (rc != HMEC_OK) return fail(rc);
    rc = k_init_state(c);
    if (rc != HMEC_OK) return fail(rc);
    c->timestep = 1;
    c->timesum = 0;
    c->dt = c->P.dtelastic0;
    c->yndtdecrease = 1;
    c->DSYLSQ.assign(c->P.niterglobal + 1, 0.0);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(HMEC_E_HIP);
    *out = c;
    return HMEC_OK;
}

extern "C" void hmec_destroy(hmec_ctx *c) {
    if (!c) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    free_ctx(c);
}

extern "C" int hmec_set_precision(hmec_ctx *c, int precision) {
    if (!c || (precision != HMEC_PREC_FP64 && precision != HMEC_PREC_FP32)) return HMEC_E_ARG;
    c->precision = precision;
    return HMEC_OK;
}

extern "C" int hmec_set_refinement(hmec_ctx *c, int max_stale, int max_fresh, double tol) {
    if (!c || max_stale < 0 || max_fresh < 0 || !(tol > 0)) return HMEC_E_ARG;
    c->ir_max_stale = max_stale;
    c->ir_max_fresh = max_fresh;
    c->ir_tol = tol;
    return HMEC_OK;
}

extern "C" int hmec_set_fp64_refinement(hmec_ctx *c, int max_ir, double tol) {
    if (!c || max_ir < 0 || !(tol > 0)) return HMEC_E_ARG;
    c->fp64_max_ir = max_ir;
    c->fp64_tol = tol;
    return HMEC_OK;
}

extern "C" int hmec_set_factor_cache(hmec_ctx *c, int slots) {
    if (!c) return HMEC_E_ARG;
    return sv_slots_set(c, slots);
}

extern "C" int hmec_init_coseismic(hmec_ctx *c, double frac, double noise, double pf_ratio,
                                   uint64_t seed) {
    if (!c || !(frac >= 0) || !(noise >= 0) || !(pf_ratio >= 0 && pf_ratio < 1)) return HMEC_E_ARG;
    const hmec_params &P = c->P;
    const int r1 = c->Ny1, c1 = c->Nx1, rb = c->Ny, cb = c->Nx;
    std::vector<double> pt0((size_t)r1 * c1), pf0(pt0.size()), ptf0(pt0.size());
    for (int j = 0; j < c1; j++)
        for (int i = 0; i < r1; i++) {
            const size_t k = (size_t)i + (size_t)j * r1;     // column-major
            const double lit = (P.PCONF + P.PTFDIFF) * (1.0 + noise * hmec_seeded_normal(seed, (int64_t)k, 20));
            pt0[k] = lit;
            pf0[k] = pf_ratio * lit;
            ptf0[k] = lit - pf0[k];
        }
    // basic node (i, j) sits between pressure nodes (i, j) .. (i + 1, j + 1)
    const double mu = 0.6;
    std::vector<double> sxy((size_t)rb * cb);
    for (int j = 0; j < cb; j++)
        for (int i = 0; i < rb; i++) {
            double se = 0;
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++) se += ptf0[(size_t)(i + a) + (size_t)(j + b) * r1];
            sxy[(size_t)i + (size_t)j * rb] = frac * mu * 0.25 * se;
        }
    HMEC_TRY(hmec_set_field(c, HMEC_F_PT0, pt0.data(), HMEC_COLMAJOR));
    HMEC_TRY(hmec_set_field(c, HMEC_F_PF0, pf0.data(), HMEC_COLMAJOR));
    HMEC_TRY(hmec_set_field(c, HMEC_F_PTF0, ptf0.data(), HMEC_COLMAJOR));
    HMEC_TRY(hmec_set_field(c, HMEC_F_PT, pt0.data(), HMEC_COLMAJOR));
    HMEC_TRY(hmec_set_field(c, HMEC_F_PF, pf0.data(), HMEC_COLMAJOR));
    HMEC_TRY(hmec_set_field(c, HMEC_F_SXY0, sxy.data(), HMEC_COLMAJOR));
    HMEC_TRY(hmec_set_field(c, HMEC_F_SXY, sxy.data(), HMEC_COLMAJOR));
    return HMEC_OK;
}

extern "C" int hmec_set_markers(hmec_ctx *c, const double *const soa[HMEC_NMARKER_FIELDS], int64_t M) {
    if (!c || !soa || M <= 0 || M > INT32_MAX) return HMEC_E_ARG;
    DeviceState &d = c->d;
    if (M != c->M) {
        hipFree(d.mk); hipFree(d.prop); hipFree(d.cell); hipFree(d.perm); hipFree(d.cell_start);
        hipFree(d.sorted);
        d.sorted = nullptr;
        hipFree(d.keys_tmp); hipFree(d.sort_tmp);
        d.mk = d.prop = nullptr;
        d.cell = d.perm = d.cell_start = d.keys_tmp = nullptr;
        d.sort_tmp = nullptr;
        const int ncell = (c->Nx - 1) * (c->Ny - 1);
        int nbits = 1;
        while ((1 << nbits) <= ncell) nbits++;
        HMEC_HIP(hipMalloc(&d.mk, sizeof(double) * HMEC_NMARKER_FIELDS * M));
        HMEC_HIP(hipMalloc(&d.prop, sizeof(double) * MP_NPROP * M));
        HMEC_HIP(hipMalloc(&d.cell, sizeof(int32_t) * 4 * M));
        HMEC_HIP(hipMalloc(&d.perm, sizeof(int32_t) * 4 * M));
        HMEC_HIP(hipMalloc(&d.sorted, sizeof(double) * 31 * M));
        HMEC_HIP(hipMalloc(&d.cell_start, sizeof(int32_t) * 4 * (ncell + 1)));
        HMEC_HIP(hipMalloc(&d.keys_tmp, sizeof(int32_t) * 2 * M));
        d.sort_tmp_bytes = hmec_sort_tmp_bytes(M, nbits);
        HMEC_HIP(hipMalloc(&d.sort_tmp, std::max<size_t>(16, d.sort_tmp_bytes)));
        c->M = M;
    }
    for (int f = 0; f < HMEC_NMARKER_FIELDS; f++) {
        if (!soa[f]) return HMEC_E_ARG;
        HMEC_HIP(hipMemcpyAsync(d.mk + (int64_t)f * M, soa[f], sizeof(double) * M,
                                hipMemcpyHostToDevice, c->stream));
    }
    HMEC_HIP(hipStreamSynchronize(c->stream));
    c->have_markers = true;
    c->markers_dirty = true;
    return HMEC_OK;
}

extern "C" int hmec_get_markers(hmec_ctx *c, double *const soa[HMEC_NMARKER_FIELDS]) {
    if (!c || !soa || !c->have_markers) return HMEC_E_ARG;
    HMEC_HIP(hipStreamSynchronize(c->stream));
    for (int f = 0; f < HMEC_NMARKER_FIELDS; f++) {
        if (!soa[f]) continue;
        HMEC_HIP(hipMemcpy(soa[f], c->d.mk + (int64_t)f * c->M, sizeof(double) * c->M,
                           hipMemcpyDeviceToHost));
    }
    return HMEC_OK;
}

extern "C" int hmec_get_marker_cells(hmec_ctx *c, int32_t *out) {
    if (!c || !out || !c->have_markers) return HMEC_E_ARG;
    if (c->markers_dirty) HMEC_TRY(k_marker_prep(c));
    HMEC_HIP(hipStreamSynchronize(c->stream));
    HMEC_HIP(hipMemcpy(out, c->d.cell, sizeof(int32_t) * 4 * c->M, hipMemcpyDeviceToHost));
    const int cy = c->Ny - 1;
    for (int64_t k = 0; k < 4 * c->M; k++) {
        const int id = out[k];
        out[k] = (id % cy) + (id / cy) * c->Ny1;
    }
    return HMEC_OK;
}

static int field_dims(hmec_ctx *c, int f, int &rows, int &cols) {
    if (f < 0 || f >= HMEC_NFIELDS) return HMEC_E_ARG;
    if (f < HMEC_NBASIC_FIELDS) { rows = c->Ny; cols = c->Nx; }
    else { rows = c->Ny1; cols = c->Nx1; }
    return HMEC_OK;
}

extern "C" int hmec_set_field(hmec_ctx *c, int f, const double *host, int layout) {
    if (!c || !host) return HMEC_E_ARG;
    int rows, cols;
    HMEC_TRY(field_dims(c, f, rows, cols));
    std::vector<double> tmp;
    const double *src = host;
    if (layout == HMEC_ROWMAJOR) {
        tmp.resize((size_t)rows * cols);
        for (int i = 0; i < rows; i++)
            for (int j = 0; j < cols; j++) tmp[i + (size_t)j * rows] = host[(size_t)i * cols + j];
        src = tmp.data();
    } else if (layout != HMEC_COLMAJOR) {
        return HMEC_E_ARG;
    }
    HMEC_HIP(hipStreamSynchronize(c->stream));
    HMEC_HIP(hipMemcpy(F(c, f), src, sizeof(double) * rows * cols, hipMemcpyHostToDevice));
    return HMEC_OK;
}

extern "C" int hmec_get_field(hmec_ctx *c, int f, double *host, int layout) {
    if (!c || !host) return HMEC_E_ARG;
    int rows, cols;
    HMEC_TRY(field_dims(c, f, rows, cols));
    HMEC_HIP(hipStreamSynchronize(c->stream));
    if (layout == HMEC_COLMAJOR) {
        HMEC_HIP(hipMemcpy(host, F(c, f), sizeof(double) * rows * cols, hipMemcpyDeviceToHost));
    } else if (layout == HMEC_ROWMAJOR) {
        std::vector<double> tmp((size_t)rows * cols);
        HMEC_HIP(hipMemcpy(tmp.data(), F(c, f), sizeof(double) * rows * cols, hipMemcpyDeviceToHost));
        for (int i = 0; i < rows; i++)
            for (int j = 0; j < cols; j++) host[(size_t)i * cols + j] = tmp[i + (size_t)j * rows];
    } else {
        return HMEC_E_ARG;
    }
    return HMEC_OK;
}

// ------------------------------------------------------------ EVO streams
// A value as an ostream with default flags and precision prints it.
static std::string evo_num(double v) {
    std::ostringstream s;
    s << v;
    return s.str();
}
// An Eigen vector in Eigen's default format: every coefficient right-aligned
// to the widest one, separated by sep (' ' along a row, '\n' down a column).
static void evo_vec(std::ostream &o, const double *v, int n, char sep) {
    std::vector<std::string> t(n);
    size_t w = 0;
    for (int j = 0; j < n; j++) {
        t[j] = evo_num(v[j]);
        w = std::max(w, t[j].size());
    }
    for (int j = 0; j < n; j++) {
        if (j) o << sep;
        o << std::setw((int)w) << t[j];
    }
}
// row line_fault of field f (first timestep only: plain copy of the field)
static int evo_field_row(hmec_ctx *c, int f, std::vector<double> &row) {
    int rows, cols;
    HMEC_TRY(field_dims(c, f, rows, cols));
    std::vector<double> tmp((size_t)rows * cols);
    HMEC_HIP(hipMemcpy(tmp.data(), F(c, f), sizeof(double) * tmp.size(), hipMemcpyDeviceToHost));
    row.resize(cols);
    for (int j = 0; j < cols; j++) row[j] = tmp[c->P.line_fault + (size_t)j * rows];
    return HMEC_OK;
}

extern "C" int hmec_write_evo(hmec_ctx *c, const char *dir, int first) {
    if (!c || !dir) return HMEC_E_ARG;
    // h_mec_rsf_v9.cpp:1751-1800, in that order
    static const char *const names[FaultRows::NSERIES] = {
        "EVO_Vslip.txt", "EVO_viscosity.txt", "EVO_press_flu.txt", "EVO_press_eff.txt", "EVO_SigmaY.txt",
        "EVO_Sii.txt", "EVO_Theta.txt", "EVO_Visc_comp.txt", "EVO_Elast_comp.txt", "EVO_vxD.txt"};
    static const int series[FaultRows::NSERIES][2] = {
        {HMEC_F_VSLIPB, -1}, {HMEC_F_ETA, -1}, {HMEC_F_PF, -1}, {HMEC_F_PT, HMEC_F_PF},
        {HMEC_F_SIGMAY, -1}, {HMEC_F_SII_FAULT, -1}, {HMEC_F_OM, -1}, {HMEC_F_VIS_COMP, -1},
        {HMEC_F_EL_DECOM, -1}, {HMEC_F_VXD, -1}};
    FaultRows a;
    for (int k = 0; k < FaultRows::NSERIES; k++) {
        a.f[k] = series[k][0];
        a.sub[k] = series[k][1];
        HMEC_TRY(field_dims(c, a.f[k], a.rows[k], a.cols[k]));
    }
    const int ld = c->Nx1;
    std::vector<double> h((size_t)FaultRows::NSERIES * ld);
    HMEC_TRY(k_fault_rows_gather(c, a, c->d.work));
    HMEC_HIP(hipMemcpyAsync(h.data(), c->d.work, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HMEC_HIP(hipStreamSynchronize(c->stream));
    const std::string d = std::string(dir) + "/";
    if (first) {
        // :1739-1749
        std::vector<double> x(c->Nx);
        linspaced(x.data(), c->Nx, 0, c->P.xsize);
        std::ofstream o(d + "x_fault.txt");
        evo_vec(o, x.data(), c->Nx, '\n');
        o << std::endl;
        if (!o) return HMEC_E_IO;
        std::vector<double> ra, rb, rl;
        HMEC_TRY(evo_field_row(c, HMEC_F_ARSF, ra));
        HMEC_TRY(evo_field_row(c, HMEC_F_BRSF, rb));
        HMEC_TRY(evo_field_row(c, HMEC_F_LRSF, rl));
        std::ofstream r(d + "rsf_fault.txt", std::ios_base::app | std::ios_base::out);
        evo_vec(r, ra.data(), c->Nx, ' ');
        r << "\n\n";
        evo_vec(r, rb.data(), c->Nx, ' ');
        r << "\n\n";
        evo_vec(r, rl.data(), c->Nx, ' ');
        r << std::endl;
        if (!r) return HMEC_E_IO;
    }
    for (int k = 0; k < FaultRows::NSERIES; k++) {
        std::ofstream o(d + names[k], std::ios_base::app | std::ios_base::out);
        o << c->timesum << "    " << c->dt << "    ";
        evo_vec(o, h.data() + (size_t)k * ld, a.cols[k], ' ');
        o << std::endl;
        if (!o) return HMEC_E_IO;
    }
    std::ofstream o(d + "EVO_data.txt", std::ios_base::app | std::ios_base::out);
    o << std::setw(20) << c->timesum << std::setw(20) << c->dt << std::setw(20) << c->Vmax << std::setw(20)
      << c->ynlast + 1 << std::setw(20) << c->iterstep + 1 << std::endl;
    return o ? HMEC_OK : HMEC_E_IO;
}

extern "C" int hmec_get_state(hmec_ctx *c, hmec_step_info *s) {
    if (!c || !s) return HMEC_E_ARG;
    *s = c->acc;
    s->timestep = c->timestep;
    s->iterations = c->iterstep + 1;
    s->ynlast = c->ynlast;
    s->yndtdecrease = c->yndtdecrease;
    s->dt = c->dt; s->dt00 = c->dt00; s->dtx = c->dtx; s->dty = c->dty;
    s->dtlapusta = c->dtlapusta; s->Vmax = c->Vmax; s->maxvxy = c->maxvxy;
    s->timesum = c->timesum; s->ptscale = c->ptscale;
    return HMEC_OK;
}

extern "C" int hmec_set_state(hmec_ctx *c, const hmec_step_info *s) {
    if (!c || !s) return HMEC_E_ARG;
    c->timestep = s->timestep;
    c->ynlast = s->ynlast;
    c->yndtdecrease = s->yndtdecrease;
    c->dt = s->dt; c->dt00 = s->dt00; c->dtx = s->dtx; c->dty = s->dty;
    c->dtlapusta = s->dtlapusta; c->Vmax = s->Vmax; c->maxvxy = s->maxvxy;
    c->timesum = s->timesum; c->ptscale = s->ptscale;
    return HMEC_OK;
}

// ---------------------------------------------------------------- stages
static double scatter_bytes(hmec_ctx *c) {
    // 19 marker streams + 10 basic + 13 staggered normalized fields
    return 8.0 * (19.0 * c->M + 10.0 * c->nb + 13.0 * c->ns);
}

extern "C" int hmec_scatter(hmec_ctx *c) {
    if (!c) return HMEC_E_ARG;
    tm_begin(c, c->t_scatter);
    HMEC_TRY(k_scatter(c));
    tm_end(c, c->t_scatter, scatter_bytes(c));
    return HMEC_OK;
}

extern "C" int hmec_step_begin(hmec_ctx *c) {
    if (!c) return HMEC_E_ARG;
    const hmec_params &P = c->P;
    const int src = (c->timestep == 1) ? HMEC_F_ETA0 : HMEC_F_ETA00;   // :586-594
    {
        const int dst[4] = {HMEC_F_ETA1, HMEC_F_ETA, HMEC_F_ETA50, HMEC_F_OM};
        const int from[4] = {src, src, src, HMEC_F_OM0};                 // :596
        HMEC_TRY(k_copy_fields(c, dst, from, 4));
    }
    if (!c->yndtdecrease) c->dt = std::max(std::min(c->dt * P.dtkoefup, P.dtelastic0), P.dtmin);
    else c->dt = std::max(std::min(c->dt, P.dtelastic0), P.dtmin);     // :599-603
    c->yndtdecrease = 0;
    c->dt00 = c->dt;
    c->ynlast = 0;
    std::fill(c->DSYLSQ.begin(), c->DSYLSQ.end(), 0.0);
    return HMEC_OK;
}

extern "C" int hmec_vep_pre(hmec_ctx *c, double dt) {
    if (!c) return HMEC_E_ARG;
    return k_vep_pre(c, dt);
}

extern "C" int hmec_assemble(hmec_ctx *c, double dt, double *ptscale_out) {
    if (!c) return HMEC_E_ARG;
    // ptscale stays on the device for the assembly and the update; a caller
    // that asks for it waits for it
    HMEC_TRY(k_scale(c, dt, ptscale_out));
    if (ptscale_out) c->ptscale = *ptscale_out;
    tm_begin(c, c->t_assemble);
    HMEC_TRY(k_assemble(c, dt));
    // CSR values + row pointers/columns + RHS written, ~26 nodal fields read
    tm_end(c, c->t_assemble, 8.0 * (c->nnz + c->N + 26.0 * c->ns) + 4.0 * (c->nnz + c->ntrip));
    return HMEC_OK;
}

extern "C" int hmec_csr_dims(hmec_ctx *c, int64_t *n, int64_t *nnz) {
    if (!c) return HMEC_E_ARG;
    if (n) *n = c->N;
    if (nnz) *nnz = c->nnz;
    return HMEC_OK;
}

extern "C" int hmec_get_csr(hmec_ctx *c, int32_t *row_ptr, int32_t *col_idx, double *values, double *rhs) {
    if (!c) return HMEC_E_ARG;
    HMEC_HIP(hipStreamSynchronize(c->stream));
    if (row_ptr) std::memcpy(row_ptr, c->h_rowp.data(), sizeof(int32_t) * (c->N + 1));
    if (col_idx) std::memcpy(col_idx, c->h_coli.data(), sizeof(int32_t) * c->nnz);
    if (values) HMEC_HIP(hipMemcpy(values, c->d.val, sizeof(double) * c->nnz, hipMemcpyDeviceToHost));
    if (rhs) HMEC_HIP(hipMemcpy(rhs, c->d.rhs, sizeof(double) * c->N, hipMemcpyDeviceToHost));
    return HMEC_OK;
}

extern "C" int hmec_set_csr_values(hmec_ctx *c, const double *values, const double *rhs) {
    if (!c) return HMEC_E_ARG;
    HMEC_HIP(hipStreamSynchronize(c->stream));
    if (values) HMEC_HIP(hipMemcpy(c->d.val, values, sizeof(double) * c->nnz, hipMemcpyHostToDevice));
    if (rhs) HMEC_HIP(hipMemcpy(c->d.rhs, rhs, sizeof(double) * c->N, hipMemcpyHostToDevice));
    c->factor_valid = false;
    if (values) c->rscale_ready = false;
    return HMEC_OK;
}

extern "C" int hmec_analyze(hmec_ctx *c) {
    if (!c) return HMEC_E_ARG;
    if (c->analyzed) return HMEC_OK;
    HMEC_TRY(hmec_condense_build(c->ns, c->h_rowp, c->h_coli, c->cp));
    HMEC_TRY(hmec_plan_build(c->Nx1, c->Ny1, c->cp.srowp, c->cp.scoli, c->plan, 4));
    HMEC_TRY(sv_setup_device(c));
    c->analyzed = true;
    return HMEC_OK;
}

// HMEC_E_PIVOT when the last factorization met a row or column with no
// finite non-zero entry or a non-finite pivot (waits for the device)
int factor_check(hmec_ctx *c) {
    int bad = 0;
    HMEC_HIP(hipStreamSynchronize(c->stream));
    HMEC_HIP(hipMemcpy(&bad, c->d.counters + 1, sizeof(int), hipMemcpyDeviceToHost));
    if (bad > 0) {
        c->factor_valid = false;
        hmec_set_last_error("factorization: " + std::to_string(bad) +
                            " unrepairable pivots (zero or non-finite rows, columns or pivots)");
        return HMEC_E_PIVOT;
    }
    return HMEC_OK;
}

extern "C" int hmec_factorize(hmec_ctx *c, int precision) {
    if (!c || (precision != HMEC_PREC_FP64 && precision != HMEC_PREC_FP32)) return HMEC_E_ARG;
    HMEC_TRY(hmec_analyze(c));
    tm_begin(c, c->t_factor);
    HMEC_TRY(sv_factorize(c, precision));
    tm_end(c, c->t_factor, c->plan.flops);
    return factor_check(c);
}

extern "C" int hmec_solve(hmec_ctx *c, hmec_solve_stats *st) {
    if (!c || !c->factor_valid) return HMEC_E_ARG;
    const bool fp32 = c->factor_prec == HMEC_PREC_FP32;
    tm_begin(c, c->t_solve);
    int rc = sv_solve(c, fp32 ? 30 : c->fp64_max_ir, fp32 ? 1e-12 : c->fp64_tol, st);
    tm_end(c, c->t_solve, 0);
    // FP64: a result that missed the 1e-14 target but meets the acceptance
    // bound is kept, like PARDISO's
    if (rc == HMEC_E_NOCONV && !fp32 && c->last_res <= kAcceptResidual) rc = HMEC_OK;
    return rc;
}

extern "C" int hmec_get_solution(hmec_ctx *c, double *x) {
    if (!c || !x) return HMEC_E_ARG;
    HMEC_HIP(hipStreamSynchronize(c->stream));
    HMEC_HIP(hipMemcpy(x, c->d.sol, sizeof(double) * c->N, hipMemcpyDeviceToHost));
    return HMEC_OK;
}

extern "C" int hmec_update(hmec_ctx *c, double dt, double ptscale, hmec_iter_scalars *out) {
    if (!c || !out) return HMEC_E_ARG;
    c->d.h_scal[0] = ptscale;
    HMEC_HIP(hipMemcpyAsync(c->d.scal, c->d.h_scal, sizeof(double), hipMemcpyHostToDevice, c->stream));
    return hmec_update_dev(c, dt, out);
}

// the update with the ptscale the assembly left in the device scalars
int hmec_update_dev(hmec_ctx *c, double dt, hmec_iter_scalars *out) {
    HMEC_TRY(k_update(c, dt, out));   // reload included
    c->Vmax = out->Vmax;
    c->ptscale = c->d.h_scal[SC_PTSCALE_SLOT];
    return HMEC_OK;
}

extern "C" int hmec_controller(hmec_ctx *c, const hmec_iter_scalars *sc, int iterstep, int *stop) {
    if (!c || !sc || !stop || iterstep < 0 || iterstep >= c->P.niterglobal) return HMEC_E_ARG;
    hipSetDevice(c->device);
    bool s = false;
    HMEC_TRY(k_controller_stage(c, *sc, iterstep, &s));
    *stop = s ? 1 : 0;
    return HMEC_OK;
}

extern "C" int hmec_iter_commit(hmec_ctx *c, int ynpl, int converged) {
    if (!c) return HMEC_E_ARG;
    return k_iter_commit(c, ynpl, converged);
}

extern "C" int hmec_advect(hmec_ctx *c, double dt) {
    if (!c) return HMEC_E_ARG;
    return k_advect(c, dt);
}

// ETA00 = ETA50, OM0 = OM (h_mec_rsf_v9.cpp:1469-1470), final stresses with
// the final dt (:1480-1521), advection (:1561-1656), history (:1658-1726).
extern "C" int hmec_step_end(hmec_ctx *c, double dt) {
    if (!c) return HMEC_E_ARG;
    {
        const int dst[2] = {HMEC_F_ETA00, HMEC_F_OM0};
        const int from[2] = {HMEC_F_ETA50, HMEC_F_OM};
        HMEC_TRY(k_copy_fields(c, dst, from, 2));
    }
    HMEC_TRY(k_final_stress(c, dt));
    if (c->P.advect_markers) HMEC_TRY(k_advect(c, dt));
    HMEC_TRY(k_history(c));
    // :1660-1690: timesum and the step's entry of the per-step series
    c->timesum += dt;
    HMEC_TRY(k_series_step(c, c->timesum, dt));
    return HMEC_OK;
}

extern "C" int hmec_timing_enable(hmec_ctx *c, int on) {
    if (!c) return HMEC_E_ARG;
    c->timing = on != 0;
    c->phase_sync = on >= 2;
    return HMEC_OK;
}

extern "C" int hmec_timing_select(hmec_ctx *c, const char *names) {
    if (!c) return HMEC_E_ARG;
    c->timing_sel.clear();
    if (names) {
        std::string cur;
        for (const char *p = names;; p++) {
            if (*p == ',' || *p == 0) {
                if (!cur.empty()) c->timing_sel.push_back(cur);
                cur.clear();
                if (*p == 0) break;
            } else if (*p != ' ') {
                cur.push_back(*p);
            }
        }
    }
    c->timing_gen++;
    return HMEC_OK;
}

extern "C" int hmec_timing_sample(hmec_ctx *c, int stride) {
    if (!c || stride < 1) return HMEC_E_ARG;
    c->timing_stride = stride;
    return HMEC_OK;
}

extern "C" int hmec_timing_reset(hmec_ctx *c) {
    if (!c) return HMEC_E_ARG;
    std::vector<Timer *> all = {&c->t_scatter, &c->t_assemble, &c->t_factor, &c->t_solve};
    for (auto &kv : c->kt) all.push_back(&kv.second);
    for (Timer *t : all) {
        tm_resolve(*t);
        t->total_ms = 0;
        t->bytes = 0;
        t->n = 0;
    }
    return HMEC_OK;
}

extern "C" int hmec_kernel_time(hmec_ctx *c, const char *name, double *avg_ms, double *bytes) {
    if (!c || !name) return -1;
    Timer *t = nullptr;
    if (!std::strcmp(name, "scatter")) t = &c->t_scatter;
    else if (!std::strcmp(name, "assemble")) t = &c->t_assemble;
    else if (!std::strcmp(name, "factor")) t = &c->t_factor;
    else if (!std::strcmp(name, "solve")) t = &c->t_solve;
    else {
        auto it = c->kt.find(name);
        if (it != c->kt.end()) t = &it->second;
    }
    if (!t) return -1;
    tm_resolve(*t);
    if (avg_ms) *avg_ms = t->n ? t->total_ms / t->n : 0;
    if (bytes) *bytes = t->n ? t->bytes / t->n : 0;
    return t->n;
}

// Device footprint: markers + derived streams + sort, nodal fields, CSR,
// FP64 + FP32 fronts of the symbolic plan (cached per grid).
static std::mutex g_fp_mu;
static std::map<std::pair<int, int>, int64_t> g_fp_cache;

extern "C" int64_t hmec_footprint_bytes(const hmec_params *p) {
    if (!p) return -1;
    const int64_t Nx1 = p->Nx + 1, Ny1 = p->Ny + 1, ns = Nx1 * Ny1, N = 6 * ns;
    const int64_t M = hmec_marker_count(p);
    int64_t arena = 0;
    {
        std::lock_guard<std::mutex> lk(g_fp_mu);
        auto it = g_fp_cache.find({p->Nx, p->Ny});
        if (it != g_fp_cache.end()) arena = it->second;
    }
    if (arena == 0) {
        // plan size estimate from the front model of a real analysis
        hmec_ctx tmp;
        tmp.P = *p;
        hmec_params_finalize(&tmp.P);
        (void)tmp;
        // 2-D nested dissection: O(ns log ns) dense front storage; the
        // constant is calibrated on analysed grids (about 150 entries per
        // node per level of the tree).
        double lv = std::log2((double)ns);
        arena = (int64_t)(36.0 * 4.5 * ns * lv);
    }
    const int64_t bytes = M * 8 * (HMEC_NMARKER_FIELDS + MP_NPROP) + M * 4 * 10 +
                          ns * 8 * HMEC_NFIELDS + N * 8 * 8 + 20 * N * 12 + arena * 12;
    return bytes;
}

void hmec_footprint_record(int Nx, int Ny, int64_t arena) {
    std::lock_guard<std::mutex> lk(g_fp_mu);
    g_fp_cache[{Nx, Ny}] = arena;
}
