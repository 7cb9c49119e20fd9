// HDF5 snapshots and restart in the reference's file layout
// (h_mec_rsf_v9.cpp:1807-1876 write, :176-282 read, cpp/hdf5.cpp:13-150).
//
// A snapshot file has the groups /Matrix, /Vector and /Value.  /Matrix holds
// the 24 basic-grid matrices (Ny x Nx) and then the 30 staggered ones
// (Ny1 x Nx1), NATIVE_DOUBLE, row-major (element [i][j] = M(i, j)); /Vector
// the six per-step series (num_timesteps entries) and the marker vectors xm,
// ym, sxxm, syym, sxym (marknum entries); /Value/values the nine scalars.
// The names and their order are the file format, so they are the
// reference's.  Unlike cpp/hdf5.cpp, which stages every dataset in a stack
// array of its size (marknum doubles overflow an 8 MB stack at 801x201),
// datasets are staged in heap buffers; the bytes on disk are the same.
//
// The HDF5 C library is opened with dlopen at first use (libhdf5.so.103, the
// 1.10 series, or libhdf5.so), so libhmec_hip does not depend on it; without
// it these calls return HMEC_E_IO.
#include <dlfcn.h>
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "hmec_internal.h"

namespace {

// HDF5 1.10 C API types (H5Ipublic.h, H5public.h)
typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;
constexpr hid_t kDefault = 0;         // H5P_DEFAULT, H5S_ALL, H5E_DEFAULT
constexpr unsigned kAccRdonly = 0x0000u, kAccTrunc = 0x0002u;

struct H5Api {
    bool tried = false, ok = false;
    herr_t (*open)();
    herr_t (*eset_auto)(hid_t, void *, void *);
    hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
    hid_t (*fopen)(const char *, unsigned, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*gcreate)(hid_t, const char *, hid_t, hid_t, hid_t);
    hid_t (*gopen)(hid_t, const char *, hid_t);
    herr_t (*gclose)(hid_t);
    hid_t (*screate)(int, const hsize_t *, const hsize_t *);
    herr_t (*sclose)(hid_t);
    int (*sndims)(hid_t);
    int (*sdims)(hid_t, hsize_t *, hsize_t *);
    hid_t (*dcreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
    hid_t (*dopen)(hid_t, const char *, hid_t);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
    hid_t (*dspace)(hid_t);
    herr_t (*dclose)(hid_t);
    hid_t native_double = -1;
};
H5Api g_h5;
std::mutex g_h5_mu;

template <typename F>
bool sym(void *h, const char *name, F &f) {
    f = reinterpret_cast<F>(dlsym(h, name));
    return f != nullptr;
}

// the HDF5 library, opened once per process
H5Api *h5() {
    std::lock_guard<std::mutex> lk(g_h5_mu);
    if (g_h5.tried) return g_h5.ok ? &g_h5 : nullptr;
    g_h5.tried = true;
    const char *names[] = {"libhdf5.so.103", "/opt/conda/lib/libhdf5.so.103", "libhdf5.so",
                           "/opt/conda/lib/libhdf5.so"};
    void *h = nullptr;
    for (const char *n : names)
        if ((h = dlopen(n, RTLD_NOW | RTLD_LOCAL))) break;
    if (!h) return nullptr;
    H5Api &a = g_h5;
    bool ok = sym(h, "H5open", a.open) && sym(h, "H5Eset_auto2", a.eset_auto) &&
              sym(h, "H5Fcreate", a.fcreate) && sym(h, "H5Fopen", a.fopen) &&
              sym(h, "H5Fclose", a.fclose) && sym(h, "H5Gcreate2", a.gcreate) &&
              sym(h, "H5Gopen2", a.gopen) && sym(h, "H5Gclose", a.gclose) &&
              sym(h, "H5Screate_simple", a.screate) && sym(h, "H5Sclose", a.sclose) &&
              sym(h, "H5Sget_simple_extent_ndims", a.sndims) &&
              sym(h, "H5Sget_simple_extent_dims", a.sdims) && sym(h, "H5Dcreate2", a.dcreate) &&
              sym(h, "H5Dopen2", a.dopen) && sym(h, "H5Dwrite", a.dwrite) &&
              sym(h, "H5Dread", a.dread) && sym(h, "H5Dget_space", a.dspace) &&
              sym(h, "H5Dclose", a.dclose);
    if (!ok || a.open() < 0) return nullptr;
    const hid_t *nd = reinterpret_cast<const hid_t *>(dlsym(h, "H5T_NATIVE_DOUBLE_g"));
    if (!nd) return nullptr;
    a.native_double = *nd;
    a.eset_auto(kDefault, nullptr, nullptr);   // errors come back as codes, not printed stacks
    a.ok = true;
    return &a;
}

int io_error(const std::string &what) {
    hmec_set_last_error(what);
    return HMEC_E_IO;
}

// dataset names and order of /Matrix and /Vector (h_mec_rsf_v9.cpp:1828-1860)
const char *kBasicNames[HMEC_H5_NBASIC] = {
    "SIIB", "OM0", "OM", "ARSF", "BRSF", "LRSF", "RHO", "ETA0", "ETA1", "ETA5", "ETA00", "IETAPLB",
    "SXY0", "YNY0", "KKK", "GGG", "COHC", "COHT", "FRIC", "FRIT", "DILC", "TTT", "EIIB", "VSLIPB"};
const int kBasicIds[HMEC_H5_NBASIC] = {
    HMEC_F_SIIB, HMEC_F_OM0, HMEC_F_OM, HMEC_F_ARSF, HMEC_F_BRSF, HMEC_F_LRSF, HMEC_F_RHO,
    HMEC_F_ETA0, HMEC_F_ETA1, HMEC_F_ETA5, HMEC_F_ETA00, HMEC_F_IETAPLB, HMEC_F_SXY0,
    HMEC_F_YNY0, HMEC_F_KKK, HMEC_F_GGG, HMEC_F_COHC, HMEC_F_COHT, HMEC_F_FRIC, HMEC_F_FRIT,
    HMEC_F_DILC, HMEC_F_TTT, HMEC_F_EIIB, HMEC_F_VSLIPB};
const char *kStagNames[HMEC_H5_NSTAG] = {
    "pt", "vxs", "vys", "pf", "vxD", "vyD", "ETAB", "ETAB0", "ETAP", "ETAP0", "POR", "GGGP", "GGGB",
    "PTF0", "PT0", "PF0", "SXX0", "SYY0", "RHOX", "RHOFX", "ETADX", "PORX", "VX0", "VXF0", "RHOY",
    "RHOFY", "ETADY", "PORY", "VY0", "VYF0"};
const int kStagIds[HMEC_H5_NSTAG] = {
    HMEC_F_PT, HMEC_F_VXS, HMEC_F_VYS, HMEC_F_PF, HMEC_F_VXD, HMEC_F_VYD, HMEC_F_ETAB,
    HMEC_F_ETAB0, HMEC_F_ETAP, HMEC_F_ETAP0, HMEC_F_POR, HMEC_F_GGGP, HMEC_F_GGGB, HMEC_F_PTF0,
    HMEC_F_PT0, HMEC_F_PF0, HMEC_F_SXX0, HMEC_F_SYY0, HMEC_F_RHOX, HMEC_F_RHOFX, HMEC_F_ETADX,
    HMEC_F_PORX, HMEC_F_VX0, HMEC_F_VXF0, HMEC_F_RHOY, HMEC_F_RHOFY, HMEC_F_ETADY, HMEC_F_PORY,
    HMEC_F_VY0, HMEC_F_VYF0};
const char *kSeriesNames[HMEC_H5_NSERIES] = {"timesumcur", "dtcur", "maxvxsmod",
                                             "minvxsmod",  "maxvysmod", "minvysmod"};
const char *kMarkerNames[HMEC_H5_NMARKER] = {"xm", "ym", "sxxm", "syym", "sxym"};
const int kMarkerIds[HMEC_H5_NMARKER] = {HMEC_M_XM, HMEC_M_YM, HMEC_M_SXXM, HMEC_M_SYYM, HMEC_M_SXYM};

// one file, its groups and datasets; every handle is closed on scope exit
struct H5File {
    H5Api *a;
    hid_t f = -1;
    std::vector<hid_t> groups;
    ~H5File() {
        for (hid_t g : groups) a->gclose(g);
        if (f >= 0) a->fclose(f);
    }
};

int write_ds(H5Api *a, hid_t g, const char *name, int rank, const hsize_t *dims, const double *buf,
             const std::string &where) {
    const hid_t sp = a->screate(rank, dims, nullptr);
    if (sp < 0) return io_error(where + ": H5Screate_simple failed");
    const hid_t ds = a->dcreate(g, name, a->native_double, sp, kDefault, kDefault, kDefault);
    if (ds < 0) {
        a->sclose(sp);
        return io_error(where + "/" + name + ": H5Dcreate2 failed");
    }
    const herr_t e = a->dwrite(ds, a->native_double, kDefault, kDefault, kDefault, buf);
    a->dclose(ds);
    a->sclose(sp);
    return e < 0 ? io_error(where + "/" + name + ": H5Dwrite failed") : HMEC_OK;
}

int read_ds(H5Api *a, hid_t g, const char *name, int rank, const hsize_t *dims, double *buf,
            const std::string &where) {
    const hid_t ds = a->dopen(g, name, kDefault);
    if (ds < 0) return io_error(where + "/" + name + ": no such dataset");
    const hid_t sp = a->dspace(ds);
    hsize_t got[2] = {0, 0};
    const bool shape_ok = sp >= 0 && a->sndims(sp) == rank && a->sdims(sp, got, nullptr) == rank &&
                          got[0] == dims[0] && (rank < 2 || got[1] == dims[1]);
    if (sp >= 0) a->sclose(sp);
    if (!shape_ok) {
        a->dclose(ds);
        return io_error(where + "/" + name + ": dataset shape differs from the grid's");
    }
    const herr_t e = a->dread(ds, a->native_double, kDefault, kDefault, kDefault, buf);
    a->dclose(ds);
    return e < 0 ? io_error(where + "/" + name + ": H5Dread failed") : HMEC_OK;
}

// column-major (rows x cols) <-> row-major (the file's [i][j] = M(i, j))
void to_rowmajor(const double *cm, int rows, int cols, double *rm) {
    for (int i = 0; i < rows; i++)
        for (int j = 0; j < cols; j++) rm[(size_t)i * cols + j] = cm[i + (size_t)j * rows];
}
void from_rowmajor(const double *rm, int rows, int cols, double *cm) {
    for (int i = 0; i < rows; i++)
        for (int j = 0; j < cols; j++) cm[i + (size_t)j * rows] = rm[(size_t)i * cols + j];
}

std::string join(const char *dir, const std::string &name) {
    std::string d = dir ? dir : ".";
    if (!d.empty() && d.back() != '/') d += '/';
    return d + name;
}

}  // namespace

extern "C" const char *hmec_h5_dataset_name(int group, int k) {
    switch (group) {
    case 0: return (k >= 0 && k < HMEC_H5_NBASIC) ? kBasicNames[k] : nullptr;
    case 1: return (k >= 0 && k < HMEC_H5_NSTAG) ? kStagNames[k] : nullptr;
    case 2: return (k >= 0 && k < HMEC_H5_NSERIES) ? kSeriesNames[k] : nullptr;
    case 3: return (k >= 0 && k < HMEC_H5_NMARKER) ? kMarkerNames[k] : nullptr;
    default: return nullptr;
    }
}

extern "C" int hmec_h5_field_id(int group, int k) {
    if (group == 0 && k >= 0 && k < HMEC_H5_NBASIC) return kBasicIds[k];
    if (group == 1 && k >= 0 && k < HMEC_H5_NSTAG) return kStagIds[k];
    return -1;
}

extern "C" int hmec_h5_write(const char *filename, const hmec_snapshot *s) {
    if (!filename || !s || s->Nx < 2 || s->Ny < 2 || s->markers < 0) return HMEC_E_ARG;
    H5Api *a = h5();
    if (!a) return io_error("HDF5 library (libhdf5.so.103) not found");
    H5File F{a};
    const std::string fn(filename);
    // create_file + add_group (cpp/hdf5.cpp:13-31)
    F.f = a->fcreate(filename, kAccTrunc, kDefault, kDefault);
    if (F.f < 0) return io_error(fn + ": cannot create");
    const char *gnames[3] = {"/Matrix", "/Vector", "/Value"};
    for (const char *g : gnames) {
        const hid_t h = a->gcreate(F.f, g, kDefault, kDefault, kDefault);
        if (h < 0) return io_error(fn + g + ": H5Gcreate2 failed");
        F.groups.push_back(h);
    }
    const int Ny = s->Ny, Nx = s->Nx, Ny1 = Ny + 1, Nx1 = Nx + 1;
    std::vector<double> rm((size_t)Ny1 * Nx1);
    const hsize_t d1[2] = {(hsize_t)Ny, (hsize_t)Nx}, d2[2] = {(hsize_t)Ny1, (hsize_t)Nx1};
    for (int k = 0; k < HMEC_H5_NBASIC; k++) {
        if (!s->basic[k]) return HMEC_E_ARG;
        to_rowmajor(s->basic[k], Ny, Nx, rm.data());
        HMEC_TRY(write_ds(a, F.groups[0], kBasicNames[k], 2, d1, rm.data(), fn + "/Matrix"));
    }
    for (int k = 0; k < HMEC_H5_NSTAG; k++) {
        if (!s->stag[k]) return HMEC_E_ARG;
        to_rowmajor(s->stag[k], Ny1, Nx1, rm.data());
        HMEC_TRY(write_ds(a, F.groups[0], kStagNames[k], 2, d2, rm.data(), fn + "/Matrix"));
    }
    const hsize_t dn[1] = {(hsize_t)HMEC_H5_SERIES_LEN}, dm[1] = {(hsize_t)s->markers},
                  dv[1] = {(hsize_t)HMEC_H5_NVALUES};
    for (int k = 0; k < HMEC_H5_NSERIES; k++) {
        if (!s->series[k]) return HMEC_E_ARG;
        HMEC_TRY(write_ds(a, F.groups[1], kSeriesNames[k], 1, dn, s->series[k], fn + "/Vector"));
    }
    for (int k = 0; k < HMEC_H5_NMARKER; k++) {
        if (!s->mk[k]) return HMEC_E_ARG;
        HMEC_TRY(write_ds(a, F.groups[1], kMarkerNames[k], 1, dm, s->mk[k], fn + "/Vector"));
    }
    return write_ds(a, F.groups[2], "values", 1, dv, s->values, fn + "/Value");
}

extern "C" int hmec_h5_read(const char *filename, hmec_snapshot *s) {
    if (!filename || !s || s->Nx < 2 || s->Ny < 2 || s->markers < 0) return HMEC_E_ARG;
    H5Api *a = h5();
    if (!a) return io_error("HDF5 library (libhdf5.so.103) not found");
    H5File F{a};
    const std::string fn(filename);
    F.f = a->fopen(filename, kAccRdonly, kDefault);
    if (F.f < 0) return io_error(fn + ": cannot open");
    const char *gnames[3] = {"Matrix", "Vector", "Value"};
    for (const char *g : gnames) {
        const hid_t h = a->gopen(F.f, g, kDefault);
        if (h < 0) return io_error(fn + ": no group " + g);
        F.groups.push_back(h);
    }
    const int Ny = s->Ny, Nx = s->Nx, Ny1 = Ny + 1, Nx1 = Nx + 1;
    std::vector<double> rm((size_t)Ny1 * Nx1);
    const hsize_t d1[2] = {(hsize_t)Ny, (hsize_t)Nx}, d2[2] = {(hsize_t)Ny1, (hsize_t)Nx1};
    for (int k = 0; k < HMEC_H5_NBASIC; k++) {
        if (!s->basic[k]) continue;
        HMEC_TRY(read_ds(a, F.groups[0], kBasicNames[k], 2, d1, rm.data(), fn + "/Matrix"));
        from_rowmajor(rm.data(), Ny, Nx, s->basic[k]);
    }
    for (int k = 0; k < HMEC_H5_NSTAG; k++) {
        if (!s->stag[k]) continue;
        HMEC_TRY(read_ds(a, F.groups[0], kStagNames[k], 2, d2, rm.data(), fn + "/Matrix"));
        from_rowmajor(rm.data(), Ny1, Nx1, s->stag[k]);
    }
    const hsize_t dn[1] = {(hsize_t)HMEC_H5_SERIES_LEN}, dm[1] = {(hsize_t)s->markers},
                  dv[1] = {(hsize_t)HMEC_H5_NVALUES};
    for (int k = 0; k < HMEC_H5_NSERIES; k++)
        if (s->series[k])
            HMEC_TRY(read_ds(a, F.groups[1], kSeriesNames[k], 1, dn, s->series[k], fn + "/Vector"));
    for (int k = 0; k < HMEC_H5_NMARKER; k++)
        if (s->mk[k]) HMEC_TRY(read_ds(a, F.groups[1], kMarkerNames[k], 1, dm, s->mk[k], fn + "/Vector"));
    return read_ds(a, F.groups[2], "values", 1, dv, s->values, fn + "/Value");
}

// --------------------------------------------------------- context level
namespace {
// host copies of a context's snapshot state (column-major fields)
struct HostSnap {
    std::vector<double> basic, stag, series, mk;
    hmec_snapshot s{};
    HostSnap(hmec_ctx *c, int64_t M) {
        basic.resize((size_t)HMEC_H5_NBASIC * c->nb);
        stag.resize((size_t)HMEC_H5_NSTAG * c->ns);
        series.resize((size_t)HMEC_H5_NSERIES * HMEC_H5_SERIES_LEN);
        mk.resize((size_t)HMEC_H5_NMARKER * M);
        s.Nx = c->Nx;
        s.Ny = c->Ny;
        s.markers = M;
        for (int k = 0; k < HMEC_H5_NBASIC; k++) s.basic[k] = basic.data() + (size_t)k * c->nb;
        for (int k = 0; k < HMEC_H5_NSTAG; k++) s.stag[k] = stag.data() + (size_t)k * c->ns;
        for (int k = 0; k < HMEC_H5_NSERIES; k++) s.series[k] = series.data() + (size_t)k * HMEC_H5_SERIES_LEN;
        for (int k = 0; k < HMEC_H5_NMARKER; k++) s.mk[k] = mk.data() + (size_t)k * M;
    }
};
}  // namespace

extern "C" int hmec_get_series(hmec_ctx *c, double *series) {
    if (!c || !series) return HMEC_E_ARG;
    HMEC_HIP(hipStreamSynchronize(c->stream));
    HMEC_HIP(hipMemcpy(series, c->d.series, sizeof(double) * HMEC_H5_NSERIES * HMEC_H5_SERIES_LEN,
                       hipMemcpyDeviceToHost));
    return HMEC_OK;
}

extern "C" int hmec_write_h5(hmec_ctx *c, const char *filename) {
    if (!c || !filename) return HMEC_E_ARG;
    HostSnap h(c, c->have_markers ? c->M : 0);
    for (int k = 0; k < HMEC_H5_NBASIC; k++) HMEC_TRY(hmec_get_field(c, kBasicIds[k], h.s.basic[k], HMEC_COLMAJOR));
    for (int k = 0; k < HMEC_H5_NSTAG; k++) HMEC_TRY(hmec_get_field(c, kStagIds[k], h.s.stag[k], HMEC_COLMAJOR));
    HMEC_TRY(hmec_get_series(c, h.series.data()));
    if (c->have_markers && c->M > 0) {
        HMEC_HIP(hipStreamSynchronize(c->stream));
        for (int k = 0; k < HMEC_H5_NMARKER; k++)
            HMEC_HIP(hipMemcpy(h.s.mk[k], c->d.mk + (int64_t)kMarkerIds[k] * c->M, sizeof(double) * c->M,
                               hipMemcpyDeviceToHost));
    }
    // :1862-1865
    const double v[HMEC_H5_NVALUES] = {c->timesum, c->dt00, c->dtx, c->dty, c->dtlapusta,
                                       c->Vmax,    c->maxvxy, c->dt, (double)c->yndtdecrease};
    std::memcpy(h.s.values, v, sizeof(v));
    return hmec_h5_write(filename, &h.s);
}

extern "C" int hmec_read_h5(hmec_ctx *c, const char *filename) {
    if (!c || !filename || !c->have_markers) return HMEC_E_ARG;
    HostSnap h(c, c->M);
    HMEC_TRY(hmec_h5_read(filename, &h.s));
    for (int k = 0; k < HMEC_H5_NBASIC; k++) HMEC_TRY(hmec_set_field(c, kBasicIds[k], h.s.basic[k], HMEC_COLMAJOR));
    for (int k = 0; k < HMEC_H5_NSTAG; k++) HMEC_TRY(hmec_set_field(c, kStagIds[k], h.s.stag[k], HMEC_COLMAJOR));
    HMEC_HIP(hipStreamSynchronize(c->stream));
    HMEC_HIP(hipMemcpy(c->d.series, h.series.data(), sizeof(double) * h.series.size(), hipMemcpyHostToDevice));
    // the five marker vectors the file keeps (:250-255); the other marker
    // properties are the caller's (the reference regenerates them, :188-209)
    {
        std::vector<std::vector<double>> soa(HMEC_NMARKER_FIELDS, std::vector<double>(c->M));
        double *ptr[HMEC_NMARKER_FIELDS];
        for (int f = 0; f < HMEC_NMARKER_FIELDS; f++) ptr[f] = soa[f].data();
        HMEC_TRY(hmec_get_markers(c, ptr));
        for (int k = 0; k < HMEC_H5_NMARKER; k++)
            std::memcpy(ptr[kMarkerIds[k]], h.s.mk[k], sizeof(double) * c->M);
        HMEC_TRY(hmec_set_markers(c, ptr, c->M));
    }
    // :257
    const double *v = h.s.values;
    c->timesum = v[0];
    c->dt00 = v[1];
    c->dtx = v[2];
    c->dty = v[3];
    c->dtlapusta = v[4];
    c->Vmax = v[5];
    c->maxvxy = v[6];
    c->dt = v[7];
    c->yndtdecrease = v[8] != 0;
    return HMEC_OK;
}

extern "C" int hmec_save(hmec_ctx *c, const char *dir) {
    if (!c) return HMEC_E_ARG;
    const int ts = c->timestep - 1;     // the timestep just completed
    if (ts < 1) return HMEC_E_ARG;
    // :1808-1811
    {
        std::ofstream o(join(dir, "file.txt"));
        if (!o) return io_error(join(dir, "file.txt") + ": cannot write");
        o << ts << std::endl;
        if (!o) return io_error(join(dir, "file.txt") + ": write failed");
    }
    return hmec_write_h5(c, join(dir, "h_mec_" + std::to_string(ts) + ".h5").c_str());
}

extern "C" int hmec_trim_evo(const char *dir, int timestep) {
    // :268-282: when EVO_data.txt already holds the restarted step's line,
    // every stream is cut back to its first timestep - 1 lines
    int lines = 0;
    {
        std::ifstream in(join(dir, "EVO_data.txt"));
        std::string line;
        while (std::getline(in, line)) lines++;
    }
    if (lines < timestep) return HMEC_OK;
    // remove_double_output (:162-174) on the reference's list, including its
    // "EVO_VxD.txt": the writer names that stream EVO_vxD.txt (:1798), so the
    // reference leaves EVO_vxD.txt as it is and writes an EVO_VxD.txt of empty
    // lines; kept as it does
    const char *files[11] = {"EVO_Vslip.txt",     "EVO_viscosity.txt", "EVO_press_flu.txt",
                             "EVO_press_eff.txt", "EVO_SigmaY.txt",    "EVO_Sii.txt",
                             "EVO_Theta.txt",     "EVO_Visc_comp.txt", "EVO_Elast_comp.txt",
                             "EVO_VxD.txt",       "EVO_data.txt"};
    for (const char *f : files) {
        const std::string path = join(dir, f);
        std::string content, line;
        {
            std::ifstream in(path);
            for (int i = 0; i < timestep - 1; i++) {
                line.clear();
                std::getline(in, line);
                content += line + '\n';
            }
        }
        std::ofstream out(path);
        if (!out) return io_error(path + ": cannot write");
        out << content;
    }
    return HMEC_OK;
}

extern "C" int hmec_restart(hmec_ctx *c, const char *dir) {
    if (!c || !c->have_markers) return HMEC_E_ARG;
    // :178-183
    int ts = 0;
    {
        std::ifstream in(join(dir, "file.txt"));
        std::string s;
        if (!in || !std::getline(in, s)) return io_error(join(dir, "file.txt") + ": cannot read");
        try {
            ts = std::stoi(s);
        } catch (...) {
            return io_error(join(dir, "file.txt") + ": no timestep");
        }
    }
    if (ts <= 0) return HMEC_OK;        // a fresh start (:283)
    HMEC_TRY(hmec_read_h5(c, join(dir, "h_mec_" + std::to_string(ts) + ".h5").c_str()));
    c->timestep = ts + 1;               // :267
    return hmec_trim_evo(dir, c->timestep);
}
