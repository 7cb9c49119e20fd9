// Host-side parameters, workload presets and the seeded marker generator.
// Parameter names and defaults mirror cpp/constants.hpp:14-151; the marker
// lattice and materials mirror h_mec_rsf_v9.cpp:188-208.
#include "hmec.h"

#include <cmath>
#include <cstring>

extern "C" const char *hmec_strerror(int code) {
    switch (code) {
    case HMEC_OK: return "ok";
    case HMEC_E_ARG: return "invalid argument or call order";
    case HMEC_E_HIP: return "HIP runtime error";
    case HMEC_E_OOM: return "device allocation failed or footprint exceeded";
    case HMEC_E_PIVOT: return "zero pivot in factorization";
    case HMEC_E_NOCONV: return "iterative refinement did not converge";
    case HMEC_E_NODEV: return "no HIP device";
    case HMEC_E_SOLVER: return "solver error";
    case HMEC_E_IO: return "output file could not be written";
    default: return "unknown error";
    }
}

extern "C" int hmec_params_size(void) { return (int)sizeof(hmec_params); }
extern "C" int hmec_step_info_size(void) { return (int)sizeof(hmec_step_info); }

extern "C" void hmec_params_default(hmec_params *p) {
    std::memset(p, 0, sizeof(*p));
    p->xsize = 40000.; p->ysize = 10000.;            // constants.hpp:23-24
    p->Nx = 401; p->Ny = 51;                         // :25-26
    p->mx_per_cell = 4; p->my_per_cell = 4;          // :81-82
    p->TS_1 = 2e3; p->TS_2 = 4e3; p->TS_3 = 34e3; p->TS_4 = 38e3;   // :29-33
    p->arsfm[0] = .025; p->arsfm[1] = .006;          // :93
    p->brsfm[0] = .001; p->brsfm[1] = .015;          // :94
    p->lrsfm[0] = .020; p->lrsfm[1] = .0085;         // :95
    p->omm[0] = 10; p->omm[1] = -10;                 // :96
    p->V0 = 1.e-9; p->alpha = 29; p->POR0 = .01;     // :97-101
    p->cohes = 0.; p->friction = .6; p->dilatation = 0.; p->tensile = 1.;
    p->shearmod = 3.e10; p->BETAFLUID = 1.e-8; p->BETASOLID = 2.e-11;
    p->gx = 0.; p->gy = 0.; p->PCONF = 1.e7; p->PTFDIFF = 3.e7;
    p->etamin = 1e-3; p->etamax = 1e50; p->kkkmin = 1e-22; p->kkkmax = 1e-12;
    p->stpmax = 2e-4; p->stpmax1 = 6e-5;
    p->bcupper = -1e-9; p->bclower = 1e-9; p->bcvyflower = 0;
    p->niterglobal = 10000; p->ynlastmax = 499; p->dtstep = 5;
    p->vratiomax = 0.001; p->dtkoef = 1.1; p->dtkoefup = 1.2; p->dtkoefv = 1.001;
    p->errmin = 1e2; p->etawt = 0.0; p->syieldmin = 1e-3; p->tyield = 1;
    p->dtelastic0 = 5e8; p->dtmin = 1e-4; p->inertia = 1.;
    p->advect_markers = 0;
    p->bisect_cap = 4000;
    p->block_override = 0;
    hmec_params_finalize(p);
}

extern "C" int hmec_params_finalize(hmec_params *p) {
    if (p->Nx < 4 || p->Ny < 4 || p->mx_per_cell < 1 || p->my_per_cell < 1 ||
        !(p->xsize > 0) || !(p->ysize > 0))
        return HMEC_E_ARG;
    p->dx = p->xsize / (p->Nx - 1);                  // constants.hpp:47
    p->dy = p->ysize / (p->Ny - 1);                  // :48
    p->faultwidth = p->dx;                           // :66
    p->line_fault = (int)((p->Ny - 1) / 2.);         // :43
    if (!p->block_override) {
        p->lower_block = p->ysize / 2. + p->dy;      // :76
        p->upper_block = p->ysize / 2. - p->dy;      // :77
    }
    if (p->bisect_cap <= 0) p->bisect_cap = 4000;
    return HMEC_OK;
}

// Workload presets.  All share a 50 km x 10 km domain with a 200 m fault
// layer at mid-depth (|y - 5 km| <= 100 m); rate-and-state a = 0.011 with
// b = 0.017 inside a central 20 km patch (tapered over 2 km) and b = 0.007
// outside; L = 0.02 m; boundary velocities +-1e-9 m/s.
extern "C" int hmec_params_config(int cfg, hmec_params *p) {
    hmec_params_default(p);
    p->xsize = 50000.; p->ysize = 10000.;
    p->mx_per_cell = 4; p->my_per_cell = 4;
    p->arsfm[0] = 0.011; p->arsfm[1] = 0.011;
    p->brsfm[0] = 0.007; p->brsfm[1] = 0.017;
    p->lrsfm[0] = 0.02;  p->lrsfm[1] = 0.02;
    p->TS_1 = 13e3; p->TS_2 = 15e3; p->TS_3 = 35e3; p->TS_4 = 37e3;
    p->block_override = 1;
    p->upper_block = 4900.; p->lower_block = 5100.;
    switch (cfg) {
    case 1: p->Nx = 101; p->Ny = 51; break;
    case 2: p->Nx = 401; p->Ny = 101; break;
    case 3: p->Nx = 801; p->Ny = 201; break;
    case 4: p->Nx = 601; p->Ny = 151; break;
    case 5: p->Nx = 201; p->Ny = 51; break;
    default: return HMEC_E_ARG;
    }
    return hmec_params_finalize(p);
}

extern "C" int64_t hmec_marker_count(const hmec_params *p) {
    return (int64_t)(p->Nx - 1) * p->mx_per_cell * (int64_t)(p->Ny - 1) * p->my_per_cell;
}

namespace {
// Counter-based generator: splitmix64 of (seed, marker, stream).
inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double uni(uint64_t seed, int64_t m, unsigned stream) {
    uint64_t h = mix64(seed ^ mix64((uint64_t)m * 16 + stream + 0x1234567ull));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);  // [0,1)
}
inline double normal(uint64_t seed, int64_t m, unsigned stream) {
    double u1 = 1.0 - uni(seed, m, stream), u2 = uni(seed, m, stream + 1);
    double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    if (z > 3.) z = 3.;
    if (z < -3.) z = -3.;
    return z;
}
}  // namespace

// seeded standard normal (clipped at 3 sigma) of the marker generator's
// counter-based stream, shared with the other seeded initializers
double hmec_seeded_normal(uint64_t seed, int64_t idx, unsigned stream) {
    return normal(seed, idx, stream);
}

extern "C" int hmec_generate_markers(const hmec_params *p, int cfg, uint64_t seed, int jitter,
                                     double *const soa[HMEC_NMARKER_FIELDS]) {
    if (!p || !soa) return HMEC_E_ARG;
    for (int f = 0; f < HMEC_NMARKER_FIELDS; f++)
        if (!soa[f]) return HMEC_E_ARG;
    const int Nxm = (p->Nx - 1) * p->mx_per_cell;    // constants.hpp:81
    const int Nym = (p->Ny - 1) * p->my_per_cell;    // :82
    const double dxms = p->xsize / (double)Nxm;      // :83
    const double dyms = p->ysize / (double)Nym;      // :84
    const int64_t M = (int64_t)Nxm * Nym;
    const double lnsig = std::log(1.05);
    // serving mix: one permeability per member, log-uniform in [1e-16, 1e-14]
    const double kmember = std::pow(10.0, -16.0 + 2.0 * uni(seed, -1, 9));
    for (int64_t m = 0; m < M; m++) {
        // h_mec_rsf_v9.cpp:190-191 lattice, optional uniform jitter in the sub-cell
        const double jx = jitter ? uni(seed, m, 0) : 0.0;
        const double jy = jitter ? uni(seed, m, 1) : 0.0;
        double xm = 0. + std::floor((double)(m / Nym)) * dxms + jx * dxms;
        double ym = 0. + (double)(m % Nym) * dyms + jy * dyms;
        // constants.hpp:153-168 defaults (fault-layer material)
        double porm = .01, kkkm = 1e-15, tm = 1, rhom = 2800, etasm = 1e21;
        double cohescm = p->cohes, cohestm = p->cohes, frictcm = 0.6, dilatcm = p->dilatation;
        double fricttm = p->tensile;
        const bool host = (ym < p->upper_block || ym > p->lower_block);   // :197
        if (host) {                                  // :198-207
            tm = -1; etasm = 1e23; rhom = 2800; kkkm = 1e-15;
            cohescm = p->cohes * 1e3; cohestm = p->cohes * 1e3;
            frictcm = .8; dilatcm = p->dilatation; fricttm = p->tensile;
        }
        if (cfg == 2) {   // +-5% seeded lognormal perturbations
            etasm *= std::exp(lnsig * normal(seed, m, 2));
            kkkm *= std::exp(lnsig * normal(seed, m, 4));
            porm *= std::exp(lnsig * normal(seed, m, 6));
        } else if (cfg == 3) {
            // fault viscosity contrast up to 1e6 (lognormal, clipped), permeability
            // log-uniform in [1e-16, 1e-14]
            if (!host) {
                double c = 3.0 + 1.0 * normal(seed, m, 2);
                if (c < 0) c = 0;
                if (c > 6) c = 6;
                etasm = 1e23 * std::pow(10.0, -c);
            }
            kkkm = std::pow(10.0, -16.0 + 2.0 * uni(seed, m, 5));
        } else if (cfg == 5) {
            kkkm = kmember;
        }
        double etam = host ? etasm : etasm * std::exp(-p->alpha * porm);   // :194,199
        soa[HMEC_M_XM][m] = xm;
        soa[HMEC_M_YM][m] = ym;
        soa[HMEC_M_PORM][m] = porm;
        soa[HMEC_M_KKKM][m] = kkkm;
        soa[HMEC_M_T_MARKER][m] = tm;
        soa[HMEC_M_COHESCM][m] = cohescm;
        soa[HMEC_M_COHESTM][m] = cohestm;
        soa[HMEC_M_FRICTCM][m] = frictcm;
        soa[HMEC_M_DILATCM][m] = dilatcm;
        soa[HMEC_M_FRICTTM][m] = fricttm;
        soa[HMEC_M_ETASM][m] = etasm;
        soa[HMEC_M_ETAM][m] = etam;
        soa[HMEC_M_RHOM][m] = rhom;
        soa[HMEC_M_RHOFM][m] = 1000;                 // constants.hpp:154
        soa[HMEC_M_ETAFM][m] = 1e-3;                 // :155
        soa[HMEC_M_GM][m] = p->shearmod;             // :153
        soa[HMEC_M_SXXM][m] = 0;
        soa[HMEC_M_SYYM][m] = 0;
        soa[HMEC_M_SXYM][m] = 0;
    }
    return HMEC_OK;
}

extern "C" int hmec_field_shape(const hmec_params *p, int field_id, int *rows, int *cols) {
    if (!p || field_id < 0 || field_id >= HMEC_NFIELDS) return HMEC_E_ARG;
    if (field_id < HMEC_NBASIC_FIELDS) {
        if (rows) *rows = p->Ny;
        if (cols) *cols = p->Nx;
    } else {
        if (rows) *rows = p->Ny + 1;
        if (cols) *cols = p->Nx + 1;
    }
    return HMEC_OK;
}
