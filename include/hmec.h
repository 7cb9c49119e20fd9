/*
 * libhmec_hip — C ABI of the MI355X (gfx950) implementation of the H-MEC 2D
 * per-timestep hot path: marker->node scatter, visco-elasto-plastic update,
 * rate-and-state friction (RSF) fault solve, fixed-pattern CSR assembly of the
 * 6-unknown-per-cell Stokes+Darcy system, its sparse LU factorization/solve,
 * and node->marker RK4 advection.
 *
 * Reference: lucadalzilio/H-MECs-2D, cpp/h_mec_rsf_v9.cpp and cpp/constants.hpp.
 * The reference has no library interface for this path: it is inline code in
 * main() (h_mec_rsf_v9.cpp:376-1726) plus Eigen's sparse-solver concept
 * (PardisoLU: analyzePattern / factorize / solve, h_mec_rsf_v9.cpp:96,1050-1052).
 * Each entry point below names the reference lines it replaces.
 *
 * Conventions
 *  - Every int-returning call returns HMEC_OK (0) or a negative HMEC_E_* code;
 *    hmec_strerror() names it.  No exceptions cross the ABI.
 *  - A context is not reentrant; distinct contexts may be used from distinct
 *    host threads.  Every context owns its device buffers and one HIP stream.
 *  - Nodal arrays use the reference's Eigen layout: column-major, element
 *    (i, j) at [i + j * rows], rows = Ny (basic grid) or Ny1 = Ny + 1
 *    (pressure / vx / vy grids).  hmec_get_field can also return row-major
 *    (the HDF5 layout of cpp/hdf5.cpp:46-53).
 *  - The global unknown index is the reference's: k = (j * Ny1 + i) * 6 + v,
 *    v = 0:Pt 1:vxs 2:vys 3:Pf 4:vxD 5:vyD (h_mec_rsf_v9.cpp:690-695).
 */
#ifndef HMEC_H
#define HMEC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ errors */
#define HMEC_OK 0
#define HMEC_E_ARG (-1)      /* invalid argument or call order               */
#define HMEC_E_HIP (-2)      /* HIP runtime error                            */
#define HMEC_E_OOM (-3)      /* device allocation failed / over footprint    */
#define HMEC_E_PIVOT (-4)    /* zero pivot that perturbation could not fix   */
#define HMEC_E_NOCONV (-5)   /* iterative refinement did not reach tolerance */
#define HMEC_E_NODEV (-6)    /* no HIP device visible                        */
#define HMEC_E_SOLVER (-7)   /* external solver error                        */
#define HMEC_E_IO (-8)       /* an output file could not be written          */
const char *hmec_strerror(int code);

/* --------------------------------------------------------------- parameters
 * Runtime mirror of cpp/constants.hpp:14-151 (same names, same defaults from
 * hmec_params_default).  Derived quantities (dx, dy, faultwidth, line_fault,
 * upper_block, lower_block) are filled by hmec_params_finalize, which follows
 * constants.hpp:36-76 unless a field was overridden (block bounds).          */
typedef struct hmec_params {
    /* grid (constants.hpp:23-26) */
    double xsize, ysize;
    int Nx, Ny;
    /* markers per basic cell in x and y (constants.hpp:81-82 use 4 and 4) */
    int mx_per_cell, my_per_cell;
    /* RSF profile tapers along x (constants.hpp:29-33) */
    double TS_1, TS_2, TS_3, TS_4;
    /* RSF parameters, index 0 = block, 1 = fault (constants.hpp:93-96) */
    double arsfm[2], brsfm[2], lrsfm[2], omm[2];
    double V0;          /* reference slip velocity (constants.hpp:97)       */
    double alpha;       /* porous-matrix viscosity exponent (:98)           */
    double POR0;        /* standard porosity (:101)                         */
    double cohes, friction, dilatation, tensile, shearmod;   /* :105-109   */
    double BETAFLUID, BETASOLID;                              /* :110-111   */
    double gx, gy, PCONF, PTFDIFF;                            /* :115-118   */
    double etamin, etamax, kkkmin, kkkmax, stpmax, stpmax1;   /* :122-127   */
    double bcupper, bclower, bcvyflower;                      /* :130-132   */
    int niterglobal, ynlastmax, dtstep;                       /* :135-137   */
    double vratiomax, dtkoef, dtkoefup, dtkoefv, errmin, etawt, syieldmin, tyield; /* :138-145 */
    double dtelastic0, dtmin, inertia;                        /* :71-74     */
    /* derived (hmec_params_finalize) */
    double dx, dy, faultwidth;       /* constants.hpp:47-48,66            */
    double upper_block, lower_block; /* constants.hpp:76-77 (fault band)  */
    int line_fault;                  /* constants.hpp:43                  */
    /* switches */
    int advect_markers;   /* 1: run the node->marker RK4 advection loop, which
                             the reference disables (h_mec_rsf_v9.cpp:1563)  */
    int bisect_cap;       /* iteration cap of the RSF stress bisection (the
                             reference has none, h_mec_rsf_v9.cpp:1224)      */
    int block_override;   /* 1: keep caller's upper_block / lower_block       */
} hmec_params;

/* sizeof(hmec_params) / sizeof(hmec_step_info), for binding checks */
int hmec_params_size(void);
int hmec_step_info_size(void);

/* constants.hpp defaults (401x51 grid, 40 km x 10 km) */
void hmec_params_default(hmec_params *p);
/* compute derived fields; returns HMEC_E_ARG on an invalid grid */
int hmec_params_finalize(hmec_params *p);
/* Workload presets.  cfg 1: 101x51, 50x10 km, 4x4 markers/cell, central
 * 20 km velocity-weakening patch.  cfg 2: 401x101 (+-5% lognormal marker
 * properties).  cfg 3: 801x201 (fault viscosity contrast up to 1e6, log-uniform
 * permeability).  cfg 4: 601x151 coseismic.  cfg 5 grid sizes come from the
 * serving mix.  Returns HMEC_E_ARG for an unknown cfg. */
int hmec_params_config(int cfg, hmec_params *p);

/* ------------------------------------------------------------------ markers
 * Marker structure-of-arrays, constants.hpp:153-173 (the 19 marker vectors). */
enum hmec_marker_field {
    HMEC_M_XM = 0, HMEC_M_YM, HMEC_M_PORM, HMEC_M_KKKM, HMEC_M_T_MARKER,
    HMEC_M_COHESCM, HMEC_M_COHESTM, HMEC_M_FRICTCM, HMEC_M_DILATCM,
    HMEC_M_FRICTTM, HMEC_M_ETASM, HMEC_M_ETAM, HMEC_M_RHOM, HMEC_M_RHOFM,
    HMEC_M_ETAFM, HMEC_M_GM, HMEC_M_SXXM, HMEC_M_SYYM, HMEC_M_SXYM,
    HMEC_NMARKER_FIELDS
};
/* number of markers of a grid: (Nx-1)*mx * (Ny-1)*my (constants.hpp:81-85) */
int64_t hmec_marker_count(const hmec_params *p);
/* Seeded synthetic marker generator (host only, no GPU).  Lattice order
 * m = ix * Ny_markers + iy as in h_mec_rsf_v9.cpp:188-191; each marker is
 * jittered uniformly inside its marker sub-cell when jitter != 0 (jitter == 0
 * reproduces the reference's regular lattice).  Materials follow
 * h_mec_rsf_v9.cpp:192-208 with the per-config perturbations of the presets
 * (cfg 5, the serving mix: one log-uniform permeability in [1e-16, 1e-14] m^2
 * per member, drawn from the seed).
 * soa[f] must hold hmec_marker_count(p) doubles for each of the 19 fields. */
int hmec_generate_markers(const hmec_params *p, int cfg, uint64_t seed,
                          int jitter, double *const soa[HMEC_NMARKER_FIELDS]);

/* ------------------------------------------------------------------- fields
 * Nodal fields; B = basic grid (Ny x Nx), S = staggered (Ny1 x Nx1).
 * Names are the reference's globals (h_mec_rsf_v9.cpp:55-109).            */
enum hmec_field {
    /* basic grid */
    HMEC_F_OM0 = 0, HMEC_F_OM, HMEC_F_OM5, HMEC_F_ARSF, HMEC_F_BRSF, HMEC_F_LRSF,
    HMEC_F_RHO, HMEC_F_ETA, HMEC_F_ETA0, HMEC_F_ETA1, HMEC_F_ETA5, HMEC_F_ETA00,
    HMEC_F_ETA50, HMEC_F_IETAPLB, HMEC_F_SXY, HMEC_F_SXY0, HMEC_F_YNY0,
    HMEC_F_YNY, HMEC_F_KKK, HMEC_F_GGG, HMEC_F_COHC, HMEC_F_COHT, HMEC_F_FRIC,
    HMEC_F_FRIT, HMEC_F_DILC, HMEC_F_TTT, HMEC_F_EIIB, HMEC_F_VSLIPB,
    HMEC_F_ESP, HMEC_F_EXY, HMEC_F_DSY, HMEC_F_SIGMAY, HMEC_F_SII_FAULT,
    HMEC_F_SIIB,
    HMEC_NBASIC_FIELDS,
    /* staggered grids */
    HMEC_F_PT = HMEC_NBASIC_FIELDS, HMEC_F_VXS, HMEC_F_VYS, HMEC_F_PF,
    HMEC_F_VXD, HMEC_F_VYD, HMEC_F_ETAB, HMEC_F_ETAB0, HMEC_F_ETAP,
    HMEC_F_ETAP0, HMEC_F_POR, HMEC_F_GGGP, HMEC_F_GGGB, HMEC_F_PTF0,
    HMEC_F_PT0, HMEC_F_PF0, HMEC_F_SXX, HMEC_F_SXX0, HMEC_F_SYY, HMEC_F_SYY0,
    HMEC_F_DILP, HMEC_F_RHOX, HMEC_F_RHOFX, HMEC_F_ETADX, HMEC_F_PORX,
    HMEC_F_VX0, HMEC_F_VXF0, HMEC_F_RHOY, HMEC_F_RHOFY, HMEC_F_ETADY,
    HMEC_F_PORY, HMEC_F_VY0, HMEC_F_VYF0, HMEC_F_EXX, HMEC_F_EYY,
    /* diagnostics of the final stress state, outputs only
     * (h_mec_rsf_v9.cpp:1522-1558): strain-rate and stress invariants,
     * dissipation, vertically averaged pressures, viscous and elastic
     * compaction */
    HMEC_F_EII, HMEC_F_EIIVP, HMEC_F_SII, HMEC_F_DSII, HMEC_F_DIS,
    HMEC_F_PT_AVE, HMEC_F_PF_AVE, HMEC_F_VIS_COMP, HMEC_F_EL_DECOM,
    HMEC_NFIELDS
};
#define HMEC_COLMAJOR 0   /* Eigen MatrixXd layout                        */
#define HMEC_ROWMAJOR 1   /* HDF5 dataset layout (cpp/hdf5.cpp:46-53)      */
/* rows/cols of a field: Ny x Nx (basic) or Ny1 x Nx1 (staggered) */
int hmec_field_shape(const hmec_params *p, int field_id, int *rows, int *cols);

/* ------------------------------------------------------------------ scalars
 * Host-visible per-step state: the reference's scalar globals
 * (h_mec_rsf_v9.cpp:99-105) and the step's outcome.                        */
typedef struct hmec_step_info {
    int timestep;          /* 1-based step just completed                   */
    int iterations;        /* Picard iterations (iterstep + 1)              */
    int ynlast;            /* iterations since last dt decrease             */
    int yndtdecrease;
    double dt, dt00, dtx, dty, dtlapusta, Vmax, maxvxy, timesum;
    double ptscale;        /* last Picard iteration's P scaling             */
    int bisect_capped;     /* RSF nodes whose bisection hit bisect_cap      */
    /* linear-solver statistics accumulated over the step */
    int factorizations;    /* numeric factorizations performed              */
    int solves;            /* linear solves (one per Picard iteration)      */
    int ir_iterations;     /* iterative-refinement sweeps (all solves)      */
    int fallbacks;         /* FP32 factor -> FP64 factor fallbacks          */
    int perturbed_pivots;  /* static-pivot perturbations                    */
    double max_rel_residual; /* max over solves of ||b-Ax||inf/||b||inf     */
    /* wall-clock split of the step (ms) */
    double ms_scatter, ms_assemble, ms_factor, ms_solve, ms_update, ms_total;
} hmec_step_info;

/* ------------------------------------------------------------------ context */
typedef struct hmec_ctx hmec_ctx;

/* Linear-solver precision (the factorization; residuals, refinement and all
 * physics stay FP64). */
#define HMEC_PREC_FP64 0  /* FP64 factor, refactored every Picard iteration
                             like Eigen PardisoLU (h_mec_rsf_v9.cpp:1050-1052) */
#define HMEC_PREC_FP32 1  /* equilibrated FP32 factor + FP64 iterative
                             refinement, factor reused across Picard iterations
                             while refinement converges; FP64 fallback         */

/* Create a context on the current HIP device; copies params.  The fresh
 * nodal state follows h_mec_rsf_v9.cpp:283-338 (replaces main()'s init). */
int hmec_create(const hmec_params *p, hmec_ctx **out);
void hmec_destroy(hmec_ctx *ctx);
/* Device footprint estimate in bytes of a context (markers, nodal state,
 * CSR, factor fill) before it is created; used for admission. */
int64_t hmec_footprint_bytes(const hmec_params *p);
int hmec_set_precision(hmec_ctx *ctx, int precision);
/* Refinement policy of HMEC_PREC_FP32: GMRES iterations allowed with a reused
 * (stale) FP32 factor before refactoring, iterations with a fresh FP32
 * factor before falling back to an FP64 factor, and the target relative
 * inf-norm residual (defaults 12, 40, 1e-12). */
int hmec_set_refinement(hmec_ctx *ctx, int max_stale, int max_fresh, double tol);
/* Refinement of solves with an FP64 factor (HMEC_PREC_FP64, and the FP32
 * policy's FP64 fallback), the counterpart of PARDISO's iparm[7] / iparm[9]
 * that Eigen's PardisoLU sets (h_mec_rsf_v9.cpp:1050-1052): at most max_ir
 * refinement sweeps after the first solve, until both the relative inf-norm
 * residual and the row-equilibrated one are <= tol or the equilibrated one
 * stops halving (defaults 3, 1e-16: rounding level, so the sweeps go on until
 * the residual stagnates; the pressures then come out within a few ulps and
 * the Darcy fluxes, their amplified differences, as accurate as PARDISO's). */
int hmec_set_fp64_refinement(hmec_ctx *ctx, int max_ir, double tol);
/* Number of FP32 factors HMEC_PREC_FP32 keeps, keyed by the dt they were
 * computed at (1..8, default 3; each costs one sweep arena). */
int hmec_set_factor_cache(hmec_ctx *ctx, int slots);
/* Coseismic initial state (the pre-stressed fault workload): lithostatic
 * pressure PT0 = PCONF + PTFDIFF with seeded relative noise of `noise`
 * (normal deviate clipped at 3 sigma) per pressure node, pore pressure
 * PF0 = pf_ratio * PT0, current PT/PF set to the same, and shear stress
 * SXY0 = SXY = frac * mu * (PT0 - PF0) averaged over the four pressure nodes
 * around each basic node (mu = the fault markers' friction 0.6).  The
 * reference starts from uniform PT0/PF0 and zero stress
 * (h_mec_rsf_v9.cpp:316-337); this replaces that start for the stress test. */
int hmec_init_coseismic(hmec_ctx *ctx, double frac, double noise, double pf_ratio, uint64_t seed);

/* Upload the 19 marker arrays (host, M doubles each); marker order kept.
 * Replaces the marker init of h_mec_rsf_v9.cpp:188-209. */
int hmec_set_markers(hmec_ctx *ctx, const double *const soa[HMEC_NMARKER_FIELDS], int64_t M);
int hmec_get_markers(hmec_ctx *ctx, double *const soa[HMEC_NMARKER_FIELDS]);
/* Per-marker anchor cells, 4 grids x M ints: basic, P, vx, vy; each value is
 * i + j * Ny1 as computed by fix()/check_bounds() (h_mec_rsf_v9.cpp:432-433,
 * 468-469, 493-494, 515-516).  Diagnostic for the cell-assignment parity. */
int hmec_get_marker_cells(hmec_ctx *ctx, int32_t *cells4xM);

int hmec_set_field(hmec_ctx *ctx, int field_id, const double *host, int layout);
int hmec_get_field(hmec_ctx *ctx, int field_id, double *host, int layout);

/* Fault-row output streams of the timestep just completed, replacing the
 * per-timestep output block of h_mec_rsf_v9.cpp:1739-1805 (the stdout summary
 * aside).  Appends one line to each of EVO_Vslip.txt, EVO_viscosity.txt,
 * EVO_press_flu.txt, EVO_press_eff.txt (pt - pf), EVO_SigmaY.txt,
 * EVO_Sii.txt, EVO_Theta.txt, EVO_Visc_comp.txt, EVO_Elast_comp.txt and
 * EVO_vxD.txt in directory dir: timesum, dt, then row line_fault of the
 * field, formatted as the reference prints them (default stream precision,
 * Eigen's column alignment), and to EVO_data.txt (timesum, dt, Vmax, global
 * and Picard iteration counts in 20-wide columns).  With first != 0 it also
 * writes x_fault.txt and appends rsf_fault.txt, as the reference does on its
 * first timestep.  The rows are gathered on the device in one launch.
 * HMEC_E_IO when a file cannot be written. */
int hmec_write_evo(hmec_ctx *ctx, const char *dir, int first);
/* ------------------------------------------------------------ HDF5 snapshots
 * The reference's restart files (write h_mec_rsf_v9.cpp:1807-1876, read
 * :176-282, helpers cpp/hdf5.cpp:13-150): one HDF5 file with the groups
 * /Matrix, /Vector and /Value.  /Matrix: the 24 basic-grid matrices (Ny x Nx)
 * then the 30 staggered ones (Ny1 x Nx1) in the order hmec_h5_dataset_name
 * gives, NATIVE_DOUBLE, row-major (element [i][j] = M(i, j), hdf5.cpp:46-53);
 * /Vector: the six per-step series timesumcur, dtcur, maxvxsmod, minvxsmod,
 * maxvysmod, minvysmod (num_timesteps = 100000 entries, constants.hpp:14),
 * then the marker vectors xm, ym, sxxm, syym, sxym (marknum entries);
 * /Value/values: timesum, dt00, dtx, dty, dtlapusta, Vmax, maxvxy, dt,
 * yndtdecrease.  Datasets are staged in heap buffers (the reference's stack
 * arrays overflow at 801x201); the file is the same.  The HDF5 C library
 * (libhdf5.so.103, 1.10) is loaded at first use: without it these calls
 * return HMEC_E_IO, as do unwritable paths, missing datasets and datasets
 * whose shape is not the grid's. */
#define HMEC_H5_NBASIC 24
#define HMEC_H5_NSTAG 30
#define HMEC_H5_NSERIES 6
#define HMEC_H5_NMARKER 5
#define HMEC_H5_NVALUES 9
#define HMEC_H5_SERIES_LEN 100000
/* dataset name k of group 0 (/Matrix basic), 1 (/Matrix staggered),
 * 2 (/Vector series) or 3 (/Vector markers); NULL out of range */
const char *hmec_h5_dataset_name(int group, int k);
/* the hmec_field id of /Matrix dataset k (group 0 or 1), -1 out of range */
int hmec_h5_field_id(int group, int k);
/* A snapshot in host memory; matrices column-major (Eigen's MatrixXd, as the
 * rest of this ABI).  hmec_h5_write reads the arrays, hmec_h5_read fills the
 * non-NULL ones (Nx, Ny and markers give the shapes it expects). */
typedef struct hmec_snapshot {
    int Nx, Ny;                           /* basic grid; staggered (Ny+1)x(Nx+1) */
    int64_t markers;                      /* marknum                              */
    double *basic[HMEC_H5_NBASIC];        /* Ny x Nx, hmec_h5_dataset_name(0, k)  */
    double *stag[HMEC_H5_NSTAG];          /* Ny1 x Nx1, hmec_h5_dataset_name(1, k) */
    double *series[HMEC_H5_NSERIES];      /* HMEC_H5_SERIES_LEN each              */
    double *mk[HMEC_H5_NMARKER];          /* xm, ym, sxxm, syym, sxym             */
    double values[HMEC_H5_NVALUES];
} hmec_snapshot;
int hmec_h5_write(const char *filename, const hmec_snapshot *snap);
int hmec_h5_read(const char *filename, hmec_snapshot *snap);
/* A context's state to a snapshot file / a snapshot file into a context
 * (fields, series, the five marker vectors, the nine scalars).  The context
 * must hold markers before hmec_read_h5: the file keeps only xm, ym, sxxm,
 * syym, sxym, and the other marker properties are regenerated by the caller,
 * as the reference's start-up does (:188-209).  Not in the file, as in the
 * reference: the current stresses SXX, SYY, SXY and the viscosity ETA; a
 * restarted step therefore sees zero current stresses where the continuous
 * run saw the last step's, which changes only the SIIB / EIIB diagnostics of
 * its first Picard iteration (:627-631), not the solution. */
int hmec_write_h5(hmec_ctx *ctx, const char *filename);
int hmec_read_h5(hmec_ctx *ctx, const char *filename);
/* The savestep block (:1807-1876): dir/file.txt holds the timestep just
 * completed, dir/h_mec_<timestep>.h5 the snapshot (the caller decides when,
 * timestep % savestep == 0 in the reference). */
int hmec_save(hmec_ctx *ctx, const char *dir);
/* The restart start-up (:178-282): timestep T from dir/file.txt; T > 0 reads
 * dir/h_mec_<T>.h5 into the context, sets the next timestep to T + 1 and trims
 * the EVO_*.txt streams (hmec_trim_evo); T == 0 leaves the fresh state. */
int hmec_restart(hmec_ctx *ctx, const char *dir);
/* :268-282: when dir/EVO_data.txt has at least timestep lines, cut the EVO
 * streams back to timestep - 1 lines (remove_double_output, :162-174, on the
 * reference's file list: it names the flux stream "EVO_VxD.txt" while the
 * writer names it "EVO_vxD.txt", so that stream is left as it is and an
 * EVO_VxD.txt of empty lines is written, as the reference does). */
int hmec_trim_evo(const char *dir, int timestep);
/* the six per-step series, 6 x HMEC_H5_SERIES_LEN doubles (entry t = step t+1) */
int hmec_get_series(hmec_ctx *ctx, double *series);

/* Scalar state (dt, timesum, ...) in/out; timestep is the next step to run. */
int hmec_get_state(hmec_ctx *ctx, hmec_step_info *info);
int hmec_set_state(hmec_ctx *ctx, const hmec_step_info *info);

/* -------------------------------------------------------------- step kernels
 * Stage-level entry points, each one HIP launch sequence on the context's
 * stream.  Together they make up hmec_step().                              */

/* Marker->node scatter and normalization, h_mec_rsf_v9.cpp:384-583. */
int hmec_scatter(hmec_ctx *ctx);
/* Start of a timestep: ETA/ETA1/ETA50 from ETA0 or ETA00, OM = OM0, dt
 * growth, h_mec_rsf_v9.cpp:586-608. */
int hmec_step_begin(hmec_ctx *ctx);
/* Picard-iteration prologue: copy_bounds(pt, pf), viscosity floor, plastic
 * strain rate, viscoplastic ETAP/ETAB, GGGB, DILP; h_mec_rsf_v9.cpp:612-669. */
int hmec_vep_pre(hmec_ctx *ctx, double dt);
/* ptscale = pfscale (h_mec_rsf_v9.cpp:671-677) and the fixed-pattern CSR
 * values + RHS of the coupled system (h_mec_rsf_v9.cpp:682-1047). */
int hmec_assemble(hmec_ctx *ctx, double dt, double *ptscale_out);
/* CSR of the last assembly (row_ptr N+1, col_idx nnz, values nnz, rhs N);
 * any pointer may be NULL.  hmec_csr_dims gives N and nnz. */
int hmec_csr_dims(hmec_ctx *ctx, int64_t *n, int64_t *nnz);
int hmec_get_csr(hmec_ctx *ctx, int32_t *row_ptr, int32_t *col_idx, double *values, double *rhs);
/* Replace the CSR values / RHS of the fixed pattern (for solver tests). */
int hmec_set_csr_values(hmec_ctx *ctx, const double *values, const double *rhs);

/* Sparse direct solver, mirroring Eigen's split (h_mec_rsf_v9.cpp:1050-1052):
 * analyze = nested-dissection ordering + symbolic multifrontal analysis,
 * done once per grid (the reference redoes it every Picard iteration);
 * factorize = numeric multifrontal LU in the given precision;
 * solve = triangular solves + FP64 iterative refinement of the last CSR.
 * Errors (Eigen's info() that the reference never reads, :1050-1052):
 *  - hmec_factorize / hmec_batch_factorize return HMEC_E_PIVOT when the
 *    (Darcy-condensed, equilibrated) matrix has a row or column with no finite
 *    non-zero entry, or elimination meets a non-finite pivot: no static-pivot
 *    perturbation can repair those.  Small finite pivots are perturbed to
 *    +-tau and counted (hmec_solve_stats.perturbed_pivots), as PARDISO's
 *    iparm[9] = 13 does.
 *  - hmec_solve returns HMEC_E_NOCONV when the final relative residual
 *    ||b - A x||inf / ||b||inf is above 1e-10 (an FP64 solve that misses its
 *    1e-14 refinement target but stays within 1e-10 is accepted), and
 *    HMEC_E_PIVOT for a factor with an unrepairable pivot. */
typedef struct hmec_solve_stats {
    int ir_iterations;
    int fallback;            /* 1: FP32 factor stalled, FP64 factor used    */
    int perturbed_pivots;
    double rel_residual;     /* ||b - A x||inf / ||b||inf after refinement   */
    double factor_flops;     /* flops of one factorization (symbolic count)  */
    int64_t factor_nnz;      /* entries of L+U                               */
    int nfronts, nlevels, max_front;
} hmec_solve_stats;
int hmec_analyze(hmec_ctx *ctx);
int hmec_factorize(hmec_ctx *ctx, int precision);
int hmec_solve(hmec_ctx *ctx, hmec_solve_stats *stats);
/* hmec_solve without any device allocation, for callers that cannot
 * allocate (tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same kernels and results as hmec_solve; the
 * refinement's two work vectors live in a file-scope device array of the
 * library (2 x 1,048,576 doubles per device, held until the process ends)
 * instead of being allocated on first use.  The context must have been
 * created, analyzed and factorized beforehand (those calls allocate).
 * HMEC_E_ARG above 1,048,576 unknowns (hmec_last_error says so).  Calls that
 * overlap on one device share that array and race: serialize them. */
int hmec_solve_noalloc(hmec_ctx *ctx, hmec_solve_stats *stats);
/* solution vector of the last solve (N doubles, reference unknown order) */
int hmec_get_solution(hmec_ctx *ctx, double *x);

/* Reload (h_mec_rsf_v9.cpp:1060-1075), stress update (:1102-1127), RSF fault
 * solve (:1152-1308) and the reductions feeding the dt controller
 * (:1310-1415).  Returns the scalars the controller (hmec_controller) reads. */
typedef struct hmec_iter_scalars {
    int ynpl;
    double ddd, dtlapusta, Vmax;       /* Vmax: VSLIPB max before the RSF   */
    double vxs_max, vxs_min, vys_max, vys_min;
    double maxvxs_band, maxvys_band;   /* |v| max inside the fault band     */
    double dtslip;                     /* min dx*stpmax/V over V > 0        */
    int bisect_capped;
} hmec_iter_scalars;
int hmec_update(hmec_ctx *ctx, double dt, double ptscale, hmec_iter_scalars *out);
/* The Picard / dt controller (h_mec_rsf_v9.cpp:1310-1437), evaluated on the
 * device as hmec_step does once per Picard iteration: the iteration error
 * DSYLSQ, the plastic, displacement, slip and Lapusta timestep candidates,
 * the new dt and the exit test, from this iteration's scalars (hmec_update's
 * output; sc->dtlapusta becomes the context's dtlapusta) and the context's
 * dt, dtx, dty, maxvxy, ynlast and the errors of this step's earlier
 * iterations.  Updates that state (hmec_get_state); ynlast counts on when
 * the loop goes on (:1455).  *stop = 1 when the exit test fires.
 * HMEC_E_ARG unless 0 <= iterstep < niterglobal. */
int hmec_controller(hmec_ctx *ctx, const hmec_iter_scalars *sc, int iterstep, int *stop);
/* Picard-iteration epilogue when not converged: ETA = clamp(ETA5), ETA50,
 * YNY0 = YNY, OM = OM5 (h_mec_rsf_v9.cpp:1440-1448); ETA = ETA0 if ynpl==0
 * (:1315-1317). */
int hmec_iter_commit(hmec_ctx *ctx, int ynpl, int converged);
/* After the Picard loop: ETA00 = ETA50, OM0 = OM, final stresses with the
 * final dt (h_mec_rsf_v9.cpp:1469-1521), RK4 advection when enabled
 * (:1561-1656) and the history update (:1658-1726). */
int hmec_step_end(hmec_ctx *ctx, double dt);
int hmec_advect(hmec_ctx *ctx, double dt);

/* One complete timestep (h_mec_rsf_v9.cpp:376-1726): the C++ host driver
 * with the reference's Picard / dt controller around the kernels above.
 * HMEC_E_NOCONV when a linear solve ends above a relative residual of 1e-10
 * (under HMEC_PREC_FP32 after the FP64 fallback), HMEC_E_PIVOT when even the
 * FP64 factor meets an unrepairable pivot; the step is then incomplete and
 * the context's state is that of the failed Picard iteration. */
int hmec_step(hmec_ctx *ctx, hmec_step_info *info);

/* ------------------------------------------------- symbolic (host only, no GPU)
 * The fixed CSR pattern of the coupled system (what setFromTriplets +
 * makeCompressed give, h_mec_rsf_v9.cpp:1046-1049) and the statistics of the
 * nested-dissection front plan used by hmec_analyze.  Call hmec_pattern with
 * NULL arrays first to get n and nnz.  elim_order (NS = 4 (Nx+1)(Ny+1) ints or
 * NULL) receives the condensed unknowns in elimination order. */
int hmec_pattern(const hmec_params *p, int64_t *n, int64_t *nnz, int32_t *row_ptr, int32_t *col_idx);
int hmec_plan_info(const hmec_params *p, hmec_solve_stats *stats, int64_t *arena_entries,
                   int32_t *elim_order);
/* The factored system is the Darcy-condensed one: vxD and vyD are eliminated
 * exactly (every Darcy row holds its diagonal plus its node's velocity and two
 * fluid pressures, or one other Darcy unknown on the boundary), leaving 4
 * unknowns per node, index node*4 + v (v = 0 Pt, 1 vx, 2 vy, 3 Pf).  Its
 * pattern (NULL arrays: sizes only); hmec_plan_info and hmec_plan_levels
 * describe the plan of this system and elim_order holds n condensed
 * unknowns. */
int hmec_condensed_pattern(const hmec_params *p, int64_t *n, int64_t *nnz, int32_t *row_ptr,
                           int32_t *col_idx);
/* per level of the plan: fronts, largest front order, largest pivot count and
 * dense flops (arrays of *nlevels entries; NULL arrays return nlevels only) */
int hmec_plan_levels(const hmec_params *p, int *nlevels, int *nfronts, int *max_front,
                     int *max_pivots, double *flops);
/* text of the last error on this thread (HIP call, plan check, ...) */
const char *hmec_last_error(void);

/* ------------------------------------------------------------------ serving
 * Factor several same-size contexts' current CSR matrices in one batched
 * launch sequence (all contexts must share Nx, Ny and the current device). */
int hmec_batch_factorize(hmec_ctx *const *ctxs, int n, int precision);
/* Advance n contexts by one timestep each, overlapping their work on the
 * device; same-size members share batched factorizations. */
int hmec_batch_step(hmec_ctx *const *ctxs, int n, hmec_step_info *infos);

/* Device bytes a context of this grid allocates under a precision policy
 * (exact, from the symbolic plan; the admission measure of the server). */
int64_t hmec_member_footprint(const hmec_params *p, int precision);

/* Ensemble server: a request loop that advances independent simulations
 * (members) one timestep at a time with continuous batching.  Requests carry
 * an arrival time; a member is admitted once it has arrived and its
 * footprint fits the HBM budget (and a slot is free), runs on its own HIP
 * stream and host thread, and frees its memory when done so later arrivals
 * are admitted while others run.  Same-size members that need a
 * factorization at the same time share one batched launch sequence
 * (hmec_batch_factorize).  Markers and factors stay device resident for the
 * member's lifetime.  No reference counterpart: the reference runs one
 * simulation per process. */
typedef struct hmec_server hmec_server;
typedef struct {
    int device;                /* HIP device of every member                   */
    int precision;             /* HMEC_PREC_FP64 / HMEC_PREC_FP32              */
    int max_active;            /* member slots (concurrent host threads)       */
    int64_t hbm_budget;        /* admission budget in bytes (<= 0: 90% of the
                                  device memory free at creation)             */
    double batch_window_ms;    /* how long a factorization waits for same-size
                                  partners; < 0: until every running same-size
                                  member asks or finishes (a deterministic
                                  rendezvous, no deadline)                     */
    int keep_fields;           /* keep the members' final fields for
                                  hmec_server_get_field                         */
} hmec_server_config;
typedef struct {
    hmec_params params;        /* grid, friction, permeability, ...            */
    int cfg;                   /* marker material layout (hmec_generate_markers) */
    uint64_t seed;             /* marker generator seed                        */
    int jitter;                /* marker jitter on / off                       */
    int nsteps;                /* timesteps to run                             */
    double arrival_ms;         /* arrival time after hmec_server_run starts    */
} hmec_request;
typedef struct {
    int status;                /* 0 done, HMEC_E_* on failure, 1 not finished   */
    int steps;
    double admit_ms, finish_ms;          /* since hmec_server_run started       */
    double p50_ms, p99_ms, mean_ms;      /* step latency of this member         */
    int64_t footprint;
    int factorizations, batched_factorizations;   /* own / inside a batch > 1  */
    int ir_iterations, fallbacks;
    double max_rel_residual;
    double timesum, dt;
} hmec_member_result;
typedef struct {
    int64_t member_steps;
    double wall_ms;
    double member_steps_per_s;
    double p50_ms, p99_ms;               /* over all member steps               */
    int members_done, members_failed;
    int max_concurrent;
    int64_t peak_bytes;                  /* peak admitted footprint              */
    int factor_batches;                  /* batched launches (> 1 member)       */
    int factor_batch_members;            /* members factored inside them        */
} hmec_server_stats;
int hmec_server_create(const hmec_server_config *cfg, hmec_server **srv);
void hmec_server_destroy(hmec_server *srv);
/* queue a request (before or while hmec_server_run runs); *id >= 0 */
int hmec_server_submit(hmec_server *srv, const hmec_request *req, int *id);
/* serve until every submitted request has finished */
int hmec_server_run(hmec_server *srv);
int hmec_server_result(hmec_server *srv, int id, hmec_member_result *res);
int hmec_server_get_stats(hmec_server *srv, hmec_server_stats *st);
/* per-step latencies (ms) of a member: NULL buf returns the count */
int hmec_server_latencies(hmec_server *srv, int id, double *buf, int cap);
/* a finished member's final field (keep_fields), as hmec_get_field */
int hmec_server_get_field(hmec_server *srv, int id, int field, double *buf, int layout);

/* ------------------------------------------------------------------ timing
 * Average device duration (ms, HIP events on the context's stream) of the
 * dominant kernels since the last reset; name = "scatter", "assemble",
 * "factor", "solve", "schur", "sweep_form", "sweep_fwd", "sweep_bwd".
 * Returns the number of samples.  hmec_timing_enable: on = 1 records the
 * events only; on = 2 also waits for the device at every phase boundary of a
 * step, so that hmec_step_info's wall-clock split is per phase (without it
 * only ms_total is the step's wall-clock time). */
int hmec_kernel_time(hmec_ctx *ctx, const char *name, double *avg_ms, double *bytes);
int hmec_timing_enable(hmec_ctx *ctx, int on);
int hmec_timing_reset(hmec_ctx *ctx);
/* Record only the named timers (comma-separated, e.g. "sweep_fwd,sweep_bwd";
 * NULL or "" records all): every event pair sits in the stream and costs
 * device time, so a throughput window keeps only the ones it reports. */
int hmec_timing_select(hmec_ctx *ctx, const char *names);
/* Record only every stride-th interval of each timer (1: all): a throughput
 * window samples its kernels' durations at a fraction of the events' cost.
 * No counterpart in the reference (measurement only). */
int hmec_timing_sample(hmec_ctx *ctx, int stride);

#ifdef __cplusplus
}
#endif
#endif /* HMEC_H */
