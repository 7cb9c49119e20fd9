/*
 * ORACLE — test infrastructure only.  A single-threaded, IEEE-strict CPU
 * restatement of the H-MEC 2D per-timestep path of lucadalzilio/H-MECs-2D
 * (cpp/h_mec_rsf_v9.cpp:188-1726, cpp/constants.hpp:14-175).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, as the
 * checker; the product (libhmec_hip) never calls it.
 *
 * The reference itself cannot be compiled here: it needs Eigen 3.3 (absent
 * from the image) and Intel icpc; so this restatement replaces Eigen with
 * plain column-major arrays (element (i, j) at [i + j * rows], Eigen's MatXd
 * layout) and Eigen's PardisoLU with a direct call of MKL PARDISO (mtype 11,
 * Eigen 3.3's pardisoInit iparm), loaded with dlopen.  Without MKL a banded
 * LU with partial pivoting is used instead (small grids only).
 *
 * Choices for the reference's quirks (each is <= a few ulp, far inside the
 * 1e-10 parity tolerance, or removes undefined behaviour):
 *  Q1  "for (auto i : {...}) i.setZero()" acts on copies (h_mec_rsf_v9.cpp:
 *      292-295, 332-334, 377-381, 1475-1477): state starts at zero and the
 *      scatter accumulators are reset every step (the reference keeps adding
 *      one marker pass per step, which changes normalized values by ulps).
 *  Q2  RK4 stage semantics kept as written (stage 3 samples stage 2's point).
 *  Q3  scatter is single-threaded in marker order (the reference's OpenMP
 *      scatter races).
 *  Q5  fix() = IEEE division + truncation; coordinates follow Eigen 3.3
 *      LinSpaced: v(i) = i == n-1 ? high : low + i * step.
 *  Q6  exit-iteration state handling exactly as written.
 *  Q7  duplicate triplets summed (Eigen setFromTriplets).
 *  Q8  the RSF bisection gets an iteration cap (params.bisect_cap); nodes
 *      that hit it are counted.
 *  Eigen's 2x2 block sums are taken in column-major order
 *  ((a00 + a10) + a01) + a11; pow(x, 2) is written x * x (pow with an
 *  exponent of 2 is the correctly rounded square, the same value).
 */
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../include/hmec.h"

#define SQ(a) ((a) * (a))
/* constants.hpp:14 num_timesteps: length of the per-step series */
#define NUM_TIMESTEPS 100000
static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

typedef struct {
    hmec_params P;
    int Nx, Ny, Nx1, Ny1, N;
    int64_t M;
    double *x, *y, *xvx, *yvx, *xvy, *yvy, *xp, *yp;
    double *F[HMEC_NFIELDS];
    double *mk[HMEC_NMARKER_FIELDS];
    /* scatter accumulators (h_mec_rsf_v9.cpp:341-359) */
    double *acc[31];
    /* scalar state (h_mec_rsf_v9.cpp:99-107) */
    int timestep, ynlast, iterstep, yndtdecrease;
    double timesum, dt, dt00, dtx, dty, dtlapusta, Vmax, maxvxy, ptscale;
    double *DSYLSQ;
    /* linear system */
    int64_t ntrip, captrip;
    int *ti, *tj;
    double *tv;
    int nnz;
    int *rowp, *coli;
    double *val, *R, *S;
    int *trip2slot;
    /* solver */
    int use_pardiso;
    void *pt[64];
    int iparm[64];
    int pardiso_ready;
    double t_solve, t_assemble, t_scatter, t_update;
    int bisect_capped;
    double *series;     /* 6 x NUM_TIMESTEPS: timesumcur, dtcur, max/minvxsmod, max/minvysmod */
    int64_t lu_nnz;
    double lu_mflops;
} orc;

/* ----------------------------------------------------------------- helpers */
/* h_mec_rsf_v9.cpp:119-121 */
int orc_fix(double t) { return t < 0. ? (int)ceil(t) : (int)floor(t); }
/* h_mec_rsf_v9.cpp:144-151 */
int orc_check_bounds(int k, int bound) {
    if (k < 0) k = 0;
    else if (k > bound - 2) k = bound - 2;
    return k;
}
/* h_mec_rsf_v9.cpp:124-131 */
static double enforce_bounds(double v, double mn, double mx) {
    if (v < mn) v = mn;
    else if (v > mx) v = mx;
    return v;
}
/* h_mec_rsf_v9.cpp:139-141 */
double orc_divplus(double a, double b) { return a / (a + b); }
/* Eigen 3.3 VectorXd::LinSpaced(n, low, high) (constants.hpp:57-64) */
void orc_linspaced(double *v, int n, double low, double high) {
    if (n == 1) { v[0] = high; return; }
    double step = (high - low) / (double)(n - 1);
    int flip = fabs(high) < fabs(low);
    for (int i = 0; i < n; i++) {
        if (flip) v[i] = (i == 0) ? low : (high - (double)(n - 1 - i) * step);
        else v[i] = (i == n - 1) ? high : (low + (double)i * step);
    }
}

#define B(o, f, i, j) ((o)->F[f][(i) + (int64_t)(j) * (o)->Ny])
#define S_(o, f, i, j) ((o)->F[f][(i) + (int64_t)(j) * (o)->Ny1])

static double blk_sum_s(orc *o, int f, int i, int j) {
    /* Eigen block(i, j, 2, 2).sum(), column-major order */
    return ((S_(o, f, i, j) + S_(o, f, i + 1, j)) + S_(o, f, i, j + 1)) + S_(o, f, i + 1, j + 1);
}

/* h_mec_rsf_v9.cpp:154-159 */
static void copy_bounds(orc *o, double *T) {
    int Ny1 = o->Ny1, Nx = o->Nx, Ny = o->Ny;
    for (int i = 0; i < Ny1; i++) T[i + 0 * Ny1] = T[i + 1 * Ny1];
    for (int i = 0; i < Ny1; i++) T[i + (int64_t)Nx * Ny1] = T[i + (int64_t)(Nx - 1) * Ny1];
    for (int j = 0; j < o->Nx1; j++) T[0 + (int64_t)j * Ny1] = T[1 + (int64_t)j * Ny1];
    for (int j = 0; j < o->Nx1; j++) T[Ny + (int64_t)j * Ny1] = T[Ny - 1 + (int64_t)j * Ny1];
}

/* ------------------------------------------------------------------ PARDISO */
typedef void (*pardiso_fn)(void *, const int *, const int *, const int *, const int *, const int *,
                           const void *, const int *, const int *, int *, const int *, int *,
                           const int *, void *, void *, int *);
static pardiso_fn g_pardiso = NULL;
static int g_pardiso_tried = 0;

static void load_pardiso(void) {
    if (g_pardiso_tried) return;
    g_pardiso_tried = 1;
    if (getenv("HMEC_ORACLE_NO_MKL")) return;
    const char *names[] = {"libmkl_rt.so", "libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so",
                           "/opt/conda/lib/libmkl_rt.so.1", NULL};
    for (int k = 0; names[k]; k++) {
        void *h = dlopen(names[k], RTLD_NOW | RTLD_GLOBAL);
        if (!h) continue;
        g_pardiso = (pardiso_fn)dlsym(h, "pardiso");
        if (g_pardiso) return;
    }
}

int orc_have_pardiso(void) {
    load_pardiso();
    return g_pardiso != NULL;
}

/* Eigen 3.3 PardisoImpl::pardisoInit for mtype 11 */
static void pardiso_init(orc *o) {
    memset(o->pt, 0, sizeof(o->pt));
    memset(o->iparm, 0, sizeof(o->iparm));
    o->iparm[0] = 1;
    o->iparm[1] = 2;
    o->iparm[7] = 2;
    o->iparm[9] = 13;
    o->iparm[10] = 1;
    o->iparm[12] = 1;
    o->iparm[17] = -1;
    o->iparm[18] = -1;
    o->iparm[34] = 1;
}

static void pardiso_release(orc *o) {
    if (!o->pardiso_ready) return;
    int maxfct = 1, mnum = 1, mtype = 11, phase = -1, n = o->N, nrhs = 1, msg = 0, err = 0;
    g_pardiso(o->pt, &maxfct, &mnum, &mtype, &phase, &n, NULL, o->rowp, o->coli, NULL, &nrhs,
              o->iparm, &msg, NULL, NULL, &err);
    o->pardiso_ready = 0;
}

/* Eigen PardisoLU analyzePattern + factorize + solve (h_mec_rsf_v9.cpp:1050-1052) */
static int pardiso_solve(orc *o) {
    int maxfct = 1, mnum = 1, mtype = 11, n = o->N, nrhs = 1, msg = 0, err = 0, phase;
    pardiso_release(o);
    pardiso_init(o);
    phase = 11;
    g_pardiso(o->pt, &maxfct, &mnum, &mtype, &phase, &n, o->val, o->rowp, o->coli, NULL, &nrhs,
              o->iparm, &msg, NULL, NULL, &err);
    o->pardiso_ready = 1;
    if (err) return err;
    phase = 22;
    g_pardiso(o->pt, &maxfct, &mnum, &mtype, &phase, &n, o->val, o->rowp, o->coli, NULL, &nrhs,
              o->iparm, &msg, NULL, NULL, &err);
    if (err) return err;
    o->lu_nnz = o->iparm[17];
    o->lu_mflops = o->iparm[18];
    phase = 33;
    g_pardiso(o->pt, &maxfct, &mnum, &mtype, &phase, &n, o->val, o->rowp, o->coli, NULL, &nrhs,
              o->iparm, &msg, o->R, o->S, &err);
    return err;
}

/* power of two nearest below 1/m (exact scaling) */
static double pow2_recip(double m) {
    if (!(m > 0) || !isfinite(m)) return 1.0;
    int e;
    frexp(m, &e);
    return ldexp(1.0, -e + 1);
}

/* Banded LU with partial pivoting (LAPACK dgbtrf/dgbtrs algorithm) on the
 * power-of-two equilibrated matrix, used when MKL is absent.  Without the
 * equilibration partial pivoting's growth destroys the solution of the
 * 101x51 system (LAPACK dgbsv fails the same way).  Band storage:
 * ab[(kl+ku+1+kl) x n], column-major. */
static int band_solve(orc *o) {
    const int n = o->N;
    double *rs = malloc(sizeof(double) * n), *cs = malloc(sizeof(double) * n);
    for (int r = 0; r < n; r++) {
        double m = 0;
        for (int p = o->rowp[r]; p < o->rowp[r + 1]; p++) m = dmax(m, fabs(o->val[p]));
        rs[r] = pow2_recip(m);
    }
    for (int c = 0; c < n; c++) cs[c] = 0;
    for (int r = 0; r < n; r++)
        for (int p = o->rowp[r]; p < o->rowp[r + 1]; p++)
            cs[o->coli[p]] = dmax(cs[o->coli[p]], fabs(o->val[p] * rs[r]));
    for (int c = 0; c < n; c++) cs[c] = pow2_recip(cs[c]);
    int kl = 0, ku = 0;
    for (int r = 0; r < n; r++)
        for (int p = o->rowp[r]; p < o->rowp[r + 1]; p++) {
            int c = o->coli[p];
            if (r - c > kl) kl = r - c;
            if (c - r > ku) ku = c - r;
        }
    const int ldab = 2 * kl + ku + 1, kv = kl + ku;
    double *ab = (double *)calloc((size_t)ldab * n, sizeof(double));
    int *ipiv = (int *)malloc(sizeof(int) * n);
    if (!ab || !ipiv) { free(ab); free(ipiv); free(rs); free(cs); return -99; }
#define AB(r, c) ab[(size_t)(kv + (r) - (c)) + (size_t)(c) * ldab]
    for (int r = 0; r < n; r++)
        for (int p = o->rowp[r]; p < o->rowp[r + 1]; p++)
            AB(r, o->coli[p]) += o->val[p] * rs[r] * cs[o->coli[p]];
    int ju = 0, info = 0;
    for (int j = 0; j < n; j++) {
        int km = kl < n - 1 - j ? kl : n - 1 - j;
        int jp = 0;
        double mx = fabs(AB(j, j));
        for (int t = 1; t <= km; t++)
            if (fabs(AB(j + t, j)) > mx) { mx = fabs(AB(j + t, j)); jp = t; }
        ipiv[j] = j + jp;
        if (AB(j + jp, j) == 0.0) { if (!info) info = j + 1; continue; }
        int jun = j + ku + jp;
        if (jun > n - 1) jun = n - 1;
        if (jun > ju) ju = jun;
        if (jp != 0)
            for (int c = j; c <= ju; c++) { double t = AB(j, c); AB(j, c) = AB(j + jp, c); AB(j + jp, c) = t; }
        double piv = AB(j, j);
        for (int t = 1; t <= km; t++) AB(j + t, j) /= piv;
        for (int c = j + 1; c <= ju; c++) {
            double u = AB(j, c);
            if (u != 0.0)
                for (int t = 1; t <= km; t++) AB(j + t, c) -= AB(j + t, j) * u;
        }
    }
    double *b = o->S;
    for (int r = 0; r < n; r++) b[r] = o->R[r] * rs[r];
    for (int j = 0; j < n; j++) {
        int km = kl < n - 1 - j ? kl : n - 1 - j;
        int l = ipiv[j];
        if (l != j) { double t = b[l]; b[l] = b[j]; b[j] = t; }
        for (int t = 1; t <= km; t++) b[j + t] -= AB(j + t, j) * b[j];
    }
    for (int j = n - 1; j >= 0; j--) {
        b[j] /= AB(j, j);
        int top = j - kv > 0 ? j - kv : 0;
        for (int r = top; r < j; r++) b[r] -= AB(r, j) * b[j];
    }
#undef AB
    for (int c = 0; c < n; c++) b[c] *= cs[c];
    free(ab);
    free(ipiv);
    free(rs);
    free(cs);
    return info ? -98 : 0;
}

/* ----------------------------------------------------------------- create */
void orc_destroy(orc *o);

orc *orc_create(const hmec_params *P, double *const *soa, int64_t M) {
    orc *o = (orc *)calloc(1, sizeof(orc));
    o->P = *P;
    o->Nx = P->Nx; o->Ny = P->Ny; o->Nx1 = P->Nx + 1; o->Ny1 = P->Ny + 1;
    o->N = o->Nx1 * o->Ny1 * 6;
    o->M = M;
    const double dx = P->dx, dy = P->dy;
    o->x = malloc(sizeof(double) * o->Nx);  orc_linspaced(o->x, o->Nx, 0, P->xsize);
    o->y = malloc(sizeof(double) * o->Ny);  orc_linspaced(o->y, o->Ny, 0, P->ysize);
    o->xvx = malloc(sizeof(double) * o->Nx1); orc_linspaced(o->xvx, o->Nx1, 0, P->xsize + dx);
    o->yvx = malloc(sizeof(double) * o->Ny1); orc_linspaced(o->yvx, o->Ny1, 0 - dy / 2., P->ysize + dy / 2.);
    o->xvy = malloc(sizeof(double) * o->Nx1); orc_linspaced(o->xvy, o->Nx1, 0 - dx / 2., P->xsize + dx / 2.);
    o->yvy = malloc(sizeof(double) * o->Ny1); orc_linspaced(o->yvy, o->Ny1, 0, P->ysize + dy);
    o->xp = malloc(sizeof(double) * o->Nx1);  orc_linspaced(o->xp, o->Nx1, 0 - dx / 2., P->xsize + dx / 2.);
    o->yp = malloc(sizeof(double) * o->Ny1);  orc_linspaced(o->yp, o->Ny1, 0 - dy / 2., P->ysize + dy / 2.);
    for (int f = 0; f < HMEC_NFIELDS; f++) {
        size_t n = f < HMEC_NBASIC_FIELDS ? (size_t)o->Ny * o->Nx : (size_t)o->Ny1 * o->Nx1;
        o->F[f] = calloc(n, sizeof(double));
    }
    for (int a = 0; a < 31; a++) o->acc[a] = calloc((size_t)o->Ny1 * o->Nx1, sizeof(double));
    for (int f = 0; f < HMEC_NMARKER_FIELDS; f++) {
        o->mk[f] = malloc(sizeof(double) * (M > 0 ? M : 1));
        if (soa && M > 0) memcpy(o->mk[f], soa[f], sizeof(double) * M);
    }
    o->DSYLSQ = calloc((size_t)P->niterglobal + 1, sizeof(double));
    o->series = calloc((size_t)6 * NUM_TIMESTEPS, sizeof(double));
    o->R = calloc(o->N, sizeof(double));
    o->S = calloc(o->N, sizeof(double));
    o->captrip = (int64_t)o->N * 10;
    o->ti = malloc(sizeof(int) * o->captrip);
    o->tj = malloc(sizeof(int) * o->captrip);
    o->tv = malloc(sizeof(double) * o->captrip);
    o->use_pardiso = orc_have_pardiso();

    /* fresh start, h_mec_rsf_v9.cpp:283-338 */
    o->timestep = 1;
    const int Ny = o->Ny, Nx = o->Nx;
    for (int j = 0; j < Nx; j++)
        for (int i = 0; i < Ny; i++) {
            B(o, HMEC_F_OM0, i, j) = P->omm[0];
            B(o, HMEC_F_ARSF, i, j) = P->arsfm[1];
            B(o, HMEC_F_BRSF, i, j) = P->brsfm[1];
            B(o, HMEC_F_LRSF, i, j) = P->lrsfm[0];
        }
    const double TS_1 = P->TS_1, TS_2 = P->TS_2, TS_3 = P->TS_3, TS_4 = P->TS_4;
    const double *arsfm = P->arsfm, *brsfm = P->brsfm, *lrsfm = P->lrsfm;
    for (int i = 0; i < Ny; i++) {
        if (o->y[i] > P->upper_block && o->y[i] < P->lower_block) {
            for (int j = 0; j < Nx; j++) {
                const double xj = o->x[j];
                B(o, HMEC_F_OM0, i, j) = P->omm[1];
                if (xj < TS_1) {
                    B(o, HMEC_F_BRSF, i, j) = brsfm[0];
                    B(o, HMEC_F_ARSF, i, j) = arsfm[0];
                }
                if (xj >= TS_1 && xj < TS_2) {
                    B(o, HMEC_F_BRSF, i, j) = brsfm[0] - (brsfm[0] - brsfm[1]) * ((xj - TS_1) / (TS_2 - TS_1));
                    B(o, HMEC_F_ARSF, i, j) = arsfm[0] - (arsfm[0] - arsfm[1]) * ((xj - TS_1) / (TS_2 - TS_1));
                    B(o, HMEC_F_LRSF, i, j) = lrsfm[0] - (lrsfm[0] - lrsfm[1]) * ((xj - TS_1) / (TS_2 - TS_1));
                }
                if (xj >= TS_2 && xj <= TS_3) B(o, HMEC_F_LRSF, i, j) = lrsfm[1];
                if (xj > TS_3 && xj <= TS_4) {
                    B(o, HMEC_F_BRSF, i, j) = brsfm[1] - (brsfm[1] - brsfm[0]) * ((xj - TS_3) / (TS_4 - TS_3));
                    B(o, HMEC_F_ARSF, i, j) = arsfm[1] - (arsfm[1] - arsfm[0]) * ((xj - TS_3) / (TS_4 - TS_3));
                    B(o, HMEC_F_LRSF, i, j) = lrsfm[1] - (lrsfm[1] - lrsfm[0]) * ((xj - TS_3) / (TS_4 - TS_3));
                }
                if (xj > TS_4) {
                    B(o, HMEC_F_BRSF, i, j) = brsfm[0];
                    B(o, HMEC_F_ARSF, i, j) = arsfm[0];
                }
            }
        }
    }
    memcpy(o->F[HMEC_F_OM], o->F[HMEC_F_OM0], sizeof(double) * Ny * Nx);
    for (int k = 0; k < o->Ny1 * o->Nx1; k++) {
        o->F[HMEC_F_PT0][k] = P->PCONF + P->PTFDIFF;
        o->F[HMEC_F_PF0][k] = P->PCONF;
    }
    o->timesum = 0;
    o->dt = P->dtelastic0;
    o->yndtdecrease = 1;
    return o;
}

void orc_destroy(orc *o) {
    if (!o) return;
    if (o->use_pardiso) pardiso_release(o);
    free(o->x); free(o->y); free(o->xvx); free(o->yvx); free(o->xvy); free(o->yvy); free(o->xp); free(o->yp);
    for (int f = 0; f < HMEC_NFIELDS; f++) free(o->F[f]);
    for (int a = 0; a < 31; a++) free(o->acc[a]);
    for (int f = 0; f < HMEC_NMARKER_FIELDS; f++) free(o->mk[f]);
    free(o->DSYLSQ); free(o->R); free(o->S); free(o->series);
    free(o->ti); free(o->tj); free(o->tv);
    free(o->rowp); free(o->coli); free(o->val); free(o->trip2slot);
    free(o);
}

double *orc_field(orc *o, int f) { return (f >= 0 && f < HMEC_NFIELDS) ? o->F[f] : NULL; }
double *orc_marker(orc *o, int f) { return (f >= 0 && f < HMEC_NMARKER_FIELDS) ? o->mk[f] : NULL; }
void orc_use_pardiso(orc *o, int on) { o->use_pardiso = on && orc_have_pardiso(); }

/* ---------------------------------------------------------------- scatter */
/* accumulator slots */
enum { A_ETA, A_RHO, A_KKK, A_TTT, A_SXY, A_GGG, A_ETA0, A_COHT, A_FRIT, A_COHC, A_FRIC, A_DILC, A_WT,
       A_ETAP, A_POR, A_SXX, A_SYY, A_GGGP, A_ETAP0, A_ETAB0, A_WTP,
       A_RHOX, A_ETADX, A_RHOFX, A_PORX, A_WTX,
       A_RHOY, A_ETADY, A_RHOFY, A_PORY, A_WTY };

/* anchor cell of marker m on grid g (0 basic, 1 P, 2 vx, 3 vy):
 * h_mec_rsf_v9.cpp:432-433, 468-469, 493-494, 515-516 */
void orc_marker_cell(orc *o, int g, int64_t m, int *ci, int *cj) {
    const double xm = o->mk[HMEC_M_XM][m], ym = o->mk[HMEC_M_YM][m];
    const double dx = o->P.dx, dy = o->P.dy;
    int j, i;
    switch (g) {
    case 0: j = orc_check_bounds(orc_fix(xm / dx), o->Nx); i = orc_check_bounds(orc_fix(ym / dy), o->Ny); break;
    case 1: j = orc_check_bounds(orc_fix((xm + dx / 2.) / dx), o->Nx); i = orc_check_bounds(orc_fix((ym + dy / 2.) / dy), o->Ny); break;
    case 2: j = orc_check_bounds(orc_fix((xm) / dx), o->Nx); i = orc_check_bounds(orc_fix((ym + dy / 2.) / dy), o->Ny); break;
    default: j = orc_check_bounds(orc_fix((xm + dx / 2.) / dx), o->Nx); i = orc_check_bounds(orc_fix((ym) / dy), o->Ny); break;
    }
    *ci = i;
    *cj = j;
}

static void add_blk(double *A, int rows, int i, int j, double v, const double w[4]) {
    /* Eigen: block(i, j, 2, 2) += v * wtm, wtm << w00, w01, w10, w11 */
    A[i + (int64_t)j * rows] += v * w[0];
    A[i + (int64_t)(j + 1) * rows] += v * w[1];
    A[i + 1 + (int64_t)j * rows] += v * w[2];
    A[i + 1 + (int64_t)(j + 1) * rows] += v * w[3];
}
static void add_wt(double *A, int rows, int i, int j, const double w[4]) {
    A[i + (int64_t)j * rows] += w[0];
    A[i + (int64_t)(j + 1) * rows] += w[1];
    A[i + 1 + (int64_t)j * rows] += w[2];
    A[i + 1 + (int64_t)(j + 1) * rows] += w[3];
}

/* h_mec_rsf_v9.cpp:384-583 */
void orc_scatter(orc *o) {
    const hmec_params *P = &o->P;
    const int Ny = o->Ny, Ny1 = o->Ny1, Nx = o->Nx, Nx1 = o->Nx1;
    const double dx = P->dx, dy = P->dy;
    for (int a = 0; a < 31; a++) memset(o->acc[a], 0, sizeof(double) * Ny1 * Nx1);
    double **mk = o->mk;
    for (int64_t m = 0; m < o->M; m++) {
        double cohescmm, cohestmm, frictcmm, dilatcmm, fricttmm, etasmm0, etamm0, etamm, rhomm, etadm;
        const double porm = mk[HMEC_M_PORM][m];
        double kkkmm = mk[HMEC_M_KKKM][m] * pow(porm / P->POR0, 3);
        kkkmm = enforce_bounds(kkkmm, P->kkkmin, P->kkkmax);
        if (mk[HMEC_M_T_MARKER][m] != 0) {
            cohescmm = mk[HMEC_M_COHESCM][m] * (1 - porm);
            cohestmm = mk[HMEC_M_COHESTM][m] * (1 - porm);
            frictcmm = mk[HMEC_M_FRICTCM][m];
            dilatcmm = mk[HMEC_M_DILATCM][m];
            fricttmm = mk[HMEC_M_FRICTTM][m];
            etasmm0 = mk[HMEC_M_ETASM][m];
            etamm0 = mk[HMEC_M_ETASM][m] * exp(-P->alpha * porm);
            etamm = mk[HMEC_M_ETAM][m];
            rhomm = mk[HMEC_M_RHOM][m] * (1 - porm) + mk[HMEC_M_RHOFM][m] * porm;
        } else {
            cohescmm = mk[HMEC_M_COHESCM][m];
            cohestmm = mk[HMEC_M_COHESTM][m];
            frictcmm = mk[HMEC_M_FRICTCM][m];
            dilatcmm = 0;
            fricttmm = mk[HMEC_M_FRICTTM][m];
            etasmm0 = mk[HMEC_M_ETASM][m];
            etamm0 = mk[HMEC_M_ETASM][m];
            etamm = mk[HMEC_M_ETAM][m];
            rhomm = mk[HMEC_M_RHOM][m];
        }
        etamm0 = enforce_bounds(etamm0, P->etamin, P->etamax);
        etamm = enforce_bounds(etamm, P->etamin, P->etamax);
        etadm = mk[HMEC_M_ETAFM][m] / kkkmm;
        const double xm = mk[HMEC_M_XM][m], ym = mk[HMEC_M_YM][m];
        const double invg = 1. / mk[HMEC_M_GM][m];
        double w[4];
        int i, j;
        double dxm, dym;
        /* basic nodes */
        j = orc_check_bounds(orc_fix(xm / dx), Nx);
        i = orc_check_bounds(orc_fix(ym / dy), Ny);
        dxm = (xm - o->x[j]) / dx;
        dym = (ym - o->y[i]) / dy;
        w[0] = (1 - dxm) * (1 - dym); w[1] = dxm * (1 - dym); w[2] = (1 - dxm) * dym; w[3] = dxm * dym;
        add_blk(o->acc[A_ETA], Ny, i, j, etamm, w);
        add_blk(o->acc[A_RHO], Ny, i, j, rhomm, w);
        add_blk(o->acc[A_KKK], Ny, i, j, kkkmm, w);
        add_blk(o->acc[A_TTT], Ny, i, j, mk[HMEC_M_T_MARKER][m], w);
        add_blk(o->acc[A_SXY], Ny, i, j, mk[HMEC_M_SXYM][m], w);
        add_blk(o->acc[A_GGG], Ny, i, j, invg, w);
        add_blk(o->acc[A_ETA0], Ny, i, j, etamm0, w);
        add_blk(o->acc[A_COHT], Ny, i, j, cohestmm, w);
        add_blk(o->acc[A_FRIT], Ny, i, j, fricttmm, w);
        add_blk(o->acc[A_COHC], Ny, i, j, cohescmm, w);
        add_blk(o->acc[A_FRIC], Ny, i, j, frictcmm, w);
        add_blk(o->acc[A_DILC], Ny, i, j, dilatcmm, w);
        add_wt(o->acc[A_WT], Ny, i, j, w);
        /* pressure nodes */
        j = orc_check_bounds(orc_fix((xm + dx / 2.) / dx), Nx);
        i = orc_check_bounds(orc_fix((ym + dy / 2.) / dy), Ny);
        dxm = (xm - o->xp[j]) / dx;
        dym = (ym - o->yp[i]) / dy;
        w[0] = (1 - dxm) * (1 - dym); w[1] = dxm * (1 - dym); w[2] = (1 - dxm) * dym; w[3] = dxm * dym;
        add_blk(o->acc[A_ETAP], Ny1, i, j, etamm, w);
        add_blk(o->acc[A_POR], Ny1, i, j, porm, w);
        add_blk(o->acc[A_SXX], Ny1, i, j, mk[HMEC_M_SXXM][m], w);
        add_blk(o->acc[A_SYY], Ny1, i, j, mk[HMEC_M_SYYM][m], w);
        add_blk(o->acc[A_GGGP], Ny1, i, j, invg, w);
        add_blk(o->acc[A_ETAP0], Ny1, i, j, etamm0, w);
        add_blk(o->acc[A_ETAB0], Ny1, i, j, etasmm0, w);
        add_wt(o->acc[A_WTP], Ny1, i, j, w);
        /* vx nodes */
        j = orc_check_bounds(orc_fix((xm) / dx), Nx);
        i = orc_check_bounds(orc_fix((ym + dy / 2.) / dy), Ny);
        dxm = (xm - o->xvx[j]) / dx;
        dym = (ym - o->yvx[i]) / dy;
        w[0] = (1 - dxm) * (1 - dym); w[1] = dxm * (1 - dym); w[2] = (1 - dxm) * dym; w[3] = dxm * dym;
        add_blk(o->acc[A_RHOX], Ny1, i, j, rhomm, w);
        add_blk(o->acc[A_ETADX], Ny1, i, j, 1. / etadm, w);
        add_blk(o->acc[A_RHOFX], Ny1, i, j, mk[HMEC_M_RHOFM][m], w);
        add_blk(o->acc[A_PORX], Ny1, i, j, porm, w);
        add_wt(o->acc[A_WTX], Ny1, i, j, w);
        /* vy nodes */
        j = orc_check_bounds(orc_fix((xm + dx / 2.) / dx), Nx);
        i = orc_check_bounds(orc_fix((ym) / dy), Ny);
        dxm = (xm - o->xvy[j]) / dx;
        dym = (ym - o->yvy[i]) / dy;
        w[0] = (1 - dxm) * (1 - dym); w[1] = dxm * (1 - dym); w[2] = (1 - dxm) * dym; w[3] = dxm * dym;
        add_blk(o->acc[A_RHOY], Ny1, i, j, rhomm, w);
        add_blk(o->acc[A_ETADY], Ny1, i, j, 1. / etadm, w);
        add_blk(o->acc[A_RHOFY], Ny1, i, j, mk[HMEC_M_RHOFM][m], w);
        add_blk(o->acc[A_PORY], Ny1, i, j, porm, w);
        add_wt(o->acc[A_WTY], Ny1, i, j, w);
    }
    /* normalization, h_mec_rsf_v9.cpp:531-583 */
#define AB_(a, i, j) (o->acc[a][(i) + (int64_t)(j) * Ny])
#define AS_(a, i, j) (o->acc[a][(i) + (int64_t)(j) * Ny1])
    for (int i = 0; i < Ny; i++) {
        for (int j = 0; j < Nx; j++) {
            double wtsum = AB_(A_WT, i, j);
            if (wtsum > 0) {
                B(o, HMEC_F_RHO, i, j) = AB_(A_RHO, i, j) / wtsum;
                B(o, HMEC_F_KKK, i, j) = AB_(A_KKK, i, j) / wtsum;
                B(o, HMEC_F_TTT, i, j) = AB_(A_TTT, i, j) / wtsum;
                B(o, HMEC_F_GGG, i, j) = 1. / (AB_(A_GGG, i, j) / wtsum);
                B(o, HMEC_F_ETA0, i, j) = AB_(A_ETA0, i, j) / wtsum;
                B(o, HMEC_F_COHT, i, j) = AB_(A_COHT, i, j) / wtsum;
                B(o, HMEC_F_FRIT, i, j) = AB_(A_FRIT, i, j) / wtsum;
                B(o, HMEC_F_COHC, i, j) = AB_(A_COHC, i, j) / wtsum;
                B(o, HMEC_F_FRIC, i, j) = AB_(A_FRIC, i, j) / wtsum;
                B(o, HMEC_F_DILC, i, j) = AB_(A_DILC, i, j) / wtsum;
            }
        }
        for (int j = 0; j < Nx1; j++) {
            double wtysum = AS_(A_WTY, i, j);
            if (wtysum > 0) {
                S_(o, HMEC_F_RHOY, i, j) = AS_(A_RHOY, i, j) / wtysum;
                S_(o, HMEC_F_ETADY, i, j) = 1. / (AS_(A_ETADY, i, j) / wtysum);
                S_(o, HMEC_F_RHOFY, i, j) = AS_(A_RHOFY, i, j) / wtysum;
                S_(o, HMEC_F_PORY, i, j) = AS_(A_PORY, i, j) / wtysum;
            }
        }
    }
    for (int i = 0; i < Ny1; i++) {
        for (int j = 0; j < Nx; j++) {
            double wtxsum = AS_(A_WTX, i, j);
            if (wtxsum > 0) {
                S_(o, HMEC_F_RHOX, i, j) = AS_(A_RHOX, i, j) / wtxsum;
                S_(o, HMEC_F_ETADX, i, j) = 1. / (AS_(A_ETADX, i, j) / wtxsum);
                S_(o, HMEC_F_RHOFX, i, j) = AS_(A_RHOFX, i, j) / wtxsum;
                S_(o, HMEC_F_PORX, i, j) = AS_(A_PORX, i, j) / wtxsum;
            }
        }
        for (int j = 0; j < Nx1; j++) {
            double wtpsum = AS_(A_WTP, i, j);
            if (wtpsum > 0) {
                S_(o, HMEC_F_ETAP, i, j) = AS_(A_ETAP, i, j) / wtpsum;
                S_(o, HMEC_F_POR, i, j) = AS_(A_POR, i, j) / wtpsum;
                S_(o, HMEC_F_ETAB0, i, j) = AS_(A_ETAB0, i, j) / wtpsum / S_(o, HMEC_F_POR, i, j);
                S_(o, HMEC_F_ETAP0, i, j) = AS_(A_ETAP0, i, j) / wtpsum;
                S_(o, HMEC_F_GGGP, i, j) = 1. / (AS_(A_GGGP, i, j) / wtpsum);
            }
        }
    }
#undef AB_
#undef AS_
}

/* raw accumulator access for tests (31 slots, Ny1 x Nx1 storage, basic
 * accumulators use leading dimension Ny) */
double *orc_accumulator(orc *o, int a) { return (a >= 0 && a < 31) ? o->acc[a] : NULL; }

/* ------------------------------------------------------------ step begin */
static void copyf(orc *o, int dst, int src) {
    size_t n = dst < HMEC_NBASIC_FIELDS ? (size_t)o->Ny * o->Nx : (size_t)o->Ny1 * o->Nx1;
    memcpy(o->F[dst], o->F[src], n * sizeof(double));
}

/* h_mec_rsf_v9.cpp:586-608 */
void orc_step_begin(orc *o) {
    const hmec_params *P = &o->P;
    int src = (o->timestep == 1) ? HMEC_F_ETA0 : HMEC_F_ETA00;
    copyf(o, HMEC_F_ETA1, src);
    copyf(o, HMEC_F_ETA, src);
    copyf(o, HMEC_F_ETA50, src);
    copyf(o, HMEC_F_OM, HMEC_F_OM0);
    if (!o->yndtdecrease) o->dt = dmax(dmin(o->dt * P->dtkoefup, P->dtelastic0), P->dtmin);
    else o->dt = dmax(dmin(o->dt, P->dtelastic0), P->dtmin);
    o->yndtdecrease = 0;
    o->dt00 = o->dt;
    o->ynlast = 0;
    memset(o->DSYLSQ, 0, sizeof(double) * (P->niterglobal + 1));
}

/* ------------------------------------------------------------- VEP */
/* h_mec_rsf_v9.cpp:612-669 */
void orc_vep_pre(orc *o) {
    const hmec_params *P = &o->P;
    const int Ny = o->Ny, Nx = o->Nx;
    const double dt = o->dt;
    const double etamincur = dt * P->shearmod * 1e-4;
    const double dy_faultw = P->dy / P->faultwidth;
    copy_bounds(o, o->F[HMEC_F_PT]);
    copy_bounds(o, o->F[HMEC_F_PF]);
    for (int i = 0; i < Ny; i++) {
        for (int j = 0; j < Nx; j++) {
            if (B(o, HMEC_F_ETA, i, j) < etamincur) B(o, HMEC_F_ETA, i, j) = etamincur;
            if (B(o, HMEC_F_ETA, i, j) < B(o, HMEC_F_ETA0, i, j)) {
                const double SXX_temp = blk_sum_s(o, HMEC_F_SXX, i, j), SYY_temp = blk_sum_s(o, HMEC_F_SYY, i, j);
                B(o, HMEC_F_SIIB, i, j) = sqrt(SQ(B(o, HMEC_F_SXY, i, j)) + (SQ(SXX_temp) + SQ(SYY_temp) + SQ(SXX_temp + SYY_temp)) / 32.);
                B(o, HMEC_F_IETAPLB, i, j) = (1. / B(o, HMEC_F_ETA, i, j) - 1. / B(o, HMEC_F_ETA0, i, j));
                B(o, HMEC_F_EIIB, i, j) = dy_faultw * B(o, HMEC_F_SIIB, i, j) * B(o, HMEC_F_IETAPLB, i, j) / 2.;
            } else {
                B(o, HMEC_F_EIIB, i, j) = 0;
                B(o, HMEC_F_IETAPLB, i, j) = 0;
            }
        }
    }
    for (int i = 1; i < Ny; i++) {
        for (int j = 1; j < Nx; j++) {
            const double IETAPL = (((B(o, HMEC_F_IETAPLB, i - 1, j - 1) + B(o, HMEC_F_IETAPLB, i, j - 1)) +
                                    B(o, HMEC_F_IETAPLB, i - 1, j)) + B(o, HMEC_F_IETAPLB, i, j)) / 4.;
            if (B(o, HMEC_F_YNY0, i - 1, j - 1) > 0 || B(o, HMEC_F_YNY0, i, j - 1) > 0 ||
                B(o, HMEC_F_YNY0, i - 1, j) > 0 || B(o, HMEC_F_YNY0, i, j) > 0) {
                S_(o, HMEC_F_ETAP, i, j) = 1. / (1. / S_(o, HMEC_F_ETAP0, i, j) + IETAPL);
                S_(o, HMEC_F_ETAB, i, j) = 1. / (1. / S_(o, HMEC_F_ETAB0, i, j) + dy_faultw * IETAPL * S_(o, HMEC_F_POR, i, j));
            } else {
                S_(o, HMEC_F_ETAP, i, j) = S_(o, HMEC_F_ETAP0, i, j);
                S_(o, HMEC_F_ETAB, i, j) = S_(o, HMEC_F_ETAB0, i, j);
            }
            if (S_(o, HMEC_F_ETAP, i, j) < etamincur) S_(o, HMEC_F_ETAP, i, j) = etamincur;
            if (S_(o, HMEC_F_ETAB, i, j) * S_(o, HMEC_F_POR, i, j) < etamincur)
                S_(o, HMEC_F_ETAB, i, j) = etamincur / S_(o, HMEC_F_POR, i, j);
            S_(o, HMEC_F_GGGB, i, j) = S_(o, HMEC_F_GGGP, i, j) / S_(o, HMEC_F_POR, i, j);
            S_(o, HMEC_F_DILP, i, j) = 0;
        }
    }
}

/* h_mec_rsf_v9.cpp:671-677 */
double orc_scale(orc *o) {
    const hmec_params *P = &o->P;
    double mn = INFINITY;
    for (int j = 0; j < o->Nx; j++)
        for (int i = 1; i < o->Ny; i++) mn = dmin(mn, S_(o, HMEC_F_ETADX, i, j));
    double pfscale;
    if (o->dt > 2e5) pfscale = mn * P->dx * 1e17 / SQ(o->dt);
    else pfscale = mn * P->dx;
    o->ptscale = pfscale;
    return pfscale;
}

/* -------------------------------------------------------- assembly */
static void T(orc *o, int r, int c, double v) {
    if (o->ntrip >= o->captrip) {
        o->captrip *= 2;
        o->ti = realloc(o->ti, sizeof(int) * o->captrip);
        o->tj = realloc(o->tj, sizeof(int) * o->captrip);
        o->tv = realloc(o->tv, sizeof(double) * o->captrip);
    }
    o->ti[o->ntrip] = r;
    o->tj[o->ntrip] = c;
    o->tv[o->ntrip] = v;
    o->ntrip++;
}

/* triplets in the reference's emission order, h_mec_rsf_v9.cpp:682-980 */
static void emit_triplets(orc *o) {
    const hmec_params *P = &o->P;
    const int Nx = o->Nx, Ny = o->Ny, Ny1 = o->Ny1, Nx1 = o->Nx1, NV = 6;
    const double dt = o->dt, dx = P->dx, dy = P->dy, dx2 = pow(dx, 2), dy2 = pow(dy, 2), dx_dy = dx * dy;
    const double gx = P->gx, gy = P->gy, inertia = P->inertia;
    const double ptscale = o->ptscale, pfscale = o->ptscale;
    const double BETASOLID = P->BETASOLID, BETAFLUID = P->BETAFLUID;
    double *R = o->R;
    o->ntrip = 0;
    memset(R, 0, sizeof(double) * o->N);
    for (int j = 0; j < Nx1; j++) {
        for (int i = 0; i < Ny1; i++) {
            const int kp = (j * Ny1 + i) * NV, kx = kp + 1, ky = kp + 2, kpf = kp + 3, kxf = kp + 4, kyf = kp + 5;
            /* 5a) vxs, :698-768 */
            if (i == 0 || i == Ny || j == 0 || j >= Nx - 1) {
                if (j == Nx) T(o, kx, kx, 1);
                if (i == 0 && j < Nx) { T(o, kx, kx, 1); T(o, kx, kx + NV, 1); R[kx] = 2 * P->bcupper; }
                if (i == Ny && j < Nx) { T(o, kx, kx, 1); T(o, kx, kx - NV, 1); R[kx] = 2 * P->bclower; }
                if (j == 0 && i > 0 && i < Ny) { T(o, kx, kx, 1); T(o, kx, kx + NV * Ny1, -1); }
                if (j == Nx - 1 && i > 0 && i < Ny) { T(o, kx, kx, 1); T(o, kx, kx - NV * Ny1, -1); }
            } else {
                double ETAXY1 = B(o, HMEC_F_ETA, i - 1, j), ETAXY2 = B(o, HMEC_F_ETA, i, j);
                double ETAXX1 = S_(o, HMEC_F_ETAP, i, j), ETAXX2 = S_(o, HMEC_F_ETAP, i, j + 1);
                const double GXY1 = B(o, HMEC_F_GGG, i - 1, j), GXY2 = B(o, HMEC_F_GGG, i, j);
                const double GXX1 = S_(o, HMEC_F_GGGP, i, j), GXX2 = S_(o, HMEC_F_GGGP, i, j + 1);
                const double KXY1 = orc_divplus(dt * GXY1, ETAXY1), KXY2 = orc_divplus(dt * GXY2, ETAXY2);
                const double KXX1 = orc_divplus(dt * GXX1, ETAXX1), KXX2 = orc_divplus(dt * GXX2, ETAXX2);
                ETAXY1 *= KXY1; ETAXY2 *= KXY2; ETAXX1 *= KXX1; ETAXX2 *= KXX2;
                const double SXY1 = B(o, HMEC_F_SXY0, i - 1, j) * (1 - KXY1), SXY2 = B(o, HMEC_F_SXY0, i, j) * (1 - KXY2);
                const double SXX1 = S_(o, HMEC_F_SXX0, i, j) * (1 - KXX1), SXX2 = S_(o, HMEC_F_SXX0, i, j + 1) * (1 - KXX2);
                const double dRHOdx = (S_(o, HMEC_F_RHOX, i, j + 1) - S_(o, HMEC_F_RHOX, i, j - 1)) / (2. * dx);
                const double dRHOdy = (B(o, HMEC_F_RHO, i, j) - B(o, HMEC_F_RHO, i - 1, j)) / dy;
                const double temp = gx * dt * dRHOdy / 4.;
                T(o, kx, kx, -(ETAXX1 + ETAXX2) / dx2 - (ETAXY1 + ETAXY2) / dy2 - gx * dt * dRHOdx - inertia * S_(o, HMEC_F_RHOX, i, j) / dt);
                T(o, kx, kx - Ny1 * NV, ETAXX1 / dx2);
                T(o, kx, kx + Ny1 * NV, ETAXX2 / dx2);
                T(o, kx, kx - NV, ETAXY1 / dy2);
                T(o, kx, kx + NV, ETAXY2 / dy2);
                T(o, kx, ky - NV, (ETAXY1 - ETAXX1) / dx_dy - temp);
                T(o, kx, ky, (ETAXX1 - ETAXY2) / dx_dy - temp);
                T(o, kx, ky - NV + Ny1 * NV, (ETAXX2 - ETAXY1) / dx_dy - temp);
                T(o, kx, ky + Ny1 * NV, (ETAXY2 - ETAXX2) / dx_dy - temp);
                T(o, kx, kp, ptscale / dx);
                T(o, kx, kp + Ny1 * NV, -ptscale / dx);
                R[kx] = -S_(o, HMEC_F_RHOX, i, j) * (inertia * S_(o, HMEC_F_VX0, i, j) / dt + gx) - (SXX2 - SXX1) / dx - (SXY2 - SXY1) / dy;
            }
            /* 5b) vys, :770-840 */
            if (j == 0 || j == Nx || i == 0 || i >= Ny - 1) {
                if (i == Ny) T(o, ky, ky, 1);
                if (j == 0) { T(o, ky, ky, 1); T(o, ky, ky + Ny1 * NV, 1); }
                if (j == Nx) { T(o, ky, ky, 1); T(o, ky, ky - Ny1 * NV, 1); }
                if (i == 0 && j > 0 && j < Nx) T(o, ky, ky, 1);
                if (i == Ny - 1 && j > 0 && j < Nx) T(o, ky, ky, 1);
            } else {
                double ETAXY1 = B(o, HMEC_F_ETA, i, j - 1), ETAXY2 = B(o, HMEC_F_ETA, i, j);
                double ETAYY1 = S_(o, HMEC_F_ETAP, i, j), ETAYY2 = S_(o, HMEC_F_ETAP, i + 1, j);
                const double GXY1 = B(o, HMEC_F_GGG, i, j - 1), GXY2 = B(o, HMEC_F_GGG, i, j);
                const double GYY1 = S_(o, HMEC_F_GGGP, i, j), GYY2 = S_(o, HMEC_F_GGGP, i + 1, j);
                const double KXY1 = orc_divplus(dt * GXY1, ETAXY1), KXY2 = orc_divplus(dt * GXY2, ETAXY2);
                const double KYY1 = orc_divplus(dt * GYY1, ETAYY1), KYY2 = orc_divplus(dt * GYY2, ETAYY2);
                ETAXY1 *= KXY1; ETAXY2 *= KXY2; ETAYY1 *= KYY1; ETAYY2 *= KYY2;
                const double SXY1 = B(o, HMEC_F_SXY0, i, j - 1) * (1 - KXY1), SXY2 = B(o, HMEC_F_SXY0, i, j) * (1 - KXY2);
                const double SYY1 = S_(o, HMEC_F_SYY0, i, j) * (1 - KYY1), SYY2 = S_(o, HMEC_F_SYY0, i + 1, j) * (1 - KYY2);
                const double dRHOdy = (S_(o, HMEC_F_RHOY, i + 1, j) - S_(o, HMEC_F_RHOY, i - 1, j)) / 2. / dy;
                const double dRHOdx = (B(o, HMEC_F_RHO, i, j) - B(o, HMEC_F_RHO, i, j - 1)) / dx;
                const double temp = gy * dt * dRHOdx / 4.;
                T(o, ky, ky, -(ETAYY1 + ETAYY2) / dy2 - (ETAXY1 + ETAXY2) / dx2 - gy * dt * dRHOdy - inertia * S_(o, HMEC_F_RHOY, i, j) / dt);
                T(o, ky, ky - Ny1 * NV, ETAXY1 / dx2);
                T(o, ky, ky + Ny1 * NV, ETAXY2 / dx2);
                T(o, ky, ky - NV, ETAYY1 / dy2);
                T(o, ky, ky + NV, ETAYY2 / dy2);
                T(o, ky, kx - Ny1 * NV, (ETAXY1 - ETAYY1) / dx_dy - temp);
                T(o, ky, kx + NV - Ny1 * NV, (ETAYY2 - ETAXY1) / dx_dy - temp);
                T(o, ky, kx, (ETAYY1 - ETAXY2) / dx_dy - temp);
                T(o, ky, kx + NV, (ETAXY2 - ETAYY2) / dx_dy - temp);
                T(o, ky, kp, ptscale / dy);
                T(o, ky, kp + NV, -ptscale / dy);
                R[ky] = -S_(o, HMEC_F_RHOY, i, j) * (inertia * S_(o, HMEC_F_VY0, i, j) / dt + gy) - (SYY2 - SYY1) / dy - (SXY2 - SXY1) / dx;
            }
            /* 5c) Pt, :842-862 */
            if (i == 0 || j == 0 || i == Ny || j == Nx) {
                T(o, kp, kp, 1);
            } else {
                const double POR = S_(o, HMEC_F_POR, i, j), ETAB = S_(o, HMEC_F_ETAB, i, j);
                const double BETADRAINED = (1. / S_(o, HMEC_F_GGGB, i, j) + BETASOLID) / (1 - POR);
                const double KBW = 1 - BETASOLID / BETADRAINED;
                T(o, kp, kx - Ny1 * NV, -1. / dx);
                T(o, kp, kx, 1. / dx);
                T(o, kp, ky - NV, -1. / dy);
                T(o, kp, ky, 1. / dy);
                T(o, kp, kp, ptscale * (1. / ETAB / (1 - POR) + BETADRAINED / dt));
                T(o, kp, kpf, -pfscale * (1. / ETAB / (1 - POR) + BETADRAINED * KBW / dt));
                R[kp] = BETADRAINED * (S_(o, HMEC_F_PT0, i, j) - KBW * S_(o, HMEC_F_PF0, i, j)) / dt + S_(o, HMEC_F_DILP, i, j);
            }
            /* 5d) vxD, :864-903 */
            if (i == 0 || i == Ny || j == 0 || j >= Nx - 1) {
                if (j == Nx) T(o, kxf, kxf, 1);
                if (i == 0 && j < Nx) { T(o, kxf, kxf, 1); T(o, kxf, kxf + NV, -1); }
                if (i == Ny && j < Nx) { T(o, kxf, kxf, 1); T(o, kxf, kxf - NV, -1); }
                if (j == 0) T(o, kxf, kxf, 1);
                if (j == Nx - 1) T(o, kxf, kxf, 1);
            } else {
                const double RHOFX = S_(o, HMEC_F_RHOFX, i, j);
                T(o, kxf, kxf, -S_(o, HMEC_F_ETADX, i, j) - RHOFX / S_(o, HMEC_F_PORX, i, j) * inertia / dt);
                T(o, kxf, kx, -RHOFX * inertia / dt);
                T(o, kxf, kpf, pfscale / dx);
                T(o, kxf, kpf + Ny1 * NV, -pfscale / dx);
                R[kxf] = -RHOFX * (inertia * S_(o, HMEC_F_VXF0, i, j) / dt + gx);
            }
            /* 5e) vyD, :905-949 */
            if (j == 0 || j == Nx || i == 0 || i >= Ny - 1) {
                if (i == Ny) T(o, kyf, kyf, 1);
                if (j == 0 && i > 0 && i < Ny - 1) { T(o, kyf, kyf, 1); T(o, kyf, kyf + Ny1 * NV, -1); }
                if (j == Nx && i > 0 && i < Ny - 1) { T(o, kyf, kyf, 1); T(o, kyf, kyf - Ny1 * NV, -1); }
                if (i == 0) { T(o, kyf, kyf, 1); R[kyf] = P->bcvyflower; }
                if (i == Ny - 1) { T(o, kyf, kyf, 1); R[kyf] = P->bcvyflower; }
            } else {
                const double RHOFY = S_(o, HMEC_F_RHOFY, i, j);
                T(o, kyf, kyf, -S_(o, HMEC_F_ETADY, i, j) - RHOFY / S_(o, HMEC_F_PORY, i, j) * inertia / dt);
                T(o, kyf, ky, -RHOFY * inertia / dt);
                T(o, kyf, kpf, pfscale / dy);
                T(o, kyf, kpf + NV, -pfscale / dy);
                R[kyf] = -RHOFY * (inertia * S_(o, HMEC_F_VYF0, i, j) / dt + gy);
            }
            /* 5f) Pf, :951-980 */
            if (j == 0 || j == Nx || i <= 1 || i >= Ny - 1) {
                if (i == 1 || i == Ny - 1) {
                    T(o, kpf, kpf, pfscale);
                    T(o, kpf, kp, -ptscale);
                    R[kpf] = -P->PTFDIFF;
                } else {
                    T(o, kpf, kpf, 1);
                }
            } else {
                const double POR = S_(o, HMEC_F_POR, i, j), ETAB = S_(o, HMEC_F_ETAB, i, j);
                const double BETADRAINED = (1 / S_(o, HMEC_F_GGGB, i, j) + BETASOLID) / (1 - POR);
                const double KBW = 1 - BETASOLID / BETADRAINED;
                const double KSK = (BETADRAINED - BETASOLID) / (BETADRAINED - BETASOLID + POR * (BETAFLUID - BETASOLID));
                T(o, kpf, kxf - Ny1 * NV, -1. / dx);
                T(o, kpf, kxf, 1. / dx);
                T(o, kpf, kyf - NV, -1. / dy);
                T(o, kpf, kyf, 1. / dy);
                T(o, kpf, kp, -ptscale * (1 / ETAB / (1 - POR) + BETADRAINED * KBW / dt));
                T(o, kpf, kpf, pfscale * (1 / ETAB / (1 - POR) + BETADRAINED * KBW / KSK / dt));
                R[kpf] = -BETADRAINED * KBW * (S_(o, HMEC_F_PT0, i, j) - 1. / KSK * S_(o, HMEC_F_PF0, i, j)) / dt - S_(o, HMEC_F_DILP, i, j);
            }
        }
    }
}

/* setFromTriplets + makeCompressed (h_mec_rsf_v9.cpp:1046-1049): CSR with
 * columns ascending, duplicates summed in emission order.  The pattern is
 * built at the first assembly and reused (it depends on Nx, Ny only). */
static void build_csr(orc *o) {
    const int n = o->N;
    const int64_t nt = o->ntrip;
    if (!o->rowp) {
        int *cnt = calloc(n + 1, sizeof(int));
        for (int64_t t = 0; t < nt; t++) cnt[o->ti[t] + 1]++;
        for (int r = 0; r < n; r++) cnt[r + 1] += cnt[r];
        int *pos = malloc(sizeof(int) * n);
        memcpy(pos, cnt, sizeof(int) * n);
        int64_t *ord = malloc(sizeof(int64_t) * nt);
        for (int64_t t = 0; t < nt; t++) ord[pos[o->ti[t]]++] = t;
        /* stable insertion sort of each row by column */
        for (int r = 0; r < n; r++) {
            for (int a = cnt[r] + 1; a < cnt[r + 1]; a++) {
                int64_t t = ord[a];
                int b = a - 1;
                while (b >= cnt[r] && o->tj[ord[b]] > o->tj[t]) { ord[b + 1] = ord[b]; b--; }
                ord[b + 1] = t;
            }
        }
        o->rowp = calloc(n + 1, sizeof(int));
        o->coli = malloc(sizeof(int) * nt);
        o->trip2slot = malloc(sizeof(int) * nt);
        int nz = 0;
        for (int r = 0; r < n; r++) {
            o->rowp[r] = nz;
            int last = -1;
            for (int a = cnt[r]; a < cnt[r + 1]; a++) {
                int64_t t = ord[a];
                if (o->tj[t] != last) { o->coli[nz] = o->tj[t]; last = o->tj[t]; nz++; }
                o->trip2slot[t] = nz - 1;
            }
        }
        o->rowp[n] = nz;
        o->nnz = nz;
        o->val = malloc(sizeof(double) * nz);
        free(cnt); free(pos); free(ord);
    }
    memset(o->val, 0, sizeof(double) * o->nnz);
    for (int64_t t = 0; t < nt; t++) o->val[o->trip2slot[t]] += o->tv[t];
}

void orc_assemble(orc *o) {
    emit_triplets(o);
    build_csr(o);
}

int orc_csr(orc *o, int **rowp, int **coli, double **val, double **rhs, int *n, int *nnz) {
    if (rowp) *rowp = o->rowp;
    if (coli) *coli = o->coli;
    if (val) *val = o->val;
    if (rhs) *rhs = o->R;
    if (n) *n = o->N;
    if (nnz) *nnz = o->nnz;
    return o->rowp ? 0 : -1;
}
int64_t orc_triplet_count(orc *o) { return o->ntrip; }
/* triplet (row, col) lists of the last assembly, emission order */
void orc_triplets(orc *o, int *ti, int *tj, double *tv) {
    if (ti) memcpy(ti, o->ti, sizeof(int) * o->ntrip);
    if (tj) memcpy(tj, o->tj, sizeof(int) * o->ntrip);
    if (tv) memcpy(tv, o->tv, sizeof(double) * o->ntrip);
}

/* ------------------------------------------------------ solve (a10, a11) */
int orc_solve(orc *o) {
    int err;
    if (o->use_pardiso) err = pardiso_solve(o);
    else err = band_solve(o);
    return err;
}
double *orc_solution(orc *o) { return o->S; }
/* x = A^-1 b with the PARDISO factor of the last orc_solve (phase 33 only,
 * with its refinement steps): the correction solve of the tests'
 * extended-precision refinement, which measures the oracle's forward error */
int orc_resolve(orc *o, const double *b, double *x) {
    if (!o->use_pardiso || !o->pardiso_ready) return -1;
    int maxfct = 1, mnum = 1, mtype = 11, n = o->N, nrhs = 1, msg = 0, err = 0, phase = 33;
    g_pardiso(o->pt, &maxfct, &mnum, &mtype, &phase, &n, o->val, o->rowp, o->coli, NULL, &nrhs,
              o->iparm, &msg, (void *)b, x, &err);
    return err;
}
void orc_lu_stats(orc *o, int64_t *nnz, double *mflops) {
    if (nnz) *nnz = o->lu_nnz;
    if (mflops) *mflops = o->lu_mflops;
}

/* h_mec_rsf_v9.cpp:1060-1075 */
void orc_reload(orc *o) {
    const int Ny1 = o->Ny1;
    for (int j = 0; j < o->Nx1; j++)
        for (int i = 0; i < Ny1; i++) {
            const int kp = (j * Ny1 + i) * 6;
            S_(o, HMEC_F_PT, i, j) = o->S[kp] * o->ptscale;
            S_(o, HMEC_F_VXS, i, j) = o->S[kp + 1];
            S_(o, HMEC_F_VYS, i, j) = o->S[kp + 2];
            S_(o, HMEC_F_PF, i, j) = o->S[kp + 3] * o->ptscale;
            S_(o, HMEC_F_VXD, i, j) = o->S[kp + 4];
            S_(o, HMEC_F_VYD, i, j) = o->S[kp + 5];
        }
}

/* ------------------------------------------ stress + RSF (a12, a13, a14) */
/* h_mec_rsf_v9.cpp:1077, 1102-1127, 1152-1308, 1344-1410 */
void orc_update(orc *o, hmec_iter_scalars *sc) {
    const hmec_params *P = &o->P;
    const int Ny = o->Ny, Nx = o->Nx, Ny1 = o->Ny1, Nx1 = o->Nx1;
    const double dt = o->dt, dx = P->dx, dy = P->dy;
    memset(sc, 0, sizeof(*sc));
    /* Vmax = VSLIPB.maxCoeff() before this iteration's RSF (:1077) */
    double vmax = -INFINITY;
    for (int k = 0; k < Ny * Nx; k++) vmax = dmax(vmax, o->F[HMEC_F_VSLIPB][k]);
    sc->Vmax = vmax;
    o->Vmax = vmax;
    /* stress, :1102-1127 */
    for (int i = 0; i < Ny; i++)
        for (int j = 0; j < Nx; j++) {
            B(o, HMEC_F_EXY, i, j) = .5 * ((S_(o, HMEC_F_VXS, i + 1, j) - S_(o, HMEC_F_VXS, i, j)) / dy +
                                          (S_(o, HMEC_F_VYS, i, j + 1) - S_(o, HMEC_F_VYS, i, j)) / dx);
            const double KXY = orc_divplus(dt * B(o, HMEC_F_GGG, i, j), B(o, HMEC_F_ETA, i, j));
            B(o, HMEC_F_SXY, i, j) = 2 * B(o, HMEC_F_ETA, i, j) * B(o, HMEC_F_EXY, i, j) * KXY + B(o, HMEC_F_SXY0, i, j) * (1 - KXY);
        }
    for (int i = 1; i < Ny; i++)
        for (int j = 1; j < Nx; j++) {
            const double dvx = S_(o, HMEC_F_VXS, i, j) - S_(o, HMEC_F_VXS, i, j - 1);
            const double dvy = S_(o, HMEC_F_VYS, i, j) - S_(o, HMEC_F_VYS, i - 1, j);
            S_(o, HMEC_F_EXX, i, j) = (2 * dvx / dx - dvy / dy) / 3.;
            S_(o, HMEC_F_EYY, i, j) = (2 * dvy / dy - dvx / dx) / 3.;
            const double KXX = orc_divplus(dt * S_(o, HMEC_F_GGGP, i, j), S_(o, HMEC_F_ETAP, i, j));
            S_(o, HMEC_F_SXX, i, j) = 2 * S_(o, HMEC_F_ETAP, i, j) * S_(o, HMEC_F_EXX, i, j) * KXX + S_(o, HMEC_F_SXX0, i, j) * (1 - KXX);
            S_(o, HMEC_F_SYY, i, j) = 2 * S_(o, HMEC_F_ETAP, i, j) * S_(o, HMEC_F_EYY, i, j) * KXX + S_(o, HMEC_F_SYY0, i, j) * (1 - KXX);
        }
    /* RSF, :1154-1308 */
    memcpy(o->F[HMEC_F_ETA5], o->F[HMEC_F_ETA0], sizeof(double) * Ny * Nx);
    memset(o->F[HMEC_F_DSY], 0, sizeof(double) * Ny * Nx);
    memset(o->F[HMEC_F_YNY], 0, sizeof(double) * Ny * Nx);
    memset(o->F[HMEC_F_SIGMAY], 0, sizeof(double) * Ny * Nx);
    memset(o->F[HMEC_F_SII_FAULT], 0, sizeof(double) * Ny * Nx);
    int ynpl = 0;
    double ddd = 0;
    double dtlapusta = 1e7;
    memcpy(o->F[HMEC_F_OM5], o->F[HMEC_F_OM], sizeof(double) * Ny * Nx);
    const double V0 = P->V0, BETASOLID = P->BETASOLID;
    int capped = 0;
    if (o->timestep > P->tyield) {
        const int i = P->line_fault;
        for (int j = 0; j < Nx; j++) {
            const double arsf = B(o, HMEC_F_ARSF, i, j), brsf = B(o, HMEC_F_BRSF, i, j), lrsf = B(o, HMEC_F_LRSF, i, j);
            const double fric = B(o, HMEC_F_FRIC, i, j), eta0 = B(o, HMEC_F_ETA0, i, j);
            const double sxx = blk_sum_s(o, HMEC_F_SXX, i, j), syy = blk_sum_s(o, HMEC_F_SYY, i, j);
            const double G = B(o, HMEC_F_GGG, i, j), ETA = B(o, HMEC_F_ETA, i, j), OMij = B(o, HMEC_F_OM, i, j);
            B(o, HMEC_F_SIIB, i, j) = sqrt(SQ(B(o, HMEC_F_SXY, i, j)) + (SQ(sxx) + SQ(syy) + SQ(sxx + syy)) / 32.);
            const double SIIBij = B(o, HMEC_F_SIIB, i, j);
            const double siiel = SIIBij / (ETA / (G * dt + ETA));
            const double prB_temp = (blk_sum_s(o, HMEC_F_PT, i, j) - blk_sum_s(o, HMEC_F_PF, i, j)) / 4.;
            double prB = prB_temp;
            if (prB < 1e3) prB = 1e3;
            double SIIB1 = SIIBij;
            double EIISLIP = V0 * sinh(dmax(SIIB1, 0.) / arsf / prB) * exp(-(brsf * OMij + fric) / arsf) / dx;
            double ETAPL = SIIB1 / 2. / EIISLIP;
            double ETAVP = 1. / (1. / eta0 + 1. / ETAPL);
            double SIIB2 = siiel * ETAVP / (G * dt + ETAVP);
            const double DSIIB1 = SIIB2 - SIIB1;
            double V = 2 * V0 * sinh(dmax(SIIB2, 0.) / arsf / prB) * exp(-(brsf * OMij + fric) / arsf);
            EIISLIP = V / dx / 2.;
            ETAPL = SIIB2 / 2. / EIISLIP;
            ETAVP = 1. / (1. / eta0 + 1. / ETAPL);
            const double DSIIB2 = siiel * ETAVP / (G * dt + ETAVP) - SIIB2;
            double SIIB4 = 0.;
            if ((DSIIB1 >= 0 && DSIIB2 <= 0) || (DSIIB1 <= 0 && DSIIB2 >= 0)) {
                double DSIIB = 1e9;
                int it = 0;
                while (fabs(DSIIB) > 1e-3) {
                    if (it++ >= P->bisect_cap) { capped++; break; }
                    SIIB4 = (SIIB1 + SIIB2) / 2.;
                    V = 2 * V0 * sinh(dmax((SIIB4), 0.) / arsf / prB) * exp(-(brsf * OMij + fric) / arsf);
                    EIISLIP = V / dx / 2.;
                    ETAPL = SIIB4 / 2. / EIISLIP;
                    ETAVP = 1. / (1. / eta0 + 1. / ETAPL);
                    DSIIB = siiel * ETAVP / (G * dt + ETAVP) - SIIB4;
                    if ((DSIIB >= 0 && DSIIB1 >= 0) || (DSIIB <= 0 && DSIIB1 <= 0)) SIIB1 = SIIB4;
                    else SIIB2 = SIIB4;
                }
            }
            if (V * dt / lrsf > 1e-6)
                B(o, HMEC_F_OM5, i, j) = log(V0 / V + (exp(B(o, HMEC_F_OM0, i, j)) - V0 / V) * exp(-V * dt / lrsf));
            else
                B(o, HMEC_F_OM5, i, j) = log(exp(B(o, HMEC_F_OM0, i, j)) * (1 - V * dt / lrsf) + V0 * dt / lrsf);
            const double syield = dmax(P->syieldmin, prB_temp * arsf * asinh(V / 2. / V0 * exp((brsf * B(o, HMEC_F_OM5, i, j) + fric) / arsf)));
            double etapl = eta0 * syield / (eta0 * V + syield);
            B(o, HMEC_F_SIGMAY, i, j) = syield;
            B(o, HMEC_F_VSLIPB, i, j) = V;
            B(o, HMEC_F_SII_FAULT, i, j) = SIIB4;
            const double g_temp = 2 * G;
            const double k = g_temp / (M_PI * (1 - ((3. / BETASOLID - g_temp) / (6. / BETASOLID + g_temp))) * dx);
            double dTETAmax;
            if ((SQ((k * lrsf / prB - brsf) / arsf - 1) / 4. - k * lrsf / arsf / prB) < 0)
                dTETAmax = dmin(1. - (brsf - arsf) * prB / (k * lrsf), .2);
            else
                dTETAmax = dmin(arsf * prB / (k * lrsf - (brsf - arsf) * prB), .2);
            dtlapusta = dmin(dtlapusta, dTETAmax * lrsf / V);
            int ynn = 0;
            if (B(o, HMEC_F_YNY0, i, j) > 0) {
                ynn = 1;
                B(o, HMEC_F_DSY, i, j) = SIIBij - syield;
                ddd += SQ(B(o, HMEC_F_DSY, i, j));
                ynpl++;
            }
            const double A = syield / siiel;
            if (A < 1) {
                etapl = dt * G * A / (1 - A);
                if (etapl < eta0) {
                    B(o, HMEC_F_ETA5, i, j) = pow(etapl, 1 - P->etawt) * pow(ETA, P->etawt);
                    B(o, HMEC_F_YNY, i, j) = 1;
                    if (!ynn) {
                        B(o, HMEC_F_DSY, i, j) = SIIBij - syield;
                        ddd += SQ(B(o, HMEC_F_DSY, i, j));
                        ynpl++;
                    }
                } else {
                    B(o, HMEC_F_ETA5, i, j) = eta0;
                }
            } else {
                B(o, HMEC_F_ETA5, i, j) = eta0;
            }
        }
    }
    o->bisect_capped += capped;
    sc->ynpl = ynpl;
    sc->ddd = ddd;
    sc->dtlapusta = dtlapusta;
    sc->bisect_capped = capped;
    /* reductions for the dt controller, :1344-1410 */
    double vxmax = -INFINITY, vxmin = INFINITY, vymax = -INFINITY, vymin = INFINITY;
    for (int k = 0; k < Ny1 * Nx1; k++) {
        vxmax = dmax(vxmax, o->F[HMEC_F_VXS][k]);
        vxmin = dmin(vxmin, o->F[HMEC_F_VXS][k]);
        vymax = dmax(vymax, o->F[HMEC_F_VYS][k]);
        vymin = dmin(vymin, o->F[HMEC_F_VYS][k]);
    }
    sc->vxs_max = vxmax; sc->vxs_min = vxmin; sc->vys_max = vymax; sc->vys_min = vymin;
    double mbx = 0, mby = 0;
    for (int i = 0; i < Ny1; i++)
        for (int j = 0; j < Nx1; j++) {
            if (o->yvx[i] >= P->upper_block && o->yvx[i] <= P->lower_block) mbx = dmax(mbx, fabs(S_(o, HMEC_F_VXS, i, j)));
            if (o->yvy[i] >= P->upper_block && o->yvy[i] <= P->lower_block) mby = dmax(mby, fabs(S_(o, HMEC_F_VYS, i, j)));
        }
    sc->maxvxs_band = mbx;
    sc->maxvys_band = mby;
    double dtslip = 1e30;
    for (int i = 0; i < Ny; i++)
        for (int j = 0; j < Nx; j++)
            if (B(o, HMEC_F_VSLIPB, i, j) > 0) dtslip = dmin(dtslip, dx * P->stpmax / B(o, HMEC_F_VSLIPB, i, j));
    sc->dtslip = dtslip;
}

/* h_mec_rsf_v9.cpp:1315-1317, 1440-1448 */
void orc_iter_commit(orc *o, int ynpl, int converged) {
    const int nb = o->Ny * o->Nx;
    if (ynpl == 0) copyf(o, HMEC_F_ETA, HMEC_F_ETA0);
    if (!converged) {
        for (int k = 0; k < nb; k++)
            o->F[HMEC_F_ETA][k] = dmax(dmin(o->F[HMEC_F_ETA5][k], o->F[HMEC_F_ETA0][k]), o->P.etamin);
        copyf(o, HMEC_F_ETA50, HMEC_F_ETA);
        copyf(o, HMEC_F_YNY0, HMEC_F_YNY);
        copyf(o, HMEC_F_OM, HMEC_F_OM5);
    }
}

/* ------------------------------------------------------ advection */
/* h_mec_rsf_v9.cpp:1561-1656 (enabled by params.advect_markers) */
void orc_advect(orc *o) {
    const hmec_params *P = &o->P;
    const int Nx = o->Nx, Ny = o->Ny;
    const double dt = o->dt, dx = P->dx, dy = P->dy;
    double *xm = o->mk[HMEC_M_XM], *ym = o->mk[HMEC_M_YM];
    double *sxxm = o->mk[HMEC_M_SXXM], *syym = o->mk[HMEC_M_SYYM], *sxym = o->mk[HMEC_M_SXYM];
    for (int64_t m = 0; m < o->M; m++) {
        double vxm[4] = {0, 0, 0, 0}, vym[4] = {0, 0, 0, 0}, spm[4] = {0, 0, 0, 0};
        const double xold = xm[m], yold = ym[m];
        for (int rk = 0; rk < 4; rk++) {
            int j = orc_check_bounds(orc_fix(xm[m] / dx), Nx);
            int i = orc_check_bounds(orc_fix((ym[m] + dy / 2.) / dy), Ny);
            double dxm = (xm[m] - o->xvx[j]) / dx, dym = (ym[m] - o->yvx[i]) / dy;
            vxm[rk] = S_(o, HMEC_F_VXS, i, j) * (1 - dxm) * (1 - dym) + S_(o, HMEC_F_VXS, i + 1, j) * (1 - dxm) * dym +
                      S_(o, HMEC_F_VXS, i, j + 1) * dxm * (1 - dym) + S_(o, HMEC_F_VXS, i + 1, j + 1) * dxm * dym;
            j = orc_check_bounds(orc_fix((xm[m] + dx / 2.) / dx), Nx);
            i = orc_check_bounds(orc_fix(ym[m] / dy), Ny);
            dxm = (xm[m] - o->xvy[j]) / dx; dym = (ym[m] - o->yvy[i]) / dy;
            vym[rk] = S_(o, HMEC_F_VYS, i, j) * (1 - dxm) * (1 - dym) + S_(o, HMEC_F_VYS, i + 1, j) * (1 - dxm) * dym +
                      S_(o, HMEC_F_VYS, i, j + 1) * dxm * (1 - dym) + S_(o, HMEC_F_VYS, i + 1, j + 1) * dxm * dym;
            j = orc_check_bounds(orc_fix((xm[m]) / dx), Nx);
            i = orc_check_bounds(orc_fix((ym[m]) / dy), Ny);
            dxm = (xm[m] - o->x[j]) / dx; dym = (ym[m] - o->y[i]) / dy;
            spm[rk] = B(o, HMEC_F_ESP, i, j) * (1 - dxm) * (1 - dym) + B(o, HMEC_F_ESP, i + 1, j) * (1 - dxm) * dym +
                      B(o, HMEC_F_ESP, i, j + 1) * dxm * (1 - dym) + B(o, HMEC_F_ESP, i + 1, j + 1) * dxm * dym;
            if (rk < 2) {
                xm[m] = xold + vxm[rk] * dt / 2.;
                ym[m] = yold + vym[rk] * dt / 2.;
            } else if (rk == 3) {
                xm[m] = xold + vxm[rk] * dt;
                ym[m] = yold + vym[rk] * dt;
            }
        }
        const double dspeff = spm[0] * dt;
        const double msxxold = sxxm[m], msyyold = syym[m], msxyold = sxym[m];
        sxym[m] = .5 * (msxxold - msyyold) * sin(2 * dspeff) + msxyold * cos(2 * dspeff);
        sxxm[m] = msxxold * SQ(cos(dspeff)) + msyyold * SQ(sin(dspeff)) - msxyold * sin(2 * dspeff);
        syym[m] = msxxold * SQ(sin(dspeff)) + msyyold * SQ(cos(dspeff)) + msxyold * sin(2 * dspeff);
        xm[m] = xold + (vxm[0] + 2 * vxm[1] + 2 * vxm[2] + vxm[3]) / 6. * dt;
        ym[m] = yold + (vym[0] + 2 * vym[1] + 2 * vym[2] + vym[3]) / 6. * dt;
        if (xm[m] < 0) xm[m] = xm[m] + P->xsize;
        if (xm[m] > P->xsize) xm[m] = xm[m] - P->xsize;
    }
}

/* square_block of h_mec_rsf_v9.cpp:134-136: the sum of squares of a 2x2
 * block in the order (0,0), (0,1), (1,0), (1,1) */
static double sqblock(double a00, double a01, double a10, double a11) {
    return ((a00 * a00 + a01 * a01) + a10 * a10) + a11 * a11;
}

/* final stresses with the final dt, h_mec_rsf_v9.cpp:1480-1521, and the
 * output diagnostics of the same loop, :1522-1558 */
static void final_stress(orc *o) {
    const hmec_params *P = &o->P;
    const int Ny = o->Ny, Nx = o->Nx;
    const double dt = o->dt, dx = P->dx, dy = P->dy;
    for (int i = 0; i < Ny; i++)
        for (int j = 0; j < Nx; j++) {
            const double dvydx = (S_(o, HMEC_F_VYS, i, j + 1) - S_(o, HMEC_F_VYS, i, j)) / dx;
            const double dvxdy = (S_(o, HMEC_F_VXS, i + 1, j) - S_(o, HMEC_F_VXS, i, j)) / dy;
            B(o, HMEC_F_ESP, i, j) = .5 * (dvydx - dvxdy);
            B(o, HMEC_F_EXY, i, j) = .5 * (dvxdy + dvydx);
            const double KXY = orc_divplus(dt * B(o, HMEC_F_GGG, i, j), B(o, HMEC_F_ETA, i, j));
            B(o, HMEC_F_SXY, i, j) = 2 * B(o, HMEC_F_ETA, i, j) * B(o, HMEC_F_EXY, i, j) * KXY + B(o, HMEC_F_SXY0, i, j) * (1 - KXY);
        }
    for (int i = 1; i < Ny; i++)
        for (int j = 1; j < Nx; j++) {
            const double dvx = S_(o, HMEC_F_VXS, i, j) - S_(o, HMEC_F_VXS, i, j - 1);
            const double dvy = S_(o, HMEC_F_VYS, i, j) - S_(o, HMEC_F_VYS, i - 1, j);
            S_(o, HMEC_F_EXX, i, j) = (2 * dvx / dx - dvy / dy) / 3.;
            S_(o, HMEC_F_EYY, i, j) = (2 * dvy / dy - dvx / dx) / 3.;
            const double KXX = orc_divplus(dt * S_(o, HMEC_F_GGGP, i, j), S_(o, HMEC_F_ETAP, i, j));
            S_(o, HMEC_F_SXX, i, j) = 2 * S_(o, HMEC_F_ETAP, i, j) * S_(o, HMEC_F_EXX, i, j) * KXX + S_(o, HMEC_F_SXX0, i, j) * (1 - KXX);
            S_(o, HMEC_F_SYY, i, j) = 2 * S_(o, HMEC_F_ETAP, i, j) * S_(o, HMEC_F_EYY, i, j) * KXX + S_(o, HMEC_F_SYY0, i, j) * (1 - KXX);
            /* :1522-1535 invariants and dissipation */
            const double sxx = S_(o, HMEC_F_SXX, i, j), syy = S_(o, HMEC_F_SYY, i, j);
            const double etap = S_(o, HMEC_F_ETAP, i, j), exx = S_(o, HMEC_F_EXX, i, j);
#define SXYq(a, b) (B(o, HMEC_F_SXY, a, b) / B(o, HMEC_F_ETA, a, b))
            S_(o, HMEC_F_EII, i, j) = sqrt(exx * exx + sqblock(B(o, HMEC_F_EXY, i - 1, j - 1), B(o, HMEC_F_EXY, i - 1, j),
                                                               B(o, HMEC_F_EXY, i, j - 1), B(o, HMEC_F_EXY, i, j)) / 4.);
            S_(o, HMEC_F_EIIVP, i, j) = sqrt(.5 * (SQ(sxx / (2 * etap)) + SQ(syy / (2 * etap))) +
                                             sqblock(SXYq(i - 1, j - 1), SXYq(i - 1, j), SXYq(i, j - 1), SXYq(i, j)) / 16.);
#undef SXYq
            S_(o, HMEC_F_SII, i, j) = sqrt(.5 * (SQ(sxx) + SQ(syy)) + sqblock(B(o, HMEC_F_SXY, i - 1, j - 1), B(o, HMEC_F_SXY, i - 1, j),
                                                                         B(o, HMEC_F_SXY, i, j - 1), B(o, HMEC_F_SXY, i, j)) / 4.);
            const double disxy = (SQ(B(o, HMEC_F_SXY, i, j)) / B(o, HMEC_F_ETA, i, j) + SQ(B(o, HMEC_F_SXY, i - 1, j)) / B(o, HMEC_F_ETA, i - 1, j) +
                                  SQ(B(o, HMEC_F_SXY, i, j - 1)) / B(o, HMEC_F_ETA, i, j - 1) + SQ(B(o, HMEC_F_SXY, i - 1, j - 1)) / B(o, HMEC_F_ETA, i - 1, j - 1)) / 2.;
            S_(o, HMEC_F_DIS, i, j) = SQ(sxx) / (2 * etap) + SQ(syy) / (2 * etap) + disxy;
            /* :1538-1557 averaged pressures and compaction */
            double pt0a, pf0a;
            if (i < Ny - 1) {
                S_(o, HMEC_F_PT_AVE, i, j) = (S_(o, HMEC_F_PT, i, j) + S_(o, HMEC_F_PT, i + 1, j)) / 2.;
                S_(o, HMEC_F_PF_AVE, i, j) = (S_(o, HMEC_F_PF, i, j) + S_(o, HMEC_F_PF, i + 1, j)) / 2.;
                pt0a = (S_(o, HMEC_F_PT0, i, j) + S_(o, HMEC_F_PT0, i + 1, j)) / 2.;
                pf0a = (S_(o, HMEC_F_PF0, i, j) + S_(o, HMEC_F_PF0, i + 1, j)) / 2.;
            } else {
                S_(o, HMEC_F_PT_AVE, i, j) = S_(o, HMEC_F_PT, i, j);
                S_(o, HMEC_F_PF_AVE, i, j) = S_(o, HMEC_F_PF, i, j);
                pt0a = S_(o, HMEC_F_PT0, i, j);
                pf0a = S_(o, HMEC_F_PF0, i, j);
            }
            const double pta = S_(o, HMEC_F_PT_AVE, i, j), pfa = S_(o, HMEC_F_PF_AVE, i, j);
            const double por = S_(o, HMEC_F_POR, i, j);
            S_(o, HMEC_F_VIS_COMP, i, j) = (pta - pfa) / (S_(o, HMEC_F_ETAB, i, j) * (1 - por));
            const double betadrained = (1 / S_(o, HMEC_F_GGGB, i, j) + P->BETASOLID) / (1 - por);
            const double kbw = 1 - P->BETASOLID / betadrained;
            S_(o, HMEC_F_EL_DECOM, i, j) = betadrained * (pta - pt0a - kbw * pfa + kbw * pf0a) / dt;
        }
}

/* h_mec_rsf_v9.cpp:1661-1690: the step's entry of the per-step series */
static void series_step(orc *o) {
    const int64_t t = o->timestep - 1;
    if (t < 0 || t >= NUM_TIMESTEPS) return;
    double *sr = o->series;
    sr[t] = o->timesum;
    sr[NUM_TIMESTEPS + t] = o->dt;
    double mx = -1e30, nx = 1e30, my = -1e30, ny = 1e30;
    for (int i = 0; i < o->Ny; i++)
        for (int j = 0; j < o->Nx; j++) {
            if (S_(o, HMEC_F_RHOX, i, j) > 2000 && i != 0) {
                mx = dmax(mx, S_(o, HMEC_F_VXS, i, j));
                nx = dmin(nx, S_(o, HMEC_F_VXS, i, j));
            }
            if (S_(o, HMEC_F_RHOY, i, j) > 2000 && j != 0) {
                my = dmax(my, S_(o, HMEC_F_VYS, i, j));
                ny = dmin(ny, S_(o, HMEC_F_VYS, i, j));
            }
        }
    sr[2 * NUM_TIMESTEPS + t] = mx;
    sr[3 * NUM_TIMESTEPS + t] = nx;
    sr[4 * NUM_TIMESTEPS + t] = my;
    sr[5 * NUM_TIMESTEPS + t] = ny;
}

double *orc_series(orc *o) { return o->series; }

/* h_mec_rsf_v9.cpp:1658-1726 */
static void history(orc *o) {
    const int ns = o->Ny1 * o->Nx1;
    o->timesum += o->dt;
    series_step(o);
    copyf(o, HMEC_F_VX0, HMEC_F_VXS);
    copyf(o, HMEC_F_VY0, HMEC_F_VYS);
    for (int k = 0; k < ns; k++) {
        if (o->F[HMEC_F_PORX][k] > 0) o->F[HMEC_F_VXF0][k] = o->F[HMEC_F_VXS][k] + o->F[HMEC_F_VXD][k] / o->F[HMEC_F_PORX][k];
        if (o->F[HMEC_F_PORY][k] > 0) o->F[HMEC_F_VYF0][k] = o->F[HMEC_F_VYS][k] + o->F[HMEC_F_VYD][k] / o->F[HMEC_F_PORY][k];
    }
    copyf(o, HMEC_F_SXX0, HMEC_F_SXX);
    copyf(o, HMEC_F_SYY0, HMEC_F_SYY);
    copyf(o, HMEC_F_SXY0, HMEC_F_SXY);
    for (int k = 0; k < ns; k++) o->F[HMEC_F_PTF0][k] = o->F[HMEC_F_PT][k] - o->F[HMEC_F_PF][k];
    copyf(o, HMEC_F_PT0, HMEC_F_PT);
    copyf(o, HMEC_F_PF0, HMEC_F_PF);
}

/* h_mec_rsf_v9.cpp:1463-1521 + advection + :1658-1726 */
void orc_step_end(orc *o) {
    if (o->dt00 > o->dt) o->yndtdecrease = 1;
    copyf(o, HMEC_F_ETA00, HMEC_F_ETA50);
    copyf(o, HMEC_F_OM0, HMEC_F_OM);
    o->dt = dmin(dmin(o->dtx, o->dty), dmin(o->dt, o->dtlapusta));
    final_stress(o);
    if (o->P.advect_markers) orc_advect(o);
    history(o);
}

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* Picard / dt controller, h_mec_rsf_v9.cpp:1310-1456.  Returns 1 when the
 * exit test fires. */
static int controller(orc *o, const hmec_iter_scalars *sc) {
    const hmec_params *P = &o->P;
    const int iterstep = o->iterstep, ynpl = sc->ynpl;
    const double dx = P->dx, dy = P->dy;
    int yn = 0;
    o->DSYLSQ[iterstep] = 0;
    if (ynpl > 0) o->DSYLSQ[iterstep] = sqrt(sc->ddd / ynpl);
    const double D_iter = o->DSYLSQ[iterstep];
    double D_iter_quot = 0;
    if (iterstep != 0) D_iter_quot = D_iter / o->DSYLSQ[iterstep - 1];
    double dtpl = o->dt;
    if (ynpl > 0 && iterstep < P->niterglobal && o->ynlast >= P->dtstep &&
        (o->ynlast > P->ynlastmax || log10(D_iter_quot) >= 0 ||
         log10(D_iter_quot) > log10(P->errmin / D_iter) / (P->ynlastmax - o->ynlast))) {
        dtpl = o->dt / P->dtkoef;
        yn = 1;
    }
    double maxvxy0 = 0;
    if (iterstep > 0) maxvxy0 = o->maxvxy;
    double maxvxs = dmax(sc->vxs_max, fabs(sc->vxs_min));
    double maxvys = dmax(sc->vys_max, fabs(sc->vys_min));
    o->maxvxy = sqrt(SQ(sc->vxs_max - sc->vxs_min) + SQ(sc->vys_max - sc->vys_min));
    double stpmaxcur = P->stpmax1;
    o->dtx = o->dt;
    if (o->dt > dx * stpmaxcur / maxvxs) { o->dtx = dx / P->dtkoefv * stpmaxcur / maxvxs; yn = 1; }
    o->dty = o->dt;
    if (o->dt > dy * stpmaxcur / maxvys) { o->dty = dy / P->dtkoefv * stpmaxcur / maxvys; yn = 1; }
    maxvxs = sc->maxvxs_band;
    maxvys = sc->maxvys_band;
    stpmaxcur = P->stpmax;
    if (o->dt > dx * stpmaxcur / maxvxs) { o->dtx = dx / P->dtkoefv * stpmaxcur / maxvxs; yn = 1; }
    if (o->dt > dy * stpmaxcur / maxvys) { o->dty = dy / P->dtkoefv * stpmaxcur / maxvys; yn = 1; }
    double dtslip = sc->dtslip;
    if (ynpl > 0 && dtslip < o->dt) { yn = 1; dtslip = dtslip / P->dtkoefv; }
    if (yn && o->dt > P->dtmin) {
        const double dtold = o->dt;
        o->dt = dmax(dmin(dmin(o->dtx, o->dty), dmin(dmin(dtpl, dtslip), o->dtlapusta)), P->dtmin);
        if (o->dt < dtold) o->ynlast = 0;
    } else {
        yn = 0;
    }
    int ynstop = 0;
    double vratio = 0;
    if (iterstep > 0) vratio = log10(o->maxvxy / maxvxy0);
    if (!yn && (ynpl == 0 || (o->DSYLSQ[iterstep] < P->errmin && iterstep > 0 && fabs(vratio) < P->vratiomax)))
        ynstop = 1;
    return ynstop;
}

/* One full timestep, h_mec_rsf_v9.cpp:376-1726 (outputs excluded). */
int orc_step(orc *o, hmec_step_info *info) {
    const hmec_params *P = &o->P;
    double t0 = now_ms();
    orc_scatter(o);
    double t1 = now_ms();
    o->t_scatter = t1 - t0;
    o->t_assemble = o->t_solve = o->t_update = 0;
    orc_step_begin(o);
    o->bisect_capped = 0;
    int err = 0;
    int nsolve = 0;
    for (o->iterstep = 0; o->iterstep < P->niterglobal; o->iterstep++) {
        double a = now_ms();
        orc_vep_pre(o);
        orc_scale(o);
        orc_assemble(o);
        double b = now_ms();
        err = orc_solve(o);
        nsolve++;
        double c = now_ms();
        if (err) break;
        orc_reload(o);
        hmec_iter_scalars sc;
        orc_update(o, &sc);
        o->dtlapusta = sc.dtlapusta;
        int stop = controller(o, &sc);
        orc_iter_commit(o, sc.ynpl, stop);
        double d = now_ms();
        o->t_assemble += b - a;
        o->t_solve += c - b;
        o->t_update += d - c;
        if (stop) break;
        o->ynlast++;
    }
    /* a loop that runs out leaves iterstep == niterglobal (:1434-1456), so
     * EVO_data.txt and the iteration count read niterglobal + 1 */
    double e = now_ms();
    orc_step_end(o);
    double f = now_ms();
    o->t_update += f - e;
    if (info) {
        memset(info, 0, sizeof(*info));
        info->timestep = o->timestep;
        info->iterations = o->iterstep + 1;
        info->ynlast = o->ynlast;
        info->yndtdecrease = o->yndtdecrease;
        info->dt = o->dt; info->dt00 = o->dt00; info->dtx = o->dtx; info->dty = o->dty;
        info->dtlapusta = o->dtlapusta; info->Vmax = o->Vmax; info->maxvxy = o->maxvxy;
        info->timesum = o->timesum; info->ptscale = o->ptscale;
        info->bisect_capped = o->bisect_capped;
        info->factorizations = nsolve;
        info->solves = nsolve;
        info->ms_scatter = o->t_scatter;
        info->ms_assemble = o->t_assemble;
        info->ms_factor = o->t_solve;
        info->ms_update = o->t_update;
        info->ms_total = f - t0;
    }
    o->timestep++;
    return err;
}

/* the controller of h_mec_rsf_v9.cpp:1310-1436 on given scalars at Picard
 * iteration iterstep (its pins drive it with constructed scalars); returns 1
 * when the exit test fires */
int orc_controller(orc *o, const hmec_iter_scalars *sc, int iterstep) {
    o->iterstep = iterstep;
    return controller(o, sc);
}
double *orc_dsylsq(orc *o) { return o->DSYLSQ; }

/* scalar state access (dt etc.) */
void orc_get_state(orc *o, hmec_step_info *info) {
    memset(info, 0, sizeof(*info));
    info->timestep = o->timestep;
    info->iterations = o->iterstep + 1;
    info->ynlast = o->ynlast;
    info->yndtdecrease = o->yndtdecrease;
    info->dt = o->dt; info->dt00 = o->dt00; info->dtx = o->dtx; info->dty = o->dty;
    info->dtlapusta = o->dtlapusta; info->Vmax = o->Vmax; info->maxvxy = o->maxvxy;
    info->timesum = o->timesum; info->ptscale = o->ptscale;
}
void orc_set_state(orc *o, const hmec_step_info *info) {
    o->timestep = info->timestep;
    o->ynlast = info->ynlast;
    o->yndtdecrease = info->yndtdecrease;
    o->dt = info->dt; o->dt00 = info->dt00; o->dtx = info->dtx; o->dty = info->dty;
    o->dtlapusta = info->dtlapusta; o->Vmax = info->Vmax; o->maxvxy = info->maxvxy;
    o->timesum = info->timesum; o->ptscale = info->ptscale;
}
void orc_set_dt(orc *o, double dt) { o->dt = dt; }
double orc_get_dt(orc *o) { return o->dt; }
int orc_sizeof_params(void) { return (int)sizeof(hmec_params); }
int orc_sizeof_step_info(void) { return (int)sizeof(hmec_step_info); }
int orc_sizeof_iter_scalars(void) { return (int)sizeof(hmec_iter_scalars); }
