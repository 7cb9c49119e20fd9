// Picard / timestep controller of h_mec_rsf_v9.cpp:1310-1436, evaluated on
// the device once per Picard iteration (a single thread): the iteration error
// DSYLSQ, the plastic, displacement, slip and Lapusta timestep candidates, the
// choice of dt and the exit test.  The branch structure and the comparisons
// are the reference's (iteration counts and dt sequences must match it).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "hmec.h"

// controller state carried from one Picard iteration to the next
struct CtlState {
    double dt, dtx, dty, dtlapusta, maxvxy;
    int ynlast;
};

// std::max / std::min semantics (first argument on ties)
__host__ __device__ inline double ctl_max(double a, double b) { return (a < b) ? b : a; }
__host__ __device__ inline double ctl_min(double a, double b) { return (b < a) ? b : a; }

// One evaluation at Picard iteration iterstep with this iteration's scalars
// sc (sc.dtlapusta already in s.dtlapusta); updates s and DSYLSQ[iterstep];
// returns true when the exit test fires.
__host__ __device__ inline bool controller_eval(const hmec_params &P, int iterstep,
                                                const hmec_iter_scalars &sc, double *DSYLSQ,
                                                CtlState &s) {
    const int ynpl = sc.ynpl;
    const double dx = P.dx, dy = P.dy;
    bool yn = false;
    // :1311-1317
    DSYLSQ[iterstep] = 0;
    if (ynpl > 0) DSYLSQ[iterstep] = sqrt(sc.ddd / ynpl);
    const double D_iter = DSYLSQ[iterstep];
    double D_iter_quot = 0;
    if (iterstep != 0) D_iter_quot = D_iter / DSYLSQ[iterstep - 1];
    // :1329-1337 plastic dt reduction
    double dtpl = s.dt;
    if (ynpl > 0 && iterstep < P.niterglobal && s.ynlast >= P.dtstep &&
        (s.ynlast > P.ynlastmax || log10(D_iter_quot) >= 0 ||
         log10(D_iter_quot) > log10(P.errmin / D_iter) / (P.ynlastmax - s.ynlast))) {
        dtpl = s.dt / P.dtkoef;
        yn = true;
    }
    // :1339-1398 displacement limits, global then in the fault band
    double maxvxy0 = 0;
    if (iterstep > 0) maxvxy0 = s.maxvxy;
    double maxvxs = ctl_max(sc.vxs_max, fabs(sc.vxs_min));
    double maxvys = ctl_max(sc.vys_max, fabs(sc.vys_min));
    s.maxvxy = sqrt((sc.vxs_max - sc.vxs_min) * (sc.vxs_max - sc.vxs_min) +
                    (sc.vys_max - sc.vys_min) * (sc.vys_max - sc.vys_min));
    double stpmaxcur = P.stpmax1;
    s.dtx = s.dt;
    if (s.dt > dx * stpmaxcur / maxvxs) {
        s.dtx = dx / P.dtkoefv * stpmaxcur / maxvxs;
        yn = true;
    }
    s.dty = s.dt;
    if (s.dt > dy * stpmaxcur / maxvys) {
        s.dty = dy / P.dtkoefv * stpmaxcur / maxvys;
        yn = true;
    }
    maxvxs = sc.maxvxs_band;
    maxvys = sc.maxvys_band;
    stpmaxcur = P.stpmax;
    if (s.dt > dx * stpmaxcur / maxvxs) {
        s.dtx = dx / P.dtkoefv * stpmaxcur / maxvxs;
        yn = true;
    }
    if (s.dt > dy * stpmaxcur / maxvys) {
        s.dty = dy / P.dtkoefv * stpmaxcur / maxvys;
        yn = true;
    }
    // :1400-1415 slip limit
    double dtslip = sc.dtslip;
    if (ynpl > 0 && dtslip < s.dt) {
        yn = true;
        dtslip = dtslip / P.dtkoefv;
    }
    // :1418-1430 the minimal timestep
    if (yn && s.dt > P.dtmin) {
        const double dtold = s.dt;
        s.dt = ctl_max(ctl_min(ctl_min(s.dtx, s.dty), ctl_min(ctl_min(dtpl, dtslip), s.dtlapusta)),
                       P.dtmin);
        if (s.dt < dtold) s.ynlast = 0;
    } else {
        yn = false;
    }
    // :1432-1437 exit test
    double vratio = 0;
    if (iterstep > 0) vratio = log10(s.maxvxy / maxvxy0);
    return !yn && (ynpl == 0 || (DSYLSQ[iterstep] < P.errmin && iterstep > 0 &&
                                 fabs(vratio) < P.vratiomax));
}
