// Row emission of the coupled 6-unknown Stokes + Darcy system for one
// staggered node (i, j), in the reference's triplet order
// (h_mec_rsf_v9.cpp:687-980; the antiplane rows :981-1042 are off).
//
// The same code runs twice: on the host with a pattern emitter (row, col only,
// values ignored) to build the fixed CSR pattern and the triplet -> CSR-slot
// map once per grid, and on the device with a value emitter that writes the
// CSR values in place every Picard iteration.  Sharing it guarantees that the
// pattern and the values agree, including the duplicate triplets that
// Eigen's setFromTriplets sums (vys rows at i == Ny, j in {0, Nx}; vxD rows
// at i in {0, Ny}, j in {0, Nx-1}).
#pragma once

#ifndef HMEC_HD
#define HMEC_HD __host__ __device__
#endif

struct AsmIn {
    const double *ETA, *ETAP, *GGG, *GGGP, *SXY0, *SXX0, *SYY0, *RHOX, *RHO, *RHOY, *VX0, *VY0;
    const double *POR, *ETAB, *GGGB, *PT0, *PF0, *DILP, *RHOFX, *ETADX, *PORX, *VXF0, *RHOFY;
    const double *ETADY, *PORY, *VYF0;
    int Nx, Ny;
    double dt, dx, dy, ptscale, gx, gy, inertia, BETASOLID, BETAFLUID;
    double bcupper, bclower, bcvyflower, PTFDIFF;
};

HMEC_HD inline double asm_divplus(double a, double b) { return a / (a + b); }

// mask selects the row sections emitted (bit s: section 5a..5f, in that
// order: the vxs, vys, Pt, vxD, vyD and Pf rows of the node); each section
// adds only to its own row, so sections emitted apart give the same sums.
constexpr int kAsmSections = 6;
template <class E>
HMEC_HD inline void emit_node(int i, int j, const AsmIn &in, E &e, int mask = 0x3F) {
    const int Nx = in.Nx, Ny = in.Ny, Ny1 = Ny + 1, NV = 6;
    const int kp = (j * Ny1 + i) * NV, kx = kp + 1, ky = kp + 2, kpf = kp + 3, kxf = kp + 4, kyf = kp + 5;
    const double dt = in.dt, dx = in.dx, dy = in.dy;
    const double dx2 = dx * dx, dy2 = dy * dy, dx_dy = dx * dy;
    const double gx = in.gx, gy = in.gy, inertia = in.inertia;
    const double ptscale = in.ptscale, pfscale = in.ptscale;
#define Bq(a, ii, jj) (in.a[(ii) + (jj) * Ny])
#define Sq(a, ii, jj) (in.a[(ii) + (jj) * Ny1])
    if (mask & 0x01) {
        // 5a) vxs
        if (i == 0 || i == Ny || j == 0 || j >= Nx - 1) {
            if (j == Nx) e.add(kx, kx, 1);
            if (i == 0 && j < Nx) { e.add(kx, kx, 1); e.add(kx, kx + NV, 1); e.rhs(kx, 2 * in.bcupper); }
            if (i == Ny && j < Nx) { e.add(kx, kx, 1); e.add(kx, kx - NV, 1); e.rhs(kx, 2 * in.bclower); }
            if (j == 0 && i > 0 && i < Ny) { e.add(kx, kx, 1); e.add(kx, kx + NV * Ny1, -1); }
            if (j == Nx - 1 && i > 0 && i < Ny) { e.add(kx, kx, 1); e.add(kx, kx - NV * Ny1, -1); }
        } else {
            double ETAXY1 = Bq(ETA, i - 1, j), ETAXY2 = Bq(ETA, i, j);
            double ETAXX1 = Sq(ETAP, i, j), ETAXX2 = Sq(ETAP, i, j + 1);
            const double GXY1 = Bq(GGG, i - 1, j), GXY2 = Bq(GGG, i, j);
            const double GXX1 = Sq(GGGP, i, j), GXX2 = Sq(GGGP, i, j + 1);
            const double KXY1 = asm_divplus(dt * GXY1, ETAXY1), KXY2 = asm_divplus(dt * GXY2, ETAXY2);
            const double KXX1 = asm_divplus(dt * GXX1, ETAXX1), KXX2 = asm_divplus(dt * GXX2, ETAXX2);
            ETAXY1 *= KXY1; ETAXY2 *= KXY2; ETAXX1 *= KXX1; ETAXX2 *= KXX2;
            const double SXY1 = Bq(SXY0, i - 1, j) * (1 - KXY1), SXY2 = Bq(SXY0, i, j) * (1 - KXY2);
            const double SXX1 = Sq(SXX0, i, j) * (1 - KXX1), SXX2 = Sq(SXX0, i, j + 1) * (1 - KXX2);
            const double dRHOdx = (Sq(RHOX, i, j + 1) - Sq(RHOX, i, j - 1)) / (2. * dx);
            const double dRHOdy = (Bq(RHO, i, j) - Bq(RHO, i - 1, j)) / dy;
            const double temp = gx * dt * dRHOdy / 4.;
            e.add(kx, kx, -(ETAXX1 + ETAXX2) / dx2 - (ETAXY1 + ETAXY2) / dy2 - gx * dt * dRHOdx - inertia * Sq(RHOX, i, j) / dt);
            e.add(kx, kx - Ny1 * NV, ETAXX1 / dx2);
            e.add(kx, kx + Ny1 * NV, ETAXX2 / dx2);
            e.add(kx, kx - NV, ETAXY1 / dy2);
            e.add(kx, kx + NV, ETAXY2 / dy2);
            e.add(kx, ky - NV, (ETAXY1 - ETAXX1) / dx_dy - temp);
            e.add(kx, ky, (ETAXX1 - ETAXY2) / dx_dy - temp);
            e.add(kx, ky - NV + Ny1 * NV, (ETAXX2 - ETAXY1) / dx_dy - temp);
            e.add(kx, ky + Ny1 * NV, (ETAXY2 - ETAXX2) / dx_dy - temp);
            e.add(kx, kp, ptscale / dx);
            e.add(kx, kp + Ny1 * NV, -ptscale / dx);
            e.rhs(kx, -Sq(RHOX, i, j) * (inertia * Sq(VX0, i, j) / dt + gx) - (SXX2 - SXX1) / dx - (SXY2 - SXY1) / dy);
        }
    }
    if (mask & 0x02) {
        // 5b) vys
        if (j == 0 || j == Nx || i == 0 || i >= Ny - 1) {
            if (i == Ny) e.add(ky, ky, 1);
            if (j == 0) { e.add(ky, ky, 1); e.add(ky, ky + Ny1 * NV, 1); }
            if (j == Nx) { e.add(ky, ky, 1); e.add(ky, ky - Ny1 * NV, 1); }
            if (i == 0 && j > 0 && j < Nx) e.add(ky, ky, 1);
            if (i == Ny - 1 && j > 0 && j < Nx) e.add(ky, ky, 1);
        } else {
            double ETAXY1 = Bq(ETA, i, j - 1), ETAXY2 = Bq(ETA, i, j);
            double ETAYY1 = Sq(ETAP, i, j), ETAYY2 = Sq(ETAP, i + 1, j);
            const double GXY1 = Bq(GGG, i, j - 1), GXY2 = Bq(GGG, i, j);
            const double GYY1 = Sq(GGGP, i, j), GYY2 = Sq(GGGP, i + 1, j);
            const double KXY1 = asm_divplus(dt * GXY1, ETAXY1), KXY2 = asm_divplus(dt * GXY2, ETAXY2);
            const double KYY1 = asm_divplus(dt * GYY1, ETAYY1), KYY2 = asm_divplus(dt * GYY2, ETAYY2);
            ETAXY1 *= KXY1; ETAXY2 *= KXY2; ETAYY1 *= KYY1; ETAYY2 *= KYY2;
            const double SXY1 = Bq(SXY0, i, j - 1) * (1 - KXY1), SXY2 = Bq(SXY0, i, j) * (1 - KXY2);
            const double SYY1 = Sq(SYY0, i, j) * (1 - KYY1), SYY2 = Sq(SYY0, i + 1, j) * (1 - KYY2);
            const double dRHOdy = (Sq(RHOY, i + 1, j) - Sq(RHOY, i - 1, j)) / 2. / dy;
            const double dRHOdx = (Bq(RHO, i, j) - Bq(RHO, i, j - 1)) / dx;
            const double temp = gy * dt * dRHOdx / 4.;
            e.add(ky, ky, -(ETAYY1 + ETAYY2) / dy2 - (ETAXY1 + ETAXY2) / dx2 - gy * dt * dRHOdy - inertia * Sq(RHOY, i, j) / dt);
            e.add(ky, ky - Ny1 * NV, ETAXY1 / dx2);
            e.add(ky, ky + Ny1 * NV, ETAXY2 / dx2);
            e.add(ky, ky - NV, ETAYY1 / dy2);
            e.add(ky, ky + NV, ETAYY2 / dy2);
            e.add(ky, kx - Ny1 * NV, (ETAXY1 - ETAYY1) / dx_dy - temp);
            e.add(ky, kx + NV - Ny1 * NV, (ETAYY2 - ETAXY1) / dx_dy - temp);
            e.add(ky, kx, (ETAYY1 - ETAXY2) / dx_dy - temp);
            e.add(ky, kx + NV, (ETAXY2 - ETAYY2) / dx_dy - temp);
            e.add(ky, kp, ptscale / dy);
            e.add(ky, kp + NV, -ptscale / dy);
            e.rhs(ky, -Sq(RHOY, i, j) * (inertia * Sq(VY0, i, j) / dt + gy) - (SYY2 - SYY1) / dy - (SXY2 - SXY1) / dx);
        }
    }
    if (mask & 0x04) {
        // 5c) Pt
        if (i == 0 || j == 0 || i == Ny || j == Nx) {
            e.add(kp, kp, 1);
        } else {
            const double POR = Sq(POR, i, j), ETAB = Sq(ETAB, i, j);
            const double BETADRAINED = (1. / Sq(GGGB, i, j) + in.BETASOLID) / (1 - POR);
            const double KBW = 1 - in.BETASOLID / BETADRAINED;
            e.add(kp, kx - Ny1 * NV, -1. / dx);
            e.add(kp, kx, 1. / dx);
            e.add(kp, ky - NV, -1. / dy);
            e.add(kp, ky, 1. / dy);
            e.add(kp, kp, ptscale * (1. / ETAB / (1 - POR) + BETADRAINED / dt));
            e.add(kp, kpf, -pfscale * (1. / ETAB / (1 - POR) + BETADRAINED * KBW / dt));
            e.rhs(kp, BETADRAINED * (Sq(PT0, i, j) - KBW * Sq(PF0, i, j)) / dt + Sq(DILP, i, j));
        }
    }
    if (mask & 0x08) {
        // 5d) vxD
        if (i == 0 || i == Ny || j == 0 || j >= Nx - 1) {
            if (j == Nx) e.add(kxf, kxf, 1);
            if (i == 0 && j < Nx) { e.add(kxf, kxf, 1); e.add(kxf, kxf + NV, -1); }
            if (i == Ny && j < Nx) { e.add(kxf, kxf, 1); e.add(kxf, kxf - NV, -1); }
            if (j == 0) e.add(kxf, kxf, 1);
            if (j == Nx - 1) e.add(kxf, kxf, 1);
        } else {
            const double RHOFX = Sq(RHOFX, i, j);
            e.add(kxf, kxf, -Sq(ETADX, i, j) - RHOFX / Sq(PORX, i, j) * inertia / dt);
            e.add(kxf, kx, -RHOFX * inertia / dt);
            e.add(kxf, kpf, pfscale / dx);
            e.add(kxf, kpf + Ny1 * NV, -pfscale / dx);
            e.rhs(kxf, -RHOFX * (inertia * Sq(VXF0, i, j) / dt + gx));
        }
    }
    if (mask & 0x10) {
        // 5e) vyD
        if (j == 0 || j == Nx || i == 0 || i >= Ny - 1) {
            if (i == Ny) e.add(kyf, kyf, 1);
            if (j == 0 && i > 0 && i < Ny - 1) { e.add(kyf, kyf, 1); e.add(kyf, kyf + Ny1 * NV, -1); }
            if (j == Nx && i > 0 && i < Ny - 1) { e.add(kyf, kyf, 1); e.add(kyf, kyf - Ny1 * NV, -1); }
            if (i == 0) { e.add(kyf, kyf, 1); e.rhs(kyf, in.bcvyflower); }
            if (i == Ny - 1) { e.add(kyf, kyf, 1); e.rhs(kyf, in.bcvyflower); }
        } else {
            const double RHOFY = Sq(RHOFY, i, j);
            e.add(kyf, kyf, -Sq(ETADY, i, j) - RHOFY / Sq(PORY, i, j) * inertia / dt);
            e.add(kyf, ky, -RHOFY * inertia / dt);
            e.add(kyf, kpf, pfscale / dy);
            e.add(kyf, kpf + NV, -pfscale / dy);
            e.rhs(kyf, -RHOFY * (inertia * Sq(VYF0, i, j) / dt + gy));
        }
    }
    if (mask & 0x20) {
        // 5f) Pf
        if (j == 0 || j == Nx || i <= 1 || i >= Ny - 1) {
            if (i == 1 || i == Ny - 1) {
                e.add(kpf, kpf, pfscale);
                e.add(kpf, kp, -ptscale);
                e.rhs(kpf, -in.PTFDIFF);
            } else {
                e.add(kpf, kpf, 1);
            }
        } else {
            const double POR = Sq(POR, i, j), ETAB = Sq(ETAB, i, j);
            const double BETADRAINED = (1 / Sq(GGGB, i, j) + in.BETASOLID) / (1 - POR);
            const double KBW = 1 - in.BETASOLID / BETADRAINED;
            const double KSK = (BETADRAINED - in.BETASOLID) / (BETADRAINED - in.BETASOLID + POR * (in.BETAFLUID - in.BETASOLID));
            e.add(kpf, kxf - Ny1 * NV, -1. / dx);
            e.add(kpf, kxf, 1. / dx);
            e.add(kpf, kyf - NV, -1. / dy);
            e.add(kpf, kyf, 1. / dy);
            e.add(kpf, kp, -ptscale * (1 / ETAB / (1 - POR) + BETADRAINED * KBW / dt));
            e.add(kpf, kpf, pfscale * (1 / ETAB / (1 - POR) + BETADRAINED * KBW / KSK / dt));
            e.rhs(kpf, -BETADRAINED * KBW * (Sq(PT0, i, j) - 1. / KSK * Sq(PF0, i, j)) / dt - Sq(DILP, i, j));
        }
    }
#undef Bq
#undef Sq
}
