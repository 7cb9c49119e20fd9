// Darcy condensation of the coupled Stokes + Darcy system (symbolic, host).
//
// The assembled rows of the Darcy velocities (h_mec_rsf_v9.cpp:899-902 vxD,
// :945-948 vyD, boundary rows :906-934, :952-960) hold the diagonal plus
// either the node's solid velocity and two fluid pressures, or (boundary
// rows) one other Darcy unknown, or nothing else.  The Darcy unknowns appear
// off the diagonal only in the fluid-pressure rows and in those boundary
// rows.  Eliminating them exactly gives x_D = g + E x_R with
//   interior d:          E[d,k] = -a_dk / a_dd,                 g_d = b_d / a_dd
//   boundary d -> n:     E[d,k] = -(a_dn / a_dd) E[n,k],         g_d = (b_d - a_dn g_n) / a_dd
// and the condensed system S x_R = b_R - A_RD g, S = A_RR + A_RD E, whose
// pattern stays within the coupled system's node graph.  This file builds
// the index structures; the values are formed on the device each
// factorization (and the vectors each solve).
#include <algorithm>
#include <vector>

#include "hmec_internal.h"

int hmec_condense_build(int64_t nnodes, const std::vector<int32_t> &rowp,
                        const std::vector<int32_t> &coli, CondPlan &cp) {
    const int64_t N = nnodes * 6;
    if ((int64_t)rowp.size() != N + 1) return HMEC_E_ARG;
    cp = CondPlan();
    cp.N = N;
    cp.NS = nnodes * 4;
    cp.ND = nnodes * 2;
    cp.s_of_full.assign(N, -1);
    cp.d_of_full.assign(N, -1);
    cp.full_of_s.assign(cp.NS, -1);
    cp.dglob.assign(cp.ND, -1);
    for (int64_t k = 0; k < N; k++) {
        const int64_t node = k / 6;
        const int v = (int)(k % 6);
        if (v < 4) {
            cp.s_of_full[k] = (int32_t)(node * 4 + v);
            cp.full_of_s[node * 4 + v] = (int32_t)k;
        } else {
            cp.d_of_full[k] = (int32_t)(node * 2 + (v - 4));
            cp.dglob[node * 2 + (v - 4)] = (int32_t)k;
        }
    }
    auto is_d = [&](int64_t k) { return cp.d_of_full[k] >= 0; };
    // Darcy rows
    const int64_t ND = cp.ND;
    cp.ddiag.assign(ND, -1);
    cp.dnb.assign(ND, -1);
    cp.dnbdiag.assign(ND, -1);
    cp.dnbe.assign(ND, -1);
    cp.des.assign(3 * ND, -1);
    cp.dec.assign(3 * ND, -1);
    std::vector<std::vector<int32_t>> own_r(ND);   // own non-Darcy entries (A slots)
    for (int64_t d = 0; d < ND; d++) {
        const int64_t r = cp.dglob[d];
        for (int p = rowp[r]; p < rowp[r + 1]; p++) {
            const int64_t c = coli[p];
            if (c == r) cp.ddiag[d] = p;
            else if (is_d(c)) {
                if (cp.dnb[d] >= 0) {
                    hmec_set_last_error("Darcy row couples to two other Darcy unknowns");
                    return HMEC_E_ARG;
                }
                cp.dnb[d] = (int32_t)c;
                cp.dnbe[d] = p;
            } else {
                own_r[d].push_back(p);
            }
        }
        if (cp.ddiag[d] < 0) {
            hmec_set_last_error("Darcy row without a diagonal entry");
            return HMEC_E_ARG;
        }
        if (cp.dnb[d] >= 0 && !own_r[d].empty()) {
            hmec_set_last_error("Darcy row couples to both a Darcy and a non-Darcy unknown");
            return HMEC_E_ARG;
        }
        if (own_r[d].size() > 3) {
            hmec_set_last_error("Darcy row with more than three non-Darcy entries");
            return HMEC_E_ARG;
        }
    }
    for (int64_t d = 0; d < ND; d++) {
        const std::vector<int32_t> *src = &own_r[d];
        if (cp.dnb[d] >= 0) {
            const int64_t n = cp.d_of_full[cp.dnb[d]];
            if (cp.dnb[n] >= 0) {
                hmec_set_last_error("Darcy boundary chain longer than one");
                return HMEC_E_ARG;
            }
            cp.dnbdiag[d] = cp.ddiag[n];
            src = &own_r[n];
        }
        for (size_t k = 0; k < src->size(); k++) {
            const int32_t p = (*src)[k];
            cp.des[3 * d + k] = p;
            cp.dec[3 * d + k] = cp.s_of_full[coli[p]];
        }
    }
    // condensed rows
    const int64_t NS = cp.NS;
    cp.srowp.assign(NS + 1, 0);
    cp.rptr.assign(NS + 1, 0);
    cp.tptr.assign(NS + 1, 0);
    std::vector<int32_t> cols;
    for (int64_t rs = 0; rs < NS; rs++) {
        const int64_t r = cp.full_of_s[rs];
        cols.clear();
        for (int p = rowp[r]; p < rowp[r + 1]; p++) {
            const int64_t c = coli[p];
            if (!is_d(c)) {
                cols.push_back(cp.s_of_full[c]);
            } else {
                const int64_t d = cp.d_of_full[c];
                for (int k = 0; k < 3; k++)
                    if (cp.dec[3 * d + k] >= 0) cols.push_back(cp.dec[3 * d + k]);
            }
        }
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        const int64_t base = cp.scoli.size();
        cp.scoli.insert(cp.scoli.end(), cols.begin(), cols.end());
        cp.ssrc.insert(cp.ssrc.end(), cols.size(), -1);
        auto slot_of = [&](int32_t col) {
            return (int32_t)(base + (std::lower_bound(cols.begin(), cols.end(), col) - cols.begin()));
        };
        for (int p = rowp[r]; p < rowp[r + 1]; p++) {
            const int64_t c = coli[p];
            if (!is_d(c)) {
                cp.ssrc[slot_of(cp.s_of_full[c])] = p;
            } else {
                const int64_t d = cp.d_of_full[c];
                cp.raslot.push_back(p);
                cp.rd.push_back((int32_t)d);
                for (int k = 0; k < 3; k++)
                    if (cp.dec[3 * d + k] >= 0) {
                        cp.taslot.push_back(p);
                        cp.tdk.push_back((int32_t)(3 * d + k));
                        cp.tsslot.push_back(slot_of(cp.dec[3 * d + k]));
                    }
            }
        }
        cp.srowp[rs + 1] = (int32_t)cp.scoli.size();
        cp.rptr[rs + 1] = (int32_t)cp.raslot.size();
        cp.tptr[rs + 1] = (int32_t)cp.taslot.size();
    }
    return HMEC_OK;
}
