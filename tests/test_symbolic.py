"""Symbolic phase on the CPU: the library's fixed CSR pattern equals the
oracle's setFromTriplets result; the Darcy condensation's pattern equals the
structure of the Schur complement S = A_RR - A_RD A_DD^-1 A_DR computed with
scipy from the oracle's matrix; and the nested-dissection elimination order
factors the equilibrated S without pivoting (checked with SuperLU in that
order, no row exchanges allowed), the full system's solution recovered by
back-substitution of the Darcy unknowns."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla


def lib_pattern(hmec, p):
    L = hmec.lib()
    n = C.c_int64()
    nnz = C.c_int64()
    assert L.hmec_pattern(C.byref(p), C.byref(n), C.byref(nnz), None, None) == 0
    rp = np.zeros(n.value + 1, np.int32)
    ci = np.zeros(nnz.value, np.int32)
    assert L.hmec_pattern(C.byref(p), C.byref(n), C.byref(nnz),
                          rp.ctypes.data_as(C.POINTER(C.c_int32)),
                          ci.ctypes.data_as(C.POINTER(C.c_int32))) == 0
    return rp, ci


def condensed_pattern(hmec, p):
    L = hmec.lib()
    n = C.c_int64()
    nnz = C.c_int64()
    assert L.hmec_condensed_pattern(C.byref(p), C.byref(n), C.byref(nnz), None, None) == 0
    rp = np.zeros(n.value + 1, np.int32)
    ci = np.zeros(nnz.value, np.int32)
    assert L.hmec_condensed_pattern(C.byref(p), C.byref(n), C.byref(nnz),
                                    rp.ctypes.data_as(C.POINTER(C.c_int32)),
                                    ci.ctypes.data_as(C.POINTER(C.c_int32))) == 0
    return rp, ci


def schur(A, b):
    """Darcy condensation with scipy: unknowns v = 4, 5 (vxD, vyD) of each node
    eliminated; condensed index node*4 + v for v < 4."""
    N = A.shape[0]
    v = np.arange(N) % 6
    D = np.where(v >= 4)[0]
    R = np.where(v < 4)[0]
    A = A.tocsr()
    ADD = A[D][:, D].tocsc()
    ADR = A[D][:, R].tocsc()
    ARD = A[R][:, D]
    E = -sla.spsolve(ADD, ADR)
    g = sla.spsolve(ADD, b[D])
    S = (A[R][:, R] + ARD @ sp.csr_matrix(E)).tocsr()
    bS = b[R] - ARD @ g
    return S, bS, E, g, R, D


def plan(hmec, p, order=True):
    L = hmec.lib()
    st = hmec.SolveStats()
    ar = C.c_int64()
    N = (p.Nx + 1) * (p.Ny + 1) * 4
    o = np.zeros(N, np.int32) if order else None
    assert L.hmec_plan_info(C.byref(p), C.byref(st), C.byref(ar),
                            o.ctypes.data_as(C.POINTER(C.c_int32)) if order else None) == 0
    return st, ar.value, o


@pytest.mark.parametrize("grid", [(13, 9), (101, 51), (30, 17)])
def test_pattern_equals_oracle(hmec, orc, grid):
    p = hmec.params(1, Nx=grid[0], Ny=grid[1])
    soa = hmec.generate_markers(p, cfg=1, seed=1)
    o = orc.Oracle(p, soa, hmec)
    o.scatter(); o.step_begin(); o.vep_pre(); o.scale(); o.assemble()
    rp, ci, _, _ = o.csr()
    rp2, ci2 = lib_pattern(hmec, p)
    assert np.array_equal(rp, rp2) and np.array_equal(ci, ci2)


@pytest.mark.parametrize("grid", [(13, 9), (101, 51)])
def test_condensed_pattern_equals_schur(hmec, orc, grid):
    p = hmec.params(1, Nx=grid[0], Ny=grid[1])
    soa = hmec.generate_markers(p, cfg=1, seed=1)
    o = orc.Oracle(p, soa, hmec)
    o.scatter(); o.step_begin(); o.vep_pre(); o.scale(); o.assemble()
    rp, ci, v, b = o.csr()
    N = len(b)
    # structure of A_RR + A_RD A_DD^-1 A_DR with boolean products (the
    # closure of A_DD's pattern gives A_DD^-1's)
    P = sp.csr_matrix((np.ones(len(ci)), ci, rp), shape=(N, N))
    vv = np.arange(N) % 6
    D = np.where(vv >= 4)[0]
    R = np.where(vv < 4)[0]
    PDD = (P[D][:, D] != 0).astype(float)
    inv = sp.identity(len(D), format="csr")
    while True:
        nxt = ((inv + inv @ PDD) != 0).astype(float)
        if (nxt != inv).nnz == 0:
            break
        inv = nxt
    want = ((P[R][:, R] + P[R][:, D] @ inv @ P[D][:, R]) != 0).astype(float).tocsr()
    want.sort_indices()
    rps, cis = condensed_pattern(hmec, p)
    assert np.array_equal(rps, want.indptr) and np.array_equal(cis, want.indices)


@pytest.mark.parametrize("grid", [(101, 51), (31, 23)])
def test_nd_order_factors_without_pivoting(hmec, orc, grid):
    p = hmec.params(1, Nx=grid[0], Ny=grid[1])
    soa = hmec.generate_markers(p, cfg=1, seed=1)
    o = orc.Oracle(p, soa, hmec)
    o.scatter(); o.step_begin(); o.vep_pre(); o.scale(); o.assemble()
    rp, ci, v, b = o.csr()
    N = len(b)
    A = sp.csr_matrix((v, ci, rp), shape=(N, N))
    S, bS, E, g, R, D = schur(A, b)
    NS = S.shape[0]
    st, arena, order = plan(hmec, p)
    assert np.array_equal(np.sort(order), np.arange(NS))
    r = 2.0 ** -np.ceil(np.log2(abs(S).max(axis=1).toarray().ravel()))
    S1 = sp.diags(r) @ S
    c = 2.0 ** -np.ceil(np.log2(abs(S1).max(axis=0).toarray().ravel()))
    Sp = (S1 @ sp.diags(c)).tocsc()[order][:, order].tocsc()
    lu = sla.splu(Sp, permc_spec="NATURAL", diag_pivot_thresh=0.0,
                  options=dict(SymmetricMode=True))
    assert np.array_equal(lu.perm_r, np.arange(NS))      # no row exchanges happened
    y = lu.solve((r * bS)[order])
    xR = np.zeros(NS)
    xR[order] = y
    xR *= c
    x = np.zeros(N)
    x[R] = xR
    x[D] = g + E @ xR                                   # back-substitution
    assert np.abs(A @ x - b).max() / np.abs(b).max() < 1e-13
    # the dense fronts hold at least the exact fill of this order
    assert st.factor_nnz >= lu.L.nnz + lu.U.nnz - NS


def test_plan_stats_scale(hmec):
    s1, a1, _ = plan(hmec, hmec.params(1), order=False)
    s3, a3, _ = plan(hmec, hmec.params(3), order=False)
    assert s3.nfronts > s1.nfronts and s3.factor_flops > 50 * s1.factor_flops
    assert s3.max_front <= 4 * 3 * 203      # a separator column plus its two neighbours
