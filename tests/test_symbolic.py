"""Symbolic phase on the CPU: the library's fixed CSR pattern equals the
oracle's setFromTriplets result, and the nested-dissection elimination order
factors the equilibrated system without pivoting (checked with SuperLU in
that order, no row exchanges allowed)."""
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


def plan(hmec, p, order=True):
    L = hmec.lib()
    st = hmec.SolveStats()
    ar = C.c_int64()
    N = (p.Nx + 1) * (p.Ny + 1) * 6
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


@pytest.mark.parametrize("grid", [(101, 51), (31, 23)])
def test_nd_order_factors_without_pivoting(hmec, orc, grid):
    p = hmec.params(1, Nx=grid[0], Ny=grid[1])
    soa = hmec.generate_markers(p, cfg=1, seed=1)
    o = orc.Oracle(p, soa, hmec)
    o.scatter(); o.step_begin(); o.vep_pre(); o.scale(); o.assemble()
    rp, ci, v, b = o.csr()
    N = len(b)
    st, arena, order = plan(hmec, p)
    assert np.array_equal(np.sort(order), np.arange(N))
    A = sp.csr_matrix((v, ci, rp), shape=(N, N))
    r = 2.0 ** -np.ceil(np.log2(abs(A).max(axis=1).toarray().ravel()))
    A1 = sp.diags(r) @ A
    c = 2.0 ** -np.ceil(np.log2(abs(A1).max(axis=0).toarray().ravel()))
    Ap = (A1 @ sp.diags(c)).tocsc()[order][:, order].tocsc()
    lu = sla.splu(Ap, permc_spec="NATURAL", diag_pivot_thresh=0.0,
                  options=dict(SymmetricMode=True))
    assert np.array_equal(lu.perm_r, np.arange(N))      # no row exchanges happened
    y = lu.solve((r * b)[order])
    x = np.zeros(N)
    x[order] = y
    x *= c
    assert np.abs(A @ x - b).max() / np.abs(b).max() < 1e-13
    # the dense fronts hold at least the exact fill of this order
    assert st.factor_nnz >= lu.L.nnz + lu.U.nnz - N


def test_plan_stats_scale(hmec):
    s1, a1, _ = plan(hmec, hmec.params(1), order=False)
    s3, a3, _ = plan(hmec, hmec.params(3), order=False)
    assert s3.nfronts > s1.nfronts and s3.factor_flops > 50 * s1.factor_flops
    assert s3.max_front <= 6 * 3 * 203      # a separator column plus its two neighbours
