"""hmec_solve_noalloc, the allocation-free variant of hmec_solve, called
through the C ABI: on two calls in a row (a 201x51 system, then a smaller
101x51 one) it returns bit-identical solutions and statistics to hmec_solve
on the same factor, in both factor precisions; a system above its scratch
array's 1,048,576 unknowns is refused with HMEC_E_ARG."""
import ctypes as C

import numpy as np
import pytest

from conftest import hm

pytestmark = pytest.mark.gpu


def factored(cfg, prec):
    p = hm.params(cfg)
    c = hm.Context(p, hm.generate_markers(p, cfg=cfg, seed=cfg), precision=prec)
    c.analyze()
    c.scatter()
    c.step_begin()
    dt = c.state().dt
    c.vep_pre(dt)
    c.assemble(dt)
    c.factorize(prec)
    return c


def solve_noalloc(c):
    st = hm.SolveStats()
    rc = hm.lib().hmec_solve_noalloc(c.h, C.byref(st))
    return rc, st


@pytest.mark.parametrize("prec", [hm.PREC_FP64, hm.PREC_FP32])
def test_noalloc_solve_matches_solve(gpu, prec):
    ctxs = [factored(5, prec), factored(1, prec)]          # 201x51, then the smaller 101x51
    refs = []
    for c in ctxs:
        st = c.solve()
        refs.append((c.solution().copy(), st.ir_iterations, st.rel_residual))
    assert len(refs[0][0]) > len(refs[1][0])
    for c, (x, ir, res) in zip(ctxs, refs):
        c.set("vxs", np.zeros_like(c.get("vxs")))           # (fields are not the solution)
        rc, st = solve_noalloc(c)
        assert rc == 0
        assert np.array_equal(c.solution(), x)
        assert st.ir_iterations == ir and st.rel_residual == res
        assert res <= 1e-12
    for c in ctxs:
        c.close()


def test_noalloc_solve_size_limit(gpu):
    p = hm.params(1, Nx=1001, Ny=201)
    assert 6 * (p.Nx + 1) * (p.Ny + 1) > (1 << 20)
    c = hm.Context(p)
    rc, _ = solve_noalloc(c)
    assert rc == hm.E_ARG
    assert b"scratch" in hm.lib().hmec_last_error()
    c.close()
