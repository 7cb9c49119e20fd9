"""Serving form: batched same-size factorizations and the ensemble server.

A batched factorization must give every member exactly the factor it gets
alone (same kernels, blockIdx.z selects the member), and a member served in
a mix with other members (admission by footprint, overlapping streams,
batched factorizations) must end bit-identical to the same request served
alone.
"""
import numpy as np
import pytest

from conftest import hm

FIELDS = ("vxs", "vys", "pt", "pf", "vxD", "vyD", "SXX", "SYY", "SXY", "ETA", "OM", "VSLIPB")


def _assembled(p, seed, prec):
    c = hm.Context(p, hm.generate_markers(p, cfg=1, seed=seed), precision=prec)
    c.analyze()
    c.scatter()
    c.step_begin()
    dt = c.state().dt
    c.vep_pre(dt)
    c.assemble(dt)
    return c


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [hm.PREC_FP64, hm.PREC_FP32])
def test_batch_factorize_bitexact(gpu, prec):
    p = hm.params(1)
    pb = hm.params(1)
    pb.brsfm[1] = 0.014          # a different member of the same grid
    ctxs = [_assembled(p, 3, prec), _assembled(p, 4, prec), _assembled(pb, 5, prec)]
    ref = []
    for c in ctxs:
        c.factorize(prec)
        st = c.solve()
        assert st.rel_residual < (1e-12 if prec == hm.PREC_FP32 else 1e-13)
        ref.append(c.solution().copy())
    hm.batch_factorize(ctxs, prec)
    for c, r in zip(ctxs, ref):
        c.solve()
        assert np.array_equal(c.solution(), r)
    for c in ctxs:
        c.close()


def _serve_alone(prec, p, nsteps, seed):
    s = hm.Server(precision=prec, max_active=1, keep_fields=True)
    i = s.submit(p, nsteps, cfg=1, seed=seed)
    s.run()
    r = s.result(i)
    assert r.status == 0 and r.steps == nsteps
    out = {f: s.get(i, f) for f in FIELDS}
    s.close()
    return out, r


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [hm.PREC_FP64, hm.PREC_FP32])
def test_server_alone_vs_mix(gpu, prec):
    nsteps = 6
    p1 = hm.params(1)
    alone, r_alone = _serve_alone(prec, p1, nsteps, 11)
    # the mix: two more members of the same grid (other seed / friction), one
    # 201x51 member; the budget admits at most two 101x51 members at a time
    pb = hm.params(1)
    pb.arsfm[1] = 0.009
    pb.brsfm[1] = 0.015
    p5 = hm.params(5)
    f1 = hm.member_footprint(p1, prec)
    f5 = hm.member_footprint(p5, prec)
    budget = int(max(2.5 * f1, f5 + 0.2 * f1))
    # a window long enough that the first requester of a factorization always
    # waits for its running same-size partner: batching then does not depend
    # on how the members' steps happen to line up
    s = hm.Server(precision=prec, max_active=3, hbm_budget=budget, batch_window_ms=200.0,
                  keep_fields=True)
    ix = s.submit(p1, nsteps, cfg=1, seed=11, arrival_ms=0.0)
    s.submit(pb, nsteps, cfg=1, seed=12, arrival_ms=0.0)
    s.submit(p1, nsteps, cfg=1, seed=13, arrival_ms=2.0)
    s.submit(p5, 3, cfg=5, seed=14, arrival_ms=4.0)
    s.run()
    st = s.stats()
    assert st.members_done == 4 and st.members_failed == 0
    assert st.peak_bytes <= budget
    assert st.max_concurrent <= 3
    assert st.member_steps == 3 * nsteps + 3
    rx = s.result(ix)
    assert rx.status == 0 and rx.steps == nsteps
    for f in FIELDS:
        assert np.array_equal(s.get(ix, f), alone[f]), f
    assert rx.timesum == r_alone.timesum and rx.dt == r_alone.dt
    if prec == hm.PREC_FP64:
        # every Picard iteration refactors: same-size members meet in batches
        assert st.factor_batches >= 1 and st.factor_batch_members >= 2
    lat = s.latencies(ix)
    assert len(lat) == nsteps and np.all(lat > 0)
    assert 0 < st.p50_ms <= st.p99_ms
    s.close()
