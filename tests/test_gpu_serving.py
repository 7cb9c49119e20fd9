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
    # rendezvous without a deadline: the first requester of a factorization
    # waits until its running same-size partner asks too (or finishes), so
    # the batching does not depend on how the members' host threads line up
    s = hm.Server(precision=prec, max_active=3, hbm_budget=budget, batch_window_ms=-1.0,
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
    # the two 101x51 members admitted together factor their first matrices in
    # one batch (under FP64 every Picard iteration refactors and they meet
    # again)
    assert st.factor_batches >= 1 and st.factor_batch_members >= 2
    lat = s.latencies(ix)
    assert len(lat) == nsteps and np.all(lat > 0)
    assert 0 < st.p50_ms <= st.p99_ms
    s.close()


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_serving_mix_64_members(gpu):
    """The serving workload at its full width with a short horizon: a Poisson
    stream of 64 members with grids drawn from {101x51, 201x51, 401x101,
    801x201}, a, b, L and permeability drawn per member, 3 timesteps each,
    16 slots, FP32 policy, and an HBM budget that holds only part of the
    mix at once.  Every member finishes, admission never exceeds the budget,
    and one member's final state is bit-identical to the same request served
    alone."""
    rng = np.random.default_rng(5)
    grids = [(101, 51), (201, 51), (401, 101), (801, 201)]
    budget = int(40e9)
    s = hm.Server(precision=hm.PREC_FP32, max_active=16, hbm_budget=budget, batch_window_ms=2.0,
                  keep_fields=True)
    t = 0.0
    reqs = []
    for k in range(64):
        t += rng.exponential(5.0)
        nx, ny = grids[int(rng.integers(len(grids)))]
        a, b, L = rng.uniform(0.008, 0.012), rng.uniform(0.012, 0.020), rng.uniform(0.01, 0.05)
        p = hm.params(1, Nx=nx, Ny=ny, arsfm=(0.011, a), brsfm=(0.007, b), lrsfm=(0.02, L))
        reqs.append((s.submit(p, 3, cfg=5, seed=1000 + k, arrival_ms=t), p, 1000 + k))
    s.run()
    st = s.stats()
    assert st.members_done == 64 and st.members_failed == 0
    assert st.member_steps == 64 * 3
    assert st.peak_bytes <= budget
    sizes = {(r[1].Nx, r[1].Ny) for r in reqs}
    assert len(sizes) == 4
    # the first 401x101 member, served alone
    i, p, seed = next(r for r in reqs if (r[1].Nx, r[1].Ny) == (401, 101))
    s1 = hm.Server(precision=hm.PREC_FP32, max_active=1, keep_fields=True)
    j = s1.submit(p, 3, cfg=5, seed=seed)
    s1.run()
    for f in FIELDS:
        assert np.array_equal(s.get(i, f), s1.get(j, f)), f
    assert s.result(i).timesum == s1.result(j).timesum
    print("serving mix:", {"member_steps_per_s": st.member_steps_per_s, "p50_ms": st.p50_ms,
                           "p99_ms": st.p99_ms, "peak_GB": st.peak_bytes / 1e9,
                           "max_concurrent": st.max_concurrent, "batches": st.factor_batches},
          flush=True)
    s1.close()
    s.close()
