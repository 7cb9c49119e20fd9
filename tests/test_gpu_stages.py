"""The stage-level entry points driven from the host in the reference's
order (h_mec_rsf_v9.cpp:376-1726: scatter, step prologue, per Picard
iteration VEP pre-update, assembly, FP64 factorization and solve, update,
controller, epilogue, then the step's end) against hmec_step on the same
markers: every field, the step state and the per-step series must agree bit
for bit, over interseismic steps and over coseismic steps with many plastic
Picard iterations."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("vxs", "vys", "pt", "pf", "vxD", "vyD", "SXX", "SYY", "SXY", "VSLIPB", "ETA", "OM",
          "ETA50", "YNY0", "SigmaY", "SII_fault", "EII", "SII")
STATE = ("timestep", "dt", "dt00", "dtx", "dty", "dtlapusta", "timesum", "maxvxy", "ynlast",
         "yndtdecrease")


def stage_step(hmec, ctx, niter):
    """One timestep through the stage entry points; returns the Picard
    iteration count."""
    ctx.scatter()
    ctx.step_begin()
    st = ctx.state()
    it = 0
    for it in range(niter):
        ctx.vep_pre(st.dt)
        ptscale = ctx.assemble(st.dt)
        ctx.factorize(hmec.PREC_FP64)
        ctx.solve()
        sc = ctx.update(st.dt, ptscale)
        stop = ctx.controller(sc, it)
        ctx.iter_commit(sc.ynpl, stop)
        st = ctx.state()
        if stop:
            break
    else:
        it = niter                       # a loop that runs out, as in the reference
    # :1464-1472
    st = ctx.state()
    if st.dt00 > st.dt:
        st.yndtdecrease = 1
    st.dt = min(st.dtx, st.dty, st.dt, st.dtlapusta)
    ctx.set_state(st)
    ctx.step_end(st.dt)
    st = ctx.state()
    st.timestep += 1
    ctx.set_state(st)
    return it + 1


@pytest.mark.parametrize("coseismic", [False, True])
def test_stage_loop_vs_step(hmec, gpu, coseismic):
    p = hmec.params(1)
    soa = hmec.generate_markers(p, cfg=1, seed=1)
    a = hmec.Context(p, soa)
    b = hmec.Context(p, soa)
    if coseismic:
        for c in (a, b):
            c.init_coseismic(frac=0.98, noise=0.02, pf_ratio=0.7, seed=4)
    a.analyze()
    iters = []
    for s in range(4):
        n = stage_step(hmec, a, p.niterglobal)
        info = b.step()
        iters.append(info.iterations)
        assert n == info.iterations, (s, n, info.iterations)
        sa, sb = a.state(), b.state()
        for k in STATE:
            assert getattr(sa, k) == getattr(sb, k), (s, k, getattr(sa, k), getattr(sb, k))
        for f in FIELDS:
            assert np.array_equal(a.get(f), b.get(f)), (s, f)
    assert np.array_equal(a.series(), b.series())
    print("coseismic" if coseismic else "interseismic", iters, flush=True)
    if coseismic:
        assert max(iters) >= 3
