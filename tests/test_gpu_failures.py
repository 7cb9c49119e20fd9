"""Failure paths, forced from the host so that every kernel still exits
normally: unrepairable pivots (a zero row, a non-finite entry, a zero Darcy
diagonal) end in HMEC_E_PIVOT, solves whose residual stays above 1e-10 end in
HMEC_E_NOCONV instead of a silent result, the RSF bisection cap is counted
like the oracle's, and a server request that can never fit its HBM budget
ends with HMEC_E_OOM while the others are served."""
import numpy as np
import pytest

from conftest import hm

pytestmark = pytest.mark.gpu


def assembled(prec):
    p = hm.params(1)
    c = hm.Context(p, hm.generate_markers(p, cfg=1, seed=1), precision=prec)
    c.analyze()
    c.scatter()
    c.step_begin()
    dt = c.state().dt
    c.vep_pre(dt)
    c.assemble(dt)
    return p, c


def row_of(p, i, j, v):
    return ((j * (p.Ny + 1)) + i) * 6 + v


@pytest.mark.parametrize("prec", [hm.PREC_FP64, hm.PREC_FP32])
@pytest.mark.parametrize("kind", ["zero_row", "nan_entry", "inf_entry", "zero_darcy_diagonal"])
def test_unrepairable_pivot(gpu, prec, kind):
    p, c = assembled(prec)
    rp, ci, v, b = c.csr()
    v0, b0 = v.copy(), b.copy()
    if kind == "zero_darcy_diagonal":
        k = row_of(p, 10, 20, 4)                 # an interior vxD row
        d = [q for q in range(rp[k], rp[k + 1]) if ci[q] == k]
        v[d[0]] = 0.0
    else:
        k = row_of(p, 10, 20, 1)                 # an interior x-Stokes row
        if kind == "zero_row":
            v[rp[k]:rp[k + 1]] = 0.0
        else:
            v[rp[k] + 2] = np.nan if kind == "nan_entry" else np.inf
    c.set_csr_values(v, b)
    with pytest.raises(hm.HmecError) as e:
        c.factorize(prec)
    assert e.value.code == hm.E_PIVOT
    # the factor is not usable afterwards
    with pytest.raises(hm.HmecError) as e:
        c.solve()
    assert e.value.code == hm.E_ARG
    # a healthy matrix factors again on the same context
    c.set_csr_values(v0, b0)
    c.factorize(prec)
    assert c.solve().rel_residual <= 1e-12


def test_fp64_solve_nonfinite_rhs_is_noconv(gpu):
    """A right-hand side with a NaN: the factorization is fine, the solve's
    residual is not finite, and the FP64 solve reports HMEC_E_NOCONV rather
    than returning the result (the refinement does not reach 1e-10)."""
    p, c = assembled(hm.PREC_FP64)
    rp, ci, v, b = c.csr()
    b[row_of(p, 10, 20, 1)] = np.nan
    c.set_csr_values(v, b)
    c.factorize(hm.PREC_FP64)
    with pytest.raises(hm.HmecError) as e:
        c.solve()
    assert e.value.code == hm.E_NOCONV


@pytest.mark.parametrize("prec", [hm.PREC_FP64, hm.PREC_FP32])
def test_step_reports_failed_solves(gpu, prec):
    """Through hmec_step: a NaN in the velocity history enters only the
    right-hand side (inertia term), no solve meets the residual bound, and
    the step returns HMEC_E_NOCONV (under FP32 after a fresh FP32 factor and
    the FP64 fallback); a NaN viscosity enters the matrix and the step
    returns HMEC_E_PIVOT."""
    p = hm.params(1)
    soa = hm.generate_markers(p, cfg=1, seed=1)
    c = hm.Context(p, soa, precision=prec)
    c.step()
    vx0 = c.get("VX0")
    vx0[10, 20] = np.nan
    c.set("VX0", vx0)
    with pytest.raises(hm.HmecError) as e:
        c.step()
    assert e.value.code == hm.E_NOCONV
    c = hm.Context(p, soa, precision=prec)
    c.step()
    eta = c.get("ETA00")
    eta[5, 20] = np.nan                          # off the fault row
    c.set("ETA00", eta)
    with pytest.raises(hm.HmecError) as e:
        c.step()
    assert e.value.code == hm.E_PIVOT


def test_bisection_cap_counted(hmec, orc, gpu):
    """With the RSF stress bisection capped at one iteration every bracketed
    fault node hits the cap.  On a fault pre-stressed to 0.98 of its
    frictional strength (the bisection brackets from the second step on) the
    GPU counts exactly the nodes the oracle counts on the first plastic
    steps, over the same Picard iterations, and up to bracket-test flips at
    equilibrium nodes afterwards.  A capped bisection never reaches the stress
    balance, so the Picard loop would run to niterglobal (10,000): it is
    bounded at 12 iterations here.  The unresolved plastic state amplifies
    rounding, so dt is held, like the other coseismic tests, to ten times the
    oracle's own solver noise (PARDISO against its banded LU, same run)."""
    from test_gpu_parity import prestress
    p = hmec.params(1, bisect_cap=1, niterglobal=12)
    soa = hmec.generate_markers(p, cfg=1, seed=1)
    ctx = hmec.Context(p, soa)
    o = orc.Oracle(p, soa, hmec)
    ob = orc.Oracle(p, soa, hmec, use_pardiso=False)
    for x in (ctx, o, ob):
        prestress(x, p)
    total = 0
    for s in range(3):
        a = ctx.step()
        b = o.step()
        c = ob.step()
        assert a.iterations == b.iterations == c.iterations, s
        if s < 2:
            assert a.bisect_capped == b.bisect_capped, (s, a.bisect_capped, b.bisect_capped)
        else:
            # deep in the plastic state a fault node at equilibrium flips the
            # bracket test (DSIIB1, DSIIB2 of opposite sign) on a last-bit
            # difference: at most Nx / 20 such nodes per Picard iteration
            flips = abs(a.bisect_capped - b.bisect_capped)
            assert flips <= max(abs(c.bisect_capped - b.bisect_capped), a.iterations * (p.Nx // 20)), \
                (s, a.bisect_capped, b.bisect_capped, c.bisect_capped)
        floor = abs(c.dt - b.dt) / b.dt
        assert abs(a.dt - b.dt) <= max(1e-9, 10 * floor) * b.dt, (s, a.dt, b.dt, floor)
        total += a.bisect_capped
    assert total > 0


def test_server_over_budget_request(gpu):
    """A request whose footprint exceeds the whole HBM budget can never be
    admitted: it ends with HMEC_E_OOM (no step run) and does not block the
    queue; the other members are served within the budget."""
    p1 = hm.params(1)
    p3 = hm.params(3)
    f1 = hm.member_footprint(p1, hm.PREC_FP32)
    f3 = hm.member_footprint(p3, hm.PREC_FP32)
    budget = int(2.5 * f1)
    assert f3 > budget
    s = hm.Server(precision=hm.PREC_FP32, max_active=2, hbm_budget=budget, keep_fields=False)
    a = s.submit(p1, 2, cfg=1, seed=1, arrival_ms=0.0)
    big = s.submit(p3, 2, cfg=3, seed=3, arrival_ms=0.0)
    c = s.submit(p1, 2, cfg=1, seed=2, arrival_ms=1.0)
    s.run()
    rb = s.result(big)
    assert rb.status == hm.E_OOM and rb.steps == 0
    for i in (a, c):
        r = s.result(i)
        assert r.status == 0 and r.steps == 2
    st = s.stats()
    assert st.members_done == 2 and st.members_failed == 1
    assert st.peak_bytes <= budget
    s.close()
