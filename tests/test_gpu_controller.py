"""The device-side Picard / dt controller (h_mec_rsf_v9.cpp:1310-1456) through
hmec_controller, against the oracle's controller on the same scalars and
state: the branch cases the oracle's own pins construct, then random
sequences of Picard iterations over the ranges a run meets.  dt, dtx, dty,
maxvxy and ynlast must agree bit for bit (the controller's arithmetic is
divisions, products and square roots, correctly rounded on both sides), and
so must the exit decision."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("dt", "dtx", "dty", "maxvxy", "ynlast", "dtlapusta")


def scalars(hmec, ynpl=0, ddd=0.0, vx=(1e-10, -1e-10), vy=(1e-10, -1e-10), band=(1e-10, 1e-10),
            dtslip=1e30, dtlapusta=1e7):
    sc = hmec.IterScalars()
    sc.ynpl, sc.ddd, sc.dtlapusta = ynpl, ddd, dtlapusta
    sc.vxs_max, sc.vxs_min = vx
    sc.vys_max, sc.vys_min = vy
    sc.maxvxs_band, sc.maxvys_band = band
    sc.dtslip = dtslip
    return sc


@pytest.fixture(scope="module")
def pair(hmec, orc, gpu):
    p = hmec.params(1)
    soa = hmec.generate_markers(p, cfg=1, seed=1)
    return p, hmec.Context(p, soa), orc.Oracle(p, soa, hmec)


def set_both(ctx, o, dt, ynlast, maxvxy, dtx=None, dty=None):
    for obj in (ctx, o):
        st = obj.state()
        st.dt, st.ynlast, st.maxvxy = dt, ynlast, maxvxy
        st.dtx = dt if dtx is None else dtx
        st.dty = dt if dty is None else dty
        obj.set_state(st)


def run_both(ctx, o, sc, it):
    """One iteration on both; the oracle's loop counts ynlast on after the
    controller when it goes on (:1455), hmec_controller does it itself."""
    st = o.state()
    st.dtlapusta = sc.dtlapusta           # the update's Lapusta bound enters the state
    o.set_state(st)
    stop_o = o.controller(sc, it)
    if not stop_o:
        st = o.state()
        st.ynlast += 1
        o.set_state(st)
    stop_g = ctx.controller(sc, it)
    return stop_g, stop_o


def same_state(ctx, o, tag):
    a, b = ctx.state(), o.state()
    for k in KEYS:
        assert getattr(a, k) == getattr(b, k), (tag, k, getattr(a, k), getattr(b, k))


def test_branch_cases(hmec, pair):
    """The cases the oracle's controller pins construct: no yielding nodes,
    the global and band displacement limits, the slip, Lapusta and dtmin
    limits, the plastic reduction after dtstep iterations."""
    p, ctx, o = pair
    vmax = 1e-5
    cases = [
        (dict(), [scalars(hmec)]),
        (dict(ynlast=7), [scalars(hmec, vx=(vmax, -vmax / 2))]),
        (dict(), [scalars(hmec, vy=(vmax / 3, -vmax))]),
        (dict(), [scalars(hmec, vx=(vmax, 0), band=(2e-5, 1e-10))]),
        (dict(), [scalars(hmec, ynpl=2, ddd=8e8, dtslip=1e5)]),
        (dict(), [scalars(hmec, ynpl=0, dtslip=1e5)]),
        (dict(), [scalars(hmec, ynpl=1, ddd=1e6, dtslip=1e5, dtlapusta=3e4)]),
        (dict(), [scalars(hmec, ynpl=1, ddd=1e6, dtslip=1e5, dtlapusta=1e-9)]),
        (dict(ynlast=3), [scalars(hmec, ynpl=1, ddd=1e8)]),
        (dict(ynlast=p.ynlastmax + 1), [scalars(hmec, ynpl=1, ddd=1e8)]),
        (dict(ynlast=p.dtstep), [scalars(hmec, ynpl=1, ddd=1e8), scalars(hmec, ynpl=1, ddd=4e8)]),
        (dict(ynlast=p.dtstep), [scalars(hmec, ynpl=1, ddd=1.0), scalars(hmec, ynpl=1, ddd=4.0)]),
        # converged: error below errmin at the second iteration, velocity ratio small
        (dict(), [scalars(hmec, ynpl=1, ddd=1e-6), scalars(hmec, ynpl=1, ddd=1e-8)]),
    ]
    stops = []
    for n, (init, seq) in enumerate(cases):
        set_both(ctx, o, 1e7, init.get("ynlast", 0), 0.0)
        for it, sc in enumerate(seq):
            g, r = run_both(ctx, o, sc, it)
            assert g == r, (n, it)
            same_state(ctx, o, (n, it))
        stops.append(g)
    assert any(stops) and not all(stops)


def test_random_sequences(hmec, pair):
    """400 random Picard sequences of 1-6 iterations with scalars and starting
    states over the ranges of the seismic-cycle runs (velocities 1e-13 to
    1e-1 m/s, slip and Lapusta bounds around dt, errors around errmin)."""
    p, ctx, o = pair
    rng = np.random.default_rng(7)

    def lu(a, b):
        return float(10.0 ** rng.uniform(a, b))

    exits = 0
    for n in range(400):
        dt = lu(-3, np.log10(p.dtelastic0))
        set_both(ctx, o, dt, int(rng.integers(0, p.ynlastmax + 3)), lu(-13, -1))
        for it in range(int(rng.integers(1, 7))):
            ynpl = int(rng.choice([0, 1, 7, 300]))
            d = lu(np.log10(p.errmin) - 3, np.log10(p.errmin) + 4)
            vx = (lu(-13, -1), -lu(-13, -1))
            vy = (lu(-14, -2), -lu(-14, -2))
            band = (lu(-13, -1) if rng.random() < 0.8 else 0.0, lu(-14, -2))
            sc = scalars(hmec, ynpl=ynpl, ddd=ynpl * d * d, vx=vx, vy=vy, band=band,
                         dtslip=lu(-2, 9), dtlapusta=lu(-3, 8))
            g, r = run_both(ctx, o, sc, it)
            assert g == r, (n, it)
            same_state(ctx, o, (n, it))
            if g:
                exits += 1
                break
    assert exits > 20


def test_argument_checks(hmec, pair):
    p, ctx, o = pair
    with pytest.raises(hmec.HmecError):
        ctx.controller(scalars(hmec), -1)
    with pytest.raises(hmec.HmecError):
        ctx.controller(scalars(hmec), p.niterglobal)
