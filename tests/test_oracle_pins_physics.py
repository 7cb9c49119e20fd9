"""Pins of the oracle's physics blocks (oracle/hmec_oracle.c) by closed forms,
inverse relations and branch specifications on constructed states: the VEP
pre-update (cpp/h_mec_rsf_v9.cpp:612-669), the RSF yield stress, plastic
viscosity, Lapusta timestep and yielding-node bookkeeping (:1252-1308), the
Picard / dt controller (:1310-1456), the step prologue and iteration epilogue
(:586-608, :1440-1448), the y-Stokes, continuity, Darcy and boundary rows of
the assembly (:699-980), and the scatter's derived marker properties and
harmonic means (:389-423, :531-583).

Each expected value comes from a relation the reference's physics fixes
(Maxwell relaxation, the slip law read backwards, textbook poroelastic
coefficients, Darcy's law, the stress deviator's second invariant, equal
weight splits on the regular marker lattice), or from the branch rule as the
reference states it, so that a dropped term, a wrong sign or index, or a
swapped branch fails a check.  CPU only.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp


def small(hmec, Nx=13, Ny=9, **over):
    base = dict(xsize=6000., ysize=4000., Nx=Nx, Ny=Ny, upper_block=1800., lower_block=2200.,
                TS_1=1000., TS_2=2000., TS_3=4000., TS_4=5000.)
    base.update(over)
    return hmec.params(1, **base)


def fresh(hmec, orc, p, soa=None, seed=1, jitter=True):
    if soa is None:
        soa = hmec.generate_markers(p, cfg=1, seed=seed, jitter=jitter)
    o = orc.Oracle(p, soa, hmec)
    return o, soa


def full(p, name, v):
    shape = (p.Ny, p.Nx) if name in BASIC else (p.Ny + 1, p.Nx + 1)
    return np.full(shape, float(v))


BASIC = set()


@pytest.fixture(autouse=True)
def _basic_names(hmec):
    BASIC.update(hmec.BASIC_FIELDS)


def j2_sqrt(sxx, syy, sxy):
    """sqrt of the second invariant of the plane-strain stress deviator
    diag(sxx, syy, -(sxx + syy)) with shear sxy: 0.5 s:s."""
    s = np.array([[sxx, sxy, 0.], [sxy, syy, 0.], [0., 0., -(sxx + syy)]])
    return math.sqrt(0.5 * float(np.sum(s * s)))


# ======================================================== VEP pre-update
def test_vep_plastic_strain_rate_and_viscoplastic_viscosities(hmec, orc):
    """ETA is built as the series combination of ETA0 with a plastic viscosity
    eta_pl(i, j) linear in its inverse q = 1/eta_pl; vep_pre must give back
    IETAPLB = q, the plastic strain rate EIIB = (dy/w) SIIB q / 2 with SIIB the
    deviator invariant of the 2x2-averaged pressure-node stresses, and, at the
    four pressure nodes around the single node with YNY0 = 1, the series
    combinations ETAP = (1/ETAP0 + qbar)^-1, ETAB = (1/ETAB0 + (dy/w) qbar POR)^-1
    with qbar the value of q at the pressure node (the mean of the linear q
    over the four basic nodes around it); everywhere else ETAP = ETAP0,
    ETAB = ETAB0; GGGB = GGGP / POR.  The grid has dy != dx (= faultwidth), so
    the localization factor dy / w = 0.8 is not 1."""
    p = small(hmec, Ny=11)
    assert p.dy / p.faultwidth == pytest.approx(0.8)
    o, _ = fresh(hmec, orc, p)
    o.scatter()
    o.step_begin()
    Ny, Nx, Ny1, Nx1 = p.Ny, p.Nx, p.Ny + 1, p.Nx + 1
    eta0 = 1e21
    ii, jj = np.meshgrid(np.arange(Ny), np.arange(Nx), indexing="ij")
    q = (1 / 5e20) * (1 + 0.1 * ii + 0.05 * jj)
    eta = 1 / (1 / eta0 + q)
    eta[2, 7] = eta0                                   # one node that is not plastic
    o.set("ETA0", full(p, "ETA0", eta0))
    o.set("ETA", eta)
    s1, s2, t = 3e6, -1e6, 2e6
    o.set("SXX", full(p, "SXX", s1))
    o.set("SYY", full(p, "SYY", s2))
    o.set("SXY", full(p, "SXY", t))
    yny0 = np.zeros((Ny, Nx))
    i0, j0 = 4, 5
    yny0[i0, j0] = 1
    o.set("YNY0", yny0)
    etap0, etab0 = 2e21, 4e22
    rng = np.random.default_rng(7)
    por = rng.uniform(0.005, 0.05, (Ny1, Nx1))
    gggp = rng.uniform(1e10, 5e10, (Ny1, Nx1))
    o.set("ETAP0", full(p, "ETAP0", etap0))
    o.set("ETAB0", full(p, "ETAB0", etab0))
    o.set("POR", por)
    o.set("GGGP", gggp)
    o.vep_pre()
    w = p.dy / p.faultwidth
    siib = j2_sqrt(s1, s2, t)
    plastic = np.ones((Ny, Nx), bool)
    plastic[2, 7] = False
    got_q = o.get("IETAPLB")
    assert np.allclose(got_q[plastic], q[plastic], rtol=1e-12, atol=0)
    assert got_q[2, 7] == 0 and o.get("EIIB")[2, 7] == 0
    assert np.allclose(o.get("SIIB")[plastic], siib, rtol=1e-14, atol=0)
    assert np.allclose(o.get("EIIB")[plastic], w * siib * q[plastic] / 2, rtol=1e-12, atol=0)
    etap, etab, gggb = o.get("ETAP"), o.get("ETAB"), o.get("GGGB")
    affected = {(i0, j0), (i0 + 1, j0), (i0, j0 + 1), (i0 + 1, j0 + 1)}
    for i in range(1, Ny):
        for j in range(1, Nx):
            if (i, j) in affected:
                qbar = (1 / 5e20) * (1 + 0.1 * (i - 0.5) + 0.05 * (j - 0.5))
                assert etap[i, j] == pytest.approx(1 / (1 / etap0 + qbar), rel=1e-12)
                assert etab[i, j] == pytest.approx(1 / (1 / etab0 + w * qbar * por[i, j]), rel=1e-12)
            else:
                assert etap[i, j] == etap0 and etab[i, j] == etab0, (i, j)
            assert gggb[i, j] == pytest.approx(gggp[i, j] / por[i, j], rel=1e-15)


def test_vep_viscosity_floors(hmec, orc):
    """The limiting viscosity etamincur = dt * shearmod * 1e-4 floors ETA and
    ETAP, and floors the bulk viscosity through ETAB * POR."""
    p = small(hmec)
    o, _ = fresh(hmec, orc, p)
    o.scatter()
    o.step_begin()
    dt = o.dt
    floor = dt * p.shearmod * 1e-4
    Ny, Nx, Ny1, Nx1 = p.Ny, p.Nx, p.Ny + 1, p.Nx + 1
    eta = full(p, "ETA", 1e21)
    eta[3, 3] = floor * 1e-3
    o.set("ETA", eta)
    o.set("ETA0", full(p, "ETA0", 1e21))
    o.set("YNY0", np.zeros((Ny, Nx)))
    etap0 = full(p, "ETAP0", 1e21)
    etap0[5, 6] = floor * 0.5
    etab0 = full(p, "ETAB0", 1e23)
    etab0[2, 2] = floor * 10          # ETAB0 * POR = floor * 0.1 < floor
    o.set("ETAP0", etap0)
    o.set("ETAB0", etab0)
    o.set("POR", full(p, "POR", 0.01))
    o.vep_pre()
    assert o.get("ETA")[3, 3] == floor
    assert o.get("ETA")[3, 4] == 1e21
    assert o.get("ETAP")[5, 6] == floor
    assert o.get("ETAB")[2, 2] == pytest.approx(floor / 0.01, rel=1e-15)
    assert o.get("ETAB")[4, 4] == 1e23


# ================================================================ RSF
def rsf_state(hmec, orc, p, sxy_row, prB=6e7, a=0.011, b=0.017, L=0.02, yny0_row=None,
              eta=None, dt=1e7, om=0.0, seed=3):
    """A constructed fault-row state: uniform pressures (effective pressure prB),
    RSF parameters a, b, L, friction 0.6, state om, the given shear-stress
    history along the fault row, zero velocities and normal-stress deviators;
    RSF active (timestep 2 > tyield)."""
    o, _ = fresh(hmec, orc, p, seed=seed)
    o.scatter()
    Ny, Nx = p.Ny, p.Nx
    i = p.line_fault
    for name, v in (("ARSF", a), ("BRSF", b), ("LRSF", L), ("FRIC", 0.6), ("GGG", 3e10),
                    ("ETA0", 1e23), ("OM", om), ("OM0", om)):
        o.set(name, full(p, name, v))
    o.set("ETA", full(p, "ETA", 1e23 if eta is None else eta))
    o.set("pt", full(p, "pt", 1e8))
    o.set("pf", full(p, "pf", 1e8 - prB))
    # the update recomputes the stresses from the velocities and the stress
    # history first (:1102-1127): zero velocities leave SXY = SXY0 (1 - K)
    for name in ("SXX", "SYY", "SXX0", "SYY0", "vxs", "vys"):
        o.set(name, full(p, name, 0.))
    sxy = np.zeros((Ny, Nx))
    sxy[i, :] = sxy_row
    o.set("SXY0", sxy)
    yny0 = np.zeros((Ny, Nx))
    if yny0_row is not None:
        yny0[i, :] = yny0_row
    o.set("YNY0", yny0)
    o.set("VSLIPB", full(p, "VSLIPB", 0.))
    st = o.state()
    st.timestep = 2
    st.dt = dt
    o.set_state(st)
    return o


def fault_row(o, p, name):
    return o.get(name)[p.line_fault]


def test_rsf_yield_stress_inverts_the_slip_law(hmec, orc):
    """SigmaY = prB a asinh(V / 2V0 exp((b OM5 + mu0) / a)) is the stress at which
    the regularized slip law gives back the node's slip rate V at the updated
    state OM5: 2 V0 sinh(SigmaY / (a prB)) exp(-(b OM5 + mu0) / a) == V."""
    p = small(hmec)
    mu_sn = 0.6 * 6e7
    o = rsf_state(hmec, orc, p, np.linspace(0.3, 1.3, p.Nx) * mu_sn)
    o.update()
    V, sy, om5 = fault_row(o, p, "VSLIPB"), fault_row(o, p, "SigmaY"), fault_row(o, p, "OM5")
    a, b = 0.011, 0.017
    assert np.all(sy > p.syieldmin) and np.all(V > 0)
    back = 2 * p.V0 * np.sinh(sy / (a * 6e7)) * np.exp(-(b * om5 + 0.6) / a)
    assert np.allclose(back, V, rtol=1e-11, atol=0)
    # the state evolved (a dropped OM5 -> OM substitution fails the check above)
    assert np.abs(om5).max() > 1e-3


@pytest.mark.parametrize("etawt", [0.0, 0.5])
def test_rsf_plastic_viscosity_relaxes_to_yield(hmec, orc, etawt):
    """A node yields (YNY = 1) exactly when the Maxwell relaxation of its elastic
    trial stress siiel = SIIB (G dt + ETA) / ETA with the intact viscosity ETA0
    stays above the yield stress; the plastic viscosity eta_pl is the one whose
    relaxation lands on the yield stress, siiel eta_pl / (G dt + eta_pl) ==
    SigmaY, and ETA5 = eta_pl^(1-etawt) ETA^etawt; non-yielding nodes keep
    ETA5 = ETA0."""
    p = small(hmec, etawt=etawt)
    mu_sn = 0.6 * 6e7
    eta = 5e22
    o = rsf_state(hmec, orc, p, np.linspace(0.3, 1.3, p.Nx) * mu_sn, eta=eta)
    o.update()
    G, dt, eta0 = 3e10, 1e7, 1e23
    siib, sy = fault_row(o, p, "SIIB"), fault_row(o, p, "SigmaY")
    yny, eta5 = fault_row(o, p, "YNY"), fault_row(o, p, "ETA5")
    siiel = siib * (G * dt + eta) / eta
    above = siiel * eta0 / (G * dt + eta0) > sy
    assert np.array_equal(yny == 1, above)
    assert 0 < above.sum() < p.Nx                       # both branches are exercised
    eta_pl = eta5[above] ** (1 / (1 - etawt)) * eta ** (-etawt / (1 - etawt))
    assert np.allclose(siiel[above] * eta_pl / (G * dt + eta_pl), sy[above], rtol=1e-10, atol=0)
    assert np.all(eta5[~above] == eta0)
    # the rest of the grid keeps ETA5 = ETA0
    e5 = o.get("ETA5")
    rows = np.arange(p.Ny) != p.line_fault
    assert np.all(e5[rows] == eta0)


def test_rsf_yielding_node_bookkeeping(hmec, orc):
    """ynpl counts the fault nodes that yielded before (YNY0 > 0) or yield now
    (YNY = 1), each once; DSY = SIIB - SigmaY on exactly those nodes, 0
    elsewhere; ddd is the sum of DSY^2 over them."""
    p = small(hmec)
    mu_sn = 0.6 * 6e7
    Nx = p.Nx
    yny0 = np.zeros(Nx)
    yny0[[0, 1, 2, Nx - 3, Nx - 2]] = 1                  # low stress and high stress nodes
    o = rsf_state(hmec, orc, p, np.linspace(0.3, 1.3, Nx) * mu_sn, yny0_row=yny0, eta=5e22)
    sc = o.update()
    yny, dsy = fault_row(o, p, "YNY"), fault_row(o, p, "DSY")
    siib, sy = fault_row(o, p, "SIIB"), fault_row(o, p, "SigmaY")
    counted = (yny0 > 0) | (yny == 1)
    assert ((yny0 > 0) & (yny == 1)).any() and ((yny0 > 0) & (yny == 0)).any()
    assert sc.ynpl == counted.sum()
    assert np.array_equal(dsy[counted], siib[counted] - sy[counted])
    assert np.all(dsy[~counted] == 0)
    assert sc.ddd == pytest.approx(np.sum(dsy[counted] ** 2), rel=1e-14)
    assert not o.get("DSY")[np.arange(p.Ny) != p.line_fault].any()


def lapusta_expected(p, a, b, L, sn, G=3e10):
    """Lapusta et al. (2000) / Lapusta & Liu (2009) state-step bound with the
    stiffness of a mode-II element k = 2G / (pi (1 - nu) dx), nu from the bulk
    modulus 1/BETASOLID and G.  The reference's discriminant reads
    (kL/(a sn) - (b + a)/a)^2 / 4 - kL/(a sn) (Lapusta & Liu print (b - a));
    the oracle follows the reference."""
    K = 1 / p.BETASOLID
    nu = (3 * K - 2 * G) / (2 * (3 * K + G))
    k = 2 * G / (math.pi * (1 - nu) * p.dx)
    x = k * L / (a * sn)
    xi = (x - (b + a) / a) ** 2 / 4 - x
    if xi < 0:
        return min(1 - (b - a) * sn / (k * L), 0.2), k
    return min(a * sn / (k * L - (b - a) * sn), 0.2), k


@pytest.mark.parametrize("case", ["branch1", "branch2", "branch1_clipped_vs_lapusta"])
def test_rsf_lapusta_timestep(hmec, orc, case):
    """dtlapusta = min over the fault row of dTETAmax * L / V; with uniform
    a, b, L and effective pressure it is dTETAmax * L / max V.  The three
    parameter sets take the xi < 0 branch unclipped (0.12), the xi >= 0 branch
    (a sn / (kL - (b - a) sn) = 0.0109), and a set where the reference's
    discriminant and Lapusta & Liu's disagree (reference: 0.2, clipped;
    Lapusta & Liu: 0.125)."""
    p = small(hmec)
    a, L = 0.011, 0.02
    if case == "branch1":
        b, sn = 0.026, 6e7
    elif case == "branch2":
        b, sn = 0.017, 1e6
    else:
        b = 3 * a
        _, k = lapusta_expected(p, a, b, L, 1.0)
        sn = k * L / (10 * a)              # kL/(a sn) = 10
    want, k = lapusta_expected(p, a, b, L, sn)
    if case == "branch1":
        assert want == pytest.approx(1 - 0.015 * sn / (k * L)) and want < 0.2
    elif case == "branch2":
        assert want == pytest.approx(a * sn / (k * L - (b - a) * sn)) and want < 0.2
    else:
        assert want == 0.2
        lap = a * sn / (k * L - (b - a) * sn)
        assert lap == pytest.approx(0.125, rel=1e-9)
    # state -10 (the reference's fault value) makes the slip fast enough that
    # the bound is below its 1e7 s start value
    o = rsf_state(hmec, orc, p, np.linspace(0.5, 0.9, p.Nx) * 0.6 * sn, prB=sn, a=a, b=b, L=L,
                  om=-10.0)
    sc = o.update()
    V = fault_row(o, p, "VSLIPB")
    assert sc.dtlapusta < 1e7
    assert sc.dtlapusta * V.max() / L == pytest.approx(want, rel=1e-12)


# ============================================================ controller
def ctl(hmec, orc, dt=1e7, ynlast=0, maxvxy=0.0, dtlapusta=1e30, iterstep=0, dsylsq_prev=None):
    p = small(hmec)
    o, _ = fresh(hmec, orc, p)
    st = o.state()
    st.dt, st.ynlast, st.maxvxy, st.dtlapusta = dt, ynlast, maxvxy, dtlapusta
    o.set_state(st)
    return p, o


def scalars(hmec, ynpl=0, ddd=0.0, vx=(1e-10, -1e-10), vy=(1e-10, -1e-10), band=(1e-10, 1e-10),
            dtslip=1e30, dtlapusta=1e7):
    sc = hmec.IterScalars()
    sc.ynpl, sc.ddd, sc.dtlapusta = ynpl, ddd, dtlapusta
    sc.vxs_max, sc.vxs_min = vx
    sc.vys_max, sc.vys_min = vy
    sc.maxvxs_band, sc.maxvys_band = band
    sc.dtslip = dtslip
    return sc


def test_controller_no_plastic_nodes_exits_at_first_iteration(hmec, orc):
    p, o = ctl(hmec, orc)
    assert o.controller(scalars(hmec), 0) is True
    st = o.state()
    assert st.dt == 1e7 and st.dtx == 1e7 and st.dty == 1e7
    assert o.dsylsq(3)[0] == 0
    # maxvxy is the norm of the velocity ranges
    assert st.maxvxy == pytest.approx(math.hypot(2e-10, 2e-10), rel=1e-15)


def test_controller_displacement_limits(hmec, orc):
    """dt > dx stpmax1 / max|vx| sets dtx = dx / dtkoefv * stpmax1 / max|vx| and
    reduces dt to it (ynlast reset, no exit); the fault-band limit with stpmax
    then overwrites dtx whenever it applies, even when the global limit was
    the smaller one (h_mec_rsf_v9.cpp:1355-1398)."""
    p, o = ctl(hmec, orc, ynlast=7)
    vmax = 1e-5
    assert o.controller(scalars(hmec, vx=(vmax, -vmax / 2)), 0) is False
    st = o.state()
    want = p.dx / p.dtkoefv * p.stpmax1 / vmax
    assert st.dtx == pytest.approx(want, rel=1e-15) and st.dty == 1e7
    assert st.dt == st.dtx and st.ynlast == 0
    # y limit from vys
    p, o = ctl(hmec, orc)
    assert o.controller(scalars(hmec, vy=(vmax / 3, -vmax)), 0) is False
    st = o.state()
    assert st.dty == pytest.approx(p.dy / p.dtkoefv * p.stpmax1 / vmax, rel=1e-15)
    assert st.dt == st.dty and st.dtx == 1e7
    # band limit overwrites the global one
    p, o = ctl(hmec, orc)
    vb = 2e-5
    o.controller(scalars(hmec, vx=(vmax, 0), band=(vb, 1e-10)), 0)
    st = o.state()
    band = p.dx / p.dtkoefv * p.stpmax / vb
    glob = p.dx / p.dtkoefv * p.stpmax1 / vmax
    assert glob < band and st.dtx == pytest.approx(band, rel=1e-15)
    assert st.dt == st.dtx


def test_controller_slip_and_lapusta_limits(hmec, orc):
    """With yielding nodes, dtslip < dt reduces dt to dtslip / dtkoefv; the
    Lapusta bound enters the same minimum; dtmin is the lower limit."""
    p, o = ctl(hmec, orc)
    assert o.controller(scalars(hmec, ynpl=2, ddd=8e8, dtslip=1e5), 0) is False
    st = o.state()
    assert st.dt == pytest.approx(1e5 / p.dtkoefv, rel=1e-15)
    assert o.dsylsq(3)[0] == pytest.approx(math.sqrt(8e8 / 2), rel=1e-15)
    # dtslip without yielding nodes is ignored: exit
    p, o = ctl(hmec, orc)
    assert o.controller(scalars(hmec, ynpl=0, dtslip=1e5), 0) is True
    assert o.state().dt == 1e7
    # the Lapusta bound is the smallest candidate
    p, o = ctl(hmec, orc, dtlapusta=3e4)
    o.controller(scalars(hmec, ynpl=1, ddd=1e6, dtslip=1e5), 0)
    assert o.state().dt == 3e4
    # dtmin floor
    p, o = ctl(hmec, orc, dtlapusta=1e-9)
    o.controller(scalars(hmec, ynpl=1, ddd=1e6, dtslip=1e5), 0)
    assert o.state().dt == p.dtmin


def test_controller_plastic_dt_reduction_after_dtstep(hmec, orc):
    """dtpl = dt / dtkoef once ynlast >= dtstep and the iteration error stalls:
    ynlast > ynlastmax, or the error grew (log10 of the ratio >= 0), or it
    decreases too slowly to reach errmin within the remaining iterations."""
    dsy_prev = 1e4
    # ynlast below dtstep: no reduction, no exit (error above errmin)
    p, o = ctl(hmec, orc, ynlast=3)
    o.dsylsq(3)                                            # (state buffer exists)
    assert o.controller(scalars(hmec, ynpl=1, ddd=dsy_prev ** 2), 0) is False
    assert o.state().dt == 1e7
    # ynlast > ynlastmax
    p, o = ctl(hmec, orc, ynlast=p.ynlastmax + 1)
    assert o.controller(scalars(hmec, ynpl=1, ddd=1e8), 0) is False
    assert o.state().dt == pytest.approx(1e7 / p.dtkoef, rel=1e-15)
    assert o.state().ynlast == 0
    # error grows from iteration 0 to 1
    p, o = ctl(hmec, orc, ynlast=p.dtstep)
    o.controller(scalars(hmec, ynpl=1, ddd=dsy_prev ** 2), 0)      # DSYLSQ[0] = 1e4
    assert o.state().dt == 1e7
    st = o.state(); st.ynlast = p.dtstep; o.set_state(st)
    assert o.controller(scalars(hmec, ynpl=1, ddd=(2 * dsy_prev) ** 2), 1) is False
    assert o.state().dt == pytest.approx(1e7 / p.dtkoef, rel=1e-15)
    # the growth condition alone (the slow-convergence one is false here: the
    # error is below errmin and only one iteration remains before ynlastmax)
    p, o = ctl(hmec, orc, ynlast=p.dtstep)
    o.controller(scalars(hmec, ynpl=1, ddd=1.0), 0)                 # DSYLSQ[0] = 1
    st = o.state(); st.ynlast = p.ynlastmax - 1; o.set_state(st)
    assert not math.log10(2.0) > math.log10(p.errmin / 2.0) / 1
    assert o.controller(scalars(hmec, ynpl=1, ddd=4.0), 1) is False   # D = 2
    assert o.state().dt == pytest.approx(1e7 / p.dtkoef, rel=1e-15)
    # error decreases, but too slowly: log10(q) > log10(errmin / D) / (ynlastmax - ynlast)
    p, o = ctl(hmec, orc, ynlast=p.dtstep)
    o.controller(scalars(hmec, ynpl=1, ddd=dsy_prev ** 2), 0)
    st = o.state(); st.ynlast = p.ynlastmax - 1; o.set_state(st)
    d = 0.999 * dsy_prev                    # rate needed: log10(1e2 / d) / 1 ~ -2
    assert math.log10(d / dsy_prev) > math.log10(p.errmin / d) / 1
    assert o.controller(scalars(hmec, ynpl=1, ddd=d ** 2), 1) is False
    assert o.state().dt == pytest.approx(1e7 / p.dtkoef, rel=1e-15)
    # ... and fast enough: no reduction
    p, o = ctl(hmec, orc, ynlast=p.dtstep)
    o.controller(scalars(hmec, ynpl=1, ddd=dsy_prev ** 2), 0)
    st = o.state(); st.ynlast = p.dtstep; o.set_state(st)
    d = 0.5 * dsy_prev
    assert not math.log10(d / dsy_prev) > math.log10(p.errmin / d) / (p.ynlastmax - p.dtstep)
    o.controller(scalars(hmec, ynpl=1, ddd=d ** 2), 1)
    assert o.state().dt == 1e7


def test_controller_velocity_ratio_exit(hmec, orc):
    """With yielding nodes the loop exits only from the second iteration on,
    when the error is below errmin and the velocity range changed by less
    than vratiomax decades since the previous iteration."""
    v = (3e-10, -1e-10)
    m0 = math.hypot(v[0] - v[1], 2e-10)
    p, o = ctl(hmec, orc)
    assert o.controller(scalars(hmec, ynpl=1, ddd=1.0, vx=v), 0) is False     # iteration 0
    for f, want in ((1 + 1e-4, True), (1.01, False), (1 / 1.01, False)):
        p, o = ctl(hmec, orc, maxvxy=m0)
        vx = (v[0] * f, v[1] * f)
        vy = (1e-10 * f, -1e-10 * f)
        assert o.controller(scalars(hmec, ynpl=1, ddd=1.0, vx=vx, vy=vy), 1) is want, f
        assert o.state().dt == 1e7
    # error above errmin: no exit even with a steady velocity
    p, o = ctl(hmec, orc, maxvxy=m0)
    assert o.controller(scalars(hmec, ynpl=1, ddd=(2 * p.errmin) ** 2, vx=v), 1) is False


# ================================================ step prologue / iteration epilogue
def test_step_begin_dt_growth_and_viscosity_source(hmec, orc):
    p = small(hmec)
    o, _ = fresh(hmec, orc, p)
    o.scatter()
    eta0 = o.get("ETA0")
    o.set("ETA00", full(p, "ETA00", 7e20))
    o.set("OM0", full(p, "OM0", 3.0))
    st = o.state()
    st.dt, st.yndtdecrease = 2e8, 0
    o.set_state(st)
    o.step_begin()                                          # timestep 1: from ETA0
    for name in ("ETA", "ETA1", "ETA50"):
        assert np.array_equal(o.get(name), eta0), name
    assert np.all(o.get("OM") == 3.0)
    st = o.state()
    assert st.dt == pytest.approx(2e8 * p.dtkoefup, rel=1e-15) and st.dt00 == st.dt
    assert st.ynlast == 0 and st.yndtdecrease == 0
    # growth is capped by dtelastic0; after a decrease dt is kept
    st.timestep, st.dt, st.yndtdecrease = 2, 4.5e8, 0
    o.set_state(st)
    o.step_begin()
    assert o.state().dt == p.dtelastic0
    assert np.all(o.get("ETA") == 7e20)                     # later steps: from ETA00
    st = o.state()
    st.dt, st.yndtdecrease = 3e3, 1
    o.set_state(st)
    o.step_begin()
    assert o.state().dt == 3e3 and o.state().yndtdecrease == 0


def test_iter_commit_branches(hmec, orc):
    """Not converged: ETA = clamp(ETA5, etamin, ETA0), ETA50 = ETA, YNY0 = YNY,
    OM = OM5.  Converged: the last ETA5 / OM5 are discarded (the reference's
    exit skips the copy).  ynpl == 0 resets ETA = ETA0 first."""
    p = small(hmec)
    o, _ = fresh(hmec, orc, p)
    rng = np.random.default_rng(2)
    eta0 = rng.uniform(1e20, 1e21, (p.Ny, p.Nx))
    eta5 = rng.uniform(1e-4, 2e21, (p.Ny, p.Nx))
    eta5[0, 0] = 1e-5
    o.set("ETA0", eta0)
    o.set("ETA5", eta5)
    o.set("ETA", full(p, "ETA", 5.))
    o.set("YNY", rng.integers(0, 2, (p.Ny, p.Nx)).astype(float))
    o.set("OM5", full(p, "OM5", -2.))
    o.iter_commit(3, False)
    want = np.maximum(np.minimum(eta5, eta0), p.etamin)
    assert want[0, 0] == p.etamin
    assert np.array_equal(o.get("ETA"), want) and np.array_equal(o.get("ETA50"), want)
    assert np.array_equal(o.get("YNY0"), o.get("YNY")) and np.all(o.get("OM") == -2.)
    o.set("ETA", full(p, "ETA", 5.))
    o.set("OM5", full(p, "OM5", 9.))
    o.iter_commit(3, True)
    assert np.all(o.get("ETA") == 5.) and np.all(o.get("OM") == -2.)
    o.iter_commit(0, True)
    assert np.array_equal(o.get("ETA"), eta0)


# ===================================================== assembly rows
def prepared(hmec, orc, p, **fields):
    o, _ = fresh(hmec, orc, p)
    o.scatter()
    o.step_begin()
    o.set("YNY0", np.zeros((p.Ny, p.Nx)))
    o.vep_pre()
    for k, v in fields.items():
        o.set(k, v)
    o.scale()
    return o


def system(o):
    o.assemble()
    rp, ci, v, b = o.csr()
    N = len(b)
    return sp.csr_matrix((v, ci, rp), shape=(N, N)), b


def test_manufactured_simple_shear_y_stokes(hmec, orc):
    """Uniform effective viscosity, zero history: vy = g x with vx = P = 0
    leaves every interior y-Stokes row with its inertia term -RHOY / dt * vy
    only (the viscous stencil cancels on a field linear in x)."""
    p = small(hmec)
    eta = 1e21
    o = prepared(hmec, orc, p, ETA=full(p, "ETA", eta), ETAP=full(p, "ETAP", eta))
    A, b = system(o)
    N = len(b)
    Ny1, Nx1 = p.Ny + 1, p.Nx + 1
    xvy = np.linspace(-p.dx / 2, p.xsize + p.dx / 2, Nx1)
    x = np.zeros(N)
    g = 1e-12
    for j in range(Nx1):
        for i in range(Ny1):
            x[(j * Ny1 + i) * 6 + 2] = g * xvy[j]
    r = A @ x - b
    scale = abs(A) @ abs(x)
    rhoy = o.get("RHOY")
    dt = o.dt
    n = 0
    for j in range(1, p.Nx):
        for i in range(1, p.Ny - 1):
            k = (j * Ny1 + i) * 6 + 2
            want = -p.inertia * rhoy[i, j] / dt * x[k]
            assert abs(r[k] - want) <= 1e-13 * scale[k], (i, j)
            assert scale[k] > 1e6 * abs(want)
            n += 1
    assert n == (p.Nx - 1) * (p.Ny - 2)


def test_momentum_rows_stress_divergence(hmec, orc):
    """Every interior x- and y-Stokes row is the discrete momentum balance
    div(sigma) - grad(pt) - inertia * rho (v - v0) / dt of a visco-elastic
    (Maxwell) body: deviatoric stresses sigma = 2 eta K e' + (1 - K) sigma0 with
    K = G dt / (G dt + eta), e'xx = -e'yy = (dvx/dx - dvy/dy) / 2 at pressure
    nodes and exy = (dvx/dy + dvy/dx) / 2 at basic nodes, evaluated here on
    the staggered grid from random velocities, pressures, viscosities, moduli,
    densities and stress / velocity histories, node by node."""
    p = small(hmec, Ny=11)
    rng = np.random.default_rng(21)
    Bs, S = (p.Ny, p.Nx), (p.Ny + 1, p.Nx + 1)
    f = dict(ETA=10 ** rng.uniform(19, 21, Bs), GGG=rng.uniform(1e10, 5e10, Bs),
             SXY0=rng.normal(0, 1e6, Bs),
             ETAP=10 ** rng.uniform(19, 21, S), GGGP=rng.uniform(1e10, 5e10, S),
             SXX0=rng.normal(0, 1e6, S), SYY0=rng.normal(0, 1e6, S),
             RHOX=rng.uniform(2e3, 3e3, S), RHOY=rng.uniform(2e3, 3e3, S),
             VX0=rng.normal(0, 1e-9, S), VY0=rng.normal(0, 1e-9, S))
    o = prepared(hmec, orc, p, **f)
    A, b = system(o)
    N = len(b)
    Ny, Nx, Ny1 = p.Ny, p.Nx, p.Ny + 1
    dt, dx, dy = o.dt, p.dx, p.dy
    ps = o.state().ptscale
    vx, vy, pt = rng.normal(0, 1e-9, S), rng.normal(0, 1e-9, S), rng.normal(1e8, 1e6, S)
    x = np.zeros(N)
    for j in range(Nx + 1):
        for i in range(Ny1):
            k = (j * Ny1 + i) * 6
            x[k], x[k + 1], x[k + 2] = pt[i, j] / ps, vx[i, j], vy[i, j]
    r = A @ x - b
    scale = abs(A) @ abs(x) + abs(b)
    KB = dt * f["GGG"] / (dt * f["GGG"] + f["ETA"])
    KP = dt * f["GGGP"] / (dt * f["GGGP"] + f["ETAP"])

    def sxx_syy(i, j):                  # pressure node (i, j)
        e = 0.5 * ((vx[i, j] - vx[i, j - 1]) / dx - (vy[i, j] - vy[i - 1, j]) / dy)
        v = 2 * f["ETAP"][i, j] * KP[i, j] * e
        return v + (1 - KP[i, j]) * f["SXX0"][i, j], -v + (1 - KP[i, j]) * f["SYY0"][i, j]

    def sxy(i, j):                      # basic node (i, j)
        e = 0.5 * ((vx[i + 1, j] - vx[i, j]) / dy + (vy[i, j + 1] - vy[i, j]) / dx)
        return 2 * f["ETA"][i, j] * KB[i, j] * e + (1 - KB[i, j]) * f["SXY0"][i, j]

    inr = p.inertia
    for j in range(1, Nx - 1):
        for i in range(1, Ny):
            want = ((sxx_syy(i, j + 1)[0] - sxx_syy(i, j)[0]) / dx + (sxy(i, j) - sxy(i - 1, j)) / dy
                    - (pt[i, j + 1] - pt[i, j]) / dx - inr * f["RHOX"][i, j] * (vx[i, j] - f["VX0"][i, j]) / dt)
            k = (j * Ny1 + i) * 6 + 1
            assert abs(r[k] - want) <= 1e-12 * scale[k], ("x", i, j, r[k], want)
    for j in range(1, Nx):
        for i in range(1, Ny - 1):
            want = ((sxx_syy(i + 1, j)[1] - sxx_syy(i, j)[1]) / dy + (sxy(i, j) - sxy(i, j - 1)) / dx
                    - (pt[i + 1, j] - pt[i, j]) / dy - inr * f["RHOY"][i, j] * (vy[i, j] - f["VY0"][i, j]) / dt)
            k = (j * Ny1 + i) * 6 + 2
            assert abs(r[k] - want) <= 1e-12 * scale[k], ("y", i, j, r[k], want)


def poro(por, gggb, BETASOLID, BETAFLUID):
    """Textbook poroelastic coefficients: drained compressibility
    beta_d = (1/K_phi + beta_s) / (1 - phi) with the pore modulus K_phi = GGGB,
    Biot-Willis alpha = 1 - beta_s / beta_d, Skempton
    B = (beta_d - beta_s) / (beta_d - beta_s + phi (beta_f - beta_s))."""
    bd = (1 / gggb + BETASOLID) / (1 - por)
    alpha = 1 - BETASOLID / bd
    B = (bd - BETASOLID) / (bd - BETASOLID + por * (BETAFLUID - BETASOLID))
    return bd, alpha, B


def test_continuity_rows_closed_form(hmec, orc):
    """Uniform pressures, random solid velocities and Darcy fluxes: the
    total-pressure row of node (i, j) reads div(v) + (pt - pf) / (ETAB (1 - phi))
    + beta_d ((pt - PT0) - alpha (pf - PF0)) / dt, the fluid-pressure row
    div(q) - (pt - pf) / (ETAB (1 - phi)) - alpha beta_d ((pt - PT0) -
    (pf - PF0) / B) / dt (solid / fluid volume balance: divergence, viscous
    compaction, poroelastic storage), each with its own node's ETAB, phi and
    pore modulus;
    the Pf rows of i = 1 and i = Ny - 1 fix pt - pf = PTFDIFF."""
    p = small(hmec)
    rng = np.random.default_rng(11)
    S = (p.Ny + 1, p.Nx + 1)
    etab = 10 ** rng.uniform(20, 23, S)
    por = rng.uniform(0.005, 0.05, S)
    gggb = 10 ** rng.uniform(11, 13, S)
    pt, pf, dpt, dpf = 1.3e8, 5e7, 2.5e5, -1e5
    o = prepared(hmec, orc, p, ETAB=etab, POR=por, GGGB=gggb,
                 PT0=np.full(S, pt - dpt), PF0=np.full(S, pf - dpf))
    A, b = system(o)
    N = len(b)
    Ny1 = S[0]
    ps = o.state().ptscale
    x = np.zeros(N)
    x[0::6] = pt / ps
    x[3::6] = pf / ps
    # solid velocities and Darcy fluxes: their divergences enter the rows
    v = {n: rng.normal(0, 1e-12, S) for n in ("vx", "vy", "qx", "qy")}
    for j in range(S[1]):
        for i in range(Ny1):
            k = (j * Ny1 + i) * 6
            x[k + 1], x[k + 2], x[k + 4], x[k + 5] = v["vx"][i, j], v["vy"][i, j], v["qx"][i, j], v["qy"][i, j]
    r = A @ x - b
    dt = o.dt
    bd, alpha, B = poro(por, gggb, p.BETASOLID, p.BETAFLUID)
    visc = (pt - pf) / (etab * (1 - por))

    def div(u, w, i, j):
        return (u[i, j] - u[i, j - 1]) / p.dx + (w[i, j] - w[i - 1, j]) / p.dy
    for j in range(1, p.Nx):
        for i in range(1, p.Ny):
            k = (j * Ny1 + i) * 6
            want = div(v["vx"], v["vy"], i, j) + visc[i, j] + bd[i, j] * (dpt - alpha[i, j] * dpf) / dt
            assert r[k] == pytest.approx(want, rel=1e-9), ("Pt", i, j)
            if 2 <= i <= p.Ny - 2:
                want = (div(v["qx"], v["qy"], i, j) - visc[i, j]
                        - alpha[i, j] * bd[i, j] * (dpt - dpf / B[i, j]) / dt)
                assert r[k + 3] == pytest.approx(want, rel=1e-9), ("Pf", i, j)
            else:
                assert r[k + 3] == pytest.approx(-(pt - pf) + p.PTFDIFF, abs=1e-7 * pt)
    # the poroelastic terms are not negligible next to the viscous ones here
    assert np.median(np.abs(bd * dpt / dt) / visc) > 1e-3


@pytest.mark.parametrize("axis", ["x", "y"])
def test_darcy_rows_uniform_gradient(hmec, orc, axis):
    """Darcy's law q = -(k / eta_f) grad pf: with pf linear in x (y), zero solid
    velocity and q = -g / ETADX (ETADY) per node, every interior Darcy row
    reduces to its fluid inertia term -RHOF / POR * inertia / dt * q."""
    p = small(hmec)
    rng = np.random.default_rng(5)
    S = (p.Ny + 1, p.Nx + 1)
    etad = 10 ** rng.uniform(11, 13, S)
    name = "ETADX" if axis == "x" else "ETADY"
    o = prepared(hmec, orc, p, **{name: etad})
    A, b = system(o)
    N = len(b)
    Ny1, Nx1 = S
    ps = o.state().ptscale
    xp = np.linspace(-p.dx / 2, p.xsize + p.dx / 2, Nx1)
    yp = np.linspace(-p.dy / 2, p.ysize + p.dy / 2, Ny1)
    g = 30.0                                        # Pa / m
    x = np.zeros(N)
    var = 4 if axis == "x" else 5
    for j in range(Nx1):
        for i in range(Ny1):
            kp = (j * Ny1 + i) * 6
            x[kp + 3] = g * (xp[j] if axis == "x" else yp[i]) / ps
            x[kp + var] = -g / etad[i, j]
    r = A @ x - b
    scale = abs(A) @ abs(x)
    rhof = o.get("RHOFX" if axis == "x" else "RHOFY")
    por = o.get("PORX" if axis == "x" else "PORY")
    dt = o.dt
    jr = range(1, p.Nx - 1) if axis == "x" else range(1, p.Nx)
    ir = range(1, p.Ny) if axis == "x" else range(1, p.Ny - 1)
    for j in jr:
        for i in ir:
            k = (j * Ny1 + i) * 6 + var
            want = -rhof[i, j] / por[i, j] * p.inertia / dt * x[k]
            assert abs(r[k] - want) <= 1e-12 * scale[k], (i, j)
            assert scale[k] > 1e3 * abs(want)


def expected_bc_rows(p, ps):
    """Boundary rows as h_mec_rsf_v9.cpp states them: vxs :699-728, vys
    :771-797, Pt :843-845, vxD :865-891, vyD :906-934, Pf :952-960.  Entries
    that a row states twice are summed (setFromTriplets)."""
    Nx, Ny, Ny1, NV = p.Nx, p.Ny, p.Ny + 1, 6
    rows = {}

    def add(r, c, v):
        d = rows.setdefault(r, [{}, 0.0])[0]
        d[c] = d.get(c, 0.0) + v

    def rhs(r, v):
        rows.setdefault(r, [{}, 0.0])[1] = v

    for j in range(Nx + 1):
        for i in range(Ny1):
            kp = (j * Ny1 + i) * NV
            kx, ky, kpf, kxf, kyf = kp + 1, kp + 2, kp + 3, kp + 4, kp + 5
            if i == 0 or i == Ny or j == 0 or j >= Nx - 1:
                if j == Nx:
                    add(kx, kx, 1)
                if i == 0 and j < Nx:
                    add(kx, kx, 1); add(kx, kx + NV, 1); rhs(kx, 2 * p.bcupper)
                if i == Ny and j < Nx:
                    add(kx, kx, 1); add(kx, kx - NV, 1); rhs(kx, 2 * p.bclower)
                if j == 0 and 0 < i < Ny:
                    add(kx, kx, 1); add(kx, kx + NV * Ny1, -1)
                if j == Nx - 1 and 0 < i < Ny:
                    add(kx, kx, 1); add(kx, kx - NV * Ny1, -1)
            if j == 0 or j == Nx or i == 0 or i >= Ny - 1:
                if i == Ny:
                    add(ky, ky, 1)
                if j == 0:
                    add(ky, ky, 1); add(ky, ky + Ny1 * NV, 1)
                if j == Nx:
                    add(ky, ky, 1); add(ky, ky - Ny1 * NV, 1)
                if i == 0 and 0 < j < Nx:
                    add(ky, ky, 1)
                if i == Ny - 1 and 0 < j < Nx:
                    add(ky, ky, 1)
            if i == 0 or j == 0 or i == Ny or j == Nx:
                add(kp, kp, 1)
            if i == 0 or i == Ny or j == 0 or j >= Nx - 1:
                if j == Nx:
                    add(kxf, kxf, 1)
                if i == 0 and j < Nx:
                    add(kxf, kxf, 1); add(kxf, kxf + NV, -1)
                if i == Ny and j < Nx:
                    add(kxf, kxf, 1); add(kxf, kxf - NV, -1)
                if j == 0:
                    add(kxf, kxf, 1)
                if j == Nx - 1:
                    add(kxf, kxf, 1)
            if j == 0 or j == Nx or i == 0 or i >= Ny - 1:
                if i == Ny:
                    add(kyf, kyf, 1)
                if j == 0 and 0 < i < Ny - 1:
                    add(kyf, kyf, 1); add(kyf, kyf + Ny1 * NV, -1)
                if j == Nx and 0 < i < Ny - 1:
                    add(kyf, kyf, 1); add(kyf, kyf - Ny1 * NV, -1)
                if i == 0:
                    add(kyf, kyf, 1); rhs(kyf, p.bcvyflower)
                if i == Ny - 1:
                    add(kyf, kyf, 1); rhs(kyf, p.bcvyflower)
            if j == 0 or j == Nx or i <= 1 or i >= Ny - 1:
                if i == 1 or i == Ny - 1:
                    add(kpf, kpf, ps); add(kpf, kp, -ps); rhs(kpf, -p.PTFDIFF)
                else:
                    add(kpf, kpf, 1)
    return rows


def test_every_boundary_row_exact(hmec, orc):
    """Every boundary row of all six unknown types equals the reference's
    statement entry for entry (values, columns and right-hand side), with
    non-zero boundary velocities and fluid flux."""
    p = small(hmec, bcupper=-2e-9, bclower=3e-9, bcvyflower=4e-12)
    o, _ = fresh(hmec, orc, p)
    o.scatter(); o.step_begin(); o.vep_pre()
    ps = o.scale()
    o.assemble()
    rp, ci, v, b = o.csr()
    want = expected_bc_rows(p, ps)
    assert len(want) > 0
    n_interior = 0
    for r in range(len(b)):
        got = dict(zip(ci[rp[r]:rp[r + 1]].tolist(), v[rp[r]:rp[r + 1]].tolist()))
        if r in want:
            assert got == want[r][0], r
            assert b[r] == want[r][1], r
        else:
            n_interior += 1
    # interior rows: Nx-2 x Ny-1 vx, Nx-1 x Ny-2 vy, (Nx-1)(Ny-1) Pt, (Nx-2)(Ny-1) vxD,
    # (Nx-1)(Ny-2) vyD, (Nx-1)(Ny-3) Pf
    Nx, Ny = p.Nx, p.Ny
    assert n_interior == ((Nx - 2) * (Ny - 1) + (Nx - 1) * (Ny - 2) + (Nx - 1) * (Ny - 1) +
                          (Nx - 2) * (Ny - 1) + (Nx - 1) * (Ny - 2) + (Nx - 1) * (Ny - 3))


# ============================================= scatter derived properties (a5/a6)
def lattice(hmec, p, **props):
    soa = hmec.generate_markers(p, cfg=1, seed=9, jitter=False)
    for k, v in props.items():
        soa[hmec.MARKER_FIELDS.index(k)] = v
    return soa


def column_parity(hmec, p, soa):
    nym = (p.Ny - 1) * p.my_per_cell
    ix = np.arange(soa.shape[1]) // nym
    return (ix % 2) == 0


def test_scatter_harmonic_and_arithmetic_means(hmec, orc):
    """On the regular lattice (4 markers per cell in x) markers of even and odd
    x-columns carry equal total weight at every interior node of every grid.
    With two-valued properties by column parity, the shear moduli (GGG, GGGP)
    and the Darcy viscosities (ETADX, ETADY) come out as the equal-weight
    harmonic mean and density as the arithmetic mean."""
    p = small(hmec)
    soa = lattice(hmec, p, porm=p.POR0, t_marker=1.0)
    even = column_parity(hmec, p, soa)
    G1, G2, k1, k2, r1, r2 = 1e10, 4e10, 1e-16, 3e-16, 2500., 3100.
    soa[hmec.MARKER_FIELDS.index("gm")] = np.where(even, G1, G2)
    soa[hmec.MARKER_FIELDS.index("kkkm")] = np.where(even, k1, k2)
    soa[hmec.MARKER_FIELDS.index("rhom")] = np.where(even, r1, r2)
    etaf = 1e-3
    soa[hmec.MARKER_FIELDS.index("etafm")] = etaf
    o, _ = fresh(hmec, orc, p, soa)
    o.scatter()
    hG = 2 / (1 / G1 + 1 / G2)
    hD = 2 / (k1 / etaf + k2 / etaf)
    phi, rf = p.POR0, soa[hmec.MARKER_FIELDS.index("rhofm")][0]
    rho = ((r1 + r2) / 2) * (1 - phi) + rf * phi
    inner_b = (slice(1, p.Ny - 1), slice(1, p.Nx - 1))
    # staggered grids: away from the last two rows / columns, where clamped
    # anchors extrapolate with negative weights (h_mec_rsf_v9.cpp:144-151)
    inner_s = (slice(2, p.Ny - 2), slice(2, p.Nx - 2))
    assert np.allclose(o.get("GGG")[inner_b], hG, rtol=1e-13, atol=0)
    assert np.allclose(o.get("GGGP")[inner_s], hG, rtol=1e-13, atol=0)
    assert np.allclose(o.get("ETADX")[inner_s], hD, rtol=1e-13, atol=0)
    assert np.allclose(o.get("ETADY")[inner_s], hD, rtol=1e-13, atol=0)
    assert np.allclose(o.get("RHO")[inner_b], rho, rtol=1e-13, atol=0)
    assert np.allclose(o.get("RHOX")[inner_s], rho, rtol=1e-13, atol=0)
    # harmonic and arithmetic means are far apart here
    assert abs(hG / ((G1 + G2) / 2) - 1) > 0.3


@pytest.mark.parametrize("phi_ratio,kkkm,want", [
    (1.5, 1e-15, 3.375e-15),          # cubic law
    (3.0, 1e-13, None),               # 2.7e-12 clamped to kkkmax
    (1.0, 1e-23, None),               # clamped to kkkmin
])
def test_scatter_permeability_cubic_law(hmec, orc, phi_ratio, kkkm, want):
    """Permeability follows k = kkkm (phi / POR0)^3 within [kkkmin, kkkmax]."""
    p = small(hmec)
    if want is None:
        want = p.kkkmax if kkkm * phi_ratio ** 3 > p.kkkmax else p.kkkmin
    soa = lattice(hmec, p, porm=phi_ratio * p.POR0, kkkm=kkkm)
    o, _ = fresh(hmec, orc, p, soa)
    o.scatter()
    K = o.get("KKK")
    recv = K != 0
    assert recv.sum() > 0
    assert np.allclose(K[recv], want, rtol=1e-13, atol=0)


@pytest.mark.parametrize("t_marker", [1.0, 0.0])
def test_scatter_porosity_weakening_and_bulk_viscosity(hmec, orc, t_marker):
    """Rock markers (t_marker != 0): matrix viscosity etas exp(-alpha phi),
    cohesion scaled by (1 - phi), density mixed with the fluid's; markers of
    type 0 keep etas, cohesion and density as given.  The bulk viscosity
    ETAB0 = etas / POR either way; ETAP is the effective viscosity etam (the
    basic-grid ETA is not a scatter output: it starts from ETA0, :586-595)."""
    p = small(hmec)
    phi, etas, etam, coh, rhos, rhof = 0.05, 3e21, 2e20, 1e6, 2700., 1000.
    soa = lattice(hmec, p, porm=phi, t_marker=t_marker, etasm=etas, etam=etam, cohescm=coh,
                  rhom=rhos, rhofm=rhof)
    o, _ = fresh(hmec, orc, p, soa)
    o.scatter()
    weak = math.exp(-p.alpha * phi) if t_marker else 1.0
    mix = (1 - phi) if t_marker else 1.0
    rho = rhos * (1 - phi) + rhof * phi if t_marker else rhos
    for name, want in (("ETA0", etas * weak), ("ETAP0", etas * weak), ("ETAP", etam), ("COHC", coh * mix), ("RHO", rho), ("ETAB0", etas / phi)):
        a = o.get(name)
        recv = o.get("POR" if name in ("ETAP0", "ETAP", "ETAB0") else "TTT") != 0 \
            if t_marker else a != 0
        assert recv.sum() > 0, name
        assert np.allclose(a[recv], want, rtol=1e-13, atol=0), name
    if t_marker:
        assert weak < 0.3
