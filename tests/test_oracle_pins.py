"""Pins of the CPU oracle (oracle/hmec_oracle.c) that do not retype its
formulas: exhaustive integer checks, closed forms, invariants, brute force on
tiny grids and an independent solver.  The reference ships no tests or golden
vectors for this path, so these are what fixes the oracle.
CPU only."""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla


def small(hmec, Nx=13, Ny=9, **over):
    base = dict(xsize=6000., ysize=4000., Nx=Nx, Ny=Ny, upper_block=1800., lower_block=2200.,
                TS_1=1000., TS_2=2000., TS_3=4000., TS_4=5000.)
    base.update(over)
    return hmec.params(1, **base)


def mk_oracle(hmec, orc, p, soa=None, cfg=1, seed=1, jitter=True):
    if soa is None:
        soa = hmec.generate_markers(p, cfg=cfg, seed=seed, jitter=jitter)
    return orc.Oracle(p, soa, hmec), soa


# ------------------------------------------------------------------ helpers
def test_fix_is_truncation(orc):
    L = orc.lib()
    xs = np.concatenate([np.arange(-12, 12, 0.25), np.random.default_rng(0).normal(0, 50, 2000)])
    for x in xs:
        assert L.orc_fix(float(x)) == math.trunc(x)
    # the marker knife edge: x exactly on a node line
    for k in range(-5, 6):
        assert L.orc_fix(float(k)) == k


def test_check_bounds_exhaustive(orc):
    L = orc.lib()
    for bound in range(2, 12):
        for k in range(-6, 20):
            assert L.orc_check_bounds(k, bound) == min(max(k, 0), bound - 2)


def test_divplus(orc):
    L = orc.lib()
    rng = np.random.default_rng(1)
    for _ in range(200):
        a, b = rng.uniform(1e-3, 1e3, 2)
        assert abs(L.orc_divplus(a, b) - a / (a + b)) <= 1e-15
    assert L.orc_divplus(1.0, 0.0) == 1.0


def test_linspaced_rule(orc):
    L = orc.lib()
    for n, lo, hi in ((401, 0.0, 40000.0), (52, -100.0, 10100.0), (7, 3.0, -1.0), (5, -9.0, 1.0)):
        v = np.zeros(n)
        L.orc_linspaced(v.ctypes.data_as(C.POINTER(C.c_double)), n, lo, hi)
        step = (hi - lo) / (n - 1)
        assert v[0] == lo and v[-1] == hi
        # equally spaced to rounding, monotone
        assert np.allclose(np.diff(v), step, rtol=1e-12, atol=1e-12 * abs(hi))
        assert np.all(np.sign(np.diff(v)) == np.sign(step))


# ------------------------------------------------------------------ scatter
def _set_marker(soa, hmec, name, value):
    soa[hmec.MARKER_FIELDS.index(name)] = value


def test_scatter_constant_fields(hmec, orc):
    """A constant marker property comes back exactly (to ulps) on every node
    that receives weight: the normalization is a partition of unity."""
    p = small(hmec)
    soa = hmec.generate_markers(p, cfg=1, seed=5)
    _set_marker(soa, hmec, "porm", 0.02)
    _set_marker(soa, hmec, "rhom", 2700.0)
    _set_marker(soa, hmec, "t_marker", 1.0)
    _set_marker(soa, hmec, "frictcm", 0.55)
    o, _ = mk_oracle(hmec, orc, p, soa)
    o.scatter()
    rho = 2700.0 * (1 - 0.02) + 1000.0 * 0.02
    for name, val in (("RHO", rho), ("RHOX", rho), ("RHOY", rho), ("FRIC", 0.55), ("POR", 0.02),
                      ("PORX", 0.02), ("PORY", 0.02), ("TTT", 1.0), ("GGG", 3e10)):
        a = o.get(name)
        recv = a != 0
        assert recv.sum() > 0
        assert np.allclose(a[recv], val, rtol=4e-15, atol=0), name


def test_scatter_linear_field_interior(hmec, orc):
    """On the regular (un-jittered) lattice a property linear in x is
    reproduced at interior basic nodes (markers symmetric about the node)."""
    p = small(hmec)
    soa = hmec.generate_markers(p, cfg=1, seed=5, jitter=False)
    xm = soa[0]
    _set_marker(soa, hmec, "porm", 0.0)
    _set_marker(soa, hmec, "t_marker", 1.0)
    _set_marker(soa, hmec, "rhom", 1000.0 + 0.01 * xm)
    o, _ = mk_oracle(hmec, orc, p, soa)
    o.scatter()
    rho = o.get("RHO")
    x = np.linspace(0, p.xsize, p.Nx)
    want = 1000.0 + 0.01 * x
    assert np.allclose(rho[1:-1, 1:-1], want[None, 1:-1], rtol=1e-13, atol=0)


def test_scatter_edge_rows_never_receive(hmec, orc):
    """check_bounds clamps anchors to [0, bound-2] on every grid: the last
    row and column of the P / vx / vy grids get no marker weight."""
    p = small(hmec)
    o, _ = mk_oracle(hmec, orc, p)
    o.scatter()
    for name in ("ETAP0", "POR", "RHOX", "RHOY", "ETADX"):
        a = o.get(name)
        assert np.all(a[-1, :] == 0) and np.all(a[:, -1] == 0), name
        assert np.all(a[:-1, :-1] != 0), name


def test_marker_cells_brute_force(hmec, orc):
    """Anchor cells equal an integer brute force on markers placed exactly on
    node lines and half-node lines (the knife edges of fix())."""
    p = small(hmec)
    soa = hmec.generate_markers(p, cfg=1, seed=2, jitter=False)
    o, _ = mk_oracle(hmec, orc, p, soa)
    cells = o.marker_cells()
    nxm, nym = (p.Nx - 1) * 4, (p.Ny - 1) * 4
    m = np.arange(soa.shape[1])
    ix, iy = m // nym, m % nym          # lattice indices: x = ix * dx / 4 exactly
    Ny1 = p.Ny + 1

    def clamp(k, b):
        return np.minimum(np.maximum(k, 0), b - 2)
    # xm / dx = ix / 4 exactly for the regular lattice
    jb = clamp(ix // 4, p.Nx)
    ib = clamp(iy // 4, p.Ny)
    jp = clamp((ix + 2) // 4, p.Nx)
    ip = clamp((iy + 2) // 4, p.Ny)
    assert np.array_equal(cells[0], ib + jb * Ny1)
    assert np.array_equal(cells[1], ip + jp * Ny1)
    assert np.array_equal(cells[2], ip + jb * Ny1)
    assert np.array_equal(cells[3], ib + jp * Ny1)


# ------------------------------------------------------------------ assembly
def assembled(hmec, orc, p, soa=None):
    o, soa = mk_oracle(hmec, orc, p, soa)
    o.scatter()
    o.step_begin()
    o.vep_pre()
    o.scale()
    o.assemble()
    return o


def test_pattern_row_lengths(hmec, orc):
    p = small(hmec)
    o = assembled(hmec, orc, p)
    rp, ci, v, b = o.csr()
    Ny1 = p.Ny + 1
    lens = np.diff(rp)
    i, j = 3, 4                      # interior node
    k = (j * Ny1 + i) * 6
    assert list(lens[k:k + 6]) == [6, 11, 11, 6, 4, 4]   # Pt vx vy Pf vxD vyD


def test_duplicate_triplets_summed(hmec, orc):
    p = small(hmec)
    o = assembled(hmec, orc, p)
    rp, ci, v, b = o.csr()
    Ny1, Nx, Ny = p.Ny + 1, p.Nx, p.Ny

    def diag(k):
        for q in range(rp[k], rp[k + 1]):
            if ci[q] == k:
                return v[q]
    for j in (0, Nx):                 # vys rows at i == Ny
        assert diag((j * Ny1 + Ny) * 6 + 2) == 2.0
    for i in (0, Ny):                 # vxD rows at j in {0, Nx-1}
        for j in (0, Nx - 1):
            assert diag((j * Ny1 + i) * 6 + 4) == 2.0
    # plain ghost rows are exactly identity rows
    k = (Nx * Ny1 + 3) * 6           # vxs ghost column j == Nx
    assert diag(k + 1) == 1.0 and rp[k + 2] - rp[k + 1] == 1


def test_manufactured_simple_shear(hmec, orc):
    """Uniform viscosity, zero stress history: the linear shear vx = g * y
    with vy = P = 0 makes every interior x-Stokes row reduce to its inertia
    term -RHOX / dt * vx (the viscous stencil cancels)."""
    p = small(hmec)
    o, soa = mk_oracle(hmec, orc, p)
    o.scatter()
    o.step_begin()
    for name in ("ETA", "ETA0"):
        o.set(name, np.full((p.Ny, p.Nx), 1e21))
    for name in ("ETAP", "ETAP0"):
        o.set(name, np.full((p.Ny + 1, p.Nx + 1), 1e21))
    o.set("YNY0", np.zeros((p.Ny, p.Nx)))
    o.vep_pre()
    o.scale()
    o.assemble()
    rp, ci, v, b = o.csr()
    N = len(b)
    A = sp.csr_matrix((v, ci, rp), shape=(N, N))
    Ny1 = p.Ny + 1
    yvx = np.linspace(-p.dy / 2, p.ysize + p.dy / 2, Ny1)
    x = np.zeros(N)
    g = 1e-12
    for j in range(p.Nx + 1):
        for i in range(Ny1):
            x[(j * Ny1 + i) * 6 + 1] = g * yvx[i]
    r = A @ x
    scale = abs(A) @ abs(x)          # size of the individual terms of each row
    rhox = o.get("RHOX")
    dt = o.dt
    for j in range(1, p.Nx - 1):
        for i in range(1, p.Ny):
            k = (j * Ny1 + i) * 6 + 1
            want = -rhox[i, j] / dt * x[k]
            # a dropped term or a wrong sign leaves an O(scale) residual
            assert abs(r[k] - want) <= 1e-13 * scale[k], (i, j)
            assert scale[k] > 1e6 * abs(want)


def test_bc_rows_exact(hmec, orc):
    p = small(hmec)
    o = assembled(hmec, orc, p)
    rp, ci, v, b = o.csr()
    Ny1 = p.Ny + 1
    j = 3
    k = (j * Ny1 + 0) * 6 + 1         # vxs upper boundary: vx(0) + vx(1) = 2 bcupper
    row = dict(zip(ci[rp[k]:rp[k + 1]], v[rp[k]:rp[k + 1]]))
    assert row == {k: 1.0, k + 6: 1.0} and b[k] == 2 * p.bcupper


# ------------------------------------------------------------------ solver
def test_pardiso_residual_cfg1(hmec, orc):
    if not orc.have_pardiso():
        pytest.skip("MKL PARDISO not loadable")
    p = hmec.params(1)
    o = assembled(hmec, orc, p)
    assert o.solve() == 0
    rp, ci, v, b = o.csr()
    x = o.solution()
    A = sp.csr_matrix((v, ci, rp), shape=(len(b), len(b)))
    assert np.abs(A @ x - b).max() / np.abs(b).max() < 1e-13


def test_band_lu_residual_tiny(hmec, orc):
    """Tiny grids give numerically singular systems (cond ~ 1e22): only the
    backward error is meaningful there."""
    p = small(hmec)
    o = assembled(hmec, orc, p)
    rp, ci, v, b = o.csr()
    A = sp.csr_matrix((v, ci, rp), shape=(len(b), len(b)))
    orc.lib().orc_use_pardiso(o.h, 0)
    assert o.solve() == 0
    xb = o.solution()
    assert np.abs(A @ xb - b).max() / np.abs(b).max() < 1e-13


def test_band_lu_matches_pardiso_cfg1(hmec, orc):
    """The oracle's banded LU (its solver without MKL) and PARDISO agree on
    the well-conditioned 101x51 system (solid velocity and pressures)."""
    if not orc.have_pardiso():
        pytest.skip("MKL PARDISO not loadable")
    p = hmec.params(1)
    o = assembled(hmec, orc, p)
    rp, ci, v, b = o.csr()
    A = sp.csr_matrix((v, ci, rp), shape=(len(b), len(b)))
    assert o.solve() == 0
    xp = o.solution()
    orc.lib().orc_use_pardiso(o.h, 0)
    assert o.solve() == 0
    xb = o.solution()
    assert np.abs(A @ xb - b).max() / np.abs(b).max() < 1e-13
    for grp in ([1, 2], [0, 3]):
        idx = np.concatenate([np.arange(g, len(b), 6) for g in grp])
        assert np.linalg.norm(xb[idx] - xp[idx]) <= 1e-10 * np.linalg.norm(xp[idx]), grp


def test_darcy_forward_error_bound(hmec, orc):
    """Measures the oracle's own forward error on the Darcy fluxes against an
    extended-precision refinement: cancellation-limited well above 1e-10
    (the fluxes are differences of nearly equal pressures), which is why the
    GPU parity holds them to a multiple of this measured floor instead."""
    p = hmec.params(1)
    o = assembled(hmec, orc, p)
    assert o.solve() == 0
    rp, ci, v, b = o.csr()
    x = o.solution()
    N = len(b)
    A = sp.csr_matrix((v, ci, rp), shape=(N, N))
    r = 1.0 / abs(A).max(axis=1).toarray().ravel()
    c = 1.0 / abs(sp.diags(r) @ A).max(axis=0).toarray().ravel()
    lu = sla.splu((sp.diags(r) @ A @ sp.diags(c)).tocsc())   # equilibrated correction solver
    vl = v.astype(np.longdouble)

    def resid(xl):
        prod = vl * xl[ci]
        out = np.add.reduceat(prod, rp[:-1])
        out[rp[:-1] == rp[1:]] = 0
        return b.astype(np.longdouble) - out
    xl = x.astype(np.longdouble)
    for _ in range(4):
        xl = xl + (c * lu.solve(r * resid(xl).astype(np.float64))).astype(np.longdouble)
    xt = xl.astype(np.float64)
    for var in (4, 5):
        e = np.linalg.norm(x[var::6] - xt[var::6]) / np.linalg.norm(xt[var::6])
        assert e < 1e-4, (var, e)
    for var in (0, 1, 3):
        e = np.linalg.norm(x[var::6] - xt[var::6]) / np.linalg.norm(xt[var::6])
        assert e < 1e-12, (var, e)


# ------------------------------------------------------------------ stress / RSF
def _shear_state(hmec, o, p, g=1e-13):
    Ny1, Nx1 = p.Ny + 1, p.Nx + 1
    yvx = np.linspace(-p.dy / 2, p.ysize + p.dy / 2, Ny1)
    o.set("vxs", np.repeat((g * yvx)[:, None], Nx1, axis=1))
    o.set("vys", np.zeros((Ny1, Nx1)))
    return g


@pytest.mark.parametrize("limit", ["viscous", "elastic"])
def test_stress_update_limits(hmec, orc, limit):
    """dt -> infinity: sigma_xy -> 2 eta eps_xy; dt -> 0: sigma_xy -> sigma0."""
    p = small(hmec)
    o, _ = mk_oracle(hmec, orc, p)
    o.scatter()
    o.step_begin()
    g = _shear_state(hmec, o, p)
    o.set("SXY0", np.full((p.Ny, p.Nx), 3e6))
    o.dt = 1e40 if limit == "viscous" else 1e-40
    o.update()
    sxy = o.get("SXY")
    eta = o.get("ETA")
    if limit == "viscous":
        assert np.allclose(sxy[:-1, :], 2 * eta[:-1, :] * (0.5 * g), rtol=1e-12, atol=0)
    else:
        assert np.allclose(sxy, 3e6, rtol=1e-12, atol=0)


def test_aging_law_against_ode(hmec, orc):
    """OM5 of every fault node equals a fine RK4 integration of the aging law
    d theta/dt = 1 - V theta / L (theta = L/V0 e^Omega) at the node's slip rate."""
    p = hmec.params(1)
    o, _ = mk_oracle(hmec, orc, p)
    o.step()
    o.step()           # RSF active from the second step
    om0 = o.get("OM0")
    st = o.state()
    o.scatter(); o.step_begin()
    o.vep_pre(); o.scale(); o.assemble()
    assert o.solve() == 0
    o.reload()
    o.update()
    V = o.get("VSLIPB")[p.line_fault]
    om5 = o.get("OM5")[p.line_fault]
    L = o.get("LRSF")[p.line_fault]
    dt = o.dt
    V0 = p.V0
    for j in range(0, p.Nx, 10):
        th = L[j] / V0 * math.exp(om0[p.line_fault, j])
        nsub = 64
        h = dt / nsub
        f = lambda t: 1.0 - V[j] * t / L[j]
        for _ in range(nsub):
            k1 = f(th); k2 = f(th + h / 2 * k1); k3 = f(th + h / 2 * k2); k4 = f(th + h * k3)
            th += h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        want = math.log(V0 * th / L[j])
        assert abs(om5[j] - want) < 1e-9 * max(1.0, abs(want)), j


def test_bisection_root(hmec, orc):
    """Where the bracket exists, the stress at the root satisfies the
    viscoplastic relation to the reference's 1e-3 Pa tolerance, evaluated
    independently here with mpmath-free float64 numpy."""
    p = hmec.params(1)
    o, _ = mk_oracle(hmec, orc, p)
    for _ in range(3):
        o.step()
    o.scatter(); o.step_begin(); o.vep_pre(); o.scale(); o.assemble()
    assert o.solve() == 0
    o.reload()
    o.update()
    i = p.line_fault
    tau = o.get("SII_fault")[i]
    V = o.get("VSLIPB")[i]
    a, bb, fr = o.get("ARSF")[i], o.get("BRSF")[i], o.get("FRIC")[i]
    om = o.get("OM")[i]
    pt, pf = o.get("pt"), o.get("pf")
    prB = ((pt[i, :-1] + pt[i + 1, :-1] + pt[i, 1:] + pt[i + 1, 1:]) -
           (pf[i, :-1] + pf[i + 1, :-1] + pf[i, 1:] + pf[i + 1, 1:]))[:p.Nx] / 4
    prB = np.maximum(prB, 1e3)
    bracket = tau != 0
    Vt = 2 * p.V0 * np.sinh(np.maximum(tau, 0) / a / prB) * np.exp(-(bb * om + fr) / a)
    assert np.allclose(Vt[bracket], V[bracket], rtol=1e-9, atol=0)


# ------------------------------------------------------------------ advection
def test_advection_uniform_translation(hmec, orc):
    p = small(hmec)
    soa = hmec.generate_markers(p, cfg=1, seed=4)
    o, _ = mk_oracle(hmec, orc, p, soa)
    Ny1, Nx1 = p.Ny + 1, p.Nx + 1
    U, W = 2e-9, -1e-9
    o.set("vxs", np.full((Ny1, Nx1), U))
    o.set("vys", np.full((Ny1, Nx1), W))
    o.set("ESP", np.zeros((p.Ny, p.Nx)))
    o.dt = 1e9
    x0, y0 = soa[0].copy(), soa[1].copy()
    o.advect()
    x1, y1 = o.marker(0), o.marker(1)
    want_x = x0 + U * 1e9
    want_x = np.where(want_x > p.xsize, want_x - p.xsize, want_x)
    assert np.allclose(x1, want_x, rtol=0, atol=1e-9)
    assert np.allclose(y1, y0 + W * 1e9, rtol=0, atol=1e-9)


def test_advection_rotation_invariants(hmec, orc):
    """A uniform spin rotates marker stresses rigidly: the trace and the
    second invariant are preserved."""
    p = small(hmec)
    soa = hmec.generate_markers(p, cfg=1, seed=4)
    rng = np.random.default_rng(3)
    M = soa.shape[1]
    soa[16] = rng.normal(0, 1e6, M)
    soa[17] = rng.normal(0, 1e6, M)
    soa[18] = rng.normal(0, 1e6, M)
    o, _ = mk_oracle(hmec, orc, p, soa)
    o.set("vxs", np.zeros((p.Ny + 1, p.Nx + 1)))
    o.set("vys", np.zeros((p.Ny + 1, p.Nx + 1)))
    o.set("ESP", np.full((p.Ny, p.Nx), 1e-10))
    o.dt = 3e9
    sxx, syy, sxy = soa[16], soa[17], soa[18]
    o.advect()
    nxx, nyy, nxy = o.marker(16), o.marker(17), o.marker(18)
    assert np.allclose(nxx + nyy, sxx + syy, rtol=0, atol=1e-6)
    j2a = ((sxx - syy) / 2) ** 2 + sxy ** 2
    j2b = ((nxx - nyy) / 2) ** 2 + nxy ** 2
    assert np.allclose(j2a, j2b, rtol=1e-12, atol=0)
    # the rotation angle is the spin times dt
    assert not np.allclose(nxy, sxy)


def test_diagnostics_closed_forms(hmec, orc):
    """Output diagnostics of the final stress state (h_mec_rsf_v9.cpp:1522-1558)
    on a constructed state: simple shear vxs = g y (EXY = g/2, EXX = 0) in the
    elastic limit (dt -> 0: SXX, SYY, SXY keep their old values), SXY0 linear
    in (i, j), pressures linear in i with a uniform step change.  For a
    linear L over a 2x2 block the sum of squares is 4 Lc^2 + b^2 + c^2 (Lc at
    the block centre), which gives EII, EIIVP, SII and DIS in closed form;
    the averaged pressures are the mid-row values (last row: its own), and
    the compaction terms follow from them."""
    p = small(hmec)
    o, _ = mk_oracle(hmec, orc, p)
    o.step()                                   # controller state (dtx, dty, dtlapusta)
    Ny, Nx, Ny1, Nx1 = p.Ny, p.Nx, p.Ny + 1, p.Nx + 1
    g = 3e-13
    yvx = np.linspace(-p.dy / 2, p.ysize + p.dy / 2, Ny1)
    o.set("vxs", np.repeat((g * yvx)[:, None], Nx1, axis=1))
    o.set("vys", np.zeros((Ny1, Nx1)))
    eta, etap, etab, por, gggb = 2e21, 3e21, 5e21, 0.01, 2e10
    o.set("ETA", np.full((Ny, Nx), eta))
    o.set("ETAP", np.full((Ny1, Nx1), etap))
    o.set("ETAB", np.full((Ny1, Nx1), etab))
    o.set("POR", np.full((Ny1, Nx1), por))
    o.set("GGGB", np.full((Ny1, Nx1), gggb))
    a, b, c = 2e6, 3e4, -7e3
    ii, jj = np.meshgrid(np.arange(Ny), np.arange(Nx), indexing="ij")
    o.set("SXY0", a + b * ii + c * jj)
    sxx0 = np.full((Ny1, Nx1), 4e5)
    syy0 = np.full((Ny1, Nx1), -4e5)
    o.set("SXX0", sxx0)
    o.set("SYY0", syy0)
    I = np.arange(Ny1)[:, None] * np.ones((1, Nx1))
    pt = 1e8 + 2e5 * I
    pf = 6e7 + 1e5 * I
    dpt, dpf = 125.0, 50.0
    o.set("pt", pt)
    o.set("pf", pf)
    o.set("PT0", pt - dpt)
    o.set("PF0", pf - dpf)
    o.dt = 1e-30
    o.step_end()
    dt = o.dt
    assert dt == 1e-30
    ci, cj = np.meshgrid(np.arange(1, Ny), np.arange(1, Nx), indexing="ij")
    inner = (slice(1, Ny), slice(1, Nx))
    lc = a + b * (ci - 0.5) + c * (cj - 0.5)
    s2 = 4 * lc ** 2 + b ** 2 + c ** 2           # sum of SXY^2 over the block
    sxx, syy = 4e5, -4e5
    want = {
        "EII": np.full(ci.shape, g / 2),
        "SII": np.sqrt(0.5 * (sxx ** 2 + syy ** 2) + s2 / 4),
        "EIIVP": np.sqrt(0.5 * ((sxx / (2 * etap)) ** 2 + (syy / (2 * etap)) ** 2) + s2 / eta ** 2 / 16),
        "DIS": (sxx ** 2 + syy ** 2) / (2 * etap) + s2 / eta / 2,
    }
    for name, w in want.items():
        got = o.get(name)[inner]
        assert np.allclose(got, w, rtol=1e-12, atol=0), (name, np.abs(got / w - 1).max())
    pta = np.where(ci < Ny - 1, 1e8 + 2e5 * (ci + 0.5), 1e8 + 2e5 * ci)
    pfa = np.where(ci < Ny - 1, 6e7 + 1e5 * (ci + 0.5), 6e7 + 1e5 * ci)
    assert np.allclose(o.get("pt_ave")[inner], pta, rtol=1e-15, atol=0)
    assert np.allclose(o.get("pf_ave")[inner], pfa, rtol=1e-15, atol=0)
    assert np.allclose(o.get("VIS_COMP")[inner], (pta - pfa) / (etab * (1 - por)), rtol=1e-14, atol=0)
    bd = (1 / gggb + p.BETASOLID) / (1 - por)
    kbw = 1 - p.BETASOLID / bd
    el = bd * (dpt - kbw * dpf) / dt
    assert np.allclose(o.get("EL_DECOM")[inner], el, rtol=1e-9, atol=0)
    # cells outside the pressure-cell range stay 0
    for name in ("EII", "SII", "DIS", "VIS_COMP", "EL_DECOM", "DSII"):
        f = o.get(name)
        assert not f[0, :].any() and not f[:, 0].any() and not f[Ny, :].any() and not f[:, Nx].any(), name
    assert not o.get("DSII").any()
