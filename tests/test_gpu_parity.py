"""GPU parity of each stage of the timestep against the CPU oracle (same
seeded inputs), called through the C ABI of libhmec_hip."""
import numpy as np
import pytest

from parity_util import GROUPS, check_step, group_errors, oracle_floor, rel, write_report

pytestmark = pytest.mark.gpu


def test_oracle_solver_is_pardiso(orc, gpu):
    """The parity tests compare against the oracle with MKL PARDISO (the
    reference's solver); its banded-LU fallback must not stand in silently."""
    assert orc.have_pardiso()


def make(hmec, orc, cfg, seed=None, **over):
    p = hmec.params(cfg, **over)
    soa = hmec.generate_markers(p, cfg=cfg, seed=seed if seed is not None else cfg)
    ctx = hmec.Context(p, soa)
    o = orc.Oracle(p, soa, hmec)
    return p, soa, ctx, o


SCATTER_FIELDS = ["RHO", "KKK", "TTT", "GGG", "ETA0", "COHT", "FRIT", "COHC", "FRIC", "DILC",
                  "ETAP", "POR", "ETAB0", "ETAP0", "GGGP", "RHOX", "ETADX", "RHOFX", "PORX",
                  "RHOY", "ETADY", "RHOFY", "PORY"]


@pytest.mark.parametrize("jitter", [True, False])
def test_marker_cells_bitexact(hmec, orc, gpu, jitter):
    """Anchor cells of all four grids, bit-exact; jitter=False is the
    reference's regular lattice, where every 4th marker sits exactly on a node
    line (the knife edge of fix(x / dx))."""
    p = hmec.params(1)
    soa = hmec.generate_markers(p, cfg=1, seed=1, jitter=jitter)
    ctx = hmec.Context(p, soa)
    o = orc.Oracle(p, soa, hmec)
    got = ctx.marker_cells()
    ref = o.marker_cells()
    assert np.array_equal(got, ref)
    if not jitter:
        on_line = (soa[0] / p.dx) == np.floor(soa[0] / p.dx)
        assert on_line.sum() >= soa.shape[1] // 4


@pytest.mark.parametrize("jitter", [True, False])
def test_scatter(hmec, orc, gpu, jitter):
    p = hmec.params(1)
    soa = hmec.generate_markers(p, cfg=1, seed=3, jitter=jitter)
    ctx = hmec.Context(p, soa)
    o = orc.Oracle(p, soa, hmec)
    ctx.scatter()
    o.scatter()
    for f in SCATTER_FIELDS:
        a, b = ctx.get(f), o.get(f)
        scale = np.abs(b).max()
        assert np.abs(a - b).max() <= 1e-13 * max(scale, 1e-300), f


def test_assembly_bitexact(hmec, orc, gpu):
    p, soa, ctx, o = make(hmec, orc, 1)
    ctx.scatter(); o.scatter()
    ctx.step_begin(); o.step_begin()
    dt = o.dt
    assert ctx.state().dt == dt
    ctx.vep_pre(dt); o.vep_pre()
    ps = ctx.assemble(dt)
    ps_o = o.scale()
    o.assemble()
    assert abs(ps - ps_o) <= 1e-13 * abs(ps_o)
    rp, ci, v, b = ctx.csr()
    rp2, ci2, v2, b2 = o.csr()
    assert np.array_equal(rp, rp2) and np.array_equal(ci, ci2)
    # inputs differ only by the scatter's summation order (ulps)
    assert np.abs(v - v2).max() <= 1e-12 * np.abs(v2).max()
    assert np.abs(b - b2).max() <= 1e-12 * np.abs(b2).max()


def test_assembly_same_inputs_bitexact(hmec, orc, gpu):
    """With identical nodal inputs the CSR values and RHS are bit-identical."""
    p, soa, ctx, o = make(hmec, orc, 1)
    o.scatter()
    o.step_begin()
    for f in hmec.FIELDS:
        ctx.set(f, o.get(f))
    ctx.step_begin()
    dt = o.dt
    ctx.vep_pre(dt); o.vep_pre()
    ps = ctx.assemble(dt)
    assert ps == o.scale()
    o.assemble()
    rp, ci, v, b = ctx.csr()
    rp2, ci2, v2, b2 = o.csr()
    assert np.array_equal(v, v2)
    assert np.array_equal(b, b2)


def test_solve_fp64_vs_pardiso(hmec, orc, gpu):
    p, soa, ctx, o = make(hmec, orc, 1)
    o.scatter(); o.step_begin(); o.vep_pre(); o.scale(); o.assemble()
    rp, ci, v, b = o.csr()
    ctx.set_csr_values(v, b)
    ctx.factorize(hmec.PREC_FP64)
    st = ctx.solve()
    x = ctx.solution()
    assert o.solve() == 0
    xr = o.solution()
    import scipy.sparse as sp
    A = sp.csr_matrix((v, ci, rp), shape=(len(b), len(b)))
    res = np.abs(A @ x - b).max() / np.abs(b).max()
    assert res < 1e-13, res
    # solid velocity and pressures: 1e-10 of the field group's norm
    for grp in ([1, 2], [0, 3]):
        idx = np.concatenate([np.arange(g, len(b), 6) for g in grp])
        assert rel(x[idx], xr[idx]) < 1e-10, (grp, rel(x[idx], xr[idx]))


def test_solve_fp32_ir(hmec, orc, gpu):
    p, soa, ctx, o = make(hmec, orc, 1)
    o.scatter(); o.step_begin(); o.vep_pre(); o.scale(); o.assemble()
    rp, ci, v, b = o.csr()
    ctx.set_csr_values(v, b)
    ctx.factorize(hmec.PREC_FP32)
    st = ctx.solve()
    assert st.rel_residual <= 1e-12
    x = ctx.solution()
    assert o.solve() == 0
    xr = o.solution()
    for grp in ([1, 2], [0, 3]):
        idx = np.concatenate([np.arange(g, len(b), 6) for g in grp])
        assert rel(x[idx], xr[idx]) < 1e-8


def tol_of(tol):
    """Group tolerance of the pairwise tests (Darcy: their floor decides)."""
    return 1e-10 if tol is None else tol


# output diagnostics of the final stress state (h_mec_rsf_v9.cpp:1522-1558);
# EL_DECOM divides the one-step pressure change (pt - PT0, a small difference
# of ~1e8 Pa values) by dt, so it inherits the pressures' error relative to
# that difference
DIAG = {"EII": 1e-10, "EIIVP": 1e-10, "SII": 1e-10, "DIS": 1e-10, "pt_ave": 1e-10,
        "pf_ave": 1e-10, "VIS_COMP": 1e-10, "EL_DECOM": 1e-6}


@pytest.mark.timeout(600)
def test_steps_cfg1(hmec, orc, gpu):
    """The whole 101x51 workload (50 timesteps) against the oracle with
    PARDISO: identical Picard iteration counts, dt to 1e-12, every group of
    fields (velocity, pressure, stress, slip rate, state, viscosity, the
    fault's yield stress and bisection stress) within 1e-10 and each field
    carrying a tenth of its group within 1e-10 on its own norm; the Darcy
    fluxes within ten times the oracle's own forward error of that step."""
    p, soa, ctx, o = make(hmec, orc, 1)
    report = []
    for s in range(50):
        a = ctx.step()
        b = o.step()
        assert a.iterations == b.iterations, s
        assert abs(a.dt - b.dt) <= 1e-12 * b.dt, s
        assert a.bisect_capped == b.bisect_capped == 0
        check_step(ctx, o, "cfg1 step %d" % s, floor=oracle_floor(o), report=report)
        for name, tol in DIAG.items():
            gb = o.get(name)
            assert np.abs(gb).max() > 0, name
            e = rel(ctx.get(name), gb)
            assert e < tol, (s, name, e)
    write_report("cfg1_fp64_50steps", report)


def prestress(obj, p, frac=0.98, mu=0.6):
    """Uniform pre-stress at `frac` of mu * (PT0 - PF0): the fault goes
    coseismic on the second step (plastic Picard iterations, dt collapse)."""
    tau = frac * mu * p.PTFDIFF
    obj.set("SXY0", np.full((p.Ny, p.Nx), tau))
    obj.set("SXY", np.full((p.Ny, p.Nx), tau))


def test_prestressed_fault_cfg1(hmec, orc, gpu):
    """Coseismic steps (10 plastic Picard iterations, dt from 1e7 s to 6e-4 s):
    the GPU path against the oracle with PARDISO, bounded by the oracle's own
    solver noise (PARDISO vs the oracle's banded LU on the same input).  The
    plastic viscosity dt*G*A/(1-A) with A -> 1 amplifies rounding, so the two
    CPU solvers already differ by ~1e-6 there; every group must stay within
    max(1e-10, 10 x that floor), iteration counts and dt must agree."""
    p, soa, ctx, o = make(hmec, orc, 1)
    ob = orc.Oracle(p, soa, hmec, use_pardiso=False)
    for x in (ctx, o, ob):
        prestress(x, p)
    plastic = 0
    for s in range(4):
        a = ctx.step()
        b = o.step()
        c = ob.step()
        assert a.iterations == b.iterations == c.iterations, (s, a.iterations, b.iterations)
        assert abs(a.dt - b.dt) <= 1e-9 * b.dt, (s, a.dt, b.dt)
        plastic += b.iterations > 1
        floor = group_errors_pair(o, ob)
        for name, (e, tol) in group_errors(ctx, o).items():
            assert e <= max(tol_of(tol), 10 * floor[name]), (s, name, e, floor[name])
    assert plastic >= 3


def group_errors_pair(x, y):
    out = {}
    for name, (fields, tol) in GROUPS.items():
        ga = np.concatenate([x.get(f).ravel() for f in fields])
        gb = np.concatenate([y.get(f).ravel() for f in fields])
        out[name] = rel(ga, gb)
    return out


def test_advection_step(hmec, orc, gpu):
    """RK4 marker advection switched on (the reference's loop body with
    markers moving): markers, marker stresses and the next scatter agree."""
    p = hmec.params(1, advect_markers=1)
    soa = hmec.generate_markers(p, cfg=1, seed=7)
    ctx = hmec.Context(p, soa)
    o = orc.Oracle(p, soa, hmec)
    for _ in range(2):
        ctx.step()
        o.step()
    mg = ctx.get_markers()
    for f in (0, 1, 16, 17, 18):
        mo = o.marker(f)
        scale = max(np.abs(mo).max(), 1e-300)
        assert np.abs(mg[f] - mo).max() <= 1e-9 * scale, f
    assert np.array_equal(ctx.marker_cells(), o.marker_cells())
