"""Fault-row output streams (EVO_*.txt, x_fault.txt, rsf_fault.txt,
h_mec_rsf_v9.cpp:1739-1805) written through the C ABI after each timestep,
checked against the fields the same context returns.  The reference prints
doubles with an ostream's default format, which the C++ standard defines as
printf's %g at precision 6, and Eigen vectors with every coefficient
right-aligned to the widest one (rows space-separated, columns one per
line)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SERIES = [("EVO_Vslip.txt", "VSLIPB", None), ("EVO_viscosity.txt", "ETA", None),
          ("EVO_press_flu.txt", "pf", None), ("EVO_press_eff.txt", "pt", "pf"),
          ("EVO_SigmaY.txt", "SigmaY", None), ("EVO_Sii.txt", "SII_fault", None),
          ("EVO_Theta.txt", "OM", None), ("EVO_Visc_comp.txt", "VIS_COMP", None),
          ("EVO_Elast_comp.txt", "EL_DECOM", None), ("EVO_vxD.txt", "vxD", None)]


def g(v):
    return "%g" % v


def eigen_vec(tokens, sep):
    w = max(len(t) for t in tokens)
    return sep.join(t.rjust(w) for t in tokens)


def test_evo_streams(hmec, gpu, tmp_path):
    p = hmec.params(1)
    soa = hmec.generate_markers(p, cfg=1, seed=1)
    ctx = hmec.Context(p, soa, precision=hmec.PREC_FP64)
    lf = p.line_fault
    expect = {name: [] for name, _, _ in SERIES}
    data = []
    for s in range(3):
        info = ctx.step()
        ctx.write_evo(tmp_path, first=(s == 0))
        for name, f, sub in SERIES:
            row = ctx.get(f)[lf]
            if sub is not None:
                row = row - ctx.get(sub)[lf]
            expect[name].append(g(info.timesum) + "    " + g(info.dt) + "    " +
                                eigen_vec([g(v) for v in row], " "))
        data.append("".join("%20s" % t for t in (g(info.timesum), g(info.dt), g(info.Vmax),
                                                  str(info.ynlast + 1), str(info.iterations))))
    for name, _, _ in SERIES:
        lines = (tmp_path / name).read_text().splitlines()
        assert lines == expect[name], name
    assert (tmp_path / "EVO_data.txt").read_text().splitlines() == data
    # written once, on the first call
    xs = (tmp_path / "x_fault.txt").read_text()
    x = np.linspace(0, p.xsize, p.Nx)
    toks = xs.split()
    assert len(toks) == p.Nx and xs.endswith("\n")
    assert xs[:-1] == eigen_vec(toks, "\n")
    assert np.allclose([float(t) for t in toks], x, rtol=1e-5, atol=0)
    rsf = (tmp_path / "rsf_fault.txt").read_text()
    want = "\n\n".join(eigen_vec([g(v) for v in ctx.get(f)[lf]], " ") for f in ("ARSF", "BRSF", "LRSF"))
    assert rsf == want + "\n"


def test_evo_unwritable_dir(hmec, gpu, tmp_path):
    p = hmec.params(1)
    soa = hmec.generate_markers(p, cfg=1, seed=1)
    ctx = hmec.Context(p, soa, precision=hmec.PREC_FP64)
    ctx.step()
    with pytest.raises(hmec.HmecError, match="output file"):
        ctx.write_evo(tmp_path / "missing" / "dir")
