"""The C-ABI library loads without a GPU and exports every entry point that
include/hmec.h declares; the Python struct mirrors have the C sizes.  No
compute call is made here (CPU only)."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "hmec.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(hmec_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_exports_every_declared_symbol(hmec):
    lib = C.CDLL(hmec.LIB_PATH)
    names = header_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_sizes(hmec):
    L = hmec.lib()
    assert L.hmec_params_size() == C.sizeof(hmec.Params)
    assert L.hmec_step_info_size() == C.sizeof(hmec.StepInfo)


def test_params_defaults_are_constants_hpp(hmec):
    p = hmec.params()
    assert (p.Nx, p.Ny, p.xsize, p.ysize) == (401, 51, 40000.0, 10000.0)
    assert p.dx == 100.0 and p.dy == 200.0 and p.faultwidth == p.dx
    assert p.line_fault == 25 and p.lower_block == 5200.0 and p.upper_block == 4800.0
    assert (p.dtelastic0, p.niterglobal, p.ynlastmax, p.dtstep) == (5e8, 10000, 499, 5)
    assert list(p.arsfm) == [.025, .006] and list(p.omm) == [10, -10]


def test_marker_lattice_without_jitter(hmec):
    """jitter = 0 reproduces the reference lattice (h_mec_rsf_v9.cpp:190-191)."""
    p = hmec.params(1)
    soa = hmec.generate_markers(p, cfg=1, seed=9, jitter=False)
    nym = (p.Ny - 1) * 4
    m = 12345
    assert soa[0][m] == (m // nym) * (p.xsize / ((p.Nx - 1) * 4))
    assert soa[1][m] == (m % nym) * (p.ysize / nym)
    # seeded: the same seed gives the same bits, another seed differs
    a = hmec.generate_markers(p, cfg=2, seed=5)
    b = hmec.generate_markers(p, cfg=2, seed=5)
    c = hmec.generate_markers(p, cfg=2, seed=6)
    assert (a == b).all() and not (a == c).all()


def test_errors_are_codes(hmec):
    L = hmec.lib()
    assert L.hmec_strerror(-4).decode().startswith("zero pivot")
    p = hmec.Params()
    assert L.hmec_params_config(99, C.byref(p)) == -1
