"""HDF5 snapshot layout of the reference's restart files
(h_mec_rsf_v9.cpp:1807-1876, cpp/hdf5.cpp:13-150), on the CPU: a snapshot
with distinct synthetic values is written by hmec_h5_write and read back
through the HDF5 C library itself (ctypes; h5py is not installed), which
checks the group and dataset names, the order the datasets were written in
(their file offsets increase in that order), their shapes, the row-major
element order and /Value/values; then hmec_h5_read round-trips it.  Also the
EVO stream trimming of a restart (:268-282) and the oracle's per-step series
(:1661-1690)."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import hm

BASIC = ["SIIB", "OM0", "OM", "ARSF", "BRSF", "LRSF", "RHO", "ETA0", "ETA1", "ETA5", "ETA00",
         "IETAPLB", "SXY0", "YNY0", "KKK", "GGG", "COHC", "COHT", "FRIC", "FRIT", "DILC", "TTT",
         "EIIB", "VSLIPB"]
STAG = ["pt", "vxs", "vys", "pf", "vxD", "vyD", "ETAB", "ETAB0", "ETAP", "ETAP0", "POR", "GGGP",
        "GGGB", "PTF0", "PT0", "PF0", "SXX0", "SYY0", "RHOX", "RHOFX", "ETADX", "PORX", "VX0",
        "VXF0", "RHOY", "RHOFY", "ETADY", "PORY", "VY0", "VYF0"]
SERIES = ["timesumcur", "dtcur", "maxvxsmod", "minvxsmod", "maxvysmod", "minvysmod"]
MARKERS = ["xm", "ym", "sxxm", "syym", "sxym"]
NT = 100000


Snapshot = hm.Snapshot


def h5lib():
    for n in ("libhdf5.so.103", "/opt/conda/lib/libhdf5.so.103"):
        try:
            L = C.CDLL(n)
            break
        except OSError:
            L = None
    if L is None:
        pytest.skip("HDF5 C library not found")
    L.H5open()
    L.H5Fopen.restype = C.c_int64
    L.H5Fopen.argtypes = [C.c_char_p, C.c_uint, C.c_int64]
    L.H5Gopen2.restype = C.c_int64
    L.H5Gopen2.argtypes = [C.c_int64, C.c_char_p, C.c_int64]
    L.H5Dopen2.restype = C.c_int64
    L.H5Dopen2.argtypes = [C.c_int64, C.c_char_p, C.c_int64]
    L.H5Dget_space.restype = C.c_int64
    L.H5Dget_space.argtypes = [C.c_int64]
    L.H5Sget_simple_extent_ndims.argtypes = [C.c_int64]
    L.H5Sget_simple_extent_dims.argtypes = [C.c_int64, C.POINTER(C.c_ulonglong), C.c_void_p]
    L.H5Dread.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_void_p]
    L.H5Dget_offset.restype = C.c_uint64
    L.H5Dget_offset.argtypes = [C.c_int64]
    L.H5Lget_name_by_idx.restype = C.c_ssize_t
    L.H5Lget_name_by_idx.argtypes = [C.c_int64, C.c_char_p, C.c_int, C.c_int, C.c_ulonglong,
                                     C.c_char_p, C.c_size_t, C.c_int64]
    for f in ("H5Dclose", "H5Sclose", "H5Gclose", "H5Fclose"):
        getattr(L, f).argtypes = [C.c_int64]
    return L


def names_in(L, f, group):
    out = []
    buf = C.create_string_buffer(256)
    for k in range(1000):
        n = L.H5Lget_name_by_idx(f, group.encode(), 0, 0, k, buf, 256, 0)
        if n < 0:
            break
        out.append(buf.value.decode())
    return out


def read_ds(L, g, name):
    d = L.H5Dopen2(g, name.encode(), 0)
    assert d >= 0, name
    sp = L.H5Dget_space(d)
    nd = L.H5Sget_simple_extent_ndims(sp)
    dims = (C.c_ulonglong * nd)()
    L.H5Sget_simple_extent_dims(sp, dims, None)
    shape = tuple(int(x) for x in dims)
    a = np.empty(shape, dtype=np.float64)
    nat = C.c_int64.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
    assert L.H5Dread(d, nat, 0, 0, 0, a.ctypes.data_as(C.c_void_p)) >= 0
    off = L.H5Dget_offset(d)
    L.H5Sclose(sp)
    L.H5Dclose(d)
    return a, off


def synthetic(Nx=21, Ny=17, M=300):
    rng = np.random.default_rng(1)
    keep = {}
    s = Snapshot()
    s.Nx, s.Ny, s.markers = Nx, Ny, M

    def arr(key, a):
        a = np.asfortranarray(a)                 # column-major like the fields API
        keep[key] = a
        return a.ctypes.data_as(C.POINTER(C.c_double))
    for k in range(24):
        s.basic[k] = arr(("b", k), 1000.0 * k + rng.normal(size=(Ny, Nx)))
    for k in range(30):
        s.stag[k] = arr(("s", k), -1000.0 * k + rng.normal(size=(Ny + 1, Nx + 1)))
    for k in range(6):
        s.series[k] = arr(("t", k), rng.normal(size=NT))
    for k in range(5):
        s.mk[k] = arr(("m", k), rng.normal(size=M))
    for k in range(9):
        s.values[k] = 0.5 + k
    return s, keep


def test_h5_names(hmec):
    L = hmec.lib()
    got = [[(L.hmec_h5_dataset_name(g, k) or b"").decode() for k in range(n)]
           for g, n in ((0, 24), (1, 30), (2, 6), (3, 5))]
    assert got == [BASIC, STAG, SERIES, MARKERS]
    assert L.hmec_h5_dataset_name(0, 24) is None
    for g, names in ((0, BASIC), (1, STAG)):
        for k, n in enumerate(names):
            assert L.hmec_h5_field_id(g, k) == hmec.FIELD_ID[n], n


def test_h5_write_layout(hmec, tmp_path):
    L = h5lib()
    s, keep = synthetic()
    fn = str(tmp_path / "h_mec_500.h5")
    assert hmec.lib().hmec_h5_write(fn.encode(), C.byref(s)) == 0
    f = L.H5Fopen(fn.encode(), 0, 0)
    assert f >= 0
    assert names_in(L, f, "/") == ["Matrix", "Value", "Vector"]
    assert sorted(names_in(L, f, "/Matrix")) == sorted(BASIC + STAG)
    assert sorted(names_in(L, f, "/Vector")) == sorted(SERIES + MARKERS)
    assert names_in(L, f, "/Value") == ["values"]
    g = L.H5Gopen2(f, b"Matrix", 0)
    offs = []
    for k, n in enumerate(BASIC):
        a, off = read_ds(L, g, n)
        assert a.shape == (17, 21), n
        # row-major: element [i][j] of the dataset is M(i, j)
        assert np.array_equal(a, keep[("b", k)]), n
        offs.append(off)
    for k, n in enumerate(STAG):
        a, off = read_ds(L, g, n)
        assert a.shape == (18, 22), n
        assert np.array_equal(a, keep[("s", k)]), n
        offs.append(off)
    L.H5Gclose(g)
    g = L.H5Gopen2(f, b"Vector", 0)
    for k, n in enumerate(SERIES):
        a, off = read_ds(L, g, n)
        assert a.shape == (NT,) and np.array_equal(a, keep[("t", k)]), n
        offs.append(off)
    for k, n in enumerate(MARKERS):
        a, off = read_ds(L, g, n)
        assert a.shape == (300,) and np.array_equal(a, keep[("m", k)]), n
        offs.append(off)
    L.H5Gclose(g)
    g = L.H5Gopen2(f, b"Value", 0)
    a, off = read_ds(L, g, "values")
    assert np.array_equal(a, 0.5 + np.arange(9))
    L.H5Gclose(g)
    L.H5Fclose(f)
    # the datasets were written in the reference's order: their data lies in
    # the file in that order (every dataset here is above HDF5's 2 KB
    # small-data threshold, under which data is packed apart; /Value/values
    # is below it)
    assert all(b > a for a, b in zip(offs, offs[1:]))


def test_h5_roundtrip_and_errors(hmec, tmp_path):
    s, keep = synthetic()
    fn = str(tmp_path / "snap.h5")
    lib = hmec.lib()
    assert lib.hmec_h5_write(fn.encode(), C.byref(s)) == 0
    r = Snapshot()
    r.Nx, r.Ny, r.markers = 21, 17, 300
    out = {}

    def arr(key, shape):
        out[key] = np.zeros(shape, order="F")
        return out[key].ctypes.data_as(C.POINTER(C.c_double))
    for k in range(24):
        r.basic[k] = arr(("b", k), (17, 21))
    for k in range(30):
        r.stag[k] = arr(("s", k), (18, 22))
    for k in range(6):
        r.series[k] = arr(("t", k), NT)
    for k in range(5):
        r.mk[k] = arr(("m", k), 300)
    assert lib.hmec_h5_read(fn.encode(), C.byref(r)) == 0
    for key in keep:
        assert np.array_equal(out[key], keep[key]), key
    assert list(r.values) == list(0.5 + np.arange(9))
    # a grid of another size does not read
    r.Nx = 22
    assert lib.hmec_h5_read(fn.encode(), C.byref(r)) == hm.E_IO
    r.Nx = 21
    r.markers = 301
    assert lib.hmec_h5_read(fn.encode(), C.byref(r)) == hm.E_IO
    # an unwritable path
    assert lib.hmec_h5_write(str(tmp_path / "no" / "x.h5").encode(), C.byref(s)) == hm.E_IO


EVO = ["EVO_Vslip.txt", "EVO_viscosity.txt", "EVO_press_flu.txt", "EVO_press_eff.txt",
       "EVO_SigmaY.txt", "EVO_Sii.txt", "EVO_Theta.txt", "EVO_Visc_comp.txt", "EVO_Elast_comp.txt",
       "EVO_vxD.txt", "EVO_data.txt"]


def test_trim_evo(hmec, tmp_path):
    """Restart at timestep 6 with streams of 10 lines: every stream of the
    reference's list is cut to 5 lines; the flux stream, which the reference
    lists as EVO_VxD.txt but writes as EVO_vxD.txt, keeps its 10 lines and an
    EVO_VxD.txt of 5 empty lines appears.  Streams shorter than the restart
    step are left alone."""
    for f in EVO:
        (tmp_path / f).write_text("".join("%s line %d\n" % (f, k) for k in range(10)))
    lib = hmec.lib()
    assert lib.hmec_trim_evo(str(tmp_path).encode(), 11) == 0      # 10 lines < 11: no change
    for f in EVO:
        assert len((tmp_path / f).read_text().splitlines()) == 10
    assert lib.hmec_trim_evo(str(tmp_path).encode(), 6) == 0
    for f in EVO:
        lines = (tmp_path / f).read_text().splitlines()
        if f == "EVO_vxD.txt":
            assert len(lines) == 10
        else:
            assert lines == ["%s line %d" % (f, k) for k in range(5)], f
    assert (tmp_path / "EVO_VxD.txt").read_text() == "\n" * 5


def test_oracle_series(hmec, orc):
    """The per-step series entry the oracle writes at the end of step t:
    timesum (after += dt), dt, and the extrema of vxs over i < Ny, j < Nx with
    RHOX > 2000 and i != 0, of vys with RHOY > 2000 and j != 0, computed here
    with numpy masks."""
    p = hmec.params(1, Nx=13, Ny=9, xsize=6000., ysize=4000., upper_block=1800., lower_block=2200.)
    o = orc.Oracle(p, hmec.generate_markers(p, cfg=1, seed=1), hmec)
    o.step()
    o.step()
    rng = np.random.default_rng(4)
    S = (p.Ny + 1, p.Nx + 1)
    vxs, vys = rng.normal(size=S), rng.normal(size=S)
    rhox, rhoy = rng.uniform(1500, 2600, S), rng.uniform(1500, 2600, S)
    vxs[0, :] = 50.0          # row i = 0 is excluded for vx
    vys[:, 0] = -50.0         # column j = 0 is excluded for vy
    vxs[p.Ny, :] = 60.0       # the last row / column are outside the loops
    vys[:, p.Nx] = -60.0
    for n, a in (("vxs", vxs), ("vys", vys), ("RHOX", rhox), ("RHOY", rhoy)):
        o.set(n, a)
    st = o.state()
    st.timestep, st.timesum = 3, 7.0
    o.set_state(st)
    o.dt = 2.5
    o.step_end()
    ser = o.series(3)
    ii = np.arange(p.Ny)[:, None]
    jj = np.arange(p.Nx)[None, :]
    mx = (rhox[:p.Ny, :p.Nx] > 2000) & (ii != 0)
    my = (rhoy[:p.Ny, :p.Nx] > 2000) & (jj != 0)
    assert ser[0, 2] == 9.5 and ser[1, 2] == 2.5
    assert ser[2, 2] == vxs[:p.Ny, :p.Nx][mx].max() and ser[3, 2] == vxs[:p.Ny, :p.Nx][mx].min()
    assert ser[4, 2] == vys[:p.Ny, :p.Nx][my].max() and ser[5, 2] == vys[:p.Ny, :p.Nx][my].min()
    # the first two steps' entries were written by the steps
    assert ser[0, 0] > 0 and ser[0, 1] > ser[0, 0]
