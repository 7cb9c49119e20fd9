"""ORACLE — test infrastructure only (tests/, smoke(), bench.py cpu_baseline).

ctypes wrapper of oracle/_build/libhmec_oracle.so, the single-threaded
IEEE-strict C restatement of cpp/h_mec_rsf_v9.cpp:188-1726 (see the header of
oracle/hmec_oracle.c).  The structs (Params, StepInfo, IterScalars) are the
C-ABI mirrors from the package's hmec.py, passed in by the caller's module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libhmec_oracle.so")

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp = C.c_void_p
        P = C.POINTER
        L.orc_create.restype = vp
        L.orc_create.argtypes = [vp, P(P(C.c_double)), C.c_int64]
        L.orc_destroy.argtypes = [vp]
        L.orc_field.restype = P(C.c_double)
        L.orc_field.argtypes = [vp, C.c_int]
        L.orc_marker.restype = P(C.c_double)
        L.orc_marker.argtypes = [vp, C.c_int]
        L.orc_accumulator.restype = P(C.c_double)
        L.orc_accumulator.argtypes = [vp, C.c_int]
        L.orc_use_pardiso.argtypes = [vp, C.c_int]
        L.orc_have_pardiso.restype = C.c_int
        L.orc_marker_cell.argtypes = [vp, C.c_int, C.c_int64, P(C.c_int), P(C.c_int)]
        L.orc_fix.argtypes = [C.c_double]
        L.orc_fix.restype = C.c_int
        L.orc_check_bounds.argtypes = [C.c_int, C.c_int]
        L.orc_check_bounds.restype = C.c_int
        L.orc_divplus.argtypes = [C.c_double, C.c_double]
        L.orc_divplus.restype = C.c_double
        L.orc_linspaced.argtypes = [P(C.c_double), C.c_int, C.c_double, C.c_double]
        for f in ("orc_scatter", "orc_step_begin", "orc_vep_pre", "orc_assemble", "orc_reload",
                  "orc_advect", "orc_step_end"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = None
        L.orc_scale.argtypes = [vp]
        L.orc_scale.restype = C.c_double
        L.orc_solve.argtypes = [vp]
        L.orc_solve.restype = C.c_int
        L.orc_solution.argtypes = [vp]
        L.orc_solution.restype = P(C.c_double)
        L.orc_csr.argtypes = [vp, P(P(C.c_int)), P(P(C.c_int)), P(P(C.c_double)),
                              P(P(C.c_double)), P(C.c_int), P(C.c_int)]
        L.orc_triplet_count.argtypes = [vp]
        L.orc_triplet_count.restype = C.c_int64
        L.orc_triplets.argtypes = [vp, P(C.c_int), P(C.c_int), P(C.c_double)]
        L.orc_update.argtypes = [vp, vp]
        L.orc_iter_commit.argtypes = [vp, C.c_int, C.c_int]
        L.orc_step.argtypes = [vp, vp]
        L.orc_step.restype = C.c_int
        L.orc_get_state.argtypes = [vp, vp]
        L.orc_set_state.argtypes = [vp, vp]
        L.orc_set_dt.argtypes = [vp, C.c_double]
        L.orc_get_dt.argtypes = [vp]
        L.orc_get_dt.restype = C.c_double
        L.orc_lu_stats.argtypes = [vp, P(C.c_int64), P(C.c_double)]
        L.orc_resolve.argtypes = [vp, P(C.c_double), P(C.c_double)]
        L.orc_resolve.restype = C.c_int
        L.orc_series.argtypes = [vp]
        L.orc_series.restype = P(C.c_double)
        L.orc_controller.argtypes = [vp, vp, C.c_int]
        L.orc_controller.restype = C.c_int
        L.orc_dsylsq.argtypes = [vp]
        L.orc_dsylsq.restype = P(C.c_double)
        _lib = L
    return _lib


def have_pardiso():
    return bool(lib().orc_have_pardiso())


class Oracle:
    """CPU restatement state for one simulation (column-major fields)."""

    def __init__(self, p, soa, hm, use_pardiso=True):
        self.p = p
        self.hm = hm            # the package's hmec module (struct mirrors)
        soa = np.ascontiguousarray(soa, dtype=np.float64)
        arr = (C.POINTER(C.c_double) * 19)(*[soa[f].ctypes.data_as(C.POINTER(C.c_double))
                                              for f in range(19)])
        self.h = lib().orc_create(C.cast(C.byref(p), C.c_void_p), arr, soa.shape[1])
        self.M = soa.shape[1]
        lib().orc_use_pardiso(self.h, int(bool(use_pardiso)))

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shape(self, name):
        return self.hm.field_shape(self.p, name)

    def get(self, name):
        r, c = self._shape(name)
        ptr = lib().orc_field(self.h, self.hm.FIELD_ID[name])
        a = np.ctypeslib.as_array(ptr, shape=(c * r,)).copy()
        return a.reshape(c, r).T

    def set(self, name, arr):
        r, c = self._shape(name)
        ptr = lib().orc_field(self.h, self.hm.FIELD_ID[name])
        dst = np.ctypeslib.as_array(ptr, shape=(c * r,))
        dst[:] = np.asarray(arr, dtype=np.float64).T.reshape(-1)

    def marker(self, f):
        ptr = lib().orc_marker(self.h, f)
        return np.ctypeslib.as_array(ptr, shape=(self.M,)).copy()

    def marker_cells(self):
        out = np.zeros((4, self.M), dtype=np.int32)
        ci = C.c_int()
        cj = C.c_int()
        Ny1 = self.p.Ny + 1
        for g in range(4):
            for m in range(self.M):
                lib().orc_marker_cell(self.h, g, m, C.byref(ci), C.byref(cj))
                out[g, m] = ci.value + cj.value * Ny1
        return out

    def scatter(self):
        lib().orc_scatter(self.h)

    def step_begin(self):
        lib().orc_step_begin(self.h)

    def vep_pre(self):
        lib().orc_vep_pre(self.h)

    def scale(self):
        return lib().orc_scale(self.h)

    def assemble(self):
        lib().orc_assemble(self.h)

    def csr(self):
        rp = C.POINTER(C.c_int)()
        ci = C.POINTER(C.c_int)()
        v = C.POINTER(C.c_double)()
        b = C.POINTER(C.c_double)()
        n = C.c_int()
        nnz = C.c_int()
        lib().orc_csr(self.h, C.byref(rp), C.byref(ci), C.byref(v), C.byref(b), C.byref(n), C.byref(nnz))
        n, nnz = n.value, nnz.value
        return (np.ctypeslib.as_array(rp, shape=(n + 1,)).copy(),
                np.ctypeslib.as_array(ci, shape=(nnz,)).copy(),
                np.ctypeslib.as_array(v, shape=(nnz,)).copy(),
                np.ctypeslib.as_array(b, shape=(n,)).copy())

    def triplets(self):
        nt = lib().orc_triplet_count(self.h)
        ti = np.zeros(nt, dtype=np.int32)
        tj = np.zeros(nt, dtype=np.int32)
        tv = np.zeros(nt, dtype=np.float64)
        lib().orc_triplets(self.h, ti.ctypes.data_as(C.POINTER(C.c_int)),
                           tj.ctypes.data_as(C.POINTER(C.c_int)),
                           tv.ctypes.data_as(C.POINTER(C.c_double)))
        return ti, tj, tv

    def solve(self):
        return lib().orc_solve(self.h)

    def solution(self):
        n = (self.p.Nx + 1) * (self.p.Ny + 1) * 6
        return np.ctypeslib.as_array(lib().orc_solution(self.h), shape=(n,)).copy()

    def series(self, n):
        """The six per-step series (timesumcur, dtcur, maxvxsmod, minvxsmod,
        maxvysmod, minvysmod), first n entries: array (6, n)."""
        a = np.ctypeslib.as_array(lib().orc_series(self.h), shape=(6 * 100000,))
        return a.reshape(6, 100000)[:, :n].copy()

    def resolve(self, b):
        """A^-1 b with PARDISO's factor of the last solve (None without it)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        rc = lib().orc_resolve(self.h, b.ctypes.data_as(C.POINTER(C.c_double)),
                               x.ctypes.data_as(C.POINTER(C.c_double)))
        return x if rc == 0 else None

    def reload(self):
        lib().orc_reload(self.h)

    def update(self):
        sc = self.hm.IterScalars()
        lib().orc_update(self.h, C.cast(C.byref(sc), C.c_void_p))
        return sc

    def iter_commit(self, ynpl, converged):
        lib().orc_iter_commit(self.h, int(ynpl), int(bool(converged)))

    def advect(self):
        lib().orc_advect(self.h)

    def step_end(self):
        lib().orc_step_end(self.h)

    def step(self):
        s = self.hm.StepInfo()
        rc = lib().orc_step(self.h, C.cast(C.byref(s), C.c_void_p))
        if rc != 0:
            raise RuntimeError(f"oracle solver error {rc}")
        return s

    def state(self):
        s = self.hm.StepInfo()
        lib().orc_get_state(self.h, C.cast(C.byref(s), C.c_void_p))
        return s

    def set_state(self, s):
        lib().orc_set_state(self.h, C.cast(C.byref(s), C.c_void_p))

    @property
    def dt(self):
        return lib().orc_get_dt(self.h)

    @dt.setter
    def dt(self, v):
        lib().orc_set_dt(self.h, float(v))

    def controller(self, sc, iterstep):
        """The Picard / dt controller on constructed scalars; True = exit."""
        return bool(lib().orc_controller(self.h, C.cast(C.byref(sc), C.c_void_p), int(iterstep)))

    def dsylsq(self, n):
        return np.ctypeslib.as_array(lib().orc_dsylsq(self.h), shape=(n,)).copy()

    def lu_stats(self):
        nnz = C.c_int64()
        mf = C.c_double()
        lib().orc_lu_stats(self.h, C.byref(nnz), C.byref(mf))
        return nnz.value, mf.value
