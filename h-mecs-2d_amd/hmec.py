"""Python host binding of libhmec_hip (ctypes over the C ABI in include/hmec.h).

The reference's host program is C++ (cpp/h_mec_rsf_v9.cpp main()); the C++
driver of this package lives in csrc/hmec_driver.cpp behind the C ABI.  This
module only marshals numpy arrays for tests, the benchmark and the serving
loop.  Every compute call goes to the HIP library; there is no Python or CPU
fallback: loading fails loudly when the library is missing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhmec_hip.so")

NMARKER = 19
MARKER_FIELDS = ["xm", "ym", "porm", "kkkm", "t_marker", "cohescm", "cohestm", "frictcm",
                 "dilatcm", "fricttm", "etasm", "etam", "rhom", "rhofm", "etafm", "gm",
                 "sxxm", "syym", "sxym"]
BASIC_FIELDS = ["OM0", "OM", "OM5", "ARSF", "BRSF", "LRSF", "RHO", "ETA", "ETA0", "ETA1",
                "ETA5", "ETA00", "ETA50", "IETAPLB", "SXY", "SXY0", "YNY0", "YNY", "KKK",
                "GGG", "COHC", "COHT", "FRIC", "FRIT", "DILC", "TTT", "EIIB", "VSLIPB",
                "ESP", "EXY", "DSY", "SigmaY", "SII_fault", "SIIB"]
STAG_FIELDS = ["pt", "vxs", "vys", "pf", "vxD", "vyD", "ETAB", "ETAB0", "ETAP", "ETAP0",
               "POR", "GGGP", "GGGB", "PTF0", "PT0", "PF0", "SXX", "SXX0", "SYY", "SYY0",
               "DILP", "RHOX", "RHOFX", "ETADX", "PORX", "VX0", "VXF0", "RHOY", "RHOFY",
               "ETADY", "PORY", "VY0", "VYF0", "EXX", "EYY",
               "EII", "EIIVP", "SII", "DSII", "DIS", "pt_ave", "pf_ave", "VIS_COMP", "EL_DECOM"]
FIELDS = BASIC_FIELDS + STAG_FIELDS
FIELD_ID = {n: k for k, n in enumerate(FIELDS)}
NBASIC = len(BASIC_FIELDS)

PREC_FP64 = 0
PREC_FP32 = 1


class Params(C.Structure):
    """Mirror of hmec_params (include/hmec.h), field order must match."""
    _fields_ = [
        ("xsize", C.c_double), ("ysize", C.c_double), ("Nx", C.c_int), ("Ny", C.c_int),
        ("mx_per_cell", C.c_int), ("my_per_cell", C.c_int),
        ("TS_1", C.c_double), ("TS_2", C.c_double), ("TS_3", C.c_double), ("TS_4", C.c_double),
        ("arsfm", C.c_double * 2), ("brsfm", C.c_double * 2), ("lrsfm", C.c_double * 2),
        ("omm", C.c_double * 2),
        ("V0", C.c_double), ("alpha", C.c_double), ("POR0", C.c_double),
        ("cohes", C.c_double), ("friction", C.c_double), ("dilatation", C.c_double),
        ("tensile", C.c_double), ("shearmod", C.c_double),
        ("BETAFLUID", C.c_double), ("BETASOLID", C.c_double),
        ("gx", C.c_double), ("gy", C.c_double), ("PCONF", C.c_double), ("PTFDIFF", C.c_double),
        ("etamin", C.c_double), ("etamax", C.c_double), ("kkkmin", C.c_double),
        ("kkkmax", C.c_double), ("stpmax", C.c_double), ("stpmax1", C.c_double),
        ("bcupper", C.c_double), ("bclower", C.c_double), ("bcvyflower", C.c_double),
        ("niterglobal", C.c_int), ("ynlastmax", C.c_int), ("dtstep", C.c_int),
        ("vratiomax", C.c_double), ("dtkoef", C.c_double), ("dtkoefup", C.c_double),
        ("dtkoefv", C.c_double), ("errmin", C.c_double), ("etawt", C.c_double),
        ("syieldmin", C.c_double), ("tyield", C.c_double),
        ("dtelastic0", C.c_double), ("dtmin", C.c_double), ("inertia", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double), ("faultwidth", C.c_double),
        ("upper_block", C.c_double), ("lower_block", C.c_double),
        ("line_fault", C.c_int), ("advect_markers", C.c_int), ("bisect_cap", C.c_int),
        ("block_override", C.c_int),
    ]


class StepInfo(C.Structure):
    """Mirror of hmec_step_info."""
    _fields_ = [
        ("timestep", C.c_int), ("iterations", C.c_int), ("ynlast", C.c_int),
        ("yndtdecrease", C.c_int),
        ("dt", C.c_double), ("dt00", C.c_double), ("dtx", C.c_double), ("dty", C.c_double),
        ("dtlapusta", C.c_double), ("Vmax", C.c_double), ("maxvxy", C.c_double),
        ("timesum", C.c_double), ("ptscale", C.c_double),
        ("bisect_capped", C.c_int),
        ("factorizations", C.c_int), ("solves", C.c_int), ("ir_iterations", C.c_int),
        ("fallbacks", C.c_int), ("perturbed_pivots", C.c_int),
        ("max_rel_residual", C.c_double),
        ("ms_scatter", C.c_double), ("ms_assemble", C.c_double), ("ms_factor", C.c_double),
        ("ms_solve", C.c_double), ("ms_update", C.c_double), ("ms_total", C.c_double),
    ]

    def as_dict(self):
        return {k: (getattr(self, k)) for k, _ in self._fields_}


class IterScalars(C.Structure):
    """Mirror of hmec_iter_scalars."""
    _fields_ = [
        ("ynpl", C.c_int), ("ddd", C.c_double), ("dtlapusta", C.c_double), ("Vmax", C.c_double),
        ("vxs_max", C.c_double), ("vxs_min", C.c_double), ("vys_max", C.c_double),
        ("vys_min", C.c_double), ("maxvxs_band", C.c_double), ("maxvys_band", C.c_double),
        ("dtslip", C.c_double), ("bisect_capped", C.c_int),
    ]

    def as_dict(self):
        return {k: (getattr(self, k)) for k, _ in self._fields_}


class SolveStats(C.Structure):
    """Mirror of hmec_solve_stats."""
    _fields_ = [
        ("ir_iterations", C.c_int), ("fallback", C.c_int), ("perturbed_pivots", C.c_int),
        ("rel_residual", C.c_double), ("factor_flops", C.c_double), ("factor_nnz", C.c_int64),
        ("nfronts", C.c_int), ("nlevels", C.c_int), ("max_front", C.c_int),
    ]

    def as_dict(self):
        return {k: (getattr(self, k)) for k, _ in self._fields_}


_lib = None


class Snapshot(C.Structure):
    """Mirror of hmec_snapshot (host arrays of an HDF5 snapshot)."""
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("markers", C.c_int64),
                ("basic", C.POINTER(C.c_double) * 24), ("stag", C.POINTER(C.c_double) * 30),
                ("series", C.POINTER(C.c_double) * 6), ("mk", C.POINTER(C.c_double) * 5),
                ("values", C.c_double * 9)]


SERIES_LEN = 100000


class ServerConfig(C.Structure):
    """hmec_server_config (include/hmec.h)."""
    _fields_ = [("device", C.c_int), ("precision", C.c_int), ("max_active", C.c_int),
                ("hbm_budget", C.c_int64), ("batch_window_ms", C.c_double), ("keep_fields", C.c_int)]


class Request(C.Structure):
    """hmec_request (include/hmec.h)."""
    _fields_ = [("params", Params), ("cfg", C.c_int), ("seed", C.c_uint64), ("jitter", C.c_int),
                ("nsteps", C.c_int), ("arrival_ms", C.c_double)]


class MemberResult(C.Structure):
    """hmec_member_result (include/hmec.h)."""
    _fields_ = [("status", C.c_int), ("steps", C.c_int), ("admit_ms", C.c_double),
                ("finish_ms", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_ms", C.c_double), ("footprint", C.c_int64), ("factorizations", C.c_int),
                ("batched_factorizations", C.c_int), ("ir_iterations", C.c_int),
                ("fallbacks", C.c_int), ("max_rel_residual", C.c_double), ("timesum", C.c_double),
                ("dt", C.c_double)]


class ServerStats(C.Structure):
    """hmec_server_stats (include/hmec.h)."""
    _fields_ = [("member_steps", C.c_int64), ("wall_ms", C.c_double),
                ("member_steps_per_s", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("members_done", C.c_int), ("members_failed", C.c_int), ("max_concurrent", C.c_int),
                ("peak_bytes", C.c_int64), ("factor_batches", C.c_int),
                ("factor_batch_members", C.c_int)]


# error codes (include/hmec.h)
E_ARG, E_HIP, E_OOM, E_PIVOT, E_NOCONV, E_NODEV, E_SOLVER, E_IO = -1, -2, -3, -4, -5, -6, -7, -8


class HmecError(RuntimeError):
    """A negative HMEC_E_* return code; .code holds it."""

    def __init__(self, msg, code=0):
        super().__init__(msg)
        self.code = code


def lib():
    """Load libhmec_hip.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HmecError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


class _Tolerant:
    """Attribute proxy: declarations of symbols the library lacks are skipped
    (tests/test_abi.py checks that every symbol of include/hmec.h exists)."""

    def __init__(self, L):
        object.__setattr__(self, "_L", L)

    def __getattr__(self, name):
        L = object.__getattribute__(self, "_L")
        try:
            return getattr(L, name)
        except AttributeError:
            return _Sink()


class _Sink:
    def __setattr__(self, k, v):
        pass


def _declare(L):
    L = _Tolerant(L)
    P = C.POINTER
    vp = C.c_void_p
    L.hmec_strerror.restype = C.c_char_p
    L.hmec_strerror.argtypes = [C.c_int]
    L.hmec_params_default.argtypes = [P(Params)]
    L.hmec_params_finalize.argtypes = [P(Params)]
    L.hmec_params_config.argtypes = [C.c_int, P(Params)]
    L.hmec_marker_count.restype = C.c_int64
    L.hmec_marker_count.argtypes = [P(Params)]
    L.hmec_generate_markers.argtypes = [P(Params), C.c_int, C.c_uint64, C.c_int, P(P(C.c_double))]
    L.hmec_field_shape.argtypes = [P(Params), C.c_int, P(C.c_int), P(C.c_int)]
    L.hmec_params_size.restype = C.c_int
    L.hmec_step_info_size.restype = C.c_int
    L.hmec_create.argtypes = [P(Params), P(vp)]
    L.hmec_destroy.argtypes = [vp]
    L.hmec_destroy.restype = None
    L.hmec_footprint_bytes.restype = C.c_int64
    L.hmec_footprint_bytes.argtypes = [P(Params)]
    L.hmec_set_precision.argtypes = [vp, C.c_int]
    L.hmec_set_refinement.argtypes = [vp, C.c_int, C.c_int, C.c_double]
    L.hmec_set_fp64_refinement.argtypes = [vp, C.c_int, C.c_double]
    L.hmec_solve_noalloc.argtypes = [vp, P(SolveStats)]
    L.hmec_last_error.restype = C.c_char_p
    L.hmec_last_error.argtypes = []
    for f in ("hmec_write_h5", "hmec_read_h5", "hmec_save", "hmec_restart"):
        getattr(L, f).argtypes = [vp, C.c_char_p]
    L.hmec_trim_evo.argtypes = [C.c_char_p, C.c_int]
    L.hmec_get_series.argtypes = [vp, P(C.c_double)]
    L.hmec_h5_dataset_name.restype = C.c_char_p
    L.hmec_h5_dataset_name.argtypes = [C.c_int, C.c_int]
    L.hmec_h5_field_id.argtypes = [C.c_int, C.c_int]
    L.hmec_h5_write.argtypes = [C.c_char_p, P(Snapshot)]
    L.hmec_h5_read.argtypes = [C.c_char_p, P(Snapshot)]
    L.hmec_set_factor_cache.argtypes = [vp, C.c_int]
    L.hmec_init_coseismic.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_uint64]
    L.hmec_set_markers.argtypes = [vp, P(P(C.c_double)), C.c_int64]
    L.hmec_get_markers.argtypes = [vp, P(P(C.c_double))]
    L.hmec_get_marker_cells.argtypes = [vp, P(C.c_int32)]
    L.hmec_set_field.argtypes = [vp, C.c_int, P(C.c_double), C.c_int]
    L.hmec_get_field.argtypes = [vp, C.c_int, P(C.c_double), C.c_int]
    L.hmec_write_evo.argtypes = [vp, C.c_char_p, C.c_int]
    L.hmec_get_state.argtypes = [vp, P(StepInfo)]
    L.hmec_set_state.argtypes = [vp, P(StepInfo)]
    L.hmec_scatter.argtypes = [vp]
    L.hmec_step_begin.argtypes = [vp]
    L.hmec_vep_pre.argtypes = [vp, C.c_double]
    L.hmec_assemble.argtypes = [vp, C.c_double, P(C.c_double)]
    L.hmec_csr_dims.argtypes = [vp, P(C.c_int64), P(C.c_int64)]
    L.hmec_get_csr.argtypes = [vp, P(C.c_int32), P(C.c_int32), P(C.c_double), P(C.c_double)]
    L.hmec_set_csr_values.argtypes = [vp, P(C.c_double), P(C.c_double)]
    L.hmec_analyze.argtypes = [vp]
    L.hmec_factorize.argtypes = [vp, C.c_int]
    L.hmec_solve.argtypes = [vp, P(SolveStats)]
    L.hmec_get_solution.argtypes = [vp, P(C.c_double)]
    L.hmec_update.argtypes = [vp, C.c_double, C.c_double, P(IterScalars)]
    L.hmec_timing_select.argtypes = [vp, C.c_char_p]
    L.hmec_timing_sample.argtypes = [vp, C.c_int]
    L.hmec_controller.argtypes = [vp, P(IterScalars), C.c_int, P(C.c_int)]
    L.hmec_iter_commit.argtypes = [vp, C.c_int, C.c_int]
    L.hmec_step_end.argtypes = [vp, C.c_double]
    L.hmec_advect.argtypes = [vp, C.c_double]
    L.hmec_step.argtypes = [vp, P(StepInfo)]
    L.hmec_batch_factorize.argtypes = [P(vp), C.c_int, C.c_int]
    L.hmec_batch_step.argtypes = [P(vp), C.c_int, P(StepInfo)]
    L.hmec_condensed_pattern.argtypes = [P(Params), P(C.c_int64), P(C.c_int64), P(C.c_int32),
                                         P(C.c_int32)]
    L.hmec_member_footprint.restype = C.c_int64
    L.hmec_member_footprint.argtypes = [P(Params), C.c_int]
    L.hmec_server_create.argtypes = [P(ServerConfig), P(vp)]
    L.hmec_server_destroy.argtypes = [vp]
    L.hmec_server_destroy.restype = None
    L.hmec_server_submit.argtypes = [vp, P(Request), P(C.c_int)]
    L.hmec_server_run.argtypes = [vp]
    L.hmec_server_result.argtypes = [vp, C.c_int, P(MemberResult)]
    L.hmec_server_get_stats.argtypes = [vp, P(ServerStats)]
    L.hmec_server_latencies.argtypes = [vp, C.c_int, P(C.c_double), C.c_int]
    L.hmec_server_get_field.argtypes = [vp, C.c_int, C.c_int, P(C.c_double), C.c_int]
    L.hmec_kernel_time.argtypes = [vp, C.c_char_p, P(C.c_double), P(C.c_double)]
    L.hmec_timing_enable.argtypes = [vp, C.c_int]
    L.hmec_timing_reset.argtypes = [vp]


def check(rc, what=""):
    if rc != 0:
        msg = lib().hmec_strerror(rc).decode()
        raise HmecError(f"{what}: {msg} ({rc})", rc)
    return rc


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def params(cfg=None, **over):
    """Preset (cfg 1..5) or constants.hpp defaults, then overrides."""
    p = Params()
    if cfg is None:
        lib().hmec_params_default(C.byref(p))
    else:
        check(lib().hmec_params_config(int(cfg), C.byref(p)), "params_config")
    block = False
    for k, v in over.items():
        if k in ("arsfm", "brsfm", "lrsfm", "omm"):
            getattr(p, k)[0] = v[0]
            getattr(p, k)[1] = v[1]
        else:
            setattr(p, k, v)
        if k in ("upper_block", "lower_block"):
            block = True
    if block:
        p.block_override = 1
    check(lib().hmec_params_finalize(C.byref(p)), "params_finalize")
    return p


def marker_count(p):
    return int(lib().hmec_marker_count(C.byref(p)))


def generate_markers(p, cfg=1, seed=1, jitter=True):
    """(19, M) float64 marker SoA from the seeded host generator."""
    M = marker_count(p)
    soa = np.zeros((NMARKER, M), dtype=np.float64)
    arr = (C.POINTER(C.c_double) * NMARKER)(*[dptr(soa[f]) for f in range(NMARKER)])
    check(lib().hmec_generate_markers(C.byref(p), int(cfg), C.c_uint64(seed), int(bool(jitter)), arr),
          "generate_markers")
    return soa


def field_shape(p, name):
    fid = FIELD_ID[name]
    return (p.Ny, p.Nx) if fid < NBASIC else (p.Ny + 1, p.Nx + 1)


class Context:
    """One simulation on the current HIP device (owns device state + stream)."""

    def __init__(self, p, markers=None, precision=PREC_FP64):
        self.p = p
        h = C.c_void_p()
        check(lib().hmec_create(C.byref(p), C.byref(h)), "hmec_create")
        self.h = h
        self.set_precision(precision)
        if markers is not None:
            self.set_markers(markers)

    def close(self):
        if self.h:
            lib().hmec_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, prec):
        check(lib().hmec_set_precision(self.h, int(prec)), "set_precision")

    def set_refinement(self, max_stale=12, max_fresh=40, tol=1e-12):
        check(lib().hmec_set_refinement(self.h, int(max_stale), int(max_fresh), float(tol)),
              "set_refinement")

    def write_h5(self, path):
        """Snapshot in the reference's restart layout (hmec_write_h5)."""
        check(lib().hmec_write_h5(self.h, os.fsencode(path)), "write_h5")

    def read_h5(self, path):
        check(lib().hmec_read_h5(self.h, os.fsencode(path)), "read_h5")

    def save(self, directory):
        """file.txt + h_mec_<timestep>.h5 of the step just completed."""
        check(lib().hmec_save(self.h, os.fsencode(directory)), "save")

    def restart(self, directory):
        check(lib().hmec_restart(self.h, os.fsencode(directory)), "restart")

    def series(self):
        """The six per-step series, array (6, SERIES_LEN)."""
        out = np.zeros((6, SERIES_LEN), dtype=np.float64)
        check(lib().hmec_get_series(self.h, dptr(out)), "get_series")
        return out

    def set_fp64_refinement(self, max_ir=3, tol=1e-14):
        check(lib().hmec_set_fp64_refinement(self.h, int(max_ir), float(tol)), "set_fp64_refinement")

    def set_factor_cache(self, slots=3):
        check(lib().hmec_set_factor_cache(self.h, int(slots)), "set_factor_cache")

    def init_coseismic(self, frac=0.98, noise=0.02, pf_ratio=0.7, seed=4):
        check(lib().hmec_init_coseismic(self.h, float(frac), float(noise), float(pf_ratio),
                                        C.c_uint64(seed)), "init_coseismic")

    def set_markers(self, soa):
        soa = np.ascontiguousarray(soa, dtype=np.float64)
        self._soa_keep = soa
        arr = (C.POINTER(C.c_double) * NMARKER)(*[dptr(soa[f]) for f in range(NMARKER)])
        check(lib().hmec_set_markers(self.h, arr, soa.shape[1]), "set_markers")

    def get_markers(self):
        M = marker_count(self.p)
        soa = np.zeros((NMARKER, M), dtype=np.float64)
        arr = (C.POINTER(C.c_double) * NMARKER)(*[dptr(soa[f]) for f in range(NMARKER)])
        check(lib().hmec_get_markers(self.h, arr), "get_markers")
        return soa

    def marker_cells(self):
        M = marker_count(self.p)
        out = np.zeros((4, M), dtype=np.int32)
        check(lib().hmec_get_marker_cells(self.h, out.ctypes.data_as(C.POINTER(C.c_int32))),
              "get_marker_cells")
        return out

    def get(self, name):
        r, c = field_shape(self.p, name)
        a = np.zeros((c, r), dtype=np.float64)   # column-major storage
        check(lib().hmec_get_field(self.h, FIELD_ID[name], dptr(a), 0), "get_field " + name)
        return a.T

    def write_evo(self, directory, first=False):
        """Fault-row EVO_*.txt streams of the last timestep (appended), plus
        x_fault.txt / rsf_fault.txt when first (h_mec_rsf_v9.cpp:1739-1805)."""
        check(lib().hmec_write_evo(self.h, os.fsencode(str(directory)), int(bool(first))), "write_evo")

    def set(self, name, arr):
        r, c = field_shape(self.p, name)
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).T)
        assert a.shape == (c, r)
        check(lib().hmec_set_field(self.h, FIELD_ID[name], dptr(a), 0), "set_field " + name)

    def state(self):
        s = StepInfo()
        check(lib().hmec_get_state(self.h, C.byref(s)), "get_state")
        return s

    def set_state(self, s):
        check(lib().hmec_set_state(self.h, C.byref(s)), "set_state")

    def scatter(self):
        check(lib().hmec_scatter(self.h), "scatter")

    def step_begin(self):
        check(lib().hmec_step_begin(self.h), "step_begin")

    def vep_pre(self, dt):
        check(lib().hmec_vep_pre(self.h, float(dt)), "vep_pre")

    def assemble(self, dt):
        s = C.c_double()
        check(lib().hmec_assemble(self.h, float(dt), C.byref(s)), "assemble")
        return s.value

    def csr(self):
        n = C.c_int64()
        nnz = C.c_int64()
        check(lib().hmec_csr_dims(self.h, C.byref(n), C.byref(nnz)), "csr_dims")
        rp = np.zeros(n.value + 1, dtype=np.int32)
        ci = np.zeros(nnz.value, dtype=np.int32)
        v = np.zeros(nnz.value, dtype=np.float64)
        b = np.zeros(n.value, dtype=np.float64)
        check(lib().hmec_get_csr(self.h, rp.ctypes.data_as(C.POINTER(C.c_int32)),
                                 ci.ctypes.data_as(C.POINTER(C.c_int32)), dptr(v), dptr(b)), "get_csr")
        return rp, ci, v, b

    def set_csr_values(self, values, rhs):
        values = np.ascontiguousarray(values, dtype=np.float64)
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        check(lib().hmec_set_csr_values(self.h, dptr(values), dptr(rhs)), "set_csr_values")

    def analyze(self):
        check(lib().hmec_analyze(self.h), "analyze")

    def factorize(self, prec=PREC_FP64):
        check(lib().hmec_factorize(self.h, int(prec)), "factorize")

    def solve(self):
        st = SolveStats()
        rc = lib().hmec_solve(self.h, C.byref(st))
        check(rc, "solve")
        return st

    def solution(self):
        n = C.c_int64()
        nnz = C.c_int64()
        check(lib().hmec_csr_dims(self.h, C.byref(n), C.byref(nnz)), "csr_dims")
        x = np.zeros(n.value, dtype=np.float64)
        check(lib().hmec_get_solution(self.h, dptr(x)), "get_solution")
        return x

    def update(self, dt, ptscale):
        sc = IterScalars()
        check(lib().hmec_update(self.h, float(dt), float(ptscale), C.byref(sc)), "update")
        return sc

    def controller(self, sc, iterstep):
        """The Picard / dt controller on the device for this iteration's
        scalars; updates the context's dt state; True when the exit test fires."""
        stop = C.c_int()
        check(lib().hmec_controller(self.h, C.byref(sc), int(iterstep), C.byref(stop)), "controller")
        return bool(stop.value)

    def iter_commit(self, ynpl, converged):
        check(lib().hmec_iter_commit(self.h, int(ynpl), int(bool(converged))), "iter_commit")

    def step_end(self, dt):
        check(lib().hmec_step_end(self.h, float(dt)), "step_end")

    def advect(self, dt):
        check(lib().hmec_advect(self.h, float(dt)), "advect")

    def step(self):
        s = StepInfo()
        check(lib().hmec_step(self.h, C.byref(s)), "step")
        return s

    def timing(self, on=True, phase_sync=False):
        """Kernel event timers; phase_sync also waits for the device at every
        phase boundary of a step (per-phase wall-clock split in StepInfo)."""
        check(lib().hmec_timing_enable(self.h, 2 if (on and phase_sync) else int(bool(on))),
              "timing_enable")

    def timing_sample(self, stride=1):
        """Record every stride-th interval of each timer (1: all)."""
        check(lib().hmec_timing_sample(self.h, int(stride)), "timing_sample")

    def timing_select(self, names=None):
        """Record only the named timers (a list or comma-separated string;
        None records all)."""
        if names is not None and not isinstance(names, str):
            names = ",".join(names)
        check(lib().hmec_timing_select(self.h, names.encode() if names else None), "timing_select")

    def timing_reset(self):
        check(lib().hmec_timing_reset(self.h), "timing_reset")

    def kernel_time(self, name):
        ms = C.c_double()
        by = C.c_double()
        n = lib().hmec_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(by))
        return n, ms.value, by.value


def batch_step(ctxs):
    n = len(ctxs)
    arr = (C.c_void_p * n)(*[c.h for c in ctxs])
    infos = (StepInfo * n)()
    check(lib().hmec_batch_step(arr, n, infos), "batch_step")
    return list(infos)


def batch_factorize(ctxs, prec=PREC_FP64):
    n = len(ctxs)
    arr = (C.c_void_p * n)(*[c.h for c in ctxs])
    check(lib().hmec_batch_factorize(arr, n, int(prec)), "batch_factorize")


def member_footprint(p, prec=PREC_FP64):
    """Device bytes a context of this grid allocates under a precision policy."""
    return int(lib().hmec_member_footprint(C.byref(p), int(prec)))


class Server:
    """Ensemble server (hmec_server_*): continuous batching of independent
    simulations, admission by device footprint, batched same-size
    factorizations.  Not in the reference (one simulation per process)."""

    def __init__(self, precision=PREC_FP32, max_active=8, hbm_budget=0, batch_window_ms=2.0,
                 keep_fields=False, device=0):
        self.cfg = ServerConfig(device, int(precision), int(max_active), int(hbm_budget),
                                float(batch_window_ms), 1 if keep_fields else 0)
        h = C.c_void_p()
        check(lib().hmec_server_create(C.byref(self.cfg), C.byref(h)), "server_create")
        self.h = h
        self.params = {}

    def submit(self, p, nsteps, cfg=1, seed=1, jitter=1, arrival_ms=0.0):
        r = Request(p, int(cfg), int(seed), int(jitter), int(nsteps), float(arrival_ms))
        i = C.c_int()
        check(lib().hmec_server_submit(self.h, C.byref(r), C.byref(i)), "server_submit")
        self.params[i.value] = p
        return i.value

    def run(self):
        check(lib().hmec_server_run(self.h), "server_run")

    def result(self, i):
        r = MemberResult()
        check(lib().hmec_server_result(self.h, int(i), C.byref(r)), "server_result")
        return r

    def stats(self):
        st = ServerStats()
        check(lib().hmec_server_get_stats(self.h, C.byref(st)), "server_stats")
        return st

    def latencies(self, i):
        n = lib().hmec_server_latencies(self.h, int(i), None, 0)
        buf = (C.c_double * max(1, n))()
        lib().hmec_server_latencies(self.h, int(i), buf, n)
        return np.array(buf[:n])

    def get(self, i, name):
        """A finished member's final field (keep_fields=True), numpy (rows, cols)."""
        r, c = field_shape(self.params[i], name)
        a = np.zeros((c, r), dtype=np.float64)   # column-major storage
        check(lib().hmec_server_get_field(self.h, int(i), FIELD_ID[name], dptr(a), 0),
              "server_get_field " + name)
        return a.T

    def close(self):
        if self.h:
            lib().hmec_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
