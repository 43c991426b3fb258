"""ctypes binding of the P3 C ABI (include/p3x.h) for Python callers.

This is the Python side of the drop-in boundary: it passes device pointers
(torch CUDA tensors or raw addresses) straight to ``p3x_main``; no compute
happens here.  The library must be built (``python -m scream_amd.build``);
there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libp3x.so")
ICE_TABLE = os.path.join(HERE, "data", "p3_lookup_table_1.dat-v4.1.1")

P3X_OK, P3X_ERR_ARG, P3X_ERR_TABLE_FILE, P3X_ERR_TABLE_VERSION = 0, 1, 2, 3
P3X_ERR_DEVICE, P3X_ERR_TEMPERATURE, P3X_ERR_NO_MEMORY = 4, 5, 6
P3X_ERR_SUBSTEPS, P3X_ERR_BUDGET, P3X_ERR_CHECK_VALUES = 7, 8, 9
P3X_ERR_SCRATCH_LIMIT = 10
# the fixed scratch array of p3x_main_noalloc, per GPU (bytes)
P3X_NOALLOC_SCRATCH_BYTES = 128 << 20
P3X_FP64, P3X_MIXED = 0, 1
# array layouts: level-major [k*ld + i] (numpy shape (nlev, ncol)) or the
# reference's column-major (column, level) views [i*ld + k] (shape (ncol, nlev))
P3X_LEVEL_MAJOR, P3X_COLUMN_MAJOR = 0, 1

RUNTIME_KEYS = (
    "max_total_ni", "autoconversion_prefactor", "autoconversion_qc_exponent",
    "autoconversion_nc_exponent", "autoconversion_radius", "accretion_prefactor",
    "accretion_qc_exponent", "accretion_qr_exponent", "rain_selfcollection_prefactor",
    "rain_selfcollection_breakup_diameter", "constant_mu_rain", "spa_ccn_to_nc_factor",
    "cldliq_to_ice_collection_factor", "rain_to_ice_collection_factor", "min_rime_rho",
    "max_rime_rho", "immersion_freezing_exponent", "deposition_nucleation_exponent",
    "ice_sedimentation_factor")
STATE = ("qc", "nc", "qr", "nr", "qi", "qm", "ni", "bm", "qv", "th")
INPUTS = ("nc_nuceat_tend", "nccn", "ni_activated", "inv_qc_relvar", "cld_frac_i",
          "cld_frac_l", "cld_frac_r", "pres", "dz", "dpres", "inv_exner", "qv_prev", "t_prev")
OUTS = ("qv2qi_depos_tend", "precip_liq_surf", "precip_ice_surf", "diag_eff_radius_qc",
        "diag_eff_radius_qi", "diag_eff_radius_qr", "rho_qi", "precip_liq_flux",
        "precip_ice_flux", "precip_total_tend", "nevapr")
HIST = ("liq_ice_exchange", "vap_liq_exchange", "vap_ice_exchange")
ALL_FIELDS = STATE + INPUTS + OUTS + HIST

# Every symbol include/p3x.h declares.
EXPORTS = ("p3x_runtime_defaults", "p3x_launch_defaults", "p3x_tables_load", "p3x_init", "p3x_reserve",
           "p3x_finalize", "p3x_main", "p3x_main_noalloc", "p3x_sync", "p3x_last_kernel_times", "p3x_strerror",
           "p3x_batch_defaults",
           "p3x_batch_create", "p3x_batch_run", "p3x_batch_tune", "p3x_batch_destroy",
           "p3x_server_defaults", "p3x_server_create", "p3x_server_add_member", "p3x_server_add_process",
           "p3x_server_submit",
           "p3x_server_poll", "p3x_server_drain", "p3x_server_stats_get", "p3x_server_destroy",
           "p3x_process_create", "p3x_process_create_layout", "p3x_process_run", "p3x_process_destroy",
           "p3x_column_budget",
           "p3x_conservation_errors", "p3x_cld_fraction", "p3x_shoc_postprocess", "p3x_ice_table_generate",
           "p3x_ice_table_write")

# p3x_atm_fields, in struct order (the P3Microphysics field names)
ATM_FIELDS = ("p_mid", "p_dry_mid", "pseudo_density", "pseudo_density_dry", "cldfrac_tot",
              "nc_nuceat_tend", "nccn", "ni_activated", "inv_qc_relvar",
              "T_mid", "qv", "qc", "nc", "qr", "nr", "qi", "qm", "ni", "bm",
              "qv_prev_micro_step", "T_prev_micro_step", "precip_liq_surf_mass", "precip_ice_surf_mass",
              "eff_radius_qc", "eff_radius_qi", "eff_radius_qr", "precip_total_tend", "nevapr", "rainfrac",
              "micro_liq_ice_exchange", "micro_vap_liq_exchange", "micro_vap_ice_exchange",
              "vapor_flux", "water_flux", "ice_flux", "heat_flux")


class Runtime(C.Structure):
    _fields_ = [(k, C.c_double) for k in RUNTIME_KEYS] + [("do_ice_production", C.c_int)]


class PrognosticState(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in STATE]


class DiagnosticInputs(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in INPUTS]


class DiagnosticOutputs(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in OUTS]


class Infrastructure(C.Structure):
    _fields_ = [("dt", C.c_double), ("it", C.c_int), ("its", C.c_int), ("ite", C.c_int),
                ("kts", C.c_int), ("kte", C.c_int), ("predictNc", C.c_int),
                ("prescribedCCN", C.c_int), ("col_location", C.c_void_p),
                ("exchange_zeroed_on_entry", C.c_int)]


class HistoryOnly(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in HIST]


class TableData(C.Structure):
    _fields_ = [("ice_table", C.c_double * 12000), ("collect_table", C.c_double * 60000),
                ("vn_table", C.c_double * 3000), ("vm_table", C.c_double * 3000),
                ("revap_table", C.c_double * 3000), ("mu_r_table", C.c_double * 150),
                ("dnu", C.c_double * 16)]


class Launch(C.Structure):
    _fields_ = [("precision", C.c_int), ("ld", C.c_int), ("columns_per_block", C.c_int),
                ("threads_per_block", C.c_int), ("sed_lanes_per_column", C.c_int),
                ("part12_waves_per_simd", C.c_int), ("stream", C.c_void_p), ("async_", C.c_int),
                ("col_status", C.c_void_p), ("sed_substeps", C.c_void_p),
                ("ice_table_idx", C.c_void_p), ("layout", C.c_int), ("ld_iface", C.c_int),
                ("check_values", C.c_void_p)]


class ColumnSet(C.Structure):
    _fields_ = [("prognostic_state", PrognosticState), ("diagnostic_inputs", DiagnosticInputs),
                ("diagnostic_outputs", DiagnosticOutputs), ("history_only", HistoryOnly),
                ("infrastructure", Infrastructure), ("nj", C.c_int), ("nk", C.c_int), ("ld", C.c_int),
                ("col_status", C.c_void_p), ("sed_substeps", C.c_void_p), ("ice_table_idx", C.c_void_p),
                ("layout", C.c_int), ("ld_iface", C.c_int), ("check_values", C.c_void_p)]


class BatchOptions(C.Structure):
    _fields_ = [("precision", C.c_int), ("columns_per_block", C.c_int), ("threads_per_block", C.c_int),
                ("sed_lanes_per_column", C.c_int), ("part12_waves_per_simd", C.c_int),
                ("blocks_per_cu", C.c_int), ("use_graph", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerOptions(C.Structure):
    _fields_ = [("precision", C.c_int), ("slots", C.c_int), ("max_batch", C.c_int),
                ("scratch_budget_bytes", C.c_double), ("columns_per_block", C.c_int),
                ("threads_per_block", C.c_int), ("sed_lanes_per_column", C.c_int),
                ("part12_waves_per_simd", C.c_int), ("hbm_budget_bytes", C.c_double)]


class Completion(C.Structure):
    _fields_ = [("tag", C.c_ulonglong), ("member", C.c_int), ("status", C.c_int),
                ("latency_us", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("completed", C.c_longlong), ("p50_us", C.c_double),
                ("p99_us", C.c_double), ("max_us", C.c_double), ("column_levels", C.c_double),
                ("busy_s", C.c_double), ("resident_bytes", C.c_double), ("scratch_bytes", C.c_double),
                ("members", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


BUDGET_FIELDS = ("pseudo_density", "T_mid", "u", "v", "qv", "qc", "qi", "qr", "ps", "phis")


class BudgetFields(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in BUDGET_FIELDS] + [("layout", C.c_int), ("ld", C.c_int)]


class AtmFields(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ATM_FIELDS]


class StepChain(C.Structure):
    """p3x_step_chain: the neighbour stages of a served process member."""
    _fields_ = [("qc2", C.c_void_p), ("cldfrac_liq", C.c_void_p), ("ice_threshold", C.c_double),
                ("ice_4out_threshold", C.c_double), ("cldfrac_ice", C.c_void_p),
                ("cldfrac_ice_4out", C.c_void_p), ("cldfrac_tot_4out", C.c_void_p),
                ("mass_err", C.c_void_p), ("energy_err", C.c_void_p), ("budget", C.c_void_p),
                ("u", C.c_void_p), ("v", C.c_void_p), ("ps", C.c_void_p), ("phis", C.c_void_p)]


class P3Error(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        super().__init__(f"{what}: p3x status {status} ({strerror(status)})")


_lib = None


def load() -> C.CDLL:
    """The compiled library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m scream_amd.build`")
        # torch ships its own libamdhip64.so.7.  Importing it first makes that the
        # process's one HIP runtime, which libp3x.so then binds to by soname; the
        # other order would load a second HIP/HSA runtime next to torch's.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(LIB_PATH)
        lib.p3x_runtime_defaults.argtypes = [C.POINTER(Runtime)]
        lib.p3x_launch_defaults.argtypes = [C.POINTER(Launch)]
        lib.p3x_tables_load.argtypes = [C.c_char_p, C.POINTER(TableData)]
        lib.p3x_tables_load.restype = C.c_int
        lib.p3x_ice_table_generate.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]
        lib.p3x_ice_table_generate.restype = C.c_int
        lib.p3x_ice_table_write.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p]
        lib.p3x_ice_table_write.restype = C.c_int
        lib.p3x_init.argtypes = [C.POINTER(TableData), C.POINTER(C.c_void_p)]
        lib.p3x_init.restype = C.c_int
        lib.p3x_finalize.argtypes = [C.c_void_p]
        lib.p3x_main.argtypes = [C.c_void_p, C.POINTER(Runtime), C.POINTER(PrognosticState),
                                 C.POINTER(DiagnosticInputs), C.POINTER(DiagnosticOutputs),
                                 C.POINTER(Infrastructure), C.POINTER(HistoryOnly), C.c_int,
                                 C.c_int, C.POINTER(Launch), C.POINTER(C.c_double)]
        lib.p3x_main.restype = C.c_int
        if hasattr(lib, "p3x_main_noalloc"):  # (every export is checked by the ABI test)
            lib.p3x_main_noalloc.argtypes = [C.c_void_p, C.POINTER(Runtime), C.POINTER(PrognosticState),
                                     C.POINTER(DiagnosticInputs), C.POINTER(DiagnosticOutputs),
                                     C.POINTER(Infrastructure), C.POINTER(HistoryOnly), C.c_int,
                                     C.c_int, C.POINTER(Launch), C.POINTER(C.c_double)]
            lib.p3x_main_noalloc.restype = C.c_int
        lib.p3x_sync.argtypes = [C.c_void_p]
        lib.p3x_sync.restype = C.c_int
        lib.p3x_last_kernel_times.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.p3x_last_kernel_times.restype = C.c_int
        lib.p3x_strerror.argtypes = [C.c_int]
        lib.p3x_strerror.restype = C.c_char_p
        lib.p3x_batch_defaults.argtypes = [C.POINTER(BatchOptions)]
        lib.p3x_batch_create.argtypes = [C.c_void_p, C.POINTER(Runtime), C.POINTER(ColumnSet), C.c_int,
                                         C.POINTER(BatchOptions), C.POINTER(C.c_void_p)]
        lib.p3x_batch_create.restype = C.c_int
        lib.p3x_batch_run.argtypes = [C.c_void_p, C.c_void_p]
        lib.p3x_batch_run.restype = C.c_int
        lib.p3x_batch_tune.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(BatchOptions),
                                       C.POINTER(C.c_double)]
        lib.p3x_batch_tune.restype = C.c_int
        lib.p3x_batch_destroy.argtypes = [C.c_void_p]
        lib.p3x_server_defaults.argtypes = [C.POINTER(ServerOptions)]
        lib.p3x_server_create.argtypes = [C.c_void_p, C.POINTER(Runtime), C.POINTER(ServerOptions),
                                          C.POINTER(C.c_void_p)]
        lib.p3x_server_create.restype = C.c_int
        lib.p3x_server_add_member.argtypes = [C.c_void_p, C.POINTER(ColumnSet), C.POINTER(C.c_int)]
        lib.p3x_server_add_member.restype = C.c_int
        lib.p3x_server_add_process.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(AtmFields),
                                               C.POINTER(StepChain), C.POINTER(C.c_int)]
        lib.p3x_server_add_process.restype = C.c_int
        lib.p3x_shoc_postprocess.argtypes = [C.c_void_p] + [C.c_void_p] * 4 + [C.c_int, C.c_int, C.c_void_p]
        lib.p3x_shoc_postprocess.restype = C.c_int
        lib.p3x_server_submit.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_ulonglong]
        lib.p3x_server_submit.restype = C.c_int
        lib.p3x_server_poll.argtypes = [C.c_void_p, C.POINTER(Completion), C.c_int, C.c_int]
        lib.p3x_server_poll.restype = C.c_int
        lib.p3x_server_drain.argtypes = [C.c_void_p]
        lib.p3x_server_drain.restype = C.c_int
        lib.p3x_server_stats_get.argtypes = [C.c_void_p, C.POINTER(ServerStats)]
        lib.p3x_server_stats_get.restype = C.c_int
        lib.p3x_server_destroy.argtypes = [C.c_void_p]
        lib.p3x_process_create.argtypes = [C.c_void_p, C.POINTER(Runtime), C.c_int, C.c_int, C.c_int,
                                           C.c_int, C.POINTER(C.c_void_p)]
        lib.p3x_process_create.restype = C.c_int
        lib.p3x_process_create_layout.argtypes = [C.c_void_p, C.POINTER(Runtime), C.c_int, C.c_int, C.c_int,
                                                  C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        lib.p3x_process_create_layout.restype = C.c_int
        lib.p3x_process_run.argtypes = [C.c_void_p, C.POINTER(AtmFields), C.c_double, C.POINTER(Launch)]
        lib.p3x_process_run.restype = C.c_int
        lib.p3x_process_destroy.argtypes = [C.c_void_p]
        lib.p3x_column_budget.argtypes = [C.c_void_p, C.POINTER(BudgetFields), C.c_int, C.c_int, C.c_void_p,
                                          C.c_void_p, C.c_void_p]
        lib.p3x_column_budget.restype = C.c_int
        lib.p3x_conservation_errors.argtypes = [C.c_void_p] + [C.c_void_p] * 8 + [C.c_int, C.c_double,
                                                                                 C.c_void_p, C.c_void_p,
                                                                                 C.c_void_p]
        lib.p3x_conservation_errors.restype = C.c_int
        lib.p3x_cld_fraction.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double,
                                         C.c_double] + [C.c_void_p] * 5
        lib.p3x_cld_fraction.restype = C.c_int
        lib.p3x_reserve.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.p3x_reserve.restype = C.c_int
        _lib = lib
    return _lib


def strerror(status: int) -> str:
    return load().p3x_strerror(int(status)).decode()


def runtime_defaults(**overrides) -> Runtime:
    rt = Runtime()
    load().p3x_runtime_defaults(C.byref(rt))
    for k, v in overrides.items():
        setattr(rt, k, v)
    return rt


def tables_load(path: str = ICE_TABLE) -> TableData:
    t = TableData()
    st = load().p3x_tables_load(path.encode(), C.byref(t))
    if st != P3X_OK:
        raise P3Error(st, f"p3x_tables_load({path})")
    return t


# Cells of the ice lookup-table generator: (rime density 5, rime fraction 4, mass 50)
ICE_TABLE_CELLS, ICE_ROW_COLS, COLL_RAIN_SIZES, COLL_ROW_COLS = 1000, 15, 30, 6


def ice_table_generate(stream=None):
    """Runs the ice lookup-table generator (create_p3_lookupTable_1.f90-v4.1) on the
    current device.  Returns (ice_rows [1000, 15], coll_rows [1000, 30, 6], wall ms)."""
    import numpy as np
    ice = np.zeros((ICE_TABLE_CELLS, ICE_ROW_COLS))
    coll = np.zeros((ICE_TABLE_CELLS, COLL_RAIN_SIZES, COLL_ROW_COLS))
    ms = C.c_double()
    st = load().p3x_ice_table_generate(ice.ctypes.data, coll.ctypes.data, stream, C.byref(ms))
    if st != P3X_OK:
        raise P3Error(st, "p3x_ice_table_generate")
    return ice, coll, ms.value


def ice_table_write(path: str, ice_rows, coll_rows) -> None:
    """Writes generator rows as a table file that tables_load reads."""
    import numpy as np
    ice = np.ascontiguousarray(ice_rows, dtype=np.float64)
    coll = np.ascontiguousarray(coll_rows, dtype=np.float64)
    if ice.shape != (ICE_TABLE_CELLS, ICE_ROW_COLS) or coll.shape != (ICE_TABLE_CELLS, COLL_RAIN_SIZES, COLL_ROW_COLS):
        raise ValueError("ice_table_write: rows of the wrong shape")
    st = load().p3x_ice_table_write(path.encode(), ice.ctypes.data, coll.ctypes.data)
    if st != P3X_OK:
        raise P3Error(st, f"p3x_ice_table_write({path})")


def _addr(x) -> int:
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()  # torch tensor on the device


class Context:
    """A p3x context on the current HIP device (tables resident in HBM)."""

    def __init__(self, tables: TableData | None = None):
        self._lib = load()
        self.tables = tables if tables is not None else tables_load()
        h = C.c_void_p()
        st = self._lib.p3x_init(C.byref(self.tables), C.byref(h))
        if st != P3X_OK:
            raise P3Error(st, "p3x_init")
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            self._lib.p3x_finalize(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def main(self, f: dict, nj: int, nk: int, dt: float, predict_nc=True, prescribed_ccn=False,
             runtime: Runtime | None = None, precision: int = P3X_FP64, ld: int = 0,
             columns_per_block: int = 0, threads_per_block: int = 0,
             sed_lanes_per_column: int = 0, part12_waves_per_simd: int = 0, stream=None,
             async_: bool = False, debug: dict | None = None, exchange_zeroed=False,
             check: bool = True, layout: int = P3X_LEVEL_MAJOR, ld_iface: int = 0,
             noalloc: bool = False) -> float:
        """One p3x_main call; ``f`` maps every field name to a device array in
        ``layout`` (level-major, shape (nlev, ncol), or column-major, shape (ncol,
        nlev), the reference's view order).  ``noalloc``: call p3x_main_noalloc
        (the step's scratch from the library's fixed per-GPU array).  Returns the
        step's device time in microseconds (0 when async)."""
        rt = runtime if runtime is not None else runtime_defaults()
        ps = PrognosticState(*[_addr(f[k]) for k in STATE])
        di = DiagnosticInputs(*[_addr(f[k]) for k in INPUTS])
        do = DiagnosticOutputs(*[_addr(f[k]) for k in OUTS])
        hi = HistoryOnly(*[_addr(f[k]) for k in HIST])
        inf = Infrastructure(dt=float(dt), it=1, its=0, ite=nj - 1, kts=0, kte=nk - 1,
                             predictNc=int(bool(predict_nc)),
                             prescribedCCN=int(bool(prescribed_ccn)), col_location=None,
                             exchange_zeroed_on_entry=int(bool(exchange_zeroed)))
        dbg = debug or {}
        lo = Launch(precision=int(precision), ld=int(ld), columns_per_block=int(columns_per_block),
                    threads_per_block=int(threads_per_block),
                    sed_lanes_per_column=int(sed_lanes_per_column),
                    part12_waves_per_simd=int(part12_waves_per_simd), stream=stream,
                    async_=int(bool(async_)), col_status=_addr(dbg.get("col_status")),
                    sed_substeps=_addr(dbg.get("sed_substeps")),
                    ice_table_idx=_addr(dbg.get("ice_table_idx")), layout=int(layout),
                    ld_iface=int(ld_iface), check_values=_addr(dbg.get("check_values")))
        us = C.c_double(0.0)
        fn = self._lib.p3x_main_noalloc if noalloc else self._lib.p3x_main
        st = fn(self.handle, C.byref(rt), C.byref(ps), C.byref(di), C.byref(do),
                C.byref(inf), C.byref(hi), int(nj), int(nk), C.byref(lo), C.byref(us))
        if check and st != P3X_OK:
            raise P3Error(st, "p3x_main_noalloc" if noalloc else "p3x_main")
        return us.value if st == P3X_OK else float("nan")

    def sync(self) -> int:
        return self._lib.p3x_sync(self.handle)

    def reserve(self, ld: int, nk: int):
        """p3x_reserve: scratch for column sets up to ld columns x nk levels, so that
        main() allocates nothing for them."""
        st = self._lib.p3x_reserve(self.handle, int(ld), int(nk))
        if st != P3X_OK:
            raise P3Error(st, "p3x_reserve")

    def last_kernel_times(self) -> tuple[float, float]:
        """Device microseconds of the last main() in two intervals: (the part 1 and part 2
        kernels, the sedimentation and part 3 kernels)."""
        a, b = C.c_double(), C.c_double()
        st = self._lib.p3x_last_kernel_times(self.handle, C.byref(a), C.byref(b))
        if st != P3X_OK:
            raise RuntimeError(f"p3x_last_kernel_times: {strerror(st)}")
        return a.value, b.value


def batch_defaults(**overrides) -> BatchOptions:
    o = BatchOptions()
    load().p3x_batch_defaults(C.byref(o))
    for k, v in overrides.items():
        setattr(o, k, v)
    return o


def column_set(f: dict, nj: int, nk: int, dt: float, predict_nc=True, prescribed_ccn=False,
               exchange_zeroed=False, ld: int = 0, debug: dict | None = None,
               layout: int = P3X_LEVEL_MAJOR, ld_iface: int = 0) -> ColumnSet:
    """The descriptor of one column set of a batch; ``f`` maps every field name to a device array
    in ``layout``."""
    dbg = debug or {}
    inf = Infrastructure(dt=float(dt), it=1, its=0, ite=nj - 1, kts=0, kte=nk - 1,
                         predictNc=int(bool(predict_nc)), prescribedCCN=int(bool(prescribed_ccn)),
                         col_location=None, exchange_zeroed_on_entry=int(bool(exchange_zeroed)))
    return ColumnSet(PrognosticState(*[_addr(f[k]) for k in STATE]),
                     DiagnosticInputs(*[_addr(f[k]) for k in INPUTS]),
                     DiagnosticOutputs(*[_addr(f[k]) for k in OUTS]),
                     HistoryOnly(*[_addr(f[k]) for k in HIST]), inf, int(nj), int(nk), int(ld),
                     _addr(dbg.get("col_status")), _addr(dbg.get("sed_substeps")),
                     _addr(dbg.get("ice_table_idx")), int(layout), int(ld_iface),
                     _addr(dbg.get("check_values")))


class Batch:
    """Many column sets advanced one P3 step each per ``run`` (p3x_batch_*)."""

    def __init__(self, ctx: Context, sets: list, options: BatchOptions | None = None,
                 runtime: Runtime | None = None):
        self._lib = load()
        self.ctx = ctx
        self.sets = (ColumnSet * len(sets))(*sets)
        rt = runtime if runtime is not None else runtime_defaults()
        opts = options if options is not None else batch_defaults()
        h = C.c_void_p()
        st = self._lib.p3x_batch_create(ctx.handle, C.byref(rt), self.sets, len(sets), C.byref(opts),
                                        C.byref(h))
        if st != P3X_OK:
            raise P3Error(st, "p3x_batch_create")
        self.handle = h
        self.options = opts

    def run(self, stream=None):
        st = self._lib.p3x_batch_run(self.handle, stream)
        if st != P3X_OK:
            raise P3Error(st, "p3x_batch_run")

    def tune(self, stream=None, reps: int = 3):
        """Autotune the launch shape on this batch's data (state left unchanged);
        returns (chosen options, best device microseconds per step)."""
        chosen, us = BatchOptions(), C.c_double(0)
        st = self._lib.p3x_batch_tune(self.handle, stream, int(reps), C.byref(chosen), C.byref(us))
        if st != P3X_OK:
            raise P3Error(st, "p3x_batch_tune")
        self.options = chosen
        return chosen, us.value

    def close(self):
        if getattr(self, "handle", None):
            self._lib.p3x_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def server_defaults(**overrides) -> ServerOptions:
    o = ServerOptions()
    load().p3x_server_defaults(C.byref(o))
    for k, v in overrides.items():
        setattr(o, k, v)
    return o


class Server:
    """Continuous batching of P3 step requests from ensemble members (p3x_server_*)."""

    def __init__(self, ctx: Context, options: ServerOptions | None = None, runtime: Runtime | None = None):
        self._lib = load()
        self.ctx = ctx
        rt = runtime if runtime is not None else runtime_defaults()
        opts = options if options is not None else server_defaults()
        h = C.c_void_p()
        st = self._lib.p3x_server_create(ctx.handle, C.byref(rt), C.byref(opts), C.byref(h))
        if st != P3X_OK:
            raise P3Error(st, "p3x_server_create")
        self.handle = h
        self._sets = []  # keep the descriptors alive

    def add_member(self, cs: ColumnSet) -> int:
        m = C.c_int(-1)
        self._sets.append(cs)
        st = self._lib.p3x_server_add_member(self.handle, C.byref(cs), C.byref(m))
        if st != P3X_OK:
            raise P3Error(st, "p3x_server_add_member")
        return m.value

    def add_process(self, proc: "P3Microphysics", fields: dict, chain: dict | None = None) -> int:
        """Register a process member: each request runs proc's whole step (pre-amble,
        p3_main, post-amble) on ``fields`` (ATM_FIELDS names), with the neighbour stages
        of ``chain`` (StepChain field names mapped to device arrays or thresholds)."""
        af = AtmFields(*[_addr(fields.get(k)) for k in ATM_FIELDS])
        ch = None
        if chain is not None:
            ch = StepChain(ice_threshold=1e-12, ice_4out_threshold=1e-5)
            for k, v in chain.items():
                setattr(ch, k, v if isinstance(v, float) else _addr(v))
        self._sets.append((af, ch, proc, fields, chain))
        m = C.c_int(-1)
        st = self._lib.p3x_server_add_process(self.handle, proc.handle, C.byref(af),
                                              C.byref(ch) if ch is not None else None, C.byref(m))
        if st != P3X_OK:
            raise P3Error(st, "p3x_server_add_process")
        return m.value

    def submit(self, member: int, dt: float, tag: int = 0):
        st = self._lib.p3x_server_submit(self.handle, int(member), float(dt), int(tag))
        if st != P3X_OK:
            raise P3Error(st, "p3x_server_submit")

    def poll(self, max_n: int = 256, wait_ms: int = 0) -> list:
        buf = (Completion * max_n)()
        n = self._lib.p3x_server_poll(self.handle, buf, max_n, int(wait_ms))
        if n < 0:
            raise P3Error(-n, "p3x_server_poll")
        return [buf[i] for i in range(n)]

    def drain(self):
        st = self._lib.p3x_server_drain(self.handle)
        if st != P3X_OK:
            raise P3Error(st, "p3x_server_drain")

    def stats(self) -> ServerStats:
        s = ServerStats()
        st = self._lib.p3x_server_stats_get(self.handle, C.byref(s))
        if st != P3X_OK:
            raise P3Error(st, "p3x_server_stats_get")
        return s

    def close(self):
        if getattr(self, "handle", None):
            self._lib.p3x_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class P3Microphysics:
    """The P3 atmosphere process (p3x_process_*): ``run(fields, dt)`` is run_impl --
    pre-amble, p3_main, post-amble -- on the field-manager fields (device arrays by
    their field names, see ATM_FIELDS; the four boundary fluxes are optional)."""

    name = "p3"

    def __init__(self, ctx: Context, ncol: int, nlev: int, predict_nc=True, prescribed_ccn=False,
                 runtime: Runtime | None = None, layout: int = P3X_LEVEL_MAJOR, ld: int = 0,
                 ld_iface: int = 0):
        self._lib = load()
        self.ctx, self.ncol, self.nlev = ctx, ncol, nlev
        self.layout = layout
        rt = runtime if runtime is not None else runtime_defaults()
        h = C.c_void_p()
        st = self._lib.p3x_process_create_layout(ctx.handle, C.byref(rt), int(ncol), int(nlev),
                                                 int(bool(predict_nc)), int(bool(prescribed_ccn)),
                                                 int(layout), int(ld), int(ld_iface), C.byref(h))
        if st != P3X_OK:
            raise P3Error(st, "p3x_process_create_layout")
        self.handle = h

    def run(self, fields: dict, dt: float, precision: int = P3X_FP64, stream=None,
            columns_per_block: int = 0, threads_per_block: int = 0, sed_lanes_per_column: int = 0,
            part12_waves_per_simd: int = 0, async_: bool = False):
        af = AtmFields(*[_addr(fields.get(k)) for k in ATM_FIELDS])
        lo = Launch(precision=int(precision), ld=0, columns_per_block=int(columns_per_block),
                    threads_per_block=int(threads_per_block),
                    sed_lanes_per_column=int(sed_lanes_per_column),
                    part12_waves_per_simd=int(part12_waves_per_simd), stream=stream,
                    async_=int(bool(async_)))
        st = self._lib.p3x_process_run(self.handle, C.byref(af), float(dt), C.byref(lo))
        if st != P3X_OK:
            raise P3Error(st, "p3x_process_run")

    def close(self):
        if getattr(self, "handle", None):
            self._lib.p3x_process_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


CLD_FRACTION_OUT = ("cldfrac_ice", "cldfrac_tot", "cldfrac_ice_for_analysis", "cldfrac_tot_for_analysis")


class CldFraction:
    """The CldFraction process (physics/cld_fraction/eamxx_cld_fraction_process_interface.cpp):
    parameters ice_cloud_threshold (default 1e-12) and ice_cloud_for_analysis_threshold
    (default 1e-5), :60-62; ``run(fields)`` reads qi and cldfrac_liq and writes the four
    CLD_FRACTION_OUT fields (device arrays by field name)."""

    name = "CldFraction"

    def __init__(self, ctx: Context, ncol: int, nlev: int, ice_cloud_threshold: float = 1e-12,
                 ice_cloud_for_analysis_threshold: float = 1e-5):
        self._lib = load()
        self.ctx, self.ncol, self.nlev = ctx, ncol, nlev
        self.ice_cloud_threshold = float(ice_cloud_threshold)
        self.ice_cloud_for_analysis_threshold = float(ice_cloud_for_analysis_threshold)

    def run(self, fields: dict, stream=None):
        st = self._lib.p3x_cld_fraction(self.ctx.handle, _addr(fields["qi"]), _addr(fields["cldfrac_liq"]),
                                        int(self.ncol), int(self.nlev), self.ice_cloud_threshold,
                                        self.ice_cloud_for_analysis_threshold,
                                        *[_addr(fields[k]) for k in CLD_FRACTION_OUT], stream)
        if st != P3X_OK:
            raise P3Error(st, "p3x_cld_fraction")


class MicrophysicsStep:
    """The stretch of the physics group around P3 that needs no other neighbour:
    CldFraction (its cldfrac_tot feeds P3), then P3Microphysics with the mass-and-
    energy column conservation check around it (the property check the atmosphere
    driver runs for a process with the check on: budgets before and after run_impl
    and the relative errors against P3's boundary fluxes).  All on one stream.

    ``run(fields, dt)`` takes the P3 fields of ATM_FIELDS (the boundary fluxes
    required), qi and cldfrac_liq for CldFraction, the four CLD_FRACTION_OUT
    outputs (cldfrac_tot is P3's input), and optionally u, v, ps, phis for the
    energy budget; it returns the per-column (mass_err, energy_err) device arrays.
    With async_ the call returns once the work is enqueued and p3x_sync reports
    P3's status; otherwise it waits and raises on an error status."""

    def __init__(self, ctx: Context, ncol: int, nlev: int, predict_nc=True, prescribed_ccn=False,
                 runtime: Runtime | None = None, ice_cloud_threshold: float = 1e-12,
                 ice_cloud_for_analysis_threshold: float = 1e-5, layout: int = P3X_LEVEL_MAJOR):
        import torch
        self.ctx, self.ncol, self.nlev, self.layout = ctx, ncol, nlev, layout
        self.cld = CldFraction(ctx, ncol, nlev, ice_cloud_threshold, ice_cloud_for_analysis_threshold)
        self.p3 = P3Microphysics(ctx, ncol, nlev, predict_nc, prescribed_ccn, runtime, layout=layout)
        dev = torch.device("cuda", torch.cuda.current_device())
        self._b = torch.zeros((6, ncol), dtype=torch.float64, device=dev)  # m0 e0 m1 e1 merr eerr

    def run(self, fields: dict, dt: float, precision: int = P3X_FP64, stream=None, async_: bool = False):
        m0, e0, m1, e1, me, ee = self._b
        self.cld.run(fields, stream=stream)
        column_budget(self.ctx, fields, self.ncol, self.nlev, m0, e0, stream=stream, layout=self.layout)
        self.p3.run(fields, dt, precision=precision, stream=stream, async_=True)
        column_budget(self.ctx, fields, self.ncol, self.nlev, m1, e1, stream=stream, layout=self.layout)
        conservation_errors(self.ctx, m0, e0, m1, e1, fields["vapor_flux"], fields["water_flux"],
                            fields["ice_flux"], fields["heat_flux"], self.ncol, dt, me, ee, stream=stream)
        if not async_:
            st = self.ctx.sync()
            if st != P3X_OK:
                raise P3Error(st, "MicrophysicsStep.run")
        return me, ee

    def close(self):
        self.p3.close()


def shoc_postprocess(ctx: Context, qc, qc2, inv_qc_relvar, ncol: int, nlev: int, cldfrac_liq=None,
                     stream=None):
    """Enqueue p3x_shoc_postprocess: SHOC's inv_qc_relvar from qc and its variance qc2
    (dense device arrays), and cldfrac_liq clipped to at most 1 when given."""
    st = load().p3x_shoc_postprocess(ctx.handle, _addr(qc), _addr(qc2), _addr(cldfrac_liq),
                                     _addr(inv_qc_relvar), int(ncol), int(nlev), stream)
    if st != P3X_OK:
        raise P3Error(st, "p3x_shoc_postprocess")


def column_budget(ctx: Context, fields: dict, ncol: int, nlev: int, mass, energy, stream=None,
                  layout: int = P3X_LEVEL_MAJOR, ld: int = 0):
    """Enqueue p3x_column_budget; fields maps BUDGET_FIELDS names to device arrays
    (u, v, ps, phis optional) in ``layout``."""
    bf = BudgetFields(*[_addr(fields.get(k)) for k in BUDGET_FIELDS], int(layout), int(ld))
    st = load().p3x_column_budget(ctx.handle, C.byref(bf), int(ncol), int(nlev), _addr(mass), _addr(energy),
                                  stream)
    if st != P3X_OK:
        raise P3Error(st, "p3x_column_budget")


def conservation_errors(ctx: Context, m0, e0, m1, e1, vapor, water, ice, heat, ncol: int, dt: float,
                        mass_err, energy_err, stream=None):
    st = load().p3x_conservation_errors(ctx.handle, *[_addr(x) for x in (m0, e0, m1, e1, vapor, water, ice,
                                                                         heat)],
                                        int(ncol), float(dt), _addr(mass_err), _addr(energy_err), stream)
    if st != P3X_OK:
        raise P3Error(st, "p3x_conservation_errors")
