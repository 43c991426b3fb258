"""Checkers for the P3 product: test infrastructure, never imported by the product.

* ``run_oracle``: the C restatement of p3_main (libp3oracle.so, built from
  p3_oracle.c), FP64, OpenMP over columns.
* ``run_ref`` / ``ref_tables`` / ``ref_lookup_*``: the reference's own Fortran
  P3 (micro_p3.F90) compiled from /root/reference into _ref/libp3ref.so; only
  present where the reference sources are (this container).

Arrays are level-major numpy float64 (shape (nlev, ncol); fluxes (nlev+1, ncol)).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libp3oracle.so")
ORACLE_FMATH_SO = os.path.join(HERE, "libp3oracle_fmath.so")
ORACLE_PMATH_SO = os.path.join(HERE, "libp3oracle_pmath.so")
# Elementary-function flavours of the restatement:
#   "glibc"    the C library, as the reference C++ on the CPU (the default);
#   "fortran"  cbrt/expm1 spelled as the non-BFB Fortran reference does, for
#              the bit-for-bit comparison with micro_p3.F90;
#   "portable" scream_amd/csrc/p3_math.h, the functions the HIP FP64 form
#              uses, for the bit-for-bit comparison with the GPU.
ORACLE_EXPUP_SO = os.path.join(HERE, "libp3oracle_expup.so")
#   "exp_ulp_up" the glibc flavour with every exp result one ulp up: a
#              conditioning probe (how far one ulp in one function moves p3_main)
FLAVOURS = {"glibc": ORACLE_SO, "fortran": ORACLE_FMATH_SO, "portable": ORACLE_PMATH_SO,
            "exp_ulp_up": ORACLE_EXPUP_SO}
REF_SO = os.path.join(HERE, "_ref", "libp3ref.so")
TABLEGEN_SO = {"glibc": os.path.join(HERE, "libp3tablegen.so"),
               "portable": os.path.join(HERE, "libp3tablegen_pmath.so")}
_tablegen = {}


def table_gen(math="glibc", cells=None, nthreads=0):
    """The ice lookup-table generator restated (p3_table_gen.c).  cells: iterable of
    flat cell indices ((rime density * 4 + rime fraction) * 50 + mass index), None =
    all 1000.  Returns (ice [1000, 15], coll [1000, 30, 6]); cells not computed are 0."""
    so = TABLEGEN_SO[math]
    if so not in _tablegen:
        if not os.path.exists(so):
            raise RuntimeError(f"{so} is not built (run `make -C oracle`)")
        lib = C.CDLL(so)
        lib.p3o_table_gen.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        lib.p3o_table_gen.restype = None
        _tablegen[so] = lib
    lib = _tablegen[so]
    ice = np.zeros((1000, 15))
    coll = np.zeros((1000, 30, 6))
    mask = None
    if cells is not None:
        mask = np.zeros(1000, np.uint8)
        mask[list(cells)] = 1
    nt = nthreads if nthreads > 0 else min(16, os.cpu_count() or 1)
    lib.p3o_table_gen(None if mask is None else mask.ctypes.data, ice.ctypes.data, coll.ctypes.data, nt)
    return ice, coll

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

RUNTIME_KEYS = (
    "max_total_ni", "autoconversion_prefactor", "autoconversion_qc_exponent",
    "autoconversion_nc_exponent", "autoconversion_radius", "accretion_prefactor",
    "accretion_qc_exponent", "accretion_qr_exponent", "rain_selfcollection_prefactor",
    "rain_selfcollection_breakup_diameter", "constant_mu_rain", "spa_ccn_to_nc_factor",
    "cldliq_to_ice_collection_factor", "rain_to_ice_collection_factor", "min_rime_rho",
    "max_rime_rho", "immersion_freezing_exponent", "deposition_nucleation_exponent",
    "ice_sedimentation_factor")

# P3Runtime defaults, components/eamxx/src/physics/p3/p3_functions.hpp:114-133
RUNTIME_DEFAULTS = dict(
    max_total_ni=740.0e3, autoconversion_prefactor=1350.0, autoconversion_qc_exponent=2.47,
    autoconversion_nc_exponent=1.79, autoconversion_radius=25.0e-6, accretion_prefactor=67.0,
    accretion_qc_exponent=1.15, accretion_qr_exponent=1.15, rain_selfcollection_prefactor=5.78,
    rain_selfcollection_breakup_diameter=0.00028, constant_mu_rain=1.0,
    spa_ccn_to_nc_factor=1.0, cldliq_to_ice_collection_factor=0.5,
    rain_to_ice_collection_factor=1.0, min_rime_rho=50.0, max_rime_rho=900.0,
    immersion_freezing_exponent=0.65, deposition_nucleation_exponent=0.304,
    ice_sedimentation_factor=1.0, do_ice_production=True)


class _Runtime(C.Structure):
    _fields_ = [(k, C.c_double) for k in RUNTIME_KEYS] + [("do_ice_production", C.c_int)]


class _Tables(C.Structure):
    _fields_ = [(k, _dp) for k in ("ice_table", "collect_table", "vn_table", "vm_table",
                                   "revap_table", "mu_r_table", "dnu")]


STATE = ("qc", "nc", "qr", "nr", "qi", "qm", "ni", "bm", "qv", "th")
INPUTS = ("nc_nuceat_tend", "nccn", "ni_activated", "inv_qc_relvar", "cld_frac_i",
          "cld_frac_l", "cld_frac_r", "pres", "dz", "dpres", "inv_exner", "qv_prev", "t_prev")
OUTS = ("qv2qi_depos_tend", "precip_liq_surf", "precip_ice_surf", "diag_eff_radius_qc",
        "diag_eff_radius_qi", "diag_eff_radius_qr", "rho_qi", "precip_liq_flux",
        "precip_ice_flux", "precip_total_tend", "nevapr")
HIST = ("liq_ice_exchange", "vap_liq_exchange", "vap_ice_exchange")


class _Fields(C.Structure):
    _fields_ = [(k, _dp) for k in STATE + INPUTS + OUTS + HIST]


# P3Microphysics fields (p3o_atm in p3_oracle.h), in struct order
ATM_IN = ("p_mid", "pseudo_density", "pseudo_density_dry", "cldfrac_tot")
ATM_STATE = ("T_mid", "qv", "qc", "nc", "qr", "nr", "qi", "qm", "ni", "bm", "qv_prev", "T_prev")
ATM_LOCAL = ("th_atm", "inv_exner", "dz", "cld_frac_l", "cld_frac_i", "cld_frac_r")
ATM_POST = ("eff_radius_qc", "eff_radius_qi", "eff_radius_qr", "precip_liq_surf",
            "precip_ice_surf", "precip_liq_surf_mass", "precip_ice_surf_mass",
            "vapor_flux", "water_flux", "ice_flux", "heat_flux")


class _Atm(C.Structure):
    _fields_ = [(k, _dp) for k in ATM_IN + ATM_STATE + ATM_LOCAL + ATM_POST]


class _Debug(C.Structure):
    _fields_ = [("col_status", _ip), ("sed_substeps", _ip), ("ice_table_idx", _ip)]


_oracle = {}
_ref = None


def load_oracle(math="glibc"):
    """The restatement built with the elementary functions of flavour ``math``."""
    so = FLAVOURS[math]
    if so not in _oracle:
        if not os.path.exists(so):
            raise RuntimeError(f"{so} is not built (run `make -C oracle`)")
        lib = C.CDLL(so)
        lib.p3o_main.restype = C.c_int
        lib.p3o_main_ex.restype = C.c_int
        lib.p3o_main.argtypes = [C.POINTER(_Runtime), C.POINTER(_Tables), C.POINTER(_Fields),
                                 C.POINTER(_Debug), C.c_int, C.c_int, C.c_double, C.c_int,
                                 C.c_int, C.c_int]
        lib.p3o_main_ex.argtypes = lib.p3o_main.argtypes + [C.c_int]
        lib.p3o_elementary.restype = C.c_double
        lib.p3o_elementary.argtypes = [C.c_int, C.c_double, C.c_double]
        lib.p3o_qv_sat_dry.restype = C.c_double
        lib.p3o_qv_sat_dry.argtypes = [C.c_double, C.c_double, C.c_int]
        lib.p3o_lookup_ice.argtypes = [C.c_double] * 4 + [_ip] * 3 + [_dp] * 3
        lib.p3o_lookup_rain.argtypes = [C.c_double, C.c_double, _ip, _dp]
        lib.p3o_apply_table_ice.restype = C.c_double
        lib.p3o_apply_table_ice.argtypes = [C.POINTER(_Tables), C.c_int] + [C.c_double] * 4
        lib.p3o_apply_table_coll.restype = C.c_double
        lib.p3o_apply_table_coll.argtypes = [C.POINTER(_Tables), C.c_int] + [C.c_double] * 6
        lib.p3o_column_budget.argtypes = [_dp] * 10 + [C.c_int, C.c_int, _dp, _dp]
        lib.p3o_conservation_errors.argtypes = [_dp] * 8 + [C.c_int, C.c_double, _dp, _dp]
        lib.p3o_cld_fraction.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_double, C.c_double] + [_dp] * 4
        lib.p3o_shoc_postprocess.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_int]
        lib.p3o_preamble.argtypes = [C.POINTER(_Atm), C.c_int, C.c_int]
        lib.p3o_postamble.argtypes = [C.POINTER(_Atm), C.c_int, C.c_int, C.c_double]
        lib.p3o_upwind_step.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                        _dp, _dp, _dp, C.POINTER(_dp), C.POINTER(_dp),
                                        C.POINTER(_dp)]
        _oracle[so] = lib
    return _oracle[so]


def _ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _runtime(rt):
    r = dict(RUNTIME_DEFAULTS)
    r.update(rt or {})
    s = _Runtime()
    for k in RUNTIME_KEYS:
        setattr(s, k, float(r[k]))
    s.do_ice_production = int(bool(r["do_ice_production"]))
    return s


def _tables(tab):
    keep = {k: np.ascontiguousarray(tab[k], dtype=np.float64)
            for k in ("ice", "coll", "vn", "vm", "revap", "mu_r", "dnu")}
    t = _Tables(_ptr(keep["ice"]), _ptr(keep["coll"]), _ptr(keep["vn"]), _ptr(keep["vm"]),
                _ptr(keep["revap"]), _ptr(keep["mu_r"]), _ptr(keep["dnu"]))
    return t, keep


def run_oracle(inputs: dict, tables: dict, dt: float, predict_nc=True, prescribed_ccn=False,
               runtime=None, nthreads=0, math="glibc", small_kernel_exits=False):
    """p3_main on copies of ``inputs``; returns (fields dict, debug dict, n_errors).
    small_kernel_exits: the early exits of the reference's small-kernel dispatch."""
    lib = load_oracle(math)
    nlev, ncol = inputs["qc"].shape
    f = {k: np.array(inputs[k], dtype=np.float64, order="C") for k in STATE + INPUTS}
    for k in OUTS + HIST:
        if k in ("precip_liq_surf", "precip_ice_surf"):
            f[k] = np.zeros(ncol)
        elif k in ("precip_liq_flux", "precip_ice_flux"):
            f[k] = np.zeros((nlev + 1, ncol))
        else:
            f[k] = np.array(inputs[k]) if k in inputs else np.zeros((nlev, ncol))
    fs = _Fields(*[_ptr(f[k]) for k in STATE + INPUTS + OUTS + HIST])
    dbg = dict(col_status=np.zeros(ncol, np.int32), sed_substeps=np.zeros((3, ncol), np.int32),
               ice_table_idx=np.zeros((nlev, ncol), np.int32))
    ds = _Debug(*[dbg[k].ctypes.data_as(_ip) for k in ("col_status", "sed_substeps",
                                                         "ice_table_idx")])
    rt = _runtime(runtime)
    tb, keep = _tables(tables)
    nerr = lib.p3o_main_ex(C.byref(rt), C.byref(tb), C.byref(fs), C.byref(ds), ncol, nlev,
                           float(dt), int(predict_nc), int(prescribed_ccn), int(nthreads),
                           int(bool(small_kernel_exits)))
    del keep
    return f, dbg, nerr


def _atm_struct(a):
    return _Atm(*[(_ptr(a[k]) if a.get(k) is not None else None)
                  for k in ATM_IN + ATM_STATE + ATM_LOCAL + ATM_POST])


def run_process(fields: dict, tables: dict, dt: float, predict_nc=True, prescribed_ccn=False,
                runtime=None, nthreads=0, math="glibc", fluxes=False):
    """P3Microphysics::run_impl on copies of the field-manager ``fields``
    (eamxx_p3_run.cpp:5-45): pre-amble, p3_main, post-amble.  ``fields`` holds
    p_mid, p_dry_mid, pseudo_density, pseudo_density_dry, cldfrac_tot,
    nc_nuceat_tend, nccn, ni_activated, inv_qc_relvar, T_mid, qv, qc, nc, qr, nr,
    qi, qm, ni, bm, qv_prev_micro_step, T_prev_micro_step, precip_liq_surf_mass,
    precip_ice_surf_mass.  Returns the updated/computed fields by their
    field-manager names."""
    lib = load_oracle(math)
    nlev, ncol = fields["qc"].shape
    f = {k: np.array(v, dtype=np.float64, order="C") for k, v in fields.items()}
    a = {k: f[k] for k in ("p_mid", "pseudo_density", "pseudo_density_dry", "cldfrac_tot", "T_mid",
                           "qv", "qc", "nc", "qr", "nr", "qi", "qm", "ni", "bm")}
    a["qv_prev"] = f["qv_prev_micro_step"]
    a["T_prev"] = f["T_prev_micro_step"]
    for k in ATM_LOCAL:
        a[k] = np.zeros((nlev, ncol))
    lib.p3o_preamble(C.byref(_atm_struct(a)), ncol, nlev)
    # p3_main on the process's views (eamxx_p3_process_interface.cpp:284-337)
    inp = {"qc": a["qc"], "nc": a["nc"], "qr": a["qr"], "nr": a["nr"], "qi": a["qi"], "qm": a["qm"],
           "ni": a["ni"], "bm": a["bm"], "qv": a["qv"], "th": a["th_atm"],
           "nc_nuceat_tend": f["nc_nuceat_tend"], "nccn": f["nccn"], "ni_activated": f["ni_activated"],
           "inv_qc_relvar": f["inv_qc_relvar"], "cld_frac_i": a["cld_frac_i"],
           "cld_frac_l": a["cld_frac_l"], "cld_frac_r": a["cld_frac_r"], "pres": f["p_dry_mid"],
           "dz": a["dz"], "dpres": f["pseudo_density_dry"], "inv_exner": a["inv_exner"],
           "qv_prev": a["qv_prev"], "t_prev": a["T_prev"]}
    out, dbg, nerr = run_oracle(inp, tables, dt, predict_nc, prescribed_ccn, runtime, nthreads, math)
    for k in ("qc", "nc", "qr", "nr", "qi", "qm", "ni", "bm", "qv"):
        a[k] = out[k]
    a["th_atm"] = out["th"]
    a["eff_radius_qc"] = out["diag_eff_radius_qc"]
    a["eff_radius_qi"] = out["diag_eff_radius_qi"]
    a["eff_radius_qr"] = out["diag_eff_radius_qr"]
    a["precip_liq_surf"] = out["precip_liq_surf"]
    a["precip_ice_surf"] = out["precip_ice_surf"]
    a["precip_liq_surf_mass"] = f["precip_liq_surf_mass"]
    a["precip_ice_surf_mass"] = f["precip_ice_surf_mass"]
    if fluxes:
        for k in ("vapor_flux", "water_flux", "ice_flux", "heat_flux"):
            a[k] = np.full(ncol, np.nan)
    lib.p3o_postamble(C.byref(_atm_struct(a)), ncol, nlev, float(dt))
    res = {k: a[k] for k in ("T_mid", "qv", "qc", "nc", "qr", "nr", "qi", "qm", "ni", "bm",
                             "eff_radius_qc", "eff_radius_qi", "eff_radius_qr",
                             "precip_liq_surf_mass", "precip_ice_surf_mass")}
    res["qv_prev_micro_step"] = a["qv_prev"]
    res["T_prev_micro_step"] = a["T_prev"]
    res["rainfrac"] = a["cld_frac_r"]
    res["precip_total_tend"] = out["precip_total_tend"]
    res["nevapr"] = out["nevapr"]
    res["micro_liq_ice_exchange"] = out["liq_ice_exchange"]
    res["micro_vap_liq_exchange"] = out["vap_liq_exchange"]
    res["micro_vap_ice_exchange"] = out["vap_ice_exchange"]
    if fluxes:
        for k in ("vapor_flux", "water_flux", "ice_flux", "heat_flux"):
            res[k] = a[k]
    return res, dbg, nerr


def column_budget(f: dict, math="glibc"):
    """Column total mass and energy of the conservation check (p3o_column_budget):
    f holds pseudo_density, T_mid, qv, qc, qi, qr and optionally u, v, ps, phis."""
    lib = load_oracle(math)
    nlev, ncol = f["qv"].shape
    a = {k: (np.ascontiguousarray(f[k], dtype=np.float64) if f.get(k) is not None else None)
         for k in ("pseudo_density", "T_mid", "u", "v", "qv", "qc", "qi", "qr", "ps", "phis")}
    mass, energy = np.zeros(ncol), np.zeros(ncol)
    lib.p3o_column_budget(*[(_ptr(a[k]) if a[k] is not None else None) for k in a], ncol, nlev,
                          _ptr(mass), _ptr(energy))
    return mass, energy


CLD_FRACTION_OUT = ("cldfrac_ice", "cldfrac_tot", "cldfrac_ice_for_analysis", "cldfrac_tot_for_analysis")


def cld_fraction(qi, cldfrac_liq, ice_threshold=1e-12, ice_4out_threshold=1e-5, math="glibc"):
    """The CldFraction process (p3o_cld_fraction) on level-major (nlev, ncol) arrays:
    a dict of its four outputs, CLD_FRACTION_OUT."""
    lib = load_oracle(math)
    qi = np.ascontiguousarray(qi, dtype=np.float64)
    liq = np.ascontiguousarray(cldfrac_liq, dtype=np.float64)
    assert qi.shape == liq.shape and qi.ndim == 2
    nlev, ncol = qi.shape
    out = {k: np.zeros_like(qi) for k in CLD_FRACTION_OUT}
    lib.p3o_cld_fraction(_ptr(qi), _ptr(liq), ncol, nlev, float(ice_threshold), float(ice_4out_threshold),
                         *[_ptr(out[k]) for k in CLD_FRACTION_OUT])
    return out


def shoc_postprocess(qc, qc2, cldfrac_liq=None, math="glibc"):
    """SHOC's post-process (p3o_shoc_postprocess) on (nlev, ncol) arrays: returns
    (inv_qc_relvar, the clipped cldfrac_liq or None)."""
    lib = load_oracle(math)
    qc = np.ascontiguousarray(qc, dtype=np.float64)
    qc2 = np.ascontiguousarray(qc2, dtype=np.float64)
    assert qc.shape == qc2.shape and qc.ndim == 2
    nlev, ncol = qc.shape
    liq = None if cldfrac_liq is None else np.array(cldfrac_liq, dtype=np.float64, order="C")
    r = np.zeros_like(qc)
    lib.p3o_shoc_postprocess(_ptr(qc), _ptr(qc2), None if liq is None else _ptr(liq), _ptr(r), ncol, nlev)
    return r, liq


def conservation_errors(m0, e0, m1, e1, vapor, water, ice, heat, dt, math="glibc"):
    lib = load_oracle(math)
    ncol = len(m0)
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (m0, e0, m1, e1, vapor, water, ice, heat)]
    me, ee = np.zeros(ncol), np.zeros(ncol)
    lib.p3o_conservation_errors(*[_ptr(x) for x in arrs], ncol, float(dt), _ptr(me), _ptr(ee))
    return me, ee


# ---------------------------------------------------------------- Fortran reference

def ref_available() -> bool:
    return os.path.exists(REF_SO)


def load_ref(table_dir: str):
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"{REF_SO} is not built (run `make -C oracle ref`)")
        lib = C.CDLL(REF_SO)
        lib.p3ref_init.restype = C.c_int
        lib.p3ref_init.argtypes = [C.c_char_p, C.c_int]
        d = table_dir.encode()
        info = lib.p3ref_init(d, len(d))
        if info != 0:
            raise RuntimeError("p3ref_init failed")
        lib.p3ref_get_tables.argtypes = [_dp] * 6
        lib.p3ref_lookup_ice.argtypes = [C.c_double] * 4 + [C.c_int] + [_ip] * 3 + [_dp] * 4
        lib.p3ref_lookup_coll.argtypes = [C.c_double] * 6 + [C.c_int, _ip, _dp, _dp]
        _ref = lib
    return _ref


def ref_tables(table_dir: str) -> dict:
    """The tables exactly as the Fortran reference holds them after p3_init."""
    lib = load_ref(table_dir)
    t = dict(ice=np.zeros((5, 4, 50, 12)), coll=np.zeros((5, 4, 50, 30, 2)),
             vn=np.zeros((300, 10)), vm=np.zeros((300, 10)), revap=np.zeros((300, 10)),
             mu_r=np.zeros(150))
    lib.p3ref_get_tables(*[_ptr(t[k]) for k in ("ice", "coll", "vn", "vm", "revap", "mu_r")])
    # dnu is a constant of micro_p3_utils_init (micro_p3_utils.F90:185-200)
    t["dnu"] = np.array([0.000, -0.557, -0.430, -0.307, -0.186, -0.067, -0.050, -0.167,
                         -0.282, -0.397, -0.512, -0.626, -0.739, -0.853, -0.966, -0.966])
    return t


def run_ref(inputs: dict, table_dir: str, dt: float, predict_nc=True, prescribed_ccn=False,
            it=1):
    """The reference Fortran p3_main on copies of ``inputs`` (FP64)."""
    lib = load_ref(table_dir)
    nlev, ncol = inputs["qc"].shape
    f = {k: np.array(inputs[k], dtype=np.float64, order="C") for k in STATE + INPUTS}
    z = lambda: np.zeros((nlev, ncol))
    for k in ("diag_eff_radius_qc", "diag_eff_radius_qi", "diag_eff_radius_qr", "rho_qi",
              "qv2qi_depos_tend", "precip_total_tend", "nevapr", "qr_evap_tend", "mu_c",
              "lamc") + HIST:
        f[k] = z()
    f["precip_liq_surf"] = np.zeros(ncol)
    f["precip_ice_surf"] = np.zeros(ncol)
    f["precip_liq_flux"] = np.zeros((nlev + 1, ncol))
    f["precip_ice_flux"] = np.zeros((nlev + 1, ncol))
    order = ("qc", "nc", "qr", "nr", "th", "qv", "qi", "qm", "ni", "bm", "pres", "dz",
             "nc_nuceat_tend", "nccn", "ni_activated", "inv_qc_relvar", "precip_liq_surf",
             "precip_ice_surf", "diag_eff_radius_qc", "diag_eff_radius_qi",
             "diag_eff_radius_qr", "rho_qi", "dpres", "inv_exner", "qv2qi_depos_tend",
             "precip_total_tend", "nevapr", "qr_evap_tend", "precip_liq_flux",
             "precip_ice_flux", "cld_frac_r", "cld_frac_l", "cld_frac_i", "mu_c", "lamc",
             "liq_ice_exchange", "vap_liq_exchange", "vap_ice_exchange", "qv_prev", "t_prev")
    lib.p3ref_main.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int] + \
        [_dp] * len(order)
    lib.p3ref_main(ncol, nlev, float(dt), int(it), int(predict_nc), int(prescribed_ccn),
                   *[_ptr(f[k]) for k in order])
    return f
