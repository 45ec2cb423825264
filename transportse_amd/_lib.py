"""ctypes binding of ``libtse_hip.so`` (the C-ABI declared in include/tse_hip.h).

The library is built in-tree by ``transportse_amd.build``. Loading fails with
an ImportError naming the missing file; there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtse_hip.so")

ABI_VERSION = 5

FORM_FAITHFUL64 = 0
FORM_PLAIN64 = 1
FORM_MIXED = 2

_dp = ctypes.c_void_p  # device pointers travel as integers


class Result(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("dt", ctypes.c_double),
                ("converged", ctypes.c_int32), ("num_it", ctypes.c_int32),
                ("num_halvings", ctypes.c_int32), ("status", ctypes.c_int32),
                ("last_diff", ctypes.c_double)]


class SorParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("sor", ctypes.c_double),
                ("omega", ctypes.c_double), ("inc_it", ctypes.c_double),
                ("max_dt", ctypes.c_double), ("min_dt", ctypes.c_double),
                ("conv", ctypes.c_double), ("max_it", ctypes.c_int32),
                ("form", ctypes.c_int32)]


P_FIELDS = ("P", "Pn", "c_f", "nue", "rho_f", "phi", "Kx", "Ky", "Kz", "BP",
            "WP", "c_r")
C_FIELDS = ("C", "Cn", "D_fx", "D_fy", "D_fz", "phi", "rho_f", "qx", "qy",
            "qz", "BC", "WC", "RC", "e", "lf_l", "lf_t")
T_FIELDS = ("T", "Tn", "l_f", "cp_f", "rho_f", "phi", "qx", "qy", "qz", "BT",
            "WT", "e", "l_r", "cp_r", "rho_r", "lt_l", "lt_t")


class SweepPlan(ctypes.Structure):
    _fields_ = [("ns", ctypes.c_int32), ("by", ctypes.c_int32),
                ("ry", ctypes.c_int32), ("zchunk", ctypes.c_int32),
                ("variant", ctypes.c_int32)]


CYCLE_FIXED, CYCLE_ADAPTIVE = 0, 1


class FastParams(ctypes.Structure):
    _fields_ = [("form", ctypes.c_int32), ("plan", SweepPlan),
                ("inner", ctypes.c_int32), ("max_sweeps", ctypes.c_int32),
                ("cycle_mode", ctypes.c_int32), ("adapt_sor", ctypes.c_int32),
                ("sor", ctypes.c_double), ("tol", ctypes.c_double)]


class FastResult(ctypes.Structure):
    _fields_ = [("sweeps", ctypes.c_int32), ("cycles", ctypes.c_int32),
                ("converged", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("initial_rel", ctypes.c_double),
                ("rel_residual", ctypes.c_double),
                ("contraction", ctypes.c_double), ("sor", ctypes.c_double)]


class EosParams(ctypes.Structure):
    _fields_ = [("rho0", ctypes.c_double), ("drho_dc", ctypes.c_double),
                ("lf0", ctypes.c_double), ("dlf_dt", ctypes.c_double),
                ("t_ref", ctypes.c_double)]


class PFields(ctypes.Structure):
    _fields_ = [(n, _dp) for n in P_FIELDS]


class CFields(ctypes.Structure):
    _fields_ = [(n, _dp) for n in C_FIELDS] + [("imblflo", ctypes.c_int32)]


class TFields(ctypes.Structure):
    _fields_ = [(n, _dp) for n in T_FIELDS] + [("imblflo", ctypes.c_int32)]


# name -> (restype, argtypes); every symbol include/tse_hip.h declares
_I, _D, _L, _V = ctypes.c_int, ctypes.c_double, ctypes.c_int64, ctypes.c_void_p
_DPH = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "tse_abi_version": (_I, []),
    "tse_sor_model": (_I, [_D, _D, _DPH]),
    "tse_last_error": (ctypes.c_char_p, []),
    "tse_ctx_create": (_I, [ctypes.POINTER(_V), _I, _V]),
    "tse_ctx_destroy": (_I, [_V]),
    "tse_ctx_set_stream": (_I, [_V, _V]),
    "tse_ctx_synchronize": (_I, [_V]),
    "tse_set_grid": (_I, [_V, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                          _DPH, _DPH, _DPH, _D, _D, _D]),
    "tse_field_elems": (_L, [_V]),
    "tse_to_split": (_I, [_V, _dp, _dp, _L, _L, _L]),
    "tse_from_split": (_I, [_V, _dp, _L, _L, _L, _dp]),
    "tse_compute_p": (_I, [_V, ctypes.POINTER(PFields),
                           ctypes.POINTER(SorParams), ctypes.POINTER(Result)]),
    "tse_compute_p_noalloc": (_I, [_V, ctypes.POINTER(PFields),
                                   ctypes.POINTER(SorParams), ctypes.POINTER(Result)]),
    "tse_compute_q": (_I, [_V] + [_dp] * 10),
    "tse_flow_error": (_I, [_V] + [_dp] * 10 + [_D, _dp]),
    "tse_compute_c": (_I, [_V, ctypes.POINTER(CFields),
                           ctypes.POINTER(SorParams), ctypes.POINTER(Result)]),
    "tse_compute_t": (_I, [_V, ctypes.POINTER(TFields),
                           ctypes.POINTER(SorParams), ctypes.POINTER(Result)]),
    "tse_l2_norm": (_I, [_V, _dp, _dp, _DPH]),
    "tse_l2_norm_async": (_I, [_V, _dp, _dp]),
    "tse_reduction_result": (_I, [_V, _DPH, ctypes.c_int32]),
    "tse_smolarkiewicz": (_I, [_V] + [_dp] * 7 + [_D, _D, _D]),
    "tse_host_l2_norm_seq": (_D, [_DPH, _DPH, _L, ctypes.POINTER(ctypes.c_int32)]),
    "tse_p_sweeps": (_I, [_V, ctypes.POINTER(PFields), _D, _D, ctypes.c_int32,
                           ctypes.c_int32]),
    "tse_fast_p_setup": (_I, [_V, ctypes.POINTER(PFields), _D]),
    "tse_fast_p_solve": (_I, [_V, ctypes.POINTER(FastParams),
                              ctypes.POINTER(FastResult)]),
    "tse_fast_p_residual": (_I, [_V, _DPH]),
    "tse_fast_p_operator": (_I, [_V, _dp, _dp, _dp, _dp, _dp]),
    "tse_fast_p_result": (_I, [_V, _dp]),
    "tse_fast_p_sweeps": (_I, [_V, ctypes.POINTER(SweepPlan), ctypes.c_int32,
                               _D, ctypes.c_int32]),
    "tse_fast_c_setup": (_I, [_V, ctypes.POINTER(CFields), _D, _D]),
    "tse_fast_t_setup": (_I, [_V, ctypes.POINTER(TFields), _D, _D]),
    "tse_fast_tr_solve": (_I, [_V, ctypes.POINTER(FastParams),
                               ctypes.POINTER(FastResult)]),
    "tse_fast_tr_residual": (_I, [_V, _DPH]),
    "tse_fast_tr_result": (_I, [_V, _dp]),
    "tse_fast_tr_sweeps": (_I, [_V, ctypes.POINTER(SweepPlan), ctypes.c_int32,
                                _D, ctypes.c_int32]),
    "tse_batch_create": (_I, [ctypes.POINTER(_V), ctypes.c_int32, _V]),
    "tse_batch_destroy": (_I, [_V]),
    "tse_batch_begin": (_I, [_V, ctypes.c_int32, ctypes.POINTER(FastParams)]),
    "tse_batch_add": (_I, [_V, _V]),
    "tse_batch_rearm": (_I, [_V, ctypes.c_int32]),
    "tse_batch_rearm_many": (_I, [_V, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]),
    "tse_batch_writeback": (_I, [_V, ctypes.c_int32]),
    "tse_batch_remove": (_I, [_V, ctypes.c_int32]),
    "tse_batch_set_sor": (_I, [_V, ctypes.c_int32, _D]),
    "tse_batch_cycle": (_I, [_V, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "tse_batch_run": (_I, [_V, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "tse_batch_status": (_I, [_V, ctypes.c_int32, ctypes.POINTER(FastResult)]),
    "tse_linear_eos": (_I, [_V, _dp, _dp, _dp, _dp, ctypes.POINTER(EosParams)]),
    "tse_graph_begin": (_I, [_V]),
    "tse_graph_end": (_I, [_V, ctypes.POINTER(_V)]),
    "tse_graph_launch": (_I, [_V, _V]),
    "tse_graph_destroy": (_I, [_V]),
    "tse_fast_mark_loaded": (_I, [_V, ctypes.c_int32]),
    "tse_fast_current": (_I, [_V, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "tse_mass_sums": (_I, [_V, _dp, _dp, _dp, _dp, _DPH]),
    "tse_mass_sums_async": (_I, [_V, _dp, _dp, _dp, _dp]),
    "tse_peclet_c": (_I, [_V, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int32, _dp, _dp,
                          _DPH]),
    "tse_peclet_t": (_I, [_V, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                          _DPH]),
    "tse_cfl_maxima": (_I, [_V, _dp, _dp, _dp, _dp, _DPH]),
}

_lib = None


class TseError(RuntimeError):
    """Error status returned by the C-ABI."""


def load():
    """Load libtse_hip.so (no fallback: raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m transportse_amd.build`"
            " (the HIP library is required; there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.tse_abi_version() != ABI_VERSION:
        raise ImportError("libtse_hip.so ABI version mismatch; rebuild it")
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().tse_last_error()
        raise TseError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
