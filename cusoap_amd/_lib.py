"""ctypes binding of libcusoap_amd.so (the C ABI declared in include/cusoap_amd.h)."""

from __future__ import annotations

import ctypes as C
import os
from typing import Optional

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcusoap_amd.so")

OK, ERR_ARG, ERR_HIP, ERR_UNSUPPORTED, ERR_CAPACITY, ERR_WORKSPACE = range(6)
RBF_GTO, RBF_POLY = 0, 1
W_NONE, W_POLY, W_POW, W_EXP = 0, 1, 2, 3
PREC_F64, PREC_F32, PREC_MIXED = 0, 1, 2
DT_F64, DT_F32 = 0, 1
MODE_PS, MODE_CC = 0, 1
MFMA_F16, MFMA_BF16 = 0, 1

p = C.c_void_p
i32 = C.c_int32
i64 = C.c_int64
f64 = C.c_double
sz = C.c_size_t


class Basis(C.Structure):
    _fields_ = [("n_species", i32), ("n_max", i32), ("l_max", i32), ("rbf", i32),
                ("eta", f64), ("cutoff", f64), ("weighting", i32), ("quad_n", i32),
                ("w_r0", f64), ("w_c", f64), ("w_d", f64), ("w_m", f64), ("w0", f64),
                ("gto_gamma", p), ("gto_pref", p), ("gto_beta", p),
                ("poly_rx", p), ("poly_rxw", p), ("poly_gss", p), ("poly_B0", p), ("poly_Bh0", p),
                ("self_l0", p), ("rad_table", p), ("rad_intervals", i32), ("rad_rmax", f64),
                ("rad_cheb", p), ("cheb_terms", i32), ("cheb_split", f64),
                ("cheb_frag", p), ("cheb_rowinv", p)]


class NLPlan(C.Structure):
    _fields_ = [("origin", f64 * 3), ("cell", f64 * 3), ("dims", i32 * 3), ("n_cells", i32),
                ("n_ext", i64), ("ws_bytes", sz)]


# name -> argtypes (restype is int for all but cusoap_last_error)
_SIGS = {
    "cusoap_version": [],
    "cusoap_max_l": [],
    "cusoap_max_n": [],
    "cusoap_nl_plan": [p, i32, i32, i32, f64, p, C.POINTER(NLPlan), p],
    "cusoap_nl_count": [C.POINTER(NLPlan), p, i32, p, i32, p, i32, f64, i32, p, sz, p,
                        C.POINTER(i64), C.POINTER(i32), p],
    "cusoap_nl_fill": [C.POINTER(NLPlan), p, i32, p, i32, p, i32, f64, i32, p, sz, p, i32, p, p],
    "cusoap_nl_build": [p, i32, p, i32, p, i32, f64, i32, p, p, sz, p, p, i64, C.POINTER(NLPlan), C.POINTER(i64),
                        C.POINTER(i32), p],
    "cusoap_coefficients": [C.POINTER(Basis), i32, p, i32, p, p, i32, p, i32, p, p, p, i32, p, p],
    "cusoap_power_spectrum": [C.POINTER(Basis), p, i32, i32, p, i32, p, i32, p],
    "cusoap_cc_projection": [C.POINTER(Basis), p, i32, i32, p, i32, p],
    "cusoap_power_spectrum_mfma": [C.POINTER(Basis), p, i32, i32, p, i32, i32, i32, p, i32, p],
    "cusoap_jacobian_workspace_size": [C.POINTER(Basis), i32, i64, i32, i32, C.POINTER(sz)],
    "cusoap_jacobian_dense": [C.POINTER(Basis), i32, p, i32, p, p, i32, p, i32, p, p, p, i64, p, p, i32,
                              p, i32, p, sz, p, i32, i32, p],
    "cusoap_vjp_workspace_size": [C.POINTER(Basis), i32, C.POINTER(sz)],
    "cusoap_vjp_ordered_workspace_size": [C.POINTER(Basis), i32, i32, i64, C.POINTER(sz)],
    "cusoap_vjp": [C.POINTER(Basis), i32, i32, p, i32, p, p, i32, p, i32, p, p, p, p, p, i32, p, p, i32, i32,
                   i32, p, sz, p, p],
    "cusoap_edge_reverse_workspace_size": [i32, C.POINTER(sz)],
    "cusoap_edge_reverse": [p, i32, p, i32, p, i32, p, p, p, sz, p, p],
    "cusoap_vjp_ordered": [C.POINTER(Basis), i32, i32, p, i32, p, p, i32, p, i32, p, p, p, p, p, i64, p, i32, p, p, i32,
                           i32, i32, p, sz, p, p],
    "cusoap_vjp_noalloc": [C.POINTER(Basis), i32, i32, p, i32, p, p, i32, p, i32, p, p, p, p, p, i32, p, p, i32,
                           i32, i32, p, p],
    "cusoap_lowdin_invsqrt_f64": [p, i32, i32, f64, p, p],
    "cusoap_radial_exact": [C.POINTER(Basis), p, i32, p, p, p],
    "cusoap_mlp_eval": [p, i32, i32, i32, p, i32, p, p, p],
    "cusoap_mlp_train_workspace_size": [i32, i32, i32, i32, C.POINTER(sz)],
    "cusoap_mlp_loss_grad": [p, i32, i32, i32, p, p, p, i32, i64, p, p, i32, C.c_uint32, i64, p, sz, p, p, p],
    "cusoap_adam_step": [p, p, p, p, i64, C.c_float, C.c_float, C.c_float, C.c_float, i64, p],
    "cusoap_mlp_refit_workspace_size": [i32, i32, i32, i32, C.POINTER(sz)],
    "cusoap_mlp_refit_output": [p, i32, i32, i32, p, p, p, i32, i64, p, f64, p, sz, p],
    "cusoap_radial_table_workspace_size": [C.POINTER(Basis), i32, C.POINTER(sz)],
    "cusoap_radial_table": [C.POINTER(Basis), p, i32, p, p, f64, i32, f64, p, sz, p, p],
    "cusoap_profile_get": [C.c_char_p, C.POINTER(f64), C.POINTER(i32)],
}

_lib: Optional[C.CDLL] = None


def load() -> C.CDLL:
    """Load the in-tree library; raise if it is missing (there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m cusoap_amd.build` "
                           "(cusoap_amd has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.cusoap_last_error.argtypes = []
    lib.cusoap_last_error.restype = C.c_char_p
    lib.cusoap_profile_enable.argtypes = [i32]
    lib.cusoap_profile_enable.restype = None
    lib.cusoap_mlp_param_count.argtypes = [i32, i32]
    lib.cusoap_mlp_param_count.restype = i64
    _lib = lib
    return lib


def exported_symbols():
    return ["cusoap_last_error", "cusoap_profile_enable", "cusoap_mlp_param_count", *_SIGS.keys()]


def profile_enable(on: bool) -> None:
    load().cusoap_profile_enable(1 if on else 0)


def profile_get(name: str):
    """(total milliseconds, launch count) of the kernels recorded under ``name``; clears them."""
    ms, n = C.c_double(0.0), C.c_int32(0)
    check(load().cusoap_profile_get(name.encode(), C.byref(ms), C.byref(n)), "profile_get")
    return ms.value, n.value


class CusoapError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    if rc == OK:
        return
    msg = load().cusoap_last_error().decode(errors="replace")
    text = f"{what}: {msg}"
    if rc == ERR_ARG:
        raise ValueError(text)
    if rc == ERR_UNSUPPORTED:
        raise NotImplementedError(text)
    raise CusoapError(text)
