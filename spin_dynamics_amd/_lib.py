"""ctypes binding of libspinchem_hip.so (include/spinchem.h).

The library is built in-tree (spin_dynamics_amd/lib/); there is no fallback:
if it is missing, or no GPU is visible, the calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libspinchem_hip.so")

SPC_OK = 0
SPC_ERR_INVALID = 1
SPC_ERR_UNSUPPORTED = 2
SPC_ERR_OOM = 3
SPC_ERR_HIP = 4
SPC_ERR_BREAKDOWN = 5
SPC_ERR_NO_DEVICE = 6
SPC_MAX_KRYLOV = 32

STATE = {"singlet": 0, "T0": 1, "Tp": 2, "Tm": 3}
VEC_SUZ, VEC_PHASE, VEC_INJECTED = 0, 1, 2
PROF_FAMILIES = ["pass1_spmm_dots", "pass2_update", "final_dots", "psi_update", "reduce", "coeffs", "expm", "step",
                 "gram", "sstep_dots", "sstep_form"]

# Exported symbols, in the order of include/spinchem.h.
EXPORTS = [
    "spc_ctx_create", "spc_ctx_destroy", "spc_last_error", "spc_version", "spc_system_set", "spc_system_set_block",
    "spc_shrink", "spc_symmetry_blocks", "spc_system_info",
    "spc_hamiltonian_copy", "spc_n_steps", "spc_trace_sample", "spc_trace_sample_noalloc", "spc_trace_sample_seeds", "spc_exact_trace", "spc_expm_small",
    "spc_krylov_basis", "spc_spmm", "spc_trace_sample_grad", "spc_orientation_batch", "spc_batch_tune", "spc_batch_tune_adaptive", "spc_pairs_batch", "spc_semiclassical_run", "spc_semiclassical_tangent", "spc_last_run_stats", "spc_profile_enable", "spc_profile_read", "spc_ini_open", "spc_ini_close",
    "spc_ini_num_sections", "spc_ini_has_section", "spc_ini_count_values", "spc_ini_get_string",
    "spc_ini_get_doubles", "spc_ini_get_int64s",
]

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


class SpinSystem(ctypes.Structure):
    _fields_ = [("n_nuc1", ctypes.c_int32), ("n_nuc2", ctypes.c_int32), ("mult", _i32p), ("hfc", _dp),
                ("J", ctypes.c_double), ("D", ctypes.c_double * 9), ("kS", ctypes.c_double),
                ("kT", ctypes.c_double), ("B", ctypes.c_double), ("field_dir", ctypes.c_double * 3),
                ("iso_rule", ctypes.c_int32)]


class SimParams(ctypes.Structure):
    _fields_ = [("n_samples", ctypes.c_int32), ("n_krylov", ctypes.c_int32), ("t_end", ctypes.c_double),
                ("dt", ctypes.c_double), ("tol", ctypes.c_double), ("init_state", ctypes.c_int32),
                ("vec_kind", ctypes.c_int32)]


class ScSystem(ctypes.Structure):
    _fields_ = [("n_nuc1", ctypes.c_int32), ("n_nuc2", ctypes.c_int32), ("mult", _i32p), ("a_iso", _dp),
                ("B", ctypes.c_double), ("field_dir", ctypes.c_double * 3), ("J", ctypes.c_double),
                ("kS", ctypes.c_double), ("kT", ctypes.c_double)]


class SpinchemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"spinchem error {code}: {msg}")
        self.code = code


_lib = None


def load() -> ctypes.CDLL:
    """Loads the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is missing: build it with `python -m spin_dynamics_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    lib.spc_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    lib.spc_ctx_destroy.argtypes = [vp]
    lib.spc_ctx_destroy.restype = None
    lib.spc_last_error.argtypes = [vp]
    lib.spc_last_error.restype = ctypes.c_char_p
    lib.spc_version.restype = ctypes.c_char_p
    lib.spc_system_set.argtypes = [vp, ctypes.POINTER(SpinSystem)]
    lib.spc_system_set_block.argtypes = [vp, ctypes.POINTER(SpinSystem)]
    lib.spc_shrink.argtypes = [ctypes.c_int32, _dp, _i32p, _dp, _i32p]
    lib.spc_symmetry_blocks.argtypes = [ctypes.c_int32, _i32p, ctypes.c_int64, _i64p, _i32p, _i64p, _i64p]
    lib.spc_system_info.argtypes = [vp, _i64p, _i64p, _i64p]
    lib.spc_hamiltonian_copy.argtypes = [vp, _i32p, _i32p, _dp]
    lib.spc_n_steps.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.spc_n_steps.restype = ctypes.c_int64
    lib.spc_trace_sample.argtypes = [vp, ctypes.POINTER(SimParams), _i64p, _dp, _dp, _dp, _dp, _dp,
                                     ctypes.c_int64, _dp, _dp]
    lib.spc_trace_sample_noalloc.argtypes = lib.spc_trace_sample.argtypes
    lib.spc_trace_sample_seeds.argtypes = [vp, ctypes.POINTER(SimParams), _i64p, _dp, _dp, _dp, _dp, ctypes.c_int64,
                                           _dp, _dp]
    lib.spc_exact_trace.argtypes = [vp, ctypes.POINTER(SimParams), _dp, _dp, _dp, _dp, ctypes.c_int64, _dp, _dp]
    lib.spc_expm_small.argtypes = [vp, ctypes.c_int32, _dp, _dp]
    lib.spc_krylov_basis.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, _dp, _dp, _dp, _dp]
    lib.spc_spmm.argtypes = [vp, ctypes.c_int32, _dp, _dp]
    lib.spc_trace_sample_grad.argtypes = [vp, ctypes.POINTER(SimParams), _i64p, _dp, ctypes.c_int32, _dp,
                                          ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.spc_orientation_batch.argtypes = [vp, ctypes.POINTER(SimParams), _i64p, ctypes.c_int32, _dp, ctypes.c_int32,
                                          ctypes.c_int32, _dp, _dp, ctypes.c_int64]
    lib.spc_batch_tune.argtypes = [vp, ctypes.POINTER(SimParams), _i64p, ctypes.c_int32, _dp, ctypes.c_double,
                                   _i32p, _i32p, _dp]
    lib.spc_batch_tune_adaptive.argtypes = [vp, ctypes.POINTER(SimParams), _i64p, ctypes.c_int32, _dp,
                                            ctypes.c_double, _i32p, _i32p, _dp, _dp]
    lib.spc_pairs_batch.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(SpinSystem), ctypes.POINTER(SimParams), _i64p,
                                    ctypes.c_int32, _dp, _i32p, _dp, _i32p, _dp]
    lib.spc_semiclassical_run.argtypes = [vp, ctypes.POINTER(ScSystem), ctypes.c_int64, _i64p, ctypes.c_double,
                                          ctypes.c_int64, _dp, _dp, _dp, _dp]
    lib.spc_semiclassical_tangent.argtypes = [vp, ctypes.POINTER(ScSystem), ctypes.c_int64, _i64p, ctypes.c_double,
                                              ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.spc_last_run_stats.argtypes = [vp, _dp, _i64p, _i64p]
    lib.spc_profile_enable.argtypes = [vp, ctypes.c_int]
    lib.spc_profile_read.argtypes = [vp, ctypes.c_int, _dp, _i64p, _dp]
    lib.spc_ini_open.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
    lib.spc_ini_close.argtypes = [vp]
    lib.spc_ini_close.restype = None
    lib.spc_ini_num_sections.argtypes = [vp]
    lib.spc_ini_has_section.argtypes = [vp, ctypes.c_char_p]
    lib.spc_ini_count_values.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p]
    lib.spc_ini_get_string.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32]
    lib.spc_ini_get_doubles.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, _dp, ctypes.c_int32]
    lib.spc_ini_get_int64s.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, _i64p, ctypes.c_int32]
    for name in EXPORTS:  # fail loudly on a stale library
        getattr(lib, name)
    _lib = lib
    return lib


def ptr(a: np.ndarray, kind=_dp):
    return None if a is None else a.ctypes.data_as(kind)


class Context:
    """Owns an spc_ctx (device, stream and device buffers)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self.h = ctypes.c_void_p()
        rc = self.lib.spc_ctx_create(device, ctypes.byref(self.h))
        if rc != SPC_OK:
            raise SpinchemError(rc, "spc_ctx_create failed" + (" (no HIP device)" if rc == SPC_ERR_NO_DEVICE else ""))

    def check(self, rc: int):
        if rc != SPC_OK:
            raise SpinchemError(rc, self.lib.spc_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib.spc_ctx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
