"""ctypes binding of liblettuce_amd.so (include/lettuce_amd.h).

The library is the only compute path of this package: every operation that
touches populations calls into it, and every call raises ``RuntimeError`` when
the library is missing or reports an error.  There is no fallback.

Replaces the pybind11 module that lettuce generates and installs per
configuration (lettuce/cuda_native/_template.py:42-53,
lettuce/cuda_native/_generator.py:86-127): here one prebuilt library serves
every configuration.
"""
import ctypes
import os
import threading

import torch

__all__ = ["lib", "LbmStepArgs", "LbmCellClass", "LbmLearned", "check",
           "stream_handle", "tuning", "STENCIL_IDS", "STORAGE_IDS", "LIB_PATH",
           "ABI_VERSION"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        "liblettuce_amd.so")
ABI_VERSION = 3

# enum values of include/lettuce_amd.h
D2Q9, D3Q19, D3Q27 = 0, 1, 2
F32, F64, F16_SHIFTED, BF16_SHIFTED = 0, 1, 2, 3
COLL_NONE, COLL_BGK, COLL_TRT, COLL_KBC, COLL_LEARNED, COLL_BGK_GUO = 0, 1, 2, 3, 4, 5
OP_COLLIDE, OP_BOUNCE_BACK, OP_EQUILIBRIUM, OP_OUTLET, OP_KEEP = 0, 1, 2, 3, 4
OP_ANTI_BOUNCE_BACK = 5
CLASS_SOLID, CLASS_NEAR_SOLID, CLASS_BC_FIELD = 0x01, 0x02, 0x04
SLAB_ALL, SLAB_INTERIOR, SLAB_FACES = 0, 1, 2

STENCIL_IDS = {"D2Q9": D2Q9, "D3Q19": D3Q19, "D3Q27": D3Q27}
STORAGE_IDS = {"f32": F32, "f64": F64, "f16_shifted": F16_SHIFTED,
               "bf16_shifted": BF16_SHIFTED}


class LbmCellClass(ctypes.Structure):
    _fields_ = [("op", ctypes.c_uint8), ("flags", ctypes.c_uint8),
                ("dir", ctypes.c_int8 * 3), ("pad", ctypes.c_uint8 * 3),
                ("no_stream", ctypes.c_uint32), ("pad2", ctypes.c_uint32),
                ("rho", ctypes.c_double), ("u", ctypes.c_double * 3)]


LEARNED_PACKED_BYTES = 17408
LEARNED_HIDDEN = 32
LEARNED_PARAMS = 1379    # [w1 (32x6), b1, w2 (32x32), b2, w3 (3x32), b3]
LEARNED_RECORD = 137     # x, a1, a2, dL/dz1, dL/dz2, dL/do per cell


class LbmLearned(ctypes.Structure):
    _fields_ = [("packed", ctypes.c_void_p),
                ("in_scale", ctypes.c_float), ("rate_lo", ctypes.c_float),
                ("rate_hi", ctypes.c_float), ("stable_lo", ctypes.c_float),
                ("stable_hi", ctypes.c_float), ("pad", ctypes.c_int32)]


class LbmStepArgs(ctypes.Structure):
    _fields_ = [("abi_version", ctypes.c_int32), ("stencil", ctypes.c_int32),
                ("storage", ctypes.c_int32), ("collision", ctypes.c_int32),
                ("streaming", ctypes.c_int32), ("n_classes", ctypes.c_int32),
                ("n", ctypes.c_int64 * 3), ("stride", ctypes.c_int64 * 4),
                ("f_in", ctypes.c_void_p), ("f_out", ctypes.c_void_p),
                ("cell_class", ctypes.c_void_p), ("classes", ctypes.c_void_p),
                ("bc_field", ctypes.c_void_p), ("tau", ctypes.c_double),
                ("tau_minus", ctypes.c_double), ("rho0_shift", ctypes.c_double),
                ("learned", ctypes.POINTER(LbmLearned)),
                ("force_tau", ctypes.c_double),
                ("accel", ctypes.c_double * 3),
                ("sr_seed", ctypes.c_uint32), ("sr_step", ctypes.c_uint32)]


_I64_3 = ctypes.c_int64 * 3
_I64_4 = ctypes.c_int64 * 4
_vp, _i32, _i64, _dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double

_SIGNATURES = {
    "lbm_abi_version": ([], ctypes.c_int),
    "lbm_last_error": ([], ctypes.c_char_p),
    "lbm_struct_size": ([ctypes.c_int], _i64),
    "lbm_step": ([ctypes.POINTER(LbmStepArgs), _vp], ctypes.c_int),
    "lbm_run": ([ctypes.POINTER(LbmStepArgs), _i64, _vp], ctypes.c_int),
    "lbm_equilibrium": ([_i32, _i32, _I64_3, _I64_4, _vp, _vp, _dbl, _vp, _vp],
                        ctypes.c_int),
    "lbm_decode": ([_i32, _i32, _I64_3, _I64_4, _vp, _dbl, _vp, _i32, _vp],
                   ctypes.c_int),
    "lbm_encode": ([_i32, _i32, _I64_3, _I64_4, _vp, _i32, _dbl, _vp, _vp],
                   ctypes.c_int),
    "lbm_moments_f64": ([_i32, _i32, _I64_3, _I64_4, _vp, _dbl, _vp, _vp, _vp],
                        ctypes.c_int),
    "lbm_reduce_workspace_bytes": ([_I64_3], _i64),
    "lbm_reduce_f64": ([_i32, _i32, _I64_3, _I64_4, _vp, _dbl, _vp, _vp, _vp,
                        _vp], ctypes.c_int),
    "lbm_enstrophy_f64": ([_i32, _I64_3, _vp, _vp, _vp, _vp], ctypes.c_int),
    "lbm_run_plan": ([ctypes.POINTER(LbmStepArgs), _i64, _vp], ctypes.c_int),
    "lbm_scan_f64": ([_i32, _i32, _I64_3, _I64_4, _vp, _dbl, _vp, _vp, _vp], ctypes.c_int),
    "lbm_link_force_f64": ([_i32, _i32, _I64_3, _I64_4, _vp, _dbl, _vp, _i64,
                            _vp, _vp, _vp], ctypes.c_int),
    "lbm_learned_pack": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "lbm_stream_copy": ([_vp, _vp, _i64, _i32, _vp], ctypes.c_int),
    "lbm_learned_train_forward": ([_i32, _i32, _vp, _dbl, _vp, _dbl, _dbl, _dbl, _vp, _vp,
                                   _vp], ctypes.c_int),
    "lbm_learned_train_backward": ([_i32, _i32, _vp, _dbl, _vp, _dbl, _dbl, _dbl, _vp, _vp,
                                    _i32, _vp, _vp, _vp], ctypes.c_int),
    "lbm_unpull": ([_i32, _i32, _vp, _vp, _vp, _vp], ctypes.c_int),
    "lbm_learned_grad_workspace": ([_i32, _i64], _i64),
    "lbm_learned_grad_reduce": ([_i32, _i64, _vp, _vp, _vp, _vp], ctypes.c_int),
    "lbm_halo_planes": ([_i32], _i64),
    "lbm_halo_pack": ([_i32, _i32, _I64_3, _I64_4, _vp, _vp, _vp, _vp], ctypes.c_int),
    "lbm_slab_step": ([ctypes.POINTER(LbmStepArgs), _i32, _vp, _vp, _vp, _vp, _vp],
                      ctypes.c_int),
    "lbm_set_tuning": ([ctypes.c_char_p, _i32], ctypes.c_int),
    "lbm_reduce_f64_static": ([_i32, _i32, _I64_3, _I64_4, _vp, _dbl, _vp, _vp, _vp],
                              ctypes.c_int),
}

_lock = threading.Lock()
_lib = None


def lib():
    """The loaded library.  torch is imported first so that the library binds
    to the HIP runtime torch already loaded (one runtime per process)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"lettuce_amd native library not found at {LIB_PATH}; build "
                f"it with `python -m lettuce_amd._build` (needs hipcc)")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = res
        if handle.lbm_abi_version() != ABI_VERSION:
            raise RuntimeError("liblettuce_amd.so ABI version mismatch")
        for which, struct in enumerate((LbmStepArgs, LbmCellClass,
                                        LbmLearned)):
            if handle.lbm_struct_size(which) != ctypes.sizeof(struct):
                raise RuntimeError(
                    f"struct layout mismatch for {struct.__name__}: library "
                    f"{handle.lbm_struct_size(which)} B, binding "
                    f"{ctypes.sizeof(struct)} B")
        _lib = handle
        return _lib


def check(rc, what):
    if rc != 0:
        msg = lib().lbm_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


class tuning:
    """Context manager that sets launch-tuning knobs of the library
    (lbm_set_tuning: two_step, two_step_planes, two_step_barriers,
    wide_offsets, post_split) and restores their start values on exit.
    The knobs are process-wide; they exist for measurements and tests."""

    def __init__(self, **knobs):
        self.knobs = knobs

    def __enter__(self):
        for name, value in self.knobs.items():
            check(lib().lbm_set_tuning(name.encode(), int(value)), "lbm_set_tuning")
        return self

    def __exit__(self, *exc):
        for name in self.knobs:
            lib().lbm_set_tuning(name.encode(), -1)
        return False


def stream_handle(device):
    """hipStream_t of torch's current stream on `device`, as a void*."""
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def require_device_tensor(t, name):
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise RuntimeError(
            f"{name} must be a tensor on a ROCm GPU; lettuce_amd computes "
            f"only through its HIP kernels (got "
            f"{getattr(t, 'device', type(t))})")
