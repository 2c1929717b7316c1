"""ctypes binding of libefwi.so, the gfx950 HIP library of this package.

torch is imported first so that the HIP runtime torch loaded is the one the
library binds to (both carry the soname libamdhip64.so.7): device pointers
and streams then mean the same thing on both sides.

There is no CPU fallback: if libefwi.so is missing, importing this module
raises, and every op that needs it fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

import torch  # noqa: F401  (must precede loading libefwi.so)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libefwi.so")

OK = 0
KERNELS = ("stress_fwd", "velocity_fwd", "stress_bwd", "velocity_bwd", "velocity_adj", "stress_adj",
           "ring", "other", "fwd_fused", "bwd_fused")
CALC_MISFIT, CALC_GRADIENT, CALC_SYNTHETIC = 0, 1, 2
TUNED_KERNELS = ("stress_fwd", "velocity_fwd", "stress_bwd", "velocity_bwd")
# efwi_autotune's candidates, in its report order: flat sweep (zdepth < 16), then z-marching
TILE_CANDIDATES = ((1, 1), (2, 1), (4, 1), (1, 2), (1, 4), (2, 2), (4, 16), (4, 32), (4, 64), (8, 16), (8, 32), (8, 64))


class EfwiError(RuntimeError):
    """Error reported by libefwi.so (the reference would have called exit())."""


class Config(C.Structure):
    _fields_ = [("nz", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("npml", C.c_int), ("nt", C.c_int),
                ("dz", C.c_float), ("dx", C.c_float), ("dy", C.c_float), ("dt", C.c_float),
                ("f0", C.c_float), ("formulation", C.c_int), ("if_win", C.c_int), ("device", C.c_int),
                ("memory_mode", C.c_int), ("checkpoint_every", C.c_int)]


class Shot(C.Structure):
    _fields_ = [("src_z", C.c_int), ("src_x", C.c_int), ("src_y", C.c_int),
                ("stf", C.POINTER(C.c_float)), ("nrec", C.c_int), ("rec", C.POINTER(C.c_int)),
                ("win_start", C.POINTER(C.c_float)), ("win_end", C.POINTER(C.c_float)),
                ("weights", C.POINTER(C.c_float)), ("src_weight", C.c_float)]


# Every symbol include/efwi.h and include/efwi_nn.h declare, with its ctypes signature.
_vp, _fp, _dp, _ip = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int)
SIGNATURES = {
    "efwi_version": (C.c_int, []),
    "efwi_last_error": (C.c_char_p, []),
    "efwi_create": (C.c_int, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
    "efwi_destroy": (None, [_vp]),
    "efwi_set_model": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp]),
    "efwi_zero_gradient": (C.c_int, [_vp, _vp]),
    "efwi_run_shot": (C.c_int, [_vp, C.POINTER(Shot), C.c_int, _vp, _vp, _vp, _dp, _vp]),
    "efwi_get_gradient": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp]),
    "efwi_run_batch": (C.c_int, [_vp, C.POINTER(Shot), C.c_int, C.c_int, C.POINTER(C.c_void_p),
                                 C.POINTER(C.c_void_p), _dp, C.c_int, _vp]),
    "efwi_batch_gradient": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, C.c_int, _vp]),
    "efwi_batch_config": (C.c_int, [_vp, C.c_int, C.c_int]),
    "efwi_batch_group": (C.c_int, [_vp, C.c_int]),
    "efwi_device_bytes": (C.c_longlong, [_vp]),
    "efwi_dot64": (C.c_int, [_vp, _vp, C.c_longlong, _dp, _vp]),
    "efwi_set_tile": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "efwi_set_fusion": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "efwi_set_resident_segments": (C.c_int, [_vp, C.c_int]),
    "efwi_get_resident_segments": (C.c_int, [_vp, _vp, _vp]),
    "efwi_set_split": (C.c_int, [_vp, C.c_int]),
    "efwi_get_split": (C.c_int, [_vp, _ip]),
    "efwi_get_tile": (C.c_int, [_vp, C.c_int, _ip, _ip]),
    "efwi_autotune": (C.c_int, [_vp, C.c_int, _dp, _vp]),
    "efwi_profile": (C.c_int, [_vp, C.c_int]),
    "efwi_profile_read": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_longlong), C.c_int]),
    "efwi_cpml_profile": (C.c_int, [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, _fp]),
    "efwi_taper_w2": (C.c_int, [C.c_int, C.c_float, C.c_float, _fp]),
    "efwi_window_factors": (C.c_int, [C.c_int, C.c_float, C.c_int, _fp, _fp, _fp, C.c_float, _fp]),
    "efwi_courant_number": (C.c_int, [C.c_double, C.c_float, C.c_float, C.c_float, C.c_float, _dp]),
    "efwi_cufd": (C.c_int, [_fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, C.c_int, C.c_int, C.c_int, _ip,
                            C.c_char_p]),
    "efwi_set_illumination": (C.c_int, [_vp, C.c_int, _vp]),
    "efwi_get_illumination": (C.c_int, [_vp, _vp, C.c_int, _vp]),
    # learned preconditioner (include/efwi_nn.h)
    "efwi_nn_conv3": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, _vp, C.c_int, C.c_int, _vp, C.c_int,
                                C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp]),
    "efwi_nn_conv3_cat": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, C.c_int, C.c_int, _vp, _vp, _vp, C.c_int,
                                    C.c_int, _vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.c_int, _vp]),
    "efwi_nn_wgrad_cat": (C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, C.c_int, _vp, C.c_int, C.c_int,
                                    _vp, C.c_long, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp]),
    "efwi_nn_loss_head_bwd": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_long, C.c_double, _vp, C.c_int,
                                        _vp, _vp, _vp, _vp]),
    "efwi_nn_flip_weights": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp]),
    "efwi_nn_flip_weights_n": (C.c_int, [C.c_int, _vp, _vp, _vp, _vp, _vp]),
    "efwi_nn_wgrad_scratch": (C.c_long, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "efwi_nn_wgrad": (C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, _vp, C.c_long, _vp, _vp,
                                C.c_int, C.c_int, C.c_int, C.c_int, _vp]),
    "efwi_nn_wgrad_noalloc": (C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, _vp, _vp, C.c_int, C.c_int,
                                        C.c_int, C.c_int, _vp]),
    "efwi_nn_chan_sum": (C.c_int, [_vp, C.c_int, C.c_int, C.c_long, _vp, _vp]),
    "efwi_nn_pool2": (C.c_int, [_vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp]),
    "efwi_nn_pool2_bwd": (C.c_int, [_vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    _vp, C.c_int, _vp]),
    "efwi_nn_up2": (C.c_int, [_vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp]),
    "efwi_nn_up2_bwd": (C.c_int, [_vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp,
                                  C.c_int, _vp]),
    "efwi_nn_add": (C.c_int, [_vp, C.c_int, _vp, C.c_int, C.c_int, C.c_long, _vp, C.c_int, _vp]),
    "efwi_nn_head": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, _vp, C.c_long, _vp]),
    "efwi_nn_head_bwd": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, _vp, C.c_int, _vp, _vp, C.c_long, _vp]),
    "efwi_nn_gate": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_long, _vp]),
    "efwi_nn_gate_bwd": (C.c_int, [_vp, _vp, _vp, _vp, C.c_long, _vp]),
    "efwi_nn_mse": (C.c_int, [_vp, _vp, _vp, C.c_long, C.c_double, _vp, _vp, _vp]),
    "efwi_nn_adam": (C.c_int, [_vp, _vp, _vp, _vp, C.c_long, C.c_float, C.c_float, C.c_float, C.c_float,
                               C.c_int, _vp, _vp]),
    "efwi_nn_to_bf16": (C.c_int, [_vp, _vp, C.c_long, _vp]),
    "efwi_nn_absmax": (C.c_int, [_vp, C.c_long, _vp, _vp]),
    "efwi_nn_pack": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, _vp, _vp]),
}

_lib = None


def lib():
    """Load libefwi.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EfwiError(f"{LIB_PATH} is missing: build it with `python efwi-3d_amd/build.py` "
                            "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != OK:
        msg = lib().efwi_last_error().decode(errors="replace")
        raise EfwiError(f"{what or 'libefwi'} failed ({rc}): {msg}")


def fptr(t) -> C.POINTER(C.c_float):
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_float))


def vptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def stream_handle(device=None):
    """The current torch stream of `device` as a hipStream_t (void*)."""
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def dot64(a, b) -> float:
    """FP64 dot product of two float32 CUDA tensors (libefwi.so k_dot64)."""
    assert a.is_cuda and b.is_cuda and a.dtype == torch.float32 and b.dtype == torch.float32
    a = a.contiguous()
    b = b.contiguous()
    assert a.numel() == b.numel()
    out = C.c_double(0.0)
    with torch.cuda.device(a.device):
        check(lib().efwi_dot64(vptr(a), vptr(b), a.numel(), C.byref(out), stream_handle(a.device)), "efwi_dot64")
    return out.value
