"""ctypes binding of ``libvlasov_amd.so`` (the C ABI in include/vlasov_amd.h).

The library must already be built (``python -m vlasov_amd.build``); this
module never falls back to anything else: a missing library or a non-GPU
tensor raises.  torch is imported first so that the library binds to the
HIP runtime torch already loaded (same SONAME), giving one runtime per
process.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch

from .build import LIB_PATH

c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_int = ctypes.c_int
c_sz = ctypes.c_size_t
c_vp = ctypes.c_void_p
I64x4 = c_i64 * 4

_SIGS = {
    "vp_error_string": (ctypes.c_char_p, [c_int]),
    "vp_abi_version": (c_int, []),
    "vp_advect_f64": (c_int, [c_vp, c_vp, I64x4, c_int, c_vp, I64x4, c_dbl, c_dbl, c_int, c_vp]),
    "vp_advect_f32": (c_int, [c_vp, c_vp, I64x4, c_int, c_vp, I64x4, c_dbl, c_dbl, c_int, c_vp]),
    "vp_advect_reflected_f64": (c_int, [c_vp, c_vp, I64x4, c_int, c_vp, I64x4, c_dbl, c_dbl, c_int, c_int, c_vp]),
    "vp_advect_kokkos_f64": (c_int, [c_vp, c_vp, I64x4, c_int, c_vp, c_vp, I64x4, c_dbl, c_dbl, c_dbl, c_dbl, c_int,
                                     c_vp]),
    "vp_advect_static_f64": (c_int, [c_vp, c_vp, I64x4, c_int, c_vp, c_vp, c_i64, I64x4, c_vp]),
    "vp_row_sum_f64": (c_int, [c_vp, c_i64, c_i64, c_dbl, c_dbl, c_dbl, c_vp, c_vp]),
    "vp_row_sum_f32": (c_int, [c_vp, c_i64, c_i64, c_dbl, c_dbl, c_dbl, c_vp, c_vp]),
    "vp_poisson_1d_workspace": (c_sz, [c_i64]),
    "vp_poisson_1d_f64": (c_int, [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "vp_poisson_2d_workspace": (c_sz, [c_i64, c_i64]),
    "vp_poisson_2d_f64": (c_int, [c_vp, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                  c_vp, c_sz, c_vp]),
    "vp_step_1d1v_workspace": (c_sz, [c_i64, c_i64]),
    "vp_step_1d1v_f64": (c_int, [c_vp, c_vp, c_i64, c_i64, c_dbl, c_vp, c_vp, c_vp, c_dbl, c_dbl,
                                 c_int, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "vp_step_1d1v_f32": (c_int, [c_vp, c_vp, c_i64, c_i64, c_dbl, c_vp, c_vp, c_vp, c_dbl, c_dbl,
                                 c_int, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "vp_step_1d1v_f64_noalloc": (c_int, [c_vp, c_vp, c_i64, c_i64, c_dbl, c_vp, c_vp, c_vp, c_dbl, c_dbl, c_int,
                                         c_vp, c_vp, c_vp, c_vp]),
    "vp_step_1d1v_shift_f64": (c_int, [c_vp, c_vp, c_i64, c_i64, c_dbl, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp,
                                       c_dbl, c_dbl, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "vp_step_2d2v_workspace": (c_sz, [I64x4]),
    "vp_step_2d2v_f64": (c_int, [c_vp, c_vp, I64x4, c_dbl, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl,
                                 c_dbl, c_dbl, c_dbl, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz,
                                 c_vp]),
    "vp_run_2d2v_workspace": (c_sz, [I64x4]),
    "vp_run_2d2v_f64": (c_int, [c_vp, I64x4, c_int, c_dbl, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_dbl,
                                c_dbl, c_dbl, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "vp_fused_workspace": (c_sz, [I64x4, c_int]),
    "vp_fused_tables_f64": (c_int, [I64x4, c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_vp, c_sz, c_vp]),
    "vp_fused_xdrift_f64": (c_int, [c_vp, c_vp, I64x4, c_int, c_vp, c_sz, c_vp]),
    "vp_fused_ydrift_f64": (c_int, [c_vp, c_vp, I64x4, c_int, c_vp, c_sz, c_vp]),
    "vp_fused_poisson_f64": (c_int, [I64x4, c_dbl, c_dbl, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                     c_sz, c_vp]),
    "vp_fused_kick_f64": (c_int, [c_vp, c_vp, I64x4, c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_vp]),
    "vp_run_2d2v_f32_workspace": (c_sz, [I64x4]),
    "vp_run_2d2v_f32": (c_int, [c_vp, I64x4, c_int, c_dbl, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_dbl,
                                c_dbl, c_dbl, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "vp_fused_tables_f32": (c_int, [I64x4, c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_vp, c_sz, c_vp]),
    "vp_fused_xdrift_f32": (c_int, [c_vp, c_vp, I64x4, c_int, c_vp, c_sz, c_vp]),
    "vp_fused_ydrift_f32": (c_int, [c_vp, c_vp, I64x4, c_int, c_vp, c_sz, c_vp]),
    "vp_fused_poisson_f32": (c_int, [I64x4, c_dbl, c_dbl, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                     c_sz, c_vp]),
    "vp_fused_kick_f32": (c_int, [c_vp, c_vp, I64x4, c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_vp]),
    "vp_dev_split_workspace": (c_sz, [I64x4]),
    "vp_dev_split_f32": (c_int, [c_vp, I64x4, c_dbl, c_dbl, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "vp_dev_join_f32": (c_int, [c_vp, c_vp, I64x4, c_vp, c_vp, c_vp]),
    "vp_run_2d2v_dev_f32": (c_int, [c_vp, c_vp, c_vp, I64x4, c_int, c_dbl, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl,
                                    c_dbl, c_dbl, c_dbl, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "vp_fused_kick_dev_f32": (c_int, [c_vp, c_vp, I64x4, c_vp, c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_vp]),
    "vp_fused_poisson_dev_f32": (c_int, [I64x4, c_dbl, c_dbl, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                         c_vp, c_sz, c_vp]),
    "vp_field_energy_f64": (c_int, [c_vp, c_vp, c_i64, c_dbl, c_dbl, c_vp, c_vp]),
    "vp_moments_workspace": (c_sz, []),
    "vp_moments_f64": (c_int, [c_vp, c_i64, c_i64, c_vp, c_vp, c_i64, c_dbl, c_dbl, c_vp, c_vp, c_sz, c_vp]),
    "vp_moments_f32": (c_int, [c_vp, c_i64, c_i64, c_vp, c_vp, c_i64, c_dbl, c_dbl, c_vp, c_vp, c_sz, c_vp]),
    "vp_ens_workspace": (c_sz, [c_int, c_int, c_int]),
    "vp_ens_run_f64": (c_int, [c_vp, c_vp, c_int, c_int, c_int, c_vp, c_int, c_dbl, c_int, c_vp, c_vp, c_sz, c_vp]),
    "vp_nn_gemm": (c_int, [c_int, c_int, c_int, c_int, ctypes.c_float, c_vp, c_i64, c_i64, c_i64, c_vp, c_i64, c_i64,
                           c_i64, ctypes.c_float, c_vp, c_i64, c_i64, c_i64, c_vp]),
    "vp_nn_f0_moments": (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_vp, c_vp]),
    "vp_nn_gelu": (c_int, [c_vp, c_vp, c_i64, c_vp]),
    "vp_nn_gelu_bwd": (c_int, [c_vp, c_vp, c_vp, c_i64, c_vp]),
    "vp_nn_bias_add": (c_int, [c_vp, c_vp, c_i64, c_int, c_vp]),
    "vp_nn_bias_grad": (c_int, [c_vp, c_vp, c_i64, c_int, c_vp]),
    "vp_nn_mse": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]),
    "vp_nn_wmse": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]),
    "vp_nn_adam": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                           ctypes.c_float, c_int, c_vp]),
    "vp_tune_set": (c_int, [c_int, c_int]),
    "vp_tune_get": (c_int, [c_int]),
    "vp_stream_f64": (c_int, [c_int, c_vp, c_vp, c_i64, c_int, c_int, c_int, c_vp]),
}

KNOBS = {"x_split": 0, "y_split": 1, "kick_prefetch": 2, "kick_kernel": 3, "drift_dma": 4, "nontemporal": 5,
         "ens_xtile": 6, "ens_ktile": 7, "ens_wgs": 8, "seg_rows": 9, "seg_contig": 10, "f32_pairs": 11,
         "poisson_dft": 12, "y_units": 13}

_lib = None
_lock = threading.Lock()


def lib():
    """The loaded library (loads it on first use; raises if it is not built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(
                        f"{LIB_PATH} is missing: build it with `python -m vlasov_amd.build` "
                        "(or __graft_entry__.build()); there is no non-HIP fallback")
                handle = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
                for name, (res, args) in _SIGS.items():
                    fn = getattr(handle, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = handle
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().vp_error_string(rc).decode()
        raise RuntimeError(f"{what} failed: {msg} (code {rc})")


def stream_handle(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def require_gpu(*tensors: torch.Tensor) -> torch.device:
    dev = None
    for t in tensors:
        if not isinstance(t, torch.Tensor):
            continue
        if t.device.type != "cuda":
            raise RuntimeError(
                "vlasov_amd runs on the GPU only: got a tensor on "
                f"{t.device}; move the grid and state to 'cuda'")
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise RuntimeError(f"tensors on different devices: {dev} and {t.device}")
    if dev is None:
        raise RuntimeError("no GPU tensor given")
    return dev


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def i64x4(vals) -> ctypes.Array:
    return I64x4(*[int(v) for v in vals])


class Workspace:
    """Per-device scratch buffers owned by torch's caching allocator.

    The library never allocates; callers hand it a slice of this buffer.  A
    buffer is reused by every call on the same device and stream, which is
    safe because those calls are ordered on that stream.
    """

    _bufs: dict = {}

    @classmethod
    def get(cls, device: torch.device, nbytes: int, tag: str = "ws") -> torch.Tensor:
        key = (device, tag, torch.cuda.current_stream(device).cuda_stream)
        buf = cls._bufs.get(key)
        if buf is None or buf.numel() < nbytes:
            buf = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)
            cls._bufs[key] = buf
        return buf

    @classmethod
    def clear(cls) -> None:
        cls._bufs.clear()
