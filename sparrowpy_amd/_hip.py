"""ctypes binding of ``libsparrow_hip.so`` (declarations in
``include/sparrow_hip.h``).

PyTorch provides device memory and the current HIP stream; every numerical
step runs in the HIP library.  There is no CPU fallback: if the library is
not built, or no GPU is visible, calls raise :class:`HipUnavailable`.

torch is imported before the library is loaded so that the library binds to
the HIP runtime torch already has in the process.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch  # noqa: F401  (loads the process's HIP runtime first)

from .build import LIB

__all__ = ["HipUnavailable", "SparrowHipError", "lib", "call", "ptr",
           "stream_ptr", "SYMBOLS"]

SPR_OK = 0
_ERRORS = {-1: "bad argument", -2: "HIP runtime error", -3: "out of range",
           -4: "unsupported", -5: "workspace too small"}


class HipUnavailable(RuntimeError):
    """The HIP library or a GPU is not available."""


class SparrowHipError(RuntimeError):
    """A sparrow_hip entry point returned an error status."""

    def __init__(self, name, code, msg):
        super().__init__(f"{name}: {_ERRORS.get(code, code)}: {msg}")
        self.code = code


_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_int = ctypes.c_int
_vp = ctypes.c_void_p

# name -> (restype, argtypes); every pointer is passed as c_void_p
SYMBOLS = {
    "spr_last_error": (ctypes.c_char_p, []),
    "spr_version": (ctypes.c_char_p, []),
    "spr_visibility_workspace": (_i64, [_i64]),
    "spr_p2p_visibility": (_int, [_i64, _vp, _i64, _vp, _vp, _vp, _vp, _i64,
                                  _vp]),
    "spr_point_visibility": (_int, [_i64, _vp, _i64, _vp, _i64, _vp, _vp, _vp,
                                    _vp, _i64, _vp]),
    "spr_visible_pairs": (_int, [_i64, _vp, _vp, _vp, _vp]),
    "spr_form_factors": (_int, [_i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp]),
    "spr_direction_map": (_int, [_i64, _vp, _vp, _i64, _vp, _vp, _vp, _vp]),
    "spr_point_direction_index": (_int, [_i64, _vp, _i64, _vp, _i64, _vp, _vp,
                                         _vp, _vp]),
    "spr_fft_dense": (_int, [_i64, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _i64,
                             _i64, _vp, _vp, _vp, _vp, _vp]),
    "spr_source_energy": (_int, [_vp, _i64, _vp, _vp, _vp, _vp, _i64, _vp,
                                 _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp,
                                 _vp]),
    "spr_padded_directions": (_i64, [_i64]),
    "spr_plan_rows": (_int, [_i64, _vp, _vp, _vp, _vp]),
    "spr_plan_from_dense": (_int, [_i64, _i64, _i64, _dbl, _dbl, _i64, _vp,
                                   _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "spr_plan_from_geometry": (_int, [_i64, _i64, _i64, _dbl, _dbl, _i64,
                                      _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                      _i64, _vp, _vp, _vp, _vp, _vp, _vp,
                                      _vp]),
    "spr_exchange_init": (_int, [_i64, _i64, _i64, _i64, _vp, _vp, _dbl,
                                 _dbl, _vp, _vp, _vp, _vp, _vp]),
    "spr_exchange_step": (_int, [_i64, _i64, _i64, _i64, _vp, _vp, _vp, _vp,
                                 _vp, _vp, _vp, _vp, _vp, _vp]),
    "spr_exchange_finalize": (_int, [_i64, _vp, _vp, _vp, _vp]),
    "spr_plan_grouped_sort": (_int, [_i64, _i64, _vp, _vp, _vp, _vp, _i64,
                                     _vp, _vp, _vp, _vp, _vp]),
    "spr_plan_grouped_fill": (_int, [_i64, _i64, _i64, _dbl, _dbl, _i64, _vp,
                                     _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                     _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                     _vp]),
    "spr_exchange_gather_grouped": (_int, [_i64, _i64, _i64, _i64, _vp, _vp,
                                           _vp, _vp, _vp, _vp, _vp, _vp]),
    "spr_exchange_spread_grouped": (_int, [_i64, _i64, _i64, _i64, _vp, _vp,
                                           _vp, _vp, _vp, _vp, _vp, _vp]),
    "spr_exchange_spread_grouped_pair": (_int, [_i64, _i64, _i64, _i64, _vp,
                                                _vp, _vp, _vp, _vp, _vp, _vp,
                                                _vp, _vp]),
    "spr_exchange_total_add": (_int, [_i64, _vp, _vp, _vp, _vp]),
    "spr_exchange_spread_ds": (_int, [_i64, _i64, _i64, _i64, _vp, _vp, _vp,
                                      _vp, _vp, _vp, _vp, _vp]),
    "spr_exchange_step_grouped": (_int, [_i64, _i64, _i64, _i64, _i64, _vp,
                                         _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                         _vp, _vp, _vp, _vp, _vp]),
    "spr_brdf_diffuse_specular": (_int, [_i64, _i64, _i64, _vp, _vp, _vp,
                                         _vp, _vp, _vp]),
    "spr_plan_ds_sort": (_int, [_i64, _i64, _i64, _vp, _vp, _vp, _vp, _i64,
                                _vp, _vp, _vp, _vp, _vp, _vp]),
    "spr_plan_ds_fill": (_int, [_i64, _i64, _i64, _dbl, _dbl, _i64, _vp, _vp,
                                _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64,
                                _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                _vp, _vp, _vp, _vp, _vp]),
    "spr_plan_tile_order": (_int, [_i64, _i64, _i64, _i64, _vp, _vp, _vp,
                                   _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "spr_plan_tile_fill": (_int, [_i64, _i64, _i64, _i64, _i64, _vp, _vp,
                                  _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                  _vp]),
    "spr_plan_tile_span": (_int, [_i64, _i64, _vp, _vp, _vp, _vp]),
    "spr_exchange_gather_tiled": (_int, [_i64, _i64, _i64, _i64, _i64, _vp,
                                         _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "spr_exchange_gather_tiled_shape": (_int, [_i64, _i64, _i64, _i64,
                                               _i64, _i64, _vp, _vp, _vp,
                                               _vp, _vp, _vp, _vp, _vp]),
    "spr_exchange_gather_tiled_ds": (_int, [_i64, _i64, _i64, _i64, _i64,
                                            _vp, _vp, _vp, _vp, _vp, _vp,
                                            _vp, _vp, _vp, _vp]),
    "spr_exchange_gather_ds": (_int, [_i64, _i64, _i64, _i64, _vp, _vp, _vp,
                                      _vp, _vp, _vp, _vp, _vp, _vp]),
    "spr_ds_weights": (_int, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "spr_ds_injection": (_int, [_i64, _i64, _i64, _vp, _vp, _vp, _i64, _vp,
                                _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "spr_plan_ds_transpose_rows": (_int, [_i64, _i64, _vp, _vp, _vp, _vp]),
    "spr_plan_ds_transpose_fill": (_int, [_i64, _i64, _i64, _i64, _vp, _vp,
                                          _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                          _vp, _vp, _vp, _vp]),
    "spr_ds_transposed_weights": (_int, [_i64, _i64, _vp, _vp, _vp, _vp, _vp,
                                         _vp]),
    "spr_collect_adjoint": (_int, [_i64, _i64, _i64, _i64, _i64, _vp, _vp,
                                   _vp, _vp, _vp, _vp, _dbl, _dbl, _vp, _vp]),
    "spr_direction_sum": (_int, [_i64, _i64, _i64, _i64, _vp, _vp, _vp]),
    "spr_exchange_adjoint_ds": (_int, [_i64, _i64, _i64, _i64, _i64, _vp, _vp,
                                       _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                       _vp, _vp, _vp]),
    "spr_ds_table_grad": (_int, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp,
                                 _vp]),
    "spr_ds_injection_grad": (_int, [_i64, _i64, _i64, _i64, _vp, _vp, _dbl,
                                     _dbl, _vp, _vp, _vp, _vp, _vp, _vp]),
    "spr_grouped_spread_adjoint": (_int, [_i64, _i64, _i64, _i64, _vp, _vp,
                                          _vp, _vp, _vp, _vp]),
    "spr_grouped_table_grad": (_int, [_i64, _i64, _i64, _i64, _vp, _vp, _vp,
                                      _vp, _vp]),
    "spr_grouped_gather_adjoint": (_int, [_i64, _i64, _i64, _i64, _i64, _vp,
                                          _vp, _vp, _vp, _vp, _vp, _vp,
                                          _vp]),
    "spr_grouped_key_grad": (_int, [_i64, _i64, _i64, _vp, _vp, _vp, _vp]),
    "spr_injection_table_grad": (_int, [_i64, _i64, _i64, _i64, _vp, _vp,
                                        _dbl, _dbl, _vp, _vp, _vp, _vp]),
    "spr_energy_exchange_workspace": (_i64, [_i64, _i64, _i64, _i64, _i64]),
    "spr_energy_exchange": (_int, [_i64, _i64, _i64, _i64, _vp, _vp, _vp,
                                   _vp, _vp, _dbl, _dbl, _i64, _i64, _vp,
                                   _vp, _vp, _i64, _vp]),
    "spr_plan_dense_hash": (_int, [_i64, _i64, _i64, _vp, _vp, _vp, _vp,
                                   _vp]),
    "spr_energy_exchange_groups": (_i64, []),
    "spr_energy_exchange_noalloc": (_int, [_i64, _i64, _i64, _i64, _vp, _vp,
                                           _vp, _vp, _vp, _dbl, _dbl, _i64,
                                           _i64, _vp, _vp, _vp]),
    "spr_receiver_energy": (_int, [_i64, _vp, _i64, _vp, _vp, _vp, _vp]),
    "spr_collect": (_int, [_i64, _i64, _i64, _i64, _i64, _vp, _vp, _vp, _vp,
                           _vp, _vp, _dbl, _dbl, _int, _vp, _vp]),
}

_lib = None
_lock = threading.Lock()


def lib(require_gpu=True):
    """Load the library (once); optionally require a visible GPU."""
    global _lib
    if require_gpu and not torch.cuda.is_available():
        raise HipUnavailable(
            "sparrowpy_amd needs an AMD GPU (torch.cuda.is_available() is "
            "False); there is no CPU fallback")
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB):
                raise HipUnavailable(
                    f"{LIB} is missing; build it with "
                    "`python -m sparrowpy_amd.build`")
            handle = ctypes.CDLL(LIB)
            for name, (res, args) in SYMBOLS.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


class DevicePointer(ctypes.c_void_p):
    """A device address that keeps its tensor alive.

    Arguments like ``ptr(as_dev(x))`` create a temporary tensor; holding it
    here keeps its storage allocated until the entry point has been called
    (kernels are then enqueued in stream order before any reuse)."""

    def __init__(self, tensor):
        super().__init__(tensor.data_ptr())
        self.tensor = tensor


_anchors = {}


def _anchor(device):
    """A small device buffer whose address stands for zero-length arrays."""
    a = _anchors.get(device)
    if a is None:
        a = _anchors[device] = torch.zeros(64, dtype=torch.uint8,
                                           device=device)
    return a


def ptr(t):
    """Device pointer of a tensor (None -> NULL).  A zero-length tensor,
    whose own address may be NULL, gets a valid address that nothing reads
    (its length, passed beside it, is 0), so entry points can keep
    rejecting NULL for arrays that are required."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    if t.numel() == 0:
        return DevicePointer(_anchor(t.device))
    return DevicePointer(t)


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def call(name, *args):
    """Call an entry point returning a status; raise on failure."""
    fn = getattr(lib(), name)
    rc = fn(*args)
    if rc != SPR_OK:
        msg = lib().spr_last_error().decode(errors="replace")
        raise SparrowHipError(name, rc, msg)
    return rc
