"""ctypes binding of ``libpds_hip.so`` (C ABI declared in ``include/pds.h``).

torch is imported first so that the process's HIP runtime is the one torch
loaded; the library's ``libamdhip64.so.7`` dependency then resolves to it.
The library must have been built (``__graft_entry__.build()``): there is no
fallback, and every entry point raises if it is missing.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime before the library)

from ._build import LIB_PATH

PDS_OK = 0
PDS_ERR_ARGUMENT = -1
PDS_ERR_WORKSPACE = -2
PDS_ERR_HIP = -3
PDS_ERR_CAPACITY = -4

MODES = {"auto": 0, "plain": 1, "binned": 2, "deterministic": 3}
STATIC_SCRATCH_BYTES = 256 << 20  # PDS_STATIC_SCRATCH_BYTES of include/pds.h


class PdsTuning(ctypes.Structure):
    """Mirror of ``pds_tuning`` (include/pds.h)."""

    _fields_ = [("tile", ctypes.c_int32), ("seg_photons", ctypes.c_int32), ("block_photons", ctypes.c_int32),
                ("capacity_pct", ctypes.c_int32), ("walk", ctypes.c_int32), ("reserved", ctypes.c_int32 * 3)]


class PdsBinsStats(ctypes.Structure):
    """Mirror of ``pds_bins_stats`` (include/pds.h)."""

    _fields_ = [("records", ctypes.c_int64), ("capacity", ctypes.c_int64), ("segments", ctypes.c_int32),
                ("overflow", ctypes.c_int32), ("tile", ctypes.c_int32), ("seg_photons", ctypes.c_int32),
                ("bins", ctypes.c_int32), ("block_photons", ctypes.c_int32), ("reserved", ctypes.c_int32 * 2)]


_lib = None

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_sz = ctypes.c_size_t
_tp = ctypes.POINTER(PdsTuning)

SIGNATURES = {
    "pds_forward_workspace_bytes": ([_i64, _i32, _i32, _i32, _i32, _i32, _tp], _sz),
    "pds_backward_workspace_bytes": ([_i64, _i32, _i32, _i32, _i32, _i32, _tp], _sz),
    "pds_forward_f32": ([_vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32, _i32, _tp, _vp, _vp, _sz, _vp], ctypes.c_int),
    "pds_backward_f32": ([_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32, _i32, _tp, _vp, _vp, _vp, _vp,
                          _sz, _vp], ctypes.c_int),
    "pds_forward_noalloc_f32": ([_vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32, _i32, _tp, _vp, _vp],
                                ctypes.c_int),
    "pds_bins_bytes": ([_i64, _i32, _i32, _i32, _i32, _i32, _tp], _sz),
    "pds_forward_bins_f32": ([_vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32, _i32, _tp, _vp, _vp, _sz, _vp],
                             ctypes.c_int),
    "pds_backward_bins_f32": ([_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32, _i32, _tp, _vp, _vp, _vp,
                               _vp, _sz, _vp], ctypes.c_int),
    "pds_forward_bins_live_f32": ([_vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32, _i32, _tp, _vp, _vp, _vp,
                                   _sz, _vp], ctypes.c_int),
    "pds_backward_bins_live_f32": ([_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32, _i32, _tp, _vp, _vp,
                                    _vp, _vp, _vp, _sz, _vp], ctypes.c_int),
    "pds_bins_stats_get": ([_vp, _sz, _i64, _i32, _i32, _i32, _i32, _i32, _tp, ctypes.POINTER(PdsBinsStats), _vp],
                           ctypes.c_int),
    "pds_last_error": ([], ctypes.c_char_p),
    "pds_version": ([], ctypes.c_char_p),
}


def lib() -> ctypes.CDLL:
    """Load ``libpds_hip.so`` once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() (no CPU fallback exists)")
        cdll = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(cdll, name)
            fn.argtypes = args
            fn.restype = res
        _lib = cdll
    return _lib


def last_error() -> str:
    return lib().pds_last_error().decode()


def check(rc: int, what: str) -> None:
    if rc != PDS_OK:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")


def tuning_struct(tuning) -> "PdsTuning | None":
    if tuning is None:
        return None
    if isinstance(tuning, PdsTuning):
        return tuning
    t = PdsTuning()
    fields = {name for name, _ in PdsTuning._fields_ if name != "reserved"}
    for k, v in dict(tuning).items():
        if k not in fields:
            raise ValueError(f"unknown pds_tuning field {k!r} (fields: {', '.join(sorted(fields))})")
        setattr(t, k, int(v))
    return t
