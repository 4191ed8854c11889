"""Drop-in replacement of the reference extension module ``PhotonDifferentialSplatting``.

Same entry points, argument order and error messages as the reference's pybind
module (``PyOptix/PhotonDifferentialSplattig.cpp:12-37``, checks from
``PyOptix/utils.hpp:37-41``); the work is done by the MI355X kernels of
``libpds_hip.so`` through its C ABI.

    pds_forward(Ep, xp, Mp, cp, radius, output_size, max_pixel_radius) -> [grid]
    pds_backward(grad_pds, Ep, xp, Mp, cp, radius, max_pixel_radius) -> [gE, gx, gM]

Deliberate, documented supersets of the reference:

* launches go to torch's current stream of the inputs' device (the reference
  uses the legacy default stream of the current device);
* a non-contiguous ``grad_pds`` is made contiguous (the reference raises);
* photons with a negative channel are skipped (the reference indexes out of
  bounds) and photons that are skipped get zero gradients (the reference
  returns uninitialised memory for them);
* the gradient at r = 0 is its finite limit (the reference divides 0/0);
* keyword-only extras after the reference arguments: ``mode`` (``"auto"``,
  ``"plain"``, ``"binned"``, ``"deterministic"``) and ``tuning``.

Only float32 photons/images and int64 channels run on the GPU; other dtypes
raise ``TypeError`` (the reference also dispatches fp16/fp64).
"""
from __future__ import annotations

import ctypes
import typing as tp

import torch

from . import _native

__all__ = ["pds_forward", "pds_backward"]


def _check_input(t: torch.Tensor, name: str) -> None:
    # CHECK_CUDA / CHECK_CONTIGUOUS, PyOptix/utils.hpp:37-41
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")


def _check_photons(Ep, xp, Mp, cp, radius, require_cuda: bool = True) -> None:
    """Validation of PyOptix/PhotonDifferentialSplattig.cpp:14-22, in order."""
    for t, name in ((Ep, "Ep"), (xp, "xp"), (Mp, "Mp"), (cp, "cp"), (radius, "radius")):
        if require_cuda:
            _check_input(t, name)
        elif not t.is_contiguous():
            raise RuntimeError(f"{name} must be contiguous")
    if not (Ep.size(0) == xp.size(1) and Ep.size(0) == Mp.size(2) and Ep.size(0) == cp.size(1)
            and Ep.size(0) == radius.size(0)):
        raise RuntimeError("Dimensions of tensors for pds_forward don't match")
    if cp.size(0) != 1:
        raise RuntimeError("Channel index must be one dimensional!")
    if xp.size(0) != 2:
        raise RuntimeError("Position index must be two dimensional!")
    if not (Mp.size(0) == 2 and Mp.size(1) == 3):
        raise RuntimeError("Change of basis matrix must be 2x3!")
    # beyond the reference: the GPU path is float32 / int64, same device
    for t, name in ((Ep, "Ep"), (xp, "xp"), (Mp, "Mp"), (radius, "radius")):
        if t.dtype != torch.float32:
            raise TypeError(f"{name} must be float32 on the GPU path (got {t.dtype})")
    if cp.dtype != torch.int64:
        raise TypeError(f"cp must be int64 (got {cp.dtype})")
    if Ep.dim() != 1 or xp.dim() != 2 or Mp.dim() != 3 or cp.dim() != 2 or radius.dim() != 1:
        raise RuntimeError("Dimensions of tensors for pds_forward don't match")
    dev = Ep.device
    for t in (xp, Mp, cp, radius):
        if t.device != dev:
            raise RuntimeError(f"all inputs must be on one device (found {dev} and {t.device})")


def _mode(mode) -> int:
    if isinstance(mode, int):
        return mode
    try:
        return _native.MODES[mode]
    except KeyError:
        raise ValueError(f"unknown mode {mode!r}; expected one of {sorted(_native.MODES)}") from None


def _ptr(t: tp.Optional[torch.Tensor]):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else ctypes.c_void_p(0)


def _workspace(query, N, C, H, W, cap, mode, tun, device):
    nbytes = query(N, C, H, W, cap, mode, ctypes.byref(tun) if tun is not None else None)
    if nbytes == 0:
        return None, 0
    return torch.empty(nbytes, dtype=torch.uint8, device=device), nbytes


def _tun_ref(tun):
    return ctypes.byref(tun) if tun is not None else None


def _resolve_tuning(tuning, N, size, cap, mode):
    """``tuning``: None (built-in defaults), a dict / PdsTuning, or ``"cached"``
    (the autotuner's cached choice for this problem class, defaults if none)."""
    if isinstance(tuning, str):
        if tuning != "cached":
            raise ValueError(f"unknown tuning {tuning!r}; expected None, a dict or 'cached'")
        from . import autotune
        tuning = autotune.lookup(N, size, cap, mode)
    return _native.tuning_struct(tuning)


def _check_live(live, device):
    if not (live.is_cuda and live.dtype == torch.int64 and live.numel() == 1 and live.device == device):
        raise RuntimeError("live must be a one-element int64 tensor on the photons' device")


def _forward(Ep, xp, Mp, cp, radius, output_size, max_pixel_radius, mode, tuning, keep_bins, live=None):
    """``live``: None, or a one-element int64 device tensor: only photons
    i < live take part (pds_forward_bins_live_f32; needs keep_bins and a
    binnable problem)."""
    _check_photons(Ep, xp, Mp, cp, radius)
    size = [int(s) for s in output_size]
    if len(size) != 3:
        raise RuntimeError(f"output_size must have 3 entries (C, H, W), got {len(size)}")
    C, H, W = size
    N = Ep.size(0)
    m = _mode(mode)
    tun = _resolve_tuning(tuning, N, size, max_pixel_radius, mode)
    lib = _native.lib()
    cap = int(max_pixel_radius)
    with torch.cuda.device(Ep.device):
        out = torch.empty(size, dtype=torch.float32, device=Ep.device)
        bins = None
        stream = ctypes.c_void_p(torch.cuda.current_stream(Ep.device).cuda_stream)
        if live is not None:
            _check_live(live, Ep.device)
            if not keep_bins or C * H * W == 0 or lib.pds_bins_bytes(N, C, H, W, cap, m, _tun_ref(tun)) == 0:
                raise RuntimeError("a live photon count needs kept bins of a binnable problem")
            bins, nbytes = _workspace(lib.pds_bins_bytes, N, C, H, W, cap, m, tun, Ep.device)
            rc = lib.pds_forward_bins_live_f32(_ptr(Ep), _ptr(xp), _ptr(Mp), _ptr(cp), _ptr(radius), N, C, H, W, cap,
                                               m, _tun_ref(tun), _ptr(live), _ptr(out), _ptr(bins), nbytes, stream)
            _native.check(rc, "pds_forward")
            return out, bins
        if keep_bins and C * H * W > 0 and lib.pds_bins_bytes(N, C, H, W, cap, m, _tun_ref(tun)) > 0:
            bins, nbytes = _workspace(lib.pds_bins_bytes, N, C, H, W, cap, m, tun, Ep.device)
            fn = lib.pds_forward_bins_f32
        else:
            ws, nbytes = _workspace(lib.pds_forward_workspace_bytes, N, C, H, W, cap, m, tun, Ep.device)
            fn = lib.pds_forward_f32
        rc = fn(_ptr(Ep), _ptr(xp), _ptr(Mp), _ptr(cp), _ptr(radius), N, C, H, W, cap, m, _tun_ref(tun), _ptr(out),
                _ptr(bins if bins is not None else ws), nbytes, stream)
    _native.check(rc, "pds_forward")
    return out, bins


def _backward(grad_pds, Ep, xp, Mp, cp, radius, max_pixel_radius, mode, tuning, bins=None, live=None):
    """``live``: the live count of the forward that kept ``bins``
    (pds_backward_bins_live_f32): gradients of photons i >= live are not
    written."""
    if not grad_pds.is_cuda:
        raise RuntimeError("grad_pds must be a CUDA tensor")
    if grad_pds.dim() != 3:
        raise RuntimeError(f"grad_pds must be [C, H, W], got {tuple(grad_pds.shape)}")
    if grad_pds.dtype != torch.float32:
        raise TypeError(f"grad_pds must be float32 on the GPU path (got {grad_pds.dtype})")
    grad_pds = grad_pds.contiguous()
    _check_photons(Ep, xp, Mp, cp, radius)
    if grad_pds.device != Ep.device:
        raise RuntimeError(f"all inputs must be on one device (found {Ep.device} and {grad_pds.device})")
    C, H, W = grad_pds.shape
    N = Ep.size(0)
    m = _mode(mode)
    tun = _resolve_tuning(tuning, N, (C, H, W), max_pixel_radius, mode)
    lib = _native.lib()
    cap = int(max_pixel_radius)
    with torch.cuda.device(Ep.device):
        gE = torch.empty_like(Ep)
        gx = torch.empty_like(xp)
        gM = torch.empty_like(Mp)
        stream = torch.cuda.current_stream(Ep.device).cuda_stream
        if live is not None:
            _check_live(live, Ep.device)
            if bins is None:
                raise RuntimeError("a live photon count needs the bins of its forward")
            rc = lib.pds_backward_bins_live_f32(_ptr(grad_pds), _ptr(Ep), _ptr(xp), _ptr(Mp), _ptr(cp), _ptr(radius),
                                                N, C, H, W, cap, m, _tun_ref(tun), _ptr(live), _ptr(gE), _ptr(gx),
                                                _ptr(gM), _ptr(bins), bins.numel(), ctypes.c_void_p(stream))
            _native.check(rc, "pds_backward")
            return [gE, gx, gM]
        if bins is not None:
            fn, ws, nbytes = lib.pds_backward_bins_f32, bins, bins.numel()
        else:
            fn = lib.pds_backward_f32
            ws, nbytes = _workspace(lib.pds_backward_workspace_bytes, N, C, H, W, cap, m, tun, Ep.device)
        rc = fn(_ptr(grad_pds), _ptr(Ep), _ptr(xp), _ptr(Mp), _ptr(cp), _ptr(radius), N, C, H, W, cap, m,
                _tun_ref(tun), _ptr(gE), _ptr(gx), _ptr(gM), _ptr(ws), nbytes, ctypes.c_void_p(stream))
    _native.check(rc, "pds_backward")
    return [gE, gx, gM]


def _bins_stats(bins, N, output_size, max_pixel_radius, mode="auto", tuning=None) -> dict:
    """Statistics of a binning kept by ``_forward(keep_bins=True)`` (synchronises)."""
    C, H, W = [int(s) for s in output_size]
    tun = _resolve_tuning(tuning, N, (C, H, W), max_pixel_radius, mode)
    st = _native.PdsBinsStats()
    rc = _native.lib().pds_bins_stats_get(_ptr(bins), bins.numel(), N, C, H, W, int(max_pixel_radius), _mode(mode),
                                          _tun_ref(tun), ctypes.byref(st),
                                          ctypes.c_void_p(torch.cuda.current_stream(bins.device).cuda_stream))
    _native.check(rc, "pds_bins_stats_get")
    return {k: getattr(st, k) for k, _ in st._fields_ if k != "reserved"}


def pds_forward(Ep: torch.Tensor, xp: torch.Tensor, Mp: torch.Tensor, cp: torch.Tensor, radius: torch.Tensor,
                output_size: tp.Sequence[int], max_pixel_radius: int, *, mode="auto", tuning=None):
    """Photon-differential splat; returns ``[grid]`` with ``grid`` of shape ``output_size``."""
    return [_forward(Ep, xp, Mp, cp, radius, output_size, max_pixel_radius, mode, tuning, keep_bins=False)[0]]


def pds_backward(grad_pds: torch.Tensor, Ep: torch.Tensor, xp: torch.Tensor, Mp: torch.Tensor, cp: torch.Tensor,
                 radius: torch.Tensor, max_pixel_radius: int, *, mode="auto", tuning=None):
    """Adjoint of :func:`pds_forward`; returns ``[grad_Ep, grad_xp, grad_Mp]``."""
    return _backward(grad_pds, Ep, xp, Mp, cp, radius, max_pixel_radius, mode, tuning)
