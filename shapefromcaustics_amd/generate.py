"""Counter-based photon source for the scale-out workload (include/pds_gen.h).

Photon i of view v is a pure function of (seed, v, i), so the backward of a
splat regenerates the photons of its forward instead of keeping them: the
recomputation checkpoint of ``ScaleOut`` (4 views x 2^24 caustic-line photons
into 12 x 2048 x 2048, cap 16; view v writes channels 3v .. 3v + 2).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _native

_vp, _i32, _i64, _u64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64
SIGNATURES = {
    "pds_gen_last_error": ([], ctypes.c_char_p),
    "pds_gen_caustic_f32": ([_u64, _i32, _i64, _i32, _i32, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _vp],
                            ctypes.c_int),
}
_bound = False


def lib():
    global _bound
    L = _native.lib()
    if not _bound:
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _bound = True
    return L


def caustic_lines(seed: int, view: int, W: int, H: int, n_lines: int = 64) -> np.ndarray:
    """[n_lines, 4] float32: centre x, y (px) ~ U(0.1, 0.9) of the image, axis
    angle ~ U(0, pi), long sigma ~ U(0.02, 0.08) W (the caustic-line mixture)."""
    rng = np.random.Generator(np.random.PCG64(seed * 100 + view))
    mu = rng.uniform(0.1, 0.9, (n_lines, 2)) * np.array([W, H])
    phi = rng.uniform(0.0, math.pi, n_lines)
    sig = rng.uniform(0.02, 0.08, n_lines) * W
    return np.ascontiguousarray(np.c_[mu, phi, sig].astype(np.float32))


def generate_caustic(seed: int, view: int, n: int, W: int, H: int, lines: torch.Tensor, out, offset: int = 0):
    """Write view ``view``'s n photons into ``out = (E, xp, M, cp, radius)`` at [offset, offset + n)."""
    E, xp, M, cp, radius = out
    stride = E.numel()
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    rc = lib().pds_gen_caustic_f32(seed, view, n, W, H, p(lines), lines.shape[0], p(E), p(xp), p(M), p(cp),
                                   p(radius), stride, offset,
                                   ctypes.c_void_p(torch.cuda.current_stream(E.device).cuda_stream))
    if rc != 0:
        raise RuntimeError(f"pds_gen_caustic_f32 failed ({rc}): {lib().pds_gen_last_error().decode()}")


class ScaleOut:
    """Views x photons-per-view caustic-line photons into [3 views, side, side]."""

    def __init__(self, views=4, per_view=1 << 24, side=2048, cap=16, seed=6, device="cuda"):
        self.views, self.per_view, self.side, self.cap, self.seed = views, per_view, side, cap, seed
        self.size = (3 * views, side, side)
        self.device = torch.device(device)
        self.lines = [torch.from_numpy(caustic_lines(seed, v, side, side)).to(self.device) for v in range(views)]

    @property
    def n(self) -> int:
        return self.views * self.per_view

    def photons(self):
        """Generate every view's photons (E, xp, M, cp, radius) on the device."""
        n, dev = self.n, self.device
        f32 = dict(dtype=torch.float32, device=dev)
        out = (torch.empty(n, **f32), torch.empty(2, n, **f32), torch.empty(2, 3, n, **f32),
               torch.empty(1, n, dtype=torch.int64, device=dev), torch.empty(n, **f32))
        for v in range(self.views):
            generate_caustic(self.seed, v, self.per_view, self.side, self.side, self.lines[v], out,
                             offset=v * self.per_view)
        return out
