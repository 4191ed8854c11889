"""Seeded synthetic photon workloads for photon-differential splatting.

Photons are generated in pixel space and mapped to the reference's input
convention, the one ``model/caustics.py:233-239`` (``splat_photons``) produces:

* ``xp = 2 p / W - 1``  (pixel-corner coordinates ``p`` -> normalised [-1, 1]);
* the footprint is the ellipse with semi-axes ``D = R(theta) diag(s1, s2)``
  pixels, so the change-of-basis matrix is ``M[:, :2] = D^-1 diag(W/2, H/2)``
  and ``M[:, 2] = 0`` (``model/caustics.py:224-230`` gives ``M Du = (2, 0)``,
  ``M Dv = (0, 2)``: the kernel support is the ellipse of half-differentials);
* ``radius = 2 max(s1, s2) / W`` (``model/caustics.py:236``: half of the
  larger differential length, in normalised units);
* ``E`` is the photon energy, ``cp`` its channel.

Every array is produced by numpy's PCG64 from a fixed seed, so the GPU path,
the CPU baseline and the oracle all see identical inputs.

Workloads (photon count, image ``[C, H, W]``, ``max_pixel_radius``):

* ``plain``      2^18 photons -> [1, 256, 256], cap 8, uniform positions,
                 theta ~ U(0, pi), s1, s2 ~ logU(0.5, 4) px, E ~ U(0.5, 1.5)/N.
* ``caustic``    2^22 photons -> [3, 1024, 1024], cap 12; 80 % of the photons on
                 64 thin elongated Gaussians (caustic lines, aspect 20:1), 20 %
                 uniform; s_major ~ logU(0.3, 6) px, condition number
                 kappa ~ U(1, 10); cp = i mod 3.
* ``batched``    16 independent splats with N_k = round(logU(2^18, 2^22)) photons,
                 each into its own 512x512 image (channel k of [16, 512, 512]),
                 distributions as ``plain``; cap 8.
* ``scaleout``   4 views x 2^24 photons -> [12, 2048, 2048], cap 16, caustic
                 distribution, cp = 3 v + i mod 3.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass
class Photons:
    """Photon batch in the reference layout (host numpy arrays)."""

    E: np.ndarray        # [N] float32
    xp: np.ndarray       # [2, N] float32
    M: np.ndarray        # [2, 3, N] float32
    cp: np.ndarray       # [1, N] int64
    radius: np.ndarray   # [N] float32
    size: tuple          # (C, H, W)
    cap: int             # max_pixel_radius
    name: str = ""

    @property
    def n(self) -> int:
        return int(self.E.shape[0])

    def arrays(self):
        return self.E, self.xp, self.M, self.cp, self.radius

    def subset(self, idx) -> "Photons":
        idx = np.asarray(idx)
        return Photons(self.E[idx].copy(), self.xp[:, idx].copy(), self.M[:, :, idx].copy(),
                       self.cp[:, idx].copy(), self.radius[idx].copy(), self.size, self.cap, self.name)

    def to_torch(self, device):
        import torch
        return tuple(torch.from_numpy(np.ascontiguousarray(a)).to(device) for a in self.arrays())


def _loguniform(rng, lo, hi, n):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), n))


def from_pixel_space(px, py, theta, s1, s2, E, cp, size, cap, name="") -> Photons:
    """Map pixel-space photons (position, axes, angle) to the reference layout."""
    C, H, W = size
    n = px.shape[0]
    c, s = np.cos(theta), np.sin(theta)
    # D = R diag(s1, s2) ; D^-1 = diag(1/s1, 1/s2) R^T
    inv00, inv01 = c / s1, s / s1
    inv10, inv11 = -s / s2, c / s2
    M = np.zeros((2, 3, n), np.float64)
    M[0, 0], M[0, 1] = inv00 * (W / 2.0), inv01 * (H / 2.0)
    M[1, 0], M[1, 1] = inv10 * (W / 2.0), inv11 * (H / 2.0)
    xp = np.stack([2.0 * px / W - 1.0, 2.0 * py / H - 1.0]).astype(np.float32)
    radius = (2.0 * np.maximum(s1, s2) / W).astype(np.float32)
    return Photons(E.astype(np.float32), xp, M.astype(np.float32), cp.reshape(1, n).astype(np.int64), radius,
                   (int(C), int(H), int(W)), int(cap), name)


def _uniform_positions(rng, n, W, H):
    return rng.uniform(0.0, W, n), rng.uniform(0.0, H, n)


def _caustic_positions(rng, n, W, H, n_lines=64, frac_uniform=0.2):
    mu = rng.uniform(0.1, 0.9, (n_lines, 2)) * np.array([W, H])
    phi = rng.uniform(0.0, math.pi, n_lines)
    sig_long = rng.uniform(0.02, 0.08, n_lines) * W
    sig_short = sig_long / 20.0
    on_line = rng.uniform(size=n) >= frac_uniform
    k = rng.integers(0, n_lines, n)
    z1, z2 = rng.standard_normal(n), rng.standard_normal(n)
    a, b = sig_long[k] * z1, sig_short[k] * z2
    cx = mu[k, 0] + np.cos(phi[k]) * a - np.sin(phi[k]) * b
    cy = mu[k, 1] + np.sin(phi[k]) * a + np.cos(phi[k]) * b
    ux, uy = _uniform_positions(rng, n, W, H)
    px = np.where(on_line, cx, ux)
    py = np.where(on_line, cy, uy)
    # keep photon centres on the image (footprints still cross the border)
    return np.clip(px, 0.0, W - 1e-3), np.clip(py, 0.0, H - 1e-3)


def _caustic_axes(rng, n):
    s_major = _loguniform(rng, 0.3, 6.0, n)
    kappa = rng.uniform(1.0, 10.0, n)
    return s_major, s_major / kappa


def plain(n=1 << 18, size=(1, 256, 256), cap=8, seed=1) -> Photons:
    """Plain splat: uniform positions, isotropic-ish differentials."""
    rng = np.random.Generator(np.random.PCG64(seed))
    C, H, W = size
    px, py = _uniform_positions(rng, n, W, H)
    theta = rng.uniform(0.0, math.pi, n)
    s1, s2 = _loguniform(rng, 0.5, 4.0, n), _loguniform(rng, 0.5, 4.0, n)
    E = rng.uniform(0.5, 1.5, n) / n
    cp = np.arange(n, dtype=np.int64) % C
    return from_pixel_space(px, py, theta, s1, s2, E, cp, size, cap, "plain")


def caustic(n=1 << 22, size=(3, 1024, 1024), cap=12, seed=2, channel_offset=0, channels=None) -> Photons:
    """Caustic-like splat: thin elongated Gaussian lines plus 20 % uniform."""
    rng = np.random.Generator(np.random.PCG64(seed))
    C, H, W = size
    nch = C if channels is None else channels
    px, py = _caustic_positions(rng, n, W, H)
    theta = rng.uniform(0.0, math.pi, n)
    s1, s2 = _caustic_axes(rng, n)
    E = rng.uniform(0.5, 1.5, n) / n
    cp = channel_offset + np.arange(n, dtype=np.int64) % nch
    return from_pixel_space(px, py, theta, s1, s2, E, cp, size, cap, "caustic")


def batched_counts(k=16, seed=3, lo=1 << 18, hi=1 << 22):
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.round(_loguniform(rng, lo, hi, k)).astype(np.int64)


def batched(k=16, side=512, cap=8, seed=3, counts=None) -> Photons:
    """Ragged batch of k independent splats; splat j writes channel j."""
    counts = batched_counts(k, seed) if counts is None else np.asarray(counts, np.int64)
    parts = []
    for j, nk in enumerate(counts):
        p = plain(int(nk), (1, side, side), cap, seed=seed * 1000 + j)
        p.cp[:] = j
        parts.append(p)
    out = concat(parts, (len(counts), side, side), cap)
    out.name = "batched"
    return out


def concat(parts, size, cap) -> Photons:
    return Photons(np.concatenate([p.E for p in parts]), np.concatenate([p.xp for p in parts], axis=1),
                   np.concatenate([p.M for p in parts], axis=2), np.concatenate([p.cp for p in parts], axis=1),
                   np.concatenate([p.radius for p in parts]), tuple(size), int(cap))


def scaleout_view(view, n=1 << 24, side=2048, cap=16, seed=6) -> Photons:
    """One of the 4 views of the scale-out workload: channels 3v..3v+2."""
    return caustic(n, (12, side, side), cap, seed=seed * 100 + view, channel_offset=3 * view, channels=3)


def upstream_grad(size, seed=11) -> np.ndarray:
    """N(0, 1) upstream gradient image."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.standard_normal(size).astype(np.float32)


WORKLOADS = {"plain": plain, "caustic": caustic, "batched": batched}
