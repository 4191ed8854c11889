"""Ragged batches of independent splats as one PDS problem, replayed from a HIP graph.

The reference runs ``num_simulations`` inner simulations one after another and
sums their images (``shape_from_caustics.py:351-353``); each call to
``splat_photons`` (``model/caustics.py:233-239``) is one ``pds_forward``.  The
reference kernel routes every photon to channel ``cp`` of the output
(``PyOptix/kernel/photon_differentials.cu:60-113``), so K independent splats
of uneven photon counts are exactly one call with ``cp = k`` for splat k and
output ``[K, H, W]``: channel k is splat k's image, and the backward of the
same call gives every splat's photon gradients.

:class:`RaggedBatch` holds static device buffers of a fixed photon capacity.
Splats are packed back to back, and the number of loaded photons lives in a
one-element device tensor that the kernels read when they run
(``pds_forward_bins_live_f32`` / ``pds_backward_bins_live_f32``): one
captured graph serves any photon counts up to the capacity without
re-capture, and a replay costs what the loaded photons cost, not the
capacity.  :class:`GraphedStep` captures forward (keeping its bins) +
backward on a side stream into a ``torch.cuda.CUDAGraph`` and replays it.
The batch's tuning defaults to the autotuner's cached choice for its
problem class (``tuning="cached"``; built-in defaults when none is cached).
"""
from __future__ import annotations

import typing as tp

import torch

from .PhotonDifferentialSplatting import _backward, _forward


class RaggedBatch:
    """K splats of uneven photon counts into channels of one ``[K, H, W]`` image."""

    def __init__(self, capacity: int, K: int, H: int, W: int, max_pixel_radius: int, device,
                 mode="auto", tuning="cached"):
        self.capacity = int(capacity)
        self.size = (int(K), int(H), int(W))
        self.cap = int(max_pixel_radius)
        self.mode = mode
        self.tuning = tuning
        self.device = torch.device(device)
        n = self.capacity
        f32 = dict(dtype=torch.float32, device=self.device)
        self.E = torch.zeros(n, **f32)
        self.xp = torch.zeros(2, n, **f32)
        self.M = torch.zeros(2, 3, n, **f32)
        self.cp = torch.full((1, n), -1, dtype=torch.int64, device=self.device)
        self.radius = torch.zeros(n, **f32)
        self.grad = torch.zeros(self.size, **f32)
        self.live = torch.zeros(1, dtype=torch.int64, device=self.device)  # photons loaded
        self.offsets = [0]

    def load(self, splats: tp.Sequence[tp.Tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]]) -> None:
        """Copy splats ``(E [n], xp [2, n], M [2, 3, n], radius [n])`` in, splat k -> channel k.

        Stream-ordered device copies on the current stream, and the live
        count set to the photons loaded; the buffers' tail is neither read
        nor written (its gradients are left as they were).
        """
        if len(splats) > self.size[0]:
            raise ValueError(f"{len(splats)} splats for {self.size[0]} channels")
        total = sum(int(s[0].numel()) for s in splats)
        if total > self.capacity:
            raise ValueError(f"{total} photons exceed the batch capacity {self.capacity}")
        o = 0
        self.offsets = [0]
        for k, (E, xp, M, radius) in enumerate(splats):
            n = int(E.numel())
            self.E[o:o + n].copy_(E, non_blocking=True)
            self.xp[:, o:o + n].copy_(xp, non_blocking=True)
            self.M[:, :, o:o + n].copy_(M, non_blocking=True)
            self.radius[o:o + n].copy_(radius, non_blocking=True)
            self.cp[:, o:o + n].fill_(k)
            o += n
            self.offsets.append(o)
        self.live.fill_(o)

    def inputs(self):
        return self.E, self.xp, self.M, self.cp, self.radius

    def forward(self):
        """Forward of the loaded photons: (image, bins kept for :meth:`backward`)."""
        return _forward(*self.inputs(), self.size, self.cap, self.mode, self.tuning, keep_bins=True, live=self.live)

    def backward(self, grad: torch.Tensor, bins=None):
        if bins is None:
            _, bins = self.forward()
        return _backward(grad, *self.inputs(), self.cap, self.mode, self.tuning, bins=bins, live=self.live)

    def step(self):
        """Forward + backward of the whole batch against ``self.grad``: (image, [gE, gx, gM])."""
        img, bins = self.forward()
        grads = self.backward(self.grad, bins)
        return img, grads

    def split(self, t: torch.Tensor) -> tp.List[torch.Tensor]:
        """Per-splat views of a per-photon tensor (last dimension = photons);
        the tail beyond the loaded photons is not part of any splat."""
        return [t[..., a:b] for a, b in zip(self.offsets[:-1], self.offsets[1:])]


class GraphedStep:
    """One HIP graph replaying :meth:`RaggedBatch.step` on the batch's static buffers.

    The graph's outputs (image, gradients) are static tensors overwritten by
    every replay.  Load new splats with :meth:`RaggedBatch.load` (and set
    ``batch.grad``) between replays.
    """

    def __init__(self, batch: RaggedBatch, warmup: int = 2):
        self.batch = batch
        side = torch.cuda.Stream(device=batch.device)
        side.wait_stream(torch.cuda.current_stream(batch.device))
        with torch.cuda.stream(side):
            for _ in range(warmup):
                batch.step()
        torch.cuda.current_stream(batch.device).wait_stream(side)
        torch.cuda.synchronize(batch.device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.image, self.grads = batch.step()

    def replay(self):
        self.graph.replay()
        return self.image, self.grads
