"""Autotuner of the binned photon-differential splatting path.

Chooses, per problem class, the knobs of ``pds_tuning`` (include/pds.h):

* ``tile``          screen tile edge (8, 16, 32, 64 px);
* ``seg_photons``   records per tile segment, i.e. the split threshold of
                    crowded tiles over several workgroups;
* ``block_photons`` photons per binning workgroup;
* ``capacity_pct``  record capacity of the workspace, per 100 photons (the
                    bin capacity): sized from the measured records per photon
                    with headroom, so the photon-order fallback stays unused
                    and the workspace no larger than needed.

None of them changes results beyond FP32 summation order; the semantic
``max_pixel_radius`` (``PyOptix/kernel/photon_differentials.cu:80-81``) is
never altered.  The tuner times ``pds_forward`` (keeping its bins) plus
``pds_backward`` -- the pair ``PDS`` runs -- on the caller's own photons with
HIP events, and caches the winner in a JSON file keyed by device
architecture, image size, cap, mode and photon-count bucket
(``$PDS_TUNE_CACHE``, default ``~/.cache/shapefromcaustics_amd/pds_tune.json``).
``tuning="cached"`` in the Python API uses the cached entry.
"""
from __future__ import annotations

import itertools
import json
import math
import os
import threading
import typing as tp

import torch

_lock = threading.Lock()

TILES = (8, 16, 32, 64)
SEGS = (512, 1024, 2048)
BLOCKS = (1024, 2048, 4096, 8192)
HEADROOM = 1.25  # capacity = measured records x HEADROOM (+ 5 per 100 photons)


def cache_path() -> str:
    p = os.environ.get("PDS_TUNE_CACHE")
    if p:
        return p
    base = os.environ.get("XDG_CACHE_HOME", os.path.join(os.path.expanduser("~"), ".cache"))
    return os.path.join(base, "shapefromcaustics_amd", "pds_tune.json")


def _arch(device=None) -> str:
    try:
        return torch.cuda.get_device_properties(device or torch.cuda.current_device()).gcnArchName.split(":")[0]
    except Exception:  # no device: the key still works for lookups made on CPU
        return "unknown"


def problem_key(N: int, size, cap: int, mode, arch: tp.Optional[str] = None) -> str:
    C, H, W = [int(s) for s in size]
    bucket = 0 if N <= 0 else int(round(math.log2(N)))
    m = mode if isinstance(mode, str) else {0: "auto", 1: "plain", 2: "binned", 3: "deterministic"}[int(mode)]
    return f"{arch or _arch()}|C{C}xH{H}xW{W}|cap{int(cap)}|N2^{bucket}|{m}"


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(N: int, size, cap: int, mode="auto") -> tp.Optional[dict]:
    """Cached tuning for this problem class, or None (built-in defaults)."""
    with _lock:
        e = _load().get(problem_key(N, size, cap, mode))
    return None if e is None else {k: int(e[k]) for k in ("tile", "seg_photons", "block_photons", "capacity_pct")}


def store(N: int, size, cap: int, mode, entry: dict) -> None:
    path = cache_path()
    with _lock:
        d = _load()
        d[problem_key(N, size, cap, mode)] = entry
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + f".tmp{os.getpid()}"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, path)


def _time_pair(inputs, grad, size, cap, mode, tuning, reps, warmup) -> float:
    from .PhotonDifferentialSplatting import _backward, _forward
    E, xp, M, cp, radius = inputs
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
    for _ in range(warmup):
        _, bins = _forward(E, xp, M, cp, radius, size, cap, mode, tuning, keep_bins=True)
        _backward(grad, E, xp, M, cp, radius, cap, mode, tuning, bins=bins)
    for r in range(reps):
        ev[2 * r].record()
        _, bins = _forward(E, xp, M, cp, radius, size, cap, mode, tuning, keep_bins=True)
        _backward(grad, E, xp, M, cp, radius, cap, mode, tuning, bins=bins)
        ev[2 * r + 1].record()
    torch.cuda.synchronize()
    times = sorted(ev[2 * r].elapsed_time(ev[2 * r + 1]) for r in range(reps))
    return times[len(times) // 2] * 1e3  # us


def capacity_for(records: int, N: int) -> int:
    """Record capacity (per 100 photons) for a measured record count."""
    if N <= 0:
        return 200
    return max(105, int(math.ceil(records * HEADROOM * 100.0 / N)) + 5)


def tune(E, xp, M, cp, radius, size, cap: int, mode="auto", grad=None, tiles=TILES, segs=SEGS, blocks=BLOCKS,
         reps=5, warmup=2, save=True, log=None) -> dict:
    """Time every candidate on these photons; return (and cache) the best.

    Returns ``{"tile", "seg_photons", "block_photons", "capacity_pct", "us",
    "records_per_photon"}``.  Candidates whose tile gives more bins than the
    LDS histogram holds are bumped to a larger tile by the library itself and
    are skipped here.
    """
    from .PhotonDifferentialSplatting import _bins_stats, _forward
    size = tuple(int(v) for v in size)
    N = int(E.size(0))
    inputs = (E, xp, M, cp, radius)
    if grad is None:
        g = torch.Generator(device=E.device)
        g.manual_seed(0)
        grad = torch.randn(size, device=E.device, dtype=torch.float32, generator=g)
    best = None
    for T in tiles:
        # records at this tile edge (generous capacity), then the capacity to use
        probe = {"tile": T, "capacity_pct": 1000}
        _, bins = _forward(E, xp, M, cp, radius, size, cap, mode, probe, keep_bins=True)
        if bins is None:
            raise RuntimeError("this problem is not binned (plain mode, tiny image or cap > 63): nothing to tune")
        st = _bins_stats(bins, N, size, cap, mode, probe)
        del bins
        if st["tile"] != T or st["overflow"]:
            continue  # tile too small for the bin histogram (the library raised it), or a pathological input
        cap_pct = capacity_for(st["records"], N)
        for seg, bp in itertools.product(segs, blocks):
            tun = {"tile": T, "seg_photons": seg, "block_photons": bp, "capacity_pct": cap_pct}
            us = _time_pair(inputs, grad, size, cap, mode, tun, reps, warmup)
            if log:
                log(f"tile {T:2d} seg {seg:5d} block {bp:5d} capacity {cap_pct}%: {us:8.1f} us")
            if best is None or us < best["us"]:
                best = dict(tun, us=us, records_per_photon=st["records"] / max(N, 1))
    if best is None:
        raise RuntimeError("no tuning candidate ran without overflow")
    if save:
        store(N, size, cap, mode, best)
    return best
