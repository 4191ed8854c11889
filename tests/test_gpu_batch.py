"""GPU: ragged batches of independent splats (one call, splat k -> channel k),
their HIP-graph replay with new photon counts (a device-side live count, so
that a replay costs the loaded photons), and the autotuner.

Each splat of a batch is checked against the FP64 oracle run on that splat
alone into a one-channel image: channel routing makes the batch exactly K
independent problems.  Graph replays are checked against the eager call on
the same buffers, replay after replay, with unrelated eager work and
synchronisations in between.
"""
import numpy as np
import pytest
import torch

from oracle import pds_oracle as O
from shapefromcaustics_amd import autotune
from shapefromcaustics_amd import workloads as WL
from shapefromcaustics_amd.batch import GraphedStep, RaggedBatch
from shapefromcaustics_amd.PhotonDifferentialSplatting import _forward

pytestmark = pytest.mark.gpu

SIDE, CAP = 128, 8


def _splats(counts, seed):
    return [WL.plain(int(n), (1, SIDE, SIDE), CAP, seed=seed * 100 + k) for k, n in enumerate(counts)]


def _load(batch, splats):
    batch.load([tuple(torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (p.E, p.xp, p.M, p.radius))
                for p in splats])


def _check_against_oracle(img, grads, batch, splats, grad_np):
    gE, gx, gM = [t.cpu().numpy() for t in grads]
    for k, p in enumerate(splats):
        ref = O.forward_f64(*p.arrays(), p.size, CAP)[0]
        got = img[k].cpu().numpy().astype(np.float64)
        m = ref > 1e-3 * ref.max()
        assert (np.abs(got - ref)[m] / ref[m]).max() <= 1e-4, f"splat {k} image"
        a, b = batch.offsets[k], batch.offsets[k + 1]
        rg = O.backward_f64(grad_np[k:k + 1], *p.arrays(), CAP)
        ok, worst = O.grads_within_tolerance([gE[a:b], gx[:, a:b], gM[:, :, a:b]], rg, rtol=1e-3)
        assert ok, f"splat {k} gradients: {worst:.3e}"
    # photons beyond the live count (earlier loads' photons, still in the
    # buffers' tail) contribute nothing: the images above match the oracle on
    # the loaded photons alone, and unused channels stay empty
    assert not img[len(splats):].any(), "unused channels"


@pytest.mark.parametrize("mode", ["auto", "deterministic"])
def test_ragged_batch_equals_independent_splats(mode):
    counts = [3000, 20000, 700, 9000, 1]
    splats = _splats(counts, seed=1)
    batch = RaggedBatch(40000, 6, SIDE, SIDE, CAP, "cuda", mode=mode)
    _load(batch, splats)
    grad_np = WL.upstream_grad(batch.size, seed=4)
    batch.grad.copy_(torch.from_numpy(grad_np))
    img, grads = batch.step()
    torch.cuda.synchronize()
    _check_against_oracle(img, grads, batch, splats, grad_np)


@pytest.mark.parametrize("mode", ["auto", "deterministic"])
def test_graph_replays_with_new_counts(mode):
    batch = RaggedBatch(60000, 4, SIDE, SIDE, CAP, "cuda", mode=mode)
    _load(batch, _splats([5000, 100, 20000, 7000], seed=2))
    grad_np = WL.upstream_grad(batch.size, seed=5)
    batch.grad.copy_(torch.from_numpy(grad_np))
    graph = GraphedStep(batch)
    for rep, counts in enumerate([[5000, 100, 20000, 7000], [30000, 2000], [1, 15000, 15000, 29000],
                                  [8000, 8000, 8000, 8000]]):
        splats = _splats(counts, seed=10 + rep)
        _load(batch, splats)
        junk = torch.randn(1 << 20, device="cuda").cumsum(0)  # unrelated eager work between replays
        img, grads = graph.replay()
        img, grads = img.clone(), [g.clone() for g in grads]
        torch.cuda.synchronize()
        e_img, e_grads = batch.step()
        torch.cuda.synchronize()
        if mode == "deterministic":
            assert torch.equal(img, e_img), f"replay {rep}: image differs from the eager call"
        else:
            torch.testing.assert_close(img, e_img, rtol=1e-5, atol=1e-7 * float(e_img.abs().max()))
        n = batch.offsets[-1]  # the loaded photons (the tail's gradients are not written)
        for g, e in zip(grads, e_grads):
            g, e = g[..., :n], e[..., :n]
            torch.testing.assert_close(g, e, rtol=1e-4, atol=1e-6 * float(e.abs().max()) + 1e-30)
        _check_against_oracle(img, grads, batch, splats, grad_np)
        del junk


def test_autotuner_picks_and_caches(tmp_path, monkeypatch):
    monkeypatch.setenv("PDS_TUNE_CACHE", str(tmp_path / "tune.json"))
    p = WL.plain(1 << 15, (1, 256, 256), 8, seed=9)
    E, xp, M, cp, radius = p.to_torch("cuda")
    best = autotune.tune(E, xp, M, cp, radius, p.size, p.cap, tiles=(16, 32), segs=(512, 1024), blocks=(4096,),
                         reps=2, warmup=1)
    assert best["tile"] in (16, 32) and best["seg_photons"] in (512, 1024) and best["capacity_pct"] >= 105
    got = autotune.lookup(p.n, p.size, p.cap, "auto")
    assert got == {k: best[k] for k in ("tile", "seg_photons", "block_photons", "capacity_pct")}
    img = _forward(E, xp, M, cp, radius, p.size, p.cap, "auto", "cached", keep_bins=False)[0]
    ref = O.forward_f64(*p.arrays(), p.size, p.cap)
    m = ref > 1e-3 * ref.max()
    assert (np.abs(img.cpu().numpy() - ref)[m] / ref[m]).max() <= 1e-4


def test_replay_costs_the_loaded_photons():
    # one graph captured for the full batched workload's capacity (19.4M
    # photons, 16 x 512^2); replays at a quarter of the load take a fraction
    # of the full-load time, because the kernels stop at the live count
    full = WL.batched()
    counts = WL.batched_counts()
    offs = np.concatenate([[0], np.cumsum(counts)])
    dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (full.E, full.xp, full.M, full.radius)]
    splat = lambda a, b: (dev[0][a:b], dev[1][:, a:b], dev[2][:, :, a:b], dev[3][a:b])
    batch = RaggedBatch(full.n, 16, 512, 512, 8, "cuda")
    batch.load([splat(a, b) for a, b in zip(offs[:-1], offs[1:])])
    batch.grad.copy_(torch.from_numpy(WL.upstream_grad(batch.size, seed=6)))
    graph = GraphedStep(batch)

    def time_replays(reps=10):
        for _ in range(2):
            graph.replay()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(reps):
            graph.replay()
        ev[1].record()
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / reps

    t_full = time_replays()
    quarter = full.n // 4
    batch.load([splat(int(a), int(min(b, quarter))) for a, b in zip(offs[:-1], offs[1:]) if a < quarter])
    t_quarter = time_replays()
    img, _ = graph.replay()
    torch.cuda.synchronize()
    print(f"\ngraph replay: full load ({full.n} photons) {t_full:.3f} ms, quarter load ({quarter}) {t_quarter:.3f} ms "
          f"({t_quarter / t_full:.2f} of full)")
    assert t_quarter <= 0.45 * t_full
    # and the quarter load's image is that of its photons only
    k = int(np.searchsorted(offs, quarter, side="right")) - 1
    a = int(offs[k])
    sub = full.subset(np.arange(a, quarter))
    ref = O.forward_f64(sub.E, sub.xp, sub.M, np.zeros((1, quarter - a), np.int64), sub.radius, (1, 512, 512), 8)[0]
    got = img[k].cpu().numpy().astype(np.float64)
    m = ref > 1e-3 * ref.max()
    assert (np.abs(got - ref)[m] / ref[m]).max() <= 1e-4
    assert not img[k + 1:].any()
