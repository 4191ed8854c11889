#!/usr/bin/env python3
"""Photon-differential splatting benchmark: photons/s of one forward + backward
pass on one MI355X, with the dominant kernel's HBM roofline fraction and a
multithreaded host-CPU baseline of the same formulae.

Default workload (the caustic-like large splat): 2^22 photons into a
3 x 1024 x 1024 fp32 image, max_pixel_radius 12, caustic-line positions,
anisotropic differentials, N(0, 1) upstream gradient; inputs resident in HBM
before the timed region.  One "step" = pds_forward + pds_backward of the whole
batch.  Prints ONE JSON line on stdout; progress goes to stderr.

    python bench.py [--steps K] [--warmup W] [--mode auto|plain|binned|deterministic]

The same line carries, under "configs", the batched inner simulations (16
ragged splats under one HIP graph vs 16 sequential plain-atomic launches),
one Landweber step of height-field reconstruction and the 4 x 2^24-photon
scale-out (atomic and deterministic, photons regenerated), each with its own
comparator and roofline; --workload batched|landweber|scaleout prints one of
them alone.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

# Algorithmic bytes (reference tensor layout, fp32 + int64 channel):
#   forward : reads E 4 + xp 8 + M 24 + cp 8 + radius 4 = 48 B per photon, writes 4 B per pixel
#   backward: reads the same 48 B, writes gE 4 + gx 8 + gM 24 = 36 B per photon, reads 4 B per pixel
FWD_B_PER_PHOTON, BWD_B_PER_PHOTON, B_PER_PIXEL_PASS = 48, 84, 4
EVENT_EVERY = 4  # timed steps per instrumented step sampled after them (see timed())


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name):
    from shapefromcaustics_amd import workloads as WL
    if name == "caustic":
        p = WL.caustic()
        desc = "caustic: 2^22 photons -> 3x1024x1024 fp32, cap 12, 64 caustic lines + 20% uniform"
    elif name == "plain":
        p = WL.plain()
        desc = "plain: 2^18 photons -> 1x256x256 fp32, cap 8, uniform"
    else:
        raise SystemExit(f"unknown workload {name}")
    return p, desc


def cpu_baseline(p, grad, budget_s=12.0):
    """OpenMP FP32 host implementation of the same formulae on the same inputs,
    built for this host (-march=native) and run on every CPU this process may
    use (the box's share for one GPU)."""
    from oracle import pds_oracle as O
    flags = O.use_native_cpu_baseline()
    host = O.host_cpu()
    threads = O.cpu_threads()
    times = []
    t_start = time.perf_counter()
    while len(times) < 60 and (time.perf_counter() - t_start) < budget_s:
        t0 = time.perf_counter()
        O.cpu_forward_f32(*p.arrays(), p.size, p.cap)
        O.cpu_backward_f32(grad, *p.arrays(), p.cap)
        times.append(time.perf_counter() - t0)
    times.sort()
    t = times[len(times) // 2]
    return {"value": p.n / t, "unit": "photons/s", "cores": threads, "kind": "port",
            "sample": f"the full benchmark batch ({p.n} photons, fwd+bwd), median of {len(times)} runs "
                      f"on {threads} OpenMP threads of the host",
            "host_cpu": host["model"], "host_logical_cpus": host["logical_cpus"],
            "usable_cpus": host["usable_cpus"], "build": flags}


def load_traffic(kernel, workload):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        e = d.get(workload, {}).get(kernel)
        return None if e is None else float(e["bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


def landweber_fields():
    """The config-4 height fields: 256^2 heights, 32 Gaussian bumps (amplitude
    <= 0.5 mm, sigma 0.3-2 mm over 10 x 10 mm) on a 1 mm base."""
    import numpy as np
    R, base = 256, 0.2

    def field(seed):
        rng = np.random.Generator(np.random.PCG64(seed))
        a, sg = rng.uniform(0.0, 0.1, 32), rng.uniform(0.06, 0.4, 32)
        mu = rng.uniform(-0.6, 0.6, (32, 2))
        lin = np.linspace(-1, 1, R)
        X, Y = np.meshgrid(lin, lin)
        h = np.full((R, R), base)
        for k in range(32):
            h += a[k] * np.exp(-((X - mu[k, 0]) ** 2 + (Y - mu[k, 1]) ** 2) / (2 * sg[k] ** 2))
        return h
    return R, base, field


def bench_landweber(steps, warmup, with_plain=True):
    """One pixel-Landweber iteration of height-field reconstruction, end to end
    on the GPU: Sobel normals, 2048^2 parallel rays traced and refracted
    through a 256^2 height field (32 Gaussian bumps, 1 mm base, n = 1.5) to a
    sensor 50 mm below, splatted into 1024^2 (cap 12); loss 1/2 |I - T|^2
    against the image of a differently seeded field; PDS adjoint, path
    adjoint (replaying the forward's ray hits) and the masked shrinkage +
    volume update, with the volume band the reference derives from the
    target's printing volume (shape_from_caustics.py:383-384).  Returns a dict
    (ms/step, rays/s, stage split, the same step with the plain-atomic splat)."""
    import torch

    from shapefromcaustics_amd import heightfield as HF
    from shapefromcaustics_amd.PhotonDifferentialSplatting import _backward, _forward

    R, base, field = landweber_fields()
    n_side, S, cap, sz = 2048, 1024, 12, -10.0
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    sc, em = HF.Scene(R, sz, (1.5,)), HF.ParallelEmitter(n_side, 2.0, 1.0)
    size = (1, S, S)
    h_target = torch.tensor(field(5), dtype=torch.float32, device=dev)
    target = HF.CausticImage.apply(h_target, sc, em, size, cap).detach()
    d0 = torch.tensor(field(4) - base, dtype=torch.float32, device=dev).contiguous()
    d = d0.clone()
    mask = torch.ones_like(d)
    params = HF.LandweberParams.for_volume(HF.printing_volume(h_target - base), R)
    N = n_side * n_side

    def timed_steps(mode):
        def step():
            # every timed iteration starts from the same field (a 256^2 copy),
            # so each measures the same physical problem
            d.copy_(d0)
            return HF.landweber_step(d, base, target, sc, em, size, cap, mask, params, mode=mode)
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    wall = timed_steps("auto")
    # stage split on one extra step
    e = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
    d.copy_(d0)
    h = (d + base).contiguous()
    e[0].record()
    n = HF.normals(h)
    Ep, xp, Mp, cp, radius, hits = HF.simulate(h, sc, em, n, keep_hits=True)
    e[1].record()
    img, bins = _forward(Ep, xp, Mp, cp, radius, size, cap, "auto", None, keep_bins=True)
    resid, lossv = HF.residual(img, target)
    e[2].record()
    gE, gx, gM = _backward(resid, Ep, xp, Mp, cp, radius, cap, "auto", None, bins=bins)
    e[3].record()
    gh = HF.adjoint(h, sc, em, gE, gx, gM, n, hits=hits)
    e[4].record()
    HF.landweber_update(d, gh, mask, params)
    e[5].record()
    torch.cuda.synchronize()
    stats = HF.path_stats(hits, sc, em)
    names = ["normals+simulate", "splat+residual", "splat adjoint", "path adjoint", "update"]
    stages = {k: round(e[i].elapsed_time(e[i + 1]) * 1e3, 1) for i, k in enumerate(names)}
    n_sensor = int((cp >= 0).sum())
    splat_bytes = 132 * N + 8 * S * S  # the PDS pair over the step's photon buffers (reference layout)
    splat_us = stages["splat+residual"] + stages["splat adjoint"]
    plain = timed_steps("plain") if with_plain else None
    log(f"[bench] landweber step {wall * 1e3:.3f} ms; stages (us) {stages}; loss {float(lossv):.4e}; "
        f"photons reaching the sensor {n_sensor} of {N}; paths {stats}; volume band "
        f"[{params.d1:.1f}, {params.d2:.1f}]" + (f"; plain-atomic splat {plain * 1e3:.3f} ms" if plain else ""))
    return {
        "workload": "landweber: 256^2 height field, 2048^2 parallel rays -> 1x1024x1024, cap 12, sensor 50 mm, "
                    "n 1.5, volume band of the target (reference default step)",
        "ms_per_step": round(wall * 1e3, 4), "steps_per_s": 1.0 / wall, "rays": N, "rays_per_s": N / wall,
        "photons_on_sensor": n_sensor, "path_outcomes": stats, "stages_us": stages,
        "volume_band": [params.d1, params.d2],
        "splat_roofline": {"bound": "hbm", "achieved": round(splat_bytes / (splat_us * 1e-6) / 1e9, 1),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(splat_bytes / (splat_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                           "algorithmic_bytes": splat_bytes, "what": "PDS forward + adjoint stages of the step"},
        "plain_atomic_ms_per_step": None if plain is None else round(plain * 1e3, 4),
        "speedup_vs_plain": None if plain is None else round(plain / wall, 3),
    }


def bench_scaleout(steps, warmup, with_plain=True):
    """Memory-bound scale-out on one GPU: 4 views x 2^24 caustic-line photons
    into 12 x 2048 x 2048 (cap 16).  A step generates the photons (counter-
    based source), splats them keeping the bins, drops them, regenerates them
    for the backward (recomputation instead of storage) and runs the adjoint.
    Timed in the atomic and the deterministic mode; the plain-atomic
    comparator keeps its photons (no generation).  Returns a dict."""
    import torch

    from shapefromcaustics_amd.generate import ScaleOut
    from shapefromcaustics_amd.PhotonDifferentialSplatting import _backward, _forward

    torch.cuda.set_device(0)
    so = ScaleOut()
    grad = torch.randn(so.size, device="cuda", generator=torch.Generator("cuda").manual_seed(11))

    def step(mode):
        ph = so.photons()
        img, bins = _forward(*ph, so.size, so.cap, mode, None, keep_bins=True)
        del ph
        ph = so.photons()
        g = _backward(grad, *ph, so.cap, mode, None, bins=bins)
        return img, g

    res = {}
    for mode in ("auto", "deterministic"):
        for _ in range(warmup):
            step(mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step(mode)
        torch.cuda.synchronize()
        res[mode] = (time.perf_counter() - t0) / steps
        log(f"[bench] scaleout {mode}: {res[mode] * 1e3:.2f} ms/step ({so.n / res[mode]:.3e} photons/s)")
    plain = None
    if with_plain:
        ph = so.photons()
        _forward(*ph, so.size, so.cap, "plain", None, keep_bins=False)
        _backward(grad, *ph, so.cap, "plain", None)
        torch.cuda.synchronize()
        reps = 2
        t0 = time.perf_counter()
        for _ in range(reps):
            _forward(*ph, so.size, so.cap, "plain", None, keep_bins=False)
            _backward(grad, *ph, so.cap, "plain", None)
        torch.cuda.synchronize()
        plain = (time.perf_counter() - t0) / reps
        del ph
        log(f"[bench] scaleout plain-atomic (photons kept, no generation): {plain * 1e3:.2f} ms/step")
    wall = res["auto"]
    nbytes = 132 * so.n + 8 * so.size[0] * so.size[1] * so.size[2]
    return {
        "workload": "scaleout: 4 views x 2^24 photons -> 12x2048x2048 fp32, cap 16, caustic lines; photons "
                    "generated on the GPU and regenerated for the backward",
        "photons": so.n, "ms_per_step": round(wall * 1e3, 4), "photons_per_s": so.n / wall,
        "deterministic_ms_per_step": round(res["deterministic"] * 1e3, 4),
        "deterministic_photons_per_s": so.n / res["deterministic"],
        "roofline": {"bound": "hbm", "achieved": round(nbytes / wall / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(nbytes / wall / 1e9 / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes": nbytes, "what": "whole step, stored-photon byte convention"},
        "plain_atomic_ms_per_step": None if plain is None else round(plain * 1e3, 4),
        "speedup_vs_plain": None if plain is None else round(plain / wall, 3),
    }


def bench_batched(steps, warmup, with_plain=True, tune=True):
    """Batched inner simulations: 16 splats of logU(2^18, 2^22) photons
    (19,365,832, seed 3) into 16 x 512^2, cap 8, as one forward + backward
    replayed from a HIP graph (RaggedBatch + GraphedStep) with the
    autotuner's tile, segment and block choice for this batch, against the
    reference's way: 16 sequential plain-atomic forward + backward launches,
    one per image.  Also a replay at a quarter of the load (same graph).
    Returns a dict."""
    import numpy as np
    import torch

    from shapefromcaustics_amd import autotune
    from shapefromcaustics_amd import workloads as WL
    from shapefromcaustics_amd.batch import GraphedStep, RaggedBatch
    from shapefromcaustics_amd.PhotonDifferentialSplatting import _backward, _forward

    torch.cuda.set_device(0)
    p = WL.batched()
    K, side, cap = p.size[0], p.size[1], p.cap
    offs = np.concatenate([[0], np.cumsum(WL.batched_counts(K))])
    dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (p.E, p.xp, p.M, p.radius)]
    splat = lambda a, b: (dev[0][a:b], dev[1][:, a:b], dev[2][:, :, a:b], dev[3][a:b])
    grad = torch.from_numpy(WL.upstream_grad(p.size, seed=4)).cuda()
    tuning = None
    if tune:
        E, xp, M, radius = dev
        cp = torch.from_numpy(np.ascontiguousarray(p.cp)).cuda()
        best = autotune.tune(E, xp, M, cp, radius, p.size, cap, grad=grad, reps=3, warmup=1, save=False)
        tuning = {k: best[k] for k in ("tile", "seg_photons", "block_photons", "capacity_pct")}
        del cp
    batch = RaggedBatch(p.n, K, side, side, cap, "cuda", tuning=tuning)
    batch.load([splat(int(a), int(b)) for a, b in zip(offs[:-1], offs[1:])])
    batch.grad.copy_(grad)
    graph = GraphedStep(batch)

    def time_replays(reps):
        for _ in range(warmup):
            graph.replay()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev[0].record()
        for _ in range(reps):
            graph.replay()
        ev[1].record()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, ev[0].elapsed_time(ev[1]) * 1e-3 / reps

    wall, gpu_t = time_replays(steps)
    quarter = p.n // 4
    batch.load([splat(int(a), int(min(b, quarter))) for a, b in zip(offs[:-1], offs[1:]) if a < quarter])
    wall_q, _ = time_replays(steps)
    plain = None
    if with_plain:
        parts = []
        for k in range(K):
            a, b = int(offs[k]), int(offs[k + 1])
            parts.append((*[t.contiguous() for t in splat(a, b)],
                          torch.zeros((1, b - a), dtype=torch.int64, device="cuda"), grad[k:k + 1].contiguous()))

        def run_plain():
            for e, x, m, r, c, g in parts:
                _forward(e, x, m, c, r, (1, side, side), cap, "plain", None, keep_bins=False)
                _backward(g, e, x, m, c, r, cap, "plain", None)
        run_plain()
        torch.cuda.synchronize()
        reps = max(2, steps // 4)
        t0 = time.perf_counter()
        for _ in range(reps):
            run_plain()
        torch.cuda.synchronize()
        plain = (time.perf_counter() - t0) / reps
    nbytes = 132 * p.n + 8 * K * side * side
    log(f"[bench] batched graph replay {wall * 1e3:.3f} ms ({p.n / wall:.3e} photons/s), quarter load "
        f"{wall_q * 1e3:.3f} ms; tuning {tuning}" +
        (f"; {K} sequential plain-atomic launches {plain * 1e3:.3f} ms" if plain else ""))
    return {
        "workload": "batched: 16 ragged splats (logU(2^18,2^22) photons, 19365832 total) -> 16x512x512 fp32, cap 8, "
                    "one forward + backward replayed from a HIP graph, autotuned",
        "photons": p.n, "ms_per_step": round(wall * 1e3, 4), "photons_per_s": p.n / wall,
        "graph_gpu_ms": round(gpu_t * 1e3, 4), "tuning": tuning,
        "quarter_load_ms_per_step": round(wall_q * 1e3, 4), "quarter_load_ratio": round(wall_q / wall, 3),
        "roofline": {"bound": "hbm", "achieved": round(nbytes / wall / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(nbytes / wall / 1e9 / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes": nbytes, "what": "whole replay (forward + backward of the batch)"},
        "plain_atomic_ms_per_step": None if plain is None else round(plain * 1e3, 4),
        "plain_what": f"{K} sequential plain-atomic forward + backward launches, one per image",
        "speedup_vs_plain": None if plain is None else round(plain / wall, 3),
    }


def summarize_profile(path, factor_fetch=1.0, factor_write=1.0):
    """Markdown per-kernel table of a rocprofv3 SQLite result (kernel trace
    and/or --pmc counters; FETCH_SIZE / WRITE_SIZE scaled by the given copy
    calibration factors and shown in MB per dispatch)."""
    import sqlite3
    from collections import defaultdict
    db = sqlite3.connect(path)
    cur = db.cursor()
    views = {r[0] for r in cur.execute("select name from sqlite_master where type in ('table','view')")}
    clean = lambda k: k.replace("(anonymous namespace)::", "").split("(")[0]
    if "kernels" in views:
        cols = [r[1] for r in cur.execute("pragma table_info(kernels)")]
        ki = cols.index("name") if "name" in cols else cols.index("kernel_name")
        si, ei = cols.index("start"), cols.index("end")
        d = defaultdict(list)
        for r in cur.execute("select * from kernels"):
            d[r[ki]].append((r[ei] - r[si]) / 1e3)
        print("| kernel | calls | avg us | total us |\n|---|---:|---:|---:|")
        for k, v in sorted(d.items(), key=lambda kv: -sum(kv[1])):
            print(f"| `{clean(k)}` | {len(v)} | {sum(v) / len(v):.1f} | {sum(v):.1f} |")
    view = next((v for v in ("counters", "pmc_events") if v in views), None)
    if view is not None:
        cols = [r[1] for r in cur.execute(f"pragma table_info({view})")]
        col = lambda *names: next(cols.index(n) for n in names if n in cols)
        ki, ci, vi, di = col("kernel_name", "name"), col("counter_name"), col("value", "counter_value"), \
            col("dispatch_id", "id")
        agg = defaultdict(lambda: defaultdict(float))
        nd = defaultdict(set)
        for r in cur.execute(f"select * from {view}"):
            agg[r[ki]][r[ci]] += r[vi]
            nd[r[ki]].add(r[di])
        names = sorted({c for dd in agg.values() for c in dd})
        if not names:
            return
        print("\n| kernel | dispatches | " + " | ".join(names) + " |\n|---|---:|" + "---:|" * len(names))
        for k, dd in agg.items():
            n = len(nd[k])
            vals = []
            for c in names:
                v = dd.get(c, 0.0) / n
                if c == "FETCH_SIZE":
                    v = v * factor_fetch * 1024 / 1e6
                elif c == "WRITE_SIZE":
                    v = v * factor_write * 1024 / 1e6
                vals.append(f"{v:.4g}")
            print(f"| `{clean(k)}` | {n} | " + " | ".join(vals) + " |")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="caustic")
    ap.add_argument("--mode", default="auto")
    ap.add_argument("--no-cpu", action="store_true", help="skip the host-CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip timing the plain-atomic comparator")
    ap.add_argument("--stateless", action="store_true",
                    help="time pds_forward + pds_backward (the backward bins again) instead of the PDS pairing")
    ap.add_argument("--graph", action="store_true", help="capture the step in a HIP graph and time its replays")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the batched, Landweber and scale-out sub-measurements of the default run")
    ap.add_argument("--eager", action="store_true", help="never use a HIP graph")
    ap.add_argument("--tuned", action="store_true", help="use the autotuner's cached tuning for this problem")
    ap.add_argument("--tuning", metavar="JSON", help='explicit pds_tuning fields, e.g. \'{"block_photons": 16384}\'')
    ap.add_argument("--summarize", metavar="DB", help="print a Markdown summary of a rocprofv3 result and exit")
    ap.add_argument("--fetch-factor", type=float, default=1.0, help="FETCH_SIZE calibration factor (--summarize)")
    ap.add_argument("--write-factor", type=float, default=1.0, help="WRITE_SIZE calibration factor (--summarize)")
    args = ap.parse_args()
    if args.summarize:
        return summarize_profile(args.summarize, args.fetch_factor, args.write_factor)
    if args.eager:
        args.graph = False
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (photon sharding plus one image "
              "all-reduce is the design, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload in ("landweber", "scaleout", "batched"):
        # one workload's own line (the default caustic run carries all three as sub-objects)
        fn = {"landweber": bench_landweber, "scaleout": bench_scaleout, "batched": bench_batched}[args.workload]
        d = fn(args.steps, args.warmup, with_plain=not args.no_plain)
        unit_v = ("steps/s", d["steps_per_s"]) if args.workload == "landweber" else ("photons/s", d["photons_per_s"])
        line = {"metric": f"{args.workload}: {unit_v[0]}", "value": unit_v[1], "unit": unit_v[0], "n_gpus": 1,
                "steps": args.steps, "warmup": args.warmup, "ms_per_step": d["ms_per_step"],
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                "data": "synthetic", "config": {"workload": d["workload"]}, "detail": d}
        print(json.dumps(line), flush=True)
        return
    if args.workload == "copy":
        # calibration for HBM byte counters: 5 device-to-device copies of 1 GiB
        # (1048576 KiB read and written per copy)
        import torch
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] copy calibration: 5 copies of 1 GiB")
        return

    import numpy as np
    import torch

    from shapefromcaustics_amd import workloads as WL
    from shapefromcaustics_amd.PhotonDifferentialSplatting import _backward, _forward

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    log(f"[bench] generating workload {args.workload}")
    p, desc = make_workload(args.workload)
    grad_np = WL.upstream_grad(p.size, seed=2)
    E, xp, M, cp, radius = p.to_torch(dev)
    grad = torch.from_numpy(grad_np).to(dev)
    C, H, W = p.size
    N = p.n
    torch.cuda.synchronize()

    def step(mode, ev=None):
        # the calls PDS.apply / PDS.backward make: the forward keeps its tile
        # binning for the backward (the stateless module API bins twice)
        keep = not args.stateless
        if ev is not None:
            ev[0].record()
        img, bins = _forward(E, xp, M, cp, radius, p.size, p.cap, mode, tuning, keep_bins=keep)
        if ev is not None:
            ev[1].record()
        g = _backward(grad, E, xp, M, cp, radius, p.cap, mode, tuning, bins=bins)
        if ev is not None:
            ev[2].record()
        return img, g

    def timed(run, steps, warmup, split=True):
        """run(ev) does one step, recording ev[0..2] (start, after forward, end)
        when ev is not None.  The wall clock times `steps` steps that carry no
        events (an event record costs a few microseconds of stream time); the
        pass durations are the medians of HIP events over steps // 4 (at least
        3) further steps, run after the timed region."""
        for _ in range(warmup):
            run(None)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            run(None)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / steps
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(max(3, steps // EVENT_EVERY))]
        for ev in evs:
            run(ev)
        torch.cuda.synchronize()
        if not split:
            tot = sorted(e[0].elapsed_time(e[2]) for e in evs)
            return wall, tot[len(tot) // 2] * 1e-3, None
        fwd = sorted(e[0].elapsed_time(e[1]) for e in evs)
        bwd = sorted(e[1].elapsed_time(e[2]) for e in evs)
        return wall, fwd[len(fwd) // 2] * 1e-3, bwd[len(bwd) // 2] * 1e-3

    tuning = "cached" if args.tuned else (json.loads(args.tuning) if args.tuning else None)
    graph = None
    if args.graph:
        # the whole step (forward keeping its bins + backward) captured once
        # and replayed: HIP graph launch instead of ~10 kernel launches
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(2):
                step(args.mode)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            step(args.mode)

    def run_main(ev):
        if graph is None:
            step(args.mode, ev)
            return
        if ev is not None:
            ev[0].record()
        graph.replay()
        if ev is not None:
            ev[1].record()
            ev[2].record()

    log(f"[bench] {desc}; mode={args.mode}{' (HIP graph)' if graph is not None else ''}; "
        f"warmup {args.warmup}, steps {args.steps}")
    if graph is None:
        wall, t_fwd, t_bwd = timed(run_main, args.steps, args.warmup)
        log(f"[bench] {args.mode}: {wall * 1e3:.3f} ms/step (fwd {t_fwd * 1e6:.1f} us, bwd {t_bwd * 1e6:.1f} us)")
    else:
        wall, t_step, _ = timed(run_main, args.steps, args.warmup, split=False)
        # forward and backward split measured on the eager sequence
        _, t_fwd, t_bwd = timed(lambda ev: step(args.mode, ev), max(3, args.steps // 4), 1)
        log(f"[bench] {args.mode} graph: {wall * 1e3:.3f} ms/step (graph {t_step * 1e6:.1f} us; eager fwd "
            f"{t_fwd * 1e6:.1f} us, bwd {t_bwd * 1e6:.1f} us)")

    plain = None
    if not args.no_plain and args.mode != "plain":
        pw, pf, pb = timed(lambda ev: step("plain", ev), max(3, args.steps // 4), 2)
        label = "plain-atomic comparator"
        plain = {"value": N / pw, "ms_per_step": pw * 1e3, "fwd_us": pf * 1e6, "bwd_us": pb * 1e6, "what": label}
        log(f"[bench] {label}: {pw * 1e3:.3f} ms/step")

    # dominant pass: the backward is one kernel launch; the forward's kernels
    # are bracketed together (zero fill + splat, or binning + tile passes)
    pix = C * H * W
    fwd_bytes = FWD_B_PER_PHOTON * N + B_PER_PIXEL_PASS * pix
    bwd_bytes = BWD_B_PER_PHOTON * N + B_PER_PIXEL_PASS * pix
    if t_bwd >= t_fwd:
        kname, kbytes, kt = "pds_backward", bwd_bytes, t_bwd
    else:
        kname, kbytes, kt = "pds_forward", fwd_bytes, t_fwd
    achieved = kbytes / kt / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(kname, args.workload),
                "kernel": kname, "algorithmic_bytes": kbytes, "avg_duration_us": round(kt * 1e6, 2),
                "step_achieved_GBps": round((fwd_bytes + bwd_bytes) / wall / 1e9, 1)}
    if roofline["traffic"]:
        # the bytes rocprof counted for this pass (profiles/), moved in the pass time measured here
        roofline["traffic_GBps"] = round(roofline["traffic"] / kt / 1e9, 1)
        roofline["traffic_frac"] = round(roofline["traffic"] / kt / 1e9 / HBM_PEAK_GBS, 4)

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (OpenMP FP32, same inputs)")
        cpu = cpu_baseline(p, grad_np)
        log(f"[bench] CPU: {cpu['value']:.3e} photons/s on {cpu['cores']} threads")

    configs = None
    if args.workload == "caustic" and not args.no_configs:
        # the other configurations' own lines, each with its comparator and roofline
        del E, xp, M, cp, radius, grad
        torch.cuda.empty_cache()
        configs = {}
        for name, fn, k, w in (("batched", bench_batched, 10, 2), ("landweber", bench_landweber, 10, 2),
                               ("scaleout", bench_scaleout, 5, 1)):
            log(f"[bench] config {name}")
            configs[name] = fn(min(args.steps, k), min(args.warmup, w), with_plain=not args.no_plain)
            torch.cuda.empty_cache()

    line = {
        "metric": "photons/sec fwd+bwd @2^22 photons->1024^2 (% HBM roofline), batched vs plain",
        "value": N / wall,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded numpy PCG64 photons and N(0,1) upstream gradient)",
        "config": {"workload": desc, "photons": N, "image": [C, H, W], "max_pixel_radius": p.cap,
                   "mode": args.mode, "hip_graph": graph is not None, "tuning": tuning or "default",
                   "parallelism": "single GPU"},
        "fwd_us": round(t_fwd * 1e6, 2),
        "bwd_us": round(t_bwd * 1e6, 2),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain_atomic": plain,
        "speedup_vs_plain": None if plain is None else round(plain["ms_per_step"] / (wall * 1e3), 3),
        "speedup_vs_cpu": None if cpu is None else round((N / wall) / cpu["value"], 1),
        "configs": configs,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
