"""Writes the golden vectors of tests/golden/ from the FP64 oracles.

    python tests/golden/make_golden.py

The reference ships no test vectors for photon-differential splatting, so
these fixtures are the oracle's outputs on small seeded inputs, committed so
that any later change of the oracle, the CPU baseline or the HIP kernels is
caught against fixed data.  Cases:

* pds_small.npz    mixed channels (incl. -1 and C), border crossing, cap 6,
                   a tiny image where the reference's window ellipse test is
                   active, E <= 0 photons: forward image and the seven
                   gradient arrays (gE, gx, gM) with their |term| magnitudes.
* heightfield.npz  a 12x12 height field, 16x16 parallel rays: Sobel normals,
                   the splat inputs of every photon that reaches the sensor,
                   the rays kept away from triangle edges, and over those
                   rays the 32x32 image, the loss against a fixed target and
                   dL/dh.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import heightfield_oracle as HO  # noqa: E402
from oracle import pds_oracle as O  # noqa: E402
from shapefromcaustics_amd import workloads as WL  # noqa: E402


def pds_cases():
    out = {}
    for name, (p, seed) in {
        "mixed": (WL.caustic(600, (3, 40, 56), 6, seed=21), 21),
        "tiny": (WL.plain(200, (1, 12, 10), 8, seed=22), 22),
    }.items():
        rng = np.random.default_rng(seed)
        if name == "mixed":
            p.cp[0, :30] = rng.integers(-1, p.size[0] + 1, 30)
            p.xp += rng.uniform(-0.05, 0.05, p.xp.shape).astype(np.float32)
            p.E[40:45] = -p.E[40:45]
        g = WL.upstream_grad(p.size, seed=seed)
        img = O.forward_f64(*p.arrays(), p.size, p.cap)
        r = O.backward_f64(g, *p.arrays(), p.cap)
        for k, v in zip(("E", "xp", "M", "cp", "radius"), p.arrays()):
            out[f"{name}_{k}"] = v
        out[f"{name}_size"] = np.array(p.size)
        out[f"{name}_cap"] = np.array(p.cap)
        out[f"{name}_grad"] = g
        out[f"{name}_image"] = img
        for k in r._fields:
            out[f"{name}_{k}"] = getattr(r, k)
    np.savez_compressed(os.path.join(HERE, "pds_small.npz"), **out)


def heightfield_case():
    import torch
    R, n_side, sz, size, cap = 12, 16, -1.5, (1, 32, 32), 6
    h = torch.from_numpy(HO.gaussian_bumps(R, seed=31, n_bumps=5, amp=(0.0, 0.06), sigma=(0.2, 0.5))).float().double()
    rays = HO.parallel_rays(n_side, 1.0)
    target = torch.from_numpy(np.random.default_rng(32).uniform(0, 1, size))
    with torch.no_grad():
        safe = HO.edge_safe_rays(HO.simulate(h, rays, 1.5, sz), n_side * n_side)
    # image, loss and dL/dh of the photons whose hits keep away from triangle
    # edges (where an FP32 trace may pick the neighbouring, equally valid triangle)
    L, g, img, path = HO.loss_and_grad(h, rays, 1.5, sz, size, cap, target, keep_rays=safe)
    normals = HO.sobel_normals(h, 2.0 / (R - 1)).permute(1, 2, 0)
    np.savez_compressed(os.path.join(HERE, "heightfield.npz"), h=h.numpy(), normals=normals.numpy(),
                        ray=path.ray.numpy(), Ep=path.Ep.detach().numpy(), xp=path.xp.detach().numpy(),
                        Mp=path.Mp.detach().numpy(), radius=path.radius.detach().numpy(), image=img.numpy(),
                        target=target.numpy(), loss=np.array(L), grad_h=g.numpy(), safe=safe.numpy(),
                        meta=np.array([R, n_side, size[1], cap]), sensor_z=np.array(sz), ior=np.array(1.5))


if __name__ == "__main__":
    pds_cases()
    heightfield_case()
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))
