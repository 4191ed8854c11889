"""Golden vectors (tests/golden/, written by tests/golden/make_golden.py).

CPU: the FP64 oracles reproduce their committed outputs, and the numpy twin
and the FP32 CPU baseline agree with them.  GPU: the HIP path, through the C
ABI, reproduces them within the stated tolerances (image 1e-4 relative on
pixels above 1e-3 max; gradients 1e-3 of the sum of |terms| plus the FP32
conditioning bound).
"""
import os

import numpy as np
import pytest

from oracle import pds_oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ("mixed", "tiny")


def _case(name):
    z = np.load(os.path.join(HERE, "pds_small.npz"))
    arrays = tuple(z[f"{name}_{k}"] for k in ("E", "xp", "M", "cp", "radius"))
    ref = O.OracleGrads(*[z[f"{name}_{k}"] for k in O.OracleGrads._fields])
    return arrays, tuple(int(v) for v in z[f"{name}_size"]), int(z[f"{name}_cap"]), z[f"{name}_grad"], \
        z[f"{name}_image"], ref


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_golden(name):
    arrays, size, cap, g, img, ref = _case(name)
    assert np.allclose(O.forward_f64(*arrays, size, cap), img, rtol=1e-13, atol=0)
    got = O.backward_f64(g, *arrays, cap)
    for a, b in zip(got, ref):
        assert np.allclose(a, b, rtol=1e-12, atol=1e-300, equal_nan=True)


@pytest.mark.parametrize("name", CASES)
def test_twin_and_cpu_baseline_match_golden(name):
    arrays, size, cap, g, img, ref = _case(name)
    m = img > 1e-3 * img.max()
    tw = O.twin_forward(*arrays, size, cap)
    assert (np.abs(tw - img)[m] / img[m]).max() <= 1e-12
    # the CPU baseline evaluates the reference's FP32 formulas literally: at a
    # width that is not a power of two, 2 px / W - 1 rounds the pixel
    # coordinate, which moves small footprints' values by up to ~2e-4
    cpu = O.cpu_forward_f32(*arrays, size, cap).astype(np.float64)
    assert (np.abs(cpu - img)[m] / img[m]).max() <= 5e-4
    # same coordinate rounding in its gradients: 2e-3 of the |term| sums here
    ok, worst = O.grads_within_tolerance(list(O.cpu_backward_f32(g, *arrays, cap)), ref, rtol=2e-3)
    assert ok, worst


def test_heightfield_oracle_reproduces_golden():
    import torch

    from oracle import heightfield_oracle as HO
    z = np.load(os.path.join(HERE, "heightfield.npz"))
    R, n_side, S, cap = (int(v) for v in z["meta"])
    h = torch.from_numpy(z["h"])
    rays = HO.parallel_rays(n_side, 1.0)
    L, g, img, path = HO.loss_and_grad(h, rays, float(z["ior"]), float(z["sensor_z"]), (1, S, S), cap,
                                       torch.from_numpy(z["target"]), keep_rays=torch.from_numpy(z["safe"]))
    assert np.array_equal(path.ray.numpy(), z["ray"])
    assert np.allclose(path.Ep.detach().numpy(), z["Ep"], rtol=1e-12)
    assert np.allclose(img.numpy(), z["image"], rtol=1e-11, atol=1e-14)
    assert abs(L - float(z["loss"])) <= 1e-12 * abs(L)
    assert np.allclose(g.numpy(), z["grad_h"], rtol=1e-10, atol=1e-12 * np.abs(z["grad_h"]).max())


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["plain", "auto", "deterministic"])
@pytest.mark.parametrize("name", CASES)
def test_gpu_matches_golden(name, mode):
    import torch

    from shapefromcaustics_amd.PhotonDifferentialSplatting import pds_backward, pds_forward
    arrays, size, cap, g, img, ref = _case(name)
    t = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]
    got = pds_forward(*t, size, cap, mode=mode)[0].cpu().numpy().astype(np.float64)
    m = img > 1e-3 * img.max()
    assert (np.abs(got - img)[m] / img[m]).max() <= 1e-4
    assert np.abs(got - img)[~m].max(initial=0.0) <= 1e-7 * img.max()
    grads = pds_backward(torch.from_numpy(g).cuda(), *t, cap, mode=mode)
    ok, worst = O.grads_within_tolerance([x.cpu().numpy() for x in grads], ref, rtol=1e-3)
    assert ok, worst


@pytest.mark.gpu
def test_gpu_heightfield_matches_golden():
    import torch

    from shapefromcaustics_amd import heightfield as HF
    z = np.load(os.path.join(HERE, "heightfield.npz"))
    R, n_side, S, cap = (int(v) for v in z["meta"])
    h = torch.from_numpy(z["h"]).float().cuda()
    assert np.abs(HF.normals(h).cpu().numpy() - z["normals"]).max() < 1e-6
    sc, em = HF.Scene(R, float(z["sensor_z"]), (float(z["ior"]),)), HF.ParallelEmitter(n_side, 1.0)
    from shapefromcaustics_amd.PhotonDifferentialSplatting import _backward, _forward
    n = HF.normals(h)
    Ep, xp, Mp, cp, radius = HF.simulate(h, sc, em, n)
    alive = np.nonzero(cp[0].cpu().numpy() >= 0)[0]
    assert np.array_equal(alive, z["ray"])
    safe = z["safe"]
    ks = safe[z["ray"]]
    assert np.abs(Ep.cpu().numpy()[alive][ks] - z["Ep"][ks]).max() <= 1e-4 * np.abs(z["Ep"]).max()
    # image and dL/dh over the edge-safe photons (others get channel -1)
    cps = torch.where(torch.from_numpy(safe).cuda()[None], cp, torch.full_like(cp, -1))
    img, bins = _forward(Ep, xp, Mp, cps, radius, (1, S, S), cap, "auto", None, keep_bins=True)
    resid = img - torch.from_numpy(z["target"]).float().cuda()
    gE, gx, gM = _backward(resid, Ep, xp, Mp, cps, radius, cap, "auto", None, bins=bins)
    gh = HF.adjoint(h, sc, em, gE, gx, gM, n).cpu().numpy()
    ref = z["image"]
    mm = ref > 1e-3 * ref.max()
    assert (np.abs(img.cpu().numpy() - ref)[mm] / ref[mm]).max() <= 1e-4
    gr = z["grad_h"]
    assert np.abs(gh - gr).max() <= 1e-3 * np.abs(gr).max()
