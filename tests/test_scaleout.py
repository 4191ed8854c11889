"""Scale-out workload: the counter-based photon source (CPU twin and GPU),
photon regeneration in the backward, and the 12-channel splat at reduced
size against the oracle and the size-independent properties (deterministic
vs atomic agreement, the adjoint identity <g, F(E)> = sum E gE)."""
import numpy as np
import pytest
import torch

from oracle import gen_twin as GT
from oracle import pds_oracle as O
from shapefromcaustics_amd import generate as G

SIDE = 256


def test_twin_distribution():
    lines = G.caustic_lines(6, 1, SIDE, SIDE)
    E, xp, M, cp, radius = GT.caustic(6, 1, 1 << 16, SIDE, SIDE, lines)
    assert np.array_equal(cp[0], 3 + np.arange(1 << 16) % 3)
    assert abs(E.sum() - 1.0) < 1e-3                       # mean energy 1/n
    assert (xp >= -1).all() and (xp < 1).all()
    s = 2.0 * np.maximum(1.0, 1.0) / SIDE
    assert (radius > 0).all() and radius.max() <= 2 * 6.0 / SIDE * 1.0001
    # about 80 % of the photons sit within 4 long sigmas of their line centre
    assert np.isfinite(M).all() and (M[:, 2] == 0).all()
    # counter-based: photons do not depend on how many are drawn
    E2, _, _, _, _ = GT.caustic(6, 1, 1 << 10, SIDE, SIDE, lines)
    assert np.array_equal(E2 * (1 << 10), E[: 1 << 10] * (1 << 16))


@pytest.mark.gpu
def test_gpu_generator_matches_twin_and_regenerates():
    lines = G.caustic_lines(6, 2, SIDE, SIDE)
    dl = torch.from_numpy(lines).cuda()
    n = 1 << 16
    mk = lambda: (torch.empty(n, device="cuda"), torch.empty(2, n, device="cuda"), torch.empty(2, 3, n, device="cuda"),
                  torch.empty(1, n, dtype=torch.int64, device="cuda"), torch.empty(n, device="cuda"))
    a, b = mk(), mk()
    G.generate_caustic(6, 2, n, SIDE, SIDE, dl, a)
    G.generate_caustic(6, 2, n, SIDE, SIDE, dl, b)
    for x, y in zip(a, b):
        assert torch.equal(x, y)                        # regeneration is bit-exact
    ref = GT.caustic(6, 2, n, SIDE, SIDE, lines)
    got = [t.cpu().numpy() for t in a]
    assert np.array_equal(got[3], ref[3])
    for g_, r_ in zip((got[0], got[1], got[2], got[4]), (ref[0], ref[1], ref[2], ref[4])):
        assert np.allclose(g_, r_, rtol=2e-5, atol=2e-6 * np.abs(r_).max())


@pytest.mark.gpu
def test_scaleout_splat_with_regenerated_photons():
    from shapefromcaustics_amd.PhotonDifferentialSplatting import _backward, _forward
    so = G.ScaleOut(views=4, per_view=1 << 16, side=SIDE, cap=16)
    grad = torch.randn(so.size, device="cuda", generator=torch.Generator("cuda").manual_seed(3))
    imgs = {}
    for mode in ("auto", "deterministic"):
        ph = so.photons()
        img, bins = _forward(*ph, so.size, so.cap, mode, None, keep_bins=True)
        kept = _backward(grad, *ph, so.cap, mode, None, bins=bins)
        del ph                                          # the photons are not kept ...
        ph2 = so.photons()                              # ... but regenerated for the backward
        regen = _backward(grad, *ph2, so.cap, mode, None, bins=bins)
        for x, y in zip(kept, regen):
            assert torch.equal(x, y)
        imgs[mode] = (img, regen, ph2)
    ia, ga, ph = imgs["auto"]
    idet, gdet, _ = imgs["deterministic"]
    ref = O.forward_f64(*[t.cpu().numpy() for t in ph], so.size, so.cap)
    for img in (ia, idet):
        got = img.cpu().numpy().astype(np.float64)
        m = ref > 1e-3 * ref.max()
        assert (np.abs(got - ref)[m] / ref[m]).max() <= 1e-4
    # adjoint identity: F is linear in E, so <g, F(E)> = sum_i E_i dL/dE_i
    lhs = float((grad.double() * ia.double()).sum())
    rhs = float((ph[0].double() * ga[0].double()).sum())
    assert abs(lhs - rhs) <= 1e-4 * abs(lhs)
