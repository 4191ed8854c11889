"""The synthetic workloads map pixel-space footprints to the reference input
convention of model/caustics.py:224-239 (checked on the CPU)."""
import numpy as np

from shapefromcaustics_amd import workloads as WL


def test_pixel_space_mapping_semiaxes():
    # M (D e_k / (W/2, H/2)) must have unit length: the footprint's semi-axes are s1, s2 px
    p = WL.plain(1000, (1, 256, 128), 8, seed=3)
    rng = np.random.default_rng(0)
    C, H, W = p.size
    for i in rng.choice(p.n, 20, replace=False):
        M = p.M[:, :2, i].astype(np.float64)
        Dn = np.linalg.inv(M)  # columns: semi-axes in normalised units
        semi_px = np.hypot(Dn[0] * W / 2, Dn[1] * H / 2)
        # radius = 2 max(s) / W, so ceil(radius W / 2) = ceil(max s)
        assert abs(semi_px.max() - p.radius[i] * W / 2) < 1e-3 * semi_px.max()
        assert 0.5 - 1e-4 <= semi_px.min() and semi_px.max() <= 4 + 1e-4
    assert not p.M[:, 2].any()


def test_caustic_distribution_properties():
    p = WL.caustic(1 << 16, (3, 256, 256), 12, seed=2)
    assert np.array_equal(p.cp[0], np.arange(p.n) % 3)
    assert (p.xp >= -1).all() and (p.xp < 1).all()
    M = p.M[:, :2].transpose(2, 0, 1).astype(np.float64)
    cond = np.linalg.cond(M)
    assert cond.min() >= 1 - 1e-4 and cond.max() <= 10 + 1e-3
    # caustic lines concentrate photons: the busiest 1 % of 16x16 tiles hold far more than 1 %
    tx = ((p.xp[0] + 1) * 128 // 16).astype(int)
    ty = ((p.xp[1] + 1) * 128 // 16).astype(int)
    counts = np.bincount(ty * 16 + tx, minlength=256)
    assert np.sort(counts)[-3:].sum() > 0.05 * p.n


def test_batched_workload_layout():
    counts = [300, 1000, 50]
    p = WL.batched(3, 64, 8, counts=counts)
    assert p.size == (3, 64, 64) and p.n == sum(counts)
    assert np.array_equal(np.bincount(p.cp[0]), counts)


def test_seeded_reproducible():
    a, b = WL.caustic(1000, seed=5), WL.caustic(1000, seed=5)
    for x, y in zip(a.arrays(), b.arrays()):
        assert np.array_equal(x, y)
