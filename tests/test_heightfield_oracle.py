"""Pins of the FP64 height-field oracle (oracle/heightfield_oracle.py) on the CPU.

Each function of the oracle is checked against what the mathematics fixes:
exact Sobel derivatives of planes, analytic ray/plane intersections, Snell's
law and its invariants for rays entering the glass, a hand-evaluated case of
the reference's exit formula, the change-of-basis invariants, the identically
zero differential normal, a flat substrate that must leave vertical rays
vertical, and -- for the whole chain -- central finite differences of the
loss of the frozen path (trace results fixed) against the autograd gradient
that the reference's detached trace defines.
"""
import math

import numpy as np
import pytest
import torch

from oracle import heightfield_oracle as HO

F64 = torch.float64


def test_sobel_exact_on_planes():
    R = 9
    c = HO.grid_coords(R)
    a, b = 0.3, -0.7
    h = a * c[0] + b * c[1] + 0.2
    n = HO.sobel_normals(h, 2.0 / (R - 1))
    want = torch.tensor([-a, -b, 1.0], dtype=F64) / math.sqrt(a * a + b * b + 1)
    inner = n[:, 1:-1, 1:-1].reshape(3, -1).T
    assert torch.allclose(inner, want.expand_as(inner), atol=1e-14)
    # replicate padding halves the derivative across the border
    half = torch.tensor([-a / 2, -b, 1.0], dtype=F64)
    assert torch.allclose(n[:, 4, 0], half / half.norm(), atol=1e-14)


def test_scene_layout():
    R = 5
    h = torch.full((R, R), 0.3, dtype=F64)
    sc = HO.build_scene(h, -2.0)
    cells = (R - 1) ** 2
    assert sc.tris.shape[0] == 2 * cells * 2 + 4 * 2 * (R - 1) + 2
    gn = HO.geometric_normal(sc.tris)
    assert (gn[:2 * cells, 2] > 0.99).all()                       # top faces up
    assert (gn[2 * cells:4 * cells, 2] < -0.99).all()             # bottom faces down
    w = 4 * cells
    e = 2 * (R - 1)
    assert (gn[w:w + e, 1] < -0.99).all() and (gn[w + e:w + 2 * e, 1] > 0.99).all()      # front, back
    assert (gn[w + 2 * e:w + 3 * e, 0] < -0.99).all() and (gn[w + 3 * e:w + 4 * e, 0] > 0.99).all()  # left, right
    assert torch.allclose(sc.vnorm.norm(dim=-1), torch.ones(sc.vnorm.shape[:2], dtype=F64))
    # lower-left triangle of cell (r, c) = (1, 2) is (p[1,2], p[1,3], p[2,2])
    c = HO.grid_coords(R)
    t = 2 * (1 * (R - 1) + 2)
    assert torch.allclose(sc.tris[t, 0, :2], c[:, 1, 2]) and torch.allclose(sc.tris[t, 1, :2], c[:, 1, 3])
    assert torch.allclose(sc.tris[t, 2, :2], c[:, 2, 2])


def _bilinear_tri_height(h, x, y):
    """Height of the piecewise-linear top surface (two triangles per cell)."""
    R = h.shape[0]
    d = 2.0 / (R - 1)
    c, r = int((x + 1) // d), int((y + 1) // d)
    fx, fy = (x + 1) / d - c, (y + 1) / d - r
    if fx + fy < 1:
        return float(h[r, c] + fx * (h[r, c + 1] - h[r, c]) + fy * (h[r + 1, c] - h[r, c])), 2 * (r * (R - 1) + c)
    return float(h[r + 1, c + 1] + (1 - fx) * (h[r + 1, c] - h[r + 1, c + 1])
                 + (1 - fy) * (h[r, c + 1] - h[r + 1, c + 1])), 2 * (r * (R - 1) + c) + 1


def test_trace_against_analytic_intersections():
    R = 7
    h = torch.from_numpy(HO.gaussian_bumps(R, seed=3, n_bumps=4))
    sc = HO.build_scene(h, -3.0)
    rng = np.random.default_rng(0)
    xy = rng.uniform(-0.95, 0.95, (40, 2))
    o = torch.tensor(np.c_[xy, np.full(40, 2.0)], dtype=F64)
    d = torch.tensor([0.0, 0.0, -1.0], dtype=F64).expand(40, 3)
    t, b, obj, prim, tri = HO.trace_bruteforce(sc, o, d)
    for k in range(40):
        z, pr = _bilinear_tri_height(h, xy[k, 0], xy[k, 1])
        assert obj[k] == 0 and prim[k] == pr
        assert abs(float(t[k]) - (2.0 - z)) < 1e-12
        V = sc.tris[tri[k]]
        p = (1 - b[k, 0] - b[k, 1]) * V[0] + b[k, 0] * V[1] + b[k, 1] * V[2]
        assert torch.allclose(p, o[k] + t[k] * d[k], atol=1e-12)
    # from inside the glass straight down: the bottom (object 0) at z = 0, then the ground (object 1)
    o2 = torch.tensor([[0.1, -0.2, 0.05]], dtype=F64)
    t2, b2, obj2, prim2, _ = HO.trace_bruteforce(sc, o2, d[:1])
    assert obj2[0] == 0 and 2 * (R - 1) ** 2 <= prim2[0] < 4 * (R - 1) ** 2 and abs(float(t2[0]) - 0.05) < 1e-12
    o3 = torch.tensor([[0.1, -0.2, -0.5]], dtype=F64)
    t3, _, obj3, _, _ = HO.trace_bruteforce(sc, o3, d[:1])
    assert obj3[0] == 1 and abs(float(t3[0]) - 2.5) < 1e-12
    # t is parametric in a non-unit direction
    t4, _, _, _, _ = HO.trace_bruteforce(sc, o3, 4 * d[:1])
    assert abs(float(t4[0]) - 2.5 / 4) < 1e-12
    # t < 1e-3 is ignored; a ray going up misses: (-2, (-1, -1), -1, -1)
    t5, b5, obj5, prim5, _ = HO.trace_bruteforce(sc, torch.tensor([[0.1, 0.1, 3.0]], dtype=F64), -d[:1])
    assert float(t5[0]) == -2.0 and (b5[0] == -1).all() and obj5[0] == -1 and prim5[0] == -1
    t6, _, obj6, _, _ = HO.trace_bruteforce(sc, torch.tensor([[0.1, -0.2, 0.0005]], dtype=F64), d[:1])
    assert obj6[0] == 1  # the bottom at t = 5e-4 < 1e-3 is skipped


def _refr(d, n, n_in=1.5):
    one = torch.ones(d.shape[0], 1, dtype=F64)
    z = torch.zeros_like(d)
    return HO.refract_reflect(d, n, z, z, z, z, one, n_in * one)


def test_refraction_entering_obeys_snell():
    rng = np.random.default_rng(1)
    n = torch.tensor(rng.normal(size=(64, 3)), dtype=F64)
    n[:, 2] = n[:, 2].abs() + 0.5
    n = HO.normalize(n)
    d = torch.tensor(rng.normal(size=(64, 3)), dtype=F64)
    d[:, 2] = -(d[:, 2].abs() + 0.3)
    d = HO.normalize(d)
    keep = HO.dot(d, n) < -0.05
    d, n = d[keep], n[keep]
    out, _, _ = _refr(d, n)
    assert torch.allclose(out.norm(dim=-1), torch.ones(out.shape[0], dtype=F64), atol=1e-12)
    sin_i = torch.linalg.cross(d, n, dim=-1).norm(dim=-1)
    sin_t = torch.linalg.cross(out, n, dim=-1).norm(dim=-1)
    assert torch.allclose(sin_i, 1.5 * sin_t, atol=1e-12)
    assert HO.dot(torch.linalg.cross(d, n, dim=-1), out).abs().max() < 1e-12   # coplanar
    assert (HO.dot(out, n) < 0).all()                                           # transmitted


def test_refraction_exit_formula_is_the_references():
    # d = (s, 0, -c) leaving through the bottom (normal (0, 0, -1)): cosi = c > 0
    # flips the normal to +z and the indices to (1.5, 1) but keeps cosi, so the
    # result is 1.5 d - (1.5 c + sqrt(1 - 2.25 s^2)) e_z  (not unit length)
    th = 0.3
    s, c = math.sin(th), math.cos(th)
    d = torch.tensor([[s, 0.0, -c]], dtype=F64)
    n = torch.tensor([[0.0, 0.0, -1.0]], dtype=F64)
    out, _, _ = _refr(d, n)
    k = 1 - 2.25 * s * s
    want = torch.tensor([[1.5 * s, 0.0, -1.5 * c - 1.5 * c - math.sqrt(k)]], dtype=F64)
    assert torch.allclose(out, want, atol=1e-14)
    # beyond the critical angle the reflection branch d - 2 cosi N uses the
    # same unflipped cosi with the flipped N: (s, 0, -c) - 2c e_z = (s, 0, -3c)
    th = 0.9
    d = torch.tensor([[math.sin(th), 0.0, -math.cos(th)]], dtype=F64)
    out, _, _ = _refr(d, n)
    assert torch.allclose(out, torch.tensor([[math.sin(th), 0.0, -3 * math.cos(th)]], dtype=F64), atol=1e-14)
    # entering from outside (cosi < 0, no flip) reflection is a true mirror
    d = torch.tensor([[0.8, 0.0, 0.6]], dtype=F64)
    one = torch.ones(1, 1, dtype=F64)
    z = torch.zeros_like(d)
    out, _, _ = HO.refract_reflect(d, n, z, z, z, z, 1.5 * one, one)   # eta 1.5, sin 0.8: k < 0
    assert torch.allclose(out, torch.tensor([[0.8, 0.0, -0.6]], dtype=F64), atol=1e-14)


def test_igehy_direction_differential_matches_fd_when_entering():
    # d(theta) and N(theta) vary; Dtheta_out must be d/dtheta of the refracted direction
    def dirs(th):
        d = torch.tensor([[math.sin(th), 0.2 * th, -1.0]], dtype=F64)
        N = torch.tensor([[0.3 * math.cos(th), 0.1, 1.0]], dtype=F64)
        return d / d.norm(), N / N.norm()

    th, h = 0.2, 1e-6
    d, N = dirs(th)
    dp, Np = dirs(th + h)
    dm, Nm = dirs(th - h)
    Dd, dN = (dp - dm) / (2 * h), (Np - Nm) / (2 * h)
    one = torch.ones(1, 1, dtype=F64)
    z = torch.zeros_like(d)
    _, Dt, _ = HO.refract_reflect(d, N, Dd, z, dN, z, one, 1.5 * one)
    fd = (_refr(dp, Np)[0] - _refr(dm, Nm)[0]) / (2 * h)
    assert torch.allclose(Dt, fd, atol=1e-8)


def test_transfer_moves_differentials_onto_the_hit_plane():
    rng = np.random.default_rng(2)
    n = HO.normalize(torch.tensor(rng.normal(size=(16, 3)), dtype=F64))
    d = HO.normalize(torch.tensor(rng.normal(size=(16, 3)), dtype=F64))
    Du = torch.tensor(rng.normal(size=(16, 3)), dtype=F64)
    z = torch.zeros_like(Du)
    t = torch.ones(16, dtype=F64)
    Du2, _, _, _ = HO.transfer(n, d, t, Du, Du, z, z)
    assert HO.dot(n, Du2).abs().max() < 1e-12
    assert torch.linalg.cross(Du2 - Du, d, dim=-1).norm(dim=-1).max() < 1e-12


def test_differential_normal_is_identically_zero():
    R = 6
    h = torch.from_numpy(HO.gaussian_bumps(R, seed=4, n_bumps=3))
    sc = HO.build_scene(h, -1.0)
    V = sc.tris[:20]
    ns = HO.shading_normal(sc.vnorm[:20], torch.full((20, 2), 0.3, dtype=F64))
    x = torch.tensor(np.random.default_rng(3).normal(size=(20, 3)), dtype=F64)
    dn = HO.differential_normal_literal(V, ns, x)
    assert dn.abs().max() < 1e-14


def test_change_of_basis_invariants():
    rng = np.random.default_rng(5)
    Du = torch.tensor(rng.normal(size=(32, 3)), dtype=F64)
    Dv = torch.tensor(rng.normal(size=(32, 3)), dtype=F64)
    ns = torch.tensor([0.0, 0.0, 1.0], dtype=F64).expand(32, 3)
    Du[:, 2] = 0
    Dv[:, 2] = 0
    p = torch.zeros(32, 3, dtype=F64)
    Ep, xp, M, radius, good = HO.splat_inputs(p, Du, Dv, ns, torch.ones(32, dtype=F64))
    M = M.permute(2, 0, 1)  # [N, 2, 3]
    assert torch.allclose(torch.einsum("nij,nj->ni", M, Du), torch.tensor([2.0, 0.0], dtype=F64).expand(32, 2))
    assert torch.allclose(torch.einsum("nij,nj->ni", M, Dv), torch.tensor([0.0, 2.0], dtype=F64).expand(32, 2))
    assert torch.allclose(torch.einsum("nij,nj->ni", M, ns), torch.zeros(32, 2, dtype=F64))
    assert torch.allclose(Ep, 4 / torch.linalg.cross(Du, Dv, dim=-1).norm(dim=-1))


def test_flat_substrate_keeps_rays_vertical():
    R = 8
    h = torch.full((R, R), 0.2, dtype=F64)
    rays = HO.parallel_rays(8, 1.0)
    path = HO.simulate(h, rays, 1.5, -2.0)
    assert path.ray.numel() == 64
    o = rays.o[path.ray]
    assert torch.allclose(path.xp.T, o[:, :2], atol=1e-14)
    pitch = 2.0 / 8
    assert torch.allclose(path.radius, torch.full((64,), pitch, dtype=F64))
    assert torch.allclose(path.Ep, torch.full((64,), (1.0 / 64) * 4 / (2 * pitch) ** 2, dtype=F64))


@pytest.fixture(scope="module")
def small_problem():
    R = 10
    h = torch.from_numpy(HO.gaussian_bumps(R, seed=4, n_bumps=6, amp=(0.0, 0.06), sigma=(0.2, 0.5)))
    rays = HO.parallel_rays(24, 1.0)
    size, cap = (1, 32, 32), 6
    target = torch.from_numpy(np.random.default_rng(7).uniform(0, 1, size))
    L, g, img, path = HO.loss_and_grad(h, rays, 1.5, -1.5, size, cap, target)
    return dict(h=h, rays=rays, size=size, cap=cap, target=target, L=L, g=g, img=img, path=path)


def test_adjoint_chain_matches_fd_of_the_frozen_path(small_problem):
    P = small_problem
    assert P["path"].ray.numel() > 300 and P["img"].max() > 0
    frozen = P["path"].hits
    rng = np.random.default_rng(11)
    R = P["h"].shape[0]
    idx = [(int(a), int(b)) for a, b in rng.integers(0, R, (12, 2))]
    g = P["g"]
    scale = g.abs().max()
    for (r, c) in idx:
        step = 1e-6
        hp, hm = P["h"].clone(), P["h"].clone()
        hp[r, c] += step
        hm[r, c] -= step
        Lp = HO.frozen_loss(hp, P["rays"], 1.5, -1.5, P["size"], P["cap"], P["target"], frozen)
        Lm = HO.frozen_loss(hm, P["rays"], 1.5, -1.5, P["size"], P["cap"], P["target"], frozen)
        fd = (Lp - Lm) / (2 * step)
        assert abs(fd - float(g[r, c])) <= 1e-5 * float(scale) + 1e-6 * abs(fd), (r, c, fd, float(g[r, c]))


def test_planted_sign_error_in_transfer_is_caught(small_problem, monkeypatch):
    P = small_problem
    orig = HO.transfer

    def bad(n, d, t, Du, Dv, Dth, Dph):
        a, b, c, e = orig(n, d, t, Du, Dv, Dth, Dph)
        return 2 * Du - a, b, c, e  # wrong sign of the projection term
    monkeypatch.setattr(HO, "transfer", bad)
    L2, g2, _, _ = HO.loss_and_grad(P["h"], P["rays"], 1.5, -1.5, P["size"], P["cap"], P["target"])
    assert (g2 - P["g"]).abs().max() > 1e-3 * P["g"].abs().max()


def test_landweber_update_pieces():
    x = torch.tensor([-0.5, -0.01, 0.0, 0.02, 0.2, 0.6], dtype=F64)
    y = HO.ext_shrink(x, 0.05, 0.0, 0.3)
    assert torch.allclose(y, torch.tensor([0.0, 0.0, 0.0, 0.0, 0.15, 0.3], dtype=F64))
    d = torch.full((6, 6), 0.1, dtype=F64)
    m = torch.ones(6, 6, dtype=F64)
    assert torch.equal(HO.vcons_heur(d, 1.0, 10.0, m, 0.1, 1), d)          # inside the volume band
    lo = HO.vcons_heur(d, 10.0, 20.0, m, 0.1, 1)                              # below: grow
    assert (lo[2:4, 2:4] > d[2:4, 2:4]).all()
    hi = HO.vcons_heur(d, 0.5, 3.0, m, 0.1, 1)                                # above: shrink
    assert (hi[2:4, 2:4] < d[2:4, 2:4]).all()
