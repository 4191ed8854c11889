"""GPU: the height-field kernels (include/hf.h, through ctypes) against the FP64 oracle.

Tolerances: normals 1e-6 absolute; trace t 1e-5 relative, barycentrics 1e-4,
object / primitive exact (rays kept 1e-3 away from triangle edges, where FP32
and FP64 may pick different but equally valid triangles); splat inputs 1e-4
relative; the image within 1e-4 relative on pixels above 1e-3 max; dL/dh
within 1e-3 of max |dL/dh| per entry and 1e-3 in relative L2 norm.
"""
import math

import numpy as np
import pytest
import torch

from oracle import heightfield_oracle as HO
from oracle import pds_oracle as O
from shapefromcaustics_amd import heightfield as HF

pytestmark = pytest.mark.gpu
F64 = torch.float64


def _h(R, seed, **kw):
    return torch.from_numpy(HO.gaussian_bumps(R, seed=seed, **kw))


def test_normals_match_sobel():
    h = _h(33, 1)
    n = HF.normals(h.float().cuda()).cpu().double()
    ref = HO.sobel_normals(h.float().double(), 2.0 / 32).permute(1, 2, 0)
    assert (n - ref).abs().max() < 1e-6


def test_trace_contract():
    R, sz = 17, -2.0
    h64 = _h(R, 2, n_bumps=6)
    h = h64.float()
    sc = HO.build_scene(h.double(), sz)
    rng = np.random.default_rng(0)
    n = 600
    o = np.zeros((n, 3))
    d = np.zeros((n, 3))
    o[:200, :2] = rng.uniform(-0.98, 0.98, (200, 2)); o[:200, 2] = 1.5; d[:200] = [0, 0, -1]     # from above
    o[200:400, :2] = rng.uniform(-0.9, 0.9, (200, 2)); o[200:400, 2] = rng.uniform(0.02, 0.15, 200)
    d[200:400] = np.c_[rng.normal(0, 0.4, (200, 2)), -np.ones(200)]                            # inside, downwards
    o[400:500, :2] = rng.uniform(-0.5, 0.5, (100, 2)); o[400:500, 2] = 0.1
    d[400:500] = np.c_[rng.choice([-1, 1], (100, 2)) * rng.uniform(2, 4, (100, 2)), rng.uniform(-0.3, 0.3, 100)]  # walls
    o[500:550, :2] = rng.uniform(-0.9, 0.9, (50, 2)); o[500:550, 2] = 2.0; d[500:550] = [0.1, 0.0, 1.0]  # up: miss
    o[550:, :2] = rng.uniform(-0.8, 0.8, (50, 2)); o[550:, 2] = -0.3
    d[550:] = np.c_[rng.normal(0, 0.05, (50, 2)), -np.ones(50)]                                # below: ground
    d[:500] /= np.linalg.norm(d[:500], axis=1, keepdims=True)
    o32, d32 = torch.tensor(o, dtype=torch.float32), torch.tensor(d, dtype=torch.float32)
    t, uv, obj, prim = [a.cpu() for a in HF.trace_rays(h.cuda(), HF.Scene(R, sz), o32.cuda(), d32.cuda())]
    rt, rb, robj, rprim, rtri = HO.trace_bruteforce(sc, o32.double(), d32.double())
    bmin = torch.minimum(torch.minimum(rb[:, 0], rb[:, 1]), 1 - rb[:, 0] - rb[:, 1])
    safe = (robj < 0) | (bmin > 1e-3)
    assert safe.sum() > 0.9 * n
    assert torch.equal(obj[safe], robj[safe]) and torch.equal(prim[safe], rprim[safe])
    hit = safe & (robj >= 0)
    assert (t[hit].double() - rt[hit]).abs().max() <= 1e-5 * rt[hit].abs().max()
    assert (uv[hit].double() - rb[hit]).abs().max() <= 1e-4
    miss = robj < 0
    assert (t[miss] == -2).all() and (uv[miss] == -1).all() and (prim[miss] == -1).all()
    assert miss.sum() >= 50 and (robj == 1).sum() >= 40 and ((robj == 0) & (rprim >= 4 * (R - 1) ** 2)).any()


@pytest.fixture(scope="module")
def problem():
    R, n_side, sz = 24, 48, -1.5
    h = _h(R, 4, n_bumps=8, amp=(0.0, 0.08), sigma=(0.15, 0.5)).float()
    size, cap = (1, 64, 64), 8
    target = torch.from_numpy(np.random.default_rng(5).uniform(0, 2, size))
    rays = HO.parallel_rays(n_side, 1.0)
    with torch.no_grad():
        base = HO.simulate(h.double(), rays, 1.5, sz)
    safe = HO.edge_safe_rays(base, n_side * n_side)
    L, g, img, path = HO.loss_and_grad(h.double(), rays, 1.5, sz, size, cap, target, keep_rays=safe)
    Lall, gall, imgall, _ = HO.loss_and_grad(h.double(), rays, 1.5, sz, size, cap, target)
    return dict(R=R, n_side=n_side, sz=sz, h=h, size=size, cap=cap, target=target, L=L, g=g, img=img, path=path,
                safe=safe, Lall=Lall, gall=gall, imgall=imgall)


def test_simulate_matches_oracle_photon_by_photon(problem):
    P = problem
    Ep, xp, Mp, cp, radius = [a.cpu() for a in HF.simulate(P["h"].cuda(), HF.Scene(P["R"], P["sz"]),
                                                           HF.ParallelEmitter(P["n_side"], 1.0))]
    path = P["path"]
    alive = torch.zeros(cp.shape[1], dtype=torch.bool)
    alive[path.ray] = True
    assert torch.equal(cp[0] >= 0, alive)
    keep = P["safe"][path.ray]
    print(f"{int(keep.sum())} of {path.ray.numel()} photons away from triangle edges")
    assert keep.float().mean() > 0.97
    r = path.ray[keep]
    rel = lambda a, b: float((a.double() - b.detach()).abs().max() / b.detach().abs().max())
    errs = {"Ep": rel(Ep[r], path.Ep[keep]), "Mp": rel(Mp[:, :, r], path.Mp[:, :, keep]),
            "radius": rel(radius[r], path.radius[keep])}
    xerr = float((xp[:, r].double() - path.xp[:, keep].detach()).abs().max())
    print("splat-input relative errors", errs, "xp abs", xerr)
    assert max(errs.values()) <= 1e-4 and xerr <= 1e-5


def _gpu_loss_grad(P, safe_only):
    from shapefromcaustics_amd.PhotonDifferentialSplatting import _backward, _forward
    sc, em = HF.Scene(P["R"], P["sz"]), HF.ParallelEmitter(P["n_side"], 1.0)
    h = P["h"].cuda()
    n = HF.normals(h)
    Ep, xp, Mp, cp, radius = HF.simulate(h, sc, em, n)
    if safe_only:
        cp = torch.where(P["safe"].cuda()[None], cp, torch.full_like(cp, -1))
    img, bins = _forward(Ep, xp, Mp, cp, radius, P["size"], P["cap"], "auto", None, keep_bins=True)
    tgt = P["target"].float().cuda()
    resid = img - tgt
    gE, gx, gM = _backward(resid, Ep, xp, Mp, cp, radius, P["cap"], "auto", None, bins=bins)
    gh = HF.adjoint(h, sc, em, gE, gx, gM, n)
    return float(0.5 * (resid.double() ** 2).sum()), gh.cpu().double(), img.cpu().double()


def test_image_and_adjoint_match_oracle(problem):
    P = problem
    L, g, img = _gpu_loss_grad(P, safe_only=True)
    ref = P["img"].detach().numpy()
    got = img.numpy()
    m = ref > 1e-3 * ref.max()
    rel = (np.abs(got - ref)[m] / ref[m]).max()
    gr = P["g"]
    err = float((g - gr).abs().max() / gr.abs().max())
    l2 = float((g - gr).norm() / gr.norm())
    print(f"image max rel err {rel:.3e}; loss {L:.6f} vs {P['L']:.6f}; dL/dh max err / max {err:.3e}, "
          f"relative L2 {l2:.3e}")
    assert rel <= 1e-4
    assert abs(L - P["L"]) <= 1e-5 * P["L"]
    assert err <= 1e-3 and l2 <= 1e-3


def test_autograd_function_and_all_photons(problem):
    # every photon (edge ones included): the autograd entry point agrees with
    # the explicit pipeline, and with the oracle to the accuracy that the
    # near-edge triangle choice allows
    P = problem
    h = P["h"].cuda().requires_grad_(True)
    img = HF.CausticImage.apply(h, HF.Scene(P["R"], P["sz"]), HF.ParallelEmitter(P["n_side"], 1.0), P["size"],
                                P["cap"])
    loss = 0.5 * ((img - P["target"].float().cuda()) ** 2).sum()
    loss.backward()
    L, g, img2 = _gpu_loss_grad(P, safe_only=False)
    assert torch.allclose(img.detach().cpu().double(), img2, rtol=1e-5, atol=1e-6 * float(img2.abs().max()))
    assert (h.grad.cpu().double() - g).abs().max() <= 1e-4 * g.abs().max()
    ref = P["imgall"].detach()
    l2 = float((img2 - ref).norm() / ref.norm())
    gl2 = float((g - P["gall"]).norm() / P["gall"].norm())
    print(f"all photons: image relative L2 {l2:.3e}, dL/dh relative L2 {gl2:.3e}")
    assert l2 <= 1e-2 and gl2 <= 5e-2


def test_landweber_update_and_residual():
    R = 40
    rng = np.random.default_rng(9)
    d = torch.tensor(rng.uniform(-0.05, 0.3, (R, R)), dtype=torch.float32)
    g = torch.tensor(rng.normal(0, 0.2, (R, R)), dtype=torch.float32)
    mask = torch.tensor(rng.uniform(size=(R, R)) > 0.2, dtype=torch.float32)
    for d1, d2 in [(-1e9, 1e9), (500.0, 900.0), (1.0, 2.0)]:
        p = HF.LandweberParams(d1=d1, d2=d2)
        want = HO.landweber_update(d.double(), g.double(), mask.double(), p.tau, p.alpha, p.lower, p.upper, d1, d2,
                                   p.gamma, p.volume_radius)
        dd = d.clone().cuda()
        HF.landweber_update(dd, g.cuda(), mask.cuda(), p)
        assert (dd.cpu().double() - want).abs().max() < 1e-5, (d1, d2)
    img = torch.tensor(rng.uniform(size=(3, 50, 70)), dtype=torch.float32).cuda()
    tgt = torch.tensor(rng.uniform(size=(3, 50, 70)), dtype=torch.float32).cuda()
    resid, loss = HF.residual(img, tgt)
    assert torch.equal(resid, img - tgt)
    want = 0.5 * ((img.double() - tgt.double()) ** 2).sum()
    assert abs(float(loss) - float(want)) <= 1e-5 * float(want)


def test_landweber_step_runs_end_to_end(problem):
    P = problem
    d = (P["h"] - 0.2).cuda().contiguous()
    d0 = d.clone()
    mask = torch.ones_like(d)
    loss, img, gh = HF.landweber_step(d, 0.2, P["target"].float().cuda(), HF.Scene(P["R"], P["sz"]),
                                      HF.ParallelEmitter(P["n_side"], 1.0), P["size"], P["cap"], mask,
                                      HF.LandweberParams())
    torch.cuda.synchronize()
    L, g, _ = _gpu_loss_grad(P, safe_only=False)
    assert abs(float(loss) - L) <= 1e-5 * L
    assert (gh.cpu().double() - g).abs().max() <= 1e-4 * g.abs().max()
    want = HO.landweber_update(d0.cpu().double(), gh.cpu().double(), mask.cpu().double(), 0.1, 2e-3, 0.0, 0.3,
                               -math.inf, math.inf, 0.1, 2)
    assert (d.cpu().double() - want).abs().max() < 1e-5


@pytest.mark.parametrize("light", [(0.1, -0.2, 3.0), (-0.5, 0.4, 2.0)])
def test_point_light_emitter_and_adjoint(light):
    # model/caustics.py:26-70: rays from a point light with direction
    # differentials (Du = Dv = 0); their transfer, refraction and adjoint
    R, n_side, sz, size, cap = 16, 40, -1.5, (1, 64, 64), 8
    h = _h(R, 9, n_bumps=6, amp=(0.0, 0.06), sigma=(0.2, 0.5)).float()
    em = HF.PointLight(n_side, light, 0.2, flux=1e-3, smoothing=1.0, seed=7)
    rays = HO.point_light_rays(n_side, 1, light, 0.2, em.diff_scale(1), em.flux, em.seed)
    with torch.no_grad():
        base = HO.simulate(h.double(), rays, 1.5, sz)
    safe = HO.edge_safe_rays(base, n_side * n_side)
    target = torch.from_numpy(np.random.default_rng(3).uniform(0, 1e-3, size))
    L, g, img, path = HO.loss_and_grad(h.double(), rays, 1.5, sz, size, cap, target, keep_rays=safe)
    assert path.ray.numel() > 1000 and (path.radius > 0).all()
    P = dict(R=R, n_side=n_side, sz=sz, h=h, size=size, cap=cap, target=target, safe=safe)
    from shapefromcaustics_amd.PhotonDifferentialSplatting import _backward, _forward
    sc = HF.Scene(R, sz)
    hc = h.cuda()
    n = HF.normals(hc)
    Ep, xp, Mp, cp, radius = HF.simulate(hc, sc, em, n)
    alive = torch.zeros(n_side * n_side, dtype=torch.bool)
    alive[base.ray] = True
    assert torch.equal((cp[0] >= 0).cpu(), alive)
    keep = safe[path.ray]
    r = path.ray[keep]
    rel = lambda a, b: float((a.cpu().double() - b.detach()).abs().max() / b.detach().abs().max())
    assert rel(Ep[r], path.Ep[keep]) <= 1e-4 and rel(Mp[:, :, r], path.Mp[:, :, keep]) <= 1e-4
    assert rel(radius[r], path.radius[keep]) <= 1e-4
    cps = torch.where(safe.cuda()[None], cp, torch.full_like(cp, -1))
    im, bins = _forward(Ep, xp, Mp, cps, radius, size, cap, "auto", None, keep_bins=True)
    resid = im - target.float().cuda()
    gE, gx, gM = _backward(resid, Ep, xp, Mp, cps, radius, cap, "auto", None, bins=bins)
    gh = HF.adjoint(hc, sc, em, gE, gx, gM, n).cpu().double()
    got = im.cpu().double().numpy()
    # the splat of these very inputs: the FP64 oracle on the GPU's FP32 splat inputs
    c_ref = O.forward_f64(*[t.cpu().numpy() for t in (Ep, xp, Mp, cps, radius)], size, cap)
    mc = c_ref > 1e-3 * c_ref.max()
    assert (np.abs(got - c_ref)[mc] / c_ref[mc]).max() <= 1e-4
    # the whole FP32 chain against the FP64 one: the faint caustic edge (pixels
    # near 1e-3 max, footprint rims where K = (1 - q)^2 is small) moves by
    # ~1.5e-4 under FP32 rounding of the photon positions and differentials
    ref = img.numpy()
    m = ref > 1e-3 * ref.max()
    assert (np.abs(got - ref)[m] / ref[m]).max() <= 5e-4
    assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)
    err = float((gh - g).abs().max() / g.abs().max())
    print(f"point light {light}: dL/dh max err / max {err:.3e}")
    assert err <= 1e-3


@pytest.mark.parametrize("case", ["parallel", "point", "grazing", "steep"])
def test_adjoint_replaying_recorded_hits_matches_tracing(case):
    # hf_simulate_hits_f32 records each photon's first hits; hf_adjoint_hits_f32
    # replays them instead of tracing: the same operations on the same values,
    # so the splat inputs equal the tracing pair's bit for bit and every
    # photon's vertex-gradient terms too; dL/dh differs only by the order of
    # the float atomic sums, which varies from run to run in either adjoint
    # (paths with more bounces than recorded slots are traced again)
    from shapefromcaustics_amd.PhotonDifferentialSplatting import _backward, _forward
    R, n_side, sz, size, cap = 24, 48, -1.5, (1, 64, 64), 8
    h = _h(R, 4, n_bumps=8, amp=(0.0, 0.12), sigma=(0.1, 0.4)).float().cuda()
    sc = HF.Scene(R, sz)
    if case == "steep":  # steep bumps: paths of 4 hits and paths longer than the record (traced again)
        n_side = 64
        h = _h(R, 4, n_bumps=12, amp=(0.4, 0.9), sigma=(0.04, 0.1)).float().cuda()
    if case in ("parallel", "steep"):
        em = HF.ParallelEmitter(n_side, 1.0)
    elif case == "point":
        em = HF.PointLight(n_side, (0.1, -0.2, 3.0), 0.2, flux=1e-3, smoothing=1.0, seed=7)
    else:  # low light far to the side: grazing rays, wall hits and internal reflections
        em = HF.PointLight(n_side, (-2.5, 0.3, 0.6), 0.1, flux=1e-3, smoothing=1.0, seed=3)
    n = HF.normals(h)
    a = HF.simulate(h, sc, em, n)
    b = HF.simulate(h, sc, em, n, keep_hits=True)
    for x, y in zip(a, b[:5]):
        assert torch.equal(x, y)
    hits = b[5]
    N = a[0].numel()
    counts = HF.hit_counts(hits, N).cpu()
    alive = (a[3][0] >= 0).cpu()
    recorded = ((counts > 0) & (counts <= 4)) | (counts == HF.HITS_RETRACE)
    assert torch.equal(recorded, alive)
    print(f"{case}: {int(alive.sum())} photons reach the sensor; hits recorded per photon:",
          {int(k): int((counts == k).sum()) for k in torch.unique(counts[alive])})
    if case == "steep":
        assert (counts == 4).any() and (counts == 255).any()
    img, bins = _forward(*a, size, cap, "auto", None, keep_bins=True)
    g = torch.from_numpy(np.random.default_rng(1).standard_normal(size).astype(np.float32)).cuda()
    gE, gx, gM = _backward(g, *a, cap, "auto", None, bins=bins)
    g0 = HF.adjoint(h, sc, em, gE, gx, gM, n)
    g0b = HF.adjoint(h, sc, em, gE, gx, gM, n)
    g1 = HF.adjoint(h, sc, em, gE, gx, gM, n, hits=hits)
    scale = float(g0.abs().max())
    assert scale > 0
    noise = float((g0 - g0b).abs().max()) / scale
    diff = float((g0 - g1).abs().max()) / scale
    print(f"{case}: run-to-run {noise:.2e}, replay vs trace {diff:.2e} (of max |dL/dh|)")
    assert diff <= 1e-5


def test_hit_buffer_keeps_the_fields_height_range():
    # the simulate stores max h and min h after the hit counts; the adjoint
    # replaying the hits takes the tracer's slab from there
    R, n_side = 33, 40
    h = (_h(R, 12, n_bumps=6, amp=(0.05, 0.2), sigma=(0.1, 0.3)) + 0.1).float().cuda()
    sc, em = HF.Scene(R, -4.0, (1.5,)), HF.ParallelEmitter(n_side, 2.0, 1.0)
    hits = HF.simulate(h, sc, em, HF.normals(h), keep_hits=True)[5]
    N = n_side * n_side
    off = ((16 * 4 * N + 255) // 256 * 256 + N + 255) // 256 * 256
    stored = hits[off:off + 8].cpu().view(torch.float32)
    assert float(stored[0]) == float(h.max()) and float(stored[1]) == float(h.min())


def _hit_numbers(path, n):
    """Hits each ray took in an oracle path (the trace results of every bounce)."""
    nh = torch.zeros(n, dtype=torch.int64)
    for ray, t, b, tri in path.hits:
        nh[ray[t >= 0]] += 1
    return nh


def test_long_paths_total_internal_reflection_and_bounce_cap():
    # a dense field of steep narrow glass pillars lit by a grazing point light:
    # rays cross several pillars, are totally internally reflected inside them
    # (caustics.py:205-210), take paths of more than 4 hits (the adjoint traces
    # them again) and more than 8, and a few are still bouncing after
    # MAX_BOUNCES = 16 hits -- dropped and counted by both the GPU and the oracle.
    #
    # Such paths focus light strongly, so some photons' splat inputs are
    # ill-conditioned: perturbing the inputs by one float32 rounding (2^-24
    # relative on h and on the ray directions; the largest change over three
    # random perturbations) moves the FP64 oracle's own result by up to ~2e-3
    # there.  Photon by photon the GPU must agree within 1e-4 relative plus
    # 256 times that sensitivity: a long float32 path rounds ~100 times per
    # hit, and 256 lies between the random-walk (~sqrt(n)) and the worst-case
    # (~n) growth of n such roundings.  The image and dL/dh are compared on the
    # edge-safe photons whose sensitivity stays below 5e-6, the photons for
    # which float32 can be held near the 1e-4 tolerance at all.
    from shapefromcaustics_amd.PhotonDifferentialSplatting import _backward, _forward
    R, n_side, sz, size, cap = 48, 96, -0.3, (1, 64, 64), 8
    light, plane_z = (-4.0, -0.5, 0.8), 0.25
    h = _h(R, 4, n_bumps=200, amp=(0.1, 0.4), sigma=(0.025, 0.04), centre=(-0.9, 0.9)).float()
    em = HF.PointLight(n_side, light, plane_z, flux=1e-3, smoothing=1.0, seed=3)
    rays = HO.point_light_rays(n_side, 1, light, plane_z, em.diff_scale(1), em.flux, em.seed)
    N = n_side * n_side
    h64 = h.double()
    perts = []
    with torch.no_grad():
        base = HO.simulate(h64, rays, 1.5, sz)
        for seed in range(3):
            rng = np.random.default_rng(seed)
            hp = h64 * (1 + torch.from_numpy(rng.uniform(-1, 1, h64.shape)) * 2.0 ** -24)
            rp = HO.Rays(*[getattr(rays, f) for f in rays.__dataclass_fields__])
            rp.d = HO.normalize(rays.d * (1 + torch.from_numpy(rng.uniform(-1, 1, rays.d.shape)) * 2.0 ** -24))
            perts.append(HO.simulate(hp, rp, 1.5, sz))
    assert HO.MAX_BOUNCES == HF.MAX_BOUNCES == 16
    nh = _hit_numbers(base, N)
    safe = HO.edge_safe_rays(base, N)
    for pt in perts:
        safe &= HO.edge_safe_rays(pt, N)
    alive = torch.zeros(N, dtype=torch.bool)
    alive[base.ray] = True
    capped = torch.zeros(N, dtype=torch.bool)
    capped[base.capped] = True
    print(f"oracle: {int(alive.sum())} of {N} reach the sensor; edge-safe paths of 5-8 hits "
          f"{int((safe & alive & (nh > 4) & (nh <= 8)).sum())}, of more than 8 "
          f"{int((safe & alive & (nh > 8)).sum())} (longest {int(nh[alive].max())}); "
          f"{int((safe & alive & base.tir).sum())} edge-safe reached rays took a TIR; {int(capped.sum())} capped "
          f"({int((safe & capped).sum())} edge-safe)")
    assert int((safe & capped).sum()) >= 1

    sc = HF.Scene(R, sz)
    hc = h.cuda()
    n = HF.normals(hc)
    Ep, xp, Mp, cp, radius, hits = HF.simulate(hc, sc, em, n, keep_hits=True)
    st = HF.path_stats(hits, sc, em)
    counts = HF.hit_counts(hits, N).cpu().long()
    g_alive = (cp[0] >= 0).cpu()
    print("GPU path outcomes:", st)
    assert sum(st.values()) == N and st["reached"] == int(g_alive.sum())
    assert st["capped"] == int((counts == HF.HITS_CAPPED).sum()) and st["capped"] >= 1
    assert st["degenerate"] == int((counts == HF.HITS_DEGENERATE).sum())
    assert st["missed"] == int((counts == 0).sum())
    # on rays whose every hit keeps away from triangle edges, the same paths
    assert torch.equal(g_alive[safe], alive[safe])
    assert torch.equal((counts == HF.HITS_CAPPED)[safe], capped[safe])
    ra = safe & alive
    want = torch.where(nh <= 4, nh, torch.full_like(nh, HF.HITS_RETRACE))
    assert torch.equal(counts[ra], want[ra])
    # splat inputs photon by photon, against the oracle and its sensitivity
    kb = {int(r): k for k, r in enumerate(base.ray.tolist())}
    kps = [{int(r): k for k, r in enumerate(pt.ray.tolist())} for pt in perts]
    rr = torch.tensor([int(r) for r in torch.nonzero(ra)[:, 0] if all(int(r) in kp for kp in kps)])
    assert rr.numel() >= 0.99 * int(ra.sum())
    ib = torch.tensor([kb[int(r)] for r in rr])
    ips = [torch.tensor([kp[int(r)] for r in rr]) for kp in kps]

    def fields(path, idx):  # per photon: Ep, radius, Mp (6), xp + 1 (positions compared on [0, 2])
        return [path.Ep[idx][None], path.radius[idx][None], path.Mp[:, :, idx].reshape(6, -1), path.xp[:, idx] + 1.0]

    got_f = [Ep[rr][None], radius[rr][None], Mp[:, :, rr].reshape(6, -1), xp[:, rr] + 1.0]
    sens = torch.zeros(rr.numel(), dtype=F64)
    errs = []
    for k, ref in enumerate(fields(base, ib)):
        ref = ref.detach()
        scale = ref.abs().max(0)[0]                          # per photon
        dev = torch.stack([(fields(pt, ip)[k].detach() - ref).abs().max(0)[0] for pt, ip in zip(perts, ips)])
        dev = dev.max(0)[0] / scale
        errs.append(((got_f[k].cpu().double() - ref).abs().max(0)[0] / scale, dev))
        sens = torch.maximum(sens, dev)
    ratio = torch.stack([e / (1e-4 + 256 * d) for e, d in errs]).max(0)[0]
    r64 = torch.stack([e / (1e-4 + 64 * d) for e, d in errs]).max(0)[0]
    print(f"splat inputs: worst error / (1e-4 + 256 sensitivity) = {float(ratio.max()):.3f}; error / (1e-4 + 64 "
          f"sensitivity): 99.9 % {float(r64.quantile(0.999)):.3f}, max {float(r64.max()):.3f}; sensitivity "
          "quantiles " + ", ".join(f"{q}: {float(sens.quantile(q)):.1e}" for q in (0.5, 0.9, 0.99, 1.0)))
    assert float(ratio.max()) <= 1.0
    well = torch.zeros(N, dtype=torch.bool)
    well[rr[sens <= 5e-6]] = True
    print(f"{int(well.sum())} well-conditioned edge-safe photons: {int((well & (nh > 8)).sum())} with more than 8 "
          f"hits, {int((well & (nh > 4)).sum())} with more than 4, {int((well & base.tir).sum())} with a TIR")
    assert int((well & (nh > 8)).sum()) >= 10 and int((well & base.tir & (nh > 4)).sum()) >= 50
    # image and dL/dh of the well-conditioned edge-safe photons against the FP64 oracle
    target = torch.from_numpy(np.random.default_rng(8).uniform(0, 1e-3, size))
    L, g, img, _ = HO.loss_and_grad(h64, rays, 1.5, sz, size, cap, target, keep_rays=well)
    cps = torch.where(well.cuda()[None], cp, torch.full_like(cp, -1))
    im, bins = _forward(Ep, xp, Mp, cps, radius, size, cap, "auto", None, keep_bins=True)
    resid = im - target.float().cuda()
    gE, gx, gM = _backward(resid, Ep, xp, Mp, cps, radius, cap, "auto", None, bins=bins)
    gh = HF.adjoint(hc, sc, em, gE, gx, gM, n, hits=hits).cpu().double()
    gh_trace = HF.adjoint(hc, sc, em, gE, gx, gM, n).cpu().double()
    got, ref = im.cpu().double().numpy(), img.numpy()
    # the splat of these very inputs: the FP64 oracle on the GPU's FP32 splat inputs
    c_ref = O.forward_f64(*[t.cpu().numpy() for t in (Ep, xp, Mp, cps, radius)], size, cap)
    mc = c_ref > 1e-3 * c_ref.max()
    assert (np.abs(got - c_ref)[mc] / c_ref[mc]).max() <= 1e-4
    m = ref > 1e-3 * ref.max()
    relp = np.abs(got - ref)[m] / ref[m]
    err_img = float(relp.max())
    l2_img = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    err = float((gh - g).abs().max() / g.abs().max())
    err_t = float((gh_trace - g).abs().max() / g.abs().max())
    print(f"image max rel err {err_img:.2e} (quantiles 0.5 {np.quantile(relp, 0.5):.1e}, 0.9 {np.quantile(relp, 0.9):.1e},"
          f" 0.99 {np.quantile(relp, 0.99):.1e}), relative L2 {l2_img:.2e}; dL/dh max err / max {err:.2e} (replaying "
          f"hits), {err_t:.2e} (tracing)")
    # the splat itself is held to 1e-4 above (c_ref); the whole float32 chain
    # of these focusing paths moves footprint rims of very bright photons, so
    # the image is compared in relative L2 (measured 4e-4) and dL/dh to 1e-3
    assert l2_img <= 1e-3
    assert err <= 1e-3 and err_t <= 1e-3
