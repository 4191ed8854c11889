"""FP64 CPU oracle of the height-field caustic pipeline around the PDS splat.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP height-field
kernels (Sobel normals, ray trace, refraction with photon differentials, splat
inputs, adjoint).  Only ``tests/`` import it; the product package never does.

It restates, in torch float64 on the CPU, the reference's differentiable
optics for one height-field substrate above a ground plane (the scene of
``model/renderable_object.py:81-193`` with ``update_from_height_field``
``:249-308``) and the bounce loop of ``compute_recursive_refraction``
(``model/caustics.py:242-291``).  Autograd of this restatement, with the hit
distance, barycentrics and triangle index detached exactly as the reference's
OptiX trace returns them (plain tensors, ``PyOptix/RayTrace.cpp:233-291``), is
the reference gradient dL/dh.  ``trace_bruteforce`` tests every triangle of the
scene (Moeller-Trumbore), so it pins the GPU tracer independently of its grid
traversal.

Restated behaviours worth knowing (both are the reference's, kept as is):

* ``differential_normal`` (``model/renderable_object.py:391-401``) forms
  ``(sum_v L_v . x) n_s`` and then projects it orthogonally to ``n_s``: the
  result is identically zero (its FP32 value in the reference is rounding
  noise).  ``differential_normal_literal`` evaluates the literal formula so a
  test can show that.
* ``refract_reflect_differentials`` (``model/caustics.py:152-212``) flips the
  normal and the indices for rays leaving the denser medium but keeps
  ``cosi = d . n`` unflipped, so ``coso = eta cosi + sqrt(k)``: the exit
  direction is not Snell's (and not unit length).  The trace then measures
  ``t`` in units of that direction (OptiX parametric distance).
"""
from __future__ import annotations

import dataclasses
import math
import typing as tp

import numpy as np
import torch

F64 = torch.float64
T_MIN = 1e-3        # scene_epsilon, PyOptix/RayTrace.cpp:48
MISS_T = -2.0       # miss program, PyOptix/kernel/ray_programs.cu:97-101


# ----------------------------------------------------------------------------- geometry helpers

def normalize(v: torch.Tensor, dim=-1) -> torch.Tensor:
    """model/utils.py:25-31: v / |v| where |v| > 0, else 0."""
    n = v.norm(dim=dim, keepdim=True)
    return torch.where(n > 0, v / torch.where(n > 0, n, torch.ones_like(n)), torch.zeros_like(v))


def dot(a, b, keepdim=False):
    return (a * b).sum(-1, keepdim=keepdim)


def grid_coords(R: int) -> torch.Tensor:
    """[2, R, R]: coords[0] = x along the width, coords[1] = y along the height
    (shape_from_caustics.py:537: stack(meshgrid(linspace, linspace)).flip(0))."""
    a = torch.linspace(-1, 1, R, dtype=F64)
    A, B = torch.meshgrid(a, a, indexing="ij")
    return torch.stack([B, A])


def sobel_normals(h: torch.Tensor, delta: float, normalize_out=True) -> torch.Tensor:
    """model/caustics.py:215-221: [3, R, R] normals from a replicate-padded Sobel
    filter scaled by -1 / (8 element_size), z = 1, then normalised."""
    hp = torch.nn.functional.pad(h[None, None], (1, 1, 1, 1), mode="replicate")[0, 0]
    R0, R1 = h.shape
    w = (1.0, 2.0, 1.0)
    gx = sum(w[a] * (hp[a:a + R0, 2:2 + R1] - hp[a:a + R0, 0:R1]) for a in range(3))
    gy = sum(w[b] * (hp[2:2 + R0, b:b + R1] - hp[0:R0, b:b + R1]) for b in range(3))
    s = -1.0 / (8.0 * delta)
    n = torch.stack([gx * s, gy * s, torch.ones_like(h)])
    return normalize(n, dim=0) if normalize_out else n


def _shuffle(p: torch.Tensor, flip=False) -> torch.Tensor:
    """model/renderable_object.py:52-78 (grid_shuffle): [r, c] grid -> triangles
    [2 (rows-1)(cols-1), 3, d]: lower-left (p00, p01, p10), upper-right
    (p01, p11, p10), interleaved; flip swaps vertices 1 and 2."""
    a, b, c, d = p[:-1, :-1], p[:-1, 1:], p[1:, :-1], p[1:, 1:]
    lo = torch.stack([a, c, b] if flip else [a, b, c], dim=-2)
    up = torch.stack([b, c, d] if flip else [b, d, c], dim=-2)
    k = lo.shape[-1]
    return torch.stack([lo.reshape(-1, 3, k), up.reshape(-1, 3, k)], dim=1).reshape(-1, 3, k)


@dataclasses.dataclass
class Scene:
    tris: torch.Tensor      # [T, 3, 3] vertices
    vnorm: torch.Tensor     # [T, 3, 3] vertex normals
    obj: torch.Tensor       # [T] object index (0 height-field substrate, 1 ground)
    prim: torch.Tensor      # [T] primitive index inside the object


def build_scene(h: torch.Tensor, sensor_z: float) -> Scene:
    """Height-field substrate (top, bottom, front, back, left, right;
    model/renderable_object.py:249-308) plus the ground plane
    (:171-193), in that triangle order."""
    R = h.shape[0]
    coords = grid_coords(R)
    delta = float(coords[0, 1, 1] - coords[0, 0, 0])
    n = sobel_normals(h, delta).permute(1, 2, 0)                      # [R, R, 3]
    xy = coords.permute(1, 2, 0)                                      # [R, R, 2]
    pts = torch.cat([xy, h[..., None]], -1)
    bot = torch.cat([xy, torch.zeros_like(h)[..., None]], -1)

    def const(v, shape):
        return torch.tensor(v, dtype=F64).expand(*shape, 3)

    parts = [
        (_shuffle(pts), _shuffle(n)),
        (_shuffle(bot, True), _shuffle(const([0, 0, -1], (R, R)), True)),
        (_shuffle(torch.cat([bot[0:1], pts[0:1]], 0)), _shuffle(const([0, -1, 0], (2, R)))),
        (_shuffle(torch.cat([bot[-1:], pts[-1:]], 0), True), _shuffle(const([0, 1, 0], (2, R)), True)),
        (_shuffle(torch.cat([bot[:, 0:1], pts[:, 0:1]], 1)), _shuffle(const([-1, 0, 0], (R, 2)))),
        (_shuffle(torch.cat([bot[:, -1:], pts[:, -1:]], 1), True), _shuffle(const([1, 0, 0], (R, 2)), True)),
    ]
    tris = torch.cat([p[0] for p in parts])
    vn = normalize(torch.cat([p[1] for p in parts]))                 # gram_schmidt normalises (:241-244)
    lo, hi = -1.0, 1.0
    g = torch.full((2, 3, 3), sensor_z, dtype=F64)
    g[0, 0, :2] = torch.tensor([lo, lo]); g[0, 1, :2] = torch.tensor([hi, lo]); g[0, 2, :2] = torch.tensor([lo, hi])
    g[1, 0, :2] = torch.tensor([lo, hi]); g[1, 1, :2] = torch.tensor([hi, lo]); g[1, 2, :2] = torch.tensor([hi, hi])
    gn = torch.tensor([0.0, 0.0, 1.0], dtype=F64).expand(2, 3, 3)
    T0 = tris.shape[0]
    return Scene(torch.cat([tris, g]), torch.cat([vn, gn]),
                 torch.cat([torch.zeros(T0, dtype=torch.int64), torch.ones(2, dtype=torch.int64)]),
                 torch.cat([torch.arange(T0), torch.arange(2)]))


# ----------------------------------------------------------------------------- trace (brute force)

def trace_bruteforce(scene: Scene, o: torch.Tensor, d: torch.Tensor, t_min=T_MIN, chunk=256):
    """Closest hit over every triangle (Moeller-Trumbore, no culling), t >= t_min,
    parametric in d.  Returns (t, b [N, 2], obj, prim, tri) with the miss
    values (-2, (-1, -1), -1, -1, -1) (PyOptix/kernel/ray_programs.cu:84-101)."""
    V = scene.tris.detach()
    e1, e2 = V[:, 1] - V[:, 0], V[:, 2] - V[:, 0]
    N = o.shape[0]
    t_out = torch.full((N,), MISS_T, dtype=F64)
    b_out = torch.full((N, 2), -1.0, dtype=F64)
    tri_out = torch.full((N,), -1, dtype=torch.int64)
    for s in range(0, N, chunk):
        oo, dd = o[s:s + chunk, None, :].detach(), d[s:s + chunk, None, :].detach()
        pv = torch.linalg.cross(dd.expand(-1, V.shape[0], -1), e2[None].expand(oo.shape[0], -1, -1), dim=-1)
        det = dot(e1[None], pv)
        ok = det.abs() > 0
        inv = torch.where(ok, 1.0 / torch.where(ok, det, torch.ones_like(det)), torch.zeros_like(det))
        tv = oo - V[None, :, 0]
        u = dot(tv, pv) * inv
        qv = torch.linalg.cross(tv, e1[None].expand(oo.shape[0], -1, -1), dim=-1)
        v = dot(dd, qv) * inv
        t = dot(e2[None], qv) * inv
        hit = ok & (u >= 0) & (v >= 0) & (u + v <= 1) & (t >= t_min)
        tt = torch.where(hit, t, torch.full_like(t, math.inf))
        best, k = tt.min(1)
        got = torch.isfinite(best)
        idx = torch.arange(oo.shape[0])
        t_out[s:s + chunk] = torch.where(got, best, t_out[s:s + chunk])
        b_out[s:s + chunk] = torch.where(got[:, None], torch.stack([u[idx, k], v[idx, k]], -1), b_out[s:s + chunk])
        tri_out[s:s + chunk] = torch.where(got, k, tri_out[s:s + chunk])
    obj = torch.where(tri_out >= 0, scene.obj[tri_out.clamp(min=0)], torch.full_like(tri_out, -1))
    prim = torch.where(tri_out >= 0, scene.prim[tri_out.clamp(min=0)], torch.full_like(tri_out, -1))
    return t_out, b_out, obj, prim, tri_out


# ----------------------------------------------------------------------------- shading, transfer, refraction

def geometric_normal(V: torch.Tensor) -> torch.Tensor:
    """model/renderable_object.py:383-389 (cw = False)."""
    return normalize(torch.linalg.cross(V[:, 1] - V[:, 0], V[:, 2] - V[:, 0], dim=-1))


def shading_normal(VN: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """model/utils.py:65-74 barycentric interpolation, renormalised
    (model/renderable_object.py:505-506)."""
    w = torch.stack([1 - b[:, 0] - b[:, 1], b[:, 0], b[:, 1]], -1)
    return normalize((w[..., None] * VN).sum(1))


def transfer(n, d, t, Du, Dv, Dth, Dph):
    """model/photon_differential.py:70-78 (Frisvad 2014): positional
    differentials moved onto the hit plane of normal n (t detached)."""
    den = dot(n, d, keepdim=True)
    Du_hit = Du - (dot(n, Du, keepdim=True) / den) * d
    Dv_hit = Dv - (dot(n, Dv, keepdim=True) / den) * d
    Dth_hit = t[:, None] * (Dth - (dot(n, Dth, keepdim=True) / den) * d)
    Dph_hit = t[:, None] * (Dph - (dot(n, Dph, keepdim=True) / den) * d)
    return Du_hit + Dth_hit, Dv_hit + Dph_hit, Dth_hit, Dph_hit


def l_planes(V: torch.Tensor) -> torch.Tensor:
    """model/renderable_object.py:211-224: Igehy barycentric planes [T, 3, 4]."""
    g = geometric_normal(V)
    orgs = V[:, [1, 2, 0]]
    L = torch.linalg.cross(g[:, None].expand_as(V), V[:, [2, 0, 1]] - orgs, dim=-1)
    dd = dot(orgs, orgs, keepdim=True)
    factor = dot(L, V, keepdim=True) + dd
    return torch.cat([L, dd], -1) / factor


def differential_normal_literal(V, ns, x):
    """model/renderable_object.py:391-401, literally: ((ns.ns) dn - (ns.dn) ns) / (ns.ns)^1.5
    with dn = (sum_v L_v[:3] . x) ns."""
    L = l_planes(V)[:, :, :3]
    s = (L * x[:, None, :]).sum(-1).sum(-1, keepdim=True)   # sum over vertices of L_v . x
    dn = s * ns
    ndn = dot(ns, ns, keepdim=True)
    return (ndn * dn - dot(ns, dn, keepdim=True) * ns) / ndn ** 1.5


def refract_reflect(d, N0, Dth, Dph, dNdt, dNdp, eta_out, eta_in, return_tir=False):
    """model/caustics.py:152-212, restated literally (cosi kept unflipped).
    return_tir: also return the mask of rays that took the total-internal-
    reflection branch (:205-210)."""
    cosi = dot(d, N0, keepdim=True).clamp(-1, 1)
    flip = cosi > 0
    etai = torch.where(flip, eta_in, eta_out)
    etat = torch.where(flip, eta_out, eta_in)
    N = torch.where(flip, -N0, N0)
    dNdt = torch.where(flip, -dNdt, dNdt)
    dNdp = torch.where(flip, -dNdp, dNdp)
    eta = etai / etat
    k = 1 - eta ** 2 * (1 - cosi ** 2)
    rm = k > 0
    ks = torch.where(rm, k, torch.ones_like(k))
    neg_sq = -torch.sqrt(ks)
    coso = eta * cosi - neg_sq
    a = eta - (eta ** 2 * cosi) / neg_sq
    d_r = eta * d - coso * N
    Dt_r = eta * Dth - (coso * dNdt + (a * (dot(Dth, N, keepdim=True) + dot(d, dNdt, keepdim=True))) * N)
    Dp_r = eta * Dph - (coso * dNdp + (a * (dot(Dph, N, keepdim=True) + dot(d, dNdp, keepdim=True))) * N)
    d_t = d - 2 * cosi * N
    Dt_t = Dth - 2 * (cosi * dNdt + (dot(Dth, N, keepdim=True) + dot(d, dNdt, keepdim=True)) * N)
    Dp_t = Dph - 2 * (cosi * dNdp + (dot(Dph, N, keepdim=True) + dot(d, dNdp, keepdim=True)) * N)
    out = torch.where(rm, d_r, d_t), torch.where(rm, Dt_r, Dt_t), torch.where(rm, Dp_r, Dp_t)
    return (*out, ~rm[:, 0]) if return_tir else out


def splat_inputs(p, Du, Dv, ns, flux):
    """model/caustics.py:224-239: (Ep, xp, Mp, radius, good); rows of Mp are
    (2/den)(Dv x n) and (2/den)(n x Du), den = Du . (Dv x n)."""
    tmp = torch.linalg.cross(Dv, ns, dim=-1)
    den = dot(Du, tmp)
    good = den != 0
    f = 2 / torch.where(good, den, torch.ones_like(den))
    M = torch.stack([f[:, None] * tmp, f[:, None] * torch.linalg.cross(ns, Du, dim=-1)], 0)   # [2, N, 3]
    radius = 0.5 * torch.maximum(Du.norm(dim=-1), Dv.norm(dim=-1))
    Ep = flux / torch.linalg.cross(Du, Dv, dim=-1).norm(dim=-1) * 4
    return Ep, p[:, :2].T, M.permute(0, 2, 1), radius, good


# ----------------------------------------------------------------------------- emitters

@dataclasses.dataclass
class Rays:
    o: torch.Tensor
    d: torch.Tensor
    Du: torch.Tensor
    Dv: torch.Tensor
    Dth: torch.Tensor
    Dph: torch.Tensor
    flux: torch.Tensor
    channel: torch.Tensor


def parallel_rays(n_side: int, z0: float, flux_total=1.0, rows=None) -> Rays:
    """Vertical rays (0, 0, -1) through the centres of an n_side^2 grid over
    [-1, 1]^2 starting at height z0; Du = 2 pitch e_x, Dv = 2 pitch e_y,
    Dtheta = Dphi = 0, flux 1 / n_rays each, channel 0.  ``rows`` selects a
    subset of ray indices (row-major j + n_side i)."""
    idx = torch.arange(n_side * n_side) if rows is None else torch.as_tensor(rows)
    i, j = idx // n_side, idx % n_side
    pitch = 2.0 / n_side
    x = -1.0 + (2 * j.to(F64) + 1) / n_side
    y = -1.0 + (2 * i.to(F64) + 1) / n_side
    n = idx.numel()
    o = torch.stack([x, y, torch.full((n,), z0, dtype=F64)], -1)
    d = torch.tensor([0.0, 0.0, -1.0], dtype=F64).expand(n, 3).clone()
    Du = torch.tensor([2 * pitch, 0.0, 0.0], dtype=F64).expand(n, 3).clone()
    Dv = torch.tensor([0.0, 2 * pitch, 0.0], dtype=F64).expand(n, 3).clone()
    z = torch.zeros(n, 3, dtype=F64)
    return Rays(o, d, Du, Dv, z, z.clone(), torch.full((n,), flux_total / (n_side * n_side), dtype=F64),
                torch.zeros(n, dtype=torch.int64))


def slerp(p0, p1, t):
    """model/utils.py:84-88."""
    w = torch.acos(dot(p0, p1, keepdim=True).clamp(-1, 1))
    return (torch.sin((1 - t) * w) * p0 + torch.sin(t * w) * p1) / torch.sin(w)


def frisvad_basis(n, eps=1e-5):
    """model/utils.py:33-53 (Frisvad 2012), n [N, 3] unit."""
    nx, ny, nz = n[:, 0], n[:, 1], n[:, 2]
    sing = nz < (-1 + eps)
    a = 1 / (1 + torch.where(sing, torch.zeros_like(nz), nz))
    b = -nx * ny * a
    b1 = torch.stack([1 - a * nx ** 2, b, -nx], -1)
    b2 = torch.stack([b, 1 - a * ny ** 2, -ny], -1)
    b1 = torch.where(sing[:, None], torch.tensor([0.0, -1.0, 0.0], dtype=F64), b1)
    b2 = torch.where(sing[:, None], torch.tensor([-1.0, 0.0, 0.0], dtype=F64), b2)
    return b1, b2


def point_light_dirs(a, b, light, plane_z) -> torch.Tensor:
    """Directions [n, 3] of the point light (model/caustics.py:38-44): the
    bilinear slerp, with coefficients a, b [n], of the unit directions from
    ``light`` to the corners of the square [-1, 1]^2 at height ``plane_z``
    (lower edge y = -1 and upper edge y = +1 interpolated by a, then the two
    by b)."""
    L = torch.tensor(light, dtype=F64)
    a = torch.as_tensor(a, dtype=F64).reshape(-1, 1)
    b = torch.as_tensor(b, dtype=F64).reshape(-1, 1)
    n = a.shape[0]

    def corner(x, y):
        return normalize(torch.tensor([x, y, plane_z], dtype=F64) - L)[None].expand(n, 3)

    lower = slerp(corner(-1, -1), corner(1, -1), a)
    upper = slerp(corner(-1, 1), corner(1, 1), a)
    return slerp(lower, upper, b)


def point_light_rays(n_side, channels, light, plane_z, diff_scale, flux, seed) -> Rays:
    """model/caustics.py:26-70 with the interpolation coefficients (a, b) of
    sample s drawn as SplitMix64(seed * 0x100000001B3 ^ (s << 2 | k)), k = 0, 1
    (the GPU emitter's counter-based draws); photon i is sample i // channels
    in channel i % channels."""
    from oracle import gen_twin as GT
    n_s = n_side * n_side
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * np.uint64(0x100000001B3)) & GT.M64
    ctr = np.arange(n_s, dtype=np.uint64) << np.uint64(2)
    a = torch.from_numpy(GT.u24(GT.splitmix64(key ^ (ctr | np.uint64(0)))).astype(np.float64))
    b = torch.from_numpy(GT.u24(GT.splitmix64(key ^ (ctr | np.uint64(1)))).astype(np.float64))
    L = torch.tensor(light, dtype=F64)
    d = point_light_dirs(a, b, light, plane_z)
    b1, b2 = frisvad_basis(d)
    rep = lambda t: t.repeat_interleave(channels, 0)
    n = n_s * channels
    z = torch.zeros(n, 3, dtype=F64)
    return Rays(L[None].expand(n, 3).clone(), rep(d), z, z.clone(), rep(diff_scale * b1), rep(diff_scale * b2),
                torch.full((n,), float(flux), dtype=F64), torch.arange(n) % channels)


# ----------------------------------------------------------------------------- the bounce loop

@dataclasses.dataclass
class Path:
    """Per-ray result: PDS inputs of the rays that reached the ground plane."""
    ray: torch.Tensor       # [n] ray index
    Ep: torch.Tensor
    xp: torch.Tensor        # [2, n]
    Mp: torch.Tensor        # [2, 3, n]
    cp: torch.Tensor        # [1, n]
    radius: torch.Tensor
    hits: list              # per bounce: (ray index, t, b, tri) -- the detached trace results
    capped: torch.Tensor = None   # rays still bouncing after max_bounces hits (dropped)
    tir: torch.Tensor = None      # rays that were totally internally reflected at least once


MAX_BOUNCES = 16  # the GPU path's cap (HF_MAX_BOUNCES of include/hf.h)


def simulate(h: torch.Tensor, rays: Rays, ior: float, sensor_z: float, max_bounces=MAX_BOUNCES,
             frozen=None) -> Path:
    """model/caustics.py:242-291 for one substrate and a ground plane (object 1).

    The reference loops until every photon has left (:255); this loop stops
    after ``max_bounces`` hits and drops the rays still bouncing (``capped``),
    as the GPU path does with the same bound.

    ``frozen``: the ``hits`` of an earlier run; the trace is replaced by those
    hits so that the result is the smooth function of h that the reference
    differentiates (t, barycentrics, triangle fixed).
    """
    scene = build_scene(h, sensor_z)
    n = rays.o.shape[0]
    ray = torch.arange(n)
    o, d, Du, Dv, Dth, Dph = rays.o, rays.d, rays.Du, rays.Dv, rays.Dth, rays.Dph
    flux, ch = rays.flux, rays.channel
    eo = torch.ones(n, 1, dtype=F64)
    ei = torch.full((n, 1), float(ior), dtype=F64)
    ground = []
    hits = []
    tir_rays = []
    for bounce in range(max_bounces):
        if ray.numel() == 0:
            break
        if frozen is None:
            t, b, obj, prim, tri = trace_bruteforce(scene, o, d)
        else:
            fr, t, b, tri = frozen[bounce]
            assert torch.equal(fr, ray), "frozen path does not match"
            obj = torch.where(tri >= 0, scene.obj[tri.clamp(min=0)], torch.full_like(tri, -1))
        hits.append((ray.clone(), t.clone(), b.clone(), tri.clone()))
        keep = t >= 0                                          # advance_path drops misses (:54-60)
        sel = lambda a: a[keep]
        ray, t, b, obj, tri = sel(ray), sel(t), sel(b), sel(obj), sel(tri)
        o, d, Du, Dv, Dth, Dph, flux, ch, eo, ei = map(sel, (o, d, Du, Dv, Dth, Dph, flux, ch, eo, ei))
        p = o + t[:, None] * d
        V, VN = scene.tris[tri], scene.vnorm[tri]
        ng = geometric_normal(V)
        ns = shading_normal(VN, b)
        Du, Dv, Dth_hit, Dph_hit = transfer(ng, d, t, Du, Dv, Dth, Dph)
        on_ground = obj == 1
        if on_ground.any():
            g = on_ground
            ground.append((ray[g], p[g], Du[g], Dv[g], ns[g], flux[g], ch[g]))
        r = ~on_ground
        ray, p, d, Du, Dv, Dth, Dph, ns, V, Dth_hit, Dph_hit, flux, ch, eo, ei = (
            a[r] for a in (ray, p, d, Du, Dv, Dth, Dph, ns, V, Dth_hit, Dph_hit, flux, ch, eo, ei))
        dNdt = differential_normal_literal(V, ns, Dth_hit)
        dNdp = differential_normal_literal(V, ns, Dph_hit)
        d, Dth, Dph, tir = refract_reflect(d, ns, Dth, Dph, dNdt, dNdp, eo, ei, return_tir=True)
        tir_rays.append(ray[tir])
        o = p
    capped = ray.clone()  # still in flight after the last allowed hit
    tir_mask = torch.zeros(n, dtype=torch.bool)
    for r in tir_rays:
        tir_mask[r] = True
    if not ground:
        z = torch.zeros(0, dtype=F64)
        return Path(torch.zeros(0, dtype=torch.int64), z, z.reshape(2, 0), z.reshape(2, 3, 0),
                    torch.zeros(1, 0, dtype=torch.int64), z, hits, capped, tir_mask)
    ray, p, Du, Dv, ns, flux, ch = (torch.cat([g[k] for g in ground]) for k in range(7))
    Ep, xp, Mp, radius, good = splat_inputs(p, Du, Dv, ns, flux)
    return Path(ray[good], Ep[good], xp[:, good], Mp[:, :, good], ch[good][None], radius[good], hits, capped,
                tir_mask)


# ----------------------------------------------------------------------------- PDS in float64 (autograd)

def pds_forward_t64(Ep, xp, Mp, cp, radius, size, cap):
    """photon_differentials.cu:60-113 in torch float64, vectorised over photons
    per window offset; differentiable in (Ep, xp, Mp) by autograd (the
    reference's backward, :136-221, is this derivative)."""
    C, H, W = size
    N = Ep.shape[0]
    out = torch.zeros(C * H * W, dtype=F64)
    if N == 0:
        return out.reshape(C, H, W)
    x = xp[0].detach().numpy().astype(np.float64)
    y = xp[1].detach().numpy().astype(np.float64)
    r = radius.detach().numpy().astype(np.float64)
    rx = np.minimum(np.ceil(r * 0.5 * W), cap).astype(np.int64)
    ry = np.minimum(np.ceil(r * 0.5 * H), cap).astype(np.int64)
    pxc = np.trunc((0.5 * x + 0.5) * W).astype(np.int64)
    pyc = np.trunc((0.5 * y + 0.5) * H).astype(np.int64)
    c = cp[0].numpy()
    for oy in range(-cap, cap + 1):
        for ox in range(-cap, cap + 1):
            px, py = pxc + ox, pyc + oy
            m = (np.abs(ox) <= rx) & (np.abs(oy) <= ry) & (px >= 0) & (px < W) & (py >= 0) & (py < H) \
                & (c >= 0) & (c < C)
            if not m.any():
                continue
            mi = torch.from_numpy(np.nonzero(m)[0])
            dx = 2.0 * torch.from_numpy(px[m]).to(F64) / W - 1.0 - xp[0, mi]
            dy = 2.0 * torch.from_numpy(py[m]).to(F64) / H - 1.0 - xp[1, mi]
            u = Mp[0, 0, mi] * dx + Mp[0, 1, mi] * dy
            v = Mp[1, 0, mi] * dx + Mp[1, 1, mi] * dy
            q = u * u + v * v
            k = torch.where(q < 1, (3.0 / math.pi) * (1 - q) ** 2, torch.zeros_like(q)) * Ep[mi]
            k = torch.where(k > 0, k, torch.zeros_like(k))
            flat = torch.from_numpy((c[m] * H + py[m]) * W + px[m])
            out = out.index_add(0, flat, k)
    return out.reshape(C, H, W)


def render(h, rays, ior, sensor_z, size, cap, frozen=None, keep_rays=None):
    """Image of the photons that reach the ground; ``keep_rays`` (bool per ray)
    drops the others' contributions (channel -1)."""
    path = simulate(h, rays, ior, sensor_z, frozen=frozen)
    cp = path.cp
    if keep_rays is not None:
        cp = torch.where(keep_rays[path.ray][None], cp, torch.full_like(cp, -1))
    return pds_forward_t64(path.Ep, path.xp, path.Mp, cp, path.radius, size, cap), path


def edge_safe_rays(path: "Path", n_rays: int, margin=1e-5) -> torch.Tensor:
    """Rays whose every hit keeps ``margin`` (barycentric) away from its
    triangle's edges: there an FP32 trace picks the same triangle as this
    FP64 one (near an edge either triangle is a valid hit of the
    piecewise-flat surface, with a different face normal)."""
    safe = torch.ones(n_rays, dtype=torch.bool)
    for ray, t, b, tri in path.hits:
        bmin = torch.minimum(torch.minimum(b[:, 0], b[:, 1]), 1 - b[:, 0] - b[:, 1])
        safe[ray[(t >= 0) & (bmin < margin)]] = False
    return safe


def loss_and_grad(h0: torch.Tensor, rays, ior, sensor_z, size, cap, target: torch.Tensor, keep_rays=None):
    """L = 1/2 sum (I - T)^2 (hyperparameter_helper.py:29-30) and dL/dh with the
    reference's detached trace.  Returns (L, dL/dh, image, path)."""
    h = h0.detach().clone().requires_grad_(True)
    img, path = render(h, rays, ior, sensor_z, size, cap, keep_rays=keep_rays)
    L = 0.5 * ((img - target) ** 2).sum()
    L.backward()
    return float(L.detach()), h.grad.detach(), img.detach(), path


def frozen_loss(h, rays, ior, sensor_z, size, cap, target, frozen) -> float:
    with torch.no_grad():
        img, _ = render(h, rays, ior, sensor_z, size, cap, frozen=frozen)
        return float(0.5 * ((img - target) ** 2).sum())


# ----------------------------------------------------------------------------- height fields, Landweber update

def gaussian_bumps(R: int, seed: int, n_bumps=32, base=0.2, amp=(0.0, 0.1), sigma=(0.06, 0.4),
                   centre=(-0.6, 0.6)) -> np.ndarray:
    """Sum of Gaussian bumps on a base thickness over the [-1, 1]^2 grid."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a = rng.uniform(*amp, n_bumps)
    s = rng.uniform(*sigma, n_bumps)
    mu = rng.uniform(*centre, (n_bumps, 2))
    c = grid_coords(R).numpy()
    h = np.full((R, R), base)
    for k in range(n_bumps):
        h += a[k] * np.exp(-((c[0] - mu[k, 0]) ** 2 + (c[1] - mu[k, 1]) ** 2) / (2 * s[k] ** 2))
    return h


def ext_shrink(x, alpha, a, b):
    """shape_from_caustics.py:184-189: clamp(max(|x| - alpha, 0) sign(x), a, b)."""
    return torch.clamp(torch.clamp(x.abs() - alpha, min=0) * torch.sign(x), max=b).clamp(min=a)


def vcons_heur(x, d1, d2, mask, gamma, volume_radius):
    """shape_from_caustics.py:228-241."""
    vx = x.sum()
    k = 2 * volume_radius + 1
    xmean = torch.nn.functional.avg_pool2d(x[None, None], k, stride=1, padding=volume_radius)[0, 0]
    if vx <= d1:
        return x + mask * xmean * torch.sin(vx / d1 * math.pi) * gamma
    if vx >= d2:
        return x - mask * xmean * torch.sin((vx / d2 - 1) * math.pi) * gamma
    return x


def landweber_update(d, grad, mask, tau, alpha, lb, ub, d1, d2, gamma, volume_radius):
    """One pixel-Landweber update (shape_from_caustics.py:387-402, default
    options): gradient step, masked extended soft shrinkage, volume heuristic."""
    x = d - tau * grad
    x = mask * ext_shrink(x, tau * alpha, lb, ub)
    return vcons_heur(x, d1, d2, mask, gamma, volume_radius)
