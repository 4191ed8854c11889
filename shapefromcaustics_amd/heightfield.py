"""Height-field caustics on the GPU: the optics around the splat, for a glass
substrate over a ground plane (include/hf.h).

Replaces, for this scene, the reference's OptiX trace and torch optics:
``get_normal_from_height_field`` (``model/caustics.py:215-221``),
``trace_rays`` (``PyOptix/RayTrace.cpp:233-291``), the bounce loop of
``compute_recursive_refraction`` (``model/caustics.py:242-291``) with
``advance_path`` (``model/photon_differential.py:53-78``) and
``refract_reflect_differentials`` (``model/caustics.py:152-212``), and
``splat_photons`` (``:224-239``) up to the PDS call.

:class:`CausticImage` is a ``torch.autograd.Function`` h -> image whose
backward runs the PDS adjoint and then the hand-written reverse sweep of the
photon paths (recomputed, not stored), with the reference's gradient
semantics (trace results are constants).  :func:`landweber_step` is one
pixel-Landweber iteration (``shape_from_caustics.py:344-420``) built from it.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import typing as tp

import torch

from . import _native
from .PhotonDifferentialSplatting import _backward, _forward

_vp, _i32, _i64, _sz, _f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_size_t, ctypes.c_float
HF_MAX_CHANNELS = 8


class HfScene(ctypes.Structure):
    _fields_ = [("R", _i32), ("channels", _i32), ("sensor_z", _f32), ("ior", _f32 * HF_MAX_CHANNELS),
                ("reserved", _i32 * 4)]


class HfEmitter(ctypes.Structure):
    _fields_ = [("n_side", _i32), ("z0", _f32), ("flux_total", _f32), ("kind", _i32), ("light", _f32 * 3),
                ("plane_z", _f32), ("diff_scale", _f32), ("seed", ctypes.c_uint32), ("reserved", _i32 * 2)]


class HfPathStats(ctypes.Structure):
    _fields_ = [("reached", _i64), ("missed", _i64), ("degenerate", _i64), ("capped", _i64),
                ("reserved", _i64 * 4)]


MAX_BOUNCES = 16  # HF_MAX_BOUNCES of include/hf.h


class HfLandweber(ctypes.Structure):
    _fields_ = [("tau", _f32), ("alpha", _f32), ("lower", _f32), ("upper", _f32), ("d1", _f32), ("d2", _f32),
                ("gamma", _f32), ("volume_radius", _i32), ("reserved", _i32 * 4)]


_SP, _EP, _LP = ctypes.POINTER(HfScene), ctypes.POINTER(HfEmitter), ctypes.POINTER(HfLandweber)

SIGNATURES = {
    "hf_last_error": ([], ctypes.c_char_p),
    "hf_workspace_bytes": ([_i32], _sz),
    "hf_normals_f32": ([_vp, _i32, _vp, _vp], ctypes.c_int),
    "hf_trace_f32": ([_vp, _vp, _SP, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _sz, _vp], ctypes.c_int),
    "hf_simulate_f32": ([_vp, _vp, _SP, _EP, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp], ctypes.c_int),
    "hf_adjoint_f32": ([_vp, _vp, _SP, _EP, _vp, _vp, _vp, _vp, _vp, _sz, _vp], ctypes.c_int),
    "hf_hits_bytes": ([_SP, _EP], _sz),
    "hf_hits_stats_get": ([_vp, _sz, _SP, _EP, ctypes.c_void_p, _vp], ctypes.c_int),
    "hf_simulate_hits_f32": ([_vp, _vp, _SP, _EP, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp, _sz, _vp], ctypes.c_int),
    "hf_adjoint_hits_f32": ([_vp, _vp, _SP, _EP, _vp, _vp, _vp, _vp, _vp, _sz, _vp, _sz, _vp], ctypes.c_int),
    "hf_residual_workspace_bytes": ([_i64], _sz),
    "hf_residual_f32": ([_vp, _vp, _i64, _vp, _vp, _vp, _sz, _vp], ctypes.c_int),
    "hf_landweber_workspace_bytes": ([_i32], _sz),
    "hf_landweber_f32": ([_vp, _vp, _vp, _i32, _LP, _vp, _sz, _vp], ctypes.c_int),
}

_bound = False


def lib():
    global _bound
    L = _native.lib()
    if not _bound:
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _bound = True
    return L


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {lib().hf_last_error().decode()}")


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else ctypes.c_void_p(0)


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


@dataclasses.dataclass
class Scene:
    """Substrate over [-1, 1]^2 (flat bottom z = 0) and a ground plane at ``sensor_z``."""

    R: int
    sensor_z: float
    iors: tp.Sequence[float] = (1.5,)

    def struct(self) -> HfScene:
        s = HfScene()
        s.R = int(self.R)
        s.channels = len(self.iors)
        s.sensor_z = float(self.sensor_z)
        if not 1 <= len(self.iors) <= HF_MAX_CHANNELS:
            raise ValueError(f"1..{HF_MAX_CHANNELS} channels supported")
        for k, v in enumerate(self.iors):
            s.ior[k] = float(v)
        return s


@dataclasses.dataclass
class ParallelEmitter:
    """Vertical rays through the centres of an n_side^2 grid, from height z0."""

    n_side: int
    z0: float
    flux_total: float = 1.0

    def struct(self) -> HfEmitter:
        e = HfEmitter()
        e.n_side, e.z0, e.flux_total = int(self.n_side), float(self.z0), float(self.flux_total)
        return e


def subtended_angle(light, tris) -> float:
    """Solid angle of triangles [T, 3, 3] seen from ``light`` (Oosterom and
    Strackee, as model/utils.py:91-108: 2 atan2(d0 . (d2 x d1), |d0||d1||d2| +
    |d0| d1.d2 + |d1| d0.d2 + |d2| d0.d1), negative (back-facing) ones dropped)."""
    import numpy as np
    total = 0.0
    for t in np.asarray(tris, np.float64):
        d0, d1, d2 = (t[k] - np.asarray(light, np.float64) for k in range(3))
        n0, n1, n2 = (np.linalg.norm(v) for v in (d0, d1, d2))
        num = float(np.dot(d0, np.cross(d2, d1)))
        den = n0 * n1 * n2 + n0 * float(np.dot(d1, d2)) + n1 * float(np.dot(d0, d2)) + n2 * float(np.dot(d0, d1))
        if not (num == 0 and den == 0):
            total += max(2.0 * math.atan2(num, den), 0.0)
    return total


@dataclasses.dataclass
class PointLight:
    """Point light at ``light`` aiming through the square [-1, 1]^2 at height
    ``plane_z`` (model/caustics.py:26-70): n_side^2 samples, flux per photon,
    Dtheta/Dphi scaled by 2 smoothing sqrt(solid angle / (pi total photons))."""

    n_side: int
    light: tp.Sequence[float]
    plane_z: float
    flux: float = 1e-5
    smoothing: float = 1.0
    seed: int = 0
    total_photons: tp.Optional[int] = None  # defaults to this emitter's photon count

    def diff_scale(self, channels: int = 1) -> float:
        z = self.plane_z
        tris = [[[-1, -1, z], [1, -1, z], [-1, 1, z]], [[-1, 1, z], [1, -1, z], [1, 1, z]]]
        total = self.total_photons or self.n_side ** 2 * channels
        return 2.0 * self.smoothing * math.sqrt(subtended_angle(self.light, tris) / (math.pi * total))

    def struct(self, channels: int = 1) -> HfEmitter:
        e = HfEmitter()
        e.n_side, e.kind, e.flux_total = int(self.n_side), 1, float(self.flux)
        for k in range(3):
            e.light[k] = float(self.light[k])
        e.plane_z, e.diff_scale, e.seed = float(self.plane_z), float(self.diff_scale(channels)), int(self.seed)
        return e


def _emitter_struct(emitter, scene) -> HfEmitter:
    return emitter.struct(len(scene.iors)) if isinstance(emitter, PointLight) else emitter.struct()


def _check_h(h: torch.Tensor, R: int):
    if not h.is_cuda or h.dtype != torch.float32 or not h.is_contiguous() or tuple(h.shape) != (R, R):
        raise RuntimeError(f"height field must be a contiguous float32 CUDA tensor of shape ({R}, {R})")


def normals(h: torch.Tensor) -> torch.Tensor:
    """Unit Sobel normals [R, R, 3] (model/caustics.py:215-221)."""
    R = h.shape[0]
    _check_h(h, R)
    out = torch.empty(R, R, 3, dtype=torch.float32, device=h.device)
    _check(lib().hf_normals_f32(_p(h), R, _p(out), _stream(h.device)), "hf_normals_f32")
    return out


def _ws(R, device):
    return torch.empty(lib().hf_workspace_bytes(R), dtype=torch.uint8, device=device)


def trace_rays(h: torch.Tensor, scene: Scene, origins: torch.Tensor, dirs: torch.Tensor):
    """The reference's trace contract for this scene: (t, uv, object, primitive)."""
    _check_h(h, scene.R)
    N = origins.shape[0]
    dev = h.device
    n = normals(h)
    t = torch.empty(N, dtype=torch.float32, device=dev)
    uv = torch.empty(N, 2, dtype=torch.float32, device=dev)
    obj = torch.empty(N, dtype=torch.int64, device=dev)
    prim = torch.empty(N, dtype=torch.int64, device=dev)
    ws = _ws(scene.R, dev)
    sc = scene.struct()
    _check(lib().hf_trace_f32(_p(h), _p(n), ctypes.byref(sc), _p(origins.contiguous()), _p(dirs.contiguous()), N,
                              _p(t), _p(uv), _p(obj), _p(prim), _p(ws), ws.numel(), _stream(dev)), "hf_trace_f32")
    return t, uv, obj, prim


HITS_DEGENERATE, HITS_CAPPED, HITS_RETRACE = 253, 254, 255


def hit_counts(hits: torch.Tensor, n_photons: int) -> torch.Tensor:
    """The per-photon hit counts inside a hit buffer from ``simulate(...,
    keep_hits=True)``: 0 = dropped, it missed the scene; 1..4 = hits
    recorded; 253 (``HITS_DEGENERATE``) = dropped by the splat's den != 0
    test; 254 (``HITS_CAPPED``) = dropped still bouncing after
    ``MAX_BOUNCES`` hits; 255
    (``HITS_RETRACE``) = more bounces than recorded (the adjoint traces that
    path again).  Layout of the buffer (csrc/hf_kernels.hip, hits_layout):
    4 x N float4 records, then the N count bytes at the next 256-byte
    boundary, then the field's maximum and minimum height and the path
    counters."""
    off = (16 * 4 * n_photons + 255) // 256 * 256
    return hits[off:off + n_photons]


def path_stats(hits: torch.Tensor, scene: Scene, emitter) -> dict:
    """How the paths of the simulate that wrote ``hits`` ended (synchronises):
    ``reached``, ``missed``, ``degenerate`` (the den != 0 mask) and
    ``capped`` (still bouncing after ``MAX_BOUNCES`` hits) photon counts."""
    sc, em = scene.struct(), _emitter_struct(emitter, scene)
    st = HfPathStats()
    _check(lib().hf_hits_stats_get(_p(hits), hits.numel(), ctypes.byref(sc), ctypes.byref(em), ctypes.byref(st),
                                   _stream(hits.device)), "hf_hits_stats_get")
    return {k: int(getattr(st, k)) for k in ("reached", "missed", "degenerate", "capped")}


def simulate(h: torch.Tensor, scene: Scene, emitter: ParallelEmitter, n: tp.Optional[torch.Tensor] = None,
             keep_hits: bool = False):
    """Splat inputs (Ep, xp, Mp, cp, radius) of every emitted photon; dropped photons have cp = -1.

    keep_hits: also return the photons' ray hits (a byte tensor), which
    :func:`adjoint` replays instead of tracing every path again."""
    _check_h(h, scene.R)
    dev = h.device
    n = normals(h) if n is None else n
    N = emitter.n_side ** 2 * len(scene.iors)
    f32 = dict(dtype=torch.float32, device=dev)
    Ep, xp, Mp = torch.empty(N, **f32), torch.empty(2, N, **f32), torch.empty(2, 3, N, **f32)
    cp, radius = torch.empty(1, N, dtype=torch.int64, device=dev), torch.empty(N, **f32)
    ws = _ws(scene.R, dev)
    sc, em = scene.struct(), _emitter_struct(emitter, scene)
    if not keep_hits:
        _check(lib().hf_simulate_f32(_p(h), _p(n), ctypes.byref(sc), ctypes.byref(em), _p(Ep), _p(xp), _p(Mp),
                                     _p(cp), _p(radius), _p(ws), ws.numel(), _stream(dev)), "hf_simulate_f32")
        return Ep, xp, Mp, cp, radius
    hits = torch.empty(lib().hf_hits_bytes(ctypes.byref(sc), ctypes.byref(em)), dtype=torch.uint8, device=dev)
    _check(lib().hf_simulate_hits_f32(_p(h), _p(n), ctypes.byref(sc), ctypes.byref(em), _p(Ep), _p(xp), _p(Mp),
                                      _p(cp), _p(radius), _p(ws), ws.numel(), _p(hits), hits.numel(), _stream(dev)),
           "hf_simulate_hits_f32")
    return Ep, xp, Mp, cp, radius, hits


def adjoint(h: torch.Tensor, scene: Scene, emitter: ParallelEmitter, gE, gx, gM,
            n: tp.Optional[torch.Tensor] = None, hits: tp.Optional[torch.Tensor] = None) -> torch.Tensor:
    """dL/dh from dL/d(Ep, xp, Mp) of :func:`simulate`'s photons (hits: the
    forward's ``keep_hits`` output for the same h, scene and emitter)."""
    _check_h(h, scene.R)
    dev = h.device
    n = normals(h) if n is None else n
    out = torch.empty_like(h)
    ws = _ws(scene.R, dev)
    sc, em = scene.struct(), _emitter_struct(emitter, scene)
    if hits is None:
        _check(lib().hf_adjoint_f32(_p(h), _p(n), ctypes.byref(sc), ctypes.byref(em), _p(gE.contiguous()),
                                    _p(gx.contiguous()), _p(gM.contiguous()), _p(out), _p(ws), ws.numel(),
                                    _stream(dev)), "hf_adjoint_f32")
    else:
        _check(lib().hf_adjoint_hits_f32(_p(h), _p(n), ctypes.byref(sc), ctypes.byref(em), _p(gE.contiguous()),
                                         _p(gx.contiguous()), _p(gM.contiguous()), _p(out), _p(ws), ws.numel(),
                                         _p(hits), hits.numel(), _stream(dev)), "hf_adjoint_hits_f32")
    return out


class CausticImage(torch.autograd.Function):
    """h [R, R] -> caustic image [C, S, S] on the ground plane, differentiable in h."""

    @staticmethod
    def forward(ctx, h, scene: Scene, emitter: ParallelEmitter, size, max_pixel_radius: int, mode="auto"):
        n = normals(h)
        grad = ctx.needs_input_grad[0]
        out = simulate(h, scene, emitter, n, keep_hits=grad)
        Ep, xp, Mp, cp, radius = out[:5]
        img, bins = _forward(Ep, xp, Mp, cp, radius, size, max_pixel_radius, mode, None, keep_bins=grad)
        ctx.save_for_backward(h, n, Ep, xp, Mp, cp, radius)
        ctx.scene, ctx.emitter, ctx.cap, ctx.mode, ctx.bins = scene, emitter, max_pixel_radius, mode, bins
        ctx.hits = out[5] if grad else None
        return img

    @staticmethod
    def backward(ctx, grad_img):
        h, n, Ep, xp, Mp, cp, radius = ctx.saved_tensors
        gE, gx, gM = _backward(grad_img.contiguous(), Ep, xp, Mp, cp, radius, ctx.cap, ctx.mode, None, bins=ctx.bins)
        ctx.bins = None
        gh = adjoint(h, ctx.scene, ctx.emitter, gE, gx, gM, n, hits=ctx.hits)
        ctx.hits = None
        return gh, None, None, None, None, None


def residual(img: torch.Tensor, target: torch.Tensor):
    """(img - target, 1/2 sum (img - target)^2 as a device scalar)."""
    n = img.numel()
    resid = torch.empty_like(img)
    loss = torch.empty((), dtype=torch.float32, device=img.device)
    ws = torch.empty(lib().hf_residual_workspace_bytes(n), dtype=torch.uint8, device=img.device)
    _check(lib().hf_residual_f32(_p(img.contiguous()), _p(target.contiguous()), n, _p(resid), _p(loss), _p(ws),
                                 ws.numel(), _stream(img.device)), "hf_residual_f32")
    return resid, loss


def pixel_area(R: int) -> float:
    """Area of one cell of the R x R height-field grid over [-1, 1]^2
    (shape_from_caustics.py:651: the product of the two grid spacings)."""
    return (2.0 / (R - 1)) ** 2


def printing_volume(d_exact: torch.Tensor) -> float:
    """Volume of the glass added on the substrate by the height offset
    ``d_exact`` [R, R] (shape_from_caustics.py:656: pixel area x sum of d)."""
    return pixel_area(int(d_exact.shape[0])) * float(d_exact.double().sum())


def volume_band(volume: float, R: int, volume_eps: float = 0.25) -> tp.Tuple[float, float]:
    """(d1, d2) of shape_from_caustics.py:383-384: the band of the summed height
    offset, volume (1 -/+ volume_eps) / pixel area (volume_eps default of
    hyperparameter_helper.py:93)."""
    a = pixel_area(R)
    return (volume - volume * volume_eps) / a, (volume + volume * volume_eps) / a


@dataclasses.dataclass
class LandweberParams:
    """Defaults of hyperparameter_helper.py:86-94 (tau, alpha, gamma, bounds, volume radius).

    ``d1``/``d2`` default to an unbounded band (the volume heuristic never
    acts); the reference derives them from the printing volume of the target,
    which :meth:`for_volume` does."""

    tau: float = 0.1
    alpha: float = 2e-3
    lower: float = 0.0
    upper: float = 0.3
    d1: float = -float("inf")
    d2: float = float("inf")
    gamma: float = 0.1
    volume_radius: int = 2

    @classmethod
    def for_volume(cls, volume: float, R: int, volume_eps: float = 0.25, **kw) -> "LandweberParams":
        """The reference's parameters for a target of this printing volume
        (:func:`printing_volume`): d1, d2 from :func:`volume_band`."""
        d1, d2 = volume_band(volume, R, volume_eps)
        return cls(d1=d1, d2=d2, **kw)

    def struct(self) -> HfLandweber:
        p = HfLandweber()
        for k in ("tau", "alpha", "lower", "upper", "d1", "d2", "gamma", "volume_radius"):
            setattr(p, k, getattr(self, k))
        return p


def landweber_update(d: torch.Tensor, grad: torch.Tensor, mask: torch.Tensor, params: LandweberParams) -> None:
    """In place: d <- VconsHeur(mask * extShrink(d - tau grad, tau alpha, lower, upper))."""
    R = d.shape[0]
    ws = torch.empty(lib().hf_landweber_workspace_bytes(R), dtype=torch.uint8, device=d.device)
    p = params.struct()
    _check(lib().hf_landweber_f32(_p(d), _p(grad.contiguous()), _p(mask.contiguous()), R, ctypes.byref(p), _p(ws),
                                  ws.numel(), _stream(d.device)), "hf_landweber_f32")


def landweber_step(d: torch.Tensor, base: float, target: torch.Tensor, scene: Scene, emitter: ParallelEmitter,
                   size, cap: int, mask: torch.Tensor, params: LandweberParams, mode="auto"):
    """One reconstruction iteration (shape_from_caustics.py:344-420): render
    h = d + base, loss 1/2 |I - T|^2, dL/dh by the adjoint, update d in place.
    Returns (loss device scalar, image, dL/dh)."""
    h = (d + base).contiguous()
    n = normals(h)
    Ep, xp, Mp, cp, radius, hits = simulate(h, scene, emitter, n, keep_hits=True)
    img, bins = _forward(Ep, xp, Mp, cp, radius, size, cap, mode, None, keep_bins=True)
    resid, loss = residual(img, target)
    gE, gx, gM = _backward(resid, Ep, xp, Mp, cp, radius, cap, mode, None, bins=bins)
    gh = adjoint(h, scene, emitter, gE, gx, gM, n, hits=hits)
    landweber_update(d, gh, mask, params)
    return loss, img, gh
