"""CPU oracle for photon-differential splatting (PDS).

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP product path
and the CPU baseline timed beside it.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import it;
the product package never does, and has no CPU fallback.

Contents
--------
* ``forward_f64`` / ``backward_f64``: ctypes bindings of the single-threaded
  FP64 C restatement in ``pds_oracle.c`` (reference
  ``PyOptix/kernel/photon_differentials.cu:60-113`` forward and ``:136-221``
  backward, window setup ``:80-87``).
* ``cpu_forward_f32`` / ``cpu_backward_f32``: the OpenMP FP32 host baseline of
  ``pds_cpu_baseline.c``.
* ``twin_forward`` / ``twin_backward``: an independent numpy restatement of the
  same reference lines, vectorised over photons per window offset instead of
  looping per photon.  It exists to cross-check the C oracle on small inputs.

Array layout is the reference's (``model/photon_differential.py:9``,
``PyOptix/PhotonDifferentialSplattig.cpp:19-22``): ``E [N]`` float32,
``xp [2, N]`` float32 (normalised coordinates in [-1, 1]),
``M [2, 3, N]`` float32, ``cp [1, N]`` int64, ``radius [N]`` float32; image
``[C, H, W]``.
"""
from __future__ import annotations

import collections
import ctypes
import os
import subprocess

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
_libs: dict = {}

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build(force: bool = False) -> None:
    """Compile the C checker libraries with the Makefile next to this file."""
    args = ["make", "-C", _DIR]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


def _lib(name: str):
    if name not in _libs:
        path = os.path.join(_DIR, name)
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        if name == "libpds_oracle.so":
            lib.pds_oracle_forward_f64.argtypes = [_f32p, _f32p, _f32p, _i64p, _f32p, ctypes.c_int64,
                                                   ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                   ctypes.c_int32, _f64p]
            lib.pds_oracle_backward_f64.argtypes = [_f32p, _f32p, _f32p, _f32p, _i64p, _f32p, ctypes.c_int64,
                                                    ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                    ctypes.c_int32, _i64p, ctypes.c_int64,
                                                    _f64p, _f64p, _f64p, _f64p, _f64p, _f64p,
                                                    _f64p, _f64p, _f64p]
            lib.pds_oracle_photon_loss_f64.argtypes = [_f32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                       ctypes.c_int32, ctypes.c_int64, ctypes.c_double,
                                                       ctypes.c_double, ctypes.c_double, _f64p, ctypes.c_double]
            lib.pds_oracle_photon_loss_f64.restype = ctypes.c_double
        else:  # libpds_cpu.so, libpds_cpu_native.so
            lib.pds_cpu_forward_f32.argtypes = [_f32p, _f32p, _f32p, _i64p, _f32p, ctypes.c_int64,
                                                ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                ctypes.c_int32, _f32p, ctypes.c_int32]
            lib.pds_cpu_backward_f32.argtypes = [_f32p, _f32p, _f32p, _f32p, _i64p, _f32p, ctypes.c_int64,
                                                 ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                 ctypes.c_int32, _f32p, _f32p, _f32p, ctypes.c_int32]
            lib.pds_cpu_threads.restype = ctypes.c_int
        _libs[name] = lib
    return _libs[name]


def _p(a: np.ndarray, ptype):
    return a.ctypes.data_as(ptype)


def _inputs(E, xp, M, cp, radius):
    E = np.ascontiguousarray(E, dtype=np.float32).reshape(-1)
    n = E.shape[0]
    xp = np.ascontiguousarray(xp, dtype=np.float32).reshape(2, n)
    M = np.ascontiguousarray(M, dtype=np.float32).reshape(2, 3, n)
    cp = np.ascontiguousarray(cp, dtype=np.int64).reshape(-1)
    radius = np.ascontiguousarray(radius, dtype=np.float32).reshape(-1)
    assert cp.shape[0] == n and radius.shape[0] == n
    return E, xp, M, cp, radius, n


def forward_f64(E, xp, M, cp, radius, size, cap) -> np.ndarray:
    """FP64 oracle forward; returns a float64 ``[C, H, W]`` image."""
    E, xp, M, cp, radius, n = _inputs(E, xp, M, cp, radius)
    C, H, W = (int(s) for s in size)
    out = np.empty((C, H, W), dtype=np.float64)
    rc = _lib("libpds_oracle.so").pds_oracle_forward_f64(
        _p(E, _f32p), _p(xp, _f32p), _p(M, _f32p), _p(cp, _i64p), _p(radius, _f32p), n, C, H, W, int(cap),
        _p(out, _f64p))
    assert rc == 0
    return out


OracleGrads = collections.namedtuple("OracleGrads", "gE gx gM aE ax aM cE cx cM")


def backward_f64(grad, E, xp, M, cp, radius, cap, idx=None) -> OracleGrads:
    """FP64 oracle backward.

    Returns ``OracleGrads(gE, gx, gM, aE, ax, aM, cE, cx, cM)`` over the photon
    subset ``idx`` (all photons if None), shaped ``[n]``, ``[2, n]``,
    ``[2, 3, n]``.  ``a*`` hold the sum of ``|term|`` of every component (the
    gradient tolerance scale), ``c*`` a first-order bound of how far an FP32
    evaluation of the same formulae can be off (large only for photons whose
    footprint barely contains a pixel).
    """
    E, xp, M, cp, radius, n_all = _inputs(E, xp, M, cp, radius)
    grad = np.ascontiguousarray(grad, dtype=np.float32)
    C, H, W = grad.shape
    if idx is None:
        n, idx_p = n_all, None
    else:
        idx = np.ascontiguousarray(idx, dtype=np.int64).reshape(-1)
        n, idx_p = idx.shape[0], _p(idx, _i64p)
    gE, aE, cE = np.empty(n), np.empty(n), np.empty(n)
    gx, ax, cx = np.empty((2, n)), np.empty((2, n)), np.empty((2, n))
    gM, aM, cM = np.empty((2, 3, n)), np.empty((2, 3, n)), np.empty((2, 3, n))
    rc = _lib("libpds_oracle.so").pds_oracle_backward_f64(
        _p(grad, _f32p), _p(E, _f32p), _p(xp, _f32p), _p(M, _f32p), _p(cp, _i64p), _p(radius, _f32p), n_all,
        C, H, W, int(cap), idx_p, n, _p(gE, _f64p), _p(gx, _f64p), _p(gM, _f64p), _p(aE, _f64p),
        _p(ax, _f64p), _p(aM, _f64p), _p(cE, _f64p), _p(cx, _f64p), _p(cM, _f64p))
    assert rc == 0
    return OracleGrads(gE, gx, gM, aE, ax, aM, cE, cx, cM)


def grads_within_tolerance(got, ref: OracleGrads, rtol=1e-3):
    """Per component: |g - g*| <= rtol * sum|terms| + FP32 conditioning bound.

    Returns ``(ok, worst)``: whether every entry passes, and the largest
    ``|g - g*| / (rtol * sum|terms| + bound)`` (ok iff worst <= 1).
    """
    worst = 0.0
    for g, r, a, c in zip(got, (ref.gE, ref.gx, ref.gM), (ref.aE, ref.ax, ref.aM), (ref.cE, ref.cx, ref.cM)):
        g = np.asarray(g, dtype=np.float64)
        err = np.abs(g - r)
        bound = rtol * a + c + 1e-300
        worst = max(worst, float(np.max(err / bound, initial=0.0)))
    return worst <= 1.0, worst


def grad_error_report(got, ref: OracleGrads, rtol=1e-3) -> dict:
    """How the gradients ``got`` compare with the oracle under the plain
    criterion |g - g*| <= rtol * sum|terms| (without the FP32 conditioning
    bound that :func:`grads_within_tolerance` adds).

    Returns ``worst_plain`` (largest |g - g*| / (rtol sum|terms|)),
    ``worst`` (the same with the bound added: ok iff <= 1) and
    ``photons_needing_bound`` (fraction of photons with an entry that passes
    only with the bound)."""
    worst_plain, worst = 0.0, 0.0
    need = None
    for g, r, a, c in zip(got, (ref.gE, ref.gx, ref.gM), (ref.aE, ref.ax, ref.aM), (ref.cE, ref.cx, ref.cM)):
        g = np.asarray(g, dtype=np.float64)
        err = np.abs(g - r)
        plain = err / (rtol * a + 1e-300)
        full = err / (rtol * a + c + 1e-300)
        worst_plain = max(worst_plain, float(np.max(plain, initial=0.0)))
        worst = max(worst, float(np.max(full, initial=0.0)))
        n = plain.shape[-1]
        flag = ((plain > 1.0) & (full <= 1.0)).reshape(-1, n).any(axis=0)
        need = flag if need is None else (need | flag)
    frac = float(need.mean()) if need is not None and need.size else 0.0
    return {"worst_plain": worst_plain, "worst": worst, "photons_needing_bound": frac}


def photon_loss_f64(grad, c, E, x, M6, radius, cap) -> float:
    """<g, F_i> of one photon with double parameters (x: 2 values, M6: 6 values)."""
    grad = np.ascontiguousarray(grad, dtype=np.float32)
    C, H, W = grad.shape
    M6 = np.ascontiguousarray(M6, dtype=np.float64).reshape(6)
    return float(_lib("libpds_oracle.so").pds_oracle_photon_loss_f64(
        _p(grad, _f32p), C, H, W, int(cap), int(c), float(E), float(x[0]), float(x[1]), _p(M6, _f64p),
        float(radius)))


def central_fd(grad, photons, i, cap, h_x, h_m_rel=1e-4, with_error=False):
    """Central finite differences of <g, F_i> for photon i.

    Returns ``(dx [2], dM [6])``: derivatives with respect to xp[:, i] (step
    ``h_x`` in normalised units) and Mp[:, :, i] (step ``h_m_rel`` times the
    largest |M| entry of the photon), all in FP64.  With ``with_error`` also
    returns ``(ex [2], eM [6])`` = |FD(h) - FD(2h)|, an estimate of the
    truncation error (large where a step crosses the kernel's support edge,
    whose second derivative jumps).
    """
    E, xp, M, cp, radius = photons
    c = int(np.asarray(cp).reshape(-1)[i])
    e = float(np.asarray(E)[i])
    x = np.asarray(xp, np.float64)[:, i].copy()
    m = np.asarray(M, np.float64)[:, :, i].reshape(6).copy()
    r = float(np.asarray(radius)[i])

    def loss(xv, mv):
        return photon_loss_f64(grad, c, e, xv, mv, r, cap)

    def fd(h_x_, h_m_):
        dx = np.zeros(2)
        for k in range(2):
            xp_, xm_ = x.copy(), x.copy()
            xp_[k] += h_x_
            xm_[k] -= h_x_
            dx[k] = (loss(xp_, m) - loss(xm_, m)) / (2 * h_x_)
        dm = np.zeros(6)
        for k in range(6):
            mp_, mm_ = m.copy(), m.copy()
            mp_[k] += h_m_
            mm_[k] -= h_m_
            dm[k] = (loss(x, mp_) - loss(x, mm_)) / (2 * h_m_)
        return dx, dm

    hm = h_m_rel * float(np.abs(m).max())
    dx, dm = fd(h_x, hm)
    if not with_error:
        return dx, dm
    dx2, dm2 = fd(2 * h_x, 2 * hm)
    return dx, dm, np.abs(dx - dx2), np.abs(dm - dm2)


_CPU_LIB = "libpds_cpu.so"
CPU_FLAGS = {"libpds_cpu.so": "gcc -O3 -march=x86-64-v3 -fopenmp",
             "libpds_cpu_native.so": "gcc -O3 -march=native -fopenmp"}


def use_native_cpu_baseline() -> str:
    """Build the CPU baseline for this host (-march=native) and time that one
    from now on; falls back to the portable build if the compiler is missing.
    Returns the compile flags of the build in use."""
    global _CPU_LIB
    try:
        subprocess.run(["make", "-C", _DIR, "-B", "libpds_cpu_native.so"], check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)
        _CPU_LIB = "libpds_cpu_native.so"
        _libs.pop(_CPU_LIB, None)
    except (OSError, subprocess.CalledProcessError):
        _CPU_LIB = "libpds_cpu.so"
    return CPU_FLAGS[_CPU_LIB]


def host_cpu() -> dict:
    """The host's CPU model, logical CPUs and the CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def cpu_threads() -> int:
    return int(_lib(_CPU_LIB).pds_cpu_threads())


def cpu_forward_f32(E, xp, M, cp, radius, size, cap, nthreads: int = 0) -> np.ndarray:
    """OpenMP FP32 host baseline forward."""
    E, xp, M, cp, radius, n = _inputs(E, xp, M, cp, radius)
    C, H, W = (int(s) for s in size)
    out = np.empty((C, H, W), dtype=np.float32)
    rc = _lib(_CPU_LIB).pds_cpu_forward_f32(
        _p(E, _f32p), _p(xp, _f32p), _p(M, _f32p), _p(cp, _i64p), _p(radius, _f32p), n, C, H, W, int(cap),
        _p(out, _f32p), int(nthreads))
    assert rc == 0
    return out


def cpu_backward_f32(grad, E, xp, M, cp, radius, cap, nthreads: int = 0):
    """OpenMP FP32 host baseline backward; returns ``(gE, gx, gM)``."""
    E, xp, M, cp, radius, n = _inputs(E, xp, M, cp, radius)
    grad = np.ascontiguousarray(grad, dtype=np.float32)
    C, H, W = grad.shape
    gE = np.empty(n, np.float32)
    gx = np.empty((2, n), np.float32)
    gM = np.empty((2, 3, n), np.float32)
    rc = _lib(_CPU_LIB).pds_cpu_backward_f32(
        _p(grad, _f32p), _p(E, _f32p), _p(xp, _f32p), _p(M, _f32p), _p(cp, _i64p), _p(radius, _f32p), n, C, H, W,
        int(cap), _p(gE, _f32p), _p(gx, _f32p), _p(gM, _f32p), int(nthreads))
    assert rc == 0
    return gE, gx, gM


# ---------------------------------------------------------------------------
# numpy twin: a second, independent restatement (offset-major, photon-vector)
# ---------------------------------------------------------------------------

def _sat_i32(d: np.ndarray) -> np.ndarray:
    """float64 -> int32 with saturation and NaN -> 0 (GPU conversion rule)."""
    d = np.where(np.isnan(d), 0.0, d)
    d = np.clip(d, -2147483648.0, 2147483647.0)
    return np.trunc(d).astype(np.int64)


def _twin_setup(E, xp, M, cp, radius, C, H, W, cap):
    """Per-photon window, photon_differentials.cu:74-89."""
    w, h = float(np.float32(W)), float(np.float32(H))
    r64 = radius.astype(np.float64)
    rx = np.minimum(_sat_i32(np.ceil(r64 * 0.5 * w)), cap)
    ry = np.minimum(_sat_i32(np.ceil(r64 * 0.5 * h)), cap)
    pxc = _sat_i32((0.5 * xp[0].astype(np.float64) + 0.5) * w)
    pyc = _sat_i32((0.5 * xp[1].astype(np.float64) + 0.5) * h)
    valid = (cp >= 0) & (cp < C) & (rx >= 0) & (ry >= 0)
    return rx, ry, pxc, pyc, valid


def _twin_offsets(rx, ry, valid, W, H):
    """Window offsets (x_off, y_off) that can be non-empty for some photon."""
    R = int(max(rx[valid].max(initial=0), ry[valid].max(initial=0)))
    for yo in range(-R, R + 1):
        for xo in range(-R, R + 1):
            # photon_differentials.cu:93, double arithmetic
            e = float(np.float32(xo * xo)) / (0.25 * W * W) + float(np.float32(yo * yo)) / (0.25 * H * H)
            if e <= 1.0:
                yield xo, yo


def twin_forward(E, xp, M, cp, radius, size, cap) -> np.ndarray:
    """numpy FP64 forward (photon_differentials.cu:60-113)."""
    E, xp, M, cp, radius, n = _inputs(E, xp, M, cp, radius)
    C, H, W = (int(s) for s in size)
    out = np.zeros(C * H * W, np.float64)
    rx, ry, pxc, pyc, valid = _twin_setup(E, xp, M, cp, radius, C, H, W, cap)
    Md = M.astype(np.float64)
    x0, x1 = xp[0].astype(np.float64), xp[1].astype(np.float64)
    for xo, yo in _twin_offsets(rx, ry, valid, W, H):
        px, py = pxc + xo, pyc + yo
        m = valid & (abs(xo) <= rx) & (abs(yo) <= ry) & (px >= 0) & (py >= 0) & (px < W) & (py < H)
        if not m.any():
            continue
        dx = 2.0 * px[m] / W - 1.0 - x0[m]
        dy = 2.0 * py[m] / H - 1.0 - x1[m]
        u = Md[0, 0, m] * dx + Md[0, 1, m] * dy + Md[0, 2, m] * 0.0
        v = Md[1, 0, m] * dx + Md[1, 1, m] * dy + Md[1, 2, m] * 0.0
        q = u * u + v * v
        k = np.where(q < 1.0, 3.0 / np.pi * (1.0 - q) ** 2, 0.0)
        val = k * E[m].astype(np.float64)
        pos = val > 0
        flat = (cp[m] * H + py[m]) * W + px[m]
        np.add.at(out, flat[pos], val[pos])
    return out.reshape(C, H, W)


def twin_backward(grad, E, xp, M, cp, radius, cap):
    """numpy FP64 backward (photon_differentials.cu:136-221), r = 0 terms -> 0."""
    E, xp, M, cp, radius, n = _inputs(E, xp, M, cp, radius)
    grad = np.asarray(grad, dtype=np.float32)
    C, H, W = grad.shape
    gflat = grad.reshape(-1).astype(np.float64)
    rx, ry, pxc, pyc, valid = _twin_setup(E, xp, M, cp, radius, C, H, W, cap)
    Md = M.astype(np.float64)
    x0, x1 = xp[0].astype(np.float64), xp[1].astype(np.float64)
    Ed = E.astype(np.float64)
    acc = np.zeros((9, n))
    for xo, yo in _twin_offsets(rx, ry, valid, W, H):
        px, py = pxc + xo, pyc + yo
        m = valid & (abs(xo) <= rx) & (abs(yo) <= ry) & (px >= 0) & (py >= 0) & (px < W) & (py < H)
        if not m.any():
            continue
        dx = 2.0 * px[m] / W - 1.0 - x0[m]
        dy = 2.0 * py[m] / H - 1.0 - x1[m]
        a, b, d, e = Md[0, 0, m], Md[0, 1, m], Md[1, 0, m], Md[1, 1, m]
        u = a * dx + b * dy + Md[0, 2, m] * 0.0
        v = d * dx + e * dy + Md[1, 2, m] * 0.0
        q = u * u + v * v
        r = np.sqrt(q)
        g = gflat[(cp[m] * H + py[m]) * W + px[m]]
        with np.errstate(invalid="ignore", divide="ignore"):
            dk = np.where(r < 1.0, -12.0 / np.pi * r * (1.0 - q), 0.0) * Ed[m] * g
            un = np.where(r == 0.0, 0.0, u / r)
            vn = np.where(r == 0.0, 0.0, v / r)
        acc[0, m] += np.where(q < 1.0, 3.0 / np.pi * (1.0 - q) ** 2, 0.0) * g
        acc[1, m] += -dk * (a * un + d * vn)
        acc[2, m] += -dk * (b * un + e * vn)
        acc[3, m] += dk * un * dx
        acc[4, m] += dk * un * dy
        acc[6, m] += dk * vn * dx
        acc[7, m] += dk * vn * dy
    gE = acc[0]
    gx = acc[1:3].copy()
    gM = acc[3:9].reshape(2, 3, n).copy()
    return gE, gx, gM
