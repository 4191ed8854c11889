"""numpy twin of the counter-based caustic-line photon source
(shapefromcaustics_amd/csrc/gen_kernels.hip).

TEST INFRASTRUCTURE ONLY: the checker of the GPU generator.  Integer work
(SplitMix64 draws, 24-bit uniforms, line choice, channels) is bit-exact;
float results follow the same float32 formulas, so they agree with the GPU to
the last-ulp differences of the transcendental functions.
"""
import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(z: np.ndarray) -> np.ndarray:
    """SplitMix64 (Steele, Lea, Flood, OOPSLA 2014) on uint64 arrays."""
    with np.errstate(over="ignore"):
        z = (z + np.uint64(0x9E3779B97F4A7C15)) & M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & M64
        return z ^ (z >> np.uint64(31))


def u24(bits: np.ndarray) -> np.ndarray:
    return (bits >> np.uint64(40)).astype(np.uint32).astype(np.float32) * np.float32(1.0 / 16777216.0)


def caustic(seed, view, n, W, H, lines):
    """(E, xp, M, cp, radius) of photons 0..n-1 of ``view`` (reference layout)."""
    f = np.float32
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * np.uint64(0x100000001B3)) & M64
    key ^= np.uint64(view) << np.uint64(56)
    i = np.arange(n, dtype=np.uint64)
    ctr = i << np.uint64(4)
    r = [splitmix64(key ^ (ctr | np.uint64(k))) for k in range(5)]
    sh = np.uint64(24)
    u_line, u_k = u24(r[0]), u24((r[0] << sh) & M64)
    u1, u2 = u24(r[1]), u24((r[1] << sh) & M64)
    ux, uy = u24(r[2]), u24((r[2] << sh) & M64)
    u_th, u_s = u24(r[3]), u24((r[3] << sh) & M64)
    u_kap, u_e = u24(r[4]), u24((r[4] << sh) & M64)
    Wf, Hf = f(W), f(H)
    k = np.minimum((u_k * f(lines.shape[0])).astype(np.int64), lines.shape[0] - 1)
    L = lines[k]
    rad = np.sqrt(f(-2.0) * np.log(f(1.0) - u1))
    tp = f(6.28318530717958648)
    z1, z2 = rad * np.cos(tp * u2), rad * np.sin(tp * u2)
    a, b = L[:, 3] * z1, (L[:, 3] / f(20.0)) * z2
    c, s = np.cos(L[:, 2]), np.sin(L[:, 2])
    on = u_line >= f(0.2)
    px = np.where(on, L[:, 0] + c * a - s * b, ux * Wf).astype(f)
    py = np.where(on, L[:, 1] + s * a + c * b, uy * Hf).astype(f)
    px = np.minimum(np.maximum(px, f(0)), Wf - f(1e-3))
    py = np.minimum(np.maximum(py, f(0)), Hf - f(1e-3))
    theta = f(3.14159265358979324) * u_th
    s1 = f(0.3) * np.exp(u_s * np.log(f(20.0)))
    s2 = s1 / (f(1.0) + f(9.0) * u_kap)
    ct, st = np.cos(theta), np.sin(theta)
    M = np.zeros((2, 3, n), np.float32)
    M[0, 0], M[0, 1] = ct / s1 * (f(0.5) * Wf), st / s1 * (f(0.5) * Hf)
    M[1, 0], M[1, 1] = -st / s2 * (f(0.5) * Wf), ct / s2 * (f(0.5) * Hf)
    E = ((f(0.5) + u_e) / f(n)).astype(f)
    xp = np.stack([f(2.0) * px / Wf - f(1.0), f(2.0) * py / Hf - f(1.0)]).astype(f)
    radius = (f(2.0) * np.maximum(s1, s2) / Wf).astype(f)
    cp = (3 * view + (np.arange(n) % 3)).astype(np.int64)[None]
    return E, xp, M, cp, radius
