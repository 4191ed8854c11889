"""GPU: pds_forward_noalloc_f32, the forward that allocates nothing (its
scratch is a fixed device array inside the library), against pds_forward_f32
with a caller-owned workspace on the same inputs, through the C ABI.

Two calls in a row, the second on a smaller problem than the first (the
shared scratch keeps the first call's contents, which the second must not
read), in the deterministic mode (bitwise equal: exact integer sums) and the
atomic mode (same image up to float summation order, and within the FP64
oracle's 1e-4 tolerance); a problem too large for the scratch is rejected
with PDS_ERR_CAPACITY before any GPU work.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import pds_oracle as O
from shapefromcaustics_amd import _native
from shapefromcaustics_amd import workloads as WL

pytestmark = pytest.mark.gpu


def _ptrs(*ts):
    return [ctypes.c_void_p(t.data_ptr()) for t in ts]


def _with_workspace(lib, p, mode, tun, stream):
    E, xp, M, cp, r = p.to_torch("cuda")
    C, H, W = p.size
    out = torch.empty(p.size, dtype=torch.float32, device="cuda")
    n = lib.pds_forward_workspace_bytes(p.n, C, H, W, p.cap, mode, tun)
    ws = torch.empty(max(n, 1), dtype=torch.uint8, device="cuda")
    rc = lib.pds_forward_f32(*_ptrs(E, xp, M, cp, r), p.n, C, H, W, p.cap, mode, tun, *_ptrs(out, ws), n, stream)
    _native.check(rc, "pds_forward_f32")
    return out


def _no_alloc(lib, p, mode, tun, stream):
    E, xp, M, cp, r = p.to_torch("cuda")
    C, H, W = p.size
    out = torch.full(p.size, float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    rc = lib.pds_forward_noalloc_f32(*_ptrs(E, xp, M, cp, r), p.n, C, H, W, p.cap, mode, tun, *_ptrs(out), stream)
    _native.check(rc, "pds_forward_noalloc_f32")
    return out


@pytest.mark.parametrize("mode", ["deterministic", "auto"])
def test_no_alloc_forward_equals_workspace_forward(mode):
    lib = _native.lib()
    m = _native.MODES[mode]
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    # first the larger problem, then a smaller one (fewer photons, a smaller
    # image, other bins) into the same scratch
    problems = [WL.caustic(1 << 20, (3, 512, 512), 12, seed=31), WL.plain(1 << 14, (1, 128, 128), 8, seed=32)]
    for k, p in enumerate(problems):
        need = lib.pds_forward_workspace_bytes(p.n, *p.size, p.cap, m, None)
        assert 0 < need <= _native.STATIC_SCRATCH_BYTES
        print(f"call {k}: {p.n} photons -> {p.size}, mode {mode}: {need} bytes of scratch", flush=True)
        got = _no_alloc(lib, p, m, None, stream)
        want = _with_workspace(lib, p, m, None, stream)
        torch.cuda.synchronize()
        assert torch.isfinite(got).all(), f"call {k}: image not fully written"
        if mode == "deterministic":
            assert torch.equal(got, want), f"call {k}: differs from pds_forward_f32"
        else:
            scale = float(want.abs().max())
            assert (got - want).abs().max() <= 1e-5 * scale, f"call {k}: differs from pds_forward_f32"
        ref = O.forward_f64(*p.arrays(), p.size, p.cap)
        g = got.cpu().numpy().astype(np.float64)
        mask = ref > 1e-3 * ref.max()
        assert (np.abs(g - ref)[mask] / ref[mask]).max() <= 1e-4, f"call {k}: vs the FP64 oracle"


def test_no_alloc_forward_rejects_problems_beyond_the_scratch():
    lib = _native.lib()
    N, C, H, W, cap = 1 << 24, 3, 1024, 1024, 12
    need = lib.pds_forward_workspace_bytes(N, C, H, W, cap, 0, None)
    assert need > _native.STATIC_SCRATCH_BYTES
    # sizes are checked before any pointer is used or any kernel launched
    fake = ctypes.c_void_p(16)
    rc = lib.pds_forward_noalloc_f32(fake, fake, fake, fake, fake, N, C, H, W, cap, 0, None, fake, None)
    assert rc == _native.PDS_ERR_CAPACITY
    assert "static scratch" in _native.last_error()
