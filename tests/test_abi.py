"""C ABI and Python binding checks that need no GPU: the library loads, exports
every function include/pds.h declares, and rejects bad arguments with the
documented status codes and messages before touching the device."""
import ctypes
import os
import re

import pytest
import torch

from shapefromcaustics_amd import _native
from shapefromcaustics_amd import PhotonDifferentialSplatting as PDSmod

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions(header, prefix):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from shapefromcaustics_amd import generate, heightfield
    lib = _native.lib()
    names = _declared_functions("pds.h", "pds")
    assert {"pds_forward_f32", "pds_backward_f32", "pds_forward_workspace_bytes", "pds_last_error"} <= set(names)
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) <= set(_native.SIGNATURES)
    hf = _declared_functions("hf.h", "hf")
    assert {"hf_trace_f32", "hf_simulate_f32", "hf_adjoint_f32", "hf_landweber_f32"} <= set(hf)
    for name in hf:
        assert hasattr(lib, name), name
    assert set(hf) <= set(heightfield.SIGNATURES)
    gen = _declared_functions("pds_gen.h", "pds_gen")
    assert gen and set(gen) <= set(generate.SIGNATURES)
    for name in gen:
        assert hasattr(lib, name), name


def test_heightfield_and_generator_argument_errors():
    from shapefromcaustics_amd import generate, heightfield
    hl, gl = heightfield.lib(), generate.lib()
    sc = heightfield.Scene(1, -1.0).struct()                       # R < 2
    rc = hl.hf_simulate_f32(None, None, ctypes.byref(sc), None, None, None, None, None, None, None, 0, None)
    assert rc == _native.PDS_ERR_ARGUMENT and b"R must be" in hl.hf_last_error()
    sc = heightfield.Scene(8, -1.0).struct()
    em = heightfield.ParallelEmitter(4, 1.0).struct()
    em.kind = 7
    h = ctypes.c_void_p(1)
    rc = hl.hf_adjoint_f32(h, h, ctypes.byref(sc), ctypes.byref(em), None, None, None, None, None, 0, None)
    assert rc == _native.PDS_ERR_ARGUMENT and b"kind" in hl.hf_last_error()
    rc = gl.pds_gen_caustic_f32(1, 0, 10, 64, 64, None, 4, None, None, None, None, None, 10, 5, None)
    assert rc == _native.PDS_ERR_ARGUMENT and b"invalid" in gl.pds_gen_last_error()
    assert hl.hf_workspace_bytes(1) == 0 and hl.hf_workspace_bytes(16) > 16 * 16 * 4


def test_heightfield_size_limits():
    # 3 R^2 normal components and n_side^2 * channels photons stay below 2^31
    from shapefromcaustics_amd import heightfield
    hl = heightfield.lib()
    h = ctypes.c_void_p(1)
    assert hl.hf_workspace_bytes(16384) > 0 and hl.hf_workspace_bytes(16385) == 0
    assert hl.hf_landweber_workspace_bytes(16385) == 0
    rc = hl.hf_normals_f32(h, 16385, h, None)
    assert rc == _native.PDS_ERR_ARGUMENT and b"16384" in hl.hf_last_error()
    sc = heightfield.Scene(16385, -1.0).struct()
    rc = hl.hf_simulate_f32(h, h, ctypes.byref(sc), None, None, None, None, None, None, None, 0, None)
    assert rc == _native.PDS_ERR_ARGUMENT and b"R must be" in hl.hf_last_error()
    sc = heightfield.Scene(8, -1.0, iors=(1.5, 1.5)).struct()
    em = heightfield.ParallelEmitter(32769, 1.0).struct()
    assert hl.hf_hits_bytes(ctypes.byref(sc), ctypes.byref(em)) == 0
    rc = hl.hf_simulate_f32(h, h, ctypes.byref(sc), ctypes.byref(em), h, h, h, h, h, h, 1 << 20, None)
    assert rc == _native.PDS_ERR_ARGUMENT and b"n_side" in hl.hf_last_error()
    em = heightfield.ParallelEmitter(32768, 1.0).struct()             # 2 channels: 2^31 photons
    assert hl.hf_hits_bytes(ctypes.byref(sc), ctypes.byref(em)) == 0
    ws = hl.hf_workspace_bytes(8)
    rc = hl.hf_adjoint_f32(h, h, ctypes.byref(sc), ctypes.byref(em), h, h, h, h, h, ws, None)
    assert rc == _native.PDS_ERR_CAPACITY and b"2^31" in hl.hf_last_error()


def test_version_and_empty_error():
    assert b"gfx950" in _native.lib().pds_version()


def _fwd(N=1, C=1, H=8, W=8, cap=4, mode=0, ptr=1, out=1, ws=None, wsb=0):
    p = ctypes.c_void_p(ptr) if ptr else None
    return _native.lib().pds_forward_f32(p, p, p, p, p, N, C, H, W, cap, mode, None,
                                         ctypes.c_void_p(out) if out else None, ws, wsb, None)


@pytest.mark.parametrize("kw,code,msg", [
    (dict(N=-1), _native.PDS_ERR_ARGUMENT, "negative"),
    (dict(C=-1), _native.PDS_ERR_ARGUMENT, "negative extent"),
    (dict(mode=7), _native.PDS_ERR_ARGUMENT, "unknown mode"),
    (dict(ptr=0), _native.PDS_ERR_ARGUMENT, "null photon"),
    (dict(out=0), _native.PDS_ERR_ARGUMENT, "null output"),
    (dict(C=4096, H=1024, W=1024), _native.PDS_ERR_CAPACITY, "2^31"),
])
def test_forward_argument_errors(kw, code, msg):
    assert _fwd(**kw) == code
    assert msg in _native.last_error()


def test_forward_workspace_too_small():
    need = _native.lib().pds_forward_workspace_bytes(1 << 16, 3, 256, 256, 8, _native.MODES["binned"], None)
    if need == 0:
        pytest.skip("binned path needs no workspace")
    assert _fwd(N=1 << 16, C=3, H=256, W=256, cap=8, mode=2, ws=None, wsb=0) == _native.PDS_ERR_WORKSPACE
    assert "workspace" in _native.last_error()


def test_empty_image_is_a_no_op():
    # C * H * W == 0: nothing to launch, success without touching pointers
    assert _fwd(C=0, out=0) == _native.PDS_OK


def test_backward_argument_errors():
    lib = _native.lib()
    p = ctypes.c_void_p(1)
    rc = lib.pds_backward_f32(p, p, p, p, p, p, 4, 1, 8, 8, 4, 0, None, None, p, p, None, 0, None)
    assert rc == _native.PDS_ERR_ARGUMENT and "gradient output" in _native.last_error()
    rc = lib.pds_backward_f32(None, p, p, p, p, p, 4, 1, 8, 8, 4, 0, None, p, p, p, None, 0, None)
    assert rc == _native.PDS_ERR_ARGUMENT and "upstream" in _native.last_error()


# ------------------------------------------------------------------ Python binding (CPU tensors)


def _photons(n=5, dtype=torch.float32):
    return (torch.ones(n, dtype=dtype), torch.zeros(2, n, dtype=dtype), torch.zeros(2, 3, n, dtype=dtype),
            torch.zeros(1, n, dtype=torch.int64), torch.ones(n, dtype=dtype))


def test_python_requires_cuda_tensors_like_reference():
    # CHECK_INPUT order and messages of PhotonDifferentialSplattig.cpp:14-18 / utils.hpp:37-41
    with pytest.raises(RuntimeError, match="^Ep must be a CUDA tensor$"):
        PDSmod.pds_forward(*_photons(), (1, 8, 8), 4)
    with pytest.raises(RuntimeError, match="^grad_pds must be a CUDA tensor$"):
        PDSmod.pds_backward(torch.zeros(1, 8, 8), *_photons(), 4)


@pytest.mark.parametrize("mutate,msg", [
    (lambda p: (p[0], p[1][:, :4].contiguous(), p[2], p[3], p[4]), "Dimensions of tensors for pds_forward don't match"),
    (lambda p: (p[0], p[1], p[2], p[3].repeat(2, 1), p[4]), "Channel index must be one dimensional!"),
    (lambda p: (p[0], p[1].repeat(2, 1), p[2], p[3], p[4]), "Position index must be two dimensional!"),
    (lambda p: (p[0], p[1], p[2][:, :2].contiguous(), p[3], p[4]), "Change of basis matrix must be 2x3!"),
    (lambda p: (p[0], p[1].t().contiguous().t(), p[2], p[3], p[4]), "xp must be contiguous"),
])
def test_python_shape_checks_match_reference_messages(mutate, msg):
    with pytest.raises(RuntimeError) as ei:
        PDSmod._check_photons(*mutate(_photons()), require_cuda=False)
    assert str(ei.value) == msg


def test_python_rejects_other_dtypes():
    with pytest.raises(TypeError, match="float32"):
        PDSmod._check_photons(*_photons(dtype=torch.float64), require_cuda=False)
    p = list(_photons())
    p[3] = p[3].int()
    with pytest.raises(TypeError, match="int64"):
        PDSmod._check_photons(*p, require_cuda=False)


def test_python_unknown_mode():
    with pytest.raises(ValueError, match="unknown mode"):
        PDSmod._mode("fastest")


def test_tuning_fields_are_checked():
    t = _native.tuning_struct({"tile": 32, "block_photons": 4096})
    assert (t.tile, t.block_photons, t.seg_photons) == (32, 4096, 0)
    with pytest.raises(ValueError, match="unknown pds_tuning field 'layout'"):
        _native.tuning_struct({"layout": 1})
    with pytest.raises(ValueError, match="unknown tuning"):
        PDSmod._resolve_tuning("fastest", 1, (1, 8, 8), 4, "auto")


def test_alias_module_name():
    import sys

    import shapefromcaustics_amd as pkg
    pkg.install_alias()
    import PhotonDifferentialSplatting as alias  # noqa: E402
    assert alias.pds_forward is PDSmod.pds_forward and "PhotonDifferentialSplatting" in sys.modules
