"""The C-ABI library loads, exports every symbol include/rubix_amd.h declares,
and rejects bad arguments with status codes (no GPU needed: argument
validation happens before any HIP call)."""
import ctypes
import os
import re

import pytest

from rubix_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "rubix_amd.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rbx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared()
    assert len(names) >= 14
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.SIGNATURES, f"{n} has no ctypes signature"


def test_struct_layouts_match_static_asserts():
    # cube.hip static_asserts sizeof(rbx_ssp) == 80, rbx_target == 40, rbx_stars == 80
    assert ctypes.sizeof(_lib.RbxSsp) == 80
    assert ctypes.sizeof(_lib.RbxTarget) == 40
    assert ctypes.sizeof(_lib.RbxStars) == 80
    # batch.hip / conv.hip index rbx_galaxy arrays on the device: 16 bytes each
    assert ctypes.sizeof(_lib.RbxGalaxy) == 16
    assert ctypes.sizeof(_lib.RbxTuning) == 16
    assert ctypes.sizeof(_lib.RbxNoiseStats) == 16  # noise.hip static_assert
    assert ctypes.sizeof(_lib.RbxExtinction) == 16


def test_version_and_errors():
    L = _lib.load()
    assert L.rbx_version() >= 100
    rc = L.rbx_spaxel_assign(None, 10, 1, None, 1, None, None)
    assert rc == -1
    assert b"rbx_spaxel_assign" in L.rbx_last_error_string()
    with pytest.raises(_lib.RbxError, match="rbx_spaxel_sort"):
        _lib.call("rbx_spaxel_sort", None, 5, 0, None, None, None, None, 0, None)
    with pytest.raises(_lib.RbxError, match="null descriptor"):
        _lib.call("rbx_cube_fwd", None, None, None, None, None, 0, None)
    s = _lib.RbxSsp()
    s.n_met, s.n_age, s.n_wave = 1, 5, 10
    t = _lib.RbxTarget()
    st = _lib.RbxStars()
    with pytest.raises(_lib.RbxError, match=">= 2 metallicities"):
        _lib.call("rbx_cube_fwd", s, t, st, None, None, 0, None)
    with pytest.raises(_lib.RbxError, match="in-place"):
        _lib.call("rbx_conv_axis", 8, 8, 1, 4, 1, 16, 3, 1, None)


def test_batch_arguments_are_validated_before_any_gpu_call():
    L = _lib.load()
    bad = _lib.RbxTuning(3, 0, 0, 0)  # chunk below 16
    assert L.rbx_batch_workspace_size(1000, 100, 50, bad) == 0
    assert L.rbx_batch_workspace_size(1000, 100, 50, None) > L.rbx_cube_workspace_size(1000, 101, 50)
    with pytest.raises(_lib.RbxError, match="waves must be"):
        st = _lib.RbxStars()
        st.n_spaxel = 4
        _lib.call("rbx_batch_fwd", _lib.RbxSsp(), _lib.RbxTarget(), st, 64, 1,
                  _lib.RbxTuning(0, 3, 0, 0), None, None, 0, None)
    with pytest.raises(_lib.RbxError, match="galaxy table"):
        _lib.call("rbx_batch_bwd", _lib.RbxSsp(), _lib.RbxTarget(), _lib.RbxStars(), None, 0, None,
                  None, None, None, None, None, None, 0, None)
    with pytest.raises(_lib.RbxError, match="rbx_conv_batch"):
        _lib.call("rbx_conv_batch", 8, 16, 24, 1, 4, 10, 2, 32, 3, 1, None)


def test_noise_arguments_are_validated_before_any_gpu_call():
    L = _lib.load()
    assert L.rbx_noise_workspace_size(2500) == 2500 * 8
    assert L.rbx_noise_workspace_size(0) == 0
    with pytest.raises(_lib.RbxError, match="unknown noise distribution 2"):
        _lib.call("rbx_noise_sample", 10, 0, 2, 8, None)
    with pytest.raises(_lib.RbxError, match="empty datacube"):
        _lib.call("rbx_apply_noise", 8, 0, 10, 8, 0, 0, 8, 8, 8, 8, None)
    with pytest.raises(_lib.RbxError, match="unknown noise distribution"):
        _lib.call("rbx_apply_noise", 8, 4, 10, 8, 0, 7, 8, 8, 8, 8, None)
    with pytest.raises(_lib.RbxError, match="null pointer"):
        _lib.call("rbx_apply_noise", 8, 4, 10, None, 0, 0, 8, 8, 8, 8, None)
    with pytest.raises(_lib.RbxError, match="workspace of 8 bytes, 32 needed"):
        _lib.call("rbx_apply_noise_bwd", 8, 8, 4, 10, 8, 8, 8, 0, 1, 8, 8, 8, None)


def test_workspace_queries():
    L = _lib.load()
    a = L.rbx_cube_workspace_size(1000, 100, 50)
    b = L.rbx_cube_workspace_size(100000, 100, 50)
    assert 0 < a < b
    assert L.rbx_cube_workspace_size(-1, 100, 50) == 0
    assert L.rbx_spaxel_sort_workspace_size(10, 4) > 0
    # one counting pass while (keys + 1) x (4096-star tiles) fits 2^24
    # counters; past that, radix passes keep three n-element index buffers
    one = L.rbx_spaxel_sort_workspace_size(10 ** 6, 2500)
    radix = L.rbx_spaxel_sort_workspace_size(10 ** 6, 100_000)
    assert one < 4 * 2 ** 22 and radix >= 12 * 10 ** 6
    assert L.rbx_spaxel_sort_workspace_size(32 * 10 ** 6, 310_000) < 4 * (2 ** 24 + 3 * 32 * 10 ** 6) + 2 ** 20
    assert L.rbx_dot_workspace_size(10 ** 7) >= 8


def test_ops_refuse_cpu_tensors():
    import torch

    from rubix_amd import ops

    with pytest.raises(RuntimeError, match="no CPU path"):
        ops.spaxel_assign(torch.zeros(4, 3), torch.tensor([0.0, 1.0]))
