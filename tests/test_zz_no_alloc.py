"""GPU: rbx_conv_axis_dot_noalloc (no workspace, block partials in the
library's own device array) against rbx_conv_axis_dot with a caller
workspace, both called through the C ABI on the same inputs.  Two calls in a
row, the second smaller than the first (fewer block partials than the
array holds from the first call), for each of the dot's kernel paths: LSF
rows, tiled PSF lines, long lines and the strided fallback.  The same
kernels run in the same order, so the results are bitwise equal."""
import pytest
import torch

from rubix_amd import _lib

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _dots(n_outer, n_axis, n_inner, n_taps, seed):
    L = _lib.load()
    g = torch.Generator(device=DEV).manual_seed(seed)
    total = n_outer * n_axis * n_inner
    a = torch.randn(total, generator=g, device=DEV)
    b = torch.randn(total, generator=g, device=DEV)
    taps = torch.randn(n_taps, generator=g, device=DEV, dtype=torch.float64)
    start = (n_taps - 1) // 2
    ws_bytes = L.rbx_dot_workspace_size(total)
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=DEV)
    r_ws = torch.full((1,), float("nan"), dtype=torch.float64, device=DEV)
    r_na = torch.full((1,), float("nan"), dtype=torch.float64, device=DEV)
    s = _lib.stream_handle()
    rc = L.rbx_conv_axis_dot(_lib.ptr(a), _lib.ptr(b), n_outer, n_axis, n_inner, _lib.ptr(taps),
                             n_taps, start, _lib.ptr(r_ws), _lib.ptr(ws), ws_bytes, s)
    assert rc == 0, L.rbx_last_error_string()
    rc = L.rbx_conv_axis_dot_noalloc(_lib.ptr(a), _lib.ptr(b), n_outer, n_axis, n_inner,
                                     _lib.ptr(taps), n_taps, start, _lib.ptr(r_na), s)
    assert rc == 0, L.rbx_last_error_string()
    torch.cuda.synchronize()
    return ws_bytes // 8, r_ws.item(), r_na.item()


@pytest.mark.parametrize("big,small", [
    # LSF rows: 2500 spectra x 3721 bins, then 100 x 1000
    ((2500, 3721, 1, 25), (100, 1000, 1, 25)),
    # tiled PSF lines (axis <= 240): x pass of 50 x 50 x 3721, then 20 x 20 x 300
    ((50, 50, 3721, 15), (20, 20, 300, 11)),
    # long lines (axis > 240): 1 x 300 x (300 x 64), then 1 x 260 x 500
    ((1, 300, 300 * 64, 7), (1, 260, 500, 5)),
    # strided fallback (inner 1, more taps than the row kernel's halo)
    ((64, 2000, 1, 301), (8, 700, 1, 301)),
])
def test_noalloc_dot_equals_workspace_dot(big, small):
    blocks = []
    for k, shape in enumerate((big, small)):
        nblk, r_ws, r_na = _dots(*shape, seed=10 + k)
        blocks.append(nblk)
        assert r_ws == r_na, (shape, r_ws, r_na)
        assert r_ws == r_ws  # finite / not NaN
    assert blocks[1] < blocks[0]


def test_noalloc_dot_rejects_bad_arguments():
    L = _lib.load()
    rc = L.rbx_conv_axis_dot_noalloc(None, None, 1, 1, 1, None, 1, 0, None, None)
    assert rc == -1
    assert b"bad arguments" in L.rbx_last_error_string()
