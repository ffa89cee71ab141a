"""GPU parity: spaxel assignment and the sort-by-spaxel index (bit-exact)."""
import numpy as np
import pytest
import torch

from oracle import ref_numpy as R
from rubix_amd import ops, synthetic as S

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _assign(coords, edges):
    ids = ops.spaxel_assign(torch.as_tensor(coords, device=DEV), torch.as_tensor(edges, device=DEV))
    return ids.cpu().numpy()


def test_known_answer():
    # tests/test_telescope_utils.py:20-35 of the reference
    coords = np.array([[0.5, 1.5], [2.5, 3.5]], np.float32)
    edges = np.array([0, 1, 2, 3, 4], np.float32)
    np.testing.assert_array_equal(_assign(coords, edges), [4, 14])


@pytest.mark.parametrize("nb,n", [(25, 10_000), (50, 1_000_000), (150, 200_000)])
def test_matches_oracle_bit_exact(nb, n):
    st = S.make_stars(n, seed=nb)
    L = S.angular_scale_kpc_per_arcsec(0.1) * nb * 0.2
    edges = np.linspace(-L / 2, L / 2, nb + 1).astype(np.float32)
    got = _assign(st["coords"], edges)
    np.testing.assert_array_equal(got, R.square_spaxel_assignment(st["coords"], edges))


def test_edge_cases():
    edges = np.array([-1.0, -0.5, 0.0, 0.5, 1.0], np.float32)
    xs = np.array([-2, -1, -0.75, -0.5, 0, 0.25, 0.5, 1.0, 1.5, np.nan, np.inf, -np.inf,
                   np.nextafter(np.float32(0.5), np.float32(0))], np.float32)
    xx, yy = np.meshgrid(xs, xs, indexing="ij")
    coords = np.stack([xx.ravel(), yy.ravel(), np.zeros(xx.size)], 1).astype(np.float32)
    np.testing.assert_array_equal(_assign(coords, edges), R.square_spaxel_assignment(coords, edges))


def test_nb_plus_two_edges():
    # the float-arange quirk of rubix/telescope/utils.py:34-36 can give nb+2 edges;
    # the stride then follows len(edges)-1 exactly as the reference does
    edges, _ = R.spatial_bin_edges(1.0, 1.0, 10)
    st = S.make_stars(5000, seed=3, h_r=0.3)
    coords = st["coords"]
    np.testing.assert_array_equal(_assign(coords, edges), R.square_spaxel_assignment(coords, edges))


def test_empty():
    out = ops.spaxel_assign(torch.zeros((0, 3), device=DEV), torch.tensor([0.0, 1.0], device=DEV))
    assert out.numel() == 0


@pytest.mark.parametrize("n,n_keys", [(0, 4), (1, 4), (63, 5), (4097, 10), (100_000, 2500),
                                      (1_000_000, 2500), (300_000, 22500)])
def test_sort_matches_stable_argsort(n, n_keys):
    rng = np.random.default_rng(n + n_keys)
    ids = rng.integers(0, n_keys, n).astype(np.int32)
    if n > 10:
        ids[: n // 3] = ids[0]  # long runs of one key
    perm, srt, off = ops.spaxel_sort(torch.as_tensor(ids, device=DEV), n_keys)
    eperm, eoff = R.sort_by_spaxel(ids, n_keys)
    np.testing.assert_array_equal(perm.cpu().numpy(), eperm)
    np.testing.assert_array_equal(srt.cpu().numpy(), ids[eperm])
    np.testing.assert_array_equal(off.cpu().numpy(), eoff)


@pytest.mark.parametrize("n,n_keys", [(1_000_000, 100_000), (16_000_000, 20_000_000)])
def test_sort_radix_passes_match_stable_argsort(n, n_keys):
    """Key counts whose (bucket, tile) count table would pass 2^24 entries
    sort in LSD radix passes (2 passes of 16 + 1 bits for 1 M stars and
    100 000 keys; 3 of 12 + 12 + 1 bits for 16 M stars and 2e7 keys), with
    out-of-range ids (negative and >= n_keys) among them: the same stable
    order, original ids and offsets as the one-pass sort."""
    rng = np.random.default_rng(n % 1000 + 11)
    ids = rng.integers(0, n_keys, n).astype(np.int32)
    ids[: n // 4] = ids[n // 2]  # a long run of one key
    bad = rng.integers(0, n, n // 50)
    ids[bad[: len(bad) // 2]] = -1
    ids[bad[len(bad) // 2:]] = n_keys + 3
    perm, srt, off = ops.spaxel_sort(torch.as_tensor(ids, device=DEV), n_keys)
    buckets = np.where((ids >= 0) & (ids < n_keys), ids, n_keys)
    eperm, eoff = R.sort_by_spaxel(buckets, n_keys)
    np.testing.assert_array_equal(perm.cpu().numpy(), eperm)
    np.testing.assert_array_equal(srt.cpu().numpy(), ids[eperm])
    np.testing.assert_array_equal(off.cpu().numpy(), eoff)


def test_sort_out_of_range_ids_go_last_in_order():
    ids = np.array([3, 7, -1, 2, 9, 3, 0, 5], np.int32)  # n_keys = 5: 7, -1, 9, 5 overflow
    perm, srt, off = ops.spaxel_sort(torch.as_tensor(ids, device=DEV), 5)
    perm = perm.cpu().numpy()
    np.testing.assert_array_equal(perm, [6, 3, 0, 5, 1, 2, 4, 7])
    np.testing.assert_array_equal(off.cpu().numpy(), [0, 1, 1, 2, 4, 4])


def test_sort_deterministic():
    rng = np.random.default_rng(7)
    ids = torch.as_tensor(rng.integers(0, 2500, 500_000).astype(np.int32), device=DEV)
    a = ops.spaxel_sort(ids, 2500)
    b = ops.spaxel_sort(ids, 2500)
    for x, y in zip(a, b):
        assert torch.equal(x, y)
