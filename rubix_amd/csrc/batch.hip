// Batched ragged mode: many galaxies with uneven star counts and spaxel
// grids in one pass of the fused path.  The galaxies' stars are
// concatenated; each star's own spaxel id becomes a batch-wide key
// (galaxy's first cube row + id, or the drop key), so the stable sort, the
// per-star kernels and the chunked cube reduction run once over the whole
// batch and write every galaxy's cube rows of one batch cube.  The per-star
// work does not depend on the galaxy, and chunks never cross a spaxel, so a
// galaxy's gradients are bitwise those of a single-galaxy run with the same
// tuning, and so is its cube when both runs use the same number of age
// bands (a batch's key count gives it one band; a single galaxy gets more
// only for template grids above one band's budget, and then its cube
// differs only by the order of the f64 chunk additions, which add float32
// terms exactly while a spaxel's terms span less than about 2^29).
#include "cube_impl.h"
#include "rbx_common.h"

namespace rbx {

// key[r] = spaxel_begin + ids[r] of star r's galaxy (the last galaxy with
// star_begin <= r), or n_rows (dropped) when the id is outside the galaxy's
// nb x nb grid, the grid does not fit the batch cube, or r precedes every
// galaxy.
__global__ void batch_keys_kernel(const int32_t* __restrict__ ids, int64_t n,
                                  const rbx_galaxy* __restrict__ gal, int n_gal, int n_rows,
                                  int32_t* __restrict__ keys) {
  for (int64_t r = gtid(); r < n; r += int64_t(gridDim.x) * blockDim.x) {
    int lo = 0, hi = n_gal;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (gal[mid].star_begin <= r) lo = mid; else hi = mid;
    }
    const rbx_galaxy G = gal[lo];
    const int64_t id = ids[r];
    const int64_t cells = int64_t(G.nb) * G.nb;
    const bool ok = r >= G.star_begin && G.nb > 0 && id >= 0 && id < cells &&
                    G.spaxel_begin >= 0 && int64_t(G.spaxel_begin) + cells <= n_rows;
    keys[r] = ok ? int32_t(G.spaxel_begin + id) : n_rows;
  }
}

static size_t keys_bytes(int64_t n) { return align256(size_t(n > 0 ? n : 1) * sizeof(int32_t)); }

static int batch_check(const rbx_stars* stars, const rbx_galaxy* gal, int32_t n_gal,
                       const char* who) {
  if (!stars || !gal || n_gal < 1) return failf(RBX_EINVAL, "%s: need a galaxy table", who);
  if (stars->n_spaxel < 1 || stars->n_spaxel > INT32_MAX - 2)
    return failf(RBX_EINVAL, "%s: n_spaxel must be the batch cube's row count", who);
  return RBX_OK;
}

// The star descriptor the fused path sees: batch keys as ids, one drop key.
static rbx_stars keyed(const rbx_stars* stars, int32_t* keys) {
  rbx_stars k = *stars;
  k.ids = keys;
  k.n_keys = stars->n_spaxel + 1;
  return k;
}

}  // namespace rbx

extern "C" {

size_t rbx_batch_workspace_size(int64_t n, int32_t n_spaxel, int32_t n_target,
                                const rbx_tuning* tune) {
  using namespace rbx;
  Tune t;
  if (n < 0 || n_spaxel < 1 || n_spaxel > INT32_MAX - 2 || tune_of(tune, &t)) return 0;
  const size_t cube = cube_workspace_impl(n, n_spaxel + 1, n_target, t);
  return cube ? keys_bytes(n) + cube : 0;
}

int rbx_batch_fwd(const rbx_ssp* ssp, const rbx_target* target, const rbx_stars* stars,
                  const rbx_galaxy* galaxies, int32_t n_gal, const rbx_tuning* tune,
                  float* cube, void* workspace, size_t workspace_bytes, void* stream) {
  using namespace rbx;
  int rc = batch_check(stars, galaxies, n_gal, "rbx_batch_fwd");
  if (rc) return rc;
  Tune t;
  if ((rc = tune_of(tune, &t))) return rc;
  const size_t kb = keys_bytes(stars->n);
  if (!workspace || workspace_bytes < kb)
    return fail(RBX_EWORKSPACE, "rbx_batch_fwd: workspace too small");
  if (stars->n > 0 && !stars->ids) return fail(RBX_EINVAL, "rbx_batch_fwd: null ids");
  hipStream_t s = as_stream(stream);
  int32_t* keys = static_cast<int32_t*>(workspace);
  if (stars->n > 0) {
    int64_t blocks = (stars->n + 255) / 256;
    if (blocks > (1 << 20)) blocks = 1 << 20;
    hipLaunchKernelGGL(batch_keys_kernel, dim3(blocks), dim3(256), 0, s, stars->ids, stars->n,
                       galaxies, n_gal, stars->n_spaxel, keys);
  }
  const rbx_stars k = keyed(stars, keys);
  return cube_fwd_impl(ssp, target, &k, nullptr, t, cube, static_cast<char*>(workspace) + kb,
                       workspace_bytes - kb, s, "rbx_batch_fwd");
}

int rbx_batch_bwd(const rbx_ssp* ssp, const rbx_target* target, const rbx_stars* stars,
                  const rbx_galaxy* galaxies, int32_t n_gal, const rbx_tuning* tune,
                  const float* dcube, float* dmass, float* dage, float* dmet, float* dvlos,
                  void* workspace, size_t workspace_bytes, void* stream) {
  using namespace rbx;
  int rc = batch_check(stars, galaxies, n_gal, "rbx_batch_bwd");
  if (rc) return rc;
  Tune t;
  if ((rc = tune_of(tune, &t))) return rc;
  const size_t kb = keys_bytes(stars->n);
  if (!workspace || workspace_bytes < kb)
    return fail(RBX_EWORKSPACE, "rbx_batch_bwd: workspace too small");
  const rbx_stars k = keyed(stars, static_cast<int32_t*>(workspace));
  return cube_bwd_impl(ssp, target, &k, nullptr, t, dcube, dmass, dage, dmet, dvlos,
                       static_cast<char*>(workspace) + kb, workspace_bytes - kb,
                       as_stream(stream), "rbx_batch_bwd");
}

}  // extern "C"
