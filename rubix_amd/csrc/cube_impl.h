// Internal entry points of the fused path (cube.hip), shared with the batched
// entry points (batch.hip).
#pragma once

#include "rbx_common.h"

namespace rbx {

// Launch tuning resolved from rbx_tuning (NULL -> defaults).
struct Tune {
  int chunk;    // stars per reduction chunk of cube_accum
  int waves;    // wavefronts per workgroup of star_norm / star_bwd
  int persist;  // 0 = one workgroup per work item, k = k workgroups per CU
  int flags;    // RBX_TUNE_TWO_PASS: no one-pass fused forward
};

int tune_of(const rbx_tuning* t, Tune* out);

size_t cube_workspace_impl(int64_t n, int32_t n_keys, int32_t nt, const Tune& t);

// ext may be NULL (no dust extinction).
int cube_fwd_impl(const rbx_ssp* ssp, const rbx_target* target, const rbx_stars* stars,
                  const rbx_extinction* ext, const Tune& t, float* cube, void* workspace,
                  size_t workspace_bytes, hipStream_t s, const char* who);

int cube_bwd_impl(const rbx_ssp* ssp, const rbx_target* target, const rbx_stars* stars,
                  const rbx_extinction* ext, const Tune& t, const float* dcube, float* dmass,
                  float* dage, float* dmet, float* dvlos, void* workspace,
                  size_t workspace_bytes, hipStream_t s, const char* who);

}  // namespace rbx
