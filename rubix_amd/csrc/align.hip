// Steps before the datacube path: aperture filter and galaxy alignment.
//
// filter_particles (rubix/core/telescope.py:129-201) masks particles outside
// the aperture, mask = x, y within [min edge, max edge] inclusive
// (rubix/telescope/utils.py:155-178), and sets every per-particle field to 0
// outside (jnp.where(mask, value, 0)).
// rotate_galaxy (rubix/galaxy/alignment.py:240-265) rotates positions and
// velocities into the principal frame of the moment of inertia tensor of the
// particles within the half-mass radius (alignment.py:66-120), then by the
// Euler rotation Rz Ry Rx (alignment.py:140-205); positions are row vectors,
// p' = (p R) E.  The inertia tensor is a fixed-order double reduction; its
// 3x3 eigendecomposition is left to the host (three numbers).
#include "rbx_common.h"

namespace rbx {

constexpr int kAlignThreads = 256;
constexpr int kAlignBlocksMax = 2048;

__device__ inline float mul_rn(float x, float y) {
#pragma clang fp contract(off)
  return x * y;
}

__device__ inline float add_rn(float x, float y) {
#pragma clang fp contract(off)
  return x + y;
}

__global__ void aperture_filter_kernel(const float* __restrict__ coords, int64_t n, int64_t stride,
                                       float lo, float hi, const float* __restrict__ in0,
                                       const float* __restrict__ in1, const float* __restrict__ in2,
                                       float* __restrict__ out0, float* __restrict__ out1,
                                       float* __restrict__ out2, uint8_t* __restrict__ mask) {
  for (int64_t i = gtid(); i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const float x = coords[i * stride], y = coords[i * stride + 1];
    const bool m = (x >= lo) && (x <= hi) && (y >= lo) && (y <= hi);
    if (mask) mask[i] = m ? 1 : 0;
    if (out0) out0[i] = m ? in0[i] : 0.f;
    if (out1) out1[i] = m ? in1[i] : 0.f;
    if (out2) out2[i] = m ? in2[i] : 0.f;
  }
}

__global__ void apply_mask_kernel(const uint8_t* __restrict__ mask, int64_t n, int64_t width,
                                  const float* __restrict__ in, float* __restrict__ out) {
  const int64_t total = n * width;
  for (int64_t i = gtid(); i < total; i += int64_t(gridDim.x) * blockDim.x)
    out[i] = mask[i / width] ? in[i] : 0.f;
}

// Block partials of the six independent entries of
// I = sum_{|r| <= R} m (|r|^2 delta_ij - r_i r_j), products and sums in double.
// |r| is evaluated in float32 as the reference does: sqrt((x^2 + y^2) + z^2).
__global__ void __launch_bounds__(kAlignThreads) inertia_partial_kernel(
    const float* __restrict__ coords, const float* __restrict__ mass, int64_t n, int64_t stride,
    float r_half, double* __restrict__ partial) {
  __shared__ double sm[6][kAlignThreads];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t i = gtid(); i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const float x = coords[i * stride], y = coords[i * stride + 1], z = coords[i * stride + 2];
    const float r2 = add_rn(add_rn(mul_rn(x, x), mul_rn(y, y)), mul_rn(z, z));
    if (!(__fsqrt_rn(r2) <= r_half)) continue;
    const double m = mass[i], dx = x, dy = y, dz = z;
    const double q = dx * dx + dy * dy + dz * dz;
    acc[0] += m * (q - dx * dx);  // xx
    acc[1] += m * (q - dy * dy);  // yy
    acc[2] += m * (q - dz * dz);  // zz
    acc[3] -= m * dx * dy;        // xy
    acc[4] -= m * dx * dz;        // xz
    acc[5] -= m * dy * dz;        // yz
  }
  for (int c = 0; c < 6; ++c) sm[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int off = kAlignThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      for (int c = 0; c < 6; ++c) sm[c][threadIdx.x] += sm[c][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x < 6) partial[int64_t(blockIdx.x) * 6 + threadIdx.x] = sm[threadIdx.x][0];
}

// Sums the block partials in block order and writes the symmetric 3x3.
__global__ void inertia_final_kernel(const double* __restrict__ partial, int nblk,
                                     double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int b = 0; b < nblk; ++b)
    for (int c = 0; c < 6; ++c) s[c] += partial[int64_t(b) * 6 + c];
  out[0] = s[0]; out[1] = s[3]; out[2] = s[4];
  out[3] = s[3]; out[4] = s[1]; out[5] = s[5];
  out[6] = s[4]; out[7] = s[5]; out[8] = s[2];
}

// out_i = (in_i R) E for row vectors, each product in float32 in the order
// of a 3-term dot (((a0 b0) + a1 b1) + a2 b2) with no fused multiply-add.
__global__ void rotate_rows_kernel(const float* __restrict__ in, int64_t n, int64_t stride,
                                   rbx_rotation rot, float* __restrict__ out) {
  for (int64_t i = gtid(); i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const float* p = in + i * stride;
    float t[3], o[3];
    for (int j = 0; j < 3; ++j)
      t[j] = add_rn(add_rn(mul_rn(p[0], rot.first[j]), mul_rn(p[1], rot.first[3 + j])),
                    mul_rn(p[2], rot.first[6 + j]));
    for (int j = 0; j < 3; ++j)
      o[j] = add_rn(add_rn(mul_rn(t[0], rot.second[j]), mul_rn(t[1], rot.second[3 + j])),
                    mul_rn(t[2], rot.second[6 + j]));
    float* q = out + i * stride;
    for (int j = 0; j < 3; ++j) q[j] = o[j];
    for (int64_t j = 3; j < stride; ++j) q[j] = p[j];
  }
}

static int grid_of(int64_t n, int threads, int cap) {
  int64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return int(g);
}

static int inertia_blocks(int64_t n) { return grid_of(n, kAlignThreads * 4, kAlignBlocksMax); }

}  // namespace rbx

extern "C" {

int rbx_filter_particles(const float* coords, int64_t n, int64_t stride, float lo, float hi,
                         const float* mass, const float* age, const float* metallicity,
                         float* mass_out, float* age_out, float* metallicity_out,
                         uint8_t* mask_out, void* stream) {
  using namespace rbx;
  if (n < 0 || stride < 2 || (n > 0 && !coords) || (mass_out && !mass) || (age_out && !age) ||
      (metallicity_out && !metallicity))
    return fail(RBX_EINVAL, "rbx_filter_particles: bad arguments");
  if (n == 0) return RBX_OK;
  hipLaunchKernelGGL(aperture_filter_kernel, dim3(grid_of(n, 256, 1 << 20)), dim3(256), 0,
                     as_stream(stream), coords, n, stride, lo, hi, mass, age, metallicity,
                     mass_out, age_out, metallicity_out, mask_out);
  return check_launch("rbx_filter_particles");
}

int rbx_apply_mask(const uint8_t* mask, int64_t n, int64_t width, const float* in, float* out,
                   void* stream) {
  using namespace rbx;
  if (n < 0 || width < 1 || (n > 0 && (!mask || !in || !out)))
    return fail(RBX_EINVAL, "rbx_apply_mask: bad arguments");
  if (n == 0) return RBX_OK;
  hipLaunchKernelGGL(apply_mask_kernel, dim3(grid_of(n * width, 256, 1 << 20)), dim3(256), 0,
                     as_stream(stream), mask, n, width, in, out);
  return check_launch("rbx_apply_mask");
}

size_t rbx_inertia_workspace_size(int64_t n) {
  if (n < 0) return 0;
  return size_t(rbx::inertia_blocks(n)) * 6 * sizeof(double);
}

int rbx_inertia_tensor(const float* coords, const float* mass, int64_t n, int64_t stride,
                       float halfmass_radius, double* out9, void* workspace,
                       size_t workspace_bytes, void* stream) {
  using namespace rbx;
  if (n < 0 || stride < 3 || !out9 || (n > 0 && (!coords || !mass)))
    return fail(RBX_EINVAL, "rbx_inertia_tensor: bad arguments");
  const int nblk = inertia_blocks(n);
  if (!workspace || workspace_bytes < size_t(nblk) * 6 * sizeof(double))
    return fail(RBX_EWORKSPACE, "rbx_inertia_tensor: workspace too small");
  hipStream_t s = as_stream(stream);
  double* partial = static_cast<double*>(workspace);
  hipLaunchKernelGGL(inertia_partial_kernel, dim3(nblk), dim3(kAlignThreads), 0, s, coords, mass,
                     n, stride, halfmass_radius, partial);
  hipLaunchKernelGGL(inertia_final_kernel, dim3(1), dim3(64), 0, s, partial, nblk, out9);
  return check_launch("rbx_inertia_tensor");
}

int rbx_rotate_rows(const float* in, int64_t n, int64_t stride, const rbx_rotation* rot,
                    float* out, void* stream) {
  using namespace rbx;
  if (n < 0 || stride < 3 || !rot || (n > 0 && (!in || !out)))
    return fail(RBX_EINVAL, "rbx_rotate_rows: bad arguments");
  if (in == out) return fail(RBX_EINVAL, "rbx_rotate_rows: in-place rotation not supported");
  if (n == 0) return RBX_OK;
  hipLaunchKernelGGL(rotate_rows_kernel, dim3(grid_of(n, 256, 1 << 20)), dim3(256), 0,
                     as_stream(stream), in, n, stride, *rot, out);
  return check_launch("rbx_rotate_rows");
}

}  // extern "C"
