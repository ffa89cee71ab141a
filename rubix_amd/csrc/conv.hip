// PSF and LSF convolutions over the (ny, nx, n_lambda) datacube.
//
// convolve_psf (rubix/telescope/psf/psf.py:33-58) convolves every lambda plane
// with a 2D kernel, mode "same"; the Gaussian kernel of
// rubix/telescope/psf/kernels.py:5-31 is outer(g, g) with g the normalised 1D
// Gaussian, so it runs as two 1D passes (x then y).  convolve_lsf
// (rubix/telescope/lsf/lsf.py:29-105) is a 25-tap full convolution along
// lambda cropped at the extend factor.  Both are one kernel: a 1D
// convolution along the middle axis of an (n_outer, n_axis, n_inner) view.
// Threads map to the flattened output in row-major order, so loads and
// stores are contiguous along lambda (the innermost, contiguous axis) for
// the spatial passes and along the convolved axis for the LSF.  The cube is
// 37 MB at 50x50x3721: these passes are HBM-bound and small next to the
// per-star kernels.
#include "rbx_common.h"

namespace rbx {

__global__ void conv_axis_kernel(const float* __restrict__ in, float* __restrict__ out,
                                 int64_t n_outer, int n_axis, int64_t n_inner,
                                 const float* __restrict__ taps, int n_taps, int start) {
  const int64_t total = n_outer * n_axis * n_inner;
  for (int64_t idx = gtid(); idx < total; idx += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = idx % n_inner;
    const int64_t ox = idx / n_inner;
    const int x = int(ox % n_axis);
    const int64_t o = ox / n_axis;
    const float* src = in + o * n_axis * n_inner + i;
    float acc = 0.f;
    for (int t = 0; t < n_taps; ++t) {
      const int xs = x + start - t;
      if (xs >= 0 && xs < n_axis) acc += taps[t] * src[int64_t(xs) * n_inner];
    }
    out[idx] = acc;
  }
}

__global__ void conv2d_planes_kernel(const float* __restrict__ in, float* __restrict__ out,
                                     int ny, int nx, int nl, const float* __restrict__ K,
                                     int ky, int kx, int sy, int sx) {
  const int64_t total = int64_t(ny) * nx * nl;
  for (int64_t idx = gtid(); idx < total; idx += int64_t(gridDim.x) * blockDim.x) {
    const int l = int(idx % nl);
    const int64_t yx = idx / nl;
    const int x = int(yx % nx);
    const int y = int(yx / nx);
    float acc = 0.f;
    for (int a = 0; a < ky; ++a) {
      const int ys = y + sy - a;
      if (ys < 0 || ys >= ny) continue;
      for (int b = 0; b < kx; ++b) {
        const int xs = x + sx - b;
        if (xs < 0 || xs >= nx) continue;
        acc += K[a * kx + b] * in[(int64_t(ys) * nx + xs) * nl + l];
      }
    }
    out[idx] = acc;
  }
}

constexpr int kDotThreads = 256;
constexpr int kDotBlocksMax = 4096;

// Per-block partial of sum(a * conv_axis(b)) in double, fixed tree order.
// The sigma-derivative taps sum to zero, so conv(dtaps, b) cancels strongly:
// taps and the convolution are double.
__global__ void __launch_bounds__(kDotThreads) conv_dot_partial_kernel(
    const float* __restrict__ a, const float* __restrict__ b, int64_t n_outer, int n_axis,
    int64_t n_inner, const double* __restrict__ taps, int n_taps, int start,
    double* __restrict__ partial) {
  __shared__ double sm[kDotThreads];
  const int64_t total = n_outer * n_axis * n_inner;
  double acc = 0.0;
  for (int64_t idx = gtid(); idx < total; idx += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = idx % n_inner;
    const int64_t ox = idx / n_inner;
    const int x = int(ox % n_axis);
    const int64_t o = ox / n_axis;
    const float* src = b + o * n_axis * n_inner + i;
    double c = 0.0;
    for (int t = 0; t < n_taps; ++t) {
      const int xs = x + start - t;
      if (xs >= 0 && xs < n_axis) c += taps[t] * double(src[int64_t(xs) * n_inner]);
    }
    acc += double(a[idx]) * c;
  }
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kDotThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sm[threadIdx.x] += sm[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
}

__global__ void __launch_bounds__(kDotThreads) dot_final_kernel(const double* __restrict__ partial,
                                                                int nblk, double* __restrict__ result) {
  __shared__ double sm[kDotThreads];
  double acc = 0.0;
  for (int k = threadIdx.x; k < nblk; k += kDotThreads) acc += partial[k];
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kDotThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sm[threadIdx.x] += sm[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *result = sm[0];
}

// ------------------------------------------------------------ batches
//
// A batch cube stacks the (nb, nb, n_lambda) cubes of several galaxies row
// after row (rbx_galaxy: first row and nb of each).  One launch convolves
// every galaxy: each output element finds its galaxy by binary search over
// the (ascending) first rows, then runs the same 1D convolution inside that
// galaxy's grid with zero padding at its edges.

struct BatchPos {
  bool ok;      // the row belongs to a galaxy whose grid fits the batch
  int64_t base; // element of coordinate u = 0 along the convolved axis
  int64_t stride;
  int u, nb;
};

__device__ inline BatchPos batch_pos(const rbx_galaxy* __restrict__ gal, int n_gal, int n_rows,
                                     int nt, int axis, int64_t idx) {
  BatchPos b;
  const int p = int(idx / nt);
  const int l = int(idx - int64_t(p) * nt);
  int lo = 0, hi = n_gal;  // last galaxy with spaxel_begin <= p
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (gal[mid].spaxel_begin <= p) lo = mid; else hi = mid;
  }
  const int64_t begin = gal[lo].spaxel_begin;
  const int nb = gal[lo].nb;
  const int64_t q = p - begin;
  b.nb = nb;
  b.ok = nb > 0 && q >= 0 && q < int64_t(nb) * nb && begin + int64_t(nb) * nb <= n_rows;
  if (!b.ok) {
    b.base = 0; b.stride = 0; b.u = 0;
    return b;
  }
  const int y = int(q / nb), x = int(q - int64_t(y) * nb);
  if (axis == 1) {
    b.u = x;
    b.stride = nt;
    b.base = (begin + int64_t(y) * nb) * nt + l;
  } else {
    b.u = y;
    b.stride = int64_t(nb) * nt;
    b.base = (begin + x) * nt + l;
  }
  return b;
}

__global__ void conv_batch_kernel(const float* __restrict__ in, float* __restrict__ out,
                                  const rbx_galaxy* __restrict__ gal, int n_gal, int n_rows,
                                  int nt, int axis, const float* __restrict__ taps, int n_taps,
                                  int start) {
  const int64_t total = int64_t(n_rows) * nt;
  for (int64_t idx = gtid(); idx < total; idx += int64_t(gridDim.x) * blockDim.x) {
    const BatchPos b = batch_pos(gal, n_gal, n_rows, nt, axis, idx);
    float acc = 0.f;
    if (b.ok) {
      for (int t = 0; t < n_taps; ++t) {
        const int us = b.u + start - t;
        if (us >= 0 && us < b.nb) acc += taps[t] * in[b.base + us * b.stride];
      }
    }
    out[idx] = acc;
  }
}

__global__ void __launch_bounds__(kDotThreads) conv_batch_dot_partial_kernel(
    const float* __restrict__ a, const float* __restrict__ bsrc,
    const rbx_galaxy* __restrict__ gal, int n_gal, int n_rows, int nt, int axis,
    const double* __restrict__ taps, int n_taps, int start, double* __restrict__ partial) {
  __shared__ double sm[kDotThreads];
  const int64_t total = int64_t(n_rows) * nt;
  double acc = 0.0;
  for (int64_t idx = gtid(); idx < total; idx += int64_t(gridDim.x) * blockDim.x) {
    const BatchPos b = batch_pos(gal, n_gal, n_rows, nt, axis, idx);
    if (!b.ok) continue;
    double c = 0.0;
    for (int t = 0; t < n_taps; ++t) {
      const int us = b.u + start - t;
      if (us >= 0 && us < b.nb) c += taps[t] * double(bsrc[b.base + us * b.stride]);
    }
    acc += double(a[idx]) * c;
  }
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kDotThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sm[threadIdx.x] += sm[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
}

static int grid_for(int64_t n, int threads, int64_t cap) {
  int64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return int(g);
}

static int dot_blocks(int64_t total) { return grid_for(total, kDotThreads * 8, kDotBlocksMax); }

int conv_axis(const float* in, float* out, int64_t n_outer, int n_axis, int64_t n_inner,
              const float* taps, int n_taps, int start, hipStream_t s) {
  const int64_t total = n_outer * n_axis * n_inner;
  if (total == 0) return RBX_OK;
  hipLaunchKernelGGL(conv_axis_kernel, dim3(grid_for(total, 256, 1 << 20)), dim3(256), 0, s, in,
                     out, n_outer, n_axis, n_inner, taps, n_taps, start);
  return check_launch("rbx_conv_axis");
}

}  // namespace rbx

extern "C" {

int rbx_conv_axis(const float* in, float* out, int64_t n_outer, int32_t n_axis, int64_t n_inner,
                  const float* taps, int32_t n_taps, int32_t start, void* stream) {
  using namespace rbx;
  if (!in || !out || !taps || n_outer < 0 || n_axis < 1 || n_inner < 1 || n_taps < 1)
    return fail(RBX_EINVAL, "rbx_conv_axis: bad arguments");
  if (in == out) return fail(RBX_EINVAL, "rbx_conv_axis: in-place convolution not supported");
  return conv_axis(in, out, n_outer, n_axis, n_inner, taps, n_taps, start, as_stream(stream));
}

int rbx_psf_sep(const float* in, float* out, float* tmp, int32_t ny, int32_t nx, int32_t nl,
                const float* gy, int32_t my, int32_t start_y, const float* gx, int32_t mx,
                int32_t start_x, void* stream) {
  using namespace rbx;
  if (!in || !out || !tmp || !gy || !gx || ny < 1 || nx < 1 || nl < 1 || my < 1 || mx < 1)
    return fail(RBX_EINVAL, "rbx_psf_sep: bad arguments");
  if (in == out || in == tmp || out == tmp)
    return fail(RBX_EINVAL, "rbx_psf_sep: in, out and tmp must be distinct");
  hipStream_t s = as_stream(stream);
  int rc = conv_axis(in, tmp, ny, nx, nl, gx, mx, start_x, s);
  if (rc) return rc;
  return conv_axis(tmp, out, 1, ny, int64_t(nx) * nl, gy, my, start_y, s);
}

int rbx_conv2d_planes(const float* in, float* out, int32_t ny, int32_t nx, int32_t nl,
                      const float* kernel, int32_t ky, int32_t kx, int32_t sy, int32_t sx,
                      void* stream) {
  using namespace rbx;
  if (!in || !out || !kernel || ny < 1 || nx < 1 || nl < 1 || ky < 1 || kx < 1 || in == out)
    return fail(RBX_EINVAL, "rbx_conv2d_planes: bad arguments");
  const int64_t total = int64_t(ny) * nx * nl;
  hipLaunchKernelGGL(conv2d_planes_kernel, dim3(grid_for(total, 256, 1 << 20)), dim3(256), 0,
                     as_stream(stream), in, out, ny, nx, nl, kernel, ky, kx, sy, sx);
  return check_launch("rbx_conv2d_planes");
}

size_t rbx_dot_workspace_size(int64_t total) {
  if (total < 0) return 0;
  return size_t(rbx::dot_blocks(total)) * sizeof(double);
}

int rbx_conv_axis_dot(const float* a, const float* b, int64_t n_outer, int32_t n_axis,
                      int64_t n_inner, const double* taps, int32_t n_taps, int32_t start,
                      double* result, void* workspace, size_t workspace_bytes, void* stream) {
  using namespace rbx;
  if (!a || !b || !taps || !result || n_outer < 0 || n_axis < 1 || n_inner < 1 || n_taps < 1)
    return fail(RBX_EINVAL, "rbx_conv_axis_dot: bad arguments");
  const int64_t total = n_outer * n_axis * n_inner;
  const int nblk = dot_blocks(total);
  if (!workspace || workspace_bytes < size_t(nblk) * sizeof(double))
    return fail(RBX_EWORKSPACE, "rbx_conv_axis_dot: workspace too small");
  hipStream_t s = as_stream(stream);
  double* partial = static_cast<double*>(workspace);
  hipLaunchKernelGGL(conv_dot_partial_kernel, dim3(nblk), dim3(kDotThreads), 0, s, a, b, n_outer,
                     n_axis, n_inner, taps, n_taps, start, partial);
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kDotThreads), 0, s, partial, nblk, result);
  return check_launch("rbx_conv_axis_dot");
}

int rbx_conv_batch(const float* in, float* out, const rbx_galaxy* galaxies, int32_t n_gal,
                   int32_t n_rows, int32_t n_target, int32_t axis, const float* taps,
                   int32_t n_taps, int32_t start, void* stream) {
  using namespace rbx;
  if (!in || !out || !galaxies || !taps || n_gal < 1 || n_rows < 0 || n_target < 1 ||
      n_taps < 1 || (axis != 0 && axis != 1))
    return fail(RBX_EINVAL, "rbx_conv_batch: bad arguments");
  if (in == out) return fail(RBX_EINVAL, "rbx_conv_batch: in-place convolution not supported");
  const int64_t total = int64_t(n_rows) * n_target;
  if (total == 0) return RBX_OK;
  hipLaunchKernelGGL(conv_batch_kernel, dim3(grid_for(total, 256, 1 << 20)), dim3(256), 0,
                     as_stream(stream), in, out, galaxies, n_gal, n_rows, n_target, axis, taps,
                     n_taps, start);
  return check_launch("rbx_conv_batch");
}

int rbx_conv_batch_dot(const float* a, const float* b, const rbx_galaxy* galaxies, int32_t n_gal,
                       int32_t n_rows, int32_t n_target, int32_t axis, const double* taps,
                       int32_t n_taps, int32_t start, double* out, void* workspace,
                       size_t workspace_bytes, void* stream) {
  using namespace rbx;
  if (!a || !b || !galaxies || !taps || !out || n_gal < 1 || n_rows < 0 || n_target < 1 ||
      n_taps < 1 || (axis != 0 && axis != 1))
    return fail(RBX_EINVAL, "rbx_conv_batch_dot: bad arguments");
  const int64_t total = int64_t(n_rows) * n_target;
  const int nblk = dot_blocks(total);
  if (!workspace || workspace_bytes < size_t(nblk) * sizeof(double))
    return fail(RBX_EWORKSPACE, "rbx_conv_batch_dot: workspace too small");
  hipStream_t s = as_stream(stream);
  double* partial = static_cast<double*>(workspace);
  hipLaunchKernelGGL(conv_batch_dot_partial_kernel, dim3(nblk), dim3(kDotThreads), 0, s, a, b,
                     galaxies, n_gal, n_rows, n_target, axis, taps, n_taps, start, partial);
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kDotThreads), 0, s, partial, nblk, out);
  return check_launch("rbx_conv_batch_dot");
}

}  // extern "C"
