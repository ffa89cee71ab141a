// PSF and LSF convolutions over the (ny, nx, n_lambda) datacube.
//
// convolve_psf (rubix/telescope/psf/psf.py:33-58) convolves every lambda plane
// with a 2D kernel, mode "same"; the Gaussian kernel of
// rubix/telescope/psf/kernels.py:5-31 is outer(g, g) with g the normalised 1D
// Gaussian, so it runs as two 1D passes (x then y).  convolve_lsf
// (rubix/telescope/lsf/lsf.py:29-105) is a 25-tap full convolution along
// lambda cropped at the extend factor.  Both are one kernel: a 1D
// convolution along the middle axis of an (n_outer, n_axis, n_inner) view.
// Work is split so that index arithmetic happens once per workgroup item,
// never per element (64-bit division per element would dominate):
// * conv_tile_kernel (n_inner > 1 and an axis of at most kTileAxis, the
//   spatial PSF passes): item = a block of 64 contiguous inner elements with
//   the whole axis staged in LDS, so each input is read once.
// * conv_lines_kernel (n_inner > 1, longer axes): item = one (outer, axis)
//   line and a chunk of up to kLineChunk contiguous inner elements; threads
//   run along the contiguous inner axis.
// * conv_rows_kernel (n_inner == 1, the LSF along lambda): item = a segment
//   of kSeg outputs of one row; the segment's inputs plus the kernel's halo
//   are staged in LDS once and every output reads its taps from LDS.
// The batched PSF (rbx_conv_batch) uses row items of the batch cube: one
// spaxel row and a chunk of lambda, its galaxy found once per item.  Every
// output sums taps in the same order t = 0 .. m-1 as the flat definition.
#include <mutex>

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
constexpr int kLineChunk = 1024;  // inner elements per line item
constexpr int kSeg = 2048;        // outputs per row item of the LSF
constexpr int kHalo = 256;        // taps - 1 supported by the row kernel
constexpr int kTileAxis = 240;    // longest axis the tiled line kernels stage (60 KB of LDS)

// One 1D convolution item of a line: out[o, x, i0 + k] for k < len.
template <class Acc, class Tap>
__device__ inline Acc line_tap_sum(const float* __restrict__ src, int x, int n_axis,
                                   int64_t n_inner, const Tap* __restrict__ taps, int n_taps,
                                   int start) {
  Acc acc = 0;
  for (int t = 0; t < n_taps; ++t) {
    const int xs = x + start - t;
    if (xs >= 0 && xs < n_axis) acc += Acc(taps[t]) * Acc(src[int64_t(xs) * n_inner]);
  }
  return acc;
}

// Items: (o, x, chunk) with chunk fastest.  Grid-strided.
__global__ void __launch_bounds__(256) conv_lines_kernel(
    const float* __restrict__ in, float* __restrict__ out, int64_t n_outer, int n_axis,
    int64_t n_inner, const float* __restrict__ taps, int n_taps, int start, int64_t n_chunks,
    int64_t n_items) {
  for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int64_t c = item % n_chunks;
    const int64_t ox = item / n_chunks;
    const int x = int(ox % n_axis);
    const int64_t o = ox / n_axis;
    const int64_t i0 = c * kLineChunk;
    const int64_t i1 = i0 + kLineChunk < n_inner ? i0 + kLineChunk : n_inner;
    const float* base = in + o * n_axis * n_inner;
    float* dst = out + (o * n_axis + x) * n_inner;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x)
      dst[i] = line_tap_sum<float>(base + i, x, n_axis, n_inner, taps, n_taps, start);
  }
}

// Rows u0, u0 + 4, u0 + 8, u0 + 12 of the convolution of a staged tile (rows
// 0 .. n-1 of 64 lanes): for each output the taps t = 0 .. m-1 in order,
// rows outside [0, n) skipped (zero padding).  Four independent sums per
// lane hide the LDS latency; each output's own order is unchanged.
template <class Acc, class Tap>
__device__ inline void tile_conv4(const float* tile, int lane, int n, const Tap* __restrict__ taps,
                                  int m, int start, int u0, Acc acc[4]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) acc[r] = Acc(0);
  for (int t = 0; t < m; ++t) {
    const Acc tv = Acc(taps[t]);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int us = u0 + 4 * r + start - t;
      if (us >= 0 && us < n) acc[r] += tv * Acc(tile[us * 64 + lane]);
    }
  }
}

// Items: (o, block of 64 contiguous inner elements).  The workgroup stages
// the block's whole axis (n_axis x 64 floats, dynamic LDS) once with
// coalesced loads; wavefront w then forms the outputs x = w, w + 4, ...
// from LDS, the taps in the order t = 0 .. m-1 of line_tap_sum (same
// values).  Every input element is read from memory once instead of once
// per tap.  n_axis <= kTileAxis.
__global__ void __launch_bounds__(256) conv_tile_kernel(
    const float* __restrict__ in, float* __restrict__ out, int n_axis, int64_t n_inner,
    const float* __restrict__ taps, int n_taps, int start, int64_t n_blocks, int64_t n_items) {
  extern __shared__ float tile[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int64_t o = item / n_blocks;
    const int64_t i = (item - o * n_blocks) * 64 + lane;
    const bool ok = i < n_inner;
    const int64_t base = o * n_axis * n_inner + (ok ? i : 0);
    __syncthreads();  // the previous item's reads are done
    for (int x = w; x < n_axis; x += 4) tile[x * 64 + lane] = ok ? in[base + int64_t(x) * n_inner] : 0.f;
    __syncthreads();
    for (int x0 = w; x0 < n_axis; x0 += 16) {
      float acc[4];
      tile_conv4<float>(tile, lane, n_axis, taps, n_taps, start, x0, acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int x = x0 + 4 * r;
        if (ok && x < n_axis) out[base + int64_t(x) * n_inner] = acc[r];
      }
    }
  }
}

// Items: (row o, segment) of a row-contiguous convolution (n_inner == 1).
// lds[k] = in[o, x0 + start - (m - 1) + k] (0 outside the row), so
// out[o, x0 + v] = sum_t taps[t] lds[v + m - 1 - t].
__global__ void __launch_bounds__(256) conv_rows_kernel(
    const float* __restrict__ in, float* __restrict__ out, int n_axis,
    const float* __restrict__ taps, int n_taps, int start, int64_t n_seg, int64_t n_items) {
  __shared__ float lds[kSeg + kHalo];
  for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int64_t o = item / n_seg;
    const int x0 = int(item - o * n_seg) * kSeg;
    const float* row = in + o * n_axis;
    const int off = x0 + start - (n_taps - 1);
    const int len = kSeg + n_taps - 1;
    __syncthreads();
    for (int k = threadIdx.x; k < len; k += blockDim.x) {
      const int xs = off + k;
      lds[k] = (xs >= 0 && xs < n_axis) ? row[xs] : 0.f;
    }
    __syncthreads();
    // lds is zero outside the row, so every tap reads it without a bounds
    // test: a tap past the row's end adds taps[t] * 0
    // outputs v, v + 256, v + 512, v + 768 of the thread as four independent
    // sums (each in the order t = 0 .. m-1), hiding the LDS latency
    for (int v0 = threadIdx.x; v0 < kSeg; v0 += 4 * 256) {
      float acc[4] = {0.f, 0.f, 0.f, 0.f};
      for (int t = 0; t < n_taps; ++t) {
        const float tv = taps[t];
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[k] += tv * lds[v0 + 256 * k + n_taps - 1 - t];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int v = v0 + 256 * k;
        if (v < kSeg && x0 + v < n_axis) out[o * n_axis + x0 + v] = acc[k];
      }
    }
  }
}

// Fixed-order block reduction of one double per thread into partial[block].
__device__ inline void block_partial(double acc, double* sm, double* __restrict__ partial) {
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kDotThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sm[threadIdx.x] += sm[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
}

// sum(a * conv(b)) per block, line items (n_inner > 1), double taps.
__global__ void __launch_bounds__(kDotThreads) conv_dot_lines_kernel(
    const float* __restrict__ a, const float* __restrict__ b, int64_t n_outer, int n_axis,
    int64_t n_inner, const double* __restrict__ taps, int n_taps, int start, int64_t n_chunks,
    int64_t n_items, double* __restrict__ partial) {
  __shared__ double sm[kDotThreads];
  double acc = 0.0;
  for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int64_t c = item % n_chunks;
    const int64_t ox = item / n_chunks;
    const int x = int(ox % n_axis);
    const int64_t o = ox / n_axis;
    const int64_t i0 = c * kLineChunk;
    const int64_t i1 = i0 + kLineChunk < n_inner ? i0 + kLineChunk : n_inner;
    const float* base = b + o * n_axis * n_inner;
    const float* av = a + (o * n_axis + x) * n_inner;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x)
      acc += double(av[i]) * line_tap_sum<double>(base + i, x, n_axis, n_inner, taps, n_taps, start);
  }
  block_partial(acc, sm, partial);
}

// sum(a * conv(b)) per block, rows (n_inner == 1), double taps.
__global__ void __launch_bounds__(kDotThreads) conv_dot_rows_kernel(
    const float* __restrict__ a, const float* __restrict__ b, int n_axis,
    const double* __restrict__ taps, int n_taps, int start, int64_t n_seg, int64_t n_items,
    double* __restrict__ partial) {
  __shared__ float lds[kSeg + kHalo];
  __shared__ double sm[kDotThreads];
  double acc = 0.0;
  for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int64_t o = item / n_seg;
    const int x0 = int(item - o * n_seg) * kSeg;
    const float* row = b + o * n_axis;
    const int off = x0 + start - (n_taps - 1);
    const int len = kSeg + n_taps - 1;
    __syncthreads();
    for (int k = threadIdx.x; k < len; k += blockDim.x) {
      const int xs = off + k;
      lds[k] = (xs >= 0 && xs < n_axis) ? row[xs] : 0.f;
    }
    __syncthreads();
    for (int v0 = threadIdx.x; v0 < kSeg; v0 += 4 * 256) {
      double c[4] = {0.0, 0.0, 0.0, 0.0};
      for (int t = 0; t < n_taps; ++t) {
        const double tv = taps[t];
#pragma unroll
        for (int k = 0; k < 4; ++k) c[k] += tv * double(lds[v0 + 256 * k + n_taps - 1 - t]);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int v = v0 + 256 * k;
        if (v < kSeg && x0 + v < n_axis) acc += double(a[o * n_axis + x0 + v]) * c[k];
      }
    }
  }
  block_partial(acc, sm, partial);
}

// sum(a * conv(b)) per block over tiled items (as conv_tile_kernel), double
// taps.  Each thread adds its outputs' a * conv(b) in item order.
__global__ void __launch_bounds__(kDotThreads) conv_dot_tile_kernel(
    const float* __restrict__ a, const float* __restrict__ b, int n_axis, int64_t n_inner,
    const double* __restrict__ taps, int n_taps, int start, int64_t n_blocks, int64_t n_items,
    double* __restrict__ partial) {
  extern __shared__ float tile[];
  __shared__ double sm[kDotThreads];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double acc = 0.0;
  for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int64_t o = item / n_blocks;
    const int64_t i = (item - o * n_blocks) * 64 + lane;
    const bool ok = i < n_inner;
    const int64_t base = o * n_axis * n_inner + (ok ? i : 0);
    __syncthreads();
    for (int x = w; x < n_axis; x += 4) tile[x * 64 + lane] = ok ? b[base + int64_t(x) * n_inner] : 0.f;
    __syncthreads();
    if (ok) {
      for (int x0 = w; x0 < n_axis; x0 += 16) {
        double c[4];
        tile_conv4<double>(tile, lane, n_axis, taps, n_taps, start, x0, c);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int x = x0 + 4 * r;
          if (x < n_axis) acc += double(a[base + int64_t(x) * n_inner]) * c[r];
        }
      }
    }
  }
  block_partial(acc, sm, partial);
}

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

// Row items: (spaxel row p of the batch cube, chunk of kLineChunk lambda);
// the galaxy and the row's (y, x) are found once per item.
__global__ void __launch_bounds__(256) conv_batch_kernel(
    const float* __restrict__ in, float* __restrict__ out, const rbx_galaxy* __restrict__ gal,
    int n_gal, int n_rows, int nt, int axis, const float* __restrict__ taps, int n_taps,
    int start, int n_chunks, int64_t n_items) {
  for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int p = int(item / n_chunks);
    const int l0 = int(item - int64_t(p) * n_chunks) * kLineChunk;
    const int l1 = l0 + kLineChunk < nt ? l0 + kLineChunk : nt;
    const BatchPos b = batch_pos(gal, n_gal, n_rows, nt, axis, int64_t(p) * nt);
    for (int l = l0 + threadIdx.x; l < l1; l += blockDim.x) {
      float acc = 0.f;
      if (b.ok) {
        for (int t = 0; t < n_taps; ++t) {
          const int us = b.u + start - t;
          if (us >= 0 && us < b.nb) acc += taps[t] * in[b.base + l + us * b.stride];
        }
      }
      out[int64_t(p) * nt + l] = acc;
    }
  }
}

// LDS-tiled batch passes.  A packed table (galaxy g's rows begin where
// galaxy g - 1's end, from row 0 to n_rows, every nb >= 1: what a batch
// layout builds) is convolved line by line: item = (line L, block of 64
// lambda), L counting the lines of galaxy 0, then galaxy 1, ...; the
// workgroup stages the line's nb x 64 inputs in LDS once (lines of at most
// kBatchTileAxis spaxels; longer ones read the inputs directly) and adds
// the taps in the order of conv_batch_kernel, so the values are the same.
// Each workgroup first builds the line prefix of the table in LDS and
// checks that it is packed; any other table runs the row items of
// conv_batch_kernel (rows of no galaxy get zeros).
constexpr int kBatchTileAxis = 100;
constexpr int kBatchGalMax = 512;  // galaxies the line prefix holds (host falls back above)

// lines[g] = number of lines of galaxies 0 .. g-1 (lines[n_gal] = total) and
// whether the table is packed; all threads of the workgroup return the same.
__device__ inline bool batch_line_prefix(const rbx_galaxy* __restrict__ gal, int n_gal,
                                         int n_rows, int* lines, int* part) {
  const int tid = threadIdx.x;
  const int per = (n_gal + 255) / 256;
  const int g0 = tid * per;
  int sum = 0, bad = 0;
  for (int k = 0; k < per; ++k) {
    const int g = g0 + k;
    if (g >= n_gal) break;
    const int64_t nb = gal[g].nb, begin = gal[g].spaxel_begin;
    const int64_t end = g + 1 < n_gal ? int64_t(gal[g + 1].spaxel_begin) : int64_t(n_rows);
    if (nb < 1 || begin + nb * nb != end || (g == 0 && begin != 0)) bad = 1;
    sum += nb > 0 ? int(nb) : 0;
  }
  part[tid] = sum;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {  // inclusive scan of the 256 partial sums
    const int v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int run = part[tid] - sum;
  for (int k = 0; k < per; ++k) {
    const int g = g0 + k;
    if (g >= n_gal) break;
    lines[g] = run;
    run += gal[g].nb > 0 ? gal[g].nb : 0;
  }
  if (tid == 255) lines[n_gal] = part[255];
  const bool packed = __syncthreads_or(bad) == 0;
  return packed;
}

// Galaxy g of line L of a packed table and the line's first element / stride.
struct BatchLine {
  int nb;
  int64_t base, stride;
};

__device__ inline BatchLine batch_line(const rbx_galaxy* __restrict__ gal, const int* lines,
                                       int n_gal, int nt, int axis, int L) {
  int lo = 0, hi = n_gal;  // last galaxy with lines[g] <= L
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (lines[mid] <= L) lo = mid; else hi = mid;
  }
  const int j = L - lines[lo];
  BatchLine b;
  b.nb = gal[lo].nb;
  const int64_t begin = gal[lo].spaxel_begin;
  if (axis == 1) {  // along x: line y = j
    b.base = (begin + int64_t(j) * b.nb) * nt;
    b.stride = nt;
  } else {          // along y: line x = j
    b.base = (begin + j) * int64_t(nt);
    b.stride = int64_t(b.nb) * nt;
  }
  return b;
}

// conv_batch_kernel's row item (row p, lambda chunk c) for tables that are
// not packed.
__device__ inline void batch_row_item(const float* __restrict__ in, float* __restrict__ out,
                                      const rbx_galaxy* __restrict__ gal, int n_gal, int n_rows,
                                      int nt, int axis, const float* __restrict__ taps,
                                      int n_taps, int start, int n_chunks, int64_t item) {
  const int p = int(item / n_chunks);
  const int l0 = int(item - int64_t(p) * n_chunks) * kLineChunk;
  const int l1 = l0 + kLineChunk < nt ? l0 + kLineChunk : nt;
  const BatchPos b = batch_pos(gal, n_gal, n_rows, nt, axis, int64_t(p) * nt);
  for (int l = l0 + threadIdx.x; l < l1; l += blockDim.x) {
    float acc = 0.f;
    if (b.ok) {
      for (int t = 0; t < n_taps; ++t) {
        const int us = b.u + start - t;
        if (us >= 0 && us < b.nb) acc += taps[t] * in[b.base + l + us * b.stride];
      }
    }
    out[int64_t(p) * nt + l] = acc;
  }
}

__global__ void __launch_bounds__(256) conv_batch_tile_kernel(
    const float* __restrict__ in, float* __restrict__ out, const rbx_galaxy* __restrict__ gal,
    int n_gal, int n_rows, int nt, int axis, const float* __restrict__ taps, int n_taps,
    int start) {
  extern __shared__ float tile[];
  __shared__ int lines[kBatchGalMax + 1];
  __shared__ int part[256];
  const bool packed = batch_line_prefix(gal, n_gal, n_rows, lines, part);
  if (!packed) {
    const int n_chunks = (nt + kLineChunk - 1) / kLineChunk;
    const int64_t n_items = int64_t(n_rows) * n_chunks;
    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x)
      batch_row_item(in, out, gal, n_gal, n_rows, nt, axis, taps, n_taps, start, n_chunks, item);
    return;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int n_lblk = (nt + 63) / 64;
  const int64_t n_items = int64_t(lines[n_gal]) * n_lblk;
  for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int L = int(item / n_lblk);
    const int l = int(item - int64_t(L) * n_lblk) * 64 + lane;
    const bool lok = l < nt;
    const BatchLine b = batch_line(gal, lines, n_gal, nt, axis, L);
    const int nb = b.nb;
    const int64_t e0 = b.base + (lok ? l : 0);
    if (nb <= kBatchTileAxis) {
      __syncthreads();  // the previous line's reads are done
      for (int u = w; u < nb; u += 4) tile[u * 64 + lane] = lok ? in[e0 + u * b.stride] : 0.f;
      __syncthreads();
      if (lok) {
        for (int u0 = w; u0 < nb; u0 += 16) {
          float acc[4];
          tile_conv4<float>(tile, lane, nb, taps, n_taps, start, u0, acc);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int u = u0 + 4 * r;
            if (u < nb) out[e0 + u * b.stride] = acc[r];
          }
        }
      }
    } else if (lok) {
      for (int u = w; u < nb; u += 4) {
        float acc = 0.f;
        for (int t = 0; t < n_taps; ++t) {
          const int us = u + start - t;
          if (us >= 0 && us < nb) acc += taps[t] * in[e0 + us * b.stride];
        }
        out[e0 + u * b.stride] = acc;
      }
    }
  }
}

// sum(a * conv(b)) per block: line items of a packed table as
// conv_batch_tile_kernel, else the row items of conv_batch_dot_partial_kernel.
__global__ void __launch_bounds__(kDotThreads) conv_batch_dot_tile_kernel(
    const float* __restrict__ a, const float* __restrict__ bsrc,
    const rbx_galaxy* __restrict__ gal, int n_gal, int n_rows, int nt, int axis,
    const double* __restrict__ taps, int n_taps, int start, double* __restrict__ partial) {
  extern __shared__ float tile[];
  __shared__ int lines[kBatchGalMax + 1];
  __shared__ int part[256];
  __shared__ double sm[kDotThreads];
  const bool packed = batch_line_prefix(gal, n_gal, n_rows, lines, part);
  double acc = 0.0;
  if (!packed) {
    const int n_chunks = (nt + kLineChunk - 1) / kLineChunk;
    const int64_t n_items = int64_t(n_rows) * n_chunks;
    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
      const int p = int(item / n_chunks);
      const int l0 = int(item - int64_t(p) * n_chunks) * kLineChunk;
      const int l1 = l0 + kLineChunk < nt ? l0 + kLineChunk : nt;
      const BatchPos b = batch_pos(gal, n_gal, n_rows, nt, axis, int64_t(p) * nt);
      if (!b.ok) continue;
      for (int l = l0 + threadIdx.x; l < l1; l += blockDim.x) {
        double c = 0.0;
        for (int t = 0; t < n_taps; ++t) {
          const int us = b.u + start - t;
          if (us >= 0 && us < b.nb) c += taps[t] * double(bsrc[b.base + l + us * b.stride]);
        }
        acc += double(a[int64_t(p) * nt + l]) * c;
      }
    }
  } else {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n_lblk = (nt + 63) / 64;
    const int64_t n_items = int64_t(lines[n_gal]) * n_lblk;
    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
      const int L = int(item / n_lblk);
      const int l = int(item - int64_t(L) * n_lblk) * 64 + lane;
      const bool lok = l < nt;
      const BatchLine b = batch_line(gal, lines, n_gal, nt, axis, L);
      const int nb = b.nb;
      const int64_t e0 = b.base + (lok ? l : 0);
      if (nb <= kBatchTileAxis) {
        __syncthreads();
        for (int u = w; u < nb; u += 4) tile[u * 64 + lane] = lok ? bsrc[e0 + u * b.stride] : 0.f;
        __syncthreads();
        if (lok) {
          for (int u0 = w; u0 < nb; u0 += 16) {
            double c[4];
            tile_conv4<double>(tile, lane, nb, taps, n_taps, start, u0, c);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int u = u0 + 4 * r;
              if (u < nb) acc += double(a[e0 + u * b.stride]) * c[r];
            }
          }
        }
      } else if (lok) {
        for (int u = w; u < nb; u += 4) {
          double c = 0.0;
          for (int t = 0; t < n_taps; ++t) {
            const int us = u + start - t;
            if (us >= 0 && us < nb) c += taps[t] * double(bsrc[e0 + us * b.stride]);
          }
          acc += double(a[e0 + u * b.stride]) * c;
        }
      }
    }
  }
  block_partial(acc, sm, partial);
}

__global__ void __launch_bounds__(kDotThreads) conv_batch_dot_partial_kernel(
    const float* __restrict__ a, const float* __restrict__ bsrc,
    const rbx_galaxy* __restrict__ gal, int n_gal, int n_rows, int nt, int axis,
    const double* __restrict__ taps, int n_taps, int start, int n_chunks, int64_t n_items,
    double* __restrict__ partial) {
  __shared__ double sm[kDotThreads];
  double acc = 0.0;
  for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int p = int(item / n_chunks);
    const int l0 = int(item - int64_t(p) * n_chunks) * kLineChunk;
    const int l1 = l0 + kLineChunk < nt ? l0 + kLineChunk : nt;
    const BatchPos b = batch_pos(gal, n_gal, n_rows, nt, axis, int64_t(p) * nt);
    if (!b.ok) continue;
    for (int l = l0 + threadIdx.x; l < l1; l += blockDim.x) {
      double c = 0.0;
      for (int t = 0; t < n_taps; ++t) {
        const int us = b.u + start - t;
        if (us >= 0 && us < b.nb) c += taps[t] * double(bsrc[b.base + l + us * b.stride]);
      }
      acc += double(a[int64_t(p) * nt + l]) * c;
    }
  }
  block_partial(acc, sm, partial);
}

static int grid_for(int64_t n, int threads, int64_t cap) {
  int64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return int(g);
}

static int dot_blocks(int64_t total) { return grid_for(total, kDotThreads * 8, kDotBlocksMax); }

static int64_t item_blocks(int64_t items, int64_t cap) { return items < cap ? (items > 0 ? items : 1) : cap; }

int conv_axis(const float* in, float* out, int64_t n_outer, int n_axis, int64_t n_inner,
              const float* taps, int n_taps, int start, hipStream_t s) {
  const int64_t total = n_outer * n_axis * n_inner;
  if (total == 0) return RBX_OK;
  if (n_inner == 1 && n_taps - 1 <= kHalo) {
    const int64_t n_seg = (n_axis + kSeg - 1) / kSeg;
    const int64_t items = n_outer * n_seg;
    hipLaunchKernelGGL(conv_rows_kernel, dim3(item_blocks(items, 1 << 20)), dim3(256), 0, s, in,
                       out, n_axis, taps, n_taps, start, n_seg, items);
  } else if (n_inner > 1 && n_axis <= kTileAxis) {
    const int64_t n_blocks = (n_inner + 63) / 64;
    const int64_t items = n_outer * n_blocks;
    hipLaunchKernelGGL(conv_tile_kernel, dim3(item_blocks(items, 1 << 20)), dim3(256),
                       size_t(n_axis) * 64 * sizeof(float), s, in, out, n_axis, n_inner, taps,
                       n_taps, start, n_blocks, items);
  } else if (n_inner > 1) {
    const int64_t n_chunks = (n_inner + kLineChunk - 1) / kLineChunk;
    const int64_t items = n_outer * n_axis * n_chunks;
    hipLaunchKernelGGL(conv_lines_kernel, dim3(item_blocks(items, 1 << 20)), dim3(256), 0, s, in,
                       out, n_outer, n_axis, n_inner, taps, n_taps, start, n_chunks, items);
  } else {
    hipLaunchKernelGGL(conv_axis_kernel, dim3(grid_for(total, 256, 1 << 20)), dim3(256), 0, s, in,
                       out, n_outer, n_axis, n_inner, taps, n_taps, start);
  }
  return check_launch("rbx_conv_axis");
}

// dot = sum(a * conv_axis(b)) into result, the block partials in `partial`
// (nblk doubles, every one written by its block before the final sum).
static int conv_axis_dot(const float* a, const float* b, int64_t n_outer, int n_axis,
                         int64_t n_inner, const double* taps, int n_taps, int start,
                         double* partial, int nblk, double* result, hipStream_t s,
                         const char* who) {
  if (n_inner == 1 && n_taps - 1 <= kHalo) {
    const int64_t n_seg = (n_axis + kSeg - 1) / kSeg;
    hipLaunchKernelGGL(conv_dot_rows_kernel, dim3(nblk), dim3(kDotThreads), 0, s, a, b, n_axis,
                       taps, n_taps, start, n_seg, n_outer * n_seg, partial);
  } else if (n_inner > 1 && n_axis <= kTileAxis) {
    const int64_t n_blocks = (n_inner + 63) / 64;
    hipLaunchKernelGGL(conv_dot_tile_kernel, dim3(nblk), dim3(kDotThreads),
                       size_t(n_axis) * 64 * sizeof(float), s, a, b, n_axis, n_inner, taps, n_taps,
                       start, n_blocks, n_outer * n_blocks, partial);
  } else if (n_inner > 1) {
    const int64_t n_chunks = (n_inner + kLineChunk - 1) / kLineChunk;
    hipLaunchKernelGGL(conv_dot_lines_kernel, dim3(nblk), dim3(kDotThreads), 0, s, a, b, n_outer,
                       n_axis, n_inner, taps, n_taps, start, n_chunks, n_outer * n_axis * n_chunks,
                       partial);
  } else {
    hipLaunchKernelGGL(conv_dot_partial_kernel, dim3(nblk), dim3(kDotThreads), 0, s, a, b, n_outer,
                       n_axis, n_inner, taps, n_taps, start, partial);
  }
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kDotThreads), 0, s, partial, nblk, result);
  return check_launch(who);
}

// Block partials of rbx_conv_axis_dot_noalloc: a file-scope device array
// instead of a caller's workspace, kDotBlocksMax doubles (32 KiB), which
// every device the library runs on holds until the process ends.  Host code
// uses its per-device address (scratch_ptr), never the array's name.
static __device__ __align__(256) double scratch[kDotBlocksMax];

// Device address of `scratch` on the current device, looked up with
// hipGetSymbolAddress on the device's first call and kept per device.
static int scratch_ptr(double** out, const char* who) {
  static std::mutex mu;
  static void* ptrs[64] = {};
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
    return failf(RBX_EHIP, "%s: no current device", who);
  std::lock_guard<std::mutex> lock(mu);
  if (!ptrs[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return failf(RBX_EHIP, "%s: hipGetSymbolAddress(scratch) failed: %s", who,
                   hipGetErrorString(e));
    ptrs[dev] = p;
  }
  *out = static_cast<double*>(ptrs[dev]);
  return RBX_OK;
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
  return conv_axis_dot(a, b, n_outer, n_axis, n_inner, taps, n_taps, start,
                       static_cast<double*>(workspace), nblk, result, as_stream(stream),
                       "rbx_conv_axis_dot");
}

int rbx_conv_axis_dot_noalloc(const float* a, const float* b, int64_t n_outer, int32_t n_axis,
                              int64_t n_inner, const double* taps, int32_t n_taps,
                              int32_t start, double* result, void* stream) {
  using namespace rbx;
  const char* who = "rbx_conv_axis_dot_noalloc";
  if (!a || !b || !taps || !result || n_outer < 0 || n_axis < 1 || n_inner < 1 || n_taps < 1)
    return failf(RBX_EINVAL, "%s: bad arguments", who);
  const int64_t total = n_outer * n_axis * n_inner;
  const int nblk = dot_blocks(total);
  if (nblk > kDotBlocksMax)
    return failf(RBX_EINVAL, "%s: %d block partials exceed the %d-double scratch", who, nblk,
                 kDotBlocksMax);
  double* partial = nullptr;
  const int rc = scratch_ptr(&partial, who);
  if (rc) return rc;
  return conv_axis_dot(a, b, n_outer, n_axis, n_inner, taps, n_taps, start, partial, nblk, result,
                       as_stream(stream), who);
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
  if (n_gal <= kBatchGalMax) {
    // persistent grid: every workgroup builds the table's line prefix once
    hipLaunchKernelGGL(conv_batch_tile_kernel, dim3(2048), dim3(256),
                       size_t(kBatchTileAxis) * 64 * sizeof(float), as_stream(stream), in, out,
                       galaxies, n_gal, n_rows, n_target, axis, taps, n_taps, start);
    return check_launch("rbx_conv_batch");
  }
  const int n_chunks = (n_target + kLineChunk - 1) / kLineChunk;
  const int64_t items = int64_t(n_rows) * n_chunks;
  hipLaunchKernelGGL(conv_batch_kernel, dim3(item_blocks(items, 1 << 20)), dim3(256), 0,
                     as_stream(stream), in, out, galaxies, n_gal, n_rows, n_target, axis, taps,
                     n_taps, start, n_chunks, items);
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
  if (n_gal <= kBatchGalMax) {
    hipLaunchKernelGGL(conv_batch_dot_tile_kernel, dim3(nblk), dim3(kDotThreads),
                       size_t(kBatchTileAxis) * 64 * sizeof(float), s, a, b, galaxies, n_gal,
                       n_rows, n_target, axis, taps, n_taps, start, partial);
  } else {
    const int n_chunks = (n_target + kLineChunk - 1) / kLineChunk;
    hipLaunchKernelGGL(conv_batch_dot_partial_kernel, dim3(nblk), dim3(kDotThreads), 0, s, a, b,
                       galaxies, n_gal, n_rows, n_target, axis, taps, n_taps, start, n_chunks,
                       int64_t(n_rows) * n_chunks, partial);
  }
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kDotThreads), 0, s, partial, nblk, out);
  return check_launch("rbx_conv_batch_dot");
}

}  // extern "C"
