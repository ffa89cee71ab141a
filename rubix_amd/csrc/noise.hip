// apply_noise (rubix/core/noise.py:17-80, rubix/telescope/noise/noise.py:8-115)
// on a datacube of n_spaxel rows x n_target bins:
//   F[p]      = sum_l cube[p, l]                                   (noise.py:62)
//   M         = median of (F > 0 ? F : NaN) over all spaxels, NaN -> 0
//               (jnp.median propagates NaN: one spaxel without positive flux
//               makes M = 0 and the added noise zero everywhere; :65-69)
//   S[p]      = F[p] > 0 ? (sqrt(M) / sn[p]) / sqrt(F[p]) : 0        (:72-78)
//   out[p, l] = cube[p, l] + cube[p, l] * (z[p, l] * S[p])          (:108-113)
// z[i] (i = flat index p * n_target + l) follows jax.random's published
// construction for key = PRNGKey(seed) = (seed >> 32, seed & 0xffffffff):
//   bits = y0 ^ y1,  (y0, y1) = Threefry-2x32-20(key, counter (i >> 32, i & 0xffffffff))
//   u01  = float((bits >> 9) | 0x3f800000) - 1                       (uniform)
//   z    = sqrt(2) * erfinv(max(lo, 2 u01 + lo)),  lo = nextafter(-1, 0)  (normal)
// Threefry-2x32 with 20 rounds is Salmon et al., "Parallel random numbers: as
// easy as 1, 2, 3" (SC 2011).  The draws are recomputed in the adjoint, never
// stored.
#include <math.h>

#include "rbx_common.h"

static_assert(sizeof(rbx_noise_stats) == 16, "rbx_noise_stats layout");

namespace rbx {

constexpr int kSelThreads = 1024;  // the one workgroup of the median / S2N kernel
constexpr int kSelStage = 8192;    // flux images up to this many spaxels sit in LDS
constexpr int kRowThreads = 256;
constexpr int kRowSeg = 1024;  // bins per elementwise item

__device__ inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

// Rotations 13 15 26 6 | 17 29 16 24, key parity constant 0x1BD11BDA, key
// injection (with the injection count added to word 1) after every 4 rounds.
__device__ inline void threefry2x32(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1,
                                    uint32_t* o0, uint32_t* o1) {
  const int rot[8] = {13, 15, 26, 6, 17, 29, 16, 24};
  const uint32_t ks[3] = {k0, k1, k0 ^ k1 ^ 0x1BD11BDAu};
  uint32_t x0 = c0 + ks[0], x1 = c1 + ks[1];
#pragma unroll
  for (int r = 0; r < 20; ++r) {
    x0 += x1;
    x1 = rotl32(x1, rot[r & 7]);
    x1 ^= x0;
    if ((r & 3) == 3) {
      const int s = (r >> 2) + 1;
      x0 += ks[s % 3];
      x1 += ks[(s + 1) % 3] + uint32_t(s);
    }
  }
  *o0 = x0;
  *o1 = x1;
}

struct NoiseKey {
  uint32_t k0, k1;
};

inline NoiseKey key_of(uint64_t seed) { return {uint32_t(seed >> 32), uint32_t(seed)}; }

__device__ inline float mul_nc(float x, float y) {
#pragma clang fp contract(off)
  return x * y;
}

__device__ inline float draw(NoiseKey key, uint64_t i, int uniform) {
  uint32_t y0, y1;
  threefry2x32(key.k0, key.k1, uint32_t(i >> 32), uint32_t(i), &y0, &y1);
  const float f = __uint_as_float(((y0 ^ y1) >> 9) | 0x3f800000u) - 1.0f;
  if (uniform) return f;
  const float lo = -0x1.fffffep-1f;          // nextafter(-1, 0)
  const float u = fmaxf(lo, f * 2.0f + lo);  // 2f is exact: one rounding, fused or not
  return mul_nc(0x1.6a09e6p+0f, erfinvf(u));  // float32(sqrt(2))
}

__global__ void __launch_bounds__(256) noise_sample_kernel(int64_t n, NoiseKey key, int uniform,
                                                           float* __restrict__ out) {
  for (int64_t i = gtid(); i < n; i += int64_t(gridDim.x) * blockDim.x)
    out[i] = draw(key, uint64_t(i), uniform);
}

// Block-wide double sum in a fixed order (wave butterflies, then waves in order).
__device__ inline double block_sum(double v, double* sm) {
  v = wave_sum(v);
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) sm[wave] = v;
  __syncthreads();
  double tot = 0.0;
  for (int w = 0; w < int(blockDim.x / kWave); ++w) tot += sm[w];
  __syncthreads();
  return tot;
}

// Flux image: F[p] = sum_l cube[p, l], one workgroup per spaxel row,
// double accumulation in a fixed order, rounded once.
__global__ void __launch_bounds__(kRowThreads) flux_rows_kernel(const float* __restrict__ cube,
                                                                int n_spaxel, int nt,
                                                                float* __restrict__ flux) {
  __shared__ double sm[kRowThreads / kWave];
  for (int p = blockIdx.x; p < n_spaxel; p += gridDim.x) {
    const int64_t row = int64_t(p) * nt;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int l = threadIdx.x;
    for (; l + 3 * kRowThreads < nt; l += 4 * kRowThreads) {
      a0 += cube[row + l];
      a1 += cube[row + l + kRowThreads];
      a2 += cube[row + l + 2 * kRowThreads];
      a3 += cube[row + l + 3 * kRowThreads];
    }
    for (; l < nt; l += kRowThreads) a0 += cube[row + l];
    const double tot = block_sum((a0 + a1) + (a2 + a3), sm);
    if (threadIdx.x == 0) flux[p] = float(tot);
  }
}

// Adjoint, first pass over spaxel row p (each draw made once): the direct
// term dcube = dout + dout * (z * S[p]) and A[p] = sum_l dout * cube * z
// (double, fixed order).  Rows with S[p] == 0 copy dout and get A = 0.
__global__ void __launch_bounds__(kRowThreads) noise_bwd_rows_kernel(
    const float* __restrict__ cube, const float* __restrict__ dout, int n_spaxel, int nt,
    const float* __restrict__ s2n, NoiseKey key, int uniform, double* __restrict__ dots,
    float* __restrict__ dcube) {
  __shared__ double sm[kRowThreads / kWave];
  for (int p = blockIdx.x; p < n_spaxel; p += gridDim.x) {
    const int64_t row = int64_t(p) * nt;
    const float s = s2n[p];
    double acc = 0.0;
    if (s != 0.f) {
      for (int l = threadIdx.x; l < nt; l += kRowThreads) {
        const float z = draw(key, uint64_t(row + l), uniform);
        const float g = dout[row + l];
        acc += double(g) * double(cube[row + l]) * double(z);
        dcube[row + l] = g + g * (z * s);
      }
    } else {
      for (int l = threadIdx.x; l < nt; l += kRowThreads) dcube[row + l] = dout[row + l];
    }
    const double tot = block_sum(acc, sm);
    if (threadIdx.x == 0) dots[p] = tot;
  }
}

// Rank-k element (0-based, ascending) of n positive floats, whose bit
// patterns order like the values: 4 radix passes of 8 bits with an LDS
// histogram.  Returns the value's bits; *tie gets k's rank among the
// elements equal to it.
__device__ uint32_t radix_select(const float* F, int n, int k, unsigned* hist, uint32_t* s_pref,
                                 int* s_k, int* tie) {
  uint32_t prefix = 0, mask = 0;
  for (int shift = 24; shift >= 0; shift -= 8) {
    for (int d = threadIdx.x; d < 256; d += blockDim.x) hist[d] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const uint32_t b = __float_as_uint(F[i]);
      if ((b & mask) == prefix) atomicAdd(&hist[(b >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kWave) {
      // wave 0: lane q owns bins 4q..4q+3; an inclusive shuffle scan finds
      // the one lane whose bins hold rank k (counts sum to more than k)
      const int lane = threadIdx.x;
      unsigned h[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) h[j] = hist[4 * lane + j];
      const unsigned own = h[0] + h[1] + h[2] + h[3];
      unsigned incl = own;
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) {
        const unsigned t = __shfl_up(incl, off, kWave);
        if (lane >= off) incl += t;
      }
      unsigned cum = incl - own;
      if (cum <= unsigned(k) && unsigned(k) < incl) {
        int d = 4 * lane;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          if (cum + h[j] > unsigned(k)) break;
          cum += h[j];
          ++d;
        }
        *s_pref = prefix | (uint32_t(d) << shift);
        *s_k = k - int(cum);
      }
    }
    __syncthreads();
    prefix = *s_pref;
    k = *s_k;
    mask |= 255u << shift;
    __syncthreads();
  }
  *tie = k;
  return prefix;
}

// Index of the rank-th (0-based) element whose bits equal `bits`, in index
// order (the element a stable sort puts at that rank).
__device__ int nth_equal(const float* F, int n, uint32_t bits, int rank,
                         unsigned* wave_cnt, int* s_idx) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int n_waves = blockDim.x / kWave;
  if (threadIdx.x == 0) *s_idx = -1;
  int running = 0;
  for (int base = 0; base < n; base += blockDim.x) {
    const int i = base + threadIdx.x;
    const bool eq = i < n && __float_as_uint(F[i]) == bits;
    const uint64_t m = __ballot(eq);
    if (lane == 0) wave_cnt[wave] = unsigned(__popcll(m));
    __syncthreads();
    int before = running;
    int total = running;
    for (int w = 0; w < n_waves; ++w) {
      if (w < wave) before += int(wave_cnt[w]);
      total += int(wave_cnt[w]);
    }
    before += __popcll(m & ((uint64_t(1) << lane) - 1));
    if (eq && before == rank) *s_idx = i;
    __syncthreads();
    running = total;
    if (running > rank) break;
  }
  __syncthreads();
  return *s_idx;
}

// One workgroup: the median of the flux image (NaN-propagating, as
// jnp.median), the spaxels that hold it (for the adjoint), and the S2N map.
__global__ void __launch_bounds__(kSelThreads) noise_s2n_kernel(
    const float* __restrict__ F, int n, const float* __restrict__ sn, float* __restrict__ s2n,
    rbx_noise_stats* __restrict__ stats) {
  __shared__ float stage[kSelStage];
  __shared__ unsigned hist[256];
  __shared__ unsigned wave_cnt[kSelThreads / kWave];
  __shared__ uint32_t s_pref;
  __shared__ int s_k, s_idx, s_any;
  if (threadIdx.x == 0) s_any = 0;
  __syncthreads();
  // the selection passes read the flux image from LDS when it fits
  const bool staged = n <= kSelStage;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const float f = F[i];
    if (!(f > 0.f)) s_any = 1;
    if (staged) stage[i] = f;
  }
  __syncthreads();
  const float* src = staged ? stage : F;
  const bool valid = s_any == 0;
  float med = 0.f;
  int lo_idx = -1, hi_idx = -1;
  if (valid) {
    // midpoint of ranks floor((n-1)/2) and ceil((n-1)/2): (lo + hi) * 0.5
    const int k_lo = (n - 1) / 2, k_hi = n / 2;
    int tie;
    const uint32_t b_lo = radix_select(src, n, k_lo, hist, &s_pref, &s_k, &tie);
    lo_idx = nth_equal(src, n, b_lo, tie, wave_cnt, &s_idx);
    uint32_t b_hi = b_lo;
    hi_idx = lo_idx;
    if (k_hi != k_lo) {
      b_hi = radix_select(src, n, k_hi, hist, &s_pref, &s_k, &tie);
      hi_idx = nth_equal(src, n, b_hi, tie, wave_cnt, &s_idx);
    }
    med = (__uint_as_float(b_lo) + __uint_as_float(b_hi)) * 0.5f;
  }
  const float root_m = sqrtf(med);
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    const float f = F[p];
    s2n[p] = f > 0.f ? (root_m / sn[p]) / sqrtf(f) : 0.f;
  }
  if (threadIdx.x == 0) {
    rbx_noise_stats st;
    st.median = med;
    st.median_lo = lo_idx;
    st.median_hi = hi_idx;
    st.any_nonpositive = valid ? 0 : 1;
    *stats = st;
  }
}

// Items (spaxel, segment of kRowSeg bins); each thread takes 4 bins.
struct RowItem {
  int p, l0;
};
__device__ inline RowItem row_item(uint32_t item, uint32_t segs) {
  const uint32_t p = item / segs;
  return {int(p), int(item - p * segs) * kRowSeg};
}

__global__ void __launch_bounds__(kRowThreads) add_noise_kernel(
    const float* cube, int nt, uint32_t n_items, uint32_t segs, const float* __restrict__ s2n,
    NoiseKey key, int uniform, float* out) {
  for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const RowItem it = row_item(item, segs);
    const float s = s2n[it.p];
    const int64_t row = int64_t(it.p) * nt;
#pragma unroll
    for (int q = 0; q < kRowSeg / kRowThreads; ++q) {
      const int l = it.l0 + q * kRowThreads + threadIdx.x;
      if (l < nt) {
        const float c = cube[row + l];
        const float zs = draw(key, uint64_t(row + l), uniform) * s;
        out[row + l] = c + mul_nc(c, zs);
      }
    }
  }
}

// Adjoint coefficients per spaxel, one workgroup.  With S[p] = sqrt(M)/sn[p]
// / sqrt(F[p]):  dS/dF = -S / (2F),  dS/dM = S / (2M), and M moves with the
// flux of the spaxel(s) at the median (weight 1/2 each of the two midpoint
// ranks).  On entry dots[p] = A[p] = sum_l dout * cube * z; on exit
// dots[p] = dL/dF[p].  M = 0 (some spaxel without positive flux) leaves the
// S map identically 0: then dL/dF = 0.
__global__ void __launch_bounds__(kSelThreads) noise_coef_kernel(
    int n, const float* __restrict__ F, const float* __restrict__ s2n,
    const rbx_noise_stats* __restrict__ stats, double* __restrict__ dots) {
  __shared__ double sm[kSelThreads / kWave];
  const rbx_noise_stats st = *stats;
  const bool live = st.any_nonpositive == 0 && st.median > 0.f;
  if (!live) {
    for (int p = threadIdx.x; p < n; p += blockDim.x) dots[p] = 0.0;
    return;
  }
  const double M = st.median;
  double acc = 0.0;
  for (int p = threadIdx.x; p < n; p += blockDim.x) acc += dots[p] * double(s2n[p]);
  const double dM = block_sum(acc, sm) / (2.0 * M);
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    const double A = dots[p];
    dots[p] = -A * double(s2n[p]) / (2.0 * double(F[p]));
  }
  __syncthreads();
  if (threadIdx.x == 0 && st.median_lo >= 0 && st.median_lo < n && st.median_hi >= 0 &&
      st.median_hi < n) {
    dots[st.median_lo] += 0.5 * dM;
    dots[st.median_hi] += 0.5 * dM;
  }
}

// Adjoint, last pass: dcube[p, l] += dL/dF[p] (the flux path).
__global__ void __launch_bounds__(kRowThreads) add_rows_kernel(float* __restrict__ dcube, int nt,
                                                               uint32_t n_items, uint32_t segs,
                                                               const double* __restrict__ dF) {
  for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const RowItem it = row_item(item, segs);
    const float b = float(dF[it.p]);
    const int64_t row = int64_t(it.p) * nt;
#pragma unroll
    for (int q = 0; q < kRowSeg / kRowThreads; ++q) {
      const int l = it.l0 + q * kRowThreads + threadIdx.x;
      if (l < nt) dcube[row + l] += b;
    }
  }
}

int row_grid(int n_spaxel) { return n_spaxel < (1 << 16) ? n_spaxel : (1 << 16); }

int item_blocks(uint32_t n_items) {
  return n_items < (1u << 20) ? int(n_items) : (1 << 20);
}

int check_noise_args(const char* who, int32_t n_spaxel, int32_t n_target, int32_t distribution) {
  if (n_spaxel < 1 || n_target < 1)
    return failf(RBX_EINVAL, "%s: empty datacube (%d spaxels x %d bins)", who, n_spaxel, n_target);
  if (distribution != RBX_NOISE_NORMAL && distribution != RBX_NOISE_UNIFORM)
    return failf(RBX_EINVAL, "%s: unknown noise distribution %d (0 = normal, 1 = uniform)", who,
                 distribution);
  const uint64_t segs = (uint64_t(n_target) + kRowSeg - 1) / kRowSeg;
  if (uint64_t(n_spaxel) * segs >= (uint64_t(1) << 32))
    return failf(RBX_EINVAL, "%s: datacube too large", who);
  return RBX_OK;
}

}  // namespace rbx

extern "C" {

int rbx_noise_sample(int64_t n, uint64_t seed, int32_t distribution, float* out, void* stream) {
  using namespace rbx;
  if (n < 0 || (n > 0 && !out)) return fail(RBX_EINVAL, "rbx_noise_sample: bad arguments");
  if (distribution != RBX_NOISE_NORMAL && distribution != RBX_NOISE_UNIFORM)
    return failf(RBX_EINVAL, "rbx_noise_sample: unknown noise distribution %d", distribution);
  if (n == 0) return RBX_OK;
  int64_t blocks = (n + 255) / 256;
  if (blocks > (1 << 20)) blocks = 1 << 20;
  hipLaunchKernelGGL(noise_sample_kernel, dim3(unsigned(blocks)), dim3(256), 0, as_stream(stream),
                     n, key_of(seed), distribution == RBX_NOISE_UNIFORM ? 1 : 0, out);
  return check_launch("rbx_noise_sample");
}

size_t rbx_noise_workspace_size(int32_t n_spaxel) {
  return n_spaxel > 0 ? size_t(n_spaxel) * sizeof(double) : 0;
}

int rbx_apply_noise(const float* cube, int32_t n_spaxel, int32_t n_target,
                    const float* signal_to_noise, uint64_t seed, int32_t distribution,
                    float* flux, float* s2n, rbx_noise_stats* stats, float* out, void* stream) {
  using namespace rbx;
  if (int rc = check_noise_args("rbx_apply_noise", n_spaxel, n_target, distribution)) return rc;
  if (!cube || !signal_to_noise || !flux || !s2n || !stats || !out)
    return fail(RBX_EINVAL, "rbx_apply_noise: null pointer");
  hipStream_t st = as_stream(stream);
  const NoiseKey key = key_of(seed);
  const int uniform = distribution == RBX_NOISE_UNIFORM ? 1 : 0;
  hipLaunchKernelGGL(flux_rows_kernel, dim3(row_grid(n_spaxel)), dim3(kRowThreads), 0, st, cube,
                     n_spaxel, n_target, flux);
  hipLaunchKernelGGL(noise_s2n_kernel, dim3(1), dim3(kSelThreads), 0, st, flux, n_spaxel,
                     signal_to_noise, s2n, stats);
  const uint32_t segs = uint32_t((n_target + kRowSeg - 1) / kRowSeg);
  const uint32_t n_items = uint32_t(n_spaxel) * segs;
  hipLaunchKernelGGL(add_noise_kernel, dim3(item_blocks(n_items)), dim3(kRowThreads), 0, st, cube,
                     n_target, n_items, segs, s2n, key, uniform, out);
  return check_launch("rbx_apply_noise");
}

int rbx_apply_noise_bwd(const float* cube, const float* dout, int32_t n_spaxel, int32_t n_target,
                        const float* flux, const float* s2n, const rbx_noise_stats* stats,
                        uint64_t seed, int32_t distribution, void* workspace,
                        size_t workspace_bytes, float* dcube, void* stream) {
  using namespace rbx;
  if (int rc = check_noise_args("rbx_apply_noise_bwd", n_spaxel, n_target, distribution))
    return rc;
  if (!cube || !dout || !flux || !s2n || !stats || !dcube || !workspace)
    return fail(RBX_EINVAL, "rbx_apply_noise_bwd: null pointer");
  if (workspace_bytes < rbx_noise_workspace_size(n_spaxel))
    return failf(RBX_EWORKSPACE, "rbx_apply_noise_bwd: workspace of %zu bytes, %zu needed",
                 workspace_bytes, rbx_noise_workspace_size(n_spaxel));
  hipStream_t st = as_stream(stream);
  const NoiseKey key = key_of(seed);
  const int uniform = distribution == RBX_NOISE_UNIFORM ? 1 : 0;
  double* dots = static_cast<double*>(workspace);
  hipLaunchKernelGGL(noise_bwd_rows_kernel, dim3(row_grid(n_spaxel)), dim3(kRowThreads), 0, st,
                     cube, dout, n_spaxel, n_target, s2n, key, uniform, dots, dcube);
  hipLaunchKernelGGL(noise_coef_kernel, dim3(1), dim3(kSelThreads), 0, st, n_spaxel, flux, s2n,
                     stats, dots);
  const uint32_t segs = uint32_t((n_target + kRowSeg - 1) / kRowSeg);
  const uint32_t n_items = uint32_t(n_spaxel) * segs;
  hipLaunchKernelGGL(add_rows_kernel, dim3(item_blocks(n_items)), dim3(kRowThreads), 0, st, dcube,
                     n_target, n_items, segs, dots);
  return check_launch("rbx_apply_noise_bwd");
}

}  // extern "C"
