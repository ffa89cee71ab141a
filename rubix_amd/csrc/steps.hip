// Plain per-step kernels: one kernel (and one hand-written adjoint) per
// rubix.core step, each writing the same RubixData field the reference step
// writes.  They materialise the per-star spectra exactly like the reference
// chain does, so they are HBM-bound by design; the fused path in cube.hip
// replaces calculate_spectra .. calculate_datacube for throughput.
//
//   calculate_spectra             rbx_ssp_lookup      (rubix/core/ifu.py:63-121)
//   scale_spectrum_by_mass        rbx_scale_rows      (rubix/core/ifu.py:149-157)
//   doppler_shift_and_resampling  rbx_resample        (rubix/core/ifu.py:269-293,
//                                                      rubix/spectra/ifu.py:190,223-266)
//   calculate_datacube            rbx_segment_cube    (rubix/spectra/ifu.py:269-288)
#include <float.h>

#include "rbx_common.h"
#include "ssp_device.h"

namespace rbx {

constexpr int kStepThreads = 256;

// Fixed-order block reduction (deterministic); every thread gets the sum.
template <class T>
__device__ inline T block_sum(T v, T* sm) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sm[w] = v;
  __syncthreads();
  T s = T(0);
  const int nw = blockDim.x >> 6;
  for (int k = 0; k < nw; ++k) s += sm[k];
  __syncthreads();
  return s;
}

// ------------------------------------------------------------ lookup

__global__ void __launch_bounds__(kStepThreads) lookup_kernel(SspView g, const float* __restrict__ met,
                                                             const float* __restrict__ age, int64_t n,
                                                             float* __restrict__ out) {
  for (int64_t r = blockIdx.x; r < n; r += gridDim.x) {
    const Cell c = make_cell(g, met[r], age[r]);
    const int nw = g.n_wave;
    float* o = out + r * nw;
    if (!c.inside) {
      for (int k = threadIdx.x; k < nw; k += blockDim.x) o[k] = 0.f;
      continue;
    }
    const float* f00 = g.flux + int64_t(c.row) * nw;
    const float* f10 = f00 + int64_t(g.n_age) * nw;
    const float* f01 = f00 + nw;
    const float* f11 = f10 + nw;
    for (int k = threadIdx.x; k < nw; k += blockDim.x)
      o[k] = c.w[0] * f00[k] + c.w[1] * f10[k] + c.w[2] * f01[k] + c.w[3] * f11[k];
  }
}

__global__ void __launch_bounds__(kStepThreads) lookup_bwd_kernel(
    SspView g, const float* __restrict__ met, const float* __restrict__ age, int64_t n,
    const float* __restrict__ dspec, float* __restrict__ dmet, float* __restrict__ dage) {
  __shared__ float sm[kStepThreads / 64];
  for (int64_t r = blockIdx.x; r < n; r += gridDim.x) {
    const float Z = met[r], t = age[r];
    const Cell c = make_cell(g, Z, t);
    if (!c.inside) {
      if (threadIdx.x == 0) {
        if (dmet) dmet[r] = 0.f;
        if (dage) dage[r] = 0.f;
      }
      continue;
    }
    const int nw = g.n_wave;
    const float* f00 = g.flux + int64_t(c.row) * nw;
    const float* f10 = f00 + int64_t(g.n_age) * nw;
    const float* f01 = f00 + nw;
    const float* f11 = f10 + nw;
    const float* d = dspec + r * nw;
    float su = 0.f, ss = 0.f;
    for (int k = threadIdx.x; k < nw; k += blockDim.x) {
      const float v00 = f00[k], v10 = f10[k], v01 = f01[k], v11 = f11[k];
      const float dk = d[k];
      su += dk * (c.gu[0] * v00 + c.gu[1] * v10 + c.gu[2] * v01 + c.gu[3] * v11);
      ss += dk * (c.gs[0] * v00 + c.gs[1] * v10 + c.gs[2] * v01 + c.gs[3] * v11);
    }
    su = block_sum(su, sm);
    ss = block_sum(ss, sm);
    if (threadIdx.x == 0) {
      if (dmet) dmet[r] = su * c.du_dZ;
      if (dage) dage[r] = ss * c.ds_dt;
    }
  }
}

// ------------------------------------------------------------ mass scaling

__global__ void scale_rows_kernel(const float* __restrict__ in, const float* __restrict__ scale,
                                  int64_t n, int nw, float* __restrict__ out) {
  const int64_t total = n * nw;
  for (int64_t idx = gtid(); idx < total; idx += int64_t(gridDim.x) * blockDim.x)
    out[idx] = in[idx] * scale[idx / nw];
}

__global__ void __launch_bounds__(kStepThreads) row_dot_kernel(const float* __restrict__ a,
                                                              const float* __restrict__ b, int64_t n,
                                                              int nw, float* __restrict__ out) {
  __shared__ float sm[kStepThreads / 64];
  for (int64_t r = blockIdx.x; r < n; r += gridDim.x) {
    float s = 0.f;
    for (int k = threadIdx.x; k < nw; k += blockDim.x) s += a[r * nw + k] * b[r * nw + k];
    s = block_sum(s, sm);
    if (threadIdx.x == 0) out[r] = s;
  }
}

// ------------------------------------------------------------ resampling

struct StarDoppler {
  float a, inv_a;
  int kl, kh;  // T window (kl >= 1)
};

__device__ inline StarDoppler doppler_of(const SspView& g, const rbx_target& tg, float v, float c) {
  StarDoppler d;
  d.a = doppler_factor(v, c);
  d.inv_a = 1.0f / d.a;
  const int klo_le = locate(g, tg.wmin, d.a, d.inv_a);
  const int klo = (klo_le >= 0 && knot(g, klo_le, d.a) == tg.wmin) ? klo_le : klo_le + 1;
  d.kl = klo < 1 ? 1 : klo;
  d.kh = locate(g, tg.wmax, d.a, d.inv_a);
  return d;
}

__device__ inline float interp_row(const float* __restrict__ S, const Seg& sg) {
  const float s0 = S[sg.i - 1], s1 = S[sg.i];
  float P = s0 + sg.theta * (s1 - s0);
  if (sg.side < 0) P = s0;
  if (sg.side > 0) P = s1;
  return P;
}

// T = sum_{k in [kl, kh]} S_k * fl(w_k - w_{k-1}) (the in-range source flux).
__device__ inline float source_total(const SspView& g, const float* __restrict__ S,
                                     const StarDoppler& d, float* sm) {
  float T = 0.f;
  for (int k = d.kl + threadIdx.x; k <= d.kh; k += blockDim.x)
    T += S[k] * (knot(g, k, d.a) - knot(g, k - 1, d.a));
  return block_sum(T, sm);
}

__global__ void __launch_bounds__(kStepThreads) resample_kernel(
    SspView g, rbx_target tg, const float* __restrict__ spectra, const float* __restrict__ vel,
    int64_t vstride, int vcomp, float c, int64_t n, float* __restrict__ out) {
  __shared__ float sm[kStepThreads / 64];
  const int ns = g.n_wave, nt = tg.n;
  for (int64_t r = blockIdx.x; r < n; r += gridDim.x) {
    const float* S = spectra + r * ns;
    float* o = out + r * nt;
    const StarDoppler d = doppler_of(g, tg, vel[r * vstride + vcomp], c);
    const float T = source_total(g, S, d, sm);
    float Nn = 0.f;
    for (int j = threadIdx.x; j < nt; j += blockDim.x) {
      const Seg sg = segment(g, tg.wave[j], d.a, d.inv_a);
      const float P = interp_row(S, sg);
      o[j] = P;
      Nn += P * tg.dwave[j];
    }
    Nn = block_sum(Nn, sm);
    float sf = T / Nn;
    if (isnan(sf)) sf = 0.f;
    else if (isinf(sf)) sf = sf > 0.f ? FLT_MAX : -FLT_MAX;
    // each thread rescales the bins it wrote itself
    for (int j = threadIdx.x; j < nt; j += blockDim.x) o[j] *= sf;
  }
}

// Adjoint of resample_kernel for one star per workgroup.
//   P_j from the input spectrum S, out_j = P_j sf, sf = T / Nn.
//   sbar = sum_j g_j P_j; Pbar_j = g_j sf + Nbar d_j with Nbar = -sbar T / Nn^2
//   (0 when T / Nn is not finite), Tbar = sbar / Nn.
//   Sbar_k gathers the interpolation weights of every target bin whose
//   segment touches knot k (first[k] = first target bin with segment index
//   >= k, built without conflicts because segment indices are monotone in
//   j), plus Tbar (w_k - w_{k-1}) inside the mask window.
//   abar = sum_j Pbar_j dP_j/da + Tbar T / a; dv = abar a / c.
// LDS: c0, c1 (float) and seg (uint16) per target bin, first[] (uint16) per
// knot; first[] is reset to n_target before it is filled, so every index read
// from it stays in [0, n_target] whatever the target ordering.
__global__ void __launch_bounds__(kStepThreads) resample_bwd_kernel(
    SspView g, rbx_target tg, const float* __restrict__ spectra, const float* __restrict__ vel,
    int64_t vstride, int vcomp, float c, int64_t n, const float* __restrict__ dout,
    float* __restrict__ dspec, float* __restrict__ dv) {
  extern __shared__ char lds[];
  __shared__ float sm[kStepThreads / 64];
  __shared__ double smd[kStepThreads / 64];
  const int ns = g.n_wave, nt = tg.n;
  float* c0 = reinterpret_cast<float*>(lds);
  float* c1 = c0 + nt;
  uint16_t* seg = reinterpret_cast<uint16_t*>(c1 + nt);
  uint16_t* first = seg + nt;  // ns + 2 entries
  for (int64_t r = blockIdx.x; r < n; r += gridDim.x) {
    const float* S = spectra + r * ns;
    const float* gj = dout + r * nt;
    float* dS = dspec ? dspec + r * ns : nullptr;
    const StarDoppler d = doppler_of(g, tg, vel[r * vstride + vcomp], c);
    const float T = source_total(g, S, d, sm);
    float Nn = 0.f, sbar = 0.f;
    for (int j = threadIdx.x; j < nt; j += blockDim.x) {
      const Seg sg = segment(g, tg.wave[j], d.a, d.inv_a);
      const float P = interp_row(S, sg);
      Nn += P * tg.dwave[j];
      sbar += gj[j] * P;
    }
    Nn = block_sum(Nn, sm);
    sbar = block_sum(sbar, sm);
    const float raw = T / Nn;
    const bool live = isfinite(raw);
    float sf = raw;
    if (isnan(sf)) sf = 0.f;
    else if (isinf(sf)) sf = sf > 0.f ? FLT_MAX : -FLT_MAX;
    const double Nbar = live ? -double(sbar) * T / (double(Nn) * Nn) : 0.0;
    const double Tbar = live ? double(sbar) / Nn : 0.0;
    double abar = 0.0;  // cancels strongly: accumulated in double
    for (int j = threadIdx.x; j < nt; j += blockDim.x) {
      const Seg sg = segment(g, tg.wave[j], d.a, d.inv_a);
      const double Pbar_d = double(gj[j]) * sf + Nbar * tg.dwave[j];
      const float Pbar = float(Pbar_d);
      float w0 = 1.f - sg.theta, w1 = sg.theta;
      if (sg.side < 0) { w0 = 1.f; w1 = 0.f; }
      if (sg.side > 0) { w0 = 0.f; w1 = 1.f; }
      c0[j] = Pbar * w0;
      c1[j] = Pbar * w1;
      seg[j] = uint16_t(sg.i);
      if (sg.side == 0) {
        const float s0 = S[sg.i - 1], s1 = S[sg.i];
        const float kappa = (g.wave[sg.i - 1] + sg.theta * (g.wave[sg.i] - g.wave[sg.i - 1])) /
                            (sg.wr - sg.wl);
        abar += Pbar_d * (-double(s1 - s0) * double(kappa));
      }
    }
    abar = block_sum(abar, smd);  // also a barrier: seg[] is complete
    if (dS) {
      for (int k = threadIdx.x; k <= ns + 1; k += blockDim.x) first[k] = uint16_t(nt);
      __syncthreads();
      for (int j = threadIdx.x; j < nt; j += blockDim.x) {
        const int lo = (j == 0) ? 0 : int(seg[j - 1]) + 1;
        for (int k = lo; k <= int(seg[j]); ++k) first[k] = uint16_t(j);
      }
      __syncthreads();
      for (int k = threadIdx.x; k < ns; k += blockDim.x) {
        float acc = 0.f;
        for (int j = first[k]; j < first[k + 1]; ++j) acc += c1[j];      // segments i = k
        for (int j = first[k + 1]; j < first[k + 2]; ++j) acc += c0[j];  // segments i = k + 1
        if (k >= d.kl && k <= d.kh) acc += float(Tbar) * (knot(g, k, d.a) - knot(g, k - 1, d.a));
        dS[k] = acc;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0 && dv) {
      const double ab = abar + Tbar * (double(T) / d.a);
      dv[r] = float(ab * d.a / c);
    }
  }
}

// ------------------------------------------------------------ segment cube

__global__ void __launch_bounds__(64) segment_accum_kernel(
    const float* __restrict__ spectra, const int32_t* __restrict__ perm, int nt,
    const ChunkDesc* __restrict__ desc, const int32_t* __restrict__ n_chunks_ptr, int n_tiles,
    double* __restrict__ partial) {
  const int tile = blockIdx.x % n_tiles;
  const int chunk = blockIdx.x / n_tiles;
  if (chunk >= *n_chunks_ptr) return;
  const ChunkDesc dd = desc[chunk];
  const int j = tile * 64 + threadIdx.x;
  if (j >= nt) return;
  double acc = 0.0;
  for (int r = dd.begin; r < dd.end; ++r) acc += spectra[int64_t(perm[r]) * nt + j];
  partial[int64_t(chunk) * nt + j] = acc;
}

__global__ void segment_bwd_kernel(const float* __restrict__ dcube, const int32_t* __restrict__ ids,
                                   int64_t n, int n_spaxel, int nt, float* __restrict__ dspec) {
  const int64_t total = n * nt;
  for (int64_t idx = gtid(); idx < total; idx += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r = idx / nt;
    const int j = int(idx - r * nt);
    const int32_t p = ids[r];
    dspec[idx] = (p >= 0 && p < n_spaxel) ? dcube[int64_t(p) * nt + j] : 0.f;
  }
}

struct SegLayout {
  int32_t *perm, *offsets, *chunk_off, *scan_blocks;
  ChunkDesc* desc;
  double* partial;
  void* sort_ws;
  size_t sort_ws_bytes;
  int64_t max_chunks;
};

template <class C>
static void seg_layout(C& c, int64_t n, int32_t n_keys, int32_t nt, SegLayout* L) {
  const int64_t max_chunks = (n + kChunk - 1) / kChunk + n_keys;
  const size_t sort_bytes = sort_workspace_size(n, n_keys);
  auto* perm = c.template take<int32_t>(n > 0 ? n : 1);
  auto* offs = c.template take<int32_t>(size_t(n_keys) + 1);
  auto* coff = c.template take<int32_t>(size_t(n_keys) + 1);
  auto* blocks = c.template take<int32_t>((size_t(n_keys) + 1) / 4096 + 2);
  auto* desc = c.template take<ChunkDesc>(max_chunks);
  auto* part = c.template take<double>(size_t(max_chunks) * nt);
  auto* sws = c.template take<char>(sort_bytes);
  if (L) {
    L->perm = reinterpret_cast<int32_t*>(perm);
    L->offsets = reinterpret_cast<int32_t*>(offs);
    L->chunk_off = reinterpret_cast<int32_t*>(coff);
    L->scan_blocks = reinterpret_cast<int32_t*>(blocks);
    L->desc = reinterpret_cast<ChunkDesc*>(desc);
    L->partial = reinterpret_cast<double*>(part);
    L->sort_ws = sws;
    L->sort_ws_bytes = sort_bytes;
    L->max_chunks = max_chunks;
  }
}

struct SizeOnly2 {
  Sizer s;
  template <class T>
  T* take(size_t count) {
    s.take<T>(count);
    return nullptr;
  }
};

static int blocks_for_rows(int64_t n) { return int(n < (1 << 20) ? (n > 0 ? n : 1) : (1 << 20)); }

static int grid_for(int64_t n, int threads) {
  int64_t gsz = (n + threads - 1) / threads;
  if (gsz < 1) gsz = 1;
  if (gsz > (1 << 20)) gsz = 1 << 20;
  return int(gsz);
}

static bool ssp_ok(const rbx_ssp* s, bool need_wave) {
  if (!s || !s->flux || !s->ax_met || !s->ax_age || s->n_met < 2 || s->n_age < 2 || s->n_wave < 2)
    return false;
  if (need_wave && (!s->wave_obs || !s->locate || s->n_locate < 1 || !(s->locate_inv_h > 0.f)))
    return false;
  return true;
}

static size_t resample_bwd_lds(int nt, int ns) {
  return size_t(nt) * 2 * 4 + size_t(nt) * 2 + size_t(ns + 2) * 2;
}

}  // namespace rbx

extern "C" {

int rbx_ssp_lookup(const rbx_ssp* ssp, const float* met, const float* age, int64_t n,
                   float* spectra, void* stream) {
  using namespace rbx;
  if (!ssp_ok(ssp, false) || n < 0 || (n > 0 && (!met || !age || !spectra)))
    return fail(RBX_EINVAL, "rbx_ssp_lookup: bad arguments");
  if (n == 0) return RBX_OK;
  hipLaunchKernelGGL(lookup_kernel, dim3(blocks_for_rows(n)), dim3(kStepThreads), 0,
                     as_stream(stream), view_of(*ssp), met, age, n, spectra);
  return check_launch("rbx_ssp_lookup");
}

int rbx_ssp_lookup_bwd(const rbx_ssp* ssp, const float* met, const float* age, int64_t n,
                       const float* dspectra, float* dmet, float* dage, void* stream) {
  using namespace rbx;
  if (!ssp_ok(ssp, false) || n < 0 || (n > 0 && (!met || !age || !dspectra)))
    return fail(RBX_EINVAL, "rbx_ssp_lookup_bwd: bad arguments");
  if (n == 0) return RBX_OK;
  hipLaunchKernelGGL(lookup_bwd_kernel, dim3(blocks_for_rows(n)), dim3(kStepThreads), 0,
                     as_stream(stream), view_of(*ssp), met, age, n, dspectra, dmet, dage);
  return check_launch("rbx_ssp_lookup_bwd");
}

int rbx_scale_rows(const float* in, const float* scale, int64_t n, int32_t nw, float* out,
                   void* stream) {
  using namespace rbx;
  if (n < 0 || nw < 1 || (n > 0 && (!in || !scale || !out)))
    return fail(RBX_EINVAL, "rbx_scale_rows: bad arguments");
  if (n == 0) return RBX_OK;
  hipLaunchKernelGGL(scale_rows_kernel, dim3(grid_for(n * nw, 256)), dim3(256), 0,
                     as_stream(stream), in, scale, n, nw, out);
  return check_launch("rbx_scale_rows");
}

int rbx_row_dot(const float* a, const float* b, int64_t n, int32_t nw, float* out, void* stream) {
  using namespace rbx;
  if (n < 0 || nw < 1 || (n > 0 && (!a || !b || !out)))
    return fail(RBX_EINVAL, "rbx_row_dot: bad arguments");
  if (n == 0) return RBX_OK;
  hipLaunchKernelGGL(row_dot_kernel, dim3(blocks_for_rows(n)), dim3(kStepThreads), 0,
                     as_stream(stream), a, b, n, nw, out);
  return check_launch("rbx_row_dot");
}

int rbx_resample(const rbx_ssp* ssp, const rbx_target* tg, const float* spectra,
                 const float* velocity, int64_t vel_stride, int32_t vel_comp, float c, int64_t n,
                 float* out, void* stream) {
  using namespace rbx;
  if (!ssp_ok(ssp, true) || !tg || tg->n < 1 || !tg->wave || !tg->dwave || n < 0 ||
      vel_comp < 0 || vel_comp >= vel_stride || !(c > 0.f) ||
      (n > 0 && (!spectra || !velocity || !out)))
    return fail(RBX_EINVAL, "rbx_resample: bad arguments");
  if (n == 0) return RBX_OK;
  hipLaunchKernelGGL(resample_kernel, dim3(blocks_for_rows(n)), dim3(kStepThreads), 0,
                     as_stream(stream), view_of(*ssp), *tg, spectra, velocity, vel_stride, vel_comp,
                     c, n, out);
  return check_launch("rbx_resample");
}

int rbx_resample_bwd(const rbx_ssp* ssp, const rbx_target* tg, const float* spectra,
                     const float* velocity, int64_t vel_stride, int32_t vel_comp, float c,
                     int64_t n, const float* dout, float* dspectra, float* dvlos, void* stream) {
  using namespace rbx;
  if (!ssp_ok(ssp, true) || !tg || tg->n < 1 || !tg->wave || !tg->dwave || n < 0 ||
      vel_comp < 0 || vel_comp >= vel_stride || !(c > 0.f) ||
      (n > 0 && (!spectra || !velocity || !dout)))
    return fail(RBX_EINVAL, "rbx_resample_bwd: bad arguments");
  const size_t lds = resample_bwd_lds(tg->n, ssp->n_wave);
  if (lds > 64 * 1024 || tg->n > 65535 || ssp->n_wave > 65533)
    return fail(RBX_EINVAL, "rbx_resample_bwd: 10*n_target + 2*n_wave bytes must fit 64 KiB of LDS");
  if (n == 0) return RBX_OK;
  hipLaunchKernelGGL(resample_bwd_kernel, dim3(blocks_for_rows(n)), dim3(kStepThreads), lds,
                     as_stream(stream), view_of(*ssp), *tg, spectra, velocity, vel_stride, vel_comp,
                     c, n, dout, dspectra, dvlos);
  return check_launch("rbx_resample_bwd");
}

size_t rbx_segment_cube_workspace_size(int64_t n, int32_t n_keys, int32_t n_wave) {
  using namespace rbx;
  if (n < 0 || n_keys < 1 || n_wave < 1) return 0;
  SizeOnly2 so;
  seg_layout(so, n, n_keys, n_wave, nullptr);
  return so.s.used;
}

int rbx_segment_cube(const float* spectra, const int32_t* ids, int64_t n, int32_t n_spaxel,
                     int32_t n_keys, int32_t n_wave, float* cube, void* workspace,
                     size_t workspace_bytes, void* stream) {
  using namespace rbx;
  if (n < 0 || n_spaxel < 1 || n_keys < n_spaxel || n_wave < 1 || !cube ||
      (n > 0 && (!spectra || !ids)))
    return fail(RBX_EINVAL, "rbx_segment_cube: bad arguments");
  if (n > INT32_MAX - 1) return fail(RBX_EINVAL, "rbx_segment_cube: n must fit int32");
  hipStream_t s = as_stream(stream);
  Carver cv(workspace, workspace_bytes);
  SegLayout L;
  seg_layout(cv, n, n_keys, n_wave, &L);
  if (!cv.ok) return fail(RBX_EWORKSPACE, "rbx_segment_cube: workspace too small");
  int rc = spaxel_sort_impl(ids, n, n_keys, L.perm, nullptr, L.offsets, L.sort_ws, L.sort_ws_bytes, s);
  if (rc) return rc;
  rc = build_chunks(L.offsets, n_spaxel, kChunk, L.chunk_off, L.scan_blocks, L.desc, s);
  if (rc) return rc;
  const int n_tiles = (n_wave + 63) / 64;
  const int64_t max_chunks = (n + kChunk - 1) / kChunk + n_spaxel;
  if (n > 0) {
    const int64_t blocks = int64_t(n_tiles) * max_chunks;
    if (blocks > INT32_MAX) return fail(RBX_EINVAL, "rbx_segment_cube: problem too large");
    hipLaunchKernelGGL(segment_accum_kernel, dim3(blocks), dim3(64), 0, s, spectra, L.perm, n_wave,
                       L.desc, L.chunk_off + n_spaxel, n_tiles, L.partial);
  }
  combine_chunks(L.partial, L.chunk_off, n_spaxel, n_wave, cube, s);
  return check_launch("rbx_segment_cube");
}

int rbx_segment_cube_bwd(const float* dcube, const int32_t* ids, int64_t n, int32_t n_spaxel,
                         int32_t n_wave, float* dspectra, void* stream) {
  using namespace rbx;
  if (n < 0 || n_spaxel < 1 || n_wave < 1 || (n > 0 && (!dcube || !ids || !dspectra)))
    return fail(RBX_EINVAL, "rbx_segment_cube_bwd: bad arguments");
  if (n == 0) return RBX_OK;
  hipLaunchKernelGGL(segment_bwd_kernel, dim3(grid_for(n * n_wave, 256)), dim3(256), 0,
                     as_stream(stream), dcube, ids, n, n_spaxel, n_wave, dspectra);
  return check_launch("rbx_segment_cube_bwd");
}

}  // extern "C"
