// Fused particle -> datacube path: forward and adjoint.
//
// Forward (replaces calculate_spectra -> scale_spectrum_by_mass ->
// doppler_shift_and_resampling -> calculate_datacube, rubix/core/ifu.py):
//   per star i with mass m, bilinear cell weights w_c over template rows f_c
//   (interp2d linear, rubix/spectra/ssp/grid.py:113-120), a = exp(v/c)
//   (rubix/spectra/ifu.py:190), shifted knots w_k = fl(wave_obs[k] * a):
//     S_k  = m * sum_c w_c f_c[k]
//     T    = sum_{k in mask, k>=1} S_k (w_k - w_{k-1})     (ifu.py:243-247)
//     P_j  = jnp.interp(target_j, w, S)                   (ifu.py:250)
//     Nn   = sum_j P_j dtarget_j                          (ifu.py:252)
//     out_j = P_j * nan_to_num(T / Nn)                    (ifu.py:255-260)
//     cube[spaxel_i, j] += out_j                          (ifu.py:286)
//   Kernels:
//     star_prep   (thread / star, sorted order): cell, weights, a, the
//                 in-range window [kl, kh] with the exact float predicate,
//                 and T in O(1) from per-template prefix sums (cum).
//     star_norm   (wavefront / star): Nn by a fixed-order wavefront reduction,
//                 then sf = nan_to_num(T / Nn).
//     cube_accum  (wavefront / (spaxel chunk, 64-bin lambda tile)): lanes own
//                 target bins; the wavefront walks the chunk's stars in sorted
//                 order accumulating P_j * sf in registers.  No atomics.
//                 Workgroups are mapped so that each XCD works on its own
//                 lambda slab (its L2 keeps that slab of the templates).
//     cube_combine: cube[p, j] = sum of the spaxel's chunk partials in chunk
//                 order (fixed order => bitwise deterministic).
//   Adjoint (star_bwd, wavefront / star): one pass over the target bins
//   gathers dL/dcube[spaxel_i, :] and accumulates, per template corner c,
//   V_c = sum_j g_j R_c(j), A_c = sum_j g_j dR_c(j)/da and the same with the
//   weights dtarget_j, where R_c(j) is template row c interpolated at
//   target_j.  Every per-star gradient (mass, age, metallicity, v_los)
//   follows in closed form from these 16 sums, T, Nn and the prefix-sum
//   window (closed form at the end of star_bwd_kernel).  Spectra are
//   recomputed, never stored.
#include <float.h>
#include <math.h>

#include "rbx_common.h"

namespace rbx {

int spaxel_sort_impl(const int32_t* ids, int64_t n, int32_t n_keys, int32_t* perm,
                     int32_t* sorted, int32_t* offsets, void* ws, size_t ws_bytes,
                     hipStream_t s);
size_t sort_workspace_size(int64_t n, int32_t n_keys);
int exclusive_scan_i32(int32_t* data, int64_t n, int32_t* blocks, hipStream_t s);
void fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);

constexpr int kChunk = 256;   // stars per reduction chunk
constexpr int kTile = 64;     // target bins per wavefront in cube_accum
constexpr float kLn10 = 2.302585092994046f;

struct __align__(16) StarState {
  int32_t row;      // flux row of corner (iz-1, ja-1); -1 = contributes nothing
  int32_t orig;     // input index of the star
  float a, inv_a;   // exp(v/c) and its reciprocal
  float w[4];       // weights of corners 00, 10 (next Z), 01 (next age), 11
  float m, T, sf, Nn;
  int32_t kl, kh;   // T window: knots kl..kh (kl >= 1)
  int32_t spaxel;   // segment id (raw id from the assignment)
  int32_t pad;
};
static_assert(sizeof(StarState) == 64, "StarState must be 64 bytes");
// The C-ABI descriptors are mirrored by ctypes structures; pin their layout.
static_assert(sizeof(rbx_ssp) == 80, "rbx_ssp layout");
static_assert(sizeof(rbx_target) == 32, "rbx_target layout");
static_assert(sizeof(rbx_stars) == 80, "rbx_stars layout");

struct ChunkDesc {
  int32_t spaxel, begin, end, pad;
};

struct SspView {
  const float* __restrict__ flux;
  const float* __restrict__ ax_met;
  const float* __restrict__ ax_age;
  const float* __restrict__ wave;
  const double* __restrict__ cum;
  const int32_t* __restrict__ loc;
  int n_met, n_age, n_wave, n_loc;
  float loc_x0, loc_inv_h;
  int log_axes;
};

static SspView view_of(const rbx_ssp& s) {
  SspView v;
  v.flux = s.flux; v.ax_met = s.ax_met; v.ax_age = s.ax_age; v.wave = s.wave_obs;
  v.cum = s.cum; v.loc = s.locate; v.n_met = s.n_met; v.n_age = s.n_age;
  v.n_wave = s.n_wave; v.n_loc = s.n_locate; v.loc_x0 = s.locate_x0;
  v.loc_inv_h = s.locate_inv_h; v.log_axes = s.log_axes;
  return v;
}

// Shifted knot, rounded like the reference's w = fl(wave * a) (no FMA).
__device__ inline float knot(const SspView& g, int k, float a) {
  return __fmul_rn(g.wave[k], a);
}

// Number of knots with fl(wave[k] * a) <= x, minus one (in [-1, n-1]):
// searchsorted(w, x, 'right') - 1 with the exact float predicate.  The
// locator table gives a start within a knot or two; the walks terminate
// after at most n steps for any input.
__device__ inline int locate(const SspView& g, float x, float a, float inv_a) {
  const float fm = (x * inv_a - g.loc_x0) * g.loc_inv_h;
  int m = 0;
  if (fm > 0.f) m = fm < float(g.n_loc - 1) ? int(fm) : g.n_loc - 1;
  int k = g.loc[m];
  const int n = g.n_wave;
  while (k + 1 < n && knot(g, k + 1, a) <= x) ++k;
  while (k >= 0 && knot(g, k, a) > x) --k;
  return k;
}

// jnp.interp bookkeeping for one target x: segment [i-1, i] with
// i = clip(k+1, 1, n-1), theta = (x - w[i-1]) / (w[i] - w[i-1]), and the
// clamp side (-1: x < w[0] -> S[0]; +1: x > w[n-1] -> S[n-1]; 0 interior).
struct Seg {
  int i;
  int side;
  float theta;
  float wl, wr;
};

__device__ inline Seg segment(const SspView& g, float x, float a, float inv_a) {
  const int k = locate(g, x, a, inv_a);
  const int n = g.n_wave;
  Seg s;
  s.i = k + 1 < 1 ? 1 : (k + 1 > n - 1 ? n - 1 : k + 1);
  s.wl = knot(g, s.i - 1, a);
  s.wr = knot(g, s.i, a);
  s.side = (k < 0) ? -1 : ((k >= n - 1 && x > s.wr) ? 1 : 0);
  s.theta = (x - s.wl) / (s.wr - s.wl);
  return s;
}

// Bilinear cell of one star (interp2d linear): corner weights, their
// derivatives w.r.t. the (possibly log10) query coordinates, and the chain
// factors to the stored metallicity / age.
struct Cell {
  int row;
  bool inside;
  float w[4], gu[4], gs[4];
  float du_dZ, ds_dt;
};

__device__ inline Cell make_cell(const SspView& g, float Z, float t) {
  Cell c;
  const float u = g.log_axes ? log10f(Z) : Z;
  const float s = g.log_axes ? log10f(t) : t;
  const int nm = g.n_met, na = g.n_age;
  c.inside = (u >= g.ax_met[0]) && (u <= g.ax_met[nm - 1]) && (s >= g.ax_age[0]) &&
             (s <= g.ax_age[na - 1]);
  int iz = upper_bound_f(g.ax_met, nm, u);
  int ja = upper_bound_f(g.ax_age, na, s);
  iz = iz < 1 ? 1 : (iz > nm - 1 ? nm - 1 : iz);
  ja = ja < 1 ? 1 : (ja > na - 1 ? na - 1 : ja);
  const float x1 = g.ax_met[iz - 1], x2 = g.ax_met[iz];
  const float y1 = g.ax_age[ja - 1], y2 = g.ax_age[ja];
  const float ax2 = x2 - u, ax1 = u - x1, ay2 = y2 - s, ay1 = s - y1;
  const float inv = 1.0f / ((x2 - x1) * (y2 - y1));
  c.w[0] = ax2 * ay2 * inv;
  c.w[1] = ax1 * ay2 * inv;
  c.w[2] = ax2 * ay1 * inv;
  c.w[3] = ax1 * ay1 * inv;
  c.gu[0] = -ay2 * inv; c.gu[1] = ay2 * inv; c.gu[2] = -ay1 * inv; c.gu[3] = ay1 * inv;
  c.gs[0] = -ax2 * inv; c.gs[1] = -ax1 * inv; c.gs[2] = ax2 * inv; c.gs[3] = ax1 * inv;
  c.row = (iz - 1) * na + (ja - 1);
  c.du_dZ = g.log_axes ? 1.0f / (Z * kLn10) : 1.0f;
  c.ds_dt = g.log_axes ? 1.0f / (t * kLn10) : 1.0f;
  return c;
}

__device__ inline int corner_row(int row, int c, int n_age) {
  return row + ((c & 1) ? n_age : 0) + ((c & 2) ? 1 : 0);
}

// sum_{k=kl..kh} f_c[k] (wave[k] - wave[k-1]) for each corner, in double.
__device__ inline void window_sums(const SspView& g, int row, int kl, int kh, double out[4]) {
  for (int c = 0; c < 4; ++c) {
    if (kh >= kl) {
      const double* cr = g.cum + int64_t(corner_row(row, c, g.n_age)) * g.n_wave;
      out[c] = cr[kh] - cr[kl - 1];
    } else {
      out[c] = 0.0;
    }
  }
}

// ------------------------------------------------------------ prep

__global__ void star_prep_kernel(SspView g, rbx_target tg, rbx_stars st,
                                 const int32_t* __restrict__ perm,
                                 const int32_t* __restrict__ skey,
                                 StarState* __restrict__ out) {
  const int64_t n = st.n;
  for (int64_t r = gtid(); r < n; r += int64_t(gridDim.x) * blockDim.x) {
    const int32_t i = perm[r];
    const int32_t key = skey[r];
    StarState s;
    s.orig = i;
    s.spaxel = key;
    s.pad = 0;
    const float m = st.mass[i];
    const float v = st.velocity[int64_t(i) * st.vel_stride + st.vel_comp];
    const float a = expf(v / st.speed_of_light);
    s.a = a;
    s.inv_a = 1.0f / a;
    s.m = m;
    const Cell c = make_cell(g, st.metallicity[i], st.age[i]);
    const bool keep = c.inside && key >= 0 && key < st.n_spaxel;
    s.row = keep ? c.row : -1;
    for (int q = 0; q < 4; ++q) s.w[q] = c.w[q];
    // in-range window of the mask (w >= min target) & (w <= max target)
    const int klo_le = locate(g, tg.wmin, a, s.inv_a);
    int klo = (klo_le >= 0 && knot(g, klo_le, a) == tg.wmin) ? klo_le : klo_le + 1;
    const int khi = locate(g, tg.wmax, a, s.inv_a);
    s.kl = klo < 1 ? 1 : klo;
    s.kh = khi;
    float T = 0.f;
    if (keep) {
      double ws[4];
      window_sums(g, c.row, s.kl, s.kh, ws);
      double acc = 0.0;
      for (int q = 0; q < 4; ++q) acc += double(c.w[q]) * ws[q];
      T = float(acc * double(a) * double(m));
    }
    s.T = T;
    s.sf = 0.f;
    s.Nn = 0.f;
    out[r] = s;
  }
}

// ------------------------------------------------------------ norm

// P_j = jnp.interp(target_j, w, S) with S = m * sum_c w_c f_c for one bin.
__device__ inline float interp_value(const SspView& g, const StarState& s, const Seg& sg) {
  const int nw = g.n_wave;
  const float* f00 = g.flux + int64_t(s.row) * nw;
  const float* f10 = f00 + int64_t(g.n_age) * nw;
  const float* f01 = f00 + nw;
  const float* f11 = f10 + nw;
  const int i0 = sg.i - 1, i1 = sg.i;
  const float F0 = s.w[0] * f00[i0] + s.w[1] * f10[i0] + s.w[2] * f01[i0] + s.w[3] * f11[i0];
  const float F1 = s.w[0] * f00[i1] + s.w[1] * f10[i1] + s.w[2] * f01[i1] + s.w[3] * f11[i1];
  const float S0 = F0 * s.m, S1 = F1 * s.m;
  float P = S0 + sg.theta * (S1 - S0);
  if (sg.side < 0) P = S0;
  if (sg.side > 0) P = S1;
  return P;
}

__global__ void __launch_bounds__(256) star_norm_kernel(SspView g, rbx_target tg, int64_t n,
                                                        StarState* __restrict__ state) {
  const int lane = threadIdx.x & 63;
  const int64_t r = int64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (r >= n) return;
  const StarState s = state[r];
  if (s.row < 0) return;  // sf = Nn = 0 already
  float acc = 0.f;
  for (int j = lane; j < tg.n; j += 64) {
    const float x = tg.wave[j];
    const Seg sg = segment(g, x, s.a, s.inv_a);
    acc += interp_value(g, s, sg) * tg.dwave[j];
  }
  acc = wave_sum(acc);
  if (lane == 0) {
    const float Nn = acc;
    float sf = s.T / Nn;
    if (isnan(sf)) sf = 0.f;
    else if (isinf(sf)) sf = sf > 0.f ? FLT_MAX : -FLT_MAX;
    state[r].sf = sf;
    state[r].Nn = Nn;
  }
}

// ------------------------------------------------------------ chunks

__global__ void chunk_count_kernel(const int32_t* __restrict__ offsets, int32_t n_spaxel,
                                   int32_t* __restrict__ chunk_off) {
  for (int64_t p = gtid(); p <= n_spaxel; p += int64_t(gridDim.x) * blockDim.x) {
    if (p == n_spaxel) {
      chunk_off[p] = 0;
    } else {
      const int32_t cnt = offsets[p + 1] - offsets[p];
      chunk_off[p] = (cnt + kChunk - 1) / kChunk;
    }
  }
}

__global__ void chunk_fill_kernel(const int32_t* __restrict__ offsets, int32_t n_spaxel,
                                  const int32_t* __restrict__ chunk_off,
                                  ChunkDesc* __restrict__ desc) {
  for (int64_t p = gtid(); p < n_spaxel; p += int64_t(gridDim.x) * blockDim.x) {
    const int32_t b = offsets[p], e = offsets[p + 1];
    int32_t o = chunk_off[p];
    for (int32_t s = b; s < e; s += kChunk, ++o) {
      ChunkDesc d;
      d.spaxel = int32_t(p);
      d.begin = s;
      d.end = s + kChunk < e ? s + kChunk : e;
      d.pad = 0;
      desc[o] = d;
    }
  }
}

// ------------------------------------------------------------ accumulate

__global__ void __launch_bounds__(64) cube_accum_kernel(
    SspView g, rbx_target tg, const StarState* __restrict__ state,
    const ChunkDesc* __restrict__ desc, const int32_t* __restrict__ n_chunks_ptr,
    int n_tiles, int tiles_per_xcd, float* __restrict__ partial) {
  const int b = blockIdx.x;
  const int xcd = b & 7;
  const int q = b >> 3;
  const int tile = xcd * tiles_per_xcd + (q % tiles_per_xcd);
  const int chunk = q / tiles_per_xcd;
  if (tile >= n_tiles || chunk >= *n_chunks_ptr) return;
  const ChunkDesc d = desc[chunk];
  const int j = tile * kTile + threadIdx.x;
  const bool jv = j < tg.n;
  const float x = tg.wave[jv ? j : tg.n - 1];
  float acc = 0.f;
  for (int r = d.begin; r < d.end; ++r) {
    const StarState s = state[r];
    if (s.row < 0) continue;
    const Seg sg = segment(g, x, s.a, s.inv_a);
    acc += interp_value(g, s, sg) * s.sf;
  }
  if (jv) partial[int64_t(chunk) * tg.n + j] = acc;
}

__global__ void cube_combine_kernel(const float* __restrict__ partial,
                                    const int32_t* __restrict__ chunk_off, int32_t n_spaxel,
                                    int32_t nt, float* __restrict__ cube) {
  const int64_t total = int64_t(n_spaxel) * nt;
  for (int64_t idx = gtid(); idx < total; idx += int64_t(gridDim.x) * blockDim.x) {
    const int32_t p = int32_t(idx / nt);
    const int32_t j = int32_t(idx - int64_t(p) * nt);
    const int32_t c0 = chunk_off[p], c1 = chunk_off[p + 1];
    float s = 0.f;
    for (int32_t c = c0; c < c1; ++c) s += partial[int64_t(c) * nt + j];
    cube[idx] = s;
  }
}

// ------------------------------------------------------------ adjoint

__global__ void __launch_bounds__(256) star_bwd_kernel(
    SspView g, rbx_target tg, rbx_stars st, const StarState* __restrict__ state,
    const float* __restrict__ dcube, float* __restrict__ dmass, float* __restrict__ dage,
    float* __restrict__ dmet, float* __restrict__ dvlos) {
  const int lane = threadIdx.x & 63;
  const int64_t r = int64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (r >= st.n) return;
  const StarState s = state[r];
  const int32_t i = s.orig;
  if (s.row < 0) {
    if (lane == 0) {
      if (dmass) dmass[i] = 0.f;
      if (dage) dage[i] = 0.f;
      if (dmet) dmet[i] = 0.f;
      if (dvlos) dvlos[i] = 0.f;
    }
    return;
  }
  const int nw = g.n_wave;
  const float* fr[4];
  for (int c = 0; c < 4; ++c) fr[c] = g.flux + int64_t(corner_row(s.row, c, g.n_age)) * nw;
  const float* grow = dcube + int64_t(s.spaxel) * tg.n;
  float Vg[4] = {0, 0, 0, 0}, Ag[4] = {0, 0, 0, 0}, Vd[4] = {0, 0, 0, 0}, Ad[4] = {0, 0, 0, 0};
  for (int j = lane; j < tg.n; j += 64) {
    const float x = tg.wave[j];
    const float gj = grow[j];
    const float dj = tg.dwave[j];
    const Seg sg = segment(g, x, s.a, s.inv_a);
    const int i0 = sg.i - 1, i1 = sg.i;
    // dtheta/da scaled: dR/da = -(f1 - f0) * kappa for interior targets
    const float kappa = (g.wave[i0] + sg.theta * (g.wave[i1] - g.wave[i0])) / (sg.wr - sg.wl);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const float f0 = fr[c][i0], f1 = fr[c][i1];
      const float df = f1 - f0;
      float R = f0 + sg.theta * df;
      float D = -df * kappa;
      if (sg.side < 0) { R = f0; D = 0.f; }
      if (sg.side > 0) { R = f1; D = 0.f; }
      Vg[c] += gj * R;
      Ag[c] += gj * D;
      Vd[c] += dj * R;
      Ad[c] += dj * D;
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    Vg[c] = wave_sum(Vg[c]);
    Ag[c] = wave_sum(Ag[c]);
    Vd[c] = wave_sum(Vd[c]);
    Ad[c] = wave_sum(Ad[c]);
  }
  if (lane != 0) return;
  const Cell cell = make_cell(g, st.metallicity[i], st.age[i]);
  double win[4];
  window_sums(g, s.row, s.kl, s.kh, win);
  const float m = s.m, a = s.a, sf = s.sf, T = s.T, Nn = s.Nn;
  float Pg_m = 0.f, Pg_u = 0.f, Pg_s = 0.f, Pg_a = 0.f;
  float Pd_u = 0.f, Pd_s = 0.f, Pd_a = 0.f;
  double W_u = 0.0, W_s = 0.0;
  for (int c = 0; c < 4; ++c) {
    Pg_m += s.w[c] * Vg[c];
    Pg_u += cell.gu[c] * Vg[c];
    Pg_s += cell.gs[c] * Vg[c];
    Pg_a += s.w[c] * Ag[c];
    Pd_u += cell.gu[c] * Vd[c];
    Pd_s += cell.gs[c] * Vd[c];
    Pd_a += s.w[c] * Ad[c];
    W_u += double(cell.gu[c]) * win[c];
    W_s += double(cell.gs[c]) * win[c];
  }
  // sum_j g_j dP_j/dtheta = m * Pg_theta, dNn/dtheta = m * Pd_theta,
  // dT/du = m a W_u, dT/ds = m a W_s, dT/da = T / a.
  const float sbar = m * Pg_m;  // sum_j g_j P_j
  const float raw = T / Nn;
  const bool live = isfinite(raw);
  float dsf_u = 0.f, dsf_s = 0.f, dsf_a = 0.f;
  if (live) {
    const float inv_n = 1.0f / Nn;
    const float t_n2 = T * inv_n * inv_n;
    dsf_u = float(double(m) * double(a) * W_u) * inv_n - t_n2 * (m * Pd_u);
    dsf_s = float(double(m) * double(a) * W_s) * inv_n - t_n2 * (m * Pd_s);
    dsf_a = (T / a) * inv_n - t_n2 * (m * Pd_a);
  }
  const float g_m = sf * Pg_m;
  const float g_u = sf * (m * Pg_u) + sbar * dsf_u;
  const float g_s = sf * (m * Pg_s) + sbar * dsf_s;
  const float g_a = sf * (m * Pg_a) + sbar * dsf_a;
  if (dmass) dmass[i] = g_m;
  if (dmet) dmet[i] = g_u * cell.du_dZ;
  if (dage) dage[i] = g_s * cell.ds_dt;
  if (dvlos) dvlos[i] = g_a * a / st.speed_of_light;
}

// ------------------------------------------------------------ ssp tables

__global__ void __launch_bounds__(256) ssp_cum_kernel(const float* __restrict__ flux,
                                                      const float* __restrict__ wave, int nw,
                                                      double* __restrict__ cum) {
  __shared__ double sm[256];
  const int64_t row = blockIdx.x;
  const float* f = flux + row * nw;
  double* o = cum + row * nw;
  double carry = 0.0;
  for (int b0 = 0; b0 < nw; b0 += 256) {
    const int k = b0 + threadIdx.x;
    double v = 0.0;
    if (k >= 1 && k < nw) v = double(f[k]) * (double(wave[k]) - double(wave[k - 1]));
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      double add = threadIdx.x >= off ? sm[threadIdx.x - off] : 0.0;
      __syncthreads();
      sm[threadIdx.x] += add;
      __syncthreads();
    }
    if (k < nw) o[k] = carry + sm[threadIdx.x];
    carry += sm[255];
    __syncthreads();
  }
}

__global__ void ssp_locate_kernel(const float* __restrict__ wave, int nw, float x0, float inv_h,
                                  int nloc, int32_t* __restrict__ loc) {
  for (int64_t m = gtid(); m < nloc; m += int64_t(gridDim.x) * blockDim.x) {
    const double xm = double(x0) + double(m) / double(inv_h);
    int lo = 0, hi = nw;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (double(wave[mid]) <= xm) lo = mid + 1; else hi = mid;
    }
    loc[m] = lo - 1;
  }
}

// ------------------------------------------------------------ host side

struct CubeLayout {
  int32_t *perm, *skey, *offsets, *chunk_off, *scan_blocks;
  StarState* state;
  ChunkDesc* desc;
  float* partial;
  void* sort_ws;
  size_t sort_ws_bytes;
  int64_t max_chunks;
};

template <class C>
static void cube_layout(C& c, int64_t n, int32_t n_keys, int32_t nt, CubeLayout* L) {
  const int64_t max_chunks = (n + kChunk - 1) / kChunk + n_keys;
  const size_t sort_bytes = sort_workspace_size(n, n_keys);
  auto* perm = c.template take<int32_t>(n > 0 ? n : 1);
  auto* skey = c.template take<int32_t>(n > 0 ? n : 1);
  auto* offs = c.template take<int32_t>(size_t(n_keys) + 1);
  auto* coff = c.template take<int32_t>(size_t(n_keys) + 1);
  auto* blocks = c.template take<int32_t>((size_t(n_keys) + 1) / 4096 + 2);
  auto* state = c.template take<StarState>(n > 0 ? n : 1);
  auto* desc = c.template take<ChunkDesc>(max_chunks > 0 ? max_chunks : 1);
  auto* part = c.template take<float>(size_t(max_chunks > 0 ? max_chunks : 1) * nt);
  auto* sws = c.template take<char>(sort_bytes);
  if (L) {
    L->perm = reinterpret_cast<int32_t*>(perm);
    L->skey = reinterpret_cast<int32_t*>(skey);
    L->offsets = reinterpret_cast<int32_t*>(offs);
    L->chunk_off = reinterpret_cast<int32_t*>(coff);
    L->scan_blocks = reinterpret_cast<int32_t*>(blocks);
    L->state = reinterpret_cast<StarState*>(state);
    L->desc = reinterpret_cast<ChunkDesc*>(desc);
    L->partial = reinterpret_cast<float*>(part);
    L->sort_ws = sws;
    L->sort_ws_bytes = sort_bytes;
    L->max_chunks = max_chunks;
  }
}

struct SizeOnly {
  Sizer s;
  template <class T>
  T* take(size_t count) {
    s.take<T>(count);
    return nullptr;
  }
};

static int grid_for(int64_t n, int threads, int64_t cap = 1 << 20) {
  int64_t gsz = (n + threads - 1) / threads;
  if (gsz < 1) gsz = 1;
  if (gsz > cap) gsz = cap;
  return int(gsz);
}

static int validate(const rbx_ssp* ssp, const rbx_target* tg, const rbx_stars* st) {
  if (!ssp || !tg || !st) return fail(RBX_EINVAL, "rbx_cube: null descriptor");
  if (ssp->n_met < 2 || ssp->n_age < 2 || ssp->n_wave < 2)
    return fail(RBX_EINVAL, "rbx_cube: the SSP grid needs >= 2 metallicities, ages and wavelengths");
  if (!ssp->flux || !ssp->ax_met || !ssp->ax_age || !ssp->wave_obs || !ssp->cum || !ssp->locate ||
      ssp->n_locate < 1 || !(ssp->locate_inv_h > 0.f))
    return fail(RBX_EINVAL, "rbx_cube: SSP descriptor incomplete (run rbx_ssp_prepare)");
  if (tg->n < 1 || !tg->wave || !tg->dwave) return fail(RBX_EINVAL, "rbx_cube: empty target grid");
  if (st->n < 0 || st->n > INT32_MAX - 1) return fail(RBX_EINVAL, "rbx_cube: bad star count");
  if (st->n > 0 && (!st->mass || !st->age || !st->metallicity || !st->velocity || !st->ids))
    return fail(RBX_EINVAL, "rbx_cube: null star array");
  if (st->vel_comp < 0 || st->vel_comp >= st->vel_stride)
    return fail(RBX_EINVAL, "rbx_cube: vel_comp out of range");
  if (st->n_spaxel < 1 || st->n_keys < st->n_spaxel)
    return fail(RBX_EINVAL, "rbx_cube: need 1 <= n_spaxel <= n_keys");
  if (!(st->speed_of_light > 0.f)) return fail(RBX_EINVAL, "rbx_cube: speed_of_light must be > 0");
  return RBX_OK;
}

}  // namespace rbx

extern "C" {

int rbx_ssp_prepare(const rbx_ssp* ssp, double* cum_out, int32_t* locate_out, void* stream) {
  using namespace rbx;
  if (!ssp || !cum_out || !locate_out || !ssp->flux || !ssp->wave_obs || ssp->n_wave < 2 ||
      ssp->n_met < 1 || ssp->n_age < 1 || ssp->n_locate < 1 || !(ssp->locate_inv_h > 0.f))
    return fail(RBX_EINVAL, "rbx_ssp_prepare: bad arguments");
  hipStream_t s = as_stream(stream);
  const int64_t rows = int64_t(ssp->n_met) * ssp->n_age;
  hipLaunchKernelGGL(ssp_cum_kernel, dim3(rows), dim3(256), 0, s, ssp->flux, ssp->wave_obs,
                     ssp->n_wave, cum_out);
  hipLaunchKernelGGL(ssp_locate_kernel, dim3(grid_for(ssp->n_locate, 256)), dim3(256), 0, s,
                     ssp->wave_obs, ssp->n_wave, ssp->locate_x0, ssp->locate_inv_h,
                     ssp->n_locate, locate_out);
  return check_launch("rbx_ssp_prepare");
}

size_t rbx_cube_workspace_size(int64_t n, int32_t n_keys, int32_t n_target) {
  using namespace rbx;
  if (n < 0 || n_keys < 1 || n_target < 1) return 0;
  SizeOnly so;
  cube_layout(so, n, n_keys, n_target, nullptr);
  return so.s.used;
}

int rbx_cube_fwd(const rbx_ssp* ssp, const rbx_target* target, const rbx_stars* stars,
                 float* cube, void* workspace, size_t workspace_bytes, void* stream) {
  using namespace rbx;
  int rc = validate(ssp, target, stars);
  if (rc) return rc;
  if (!cube) return fail(RBX_EINVAL, "rbx_cube_fwd: null cube");
  hipStream_t s = as_stream(stream);
  const int64_t n = stars->n;
  const int32_t nt = target->n;
  Carver c(workspace, workspace_bytes);
  CubeLayout L;
  cube_layout(c, n, stars->n_keys, nt, &L);
  if (!c.ok) return fail(RBX_EWORKSPACE, "rbx_cube_fwd: workspace too small");
  const SspView g = view_of(*ssp);
  rc = spaxel_sort_impl(stars->ids, n, stars->n_keys, L.perm, L.skey, L.offsets, L.sort_ws,
                        L.sort_ws_bytes, s);
  if (rc) return rc;
  if (n > 0) {
    hipLaunchKernelGGL(star_prep_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, g, *target,
                       *stars, L.perm, L.skey, L.state);
    hipLaunchKernelGGL(star_norm_kernel, dim3((n + 3) / 4), dim3(256), 0, s, g, *target, n,
                       L.state);
  }
  const int32_t ns = stars->n_spaxel;
  hipLaunchKernelGGL(chunk_count_kernel, dim3(grid_for(int64_t(ns) + 1, 256)), dim3(256), 0, s,
                     L.offsets, ns, L.chunk_off);
  rc = exclusive_scan_i32(L.chunk_off, int64_t(ns) + 1, L.scan_blocks, s);
  if (rc) return rc;
  hipLaunchKernelGGL(chunk_fill_kernel, dim3(grid_for(ns, 256)), dim3(256), 0, s, L.offsets, ns,
                     L.chunk_off, L.desc);
  const int n_tiles = (nt + kTile - 1) / kTile;
  const int tpx = (n_tiles + 7) / 8;
  const int64_t max_chunks = (n + kChunk - 1) / kChunk + ns;
  if (n > 0) {
    const int64_t blocks = int64_t(8) * tpx * max_chunks;
    if (blocks > INT32_MAX) return fail(RBX_EINVAL, "rbx_cube_fwd: problem too large for one launch");
    hipLaunchKernelGGL(cube_accum_kernel, dim3(blocks), dim3(64), 0, s, g, *target, L.state,
                       L.desc, L.chunk_off + ns, n_tiles, tpx, L.partial);
  }
  hipLaunchKernelGGL(cube_combine_kernel, dim3(grid_for(int64_t(ns) * nt, 256)), dim3(256), 0, s,
                     L.partial, L.chunk_off, ns, nt, cube);
  return check_launch("rbx_cube_fwd");
}

int rbx_cube_bwd(const rbx_ssp* ssp, const rbx_target* target, const rbx_stars* stars,
                 const float* dcube, float* dmass, float* dage, float* dmet, float* dvlos,
                 void* workspace, size_t workspace_bytes, void* stream) {
  using namespace rbx;
  int rc = validate(ssp, target, stars);
  if (rc) return rc;
  if (!dcube) return fail(RBX_EINVAL, "rbx_cube_bwd: null dcube");
  hipStream_t s = as_stream(stream);
  const int64_t n = stars->n;
  Carver c(workspace, workspace_bytes);
  CubeLayout L;
  cube_layout(c, n, stars->n_keys, target->n, &L);
  if (!c.ok) return fail(RBX_EWORKSPACE, "rbx_cube_bwd: workspace too small");
  if (n == 0) return RBX_OK;
  const SspView g = view_of(*ssp);
  hipLaunchKernelGGL(star_bwd_kernel, dim3((n + 3) / 4), dim3(256), 0, s, g, *target, *stars,
                     L.state, dcube, dmass, dage, dmet, dvlos);
  return check_launch("rbx_cube_bwd");
}

}  // extern "C"
