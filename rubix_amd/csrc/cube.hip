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
#include "ssp_device.h"

namespace rbx {


constexpr int kTile = 64;    // target bins per wavefront tile
constexpr int kSlabs = 8;    // lambda slabs, one per XCD

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
    const float a = doppler_factor(v, st.speed_of_light);
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

// ------------------------------------------------------------ knot windows
//
// The hot loops evaluate P_j = jnp.interp(target_j, w, S) for the 64 target
// bins of a tile.  Target bins are increasing, so the knots their segments
// use form a short window [kb, kb + nk) (about 0.8 knots per bin at the
// benchmark shapes).  Lane t loads knot kb + t once (its shifted wavelength
// and its template values); every target lane then finds its segment with a
// 6-step binary search over the wave's knots (ds_bpermute) and fetches the
// two knots it needs the same way.  Per tile that is one template load per
// corner per lane instead of two per target, and no data-dependent walks.
// Tiles whose window does not fit 64 knots, or targets that are not
// increasing, take the per-target path (segment() + direct loads).

struct Win {
  int kb, nk;  // first knot, number of knots
  bool ok;     // window valid and <= 64 knots (uniform over the wave)
  int k;       // this lane's knot index
  float w;     // fl(wave_obs[k] * a)
};

__device__ inline int table_guess(const SspView& g, float x, float inv_a) {
  const float fm = (x * inv_a - g.loc_x0) * g.loc_inv_h;
  int m = 0;
  if (fm > 0.f) m = fm < float(g.n_loc - 1) ? int(fm) : g.n_loc - 1;
  return g.loc[m];
}

// Window for targets in [x_first, x_last] (uniform values).  The locator
// table bounds the first/last needed knot within a cell; margins of 2 and 3
// knots absorb the rounding of x * inv_a, and the result is verified with
// the exact predicate: knot kb must be <= x_first (or kb = 0) and knot ke
// must be > x_last (or ke = n - 1), which makes the windowed segment search
// identical to searchsorted over all knots.
__device__ inline Win make_window(const SspView& g, float x_first, float x_last, float a,
                                  float inv_a, bool ascending) {
  Win W;
  const int n = g.n_wave;
  int kb = table_guess(g, x_first, inv_a) - 2;
  int ke = table_guess(g, x_last, inv_a) + 3;
  kb = kb < 0 ? 0 : (kb > n - 2 ? n - 2 : kb);
  ke = ke > n - 1 ? n - 1 : ke;
  if (ke < kb + 1) ke = kb + 1;
  W.kb = kb;
  W.nk = ke - kb + 1;
  const int lane = threadIdx.x & 63;
  W.k = kb + (lane < W.nk ? lane : W.nk - 1);
  if (W.k > n - 1) W.k = n - 1;
  W.w = knot(g, W.k, a);
  const float w_first = __shfl(W.w, 0, 64);
  const float w_last = __shfl(W.w, (W.nk - 1) & 63, 64);
  W.ok = ascending && W.nk <= 64 && (kb == 0 || w_first <= x_first) &&
         (ke == n - 1 || w_last > x_last);
  return W;
}

// Number of window knots <= x (window sorted; branch-free binary search).
__device__ inline int win_count(const Win& W, float x) {
  int cnt = 0;
#pragma unroll
  for (int step = 32; step >= 1; step >>= 1) {
    const int probe = cnt + step - 1;
    const float wp = __shfl(W.w, probe & 63, 64);
    if (probe < W.nk && wp <= x) cnt += step;
  }
  return cnt;
}

// jnp.interp segment of target x from the window: local knot indices
// li-1, li of the segment, interpolation weight and clamp side.
struct WinSeg {
  int li;
  int side;
  float theta;
  float wl, wr;
};

__device__ inline WinSeg win_segment(const SspView& g, const Win& W, float x) {
  const int n = g.n_wave;
  const int kstar = W.kb - 1 + win_count(W, x);  // last knot <= x, in [-1, n-1]
  int i = kstar + 1;
  i = i < 1 ? 1 : (i > n - 1 ? n - 1 : i);
  WinSeg s;
  s.li = i - W.kb;  // in [1, nk - 1] for a valid window
  s.wl = __shfl(W.w, (s.li - 1) & 63, 64);
  s.wr = __shfl(W.w, s.li & 63, 64);
  s.side = kstar < 0 ? -1 : ((kstar >= n - 1 && x > s.wr) ? 1 : 0);
  s.theta = (x - s.wl) / (s.wr - s.wl);
  return s;
}

// P_j = jnp.interp(target_j, w, S) with S = m * sum_c w_c f_c, per-target
// path (direct loads of both knots).
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

// P for this lane's target x in the tile spanning [x_first, x_last].
__device__ inline float forward_value(const SspView& g, const StarState& s, float x,
                                      float x_first, float x_last, bool ascending) {
  const Win W = make_window(g, x_first, x_last, s.a, s.inv_a, ascending);
  if (W.ok) {
    const int nw = g.n_wave;
    const float* f00 = g.flux + int64_t(s.row) * nw;
    const float* f10 = f00 + int64_t(g.n_age) * nw;
    const float* f01 = f00 + nw;
    const float* f11 = f10 + nw;
    const int k = W.k;
    const float Fk = s.w[0] * f00[k] + s.w[1] * f10[k] + s.w[2] * f01[k] + s.w[3] * f11[k];
    const WinSeg sg = win_segment(g, W, x);
    const float S0 = __shfl(Fk, (sg.li - 1) & 63, 64) * s.m;
    const float S1 = __shfl(Fk, sg.li & 63, 64) * s.m;
    float P = S0 + sg.theta * (S1 - S0);
    if (sg.side < 0) P = S0;
    if (sg.side > 0) P = S1;
    return P;
  }
  const Seg sg = segment(g, x, s.a, s.inv_a);
  return interp_value(g, s, sg);
}

struct TileX {
  int j;
  bool valid;
  float x, x_first, x_last;
};

__device__ inline TileX tile_x(const rbx_target& tg, int tile) {
  TileX t;
  const int j0 = tile * kTile;
  t.j = j0 + (threadIdx.x & 63);
  t.valid = t.j < tg.n;
  t.x = tg.wave[t.valid ? t.j : tg.n - 1];
  t.x_first = tg.wave[j0];
  t.x_last = tg.wave[j0 + kTile - 1 < tg.n ? j0 + kTile - 1 : tg.n - 1];
  return t;
}

// ------------------------------------------------------------ norm

// Slab-mapped: workgroup b runs on XCD b % 8 and handles lambda slab b % 8
// (tiles [slab * tps, slab * tps + tps)) of four stars, one per wavefront, so
// each XCD's L2 holds only its slab of the templates.  Writes the slab's
// partial Nn; star_sf_kernel adds the slabs in a fixed order.
__global__ void __launch_bounds__(256) star_norm_kernel(SspView g, rbx_target tg, int ascending,
                                                        int64_t n, const StarState* __restrict__ state,
                                                        int n_tiles, int tps,
                                                        float* __restrict__ nn_part) {
  const int slab = blockIdx.x & 7;
  const int64_t r = int64_t(blockIdx.x >> 3) * 4 + (threadIdx.x >> 6);
  if (r >= n) return;
  const StarState s = state[r];
  float acc = 0.f;
  if (s.row >= 0) {
    const int t0 = slab * tps;
    const int t1 = t0 + tps < n_tiles ? t0 + tps : n_tiles;
    for (int t = t0; t < t1; ++t) {
      const TileX tx = tile_x(tg, t);
      const float P = forward_value(g, s, tx.x, tx.x_first, tx.x_last, ascending != 0);
      if (tx.valid) acc += P * tg.dwave[tx.j];
    }
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) nn_part[r * kSlabs + slab] = acc;
}

__global__ void star_sf_kernel(int64_t n, const float* __restrict__ nn_part,
                               StarState* __restrict__ state) {
  for (int64_t r = gtid(); r < n; r += int64_t(gridDim.x) * blockDim.x) {
    if (state[r].row < 0) continue;  // sf = Nn = 0 from star_prep
    float Nn = 0.f;
    for (int q = 0; q < kSlabs; ++q) Nn += nn_part[r * kSlabs + q];
    float sf = state[r].T / Nn;
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
    SspView g, rbx_target tg, int ascending, const StarState* __restrict__ state,
    const ChunkDesc* __restrict__ desc, const int32_t* __restrict__ n_chunks_ptr,
    int n_tiles, int tiles_per_xcd, double* __restrict__ partial) {
  const int b = blockIdx.x;
  const int xcd = b & 7;
  const int q = b >> 3;
  const int tile = xcd * tiles_per_xcd + (q % tiles_per_xcd);
  const int chunk = q / tiles_per_xcd;
  if (tile >= n_tiles || chunk >= *n_chunks_ptr) return;
  const ChunkDesc d = desc[chunk];
  const TileX tx = tile_x(tg, tile);
  // double accumulator: the cube is a sum of up to chunk-size positive
  // terms per lane; accumulating in double keeps the forward's rounding at
  // the level of the individual float32 terms.
  double acc = 0.0;
  for (int r = d.begin; r < d.end; ++r) {
    const StarState s = state[r];
    if (s.row < 0) continue;
    const float P = forward_value(g, s, tx.x, tx.x_first, tx.x_last, ascending != 0);
    acc += double(P * s.sf);
  }
  if (tx.valid) partial[int64_t(chunk) * tg.n + tx.j] = acc;
}

__global__ void cube_combine_kernel(const double* __restrict__ partial,
                                    const int32_t* __restrict__ chunk_off, int32_t n_spaxel,
                                    int32_t nt, float* __restrict__ cube) {
  const int64_t total = int64_t(n_spaxel) * nt;
  for (int64_t idx = gtid(); idx < total; idx += int64_t(gridDim.x) * blockDim.x) {
    const int32_t p = int32_t(idx / nt);
    const int32_t j = int32_t(idx - int64_t(p) * nt);
    const int32_t c0 = chunk_off[p], c1 = chunk_off[p + 1];
    double s = 0.0;
    for (int32_t c = c0; c < c1; ++c) s += partial[int64_t(c) * nt + j];
    cube[idx] = float(s);
  }
}

// ------------------------------------------------------------ adjoint

// Per-slab sums of one star (see the file comment): V = sum_j w_j R_c(j) and
// A = sum_j w_j dR_c(j)/da for the weights w = dL/dC (g) and dtarget (d).
// The Doppler sums (A) cancel strongly (sum_j g_j dP_j/da is a small
// difference of terms ~ slope * lambda / dlambda; a float32 autograd of the
// reference graph loses ~3e-4 on d/dv_los), so they are accumulated in
// double; the value sums (V) do not cancel and stay float32.
struct BwdSums {
  float Vg[4], Vd[4];
  double Ag[4], Ad[4];
};

__device__ inline void bwd_accumulate(BwdSums& S, float gj, float dj, float theta, int side,
                                      float kappa, const float f0[4], const float f1[4]) {
  const double gk = side == 0 ? -double(gj) * double(kappa) : 0.0;
  const double dk = side == 0 ? -double(dj) * double(kappa) : 0.0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float df = f1[c] - f0[c];
    float R = f0[c] + theta * df;
    if (side < 0) R = f0[c];
    if (side > 0) R = f1[c];
    S.Vg[c] += gj * R;
    S.Vd[c] += dj * R;
    S.Ag[c] += gk * double(df);
    S.Ad[c] += dk * double(df);
  }
}

// Slab-mapped like star_norm_kernel; writes the 16 sums of (star, slab) to
// vpart / apart for star_grad_kernel.
__global__ void __launch_bounds__(256) star_bwd_kernel(
    SspView g, rbx_target tg, int ascending, int64_t n, const StarState* __restrict__ state,
    const float* __restrict__ dcube, int n_tiles, int tps, float* __restrict__ vpart,
    double* __restrict__ apart) {
  const int slab = blockIdx.x & 7;
  const int64_t r = int64_t(blockIdx.x >> 3) * 4 + (threadIdx.x >> 6);
  if (r >= n) return;
  const StarState s = state[r];
  BwdSums S;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    S.Vg[c] = 0.f; S.Vd[c] = 0.f; S.Ag[c] = 0.0; S.Ad[c] = 0.0;
  }
  if (s.row >= 0) {
    const int nw = g.n_wave;
    const float* fr[4];
    for (int c = 0; c < 4; ++c) fr[c] = g.flux + int64_t(corner_row(s.row, c, g.n_age)) * nw;
    const float* grow = dcube + int64_t(s.spaxel) * tg.n;
    const int t0 = slab * tps;
    const int t1 = t0 + tps < n_tiles ? t0 + tps : n_tiles;
    for (int t = t0; t < t1; ++t) {
      const TileX tx = tile_x(tg, t);
      const float gj = tx.valid ? grow[tx.j] : 0.f;
      const float dj = tx.valid ? tg.dwave[tx.j] : 0.f;
      const Win W = make_window(g, tx.x_first, tx.x_last, s.a, s.inv_a, ascending != 0);
      float f0[4], f1[4];
      float theta, kappa;
      int side;
      if (W.ok) {
        float fk[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) fk[c] = fr[c][W.k];
        const float lam = g.wave[W.k];
        const WinSeg sg = win_segment(g, W, tx.x);
        const int l0 = (sg.li - 1) & 63, l1 = sg.li & 63;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          f0[c] = __shfl(fk[c], l0, 64);
          f1[c] = __shfl(fk[c], l1, 64);
        }
        const float lam0 = __shfl(lam, l0, 64), lam1 = __shfl(lam, l1, 64);
        theta = sg.theta;
        side = sg.side;
        kappa = (lam0 + theta * (lam1 - lam0)) / (sg.wr - sg.wl);
      } else {
        const Seg sg = segment(g, tx.x, s.a, s.inv_a);
        const int i0 = sg.i - 1, i1 = sg.i;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          f0[c] = fr[c][i0];
          f1[c] = fr[c][i1];
        }
        theta = sg.theta;
        side = sg.side;
        kappa = (g.wave[i0] + theta * (g.wave[i1] - g.wave[i0])) / (sg.wr - sg.wl);
      }
      bwd_accumulate(S, gj, dj, theta, side, kappa, f0, f1);
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    S.Vg[c] = wave_sum(S.Vg[c]);
    S.Vd[c] = wave_sum(S.Vd[c]);
    S.Ag[c] = wave_sum(S.Ag[c]);
    S.Ad[c] = wave_sum(S.Ad[c]);
  }
  if ((threadIdx.x & 63) == 0) {
    float* vp = vpart + (r * kSlabs + slab) * 8;
    double* ap = apart + (r * kSlabs + slab) * 8;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      vp[c] = S.Vg[c];
      vp[4 + c] = S.Vd[c];
      ap[c] = S.Ag[c];
      ap[4 + c] = S.Ad[c];
    }
  }
}

// Per star: add the slab sums in slab order and form every gradient in
// closed form.  With P_j = m sum_c w_c R_c(j) and sf = T / Nn:
//   sum_j g_j dP_j/dtheta = m Pg_theta, dNn/dtheta = m Pd_theta,
//   dT/du = m a W_u, dT/ds = m a W_s, dT/da = T / a, dT/dm = T / m,
//   dL/dtheta = sf sum_j g_j dP_j/dtheta + sbar dsf/dtheta, sbar = sum_j g_j P_j
// (dsf/dm = 0; dsf = 0 where T / Nn is not finite, as nan_to_num's gradient).
__global__ void star_grad_kernel(SspView g, rbx_stars st, const StarState* __restrict__ state,
                                 const float* __restrict__ vpart, const double* __restrict__ apart,
                                 float* __restrict__ dmass, float* __restrict__ dage,
                                 float* __restrict__ dmet, float* __restrict__ dvlos) {
  for (int64_t r = gtid(); r < st.n; r += int64_t(gridDim.x) * blockDim.x) {
    const StarState s = state[r];
    const int32_t i = s.orig;
    if (s.row < 0) {
      if (dmass) dmass[i] = 0.f;
      if (dage) dage[i] = 0.f;
      if (dmet) dmet[i] = 0.f;
      if (dvlos) dvlos[i] = 0.f;
      continue;
    }
    double Vg[4] = {0, 0, 0, 0}, Vd[4] = {0, 0, 0, 0}, Ag[4] = {0, 0, 0, 0}, Ad[4] = {0, 0, 0, 0};
    for (int q = 0; q < kSlabs; ++q) {
      const float* vp = vpart + (r * kSlabs + q) * 8;
      const double* ap = apart + (r * kSlabs + q) * 8;
      for (int c = 0; c < 4; ++c) {
        Vg[c] += vp[c];
        Vd[c] += vp[4 + c];
        Ag[c] += ap[c];
        Ad[c] += ap[4 + c];
      }
    }
    const Cell cell = make_cell(g, st.metallicity[i], st.age[i]);
    double win[4];
    window_sums(g, s.row, s.kl, s.kh, win);
    const double m = s.m, a = s.a, sf = s.sf, T = s.T, Nn = s.Nn;
    double Pg_m = 0, Pg_u = 0, Pg_s = 0, Pg_a = 0, Pd_u = 0, Pd_s = 0, Pd_a = 0, W_u = 0, W_s = 0;
    for (int c = 0; c < 4; ++c) {
      const double w = s.w[c], gu = cell.gu[c], gs = cell.gs[c];
      Pg_m += w * Vg[c];
      Pg_u += gu * Vg[c];
      Pg_s += gs * Vg[c];
      Pg_a += w * Ag[c];
      Pd_u += gu * Vd[c];
      Pd_s += gs * Vd[c];
      Pd_a += w * Ad[c];
      W_u += gu * win[c];
      W_s += gs * win[c];
    }
    const double sbar = m * Pg_m;
    const bool live = isfinite(s.T / s.Nn);
    double dsf_u = 0, dsf_s = 0, dsf_a = 0;
    if (live) {
      const double t_n2 = T / (Nn * Nn);
      dsf_u = m * a * W_u / Nn - t_n2 * (m * Pd_u);
      dsf_s = m * a * W_s / Nn - t_n2 * (m * Pd_s);
      dsf_a = (T / a) / Nn - t_n2 * (m * Pd_a);
    }
    const double g_m = sf * Pg_m;
    const double g_u = sf * (m * Pg_u) + sbar * dsf_u;
    const double g_s = sf * (m * Pg_s) + sbar * dsf_s;
    const double g_a = sf * (m * Pg_a) + sbar * dsf_a;
    if (dmass) dmass[i] = float(g_m);
    if (dmet) dmet[i] = float(g_u * cell.du_dZ);
    if (dage) dage[i] = float(g_s * cell.ds_dt);
    if (dvlos) dvlos[i] = float(g_a * a / double(st.speed_of_light));
  }
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
  double* partial;
  float* nn_part;  // (n, kSlabs) slab partials of Nn
  float* vpart;    // (n, kSlabs, 8) adjoint value sums
  double* apart;   // (n, kSlabs, 8) adjoint Doppler sums
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
  auto* part = c.template take<double>(size_t(max_chunks > 0 ? max_chunks : 1) * nt);
  auto* sws = c.template take<char>(sort_bytes);
  const size_t ns = size_t(n > 0 ? n : 1) * kSlabs;
  auto* nnp = c.template take<float>(ns);
  auto* vp = c.template take<float>(ns * 8);
  auto* ap = c.template take<double>(ns * 8);
  if (L) {
    L->nn_part = reinterpret_cast<float*>(nnp);
    L->vpart = reinterpret_cast<float*>(vp);
    L->apart = reinterpret_cast<double*>(ap);
    L->perm = reinterpret_cast<int32_t*>(perm);
    L->skey = reinterpret_cast<int32_t*>(skey);
    L->offsets = reinterpret_cast<int32_t*>(offs);
    L->chunk_off = reinterpret_cast<int32_t*>(coff);
    L->scan_blocks = reinterpret_cast<int32_t*>(blocks);
    L->state = reinterpret_cast<StarState*>(state);
    L->desc = reinterpret_cast<ChunkDesc*>(desc);
    L->partial = reinterpret_cast<double*>(part);
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

int build_chunks(const int32_t* offsets, int32_t ns, int32_t* chunk_off, int32_t* scan_blocks,
                 ChunkDesc* desc, hipStream_t s) {
  hipLaunchKernelGGL(chunk_count_kernel, dim3(grid_for(int64_t(ns) + 1, 256)), dim3(256), 0, s,
                     offsets, ns, chunk_off);
  int rc = exclusive_scan_i32(chunk_off, int64_t(ns) + 1, scan_blocks, s);
  if (rc) return rc;
  hipLaunchKernelGGL(chunk_fill_kernel, dim3(grid_for(ns, 256)), dim3(256), 0, s, offsets, ns,
                     chunk_off, desc);
  return check_launch("build_chunks");
}

void combine_chunks(const double* partial, const int32_t* chunk_off, int32_t ns, int32_t nt,
                    float* cube, hipStream_t s) {
  hipLaunchKernelGGL(cube_combine_kernel, dim3(grid_for(int64_t(ns) * nt, 256)), dim3(256), 0, s,
                     partial, chunk_off, ns, nt, cube);
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
    const int n_tiles = (nt + kTile - 1) / kTile;
    const int tps = (n_tiles + kSlabs - 1) / kSlabs;
    const int64_t blocks = int64_t(kSlabs) * ((n + 3) / 4);
    if (blocks > INT32_MAX) return fail(RBX_EINVAL, "rbx_cube_fwd: too many stars for one launch");
    hipLaunchKernelGGL(star_norm_kernel, dim3(blocks), dim3(256), 0, s, g, *target,
                       target->ascending, n, L.state, n_tiles, tps, L.nn_part);
    hipLaunchKernelGGL(star_sf_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, L.nn_part,
                       L.state);
  }
  const int32_t ns = stars->n_spaxel;
  rc = build_chunks(L.offsets, ns, L.chunk_off, L.scan_blocks, L.desc, s);
  if (rc) return rc;
  const int n_tiles = (nt + kTile - 1) / kTile;
  const int tpx = (n_tiles + 7) / 8;
  const int64_t max_chunks = (n + kChunk - 1) / kChunk + ns;
  if (n > 0) {
    const int64_t blocks = int64_t(8) * tpx * max_chunks;
    if (blocks > INT32_MAX) return fail(RBX_EINVAL, "rbx_cube_fwd: problem too large for one launch");
    hipLaunchKernelGGL(cube_accum_kernel, dim3(blocks), dim3(64), 0, s, g, *target,
                       target->ascending, L.state, L.desc, L.chunk_off + ns, n_tiles, tpx,
                       L.partial);
  }
  combine_chunks(L.partial, L.chunk_off, ns, nt, cube, s);
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
  const int32_t nt = target->n;
  const int n_tiles = (nt + kTile - 1) / kTile;
  const int tps = (n_tiles + kSlabs - 1) / kSlabs;
  const int64_t blocks = int64_t(kSlabs) * ((n + 3) / 4);
  if (blocks > INT32_MAX) return fail(RBX_EINVAL, "rbx_cube_bwd: too many stars for one launch");
  hipLaunchKernelGGL(star_bwd_kernel, dim3(blocks), dim3(256), 0, s, g, *target,
                     target->ascending, n, L.state, dcube, n_tiles, tps, L.vpart, L.apart);
  hipLaunchKernelGGL(star_grad_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, g, *stars,
                     L.state, L.vpart, L.apart, dmass, dage, dmet, dvlos);
  return check_launch("rbx_cube_bwd");
}

}  // extern "C"
