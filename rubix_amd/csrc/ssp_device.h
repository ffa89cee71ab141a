// Device-side helpers for the SSP grid, shared by the fused and the
// per-step kernels: bilinear cell (interp2d linear,
// rubix/spectra/ssp/grid.py:113-120), shifted knots w_k = fl(wave_obs[k] * a)
// (rubix/spectra/ifu.py:190) and the jnp.interp segment bookkeeping
// (rubix/spectra/ifu.py:250).
#pragma once

#include "rbx_common.h"

namespace rbx {

constexpr float kLn10 = 2.302585092994046f;
constexpr int kChunk = 256;  // stars per reduction chunk of the segmented sums

struct ChunkDesc;
// Per-spaxel chunk descriptors (at most `chunk` stars each) from the sort
// offsets (cube.hip); chunk_off gets n_spaxel + 1 entries, the last one the
// number of chunks.
int build_chunks(const int32_t* offsets, int32_t ns, int chunk, int32_t* chunk_off,
                 int32_t* scan_blocks, ChunkDesc* desc, hipStream_t s);
// cube[p, j] = sum of partial[c, j] over the chunks of spaxel p, in order;
// with bands > 1 the chunk offsets are indexed by band * ns + p and each
// spaxel's bands are added in band order.
void combine_chunks(const double* partial, const int32_t* chunk_off, int32_t ns, int32_t nt,
                    float* cube, hipStream_t s, int32_t bands = 1);
int spaxel_sort_impl(const int32_t* ids, int64_t n, int32_t n_keys, int32_t* perm,
                     int32_t* sorted, int32_t* offsets, void* ws, size_t ws_bytes,
                     hipStream_t s);
size_t sort_workspace_size(int64_t n, int32_t n_keys);
size_t sort_workspace_bound(int64_t n, int32_t max_keys);  // any key count <= max_keys
int exclusive_scan_i32(int32_t* data, int64_t n, int32_t* blocks, hipStream_t s);
void fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);

struct ChunkDesc {
  int32_t spaxel, begin, end, pad;
};

struct SspView {
  const float* __restrict__ flux;
  const float* __restrict__ ax_met;
  const float* __restrict__ ax_age;
  const float* __restrict__ wave;
  const double* __restrict__ cum;
  const int32_t* __restrict__ loc;  // n_loc entries {k, wave[b], wave[b+1], wave[b+2]}
  int n_met, n_age, n_wave, n_loc;
  float loc_x0, loc_inv_h;
  int log_axes;
  int loc_multi;  // cells may hold several knots (coarse locator)
};

inline SspView view_of(const rbx_ssp& s) {
  SspView v;
  v.flux = s.flux; v.ax_met = s.ax_met; v.ax_age = s.ax_age; v.wave = s.wave_obs;
  v.cum = s.cum; v.loc = s.locate; v.n_met = s.n_met; v.n_age = s.n_age;
  v.n_wave = s.n_wave; v.n_loc = s.n_locate; v.loc_x0 = s.locate_x0;
  v.loc_inv_h = s.locate_inv_h; v.log_axes = s.log_axes;
  v.loc_multi = s.locate_multi;
  return v;
}

// Doppler factor a = exp(v / c) (rubix/spectra/ifu.py:190): v / c in float32
// as the reference divides, exp evaluated in double and rounded once, i.e.
// the correctly rounded float32 exponential.
__device__ inline float doppler_factor(float v, float c) {
  return float(exp(double(v / c)));
}

// Correctly rounded float product that the compiler may not fuse into an
// FMA with a later add (HIP's __fmul_rn is a plain, contractable multiply).
__device__ inline float mul_nocontract(float x, float y) {
#pragma clang fp contract(off)
  return x * y;
}

// Shifted knot, rounded like the reference's w = fl(wave * a): the product
// is rounded before any use (x - w must not become fma(-wave, a, x)).
__device__ inline float knot(const SspView& g, int k, float a) {
  return mul_nocontract(g.wave[k], a);
}

// Locator cell of an unshifted wavelength y (cell 0 = below the first knot;
// NaN maps to cell 0).
__device__ inline int loc_cell(const SspView& g, float y) {
  float fm = (y - g.loc_x0) * g.loc_inv_h + 1.0f;
  fm = fminf(fmaxf(fm, 0.f), float(g.n_loc - 1));
  return int(fm);
}

// Number of knots with fl(wave[k] * a) <= x, minus one (in [-1, n-1]):
// searchsorted(w, x, 'right') - 1 with the exact float predicate.  The
// locator table gives a start within a knot or two; the walks terminate
// after at most n steps for any input.
__device__ inline int locate(const SspView& g, float x, float a, float inv_a) {
  int k = g.loc[4 * loc_cell(g, x * inv_a)];
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
  // fractional position in the cell, fu = (u - x1)/(x2 - x1), fs = (s - y1)/(y2 - y1),
  // and the reciprocal cell sizes: with A = f10 - f00, C = f01 - f00 and
  // E = f11 - f10 - f01 + f00 the corner sums are
  //   sum_c w_c f_c  = f00 + fu A + fs (C + fu E),
  //   sum_c gu_c f_c = (A + fs E) inv_dx,   sum_c gs_c f_c = (C + fu E) inv_dy.
  float fu, fs, inv_dx, inv_dy;
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
  c.fu = ax1 / (x2 - x1);
  c.fs = ay1 / (y2 - y1);
  c.inv_dx = 1.0f / (x2 - x1);
  c.inv_dy = 1.0f / (y2 - y1);
  c.row = (iz - 1) * na + (ja - 1);
  c.du_dZ = g.log_axes ? 1.0f / (Z * kLn10) : 1.0f;
  c.ds_dt = g.log_axes ? 1.0f / (t * kLn10) : 1.0f;
  return c;
}

__device__ inline int corner_row(int row, int c, int n_age) {
  return row + ((c & 1) ? n_age : 0) + ((c & 2) ? 1 : 0);
}

}  // namespace rbx
