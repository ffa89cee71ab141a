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
//     cube_fused  (workgroup of 16 wavefronts / spaxel chunk, two per lambda
//                 slab): per star of the chunk, P_j on every bin once from the
//                 knot window of each wavefront's 64 NB bins staged in its LDS,
//                 Nn from 64 row shares after a barrier (one per two stars),
//                 sf = nan_to_num(T/Nn), then P_j * sf into f64 registers.
//                 No atomics.  (Ascending targets of at most 8192 bins.)
//     otherwise   star_norm (wavefront / (star, lambda slab)): partial Nn of
//                 the slab's bins, star_sf adds the 8 slabs in order; then
//                 cube_accum (wavefront / (spaxel chunk, 64-bin tile)) walks
//                 the chunk's stars accumulating P_j * sf in registers.
//     cube_combine: cube[p, j] = sum of the spaxel's chunk partials in chunk
//                 order (fixed order => bitwise deterministic).
//   The per-(star, slab) kernels put lambda slab b % S on XCD b % 8 (S = 8
//   for star_norm, 2 for the adjoint), so each XCD's L2 holds only its
//   slab's share of the templates.
//   Adjoint: star_bwd (wavefront / (star, lambda slab)) gathers
//   dL/dcube[spaxel_i, :] over the slab's bins and accumulates seven sums of
//   the combined template rows R_w (bilinear weights), R_u, R_s (their
//   derivatives in log Z, log t) interpolated at target_j and of dR_w/da,
//   weighted by g_j and dtarget_j; star_grad adds the slabs in order and
//   forms every per-star gradient (mass, age, metallicity, v_los) in closed
//   form from these sums, T, Nn and the prefix-sum window.  Spectra are
//   recomputed, never stored.
#include <float.h>
#include <math.h>

#include <type_traits>

#include "rbx_common.h"
#include "ssp_device.h"
#include "cube_impl.h"

namespace rbx {


constexpr int kTile = 64;    // target bins per wavefront tile
constexpr int kSlabs = 8;    // lambda slabs, one per XCD
constexpr int kBwdSlabs = 2; // lambda slabs of the adjoint's (star, slab) items

struct __align__(16) StarState {
  int32_t row;      // flux row of corner (iz-1, ja-1); -1 = contributes nothing
  int32_t orig;     // input index of the star
  float a, inv_a;   // exp(v/c) and its reciprocal
  float w[4];       // weights of corners 00, 10 (next Z), 01 (next age), 11
  float m, T, sf, Nn;
  int32_t kl, kh;   // T window: knots kl..kh (kl >= 1)
  int32_t spaxel;   // segment id (raw id from the assignment)
  float av;         // dust A_V of the star (0 without extinction)
};
static_assert(sizeof(StarState) == 64, "StarState must be 64 bytes");
// The C-ABI descriptors are mirrored by ctypes structures; pin their layout.
static_assert(sizeof(rbx_ssp) == 80, "rbx_ssp layout");
static_assert(sizeof(rbx_target) == 40, "rbx_target layout");
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
                                 const int32_t* __restrict__ skey, const float* __restrict__ av,
                                 StarState* __restrict__ out, float* __restrict__ gradw) {
  const int64_t n = st.n;
  for (int64_t r = gtid(); r < n; r += int64_t(gridDim.x) * blockDim.x) {
    const int32_t i = perm[r];
    const int32_t key = skey[r];
    StarState s;
    s.orig = i;
    s.spaxel = key >= 0 ? key % st.n_spaxel : -1;  // key = band * n_spaxel + spaxel
    s.av = av ? av[i] : 0.f;
    const float m = st.mass[i];
    const float v = st.velocity[int64_t(i) * st.vel_stride + st.vel_comp];
    const float a = doppler_factor(v, st.speed_of_light);
    s.a = a;
    s.inv_a = 1.0f / a;
    s.m = m;
    const Cell c = make_cell(g, st.metallicity[i], st.age[i]);
    const bool keep = c.inside && key >= 0;
    s.row = keep ? c.row : -1;
    for (int q = 0; q < 4; ++q) s.w[q] = c.w[q];
    // the fractional cell position the adjoint's combined rows use
    gradw[r * 2] = c.fu;
    gradw[r * 2 + 1] = c.fs;
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

// ------------------------------------------------------------ age bands
//
// The fused forward walks spaxel chunks with all 8 lambda slabs of a star in
// one workgroup, so an XCD's L2 sees every slab of the templates the stars
// use.  The sort key band * n_spaxel + spaxel orders the stars by the age
// band of their template cell first: chunks, and so the workgroups in
// flight at any time, come from one band, whose template rows
// (n_met x (n_age / bands + 1)) fit an XCD's L2.  Stars of one spaxel in
// different bands form separate chunks; cube_combine adds a spaxel's chunks
// band by band in a fixed order.
constexpr int kMaxBands = 8;
constexpr int32_t kBandKeys = 8192;          // key counts up to this get bands
constexpr int64_t kBandBytes = int64_t(3) << 20;  // template bytes per band

// Bands the workspace is laid out for: a function of the key count only.
static int band_cap(int32_t n_keys) { return n_keys <= kBandKeys ? kMaxBands : 1; }

// Bands for a template grid: about kBandBytes of templates each.
static int band_count(const rbx_ssp& ssp, int32_t n_keys) {
  const int64_t bytes = int64_t(ssp.n_met) * ssp.n_age * ssp.n_wave * int64_t(sizeof(float));
  int64_t b = (bytes + kBandBytes - 1) / kBandBytes;
  if (b > band_cap(n_keys)) b = band_cap(n_keys);
  if (b > ssp.n_age - 1) b = ssp.n_age - 1;
  return b < 1 ? 1 : int(b);
}

// key = band * n_spaxel + id for ids in [0, n_spaxel), -1 (sorted last,
// dropped) otherwise; band from the star's age cell (make_cell's index).
__global__ void band_key_kernel(SspView g, rbx_stars st, int bands, int32_t* __restrict__ keys) {
  for (int64_t r = gtid(); r < st.n; r += int64_t(gridDim.x) * blockDim.x) {
    const int32_t id = st.ids[r];
    int32_t key = -1;
    if (id >= 0 && id < st.n_spaxel) {
      int band = 0;
      if (bands > 1) {
        const float t = st.age[r];
        const float sv = g.log_axes ? log10f(t) : t;
        int ja = upper_bound_f(g.ax_age, g.n_age, sv);
        ja = ja < 1 ? 1 : (ja > g.n_age - 1 ? g.n_age - 1 : ja);
        band = ((ja - 1) * bands) / (g.n_age - 1);
        band = band > bands - 1 ? bands - 1 : band;
      }
      key = band * st.n_spaxel + id;
    }
    keys[r] = key;
  }
}

// ------------------------------------------------------------ segments
//
// The hot loops evaluate P_j = jnp.interp(target_j, w, S) with one lane per
// target bin.  The segment starts from one probe of the locator table: the
// cell of y = x / a gives k0 = (knots <= cell start) - 1 and the next knot,
// and k = k0 + (next <= y).  That guess can be one knot off when x lies
// within rounding distance of a knot, so it is then corrected with the exact
// float32 predicate on the shifted knots w = fl(wave * a) that are loaded
// anyway: step right while x >= w_r, left while x < w_l.  The result is
// searchsorted(w, x, 'right') - 1 exactly, as the reference computes it
// (the segment decides dP/da, which jumps at a knot).  The correction loop
// also covers coarse tables (cells with several knots).

// Element idx of an array with a wave-uniform base: the byte offset is formed
// in 32 bits, so the load can use a scalar base plus a 32-bit vector offset
// (every array read this way is below 4 GiB: checked by validate()).
template <class T>
__device__ inline T ld_at(const T* base, uint32_t idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + idx * uint32_t(sizeof(T)));
}

// Elements idx and idx + 1 in one 8-byte load (4-byte alignment is enough
// for global memory on this target).
struct alignas(4) Pair {
  float lo, hi;
};

__device__ inline Pair ld_pair(const float* base, uint32_t idx) {
  return *reinterpret_cast<const Pair*>(reinterpret_cast<const char*>(base) + idx * 4u);
}

// Elements idx .. idx + 3 in one 16-byte load (4-byte alignment, as above).
struct alignas(4) Quad {
  float v[4];
};

__device__ inline Quad ld_quad(const float* base, uint32_t idx) {
  return *reinterpret_cast<const Quad*>(reinterpret_cast<const char*>(base) + idx * 4u);
}

// Per-star constants of the search: a, 1/a and the locator cell of target x
// as x * c1 + c0.
struct Shift {
  float a, inv_a, c1, c0;
};

__device__ inline Shift shift_of(const SspView& g, float a, float inv_a) {
  Shift h;
  h.a = a;
  h.inv_a = inv_a;
  h.c1 = inv_a * g.loc_inv_h;
  h.c0 = 1.0f - g.loc_x0 * g.loc_inv_h;
  return h;
}

// jnp.interp segment [i0, i0 + 1] of target x: interpolation weight theta
// (0 / 1 at the left / right clamp, 0 for a zero-width segment, as
// jnp.interp's dx == 0 branch) and, for the adjoint, kappa = lambda(x) /
// (w_r - w_l) with lambda(x) = l0 + theta (l1 - l0) on the unshifted knots,
// so that dP/da = -kappa (S1 - S0) inside the grid (0 at the clamps).
struct Knots {
  uint32_t i0;
  float theta;
  float kappa;
};

template <bool kKappa>
__device__ inline Knots knots_of(const SspView& g, float x, const Shift& h) {
  const int n = g.n_wave;
  float fm = fmaf(x, h.c1, h.c0);
  fm = fminf(fmaxf(fm, 0.f), float(g.n_loc - 1));
  const int4 e = ld_at(reinterpret_cast<const int4*>(g.loc), uint32_t(fm));
  const float y = x * h.inv_a;
  const float e1 = __int_as_float(e.y), e2 = __int_as_float(e.z), e3 = __int_as_float(e.w);
  const int k = e.x + (e2 <= y ? 1 : 0);  // e2 = the knot after k0 (see ssp_locate_kernel)
  int i = k + 1 < 1 ? 1 : (k + 1 > n - 1 ? n - 1 : k + 1);
  // the segment's knots from the entry: i - 1 - b is 0 or 1
  const int b = e.x < n - 2 ? e.x : n - 2;
  const bool hi = (i - 1 - b) > 0;
  Pair lp;
  lp.lo = hi ? e2 : e1;
  lp.hi = hi ? e3 : e2;
  float wl = mul_nocontract(lp.lo, h.a), wr = mul_nocontract(lp.hi, h.a);
  // exact segment: i in [1, n-1] moves monotonically, so the loop ends.
  // Corrections are rare; a wave vote keeps the common case to one
  // compare and a uniform branch.
  bool step = (x >= wr && i < n - 1) || (x < wl && i > 1);
  if (__ballot(step)) {
    while (step) {
      i += x >= wr ? 1 : -1;
      lp = ld_pair(g.wave, uint32_t(i - 1));
      wl = mul_nocontract(lp.lo, h.a);
      wr = mul_nocontract(lp.hi, h.a);
      step = (x >= wr && i < n - 1) || (x < wl && i > 1);
    }
  }
  const uint32_t i0 = uint32_t(i - 1);
  const float l0 = lp.lo, l1 = lp.hi;
  const float dx = wr - wl;
  const float r = __builtin_amdgcn_rcpf(dx);
  const bool flat = !(dx > 0.f);
  const float th = flat ? 0.f : (x - wl) * r;
  const bool left = (i == 1) && (x < wl);
  const bool right = (i == n - 1) && (x > wr);
  Knots K;
  K.i0 = i0;
  K.theta = left ? 0.f : (right ? 1.f : th);
  K.kappa = 0.f;
  if (kKappa) K.kappa = (left || right || flat) ? 0.f : (l0 + th * (l1 - l0)) * r;
  return K;
}

// Row pointers of a star's four template corners (uniform over the wave).
struct Rows {
  const float* f[4];
};

__device__ inline Rows rows_of(const SspView& g, int row) {
  Rows R;
  const int nw = g.n_wave;
#pragma unroll
  for (int c = 0; c < 4; ++c) R.f[c] = g.flux + int64_t(corner_row(row, c, g.n_age)) * nw;
  return R;
}

// P = S0 + theta (S1 - S0), S = m * sum_c w_c f_c (the spectrum of
// calculate_spectra * mass at the two knots).
__device__ inline float star_value(const Rows& R, const StarState& s, const Knots& K) {
  float F0 = 0.f, F1 = 0.f;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const Pair f = ld_pair(R.f[c], K.i0);
    F0 += s.w[c] * f.lo;
    F1 += s.w[c] * f.hi;
  }
  const float S0 = F0 * s.m, S1 = F1 * s.m;
  return S0 + K.theta * (S1 - S0);
}

// ------------------------------------------------------------ LDS windows
//
// A (star, lambda slab) item touches only the knots between the segments of
// the slab's first and last bins (targets ascending).  The window's
// unshifted knots and the star's combined template F_k = sum_c w_c f_c[k]
// are loaded once with coalesced loads into the wave's LDS; the bins then
// interpolate from LDS instead of gathering four template pairs each
// through the L1.
constexpr int kWindow = 640;  // knots per staged window (LDS floats per array)

// Order a wave's LDS stores before its later loads (the wave's own data;
// LDS instructions of one wave complete in order).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Segment of target x inside a staged window of knots [k_lo, k_lo + nk):
// the locator guess clamped into the window, then the exact correction on
// the window's knots (for ascending targets the true segment lies inside).
struct WinKnots {
  int li;       // local left knot, in [0, nk - 2]
  float theta;  // as Knots::theta
  float kappa;  // as Knots::kappa (only when requested)
};

// kInterior: the window lies strictly inside the grid (k_lo >= 1 and
// k_lo + nk < n_wave) and was built by slab_window from locator guesses
// (locate_multi == 0), so no segment of it touches knot 0 or n_wave - 1:
// the grid-end bounds of the correction and the left / right clamps cannot
// apply and are left out (same result, fewer operations per bin).  The
// guess needs no clamp either: k0 is monotone in x, so for the slab's bins
// k0(x_first) <= k0 <= k0(x_last), and the window's margins (3 knots below
// the first guess, 4 above the last) put li = k0 - k_lo in [2, nk - 5].
// The search then ends on the exact segment w_l <= x < w_r, whose width
// is positive.
// Locator guess k0 of target x: the compact locator copy (4 bytes per cell)
// holds the number of knots <= the cell start, minus one.
__device__ inline int loc_k0(const SspView& g, float x, const Shift& h) {
  const float fm = __builtin_amdgcn_fmed3f(fmaf(x, h.c1, h.c0), 0.f, float(g.n_loc - 1));
  return ld_at(g.loc + 4 * g.n_loc, uint32_t(fm));
}

// win_knots with the locator guess already loaded (k0 = loc_k0(g, x, h)).
template <bool kKappa, bool kInterior = false>
__device__ inline WinKnots win_knots_k0(const SspView& g, float x, const Shift& h, int k_lo,
                                        int nk, const float* lam, int k0) {
  const int n = g.n_wave;
  const float y = x * h.inv_a;
  const int li_max = nk - 2, li_min = 0;
  int li = k0 - k_lo;
  if (!kInterior) li = li < 0 ? 0 : (li > li_max ? li_max : li);
  // next-knot test, kept inside [0, nk - 2] (lam[li + 1] is in the window)
  li += lam[li + 1] <= y ? 1 : 0;
  if (!kInterior) li = li > li_max ? li_max : li;
  float l0 = lam[li], l1 = lam[li + 1];
  float wl = mul_nocontract(l0, h.a), wr = mul_nocontract(l1, h.a);
  // the global segment [li + k_lo, li + k_lo + 1] moves within [0, n - 1]
  auto needs_step = [&](int l, float lo, float hi) {
    if (kInterior) return (x >= hi && l < li_max) || (x < lo && l > li_min);
    return (x >= hi && l < li_max && l + k_lo + 1 < n - 1) || (x < lo && l > li_min && l + k_lo > 0);
  };
  bool step = needs_step(li, wl, wr);
  if (__ballot(step)) {
    while (step) {
      li += x >= wr ? 1 : -1;
      l0 = lam[li];
      l1 = lam[li + 1];
      wl = mul_nocontract(l0, h.a);
      wr = mul_nocontract(l1, h.a);
      step = needs_step(li, wl, wr);
    }
  }
  const int i = li + k_lo + 1;  // global right knot
  const float dx = wr - wl;
  const float r = __builtin_amdgcn_rcpf(dx);
  const bool flat = !kInterior && !(dx > 0.f);
  const float th = flat ? 0.f : (x - wl) * r;
  const bool left = !kInterior && (i == 1) && (x < wl);
  const bool right = !kInterior && (i == n - 1) && (x > wr);
  WinKnots K;
  K.li = li;
  K.theta = left ? 0.f : (right ? 1.f : th);
  K.kappa = 0.f;
  if (kKappa) K.kappa = (left || right || flat) ? 0.f : (l0 + th * (l1 - l0)) * r;
  return K;
}

template <bool kKappa, bool kInterior = false>
__device__ inline WinKnots win_knots(const SspView& g, float x, const Shift& h, int k_lo, int nk,
                                     const float* lam) {
  return win_knots_k0<kKappa, kInterior>(g, x, h, k_lo, nk, lam, loc_k0(g, x, h));
}

// P_j of NB bins of a lane from a staged window of shifted knots W[k] =
// w_{k_lo+k} and spectrum values S[k] = S_{k_lo+k}, the bins interleaved:
// the locator guess k0, the next-knot test on the shifted knot, then the
// exact correction (one wave vote for the rare step) to the segment with
// w_l <= x < w_r, as jnp.interp's searchsorted, and S_l + theta (S_r - S_l).
// The search ends on the same segment from any guess inside the window, so
// the values are those of win_knots_k0 on the unshifted knots.
// kInterior: the window lies strictly inside the grid (k_lo >= 1 and
// k_lo + nk < n_wave) and holds every bin's guess with a margin of 3 knots
// below and 5 above (the guesses' error bound), so no segment touches knot
// 0 or n_wave - 1: the grid-end bounds, the flat-segment case and the
// clamps cannot apply and are left out.
template <int NB, bool kInterior>
__device__ inline void win_values(const float (&x)[NB], const int (&k0)[NB], int k_lo, int nk,
                                  int n_wave, const float* W, const float* S, float (&P)[NB]) {
  const int li_max = nk - 2;
  int li[NB];
  float wl[NB], wr[NB];
  auto needs_step = [&](int l, float xv, float lo, float hi) {
    if (kInterior) return (xv >= hi && l < li_max) | (xv < lo && l > 0);
    return (xv >= hi && l < li_max && l + k_lo + 1 < n_wave - 1) |
           (xv < lo && l > 0 && l + k_lo > 0);
  };
  bool any = false;
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    int l = k0[t] - k_lo;
    if (kInterior) {
      // the three knots of both candidate segments in one round of reads,
      // then the next-knot test picks one (no read waits on the test)
      const float w0 = W[l], w1 = W[l + 1], w2 = W[l + 2];
      const bool step = w1 <= x[t];
      li[t] = l + (step ? 1 : 0);
      wl[t] = step ? w1 : w0;
      wr[t] = step ? w2 : w1;
    } else {
      l = l < 0 ? 0 : (l > li_max ? li_max : l);
      l += W[l + 1] <= x[t] ? 1 : 0;
      l = l > li_max ? li_max : l;
      li[t] = l;
      wl[t] = W[l];
      wr[t] = W[l + 1];
    }
    // interior: l lies in [3, nk - 5] (the window's margins), inside both
    // bounds of the step, so a step is needed exactly when x is off [w_l, w_r)
    if (kInterior)
      any |= (x[t] >= wr[t]) | (x[t] < wl[t]);
    else
      any |= needs_step(l, x[t], wl[t], wr[t]);
  }
  if (__ballot(any)) {
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      while (needs_step(li[t], x[t], wl[t], wr[t])) {
        li[t] += x[t] >= wr[t] ? 1 : -1;
        wl[t] = W[li[t]];
        wr[t] = W[li[t] + 1];
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    const float dx = wr[t] - wl[t];
    const float r = __builtin_amdgcn_rcpf(dx);
    float th = (x[t] - wl[t]) * r;
    if (!kInterior) {
      const int i = li[t] + k_lo + 1;  // global right knot
      const bool flat = !(dx > 0.f);
      const bool left = (i == 1) && (x[t] < wl[t]);
      const bool right = (i == n_wave - 1) && (x[t] > wr[t]);
      th = left ? 0.f : (right ? 1.f : (flat ? 0.f : th));
    }
    const float S0 = S[li[t]], S1 = S[li[t] + 1];
    P[t] = S0 + th * (S1 - S0);
  }
}

// Locator guess of the left knot of x's segment: k0 of x's cell plus the
// next-knot test on y = x / a.  With at most one knot per cell
// (locate_multi == 0) it is within 3 knots of the exact segment: one cell
// of rounding in the cell index, one knot of rounding in y, one of the
// shifted knots' rounding.
__device__ inline int locate_guess(const SspView& g, float x, const Shift& h) {
  float fm = fmaf(x, h.c1, h.c0);
  fm = fminf(fmaxf(fm, 0.f), float(g.n_loc - 1));
  const int k0 = ld_at(g.loc + 4 * g.n_loc, uint32_t(fm));
  const int kn = k0 + 1 < g.n_wave - 1 ? k0 + 1 : g.n_wave - 1;
  return k0 + (ld_at(g.wave, uint32_t(kn < 0 ? 0 : kn)) <= x * h.inv_a ? 1 : 0);
}

// Knot window [k_lo, k_lo + nk) holding every segment of the bins [j0, j1)
// (ascending targets).  With a fine locator: the guesses of the first and
// last bin widened by the guess's error bound (3 knots left, 4 right); the
// bins' exact search then runs inside it.  With a coarse locator: from the
// first bin's exact left knot to the last bin's exact right knot.
__device__ inline void slab_window(const SspView& g, const rbx_target& tg, uint32_t j0,
                                   uint32_t j1, const Shift& h, int* k_lo, int* nk) {
  const float xf = ld_at(tg.wave, j0), xl = ld_at(tg.wave, j1 - 1);
  int lo, hi;
  if (!g.loc_multi) {
    lo = locate_guess(g, xf, h) - 3;
    hi = locate_guess(g, xl, h) + 4;
    lo = lo < 0 ? 0 : lo;
    hi = hi > g.n_wave - 1 ? g.n_wave - 1 : hi;
  } else {
    lo = int(knots_of<false>(g, xf, h).i0);
    hi = int(knots_of<false>(g, xl, h).i0) + 1;
  }
  *k_lo = __builtin_amdgcn_readfirstlane(lo);
  *nk = __builtin_amdgcn_readfirstlane(hi - lo + 1);
}

// Work items of the slab-mapped per-star kernels: item q is lambda slab
// q % S of the W stars [(q / S) W, (q / S) W + W), one per wavefront of a
// workgroup of W waves (S = 8 slabs for star_norm, 2 for the adjoint).
// Workgroup b takes items b, b + G, b + 2G, ... with G = gridDim.x a
// multiple of 8 (G = number of items, or a persistent grid of a few
// workgroups per CU), so it always holds slab b % S and runs on XCD b % 8:
// each XCD's L2 holds only one slab of the templates.
struct SlabTask {
  int slab, wave, lane;
  int64_t r;
};

template <int W, int S = kSlabs>
__device__ inline SlabTask slab_task(int64_t q) {
  static_assert((S & (S - 1)) == 0 && S <= 8, "slab count: a power of two dividing 8");
  SlabTask t;
  t.slab = int(q & (S - 1));
  t.wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  t.lane = threadIdx.x & 63;
  t.r = (q / S) * W + t.wave;
  return t;
}

// ------------------------------------------------------------ norm

// Partial Nn of (star, slab); star_sf_kernel adds the slabs in a fixed order.
template <int W>
__global__ void __launch_bounds__(64 * W) star_norm_kernel(SspView g, rbx_target tg, int64_t n,
                                                           int64_t n_items,
                                                           const StarState* __restrict__ state,
                                                           int slab_bins,
                                                           float* __restrict__ nn_part) {
  __shared__ float lam_s[W][kWindow];
  __shared__ float F_s[W][kWindow];
  for (int64_t q = blockIdx.x; q < n_items; q += gridDim.x) {
    const SlabTask t = slab_task<W>(q);
    if (t.r >= n) continue;
    const StarState s = state[t.r];
    float acc = 0.f;
    const uint32_t j0 = uint32_t(t.slab * slab_bins);
    const uint32_t j1 = j0 + slab_bins < uint32_t(tg.n) ? j0 + slab_bins : uint32_t(tg.n);
    if (s.row >= 0 && j0 < j1) {
      const Rows R = rows_of(g, __builtin_amdgcn_readfirstlane(s.row));
      const Shift h = shift_of(g, s.a, s.inv_a);
      int k_lo = 0, nk = kWindow + 1;
      if (tg.ascending) slab_window(g, tg, j0, j1, h, &k_lo, &nk);
      if (nk <= kWindow) {
        float* lam = lam_s[t.wave];
        float* F = F_s[t.wave];
        for (int k = t.lane; k < nk; k += 64) {
          const uint32_t kk = uint32_t(k_lo + k);
          lam[k] = ld_at(g.wave, kk);
          float f = 0.f;
#pragma unroll
          for (int c = 0; c < 4; ++c) f += s.w[c] * ld_at(R.f[c], kk);
          F[k] = f * s.m;  // S_k = F_k m, rounded once per knot as the reference's spectra
        }
        wave_lds_sync();
        auto bins = [&](auto interior) {
#pragma unroll 2
          for (uint32_t j = j0 + t.lane; j < j1; j += 64) {
            const Pair xd = ld_pair(tg.pairs, 2 * j);
            const WinKnots K = win_knots<false, decltype(interior)::value>(g, xd.lo, h, k_lo, nk, lam);
            const float S0 = F[K.li], S1 = F[K.li + 1];
            acc += (S0 + K.theta * (S1 - S0)) * xd.hi;
          }
        };
        if (k_lo >= 1 && k_lo + nk < g.n_wave && !g.loc_multi) bins(std::true_type{});
        else bins(std::false_type{});
        wave_lds_sync();
      } else {
#pragma unroll 2
        for (uint32_t j = j0 + t.lane; j < j1; j += 64) {
          const Pair xd = ld_pair(tg.pairs, 2 * j);
          const Knots K = knots_of<false>(g, xd.lo, h);
          acc += star_value(R, s, K) * xd.hi;
        }
      }
      acc = wave_sum(acc);
    }
    if (t.lane == 0) nn_part[t.r * kSlabs + t.slab] = acc;
  }
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
                                   int chunk, int32_t* __restrict__ chunk_off) {
  for (int64_t p = gtid(); p <= n_spaxel; p += int64_t(gridDim.x) * blockDim.x) {
    if (p == n_spaxel) {
      chunk_off[p] = 0;
    } else {
      const int32_t cnt = offsets[p + 1] - offsets[p];
      chunk_off[p] = (cnt + chunk - 1) / chunk;
    }
  }
}

__global__ void chunk_fill_kernel(const int32_t* __restrict__ offsets, int32_t n_spaxel,
                                  int chunk, const int32_t* __restrict__ chunk_off,
                                  ChunkDesc* __restrict__ desc) {
  for (int64_t p = gtid(); p < n_spaxel; p += int64_t(gridDim.x) * blockDim.x) {
    const int32_t b = offsets[p], e = offsets[p + 1];
    int32_t o = chunk_off[p];
    for (int32_t s = b; s < e; s += chunk, ++o) {
      ChunkDesc d;
      d.spaxel = int32_t(p);
      d.begin = s;
      d.end = s + chunk < e ? s + chunk : e;
      d.pad = 0;
      desc[o] = d;
    }
  }
}

// ------------------------------------------------------------ accumulate

// Dust attenuation of bin j for a star with extinction A_V:
// 10^((-0.4 A(lambda_j)/A(V)) * A_V), the factor the calculate_extinction
// step multiplies into the resampled spectra (each product rounded on its own).
__device__ inline float dust_tau(const float* axav, uint32_t j) {
  return mul_nocontract(-0.4f, ld_at(axav, j));
}
// 10^y = 2^(y log2 10): one multiply and the hardware exp2 (v_exp_f32,
// about 1 ulp).  The rounding of y log2 10 adds |y log2 10| 2^-24 relative
// error, under 2e-6 for |y| <= 10 (the exponents here are a few units), so
// the factor stays far inside the 1e-5 parity bound.  NaN propagates.
__device__ inline float exp10_fast(float y) {
  return __builtin_amdgcn_exp2f(y * 3.321928094887362f);
}
__device__ inline float dust_factor(float tau, float av) {
  return exp10_fast(mul_nocontract(tau, av));
}

// Item q: 64-bin lambda tile of XCD q % 8's slab and one spaxel chunk; the
// items are strided over the grid like the per-star kernels' (G a multiple
// of 8), so workgroup b stays on slab b % 8.
template <bool kDust>
__global__ void __launch_bounds__(64) cube_accum_kernel(
    SspView g, rbx_target tg, const StarState* __restrict__ state,
    const ChunkDesc* __restrict__ desc, const int32_t* __restrict__ n_chunks_ptr,
    int n_tiles, int tiles_per_xcd, int64_t n_items, const float* __restrict__ axav,
    double* __restrict__ partial) {
  const int n_chunks = *n_chunks_ptr;
  for (int64_t q = blockIdx.x; q < n_items; q += gridDim.x) {
    const int xcd = int(q & 7);
    const uint32_t w = uint32_t(q >> 3);  // n_items < 2^34 (checked on the host)
    const uint32_t chunk = w / uint32_t(tiles_per_xcd);
    const int tile = xcd * tiles_per_xcd + int(w - chunk * uint32_t(tiles_per_xcd));
    if (tile >= n_tiles || int64_t(chunk) >= n_chunks) continue;
    const ChunkDesc d = desc[chunk];
    const int j = tile * kTile + int(threadIdx.x);
    const bool valid = j < tg.n;
    const float x = ld_at(tg.wave, uint32_t(valid ? j : tg.n - 1));
    const float tau = kDust ? dust_tau(axav, uint32_t(valid ? j : tg.n - 1)) : 0.f;
    // double accumulator: the cube is a sum of up to chunk-size positive
    // terms per lane; accumulating in double keeps the forward's rounding
    // at the level of the individual float32 terms.
    double acc = 0.0;
#pragma unroll 2
    for (int r = d.begin; r < d.end; ++r) {
      const StarState s = state[r];
      if (s.row < 0) continue;
      const Rows R = rows_of(g, s.row);
      const Knots K = knots_of<false>(g, x, shift_of(g, s.a, s.inv_a));
      float v = star_value(R, s, K) * s.sf;
      if (kDust && s.av != 0.f) v *= dust_factor(tau, s.av);
      acc += double(v);
    }
    if (valid) partial[int64_t(chunk) * tg.n + j] = acc;
  }
}

// Bins per lane of the fused pass: NB = 4 for targets of up to 4096 bins
// (8 slabs x 2 wavefronts x 64 lanes x 4), up to 8 for 8192.
constexpr int kMaxFusedBins = 8;

// Normalisation and accumulation in one pass for ascending targets: item q
// is spaxel chunk q, one workgroup of 16 wavefronts, two per lambda slab
// (wavefront 2b + h takes the bins j0 + 64 NB h + lane + 64 t, t < NB, of
// slab b; NB = 4 up to 4096 target bins).  The chunk's star states are
// staged in LDS, 64 at a time.  Per star each wavefront takes its knot
// window from the locator guesses of its first and last bin, stages the
// window's shifted knots and spectrum values in its own LDS (no other
// wavefront reads them) and evaluates P_j on its bins once (the per-bin
// arithmetic of star_norm_kernel and cube_accum_kernel), then posts four
// shares of Nn = sum_j P_j dtarget_j; after one barrier (one per two stars,
// both stars' shares posted first) every wavefront adds the 64 shares in a
// fixed order, forms sf = nan_to_num(T / Nn) and
// accumulates P_j sf into its f64 registers.  The shares are
// double-buffered, so a wavefront may post the next star's share while
// another still reads this one's.  Each (star, bin) value is computed once
// instead of twice.  The first wavefront stores sf and Nn for the adjoint.
constexpr int kFusedWaves = 2 * kSlabs;
static_assert(kFusedWaves == 16, "the Nn shares fill one 64-lane wave: 16 waves x 4 rows");

// v + the row-of-16 DPP move of v (every lane of the wave active).
template <int kCtrl>
__device__ inline float plus_dpp(float v) {
  return v + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), kCtrl, 0xF, 0xF, false));
}

// Every lane gets the sum of its 16-lane row, by the same fixed tree in
// every row: pairs (xor 1), quads (xor 2), then rotations of the row by 4
// and 8.  All 64 lanes must be active.
__device__ inline float row_sum16(float v) {
  v = plus_dpp<0xB1>(v);   // quad_perm(1, 0, 3, 2)
  v = plus_dpp<0x4E>(v);   // quad_perm(2, 3, 0, 1)
  v = plus_dpp<0x124>(v);  // row_ror:4
  return plus_dpp<0x128>(v);  // row_ror:8
}

// Sum over the 64 lanes in a fixed order, broadcast: the four row sums of
// row_sum16, added row 0 to row 3.  DPP moves and lane reads only (no LDS
// round trips).  All 64 lanes must be active.
__device__ inline float wave_sum_dpp(float v) {
  v = row_sum16(v);
  const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
  const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
  const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
  const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
  return ((r0 + r1) + r2) + r3;
}
// Knots of a wavefront's staged window: 80 per bin of a lane (the 64 NB
// bins of a wavefront span ~64 NB x 1.25 A, ~57 NB knots of a 1.4 A grid,
// plus the margins); wider windows take the global search.
template <int NB>
constexpr int fused_window() { return 80 * NB; }

// Star states staged in LDS per batch: the per-star fields come from LDS
// (broadcast reads) instead of a global round trip ahead of every star.
constexpr int kStateBatch = 64;

// A wave-uniform float (the lanes hold the same value) in a scalar register.
__device__ inline float uniform_f(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// Up to 4 bins per lane the kernel holds to 64 VGPRs (8 wavefronts per
// SIMD); more bins per lane take 128 (4 per SIMD).
template <bool kDust, int NB>
__global__ void __launch_bounds__(64 * kFusedWaves)
__attribute__((amdgpu_waves_per_eu(NB <= 4 ? 8 : 4)))
cube_fused_kernel(SspView g, rbx_target tg, StarState* __restrict__ state,
                  const ChunkDesc* __restrict__ desc, const int32_t* __restrict__ n_chunks_ptr,
                  int slab_bins, int64_t n_items, const float* __restrict__ axav,
                  double* __restrict__ partial) {
  constexpr int kHalfWindow = fused_window<NB>();
  __shared__ float W_s[kFusedWaves][kHalfWindow];  // shifted knots of each wavefront's window
  __shared__ float S_s[kFusedWaves][kHalfWindow];  // the star's spectrum at those knots
  __shared__ float4 st_s[4 * kStateBatch];
  __shared__ float nn_s[2][2][4 * kFusedWaves];  // [buffer][star of the pair][share]
  __shared__ float x_s[kFusedWaves][NB][64];   // the lanes' target wavelengths
  __shared__ float dt_s[kFusedWaves][NB][64];  // and bin widths
  static_assert(sizeof(StarState) == 4 * sizeof(float4), "a star state is four float4");
  const int n_chunks = *n_chunks_ptr;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int slab = wave >> 1, half = wave & 1;
  const int lane = threadIdx.x & 63;
  float* W = W_s[wave];
  float* S = S_s[wave];
  const uint32_t j0 = uint32_t(slab * slab_bins);
  const uint32_t j1 = j0 + slab_bins < uint32_t(tg.n) ? j0 + slab_bins : uint32_t(tg.n);
  // this wavefront's bins [jh, jh_end)
  const uint32_t jh = j0 + 64u * NB * half;
  const uint32_t jh_end = jh + 64u * NB < j1 ? jh + 64u * NB : j1;
  // the lane's bins, the same for every chunk: a bin past the end takes the
  // last bin's x (inside the window) and weight 0, so it adds nothing to Nn
  // and its sum is never stored.  x and the widths wait in LDS and are read
  // where used (fewer registers live through the staging loop).
  float tau[NB];
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    const uint32_t j = jh + lane + 64u * t;
    const uint32_t jc = j < j1 ? j : (j1 > j0 ? j1 - 1 : 0);
    const Pair xd = ld_pair(tg.pairs, 2 * jc);
    x_s[wave][t][lane] = xd.lo;
    dt_s[wave][t][lane] = j < jh_end ? xd.hi : 0.f;
    tau[t] = kDust ? dust_tau(axav, jc) : 0.f;
  }
  for (int64_t q = blockIdx.x; q < n_items; q += gridDim.x) {
    if (q >= n_chunks) break;  // chunk count is uniform: the whole workgroup leaves
    const ChunkDesc d = desc[q];
    double acc[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) acc[t] = 0.0;
    int par = 0;
    for (int b0 = d.begin; b0 < d.end; b0 += kStateBatch) {
      const int nb_st = d.end - b0 < kStateBatch ? d.end - b0 : kStateBatch;
      __syncthreads();  // the previous batch's states are read
      const float4* src = reinterpret_cast<const float4*>(state + b0);
      for (int i = threadIdx.x; i < 4 * nb_st; i += 64 * kFusedWaves) st_s[i] = src[i];
      __syncthreads();
      // P_j of star i of the batch on this wavefront's bins; false (P = 0)
      // for a dropped star (uniform: the same star for every wavefront)
      auto star_values = [&](int i, float (&P)[NB], float& T, float& av) -> bool {
        const float4 s0 = st_s[4 * i], s1 = st_s[4 * i + 1], s2 = st_s[4 * i + 2];
        const int row = __builtin_amdgcn_readfirstlane(__float_as_int(s0.x));
#pragma unroll
        for (int t = 0; t < NB; ++t) P[t] = 0.f;
        if (row < 0) return false;
        StarState s;
        s.row = row;
        s.a = uniform_f(s0.z);
        s.inv_a = uniform_f(s0.w);
        s.w[0] = uniform_f(s1.x);
        s.w[1] = uniform_f(s1.y);
        s.w[2] = uniform_f(s1.z);
        s.w[3] = uniform_f(s1.w);
        s.m = uniform_f(s2.x);
        T = uniform_f(s2.y);
        av = kDust ? uniform_f(st_s[4 * i + 3].w) : 0.f;
        if (jh < jh_end) {
          const Rows R = rows_of(g, s.row);
          const Shift h = shift_of(g, s.a, s.inv_a);
          int k0[NB];
#pragma unroll
          for (int t = 0; t < NB; ++t) k0[t] = loc_k0(g, x_s[wave][t][lane], h);
          int k_lo, nk;
          if (!g.loc_multi) {
            // the guesses are monotone in x: lane 0's first and lane 63's
            // last bin bound the wavefront's, widened by the error bound
            const int lo = __builtin_amdgcn_readfirstlane(k0[0]) - 3;
            const int hi = __builtin_amdgcn_readlane(k0[NB - 1], 63) + 5;
            k_lo = lo < 0 ? 0 : lo;
            nk = (hi > g.n_wave - 1 ? g.n_wave - 1 : hi) - k_lo + 1;
          } else {
            slab_window(g, tg, jh, jh_end, h, &k_lo, &nk);
          }
          if (nk <= kHalfWindow) {
            // shifted knot as the reference's wave * a; S_k = F_k m,
            // rounded once per knot as the reference's spectra
            if (nk > 128) {
              // wide windows (knots denser than the bins): four knots per
              // lane in 16-byte loads, one round of loads up to 256 knots.
              // The last group ends on the window's last knot (b = nk - 4),
              // so no load leaves the grid; its knots overlap the lane
              // before, which stores the same values.
              for (int k = 4 * lane; k < nk; k += 256) {
                const int b = k + 4 <= nk ? k : nk - 4;
                const uint32_t kk = uint32_t(k_lo + b);
                const Quad lw = ld_quad(g.wave, kk);
                Quad fc[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) fc[c] = ld_quad(R.f[c], kk);
                // the five loads in flight together, then the arithmetic
                __builtin_amdgcn_sched_group_barrier(0x020, 5, 0);
                __builtin_amdgcn_sched_group_barrier(0x002, 32, 0);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                  float f = 0.f;
#pragma unroll
                  for (int c = 0; c < 4; ++c) f += s.w[c] * fc[c].v[e];
                  W[b + e] = mul_nocontract(lw.v[e], s.a);
                  S[b + e] = f * s.m;
                }
              }
            } else {
              // narrow windows: one knot per lane (a quarter of the
              // arithmetic per round when the window fits 64 or 128 lanes)
              for (int k = lane; k < nk; k += 64) {
                const uint32_t kk = uint32_t(k_lo + k);
                const float lw = ld_at(g.wave, kk);
                float fc[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) fc[c] = ld_at(R.f[c], kk);
                __builtin_amdgcn_sched_group_barrier(0x020, 5, 0);
                __builtin_amdgcn_sched_group_barrier(0x002, 8, 0);
                float f = 0.f;
#pragma unroll
                for (int c = 0; c < 4; ++c) f += s.w[c] * fc[c];
                W[k] = mul_nocontract(lw, s.a);
                S[k] = f * s.m;
              }
            }
            wave_lds_sync();
            float x[NB];
#pragma unroll
            for (int t = 0; t < NB; ++t) x[t] = x_s[wave][t][lane];
            // every bin of the lane, also one past the end (its x lies in the window)
            if (k_lo >= 1 && k_lo + nk < g.n_wave && !g.loc_multi)
              win_values<NB, true>(x, k0, k_lo, nk, g.n_wave, W, S, P);
            else
              win_values<NB, false>(x, k0, k_lo, nk, g.n_wave, W, S, P);
            wave_lds_sync();  // this window's reads done before the next star's stores
          } else {
#pragma unroll
            for (int t = 0; t < NB; ++t) {
              if (jh + lane + 64u * t < jh_end) {
                const Knots K = knots_of<false>(g, x_s[wave][t][lane], h);
                P[t] = star_value(R, s, K);
              }
            }
          }
        }
        return true;
      };
      // this wavefront's four row sums of Nn = sum_j P_j dtarget_j
      auto post_share = [&](const float (&P)[NB], float* slot) {
        float nn = 0.f;
#pragma unroll
        for (int t = 0; t < NB; ++t) nn += P[t] * dt_s[wave][t][lane];
        nn = row_sum16(nn);
        if ((lane & 15) == 0) slot[4 * wave + (lane >> 4)] = nn;
      };
      // sf = nan_to_num(T / Nn) from the 64 shares (added in a fixed order),
      // stored for the adjoint, and P_j sf into the f64 sums
      auto accumulate = [&](const float (&P)[NB], float T, float av, const float* slot, int r) {
        const float Nn = wave_sum_dpp(slot[lane]);
        // T / Nn with the hardware reciprocal (within one ulp of the division;
        // Nn = 0 still gives inf or nan for nan_to_num below)
        float sf = T * __builtin_amdgcn_rcpf(Nn);
        if (isnan(sf)) sf = 0.f;
        else if (isinf(sf)) sf = sf > 0.f ? FLT_MAX : -FLT_MAX;
        if (wave == 0 && lane == 0) {
          state[r].sf = sf;
          state[r].Nn = Nn;
        }
#pragma unroll
        for (int t = 0; t < NB; ++t) {
          float v = P[t] * sf;
          if (kDust && av != 0.f) v *= dust_factor(tau[t], av);
          acc[t] += double(v);
        }
      };
      // two stars per barrier: both stars' shares are posted, then one
      // barrier; the stars are added in batch order as one at a time would
      for (int i = 0; i < nb_st; i += 2) {
        float PA[NB], PB[NB];
        float TA = 0.f, TB = 0.f, avA = 0.f, avB = 0.f;
        const bool va = star_values(i, PA, TA, avA);
        const bool vb = i + 1 < nb_st && star_values(i + 1, PB, TB, avB);
        if (!va && !vb) continue;
        if (!vb) {
#pragma unroll
          for (int t = 0; t < NB; ++t) PB[t] = 0.f;
        }
        post_share(PA, nn_s[par][0]);
        post_share(PB, nn_s[par][1]);
        __syncthreads();  // every share of both stars posted
        if (va) accumulate(PA, TA, avA, nn_s[par][0], b0 + i);
        if (vb) accumulate(PB, TB, avB, nn_s[par][1], b0 + i + 1);
        par ^= 1;
      }
    }
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      const uint32_t j = jh + lane + 64u * t;
      if (j < jh_end) partial[q * tg.n + j] = acc[t];
    }
  }
}

// chunk_off is indexed by key band * n_spaxel + p: the chunks of spaxel p
// are added band by band, each band's in chunk order.
__global__ void cube_combine_kernel(const double* __restrict__ partial,
                                    const int32_t* __restrict__ chunk_off, int32_t n_spaxel,
                                    int32_t bands, int32_t nt, float* __restrict__ cube) {
  const int64_t total = int64_t(n_spaxel) * nt;
  for (int64_t idx = gtid(); idx < total; idx += int64_t(gridDim.x) * blockDim.x) {
    const int32_t p = int32_t(idx / nt);
    const int32_t j = int32_t(idx - int64_t(p) * nt);
    double s = 0.0;
    for (int32_t b = 0; b < bands; ++b) {
      const int64_t k = int64_t(b) * n_spaxel + p;
      const int32_t c0 = chunk_off[k], c1 = chunk_off[k + 1];
      for (int32_t c = c0; c < c1; ++c) s += partial[int64_t(c) * nt + j];
    }
    cube[idx] = float(s);
  }
}

// ------------------------------------------------------------ adjoint

// Per-slab sums of one star: value sums of interpolated template rows and
// Doppler sums of their a-derivatives, weighted by dL/dC (g) and dtarget (d).
// Each lane accumulates its bins in float32; the wave's 64 lane sums are
// added in double in a fixed order (lds_sum), so the Doppler sums, which
// cancel strongly across bins, lose nothing beyond the float32 per-bin terms.
constexpr int kVSums = 5;                // value sums per (star, slab): Vg_w, Vg_u, Vg_s, Vd_u, Vd_s
constexpr int kASums = 2;                // Doppler sums per (star, slab): Ag_w, Ad_w
constexpr int kRedRow = 68;              // LDS floats per value row (64 lanes, 16-byte rows)
constexpr int kRedFloats = 8 * kRedRow;  // one wavefront's reduction buffer

// The double of lane f(l) for a row DPP control (both halves moved).
template <int kCtrl>
__device__ inline double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp(int(b & 0xffffffffll), kCtrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(int(b >> 32), kCtrl, 0xF, 0xF, false);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<long long>(static_cast<unsigned>(lo)));
}

// v[0..K) of all 64 lanes -> lane 8q (q < K) holds the wave total of v[q]
// (double).  Lane l stores v[q] at row q, column l; lane 8q + p adds the
// row's columns 8p .. 8p + 7 in order (two 16-byte LDS reads), then the
// eight parts are combined by DPP moves over 1, 2 and 4 lanes: a fixed order.
// buf is the wave's own LDS (kRedFloats floats, 16-byte aligned).
template <int K>
__device__ inline double lds_sum(const float v[K], float* buf, int lane) {
  static_assert(K <= 8, "at most 8 values");
  wave_lds_sync();
#pragma unroll
  for (int q = 0; q < K; ++q) buf[q * kRedRow + lane] = v[q];
  wave_lds_sync();
  const int q = lane >> 3, p = lane & 7;
  // lanes past the K rows read row 0 (their totals are not used)
  const float4* src = reinterpret_cast<const float4*>(buf + (q < K ? q : 0) * kRedRow + 8 * p);
  const float4 a = src[0], b = src[1];
  double acc = double(a.x);
  acc += double(a.y);
  acc += double(a.z);
  acc += double(a.w);
  acc += double(b.x);
  acc += double(b.y);
  acc += double(b.z);
  acc += double(b.w);
  acc += dpp_f64<0xB1>(acc);   // lane ^ 1: quad_perm(1, 0, 3, 2)
  acc += dpp_f64<0x4E>(acc);   // lane ^ 2: quad_perm(2, 3, 0, 1)
  acc += dpp_f64<0x12C>(acc);  // lane + 4 within the row: row_ror:12
  return acc;
}

// The three combined rows of one knot from its four corner values, in the
// difference form of Cell (A = f10 - f00, C = f01 - f00, E = f11 - f10 -
// f01 + f00): w = sum_c w_c f_c, and u, s the corner-derivative sums
// without their reciprocal cell sizes (star_grad_kernel applies those).
struct RowVals {
  float w, u, s;
};

__device__ inline RowVals combine_rows(float f0, float f1, float f2, float f3, float fu, float fs) {
  const float A = f1 - f0, C = f2 - f0, E = (f3 - f1) - C;
  RowVals r;
  r.s = fmaf(fu, E, C);
  r.u = fmaf(fs, E, A);
  r.w = fmaf(fs, r.s, fmaf(fu, A, f0));
  return r;
}

// One bin of star_bwd_kernel's sums, from gathered template pairs.
struct BwdRows {
  float fu, fs;  // fractional cell position (Cell::fu, Cell::fs)
};

__device__ inline void bwd_bin_gather(const SspView& g, const Rows& R, const StarState& s,
                                      const BwdRows& B, const Shift& h, float x, float gj,
                                      float dj, float v[7]) {
  const Knots K = knots_of<true>(g, x, h);
  const float gk = -gj * K.kappa, dk = -dj * K.kappa;
  Pair f[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) f[c] = ld_pair(R.f[c], K.i0);
  // rows at the left knot, and their change to the right knot from the
  // corners' own knot-to-knot differences (exact in float32 for neighbouring
  // values within a factor 2 of each other), so that the Doppler sums, which
  // cancel strongly across bins, are not left with the rounding of the rows
  // themselves (it dominates where the knots are denser than the bins)
  const RowVals a = combine_rows(f[0].lo, f[1].lo, f[2].lo, f[3].lo, B.fu, B.fs);
  const RowVals d = combine_rows(f[0].hi - f[0].lo, f[1].hi - f[1].lo, f[2].hi - f[2].lo,
                                 f[3].hi - f[3].lo, B.fu, B.fs);
  const float Rw = a.w + K.theta * d.w;
  const float Ru = a.u + K.theta * d.u;
  const float Rs = a.s + K.theta * d.s;
  v[0] += gj * Rw;
  v[1] += gj * Ru;
  v[2] += gj * Rs;
  v[3] += dj * Ru;
  v[4] += dj * Rs;
  v[5] += gk * d.w;
  v[6] += dk * d.w;
}

// Slab-mapped like star_norm_kernel.  The per-star gradients need the four
// corner rows only through three weightings: the bilinear weights w_c and
// their derivatives gu_c = dw_c/d(log Z), gs_c = dw_c/d(log t), formed as
// combine_rows does from corner differences.  Per bin the kernel gathers the
// four corner pairs (8-byte loads from scalar bases), forms the combined rows
// R_w, R_u, R_s at the bin and accumulates the seven sums star_grad_kernel
// uses:
//   Vg_w, Vg_u, Vg_s = sum_j g_j R_{w,u,s}(j),  Vd_u, Vd_s = sum_j d_j R_{u,s}(j),
//   Ag_w, Ad_w = sum_j {g_j, d_j} dR_w(j)/da      (d = dtarget),
// float32 per lane, added across lanes in double in a fixed order; vpart
// gets the five value sums, apart the two Doppler sums.  Combining the
// corners per knot takes differences of neighbouring templates while they
// are small and nearly exact, before any sum over bins.  (Staging the rows
// in LDS as the forward kernels do cost more occupancy than it saved here.)
template <int W, bool kDust>
__global__ void __launch_bounds__(64 * W) star_bwd_kernel(
    SspView g, rbx_target tg, int64_t n, int64_t n_items, const StarState* __restrict__ state,
    const float* __restrict__ gradw, const float* __restrict__ dcube, int slab_bins,
    const float* __restrict__ axav, float* __restrict__ vpart, double* __restrict__ apart) {
  __shared__ __align__(16) float red[W][kRedFloats];
  for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const SlabTask t = slab_task<W, kBwdSlabs>(item);
    const bool live = t.r < n;
    float v[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) v[q] = 0.f;
    StarState s;
    s.row = -1;
    if (live) s = state[t.r];
    if (s.row >= 0) {
      const Rows R = rows_of(g, __builtin_amdgcn_readfirstlane(s.row));
      const float* grow = dcube + int64_t(s.spaxel) * tg.n;
      const Shift h = shift_of(g, s.a, s.inv_a);
      BwdRows B;
      B.fu = gradw[t.r * 2];
      B.fs = gradw[t.r * 2 + 1];
      const uint32_t j0 = uint32_t(t.slab * slab_bins);
      const uint32_t j1 = j0 + slab_bins < uint32_t(tg.n) ? j0 + slab_bins : uint32_t(tg.n);
#pragma unroll 2
      for (uint32_t j = j0 + t.lane; j < j1; j += 64) {
        const Pair xd = ld_pair(tg.pairs, 2 * j);
        float gj = ld_at(grow, j);
        // with dust the cube holds sf P_j e_j: every g-sum takes g_j e_j
        if (kDust && s.av != 0.f) gj *= dust_factor(dust_tau(axav, j), s.av);
        bwd_bin_gather(g, R, s, B, h, xd.lo, gj, xd.hi, v);
      }
    }
    const double tot = lds_sum<7>(v, red[t.wave], t.lane);
    const int q = t.lane >> 3;
    if (live && (t.lane & 7) == 0 && q < 7) {
      const int64_t it = t.r * kBwdSlabs + t.slab;
      if (q < kVSums) vpart[it * kVSums + q] = float(tot);
      else apart[it * kASums + q - kVSums] = tot;
    }
  }
}

// v of lane l (0 <= l < 64) by one ds_bpermute (no lane wrapping needed).
__device__ inline float lane_of(float v, int l) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(l << 2, __float_as_int(v)));
}

// The next lane's v (DPP wave shift left by one); lane 63 reads 0.
__device__ inline float next_lane(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xF, 0xF, true));
}

// Change of the combined row R_w from this lane's knot to the next lane's,
// sum_c w_c (f_c[k + 1] - f_c[k]) from the corners' own knot-to-knot
// differences: exact in float32 for neighbouring values within a factor 2,
// unlike the difference of the two rounded rows, which loses the digits
// the Doppler sums need where the rows change little from knot to knot.
// The exact differences are weighted with the star's bilinear corner weights
// wc (four products instead of the difference form's seven operations).
// Lanes whose next lane does not hold the next knot (the window's last, and
// lanes past it) give a value that is never used.
__device__ inline float w_change(const float (&f)[4], const float (&wc)[4]) {
  float r = 0.f;
#pragma unroll
  for (int c = 0; c < 4; ++c) r += wc[c] * (next_lane(f[c]) - f[c]);
  return r;
}

// star_bwd_kernel for ascending targets with a register window: for each
// group of 64 consecutive bins (one per lane) the knots the group can touch,
// [k0(first bin) - 1, k0(last bin) + 2] from the compact locator, are
// loaded one per lane with coalesced loads (the knot and its four corner
// values, combined into R_w, R_u, R_s in the lane), and every bin takes its
// segment's two knots from the owning lanes by cross-lane permutes instead
// of gathering the template pairs.  The segment is the exact one of
// win_knots, searched on the permuted knots with a wave-uniform correction
// loop.  A group whose knots do not fit one lane each, or where a bin's
// segment would leave the window, is done with the gathers instead.
//
// Full groups of a slab strictly inside the grid run first on a short path:
// the locator guesses are monotone in the bin, so every lane's guess lies in
// the window [1, span - 3] without clamps, a lane that does not end on
// w_l <= x < w_r after the correction steps is a segment outside the window,
// and no grid-end case can apply.  A group that does not fit that path
// (window over 64 knots or reaching a grid end, a segment outside it) is
// marked and redone after the loop with the general path; nothing of it was
// added, so each group's terms are added exactly once, in group order within
// each of the two loops.
template <int W, bool kDust>
__global__ void __launch_bounds__(64 * W) star_bwd_reg_kernel(
    SspView g, rbx_target tg, int64_t n, int64_t n_items, const StarState* __restrict__ state,
    const float* __restrict__ gradw, const float* __restrict__ dcube, int slab_bins,
    const float* __restrict__ axav, float* __restrict__ vpart, double* __restrict__ apart) {
  __shared__ __align__(16) float red[W][kRedFloats];
  const int n_w = g.n_wave;
  for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const SlabTask t = slab_task<W, kBwdSlabs>(item);
    const bool live = t.r < n;
    float v[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) v[q] = 0.f;
    StarState s;
    s.row = -1;
    if (live) s = state[t.r];
    if (s.row >= 0) {
      const Rows R = rows_of(g, __builtin_amdgcn_readfirstlane(s.row));
      const float* grow = dcube + int64_t(s.spaxel) * tg.n;
      const Shift h = shift_of(g, s.a, s.inv_a);
      BwdRows B;
      B.fu = gradw[t.r * 2];
      B.fs = gradw[t.r * 2 + 1];
      float wc[4];  // the corner weights, wave-uniform
#pragma unroll
      for (int c = 0; c < 4; ++c) wc[c] = uniform_f(s.w[c]);
      const uint32_t j0 = uint32_t(t.slab * slab_bins);
      const uint32_t j1 = j0 + slab_bins < uint32_t(tg.n) ? j0 + slab_bins : uint32_t(tg.n);
      // interior slab: every bin's segment has knots in [1, n - 2] (x ascending
      // from x_first >= w_1 to x_last < w_{n-2}), so no grid-end bound or
      // clamp can apply to it
      bool interior = false;
      if (j0 < j1) {
        const float w1 = mul_nocontract(ld_at(g.wave, 1u), h.a);
        const float wn = mul_nocontract(ld_at(g.wave, uint32_t(n_w - 2)), h.a);
        interior = n_w >= 4 && ld_at(tg.wave, j0) >= w1 && ld_at(tg.wave, j1 - 1) < wn;
      }
      // the short path of a full interior group, given its bins' (x, dtarget)
      // pair, dL/dcube value and locator guess; false (nothing added) when
      // the group needs the general path.  mid() runs once the group's knots
      // are in (the next group's locator guess goes there).
      auto fast_group = [&](uint32_t jb, Pair xd, float gj, int k0, auto&& mid) -> bool {
        const uint32_t j = jb + uint32_t(t.lane);
        const float x = xd.lo, dj = xd.hi;
        if (kDust && s.av != 0.f) gj *= dust_factor(dust_tau(axav, j), s.av);
        const int wb = __builtin_amdgcn_readfirstlane(k0) - 1;
        const int we = __builtin_amdgcn_readlane(k0, 63) + 2;
        const int span = we - wb + 1;
        if (wb < 0 || we > n_w - 1 || span > 64) return false;
        const int kl = wb + t.lane;
        const uint32_t kk = uint32_t(kl < we ? kl : we);
        const float lamk = ld_at(g.wave, kk);
        float fc[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) fc[c] = ld_at(R.f[c], kk);
        const RowVals rv = combine_rows(fc[0], fc[1], fc[2], fc[3], B.fu, B.fs);
        // the change of R_w to the next knot (the next lane's) from the
        // corners' own knot-to-knot differences (exact in float32 for
        // neighbouring values), not from the rounded rows: the Doppler sums
        // cancel strongly across bins
        const float dwk = w_change(fc, wc);
        mid();
        const float y = x * h.inv_a;
        int li = k0 - wb;  // in [1, span - 3]: the guesses are monotone in the bin
        // the next-knot test on knot k0 + 1; the segment's other knot (k0 + 2
        // after a step, k0 otherwise) in one more permute
        const float ln = lane_of(lamk, li + 1);
        const bool st = ln <= y;
        const float lo2 = lane_of(lamk, st ? li + 2 : li);
        li += st ? 1 : 0;
        float l0 = st ? ln : lo2, l1 = st ? lo2 : ln;
        // the segment's row values, read before the (rare) correction vote so
        // that their permutes overlap the knots' (read again after a correction)
        float w0 = lane_of(rv.w, li), dw = lane_of(dwk, li);
        float u0 = lane_of(rv.u, li), u1 = lane_of(rv.u, li + 1);
        float s0 = lane_of(rv.s, li), s1 = lane_of(rv.s, li + 1);
        float wl = mul_nocontract(l0, h.a), wr = mul_nocontract(l1, h.a);
        if (__ballot(x >= wr || x < wl)) {  // rare: correction steps inside the window
          bool step = (x >= wr && li < span - 2) || (x < wl && li > 0);
          while (__ballot(step)) {
            const int nl = step ? li + (x >= wr ? 1 : -1) : li;
            const float a0 = lane_of(lamk, nl), a1 = lane_of(lamk, nl + 1);
            if (step) {
              li = nl;
              l0 = a0;
              l1 = a1;
              wl = mul_nocontract(l0, h.a);
              wr = mul_nocontract(l1, h.a);
            }
            step = (x >= wr && li < span - 2) || (x < wl && li > 0);
          }
          // still outside: the segment lies past the window's edge
          if (__ballot(x >= wr || x < wl)) return false;
          w0 = lane_of(rv.w, li);
          dw = lane_of(dwk, li);
          u0 = lane_of(rv.u, li);
          u1 = lane_of(rv.u, li + 1);
          s0 = lane_of(rv.s, li);
          s1 = lane_of(rv.s, li + 1);
        }
        const float r = __builtin_amdgcn_rcpf(wr - wl);
        const float th = (x - wl) * r;
        const float kap = y * r;  // lambda(x) / (w_r - w_l), lambda(x) = x / a
        const float Ruv = u0 + th * (u1 - u0);
        const float Rsv = s0 + th * (s1 - s0);
        v[0] += gj * (w0 + th * dw);
        v[1] += gj * Ruv;
        v[2] += gj * Rsv;
        v[3] += dj * Ruv;
        v[4] += dj * Rsv;
        const float kd = kap * dw;  // the Doppler change of R_w across the segment
        v[5] -= gj * kd;
        v[6] -= dj * kd;
        return true;
      };
      // the general path of one group (any slab, full or the last partial one)
      auto group = [&](uint32_t jb, auto interior_c, auto full_c) {
        constexpr bool kIn = decltype(interior_c)::value;
        constexpr bool kFull = decltype(full_c)::value;
        const uint32_t j = jb + uint32_t(t.lane);
        const bool valid = kFull || j < j1;
        const uint32_t jc = valid ? j : j1 - 1;
        const Pair xd = ld_pair(tg.pairs, 2 * jc);
        const float x = xd.lo;
        float gj = valid ? ld_at(grow, jc) : 0.f;
        if (kDust && s.av != 0.f) gj *= dust_factor(dust_tau(axav, jc), s.av);
        const float dj = valid ? xd.hi : 0.f;
        // locator guess (compact copy: k0 = knots <= cell start - 1)
        const float fm = __builtin_amdgcn_fmed3f(fmaf(x, h.c1, h.c0), 0.f, float(g.n_loc - 1));
        const int k0 = ld_at(g.loc + 4 * g.n_loc, uint32_t(fm));
        const int last = int(j1 - jb < 64u ? j1 - jb - 1 : 63u);
        int wb = __builtin_amdgcn_readfirstlane(k0) - 1;
        int we = __builtin_amdgcn_readlane(k0, last) + 2;
        wb = wb < 0 ? 0 : wb;
        we = we > n_w - 1 ? n_w - 1 : we;
        const int span = we - wb + 1;
        bool gather = span > 64 || span < 2;
        if (!gather) {
          // this lane's knot of the window and its combined rows
          const int kl = wb + t.lane;
          const uint32_t kk = uint32_t(kl < we ? kl : we);
          const float lamk = ld_at(g.wave, kk);
          // the knot and the next one (the last knot of the grid pairs with
          // the one before it; its change to a next knot is never used)
          float fc[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) fc[c] = ld_at(R.f[c], kk);
          const RowVals rv = combine_rows(fc[0], fc[1], fc[2], fc[3], B.fu, B.fs);
          const float dwk = w_change(fc, wc);  // exact change to the next knot, as above
          // exact segment on the window's knots (as win_knots)
          const float y = x * h.inv_a;
          int li = k0 - wb;
          li = li < 0 ? 0 : (li > span - 2 ? span - 2 : li);
          const float ln = lane_of(lamk, li + 1);
          li += ln <= y ? 1 : 0;
          li = li > span - 2 ? span - 2 : li;
          float l0 = lane_of(lamk, li), l1 = lane_of(lamk, li + 1);
          float wl = mul_nocontract(l0, h.a), wr = mul_nocontract(l1, h.a);
          auto needs_step = [&](int l, float lo, float hi) {
            if (kIn) return (x >= hi && l < span - 2) || (x < lo && l > 0);
            return (x >= hi && l < span - 2 && l + wb + 1 < n_w - 1) ||
                   (x < lo && l > 0 && l + wb > 0);
          };
          bool step = needs_step(li, wl, wr);
          if (__ballot(step)) {  // rare: the common path carries no loop state
            do {
              const int nl = step ? li + (x >= wr ? 1 : -1) : li;
              const float a0 = lane_of(lamk, nl), a1 = lane_of(lamk, nl + 1);
              if (step) {
                li = nl;
                l0 = a0;
                l1 = a1;
                wl = mul_nocontract(l0, h.a);
                wr = mul_nocontract(l1, h.a);
              }
              step = needs_step(li, wl, wr);
            } while (__ballot(step));
          }
          // a segment pinned at a window edge that is not a grid end: the
          // window was too narrow for this lane
          const bool off =
              valid && (kIn ? ((x >= wr && li == span - 2) || (x < wl && li == 0))
                            : ((x >= wr && li == span - 2 && li + wb + 1 < n_w - 1) ||
                               (x < wl && li == 0 && li + wb > 0)));
          gather = __ballot(off) != 0;
          if (!gather) {
            const int i = li + wb + 1;
            const float dx = wr - wl;
            const float r = __builtin_amdgcn_rcpf(dx);
            // interior: a lane that is not off ends on w_l <= x < w_r (the
            // lanes past the slab repeat its last bin), so dx > 0
            const bool flat = !kIn && !(dx > 0.f);
            const float th0 = flat ? 0.f : (x - wl) * r;
            const bool left = !kIn && (i == 1) && (x < wl);
            const bool right = !kIn && (i == n_w - 1) && (x > wr);
            const float th = left ? 0.f : (right ? 1.f : th0);
            const float kap = (left || right || flat) ? 0.f : (l0 + th0 * (l1 - l0)) * r;
            const float w0 = lane_of(rv.w, li), dw = lane_of(dwk, li);
            const float u0 = lane_of(rv.u, li), u1 = lane_of(rv.u, li + 1);
            const float s0 = lane_of(rv.s, li), s1 = lane_of(rv.s, li + 1);
            const float kd = kap * dw;  // the Doppler change of R_w across the segment
            const float Rwv = w0 + th * dw;
            const float Ruv = u0 + th * (u1 - u0);
            const float Rsv = s0 + th * (s1 - s0);
            v[0] += gj * Rwv;
            v[1] += gj * Ruv;
            v[2] += gj * Rsv;
            v[3] += dj * Ruv;
            v[4] += dj * Rsv;
            v[5] -= gj * kd;
            v[6] -= dj * kd;
          }
        }
        if (gather && valid) bwd_bin_gather(g, R, s, B, h, x, gj, dj, v);
      };
      auto general = [&](uint32_t jb) {
        const bool full = jb + 64 <= j1;
        if (interior) {
          if (full) group(jb, std::true_type{}, std::true_type{});
          else group(jb, std::true_type{}, std::false_type{});
        } else {
          if (full) group(jb, std::false_type{}, std::true_type{});
          else group(jb, std::false_type{}, std::false_type{});
        }
      };
      // the short path over the first (up to 64) full groups of an interior
      // slab, marking those it leaves; then the general path for the marked
      // groups in order, and for every group after the short path's
      uint64_t redo = 0;
      uint32_t jb = j0;
      if (interior) {
        const uint32_t full = (j1 - j0) / 64u;
        const uint32_t ng = full < 64u ? full : 64u;
        // the loads that do not depend on the group before run ahead: the
        // next group's (x, dtarget) pairs and dL/dcube values are issued with
        // this group's knot loads, its locator guesses (which need x) once
        // these knots are in, so a group waits on one round of loads instead
        // of three in a row.  The next group's index is clamped to the last
        // full group (in bounds; the extra loads are not used).  Two groups
        // per pass, the inputs alternating between the A and B registers
        // (no copies between groups).
        const uint32_t jl = j0 + uint32_t(t.lane);
        auto guess = [&](float x) {
          const float fm = __builtin_amdgcn_fmed3f(fmaf(x, h.c1, h.c0), 0.f, float(g.n_loc - 1));
          return int(ld_at(g.loc + 4 * g.n_loc, uint32_t(fm)));
        };
        auto run = [&](uint32_t gi, const Pair& xd, float gj, int k0, Pair& xdn, float& gjn,
                       int& k0n) {
          const uint32_t bn = jl + 64u * (gi + 1 < ng ? gi + 1 : ng - 1);
          xdn = ld_pair(tg.pairs, 2 * bn);
          gjn = ld_at(grow, bn);
          bool guessed = false;
          auto mid = [&] {
            k0n = guess(xdn.lo);
            guessed = true;
          };
          if (!fast_group(j0 + 64u * gi, xd, gj, k0, mid)) redo |= uint64_t(1) << gi;  // rare, uniform
          if (!guessed) k0n = guess(xdn.lo);
        };
        if (ng > 0) {
          Pair xdA = ld_pair(tg.pairs, 2 * jl), xdB;
          float gjA = ld_at(grow, jl), gjB;
          int k0A = guess(xdA.lo), k0B;
          uint32_t gi = 0;
          for (; gi + 1 < ng; gi += 2) {
            run(gi, xdA, gjA, k0A, xdB, gjB, k0B);
            run(gi + 1, xdB, gjB, k0B, xdA, gjA, k0A);
          }
          if (gi < ng) run(gi, xdA, gjA, k0A, xdB, gjB, k0B);
        }
        jb = j0 + 64u * ng;
      }
      while (redo) {
        const int gi = __builtin_ctzll(redo);
        redo &= redo - 1;
        general(j0 + 64u * uint32_t(gi));
      }
      for (; jb < j1; jb += 64) general(jb);
    }
    const double tot = lds_sum<7>(v, red[t.wave], t.lane);
    const int q = t.lane >> 3;
    if (live && (t.lane & 7) == 0 && q < 7) {
      const int64_t it = t.r * kBwdSlabs + t.slab;
      if (q < kVSums) vpart[it * kVSums + q] = float(tot);
      else apart[it * kASums + q - kVSums] = tot;
    }
  }
}

// Per star: add the slab sums in slab order and form every gradient in
// closed form.  With P_j = m sum_c w_c R_c(j) and sf = T / Nn:
//   sum_j g_j dP_j/dtheta = m Pg_theta, dNn/dtheta = m Pd_theta,
//   dT/du = m a W_u, dT/ds = m a W_s, dT/da = T / a, dT/dm = T / m,
//   dL/dtheta = sf sum_j g_j dP_j/dtheta + sbar dsf/dtheta, sbar = sum_j g_j P_j
// (dsf/dm = 0; dsf = 0 where T / Nn is not finite, as nan_to_num's gradient).
__global__ void __launch_bounds__(256) star_grad_kernel(SspView g, rbx_stars st, const StarState* __restrict__ state,
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
    // slab sums in slab order: Vg_w, Vg_u, Vg_s, Vd_u, Vd_s | Ag_w, Ad_w
    double Pg_m = 0, Pg_u = 0, Pg_s = 0, Pd_u = 0, Pd_s = 0, Pg_a = 0, Pd_a = 0;
    for (int q = 0; q < kBwdSlabs; ++q) {
      const float* vp = vpart + (r * kBwdSlabs + q) * kVSums;
      const double* ap = apart + (r * kBwdSlabs + q) * kASums;
      Pg_m += vp[0];
      Pg_u += vp[1];
      Pg_s += vp[2];
      Pd_u += vp[3];
      Pd_s += vp[4];
      Pg_a += ap[0];
      Pd_a += ap[1];
    }
    const Cell cell = make_cell(g, st.metallicity[i], st.age[i]);
    // the kernels' u and s sums are of (A + fs E) and (C + fu E): times the
    // reciprocal cell sizes they are the sums of the corner derivatives
    Pg_u *= double(cell.inv_dx);
    Pd_u *= double(cell.inv_dx);
    Pg_s *= double(cell.inv_dy);
    Pd_s *= double(cell.inv_dy);
    double win[4];
    window_sums(g, s.row, s.kl, s.kh, win);
    const double m = s.m, a = s.a, sf = s.sf, T = s.T, Nn = s.Nn;
    double W_u = 0, W_s = 0;
    for (int c = 0; c < 4; ++c) {
      W_u += double(cell.gu[c]) * win[c];
      W_s += double(cell.gs[c]) * win[c];
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

// Locator entries (layout in the rbx_ssp comment of the C header): entry 0
// covers queries below x0, entry m >= 1 the cell starting at x0 + (m - 1) h;
// k = last knot <= the cell start (double comparison), then the three knots
// from b = min(max(k, -1), n - 2), -inf / +inf outside the grid.
__global__ void ssp_locate_kernel(const float* __restrict__ wave, int nw, float x0, float inv_h,
                                  int nloc, int32_t* __restrict__ loc) {
  for (int64_t m = gtid(); m < nloc; m += int64_t(gridDim.x) * blockDim.x) {
    int k = -1;
    if (m > 0) {
      const double xm = double(x0) + double(m - 1) / double(inv_h);
      int lo = 0, hi = nw;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (double(wave[mid]) <= xm) lo = mid + 1; else hi = mid;
      }
      k = lo - 1;
    }
    const int b = k < nw - 2 ? k : nw - 2;
    loc[4 * m] = k;
    loc[4 * int64_t(nloc) + m] = k;  // compact copy
    for (int q = 0; q < 3; ++q) {
      const int kk = b + q;
      const float v = kk < 0 ? -INFINITY : (kk >= nw ? INFINITY : wave[kk]);
      loc[4 * m + 1 + q] = __float_as_int(v);
    }
  }
}

// ------------------------------------------------------------ host side

// Workspace layout.  The per-star arrays the adjoint reads (perm, skey,
// state and the slab partials) come first and do not depend on the chunk
// size, so the adjoint may run with any tuning.
struct CubeLayout {
  int32_t *perm, *skey, *offsets, *chunk_off, *scan_blocks, *keys;
  StarState* state;
  ChunkDesc* desc;
  double* partial;
  float* gradw;    // (n, 2) fractional cell position (fu, fs) per sorted star
  float* nn_part;  // (n, kSlabs) slab partials of Nn
  float* vpart;    // (n, kBwdSlabs, kVSums) adjoint value sums
  double* apart;   // (n, kBwdSlabs, kASums) adjoint Doppler sums
  void* sort_ws;
  size_t sort_ws_bytes;
  int64_t max_chunks;
};

// per_star_only: carve just the per-star arrays (what the adjoint reads).
template <class C>
static void cube_layout(C& c, int64_t n, int32_t n_keys, int32_t nt, int chunk, CubeLayout* L,
                        bool per_star_only = false) {
  // sort keys band * n_spaxel + spaxel: up to band_cap(n_keys) * n_keys
  const int64_t nk = int64_t(band_cap(n_keys)) * n_keys;
  const int64_t max_chunks = (n + chunk - 1) / chunk + nk;
  const size_t sort_bytes = sort_workspace_bound(n, int32_t(nk));
  const size_t n1 = size_t(n > 0 ? n : 1);
  auto* perm = c.template take<int32_t>(n1);
  auto* skey = c.template take<int32_t>(n1);
  auto* state = c.template take<StarState>(n1);
  auto* gw = c.template take<float>(n1 * 2);
  auto* nnp = c.template take<float>(n1 * kSlabs);
  auto* vp = c.template take<float>(n1 * kBwdSlabs * kVSums);
  auto* ap = c.template take<double>(n1 * kBwdSlabs * kASums);
  if (per_star_only) {
    if (L) {
      L->perm = reinterpret_cast<int32_t*>(perm);
      L->skey = reinterpret_cast<int32_t*>(skey);
      L->state = reinterpret_cast<StarState*>(state);
      L->gradw = reinterpret_cast<float*>(gw);
      L->nn_part = reinterpret_cast<float*>(nnp);
      L->vpart = reinterpret_cast<float*>(vp);
      L->apart = reinterpret_cast<double*>(ap);
    }
    return;
  }
  auto* offs = c.template take<int32_t>(size_t(nk) + 1);
  auto* coff = c.template take<int32_t>(size_t(nk) + 1);
  auto* blocks = c.template take<int32_t>((size_t(nk) + 1) / 4096 + 2);
  auto* keys = c.template take<int32_t>(n1);
  auto* desc = c.template take<ChunkDesc>(size_t(max_chunks > 0 ? max_chunks : 1));
  auto* part = c.template take<double>(size_t(max_chunks > 0 ? max_chunks : 1) * nt);
  auto* sws = c.template take<char>(sort_bytes);
  if (L) {
    L->perm = reinterpret_cast<int32_t*>(perm);
    L->skey = reinterpret_cast<int32_t*>(skey);
    L->state = reinterpret_cast<StarState*>(state);
    L->gradw = reinterpret_cast<float*>(gw);
    L->nn_part = reinterpret_cast<float*>(nnp);
    L->vpart = reinterpret_cast<float*>(vp);
    L->apart = reinterpret_cast<double*>(ap);
    L->offsets = reinterpret_cast<int32_t*>(offs);
    L->chunk_off = reinterpret_cast<int32_t*>(coff);
    L->scan_blocks = reinterpret_cast<int32_t*>(blocks);
    L->keys = reinterpret_cast<int32_t*>(keys);
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

int tune_of(const rbx_tuning* t, Tune* out) {
  Tune d;
  d.chunk = 256;
  d.waves = 4;
  d.persist = 0;
  d.flags = 0;
  if (t) {
    if (t->chunk != 0) d.chunk = t->chunk;
    if (t->waves != 0) d.waves = t->waves;
    d.persist = t->persist;
    d.flags = t->flags;
  }
  if (d.chunk < 16 || d.chunk > 65536)
    return fail(RBX_EINVAL, "rbx_tuning: chunk must be in [16, 65536] (0 = default)");
  if (d.waves != 1 && d.waves != 2 && d.waves != 4 && d.waves != 8)
    return fail(RBX_EINVAL, "rbx_tuning: waves must be 1, 2, 4 or 8 (0 = default)");
  if (d.persist < 0 || d.persist > 64)
    return fail(RBX_EINVAL, "rbx_tuning: persist must be in [0, 64]");
  if (d.flags & ~RBX_TUNE_TWO_PASS)
    return fail(RBX_EINVAL, "rbx_tuning: unknown flags");
  *out = d;
  return RBX_OK;
}

// Compute units of the current device (cached per device).
static int n_cus() {
  static int cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  if (cache[dev] <= 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        v < 1)
      v = 256;
    cache[dev] = v;
  }
  return cache[dev];
}

// Workgroups of `threads` for `items` slab-mapped work items (a multiple of
// 8): one per item, or a persistent grid of `persist` workgroups per CU;
// always a multiple of 8, and capped so that the launch has fewer than 2^32
// threads (one dispatch packet, also when captured into a graph); the
// kernels stride over the remaining items.
static int64_t item_grid(int64_t items, int persist, int threads) {
  int64_t g = items;
  if (persist > 0) {
    const int64_t p = int64_t(persist) * n_cus();
    if (p < g) g = p;
  }
  const int64_t cap = ((int64_t(1) << 32) - 1) / threads / 8 * 8;
  if (g > cap) g = cap;
  g = (g + 7) / 8 * 8;
  if (g < 8) g = 8;
  return g;
}

template <int W>
static void launch_norm(int64_t grid, hipStream_t s, const SspView& g, const rbx_target& tg,
                        int64_t n, int64_t items, const StarState* st, int slab_bins,
                        float* nn) {
  hipLaunchKernelGGL(star_norm_kernel<W>, dim3(grid), dim3(64 * W), 0, s, g, tg, n, items, st,
                     slab_bins, nn);
}

template <int W>
static void launch_bwd(int64_t grid, hipStream_t s, const SspView& g, const rbx_target& tg,
                       int64_t n, int64_t items, const StarState* st, const float* gradw,
                       const float* dcube, int slab_bins, const float* axav, float* vp,
                       double* ap) {
  if (tg.ascending) {
    if (axav)
      hipLaunchKernelGGL((star_bwd_reg_kernel<W, true>), dim3(grid), dim3(64 * W), 0, s, g, tg, n,
                         items, st, gradw, dcube, slab_bins, axav, vp, ap);
    else
      hipLaunchKernelGGL((star_bwd_reg_kernel<W, false>), dim3(grid), dim3(64 * W), 0, s, g, tg,
                         n, items, st, gradw, dcube, slab_bins, nullptr, vp, ap);
    return;
  }
  if (axav)
    hipLaunchKernelGGL((star_bwd_kernel<W, true>), dim3(grid), dim3(64 * W), 0, s, g, tg, n, items,
                       st, gradw, dcube, slab_bins, axav, vp, ap);
  else
    hipLaunchKernelGGL((star_bwd_kernel<W, false>), dim3(grid), dim3(64 * W), 0, s, g, tg, n,
                       items, st, gradw, dcube, slab_bins, nullptr, vp, ap);
}

static int validate(const rbx_ssp* ssp, const rbx_target* tg, const rbx_stars* st) {
  if (!ssp || !tg || !st) return fail(RBX_EINVAL, "rbx_cube: null descriptor");
  if (ssp->n_met < 2 || ssp->n_age < 2 || ssp->n_wave < 2)
    return fail(RBX_EINVAL, "rbx_cube: the SSP grid needs >= 2 metallicities, ages and wavelengths");
  if (!ssp->flux || !ssp->ax_met || !ssp->ax_age || !ssp->wave_obs || !ssp->cum || !ssp->locate ||
      ssp->n_locate < 1 || !(ssp->locate_inv_h > 0.f))
    return fail(RBX_EINVAL, "rbx_cube: SSP descriptor incomplete (run rbx_ssp_prepare)");
  if (tg->n < 1 || !tg->wave || !tg->dwave || !tg->pairs)
    return fail(RBX_EINVAL, "rbx_cube: empty target grid (wave, dwave and pairs are required)");
  if (st->n < 0 || st->n > INT32_MAX - 1) return fail(RBX_EINVAL, "rbx_cube: bad star count");
  if (st->n > 0 && (!st->mass || !st->age || !st->metallicity || !st->velocity || !st->ids))
    return fail(RBX_EINVAL, "rbx_cube: null star array");
  if (st->vel_comp < 0 || st->vel_comp >= st->vel_stride)
    return fail(RBX_EINVAL, "rbx_cube: vel_comp out of range");
  if (tg->n >= (1 << 28) || ssp->n_wave >= (1 << 28) || ssp->n_locate >= (1 << 27))
    return fail(RBX_EINVAL, "rbx_cube: wavelength axes or locator table too long");
  if (st->n_spaxel < 1 || st->n_keys < st->n_spaxel)
    return fail(RBX_EINVAL, "rbx_cube: need 1 <= n_spaxel <= n_keys");
  if (!(st->speed_of_light > 0.f)) return fail(RBX_EINVAL, "rbx_cube: speed_of_light must be > 0");
  return RBX_OK;
}

int build_chunks(const int32_t* offsets, int32_t ns, int chunk, int32_t* chunk_off,
                 int32_t* scan_blocks, ChunkDesc* desc, hipStream_t s) {
  hipLaunchKernelGGL(chunk_count_kernel, dim3(grid_for(int64_t(ns) + 1, 256)), dim3(256), 0, s,
                     offsets, ns, chunk, chunk_off);
  int rc = exclusive_scan_i32(chunk_off, int64_t(ns) + 1, scan_blocks, s);
  if (rc) return rc;
  hipLaunchKernelGGL(chunk_fill_kernel, dim3(grid_for(ns, 256)), dim3(256), 0, s, offsets, ns,
                     chunk, chunk_off, desc);
  return check_launch("build_chunks");
}

void combine_chunks(const double* partial, const int32_t* chunk_off, int32_t ns, int32_t nt,
                    float* cube, hipStream_t s, int32_t bands) {
  hipLaunchKernelGGL(cube_combine_kernel, dim3(grid_for(int64_t(ns) * nt, 256)), dim3(256), 0, s,
                     partial, chunk_off, ns, bands, nt, cube);
}

size_t cube_workspace_impl(int64_t n, int32_t n_keys, int32_t nt, const Tune& t) {
  if (n < 0 || n_keys < 1 || nt < 1) return 0;
  SizeOnly so;
  cube_layout(so, n, n_keys, nt, t.chunk, nullptr);
  return so.s.used;
}

struct FusedLaunch {
  int64_t grid;
  hipStream_t s;
  SspView g;
  rbx_target tg;
  StarState* state;
  const ChunkDesc* desc;
  const int32_t* n_chunks;
  int slab_bins;
  int64_t items;
  const float* axav;
  double* partial;
};

template <int NB>
void launch_fused(const FusedLaunch& f) {
  if (f.axav)
    hipLaunchKernelGGL((cube_fused_kernel<true, NB>), dim3(f.grid), dim3(64 * kFusedWaves), 0,
                       f.s, f.g, f.tg, f.state, f.desc, f.n_chunks, f.slab_bins, f.items, f.axav,
                       f.partial);
  else
    hipLaunchKernelGGL((cube_fused_kernel<false, NB>), dim3(f.grid), dim3(64 * kFusedWaves), 0,
                       f.s, f.g, f.tg, f.state, f.desc, f.n_chunks, f.slab_bins, f.items,
                       nullptr, f.partial);
}

int cube_fwd_impl(const rbx_ssp* ssp, const rbx_target* target, const rbx_stars* stars,
                  const rbx_extinction* ext, const Tune& t, float* cube, void* workspace,
                  size_t workspace_bytes, hipStream_t s, const char* who) {
  int rc = validate(ssp, target, stars);
  if (rc) return rc;
  if (ext && ((stars->n > 0 && !ext->av) || !ext->axav))
    return failf(RBX_EINVAL, "%s: extinction descriptor needs av and axav", who);
  const float* axav = ext ? ext->axav : nullptr;
  if (!cube) return failf(RBX_EINVAL, "%s: null cube", who);
  const int64_t n = stars->n;
  const int32_t nt = target->n;
  Carver c(workspace, workspace_bytes);
  CubeLayout L;
  cube_layout(c, n, stars->n_keys, nt, t.chunk, &L);
  if (!c.ok) return failf(RBX_EWORKSPACE, "%s: workspace too small", who);
  const SspView g = view_of(*ssp);
  const int32_t ns = stars->n_spaxel;
  const int bands = band_count(*ssp, stars->n_keys);
  const int32_t n_band_keys = bands * ns;  // <= band_cap(n_keys) * n_keys
  if (n > 0)
    hipLaunchKernelGGL(band_key_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, g, *stars, bands,
                       L.keys);
  rc = spaxel_sort_impl(L.keys, n, n_band_keys, L.perm, L.skey, L.offsets, L.sort_ws,
                        L.sort_ws_bytes, s);
  if (rc) return rc;
  const int n_tiles = (nt + kTile - 1) / kTile;
  const int tps = (n_tiles + kSlabs - 1) / kSlabs;
  // ascending targets with at most kMaxFusedBins bins per lane: one fused
  // pass (cube_fused_kernel, NB bins per lane); otherwise (or when the
  // tuning asks for it) star_norm + star_sf, then the per-tile accumulation
  const int nb_lane = (tps + 1) / 2 < 4 ? 4 : (tps + 1) / 2;
  const bool fused = target->ascending && nb_lane <= kMaxFusedBins &&
                     !(t.flags & RBX_TUNE_TWO_PASS);
  if (n > 0) {
    hipLaunchKernelGGL(star_prep_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, g, *target,
                       *stars, L.perm, L.skey, ext ? ext->av : nullptr, L.state, L.gradw);
  }
  if (n > 0 && !fused) {
    const int64_t items = int64_t(kSlabs) * ((n + t.waves - 1) / t.waves);
    const int64_t grid = item_grid(items, t.persist, 64 * t.waves);
    switch (t.waves) {
      case 1: launch_norm<1>(grid, s, g, *target, n, items, L.state, tps * kTile, L.nn_part); break;
      case 2: launch_norm<2>(grid, s, g, *target, n, items, L.state, tps * kTile, L.nn_part); break;
      case 8: launch_norm<8>(grid, s, g, *target, n, items, L.state, tps * kTile, L.nn_part); break;
      default: launch_norm<4>(grid, s, g, *target, n, items, L.state, tps * kTile, L.nn_part);
    }
    if ((rc = check_launch(who))) return rc;
    hipLaunchKernelGGL(star_sf_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, L.nn_part,
                       L.state);
  }
  rc = build_chunks(L.offsets, n_band_keys, t.chunk, L.chunk_off, L.scan_blocks, L.desc, s);
  if (rc) return rc;
  const int tpx = (n_tiles + 7) / 8;
  if (n > 0 && fused) {
    const int64_t items = L.max_chunks;
    const int64_t grid = item_grid(items, t.persist, 64 * kFusedWaves);
    const FusedLaunch f{grid, s, g, *target, L.state, L.desc, L.chunk_off + n_band_keys,
                        tps * kTile, items, axav, L.partial};
    switch (nb_lane) {
      case 4: launch_fused<4>(f); break;
      case 5: launch_fused<5>(f); break;
      case 6: launch_fused<6>(f); break;
      case 7: launch_fused<7>(f); break;
      default: launch_fused<8>(f);
    }
    if ((rc = check_launch(who))) return rc;
  } else if (n > 0) {
    const int64_t items = int64_t(8) * tpx * L.max_chunks;
    if (items >= (int64_t(1) << 34))
      return failf(RBX_EINVAL, "%s: problem too large for one launch", who);
    const int64_t grid = item_grid(items, t.persist, 64);
    if (axav)
      hipLaunchKernelGGL(cube_accum_kernel<true>, dim3(grid), dim3(64), 0, s, g, *target, L.state,
                         L.desc, L.chunk_off + n_band_keys, n_tiles, tpx, items, axav, L.partial);
    else
      hipLaunchKernelGGL(cube_accum_kernel<false>, dim3(grid), dim3(64), 0, s, g, *target,
                         L.state, L.desc, L.chunk_off + n_band_keys, n_tiles, tpx, items, nullptr,
                         L.partial);
    if ((rc = check_launch(who))) return rc;
  }
  combine_chunks(L.partial, L.chunk_off, ns, nt, cube, s, bands);
  return check_launch(who);
}

int cube_bwd_impl(const rbx_ssp* ssp, const rbx_target* target, const rbx_stars* stars,
                  const rbx_extinction* ext, const Tune& t, const float* dcube, float* dmass,
                  float* dage, float* dmet, float* dvlos, void* workspace,
                  size_t workspace_bytes, hipStream_t s, const char* who) {
  int rc = validate(ssp, target, stars);
  if (rc) return rc;
  if (ext && !ext->axav)
    return failf(RBX_EINVAL, "%s: extinction descriptor needs axav", who);
  const float* axav = ext ? ext->axav : nullptr;
  if (!dcube) return failf(RBX_EINVAL, "%s: null dcube", who);
  const int64_t n = stars->n;
  Carver c(workspace, workspace_bytes);
  CubeLayout L;
  cube_layout(c, n, stars->n_keys, target->n, t.chunk, &L, true);
  if (!c.ok) return failf(RBX_EWORKSPACE, "%s: workspace too small", who);
  if (n == 0) return RBX_OK;
  const SspView g = view_of(*ssp);
  const int32_t nt = target->n;
  const int n_tiles = (nt + kTile - 1) / kTile;
  const int tps = (n_tiles + kBwdSlabs - 1) / kBwdSlabs;
  const int64_t items = int64_t(kBwdSlabs) * ((n + t.waves - 1) / t.waves);
  const int64_t grid = item_grid(items, t.persist, 64 * t.waves);
  switch (t.waves) {
    case 1: launch_bwd<1>(grid, s, g, *target, n, items, L.state, L.gradw, dcube, tps * kTile, axav, L.vpart, L.apart); break;
    case 2: launch_bwd<2>(grid, s, g, *target, n, items, L.state, L.gradw, dcube, tps * kTile, axav, L.vpart, L.apart); break;
    case 8: launch_bwd<8>(grid, s, g, *target, n, items, L.state, L.gradw, dcube, tps * kTile, axav, L.vpart, L.apart); break;
    default: launch_bwd<4>(grid, s, g, *target, n, items, L.state, L.gradw, dcube, tps * kTile, axav, L.vpart, L.apart);
  }
  if ((rc = check_launch(who))) return rc;
  hipLaunchKernelGGL(star_grad_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, g, *stars,
                     L.state, L.vpart, L.apart, dmass, dage, dmet, dvlos);
  return check_launch(who);
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
  rbx::Tune t;
  rbx::tune_of(nullptr, &t);
  return rbx::cube_workspace_impl(n, n_keys, n_target, t);
}

int rbx_cube_fwd(const rbx_ssp* ssp, const rbx_target* target, const rbx_stars* stars,
                 float* cube, void* workspace, size_t workspace_bytes, void* stream) {
  rbx::Tune t;
  rbx::tune_of(nullptr, &t);
  return rbx::cube_fwd_impl(ssp, target, stars, nullptr, t, cube, workspace, workspace_bytes,
                            rbx::as_stream(stream), "rbx_cube_fwd");
}

int rbx_cube_fwd_ext(const rbx_ssp* ssp, const rbx_target* target, const rbx_stars* stars,
                     const rbx_extinction* ext, float* cube, void* workspace,
                     size_t workspace_bytes, void* stream) {
  rbx::Tune t;
  rbx::tune_of(nullptr, &t);
  return rbx::cube_fwd_impl(ssp, target, stars, ext, t, cube, workspace, workspace_bytes,
                            rbx::as_stream(stream), "rbx_cube_fwd_ext");
}

int rbx_cube_bwd(const rbx_ssp* ssp, const rbx_target* target, const rbx_stars* stars,
                 const float* dcube, float* dmass, float* dage, float* dmet, float* dvlos,
                 void* workspace, size_t workspace_bytes, void* stream) {
  rbx::Tune t;
  rbx::tune_of(nullptr, &t);
  return rbx::cube_bwd_impl(ssp, target, stars, nullptr, t, dcube, dmass, dage, dmet, dvlos,
                            workspace, workspace_bytes, rbx::as_stream(stream), "rbx_cube_bwd");
}

int rbx_cube_bwd_ext(const rbx_ssp* ssp, const rbx_target* target, const rbx_stars* stars,
                     const rbx_extinction* ext, const float* dcube, float* dmass, float* dage,
                     float* dmet, float* dvlos, void* workspace, size_t workspace_bytes,
                     void* stream) {
  rbx::Tune t;
  rbx::tune_of(nullptr, &t);
  return rbx::cube_bwd_impl(ssp, target, stars, ext, t, dcube, dmass, dage, dmet, dvlos,
                            workspace, workspace_bytes, rbx::as_stream(stream),
                            "rbx_cube_bwd_ext");
}

}  // extern "C"
