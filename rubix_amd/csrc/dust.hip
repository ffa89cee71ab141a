// Dust extinction of the calc_dusty_ifu chain (rubix/core/dust.py:16-65 over
// rubix/spectra/dust/dust_extinction.py:163-358 and extinction_models.py).
//
// Per gas cell: the dust-to-gas ratio (Remy-Ruyer et al. 2014, table 1) of
// log(O/H) + 12, the dust mass and its V-band extinction per spaxel area.
//
// Per star: A_V = the cumulative extinction of its spaxel's gas along z,
// interpolated at the star's z.  The reference (dust_extinction.py:287-327)
// evaluates this for spaxel s as jnp.interp(z_star, xp, fp, left=
// "extrapolate") over ALL gas cells in (pixel, z) lexsort order, where
//   key  = z (own cells)  or  fl(z * 1e30) (cells of other spaxels),
//   value = the cumulative sum of the own extinction (own cells) or 0,
// stably sorted by key.  Cells of other spaxels therefore sit at huge
// |keys| (or at 0 when z == 0) and still decide the neighbours of a star
// below or above its own spaxel's gas.  This file computes the same
// neighbours per star without building the merged array: binary searches in
// the star's own segment and in one global z order, where "the nearest cell
// of another spaxel" is one step through a table of equal-pixel runs.
//
// Per (star, wavelength): f = 10^(-0.4 A(lambda)/A(V) * A_V) with the
// Cardelli89 or Gordon23 curve, spectra *= f (dust_extinction.py:330-358).
#include <math.h>

#include "rbx_common.h"
#include "ssp_device.h"

namespace rbx {

constexpr int kDustThreads = 256;

__device__ inline float mul_nc32(float x, float y) {
#pragma clang fp contract(off)
  return x * y;
}
__device__ inline float add_nc32(float x, float y) {
#pragma clang fp contract(off)
  return x + y;
}

// ------------------------------------------------------------ curves
// A(lambda)/A(V); the argument is wavelength / 1e4 as the reference passes it.

__device__ inline float cardelli89_axav(float w, float rv) {
  // extinction_models.py:104-178; the branches treat w as a wavenumber.
  float a = 0.f, b = 0.f;
  if (0.3f <= w && w < 1.1f) {
    const float p = powf(w, 1.61f);
    a = 0.574f * p;
    b = -0.527f * p;
  }
  if (1.1f <= w && w < 3.3f) {
    const float y = w - 1.82f, y2 = y * y, y3 = y2 * y, y4 = y3 * y, y5 = y4 * y, y6 = y5 * y,
                y7 = y6 * y;
    a = 1.f + 0.17699f * y - 0.50447f * y2 - 0.02427f * y3 + 0.72085f * y4 + 0.01979f * y5 -
        0.77530f * y6 + 0.32999f * y7;
    b = 1.41338f * y + 2.28305f * y2 + 1.07233f * y3 - 5.38434f * y4 - 0.62251f * y5 +
        5.30260f * y6 - 2.09002f * y7;
  }
  if (3.3f <= w && w <= 8.0f) {
    const float da = w - 4.67f, db = w - 4.62f;
    a = 1.752f - 0.316f * w - 0.104f / (da * da + 0.341f);
    b = -3.09f + 1.825f * w + 1.206f / (db * db + 0.263f);
  }
  if (5.9f <= w && w <= 8.f) {
    const float y = w - 5.9f, y2 = y * y, y3 = y2 * y;
    a = a + (-0.04473f * y2 - 0.009779f * y3);
    b = b + (0.2130f * y2 + 0.1207f * y3);
  }
  if (8.f < w && w <= 10.f) {
    const float y = w - 8.f, y2 = y * y, y3 = y2 * y;
    a = -1.073f - 0.628f * y + 0.137f * y2 - 0.070f * y3;
    b = 13.670f + 4.257f * y - 0.420f * y2 + 0.374f * y3;
  }
  return a + b / rv;
}

__device__ inline float smoothstep1(float x, float lo, float hi) {
  float t = (x - lo) / (hi - lo);
  t = fminf(fmaxf(t, 0.f), 1.f);
  return (3.f - 2.f * t) * t * t;
}

__device__ inline float drude(float x, float amp, float x0, float fwhm) {
  const float r = fwhm / x0, d = x / x0 - x0 / x;
  return amp * (r * r) / (d * d + r * r);
}

__device__ inline float modified_drude(float x, float scale, float xo, float gamma_o, float asym) {
  const float g = 2.f * gamma_o / (1.f + expf(asym * (x - xo)));
  const float r = g / xo, d = x / xo - xo / x;
  return scale * (r * r) / (d * d + r * r);
}

__device__ inline float fm90(float x, float c1, float c2, float c3, float c4, float xo,
                             float gamma) {
  float e = c1 + c2 * x;
  const float x2 = x * x, q = x2 - xo * xo;
  e = e + c3 * (x2 / (q * q + x2 * (gamma * gamma)));
  if (x >= 5.9f) {
    const float y = x - 5.9f;
    e = e + c4 * (0.5392f * (y * y) + 0.05644f * (y * y * y));
  }
  return e;
}

// extinction_models.py:315-351 with 5 polynomial coefficients and 3 Drudes.
__device__ inline float g23_optical(float x, const float* p) {
  float poly = p[4];
  for (int i = 3; i >= 0; --i) poly = p[i] + poly * x;
  return poly + drude(x, p[5], p[6], p[7]) + drude(x, p[8], p[9], p[10]) +
         drude(x, p[11], p[12], p[13]);
}

// extinction_models.py:401-445
__device__ inline float g23_nirmir(float w) {
  const float scale = 0.38526f, alpha = 1.68467f, alpha2 = 0.78791f, swave = 4.30578f,
              swidth = 4.78338f;
  const float pow1 = scale * powf(w, -alpha);
  const float ratio = powf(swave, -alpha) / powf(swave, -alpha2);
  const float pow2 = scale * ratio * powf(w, -alpha2);
  const float wt = smoothstep1(w, swave - swidth / 2.f, swave + swidth / 2.f);
  float ax = pow1 * (1.f - wt) + pow2 * wt;
  ax += modified_drude(w, 0.06652f, 9.8434f, 2.21205f, -0.24703f);
  ax += modified_drude(w, 0.0267f, 19.58294f, 17.f, -0.27f);
  return ax;
}

__device__ inline float gordon23_axav(float w, float rv) {
  // extinction_models.py:255-398, w in microns
  const float opt_a[14] = {-0.35848f, 0.7122f, 0.08746f, -0.05403f, 0.00674f,
                           0.03893f,  2.288f,  0.243f,   0.02965f,  2.054f,
                           0.179f,    0.01747f, 1.587f,  0.243f};
  const float opt_b[14] = {0.12354f, -2.68335f, 2.01901f, -0.39299f, 0.03355f,
                           0.18453f, 2.288f,    0.243f,   0.19728f,  2.054f,
                           0.179f,   0.1713f,   1.587f,   0.243f};
  float a = 0.f, b = 0.f;
  const float x = 1.f / w;
  const bool ir = 1.0f <= w && w < 35.f, opt = 0.3f <= w && w < 1.1f,
             uv = 0.09f <= w && w <= 0.3f;
  const bool optir = w >= 0.9f && w <= 1.1f, uvopt = w >= 0.3f && w <= 0.33f;
  const float ir_b = -1.01251f * powf(w / 1.f, 1.06099f);
  if (ir) {
    a = g23_nirmir(w);
    b = ir_b;
  }
  if (opt) {
    a = g23_optical(x, opt_a);
    b = g23_optical(x, opt_b);
  }
  if (optir) {
    const float wt = smoothstep1(w, 0.9f, 1.1f);
    a = (1.f - wt) * g23_optical(x, opt_a) + wt * g23_nirmir(w);
    b = (1.f - wt) * g23_optical(x, opt_b) + wt * ir_b;
  }
  if (uv) {
    a = fm90(x, 0.81297f, 0.2775f, 1.06295f, 0.11303f, 4.60f, 0.99f);
    b = fm90(x, -2.97868f, 1.89808f, 3.10334f, 0.65484f, 4.60f, 0.99f);
  }
  if (uvopt) {
    const float wt = smoothstep1(w, 0.3f, 0.33f);
    a = (1.f - wt) * fm90(x, 0.81297f, 0.2775f, 1.06295f, 0.11303f, 4.60f, 0.99f) +
        wt * g23_optical(x, opt_a);
    b = (1.f - wt) * fm90(x, -2.97868f, 1.89808f, 3.10334f, 0.65484f, 4.60f, 0.99f) +
        wt * g23_optical(x, opt_b);
  }
  return a + b * (1.f / rv - 1.f / 3.1f);
}

__global__ void __launch_bounds__(256) ext_curve_kernel(int model, float rv,
                                                        const float* __restrict__ wave, int n,
                                                        float* __restrict__ axav) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float w = wave[i] / 1e4f;
    axav[i] = model == RBX_EXT_GORDON23 ? gordon23_axav(w, rv) : cardelli89_axav(w, rv);
  }
}

// ------------------------------------------------------------ gas

struct DustParams {
  int model;  // RBX_DTG_*
  int xco;    // RBX_XCO_*
  float num;  // float32(3 pi) * ... see gas_ext_kernel
  float conv; // MSUN_TO_GRAMS / KPC_TO_CM^2 as float32
  float den;  // float32 of (0.4 ln 10) * 5448e-8 * grain density, rounded step by step
  float area; // spaxel area (kpc^2)
};

// log(O/H) + 12 from metals[:, 4] / (16 metals[:, 0]), D/G, dust mass and
// 3 pi Sigma C / (0.4 ln 10 lambda_V rho) / area (dust_extinction.py:13-160,
// 269-287), float32 step by step.
__global__ void __launch_bounds__(256) gas_ext_kernel(const float* __restrict__ mass,
                                                      const float* __restrict__ metals,
                                                      int64_t n, int64_t mstride, DustParams p,
                                                      float* __restrict__ ext) {
  for (int64_t i = gtid(); i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const float h = metals[i * mstride + 0], o = metals[i * mstride + 4];
    const float log_oh = 12.f + log10f(o / (16.f * h));
    const float xs = 8.69f - log_oh;
    float e;
    if (p.model == RBX_DTG_POWER_LAW_FREE) {
      const float alpha = p.xco == RBX_XCO_MW ? 1.62f : 2.02f;
      e = 2.21f + alpha * xs;
    } else {
      const float xt = p.xco == RBX_XCO_MW ? 7.96f : 8.10f;
      const float bl = p.xco == RBX_XCO_MW ? 0.68f : 0.96f;
      const float al = p.xco == RBX_XCO_MW ? 3.08f : 3.10f;
      e = log_oh > xt ? 2.21f + 1.0f * xs : bl + al * xs;
    }
    const float dg = 1.f / exp10f(e);
    const float dust = mass[i] * dg;
    ext[i] = mul_nc32(mul_nc32(p.num, dust), p.conv) / p.den / p.area;
  }
}

// ------------------------------------------------------------ sort helpers

// Order-preserving bits of z with jax's sort conventions: -0 == +0, every
// NaN after +inf.
__device__ inline uint32_t z_order_key(float z) {
  if (z != z) return 0xFFFFFFFFu;
  if (z == 0.f) z = 0.f;
  const uint32_t b = __float_as_uint(z);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ void z_keys_kernel(const float* __restrict__ coords, int64_t stride, int64_t n,
                              uint32_t* __restrict__ key, int32_t* __restrict__ perm) {
  for (int64_t i = gtid(); i < n; i += int64_t(gridDim.x) * blockDim.x) {
    key[i] = z_order_key(coords[i * stride + 2]);
    perm[i] = int32_t(i);
  }
}

__global__ void digit_kernel(const uint32_t* __restrict__ key, const int32_t* __restrict__ perm,
                             int64_t n, int shift, int32_t* __restrict__ digit) {
  for (int64_t i = gtid(); i < n; i += int64_t(gridDim.x) * blockDim.x)
    digit[i] = int32_t((key[perm[i]] >> shift) & 255u);
}

__global__ void gather_pix_kernel(const int32_t* __restrict__ pix, const int32_t* __restrict__ perm,
                                  int64_t n, int32_t* __restrict__ out) {
  for (int64_t i = gtid(); i < n; i += int64_t(gridDim.x) * blockDim.x) out[i] = pix[perm[i]];
}

__global__ void compose_kernel(const int32_t* __restrict__ perm, const int32_t* __restrict__ pp,
                               int64_t n, int32_t* __restrict__ out) {
  for (int64_t i = gtid(); i < n; i += int64_t(gridDim.x) * blockDim.x) out[i] = perm[pp[i]];
}

// Global z order: zs, ps, run-start flags; zero-z pixel range and the
// non-finite-extinction flag (stats: min pixel, max pixel, any non-finite).
__global__ void zorder_kernel(const float* __restrict__ coords, int64_t stride,
                              const int32_t* __restrict__ pix, const float* __restrict__ ext,
                              const int32_t* __restrict__ perm_z, int64_t n,
                              float* __restrict__ zs, int32_t* __restrict__ ps,
                              int32_t* __restrict__ flag, int32_t* __restrict__ stats) {
  for (int64_t i = gtid(); i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int32_t g = perm_z[i];
    const float z = coords[int64_t(g) * stride + 2];
    const int32_t p = pix[g];
    zs[i] = z;
    ps[i] = p;
    flag[i] = (i == 0 || pix[perm_z[i - 1]] != p) ? 1 : 0;
    if (z == 0.f) {
      atomicMin(&stats[0], p);
      atomicMax(&stats[1], p);
    }
    if (!isfinite(ext[g])) atomicOr(&stats[2], 1);
  }
}

__global__ void init_stats_kernel(int32_t* stats) {
  if (threadIdx.x == 0) {
    stats[0] = INT32_MAX;
    stats[1] = INT32_MIN;
    stats[2] = 0;
    stats[3] = 0;
  }
}

// run_id[i] = inclusive count of run starts - 1 (flag holds the exclusive
// scan on entry, start_flag recomputed); run_start[r] = first index of run
// r, run_start[n_runs] = n.
__global__ void runs_kernel(const int32_t* __restrict__ ps, const int32_t* __restrict__ excl,
                            int64_t n, int32_t* __restrict__ run_id,
                            int32_t* __restrict__ run_start) {
  for (int64_t i = gtid(); i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const bool start = i == 0 || ps[i - 1] != ps[i];
    const int32_t r = excl[i] + (start ? 1 : 0) - 1;
    run_id[i] = r;
    if (start) run_start[r] = int32_t(i);
    if (i == n - 1) run_start[r + 1] = int32_t(n);
  }
}

// One workgroup per spaxel segment: own z in lexsort order and the inclusive
// cumulative extinction (double, rounded once per entry).
__global__ void __launch_bounds__(kDustThreads) segment_cum_kernel(
    const float* __restrict__ coords, int64_t stride, const float* __restrict__ ext,
    const int32_t* __restrict__ sorted, const int32_t* __restrict__ off, int n_seg,
    float* __restrict__ seg_z, float* __restrict__ cum) {
  __shared__ double sm[kDustThreads];
  for (int s = blockIdx.x; s < n_seg; s += gridDim.x) {
    const int b = off[s], e = off[s + 1];
    double carry = 0.0;
    for (int base = b; base < e; base += kDustThreads) {
      const int j = base + threadIdx.x;
      double v = 0.0;
      if (j < e) {
        const int32_t g = sorted[j];
        seg_z[j] = coords[int64_t(g) * stride + 2];
        v = ext[g];
      }
      sm[threadIdx.x] = v;
      __syncthreads();
      for (int off2 = 1; off2 < kDustThreads; off2 <<= 1) {
        const double t = threadIdx.x >= off2 ? sm[threadIdx.x - off2] : 0.0;
        __syncthreads();
        sm[threadIdx.x] += t;
        __syncthreads();
      }
      if (j < e) cum[j] = float(carry + sm[threadIdx.x]);
      carry += sm[kDustThreads - 1];
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------ per star

struct Elem {
  float key, val;
  int32_t pix;
  bool own, ok;
};

struct ColumnView {
  const int32_t* off;
  const float* seg_z;
  const float* cum;
  const float* zs;
  const int32_t* ps;
  const int32_t* run_id;
  const int32_t* run_start;
  const int32_t* stats;
  int32_t n;  // gas cells
};

constexpr float kBig = 1e30f;

__device__ inline float big_key(float z) { return mul_nc32(z, kBig); }

// number of a[lo..hi) with key(a) <= x (a ascending in key)
template <bool kBigKey>
__device__ inline int count_le(const float* a, int lo, int hi, float x) {
  int l = lo, h = hi;
  while (l < h) {
    const int m = l + ((h - l) >> 1);
    const float k = kBigKey ? big_key(a[m]) : a[m];
    if (k <= x) l = m + 1;
    else h = m;
  }
  return l - lo;
}

__device__ inline Elem own_at(const ColumnView& v, int b, int e, int k) {
  Elem r{0.f, 0.f, 0, true, false};
  if (k >= 0 && k < e - b) {
    r.key = v.seg_z[b + k];
    r.val = v.cum[b + k];
    r.ok = true;
  }
  return r;
}

__device__ inline Elem foreign_at(const ColumnView& v, int q) {
  Elem r{0.f, 0.f, 0, false, false};
  if (q >= 0 && q < v.n) {
    r.key = big_key(v.zs[q]);
    r.pix = v.ps[q];
    r.ok = true;
  }
  return r;
}

// nearest position >= q (<= q) in the global z order whose pixel is not s
__device__ inline int foreign_fwd(const ColumnView& v, int q, int s) {
  if (q >= v.n) return -1;
  if (v.ps[q] != s) return q;
  const int nq = v.run_start[v.run_id[q] + 1];
  return nq < v.n ? nq : -1;
}
__device__ inline int foreign_bwd(const ColumnView& v, int q, int s) {
  if (q < 0) return -1;
  if (v.ps[q] != s) return q;
  return v.run_start[v.run_id[q]] - 1;
}

// Stable-sort order between an own and a foreign element of equal key: the
// merged array keeps the lexsort (pixel, z) order, so own (pixel s) precedes
// a foreign cell of pixel p iff s < p.  For key 0 every foreign zero counts
// (stats hold the pixel range of the z == 0 cells); other equal keys (a
// foreign |z| below 1e-29) compare the candidate's pixel only.
__device__ inline bool own_last_among_equal(const ColumnView& v, const Elem& f, int s) {
  const int pmax = f.key == 0.f ? v.stats[1] : f.pix;
  return !(pmax > s);
}
__device__ inline bool foreign_first_among_equal(const ColumnView& v, const Elem& f, int s) {
  const int pmin = f.key == 0.f ? v.stats[0] : f.pix;
  return pmin < s;
}

// the later (in merged order) of an own and a foreign candidate
__device__ inline Elem later(const ColumnView& v, const Elem& o, const Elem& f, int s) {
  if (!o.ok) return f;
  if (!f.ok) return o;
  if (o.key != f.key) return o.key > f.key ? o : f;
  return own_last_among_equal(v, f, s) ? o : f;
}
// the earlier of an own and a foreign candidate
__device__ inline Elem earlier(const ColumnView& v, const Elem& o, const Elem& f, int s) {
  if (!o.ok) return f;
  if (!f.ok) return o;
  if (o.key != f.key) return o.key < f.key ? o : f;
  return foreign_first_among_equal(v, f, s) ? f : o;
}

// jnp.interp between merged neighbours e0 (position i-1) and e1 (position i)
__device__ inline float interp_pair(float x, const Elem& e0, const Elem& e1) {
  const float df = e1.val - e0.val;
  const float dx = e1.key - e0.key;
  const float delta = x - e0.key;
  if (fabsf(dx) <= 0x1p-46f) return e0.val;  // spacing(eps(float32))
  return add_nc32(e0.val, mul_nc32(delta / dx, df));
}

__device__ float column_av(const ColumnView& v, int s, float x) {
  const int b = v.off[s], e = v.off[s + 1], len = e - b;
  const int n = v.n;
  if (n == 0) return 0.f;
  if (n == 1) {
    const Elem o = own_at(v, b, e, 0);
    return o.ok ? o.val : 0.f;
  }
  const int own_le = count_le<false>(v.seg_z, b, e, x);
  const int own_le_big = count_le<true>(v.seg_z, b, e, x);
  const int all_le_big = count_le<true>(v.zs, 0, n, x);
  const int count = own_le + (all_le_big - own_le_big);
  if (count == 0) {
    // left extrapolation from the first two merged elements
    const Elem o0 = own_at(v, b, e, 0);
    const int q0 = foreign_fwd(v, 0, s);
    const Elem f0 = foreign_at(v, q0);
    const Elem e0 = earlier(v, o0, f0, s);
    Elem e1;
    if (e0.own) e1 = earlier(v, own_at(v, b, e, 1), f0, s);
    else e1 = earlier(v, o0, foreign_at(v, foreign_fwd(v, q0 + 1, s)), s);
    return interp_pair(x, e0, e1);
  }
  if (count == n) {
    const Elem ol = own_at(v, b, e, len - 1);
    const int ql = foreign_bwd(v, n - 1, s);
    const Elem fl = foreign_at(v, ql);
    const Elem l1 = later(v, ol, fl, s);
    if (x > l1.key) return l1.val;
    Elem l2;
    if (l1.own) l2 = later(v, own_at(v, b, e, len - 2), fl, s);
    else l2 = later(v, ol, foreign_at(v, foreign_bwd(v, ql - 1, s)), s);
    return interp_pair(x, l2, l1);
  }
  const Elem l1 = later(v, own_at(v, b, e, own_le - 1),
                        foreign_at(v, foreign_bwd(v, all_le_big - 1, s)), s);
  const Elem r1 = earlier(v, own_at(v, b, e, own_le),
                          foreign_at(v, foreign_fwd(v, all_le_big, s)), s);
  return interp_pair(x, l1, r1);
}

__global__ void __launch_bounds__(256) star_av_kernel(ColumnView v, const float* __restrict__ sc,
                                                      int64_t sstride,
                                                      const int32_t* __restrict__ spix,
                                                      int64_t n_star, int n_spaxels,
                                                      float* __restrict__ av) {
  const bool poisoned = v.stats[2] != 0;
  for (int64_t i = gtid(); i < n_star; i += int64_t(gridDim.x) * blockDim.x) {
    const int s = spix[i];
    float a = 0.f;
    if (s >= 0 && s < n_spaxels) a = poisoned ? __int_as_float(0x7fc00000) : column_av(v, s, sc[i * sstride + 2]);
    av[i] = a;
  }
}

// ------------------------------------------------------------ apply

// out[i, l] = in[i, l] * 10^((-0.4 axav[l]) * av[i]), one workgroup per row
// (also the adjoint: the factor does not depend on the spectra).  in and out
// may alias.
__global__ void __launch_bounds__(256) apply_ext_kernel(const float* in, const float* __restrict__ av,
                                                        const float* __restrict__ axav, int64_t n,
                                                        int nw, float* out) {
  for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
    const float a = av[i];
    const int64_t row = i * nw;
    for (int l = threadIdx.x; l < nw; l += blockDim.x) {
      const float t = mul_nc32(-0.4f, axav[l]);
      out[row + l] = in[row + l] * exp10f(mul_nc32(t, a));
    }
  }
}

int grid_of(int64_t n, int threads = 256, int64_t cap = 1 << 20) {
  int64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return int(g);
}

struct DustLayout {
  uint32_t* key;
  int32_t *digit, *perm_a, *perm_b, *pp, *offs_tmp, *offsets, *flag, *blocks;
  float *seg_z, *cum, *zs;
  int32_t *ps, *run_id, *run_start, *stats;
  void* sort_ws;
  size_t sort_ws_bytes;
};

template <class C>
void dust_layout(C& c, int64_t n, int32_t n_keys, DustLayout* L) {
  const int32_t kmax = n_keys > 256 ? n_keys : 256;
  L->key = c.template take<uint32_t>(n);
  L->digit = c.template take<int32_t>(n);
  L->perm_a = c.template take<int32_t>(n);
  L->perm_b = c.template take<int32_t>(n);
  L->pp = c.template take<int32_t>(n);
  L->offs_tmp = c.template take<int32_t>(kmax + 1);
  L->offsets = c.template take<int32_t>(int64_t(n_keys) + 1);
  L->flag = c.template take<int32_t>(n);
  L->blocks = c.template take<int32_t>(n / 4096 + 2);
  L->seg_z = c.template take<float>(n);
  L->cum = c.template take<float>(n);
  L->zs = c.template take<float>(n);
  L->ps = c.template take<int32_t>(n);
  L->run_id = c.template take<int32_t>(n);
  L->run_start = c.template take<int32_t>(n + 1);
  L->stats = c.template take<int32_t>(4);
  const size_t a = sort_workspace_size(n, 256), b = sort_workspace_size(n, n_keys);
  L->sort_ws_bytes = a > b ? a : b;
  L->sort_ws = c.template take<char>(L->sort_ws_bytes);
}

struct SizeOnly {
  Sizer s;
  template <class T>
  T* take(size_t count) {
    s.take<T>(count);
    return nullptr;
  }
};

}  // namespace rbx

extern "C" {

int rbx_extinction_curve(int32_t model, float rv, const float* wave, int32_t n, float* axav,
                         void* stream) {
  using namespace rbx;
  if (model != RBX_EXT_CARDELLI89 && model != RBX_EXT_GORDON23)
    return failf(RBX_EINVAL, "rbx_extinction_curve: unknown model %d (0 = Cardelli89, 1 = Gordon23)",
                 model);
  if (n < 0 || (n > 0 && (!wave || !axav)))
    return fail(RBX_EINVAL, "rbx_extinction_curve: bad arguments");
  if (n == 0) return RBX_OK;
  hipLaunchKernelGGL(ext_curve_kernel, dim3(grid_of(n)), dim3(256), 0, as_stream(stream), model,
                     rv, wave, n, axav);
  return check_launch("rbx_extinction_curve");
}

int rbx_gas_extinction(const float* mass, const float* metals, int64_t n, int64_t metals_stride,
                       int32_t dtg_model, int32_t xco, float grain_density, float spaxel_area,
                       float* ext, void* stream) {
  using namespace rbx;
  if (dtg_model != RBX_DTG_POWER_LAW_FREE && dtg_model != RBX_DTG_BROKEN_POWER_LAW)
    return failf(RBX_EINVAL, "rbx_gas_extinction: unknown dust-to-gas model %d", dtg_model);
  if (xco != RBX_XCO_MW && xco != RBX_XCO_Z)
    return failf(RBX_EINVAL, "rbx_gas_extinction: unknown Xco %d (0 = MW, 1 = Z)", xco);
  if (n < 0 || metals_stride < 5 || (n > 0 && (!mass || !metals || !ext)))
    return fail(RBX_EINVAL, "rbx_gas_extinction: bad arguments (metals need >= 5 columns)");
  if (n == 0) return RBX_OK;
  DustParams p;
  p.model = dtg_model;
  p.xco = xco;
  // 3 * pi is a Python float multiplied into a float32 array: float32(3 pi);
  // the constant and the denominator are float32 values as the reference's
  // weakly typed scalars become (dust_extinction.py:150-158).
  p.num = float(3.0 * M_PI);
  p.conv = float(1.989e33 / (3.08568e21 * 3.08568e21));
  const float ln10 = logf(10.f);
  p.den = (0.4f * ln10) * float(5448 * 1e-8);
  p.den = p.den * grain_density;
  p.area = spaxel_area;
  hipLaunchKernelGGL(gas_ext_kernel, dim3(grid_of(n)), dim3(256), 0, as_stream(stream), mass,
                     metals, n, metals_stride, p, ext);
  return check_launch("rbx_gas_extinction");
}

size_t rbx_column_av_workspace_size(int64_t n_gas, int32_t n_keys) {
  using namespace rbx;
  if (n_gas < 0 || n_keys < 1) return 0;
  SizeOnly s;
  DustLayout L;
  dust_layout(s, n_gas, n_keys, &L);
  return s.s.used;
}

int rbx_column_av(const float* gas_coords, int64_t gas_stride, const int32_t* gas_pix,
                  const float* gas_ext, int64_t n_gas, const float* star_coords,
                  int64_t star_stride, const int32_t* star_pix, int64_t n_star,
                  int32_t n_spaxels, int32_t n_keys, void* workspace, size_t workspace_bytes,
                  float* av, void* stream) {
  using namespace rbx;
  if (n_gas < 0 || n_star < 0 || gas_stride < 3 || star_stride < 3 || n_keys < 1 ||
      n_spaxels < 0 || n_spaxels > n_keys || (n_gas > 0 && (!gas_coords || !gas_pix || !gas_ext)) ||
      (n_star > 0 && (!star_coords || !star_pix || !av)))
    return fail(RBX_EINVAL, "rbx_column_av: bad arguments");
  if (n_gas > INT32_MAX - 1) return fail(RBX_EINVAL, "rbx_column_av: n_gas must fit int32");
  if (int64_t(n_keys) + 1 > (int64_t(1) << 30))
    return fail(RBX_EINVAL, "rbx_column_av: n_keys too large");
  Carver c(workspace, workspace_bytes);
  DustLayout L;
  dust_layout(c, n_gas, n_keys, &L);
  if (!c.ok) return fail(RBX_EWORKSPACE, "rbx_column_av: workspace too small");
  hipStream_t st = as_stream(stream);
  const int64_t n = n_gas;
  hipLaunchKernelGGL(init_stats_kernel, dim3(1), dim3(64), 0, st, L.stats);
  if (n > 0) {
    // lexsort (pixel, z): 4 stable 8-bit passes over the z order keys, then
    // the stable counting sort by pixel
    hipLaunchKernelGGL(z_keys_kernel, dim3(grid_of(n)), dim3(256), 0, st, gas_coords, gas_stride,
                       n, L.key, L.perm_a);
    int32_t* cur = L.perm_a;
    int32_t* nxt = L.perm_b;
    for (int shift = 0; shift < 32; shift += 8) {
      hipLaunchKernelGGL(digit_kernel, dim3(grid_of(n)), dim3(256), 0, st, L.key, cur, n, shift,
                         L.digit);
      int rc = spaxel_sort_impl(L.digit, n, 256, L.pp, nullptr, L.offs_tmp, L.sort_ws,
                                L.sort_ws_bytes, st);
      if (rc) return rc;
      hipLaunchKernelGGL(compose_kernel, dim3(grid_of(n)), dim3(256), 0, st, cur, L.pp, n, nxt);
      int32_t* t = cur;
      cur = nxt;
      nxt = t;
    }
    // cur = global z order (ties: input order)
    hipLaunchKernelGGL(zorder_kernel, dim3(grid_of(n)), dim3(256), 0, st, gas_coords, gas_stride,
                       gas_pix, gas_ext, cur, n, L.zs, L.ps, L.flag, L.stats);
    int rc = exclusive_scan_i32(L.flag, n, L.blocks, st);
    if (rc) return rc;
    hipLaunchKernelGGL(runs_kernel, dim3(grid_of(n)), dim3(256), 0, st, L.ps, L.flag, n, L.run_id,
                       L.run_start);
    hipLaunchKernelGGL(gather_pix_kernel, dim3(grid_of(n)), dim3(256), 0, st, gas_pix, cur, n,
                       L.digit);
    rc = spaxel_sort_impl(L.digit, n, n_keys, L.pp, nullptr, L.offsets, L.sort_ws,
                          L.sort_ws_bytes, st);
    if (rc) return rc;
    hipLaunchKernelGGL(compose_kernel, dim3(grid_of(n)), dim3(256), 0, st, cur, L.pp, n, nxt);
    // nxt = lexsort order; segments off[s] .. off[s + 1]
    const int n_seg = n_spaxels;
    if (n_seg > 0)
      hipLaunchKernelGGL(segment_cum_kernel, dim3(n_seg < 65536 ? n_seg : 65536),
                         dim3(kDustThreads), 0, st, gas_coords, gas_stride, gas_ext, nxt,
                         L.offsets, n_seg, L.seg_z, L.cum);
  } else {
    fill_i32(L.offsets, int64_t(n_keys) + 1, 0, st);
  }
  if (n_star > 0) {
    ColumnView v{L.offsets, L.seg_z, L.cum, L.zs, L.ps, L.run_id, L.run_start, L.stats,
                 int32_t(n)};
    hipLaunchKernelGGL(star_av_kernel, dim3(grid_of(n_star)), dim3(256), 0, st, v, star_coords,
                       star_stride, star_pix, n_star, n_spaxels, av);
  }
  return check_launch("rbx_column_av");
}

int rbx_apply_extinction(const float* in, const float* av, const float* axav, int64_t n,
                         int32_t n_wave, float* out, void* stream) {
  using namespace rbx;
  if (n < 0 || n_wave < 1 || (n > 0 && (!in || !av || !axav || !out)))
    return fail(RBX_EINVAL, "rbx_apply_extinction: bad arguments");
  if (n == 0) return RBX_OK;
  const int blocks = int(n < (int64_t(1) << 20) ? n : (int64_t(1) << 20));
  hipLaunchKernelGGL(apply_ext_kernel, dim3(blocks), dim3(256), 0, as_stream(stream), in, av, axav,
                     n, n_wave, out);
  return check_launch("rbx_apply_extinction");
}

}  // extern "C"
