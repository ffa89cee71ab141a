/*
 * rubix_amd: C ABI of the MI355X (gfx950) particle-to-datacube path.
 *
 * Plain pointers and sizes only.  All array pointers are DEVICE pointers;
 * every call is asynchronous on the given HIP stream (hipStream_t passed as
 * void*; NULL = default stream) and returns 0 on success or a negative
 * RBX_E* code; rbx_last_error_string() describes the last failure of the
 * calling thread.  The library never allocates device memory: callers pass
 * workspaces sized by the *_workspace_size queries.
 *
 * Layouts (all row-major, float32 unless stated):
 *   coords, velocity  (n, stride) - component c of star i at [i*stride + c]
 *   flux              (n_met, n_age, n_wave)      rubix/spectra/ssp/grid.py:31
 *   datacube          (nb, nb, n_target): spaxel id p = x + nb*y is row p,
 *                     lambda contiguous            rubix/spectra/ifu.py:286-287
 *
 * Each entry point names the reference function it replaces (paths relative
 * to the AstroAI-Lab/rubix tree).
 */
#ifndef RUBIX_AMD_H
#define RUBIX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RBX_OK 0
#define RBX_EINVAL (-1)    /* bad argument */
#define RBX_EHIP (-2)      /* HIP runtime error (launch / memset) */
#define RBX_EWORKSPACE (-3) /* workspace too small or misaligned */

/* Version (major*10000 + minor*100 + patch) and diagnostics. */
int rbx_version(void);
const char* rbx_last_error_string(void);

/* ---------------------------------------------------------------- a4 / a6
 * Spaxel assignment.  Replaces square_spaxel_assignment
 * (rubix/telescope/utils.py:77-152) as called by the spaxel_assignment step
 * (rubix/core/telescope.py:108-124):
 *   ix = clip(searchsorted(edges, x, 'right') - 1, 0, nbins - 1), same for y,
 *   out = ix + nbins * iy, nbins = n_edges - 1; NaN sorts after every edge.
 * Coordinates x, y are components 0 and 1 of each star's row.
 */
int rbx_spaxel_assign(const float* coords, int64_t n, int64_t stride,
                      const float* edges, int32_t n_edges, int32_t* out_ids,
                      void* stream);

/* Stable sort-by-spaxel index (new; the reference analogue is the lexsort +
 * searchsorted at rubix/spectra/dust/dust_extinction.py:237-267).
 *   perm     (n)          stable argsort of ids (ids outside [0, n_keys) last,
 *                         in input order)
 *   sorted   (n)          ids[perm] (may be NULL)
 *   offsets  (n_keys + 1) offsets[s] = first sorted position with id >= s,
 *                         offsets[n_keys] = number of in-range ids.
 * Deterministic (no floating point; integer atomics only on per-tile
 * counters owned by one wavefront).
 */
size_t rbx_spaxel_sort_workspace_size(int64_t n, int32_t n_keys);
int rbx_spaxel_sort(const int32_t* ids, int64_t n, int32_t n_keys,
                    int32_t* perm, int32_t* sorted, int32_t* offsets,
                    void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------- SSP grid
 * Template grid + the tables the fused kernels use.  wave_obs is the SSP
 * wavelength already shifted by the cosmological (1+z) factor, in float32,
 * as rubix/core/ifu.py:261 computes it.  log_axes != 0 means ax_met/ax_age
 * hold log10 values and queries are log10'd before interpolation.
 * rbx_ssp_prepare fills cum (double, n_met*n_age*n_wave: per-template
 * prefix sums of flux[k]*(wave_obs[k]-wave_obs[k-1]), 0 at k = 0) and the
 * knot locator: n_locate entries of 4 int32 followed by the n_locate k
 * values again as a compact int32 array (5 n_locate int32 in all).  With
 * h = 1/locate_inv_h, entry
 * m >= 1 describes the cell starting at locate_x0 + (m-1) h and entry 0 every
 * query below locate_x0: {k, bits of wave_obs[b], wave_obs[b+1],
 * wave_obs[b+2] as float} with k = (number of wave_obs <= cell start) - 1
 * (-1 for entry 0) and b = min(max(k, -1), n_wave-2); knots outside
 * [0, n_wave) read as -inf / +inf.  When h is at most the smallest knot
 * spacing every cell holds at most one knot and locate_multi = 0; otherwise
 * locate_multi = 1 and the kernels step further from the entry.
 */
typedef struct rbx_ssp {
  const float* flux;
  const float* ax_met;
  const float* ax_age;
  const float* wave_obs;
  const double* cum;
  const int32_t* locate;
  int32_t n_met, n_age, n_wave, n_locate;
  float locate_x0, locate_inv_h;
  int32_t log_axes;
  int32_t locate_multi;
} rbx_ssp;

int rbx_ssp_prepare(const rbx_ssp* ssp, double* cum_out, int32_t* locate_out,
                    void* stream);

/* Telescope wavelength axis (rubix/core/ifu.py:254-255).  dwave is
 * calculate_diff(wave) (first element 0, rubix/spectra/ifu.py:83-102);
 * pairs holds (wave[j], dwave[j]) interleaved (2n floats, read by the fused
 * kernels); wmin/wmax are min/max of wave; ascending != 0 when wave is
 * non-decreasing (required by rbx_resample_bwd; the fused path accepts any
 * order). */
typedef struct rbx_target {
  const float* wave;
  const float* dwave;
  int32_t n;
  float wmin, wmax;
  int32_t ascending;
  const float* pairs;
} rbx_target;

/* Per-star inputs of the fused path (pointers to (n,) arrays unless noted).
 * velocity is (n, vel_stride); the line-of-sight component vel_comp is used
 * (rubix/config/rubix_config.yml:245: "z" -> 2).  ids are spaxel ids from
 * rbx_spaxel_assign; n_spaxel = nb*nb segments are kept, ids >= n_spaxel are
 * dropped as segment_sum does (rubix/spectra/ifu.py:286); n_keys bounds the
 * ids (nbins^2 of the edge array). */
typedef struct rbx_stars {
  int64_t n;
  const float* mass;
  const float* age;
  const float* metallicity;
  const float* velocity;
  int64_t vel_stride;
  int32_t vel_comp;
  int32_t n_spaxel;
  int32_t n_keys;
  int32_t reserved;
  const int32_t* ids;
  float speed_of_light;
  float reserved_f;
} rbx_stars;

/* ---------------------------------------------------------------- fused cube
 * Forward: replaces the calculate_spectra -> scale_spectrum_by_mass ->
 * doppler_shift_and_resampling -> calculate_datacube chain
 * (rubix/core/ifu.py:63-121, 149-157, 269-293, 327-339;
 *  rubix/spectra/ifu.py:190, 223-266, 269-288; interp2d linear at
 *  rubix/spectra/ssp/grid.py:113-120) with one fused, atomic-free,
 * deterministic path.  Writes cube (n_spaxel, target.n).  The workspace
 * keeps the per-star state that rbx_cube_bwd reads: keep it untouched
 * between the two calls.  Per-star spectra are never materialised.
 * For n_keys <= 8192 the workspace also holds the chunk partials of up to 8
 * template-age bands per spaxel (the stars are reduced band by band when the
 * template grid exceeds one band's ~3 MB, which keeps each XCD's L2 on one
 * band's templates), so it grows with 8 * n_keys rows of n_target doubles.
 */
size_t rbx_cube_workspace_size(int64_t n, int32_t n_keys, int32_t n_target);
int rbx_cube_fwd(const rbx_ssp* ssp, const rbx_target* target,
                 const rbx_stars* stars, float* cube, void* workspace,
                 size_t workspace_bytes, void* stream);

/* Adjoint of rbx_cube_fwd (the reference has none:
 * rubix/core/pipeline.py:203-208).  dcube (n_spaxel, target.n) is dL/dcube;
 * writes dL/dmass, dL/dage, dL/dmetallicity (n,) and dL/dv_los (n,) - the
 * gradient w.r.t. velocity[:, vel_comp].  Any output pointer may be NULL.
 * Recomputes per-star spectra on the fly (no N x n_wave storage). */
int rbx_cube_bwd(const rbx_ssp* ssp, const rbx_target* target,
                 const rbx_stars* stars, const float* dcube, float* dmass,
                 float* dage, float* dmet, float* dvlos, void* workspace,
                 size_t workspace_bytes, void* stream);

/* Dust extinction inside the fused path (the calc_dusty_ifu chain of
 * rubix/config/pipeline_config.yml:61-131 without materialised spectra):
 * each (star, bin) value is multiplied by 10^((-0.4 axav[j]) * av[star]),
 * the factor rbx_apply_extinction applies to resampled spectra
 * (rubix/spectra/dust/dust_extinction.py:330-358).  av (n, input order) from
 * rbx_column_av, axav (target.n) from rbx_extinction_curve.  The adjoint
 * takes the same descriptor (it reads av from the forward's workspace state;
 * A_V is a constant of the differentiation).  NULL = no extinction. */
typedef struct rbx_extinction {
  const float* av;
  const float* axav;
} rbx_extinction;

int rbx_cube_fwd_ext(const rbx_ssp* ssp, const rbx_target* target,
                     const rbx_stars* stars, const rbx_extinction* ext, float* cube,
                     void* workspace, size_t workspace_bytes, void* stream);
int rbx_cube_bwd_ext(const rbx_ssp* ssp, const rbx_target* target,
                     const rbx_stars* stars, const rbx_extinction* ext,
                     const float* dcube, float* dmass, float* dage, float* dmet,
                     float* dvlos, void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------- batches
 * Many galaxies with uneven star counts and spaxel grids in one pass (the
 * reference renders one galaxy per pipeline run, rubix/core/pipeline.py;
 * this has no counterpart there).  The stars of all galaxies are
 * concatenated in one rbx_stars (n = total stars, ids = each galaxy's own
 * spaxel ids from rbx_spaxel_assign, n_spaxel = total cube rows; n_keys is
 * not used).  galaxies[g] (a device array of n_gal entries, star_begin and
 * spaxel_begin ascending) owns stars [star_begin, next star_begin or n)
 * and rows [spaxel_begin, spaxel_begin + nb*nb) of the batch cube
 * (n_spaxel rows x target.n); its ids outside [0, nb*nb) are dropped like
 * segment_sum drops them.  Each fused stage is one launch over the whole
 * batch.  rbx_batch_bwd reads the per-star state rbx_batch_fwd left in the
 * workspace and may use a different tuning. */
typedef struct rbx_galaxy {
  int64_t star_begin;
  int32_t spaxel_begin;
  int32_t nb;
} rbx_galaxy;

/* Launch tuning of the fused kernels (NULL = defaults): chunk = stars per
 * reduction chunk of the cube accumulation (16..65536, 0 = 256); waves =
 * wavefronts per workgroup of the per-star kernels (1, 2, 4, 8; 0 = 4);
 * persist = 0 for one workgroup per work item, k > 0 for persistent grids of
 * k workgroups per compute unit; flags: RBX_TUNE_TWO_PASS runs the forward
 * as star_norm + per-tile accumulation (each (star, bin) evaluated twice)
 * even where the one-pass fused kernel applies (the same values up to the
 * grouping of the cube's float64 chunk sums). */
typedef struct rbx_tuning {
  int32_t chunk;
  int32_t waves;
  int32_t persist;
  int32_t flags;
} rbx_tuning;
#define RBX_TUNE_TWO_PASS 1

size_t rbx_batch_workspace_size(int64_t n, int32_t n_spaxel, int32_t n_target,
                                const rbx_tuning* tune);
int rbx_batch_fwd(const rbx_ssp* ssp, const rbx_target* target,
                  const rbx_stars* stars, const rbx_galaxy* galaxies,
                  int32_t n_gal, const rbx_tuning* tune, float* cube,
                  void* workspace, size_t workspace_bytes, void* stream);
int rbx_batch_bwd(const rbx_ssp* ssp, const rbx_target* target,
                  const rbx_stars* stars, const rbx_galaxy* galaxies,
                  int32_t n_gal, const rbx_tuning* tune, const float* dcube,
                  float* dmass, float* dage, float* dmet, float* dvlos,
                  void* workspace, size_t workspace_bytes, void* stream);

/* One pass of the separable PSF over every galaxy of a batch cube
 * (n_rows = total rows, each row n_target long; galaxy g's rows are its
 * (nb, nb) spaxel grid, cube[y, x, l] layout as rubix/core/psf.py:56-57
 * convolves it): axis 1 convolves along x, axis 0 along y,
 * out[.., u, ..] = sum_t taps[t] in[.., u + start - t, ..] with zero padding
 * inside each galaxy; a row that belongs to no galaxy (a table whose
 * galaxies are not packed row after row from row 0) is written as zeros.
 * Packed tables of up to 512 galaxies take an LDS-tiled path with the same
 * values.  rbx_conv_batch_dot returns
 * out[0] = sum a * conv(b) (double taps, double accumulation, fixed order);
 * its workspace is rbx_dot_workspace_size(n_rows * n_target). */
int rbx_conv_batch(const float* in, float* out, const rbx_galaxy* galaxies,
                   int32_t n_gal, int32_t n_rows, int32_t n_target, int32_t axis,
                   const float* taps, int32_t n_taps, int32_t start, void* stream);
int rbx_conv_batch_dot(const float* a, const float* b, const rbx_galaxy* galaxies,
                       int32_t n_gal, int32_t n_rows, int32_t n_target, int32_t axis,
                       const double* taps, int32_t n_taps, int32_t start,
                       double* out, void* workspace, size_t workspace_bytes,
                       void* stream);

/* ---------------------------------------------------------------- per step
 * One kernel per rubix.core step, materialising the same per-star arrays the
 * reference chain stores (the "plain" path; HBM-bound by construction).
 *
 * calculate_spectra (rubix/core/ifu.py:63-121 with the interp2d linear lookup
 * of rubix/spectra/ssp/grid.py:113-120): spectra (n, n_wave) of star i at
 * (metallicity[i], age[i]); 0 outside the grid.  Needs flux and axes only. */
int rbx_ssp_lookup(const rbx_ssp* ssp, const float* met, const float* age, int64_t n,
                   float* spectra, void* stream);
/* Adjoint: dmet[i], dage[i] from dspectra (n, n_wave); 0 outside the grid. */
int rbx_ssp_lookup_bwd(const rbx_ssp* ssp, const float* met, const float* age, int64_t n,
                       const float* dspectra, float* dmet, float* dage, void* stream);

/* scale_spectrum_by_mass (rubix/core/ifu.py:149-157): out[i, k] = in[i, k] *
 * scale[i] (out may alias in).  Its adjoint is the same call with the
 * incoming gradient, plus rbx_row_dot for d/dscale. */
int rbx_scale_rows(const float* in, const float* scale, int64_t n, int32_t n_wave, float* out,
                   void* stream);
/* out[i] = sum_k a[i, k] * b[i, k] (fixed-order reduction). */
int rbx_row_dot(const float* a, const float* b, int64_t n, int32_t n_wave, float* out,
                void* stream);

/* doppler_shift_and_resampling (rubix/core/ifu.py:269-293 with
 * velocity_doppler_shift rubix/spectra/ifu.py:190 and resample_spectrum
 * rubix/spectra/ifu.py:223-266): spectra (n, ssp->n_wave) on the knots
 * wave_obs * exp(v / c) -> out (n, target->n).  Uses ssp->wave_obs and the
 * locator table only. */
int rbx_resample(const rbx_ssp* ssp, const rbx_target* target, const float* spectra,
                 const float* velocity, int64_t vel_stride, int32_t vel_comp,
                 float speed_of_light, int64_t n, float* out, void* stream);
/* Adjoint: dspectra (n, ssp->n_wave) and dvlos (n) from dout (n, target->n).
 * Target wavelengths must be increasing; 10*n_target + 2*n_wave bytes of LDS
 * per workgroup (<= 64 KiB). */
int rbx_resample_bwd(const rbx_ssp* ssp, const rbx_target* target, const float* spectra,
                     const float* velocity, int64_t vel_stride, int32_t vel_comp,
                     float speed_of_light, int64_t n, const float* dout, float* dspectra,
                     float* dvlos, void* stream);

/* calculate_datacube (rubix/spectra/ifu.py:269-288, segment_sum into
 * n_spaxel segments, ids outside [0, n_spaxel) dropped): deterministic
 * sort-by-spaxel + chunked fixed-order sums.  cube (n_spaxel, n_wave). */
size_t rbx_segment_cube_workspace_size(int64_t n, int32_t n_keys, int32_t n_wave);
int rbx_segment_cube(const float* spectra, const int32_t* ids, int64_t n, int32_t n_spaxel,
                     int32_t n_keys, int32_t n_wave, float* cube, void* workspace,
                     size_t workspace_bytes, void* stream);
/* Adjoint (gather): dspectra[i] = dcube[ids[i]] (0 for dropped ids). */
int rbx_segment_cube_bwd(const float* dcube, const int32_t* ids, int64_t n, int32_t n_spaxel,
                         int32_t n_wave, float* dspectra, void* stream);

/* ---------------------------------------------------------------- PSF / LSF
 * 1D convolution along the middle axis of an (n_outer, n_axis, n_inner)
 * array with zero padding and a crop start:
 *   out[o, x, i] = sum_t taps[t] * in[o, x + start - t, i].
 * With start = (len-1)//2 this is one axis of convolve2d(mode="same")
 * (rubix/telescope/psf/psf.py:9-11); with start = extend_factor it is the
 * LSF full-convolution + crop (rubix/telescope/lsf/lsf.py:29-65).  The
 * adjoint is the same call with reversed taps and start' = len-1-start.
 */
int rbx_conv_axis(const float* in, float* out, int64_t n_outer, int32_t n_axis,
                  int64_t n_inner, const float* taps, int32_t n_taps,
                  int32_t start, void* stream);

/* Separable PSF: out = conv_y(conv_x(in)) over a (ny, nx, n_lambda) cube;
 * tmp (same size as in) receives conv_x(in).  Replaces apply_psf
 * (rubix/telescope/psf/psf.py:33-58) for kernels K = outer(gy, gx), which
 * is exact for gaussian_kernel_2d (rubix/telescope/psf/kernels.py:5-31). */
int rbx_psf_sep(const float* in, float* out, float* tmp, int32_t ny, int32_t nx,
                int32_t nl, const float* gy, int32_t my, int32_t start_y,
                const float* gx, int32_t mx, int32_t start_x, void* stream);

/* General (non-separable) 2D kernel, per lambda plane, mode "same":
 *   out[y, x, l] = sum_ab K[a, b] * in[y + sy - a, x + sx - b, l]. */
int rbx_conv2d_planes(const float* in, float* out, int32_t ny, int32_t nx,
                      int32_t nl, const float* kernel, int32_t ky, int32_t kx,
                      int32_t sy, int32_t sx, void* stream);

/* dot = sum(a * conv_axis(b)) with double taps, a double convolution and a
 * fixed-order double reduction (used for dL/dsigma of the PSF and LSF, whose
 * derivative taps sum to zero).  workspace: rbx_dot_workspace_size(total)
 * bytes; result: one device double. */
size_t rbx_dot_workspace_size(int64_t total);
int rbx_conv_axis_dot(const float* a, const float* b, int64_t n_outer,
                      int32_t n_axis, int64_t n_inner, const double* taps,
                      int32_t n_taps, int32_t start, double* result,
                      void* workspace, size_t workspace_bytes, void* stream);

/* rbx_conv_axis_dot without a workspace argument and without allocating:
 * no counterpart in the reference; for callers that cannot allocate (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * The block partials (at most 4096 doubles, 32 KiB, for any input size; a
 * larger count is rejected with RBX_EINVAL) live in a file-scope device array
 * that each device keeps until the process ends; its address is looked up
 * on the first call on a device.  Every call on one device shares that
 * array: calls that overlap on one device race (serialise them on one
 * stream).  Same kernels and result as rbx_conv_axis_dot. */
int rbx_conv_axis_dot_noalloc(const float* a, const float* b, int64_t n_outer,
                              int32_t n_axis, int64_t n_inner, const double* taps,
                              int32_t n_taps, int32_t start, double* result,
                              void* stream);

/* ---------------------------------------------------------------- alignment
 * The steps before spaxel assignment in the reference chain
 * (rubix/config/pipeline_config.yml: rotate_galaxy -> filter_particles).
 *
 * rbx_filter_particles replaces mask_particles_outside_aperture +
 * filter_particles (rubix/telescope/utils.py:155-178,
 * rubix/core/telescope.py:129-201): mask = coords x and y within [lo, hi]
 * (the min / max spatial bin edge, inclusive); each non-NULL *_out gets the
 * field where the mask holds and 0 elsewhere (may alias its input); mask_out
 * (uint8, may be NULL) the mask.  rbx_apply_mask: out = mask ? in : 0 over
 * (n, width) rows (the adjoint of the filter).
 *
 * rbx_inertia_tensor replaces moment_of_inertia_tensor
 * (rubix/galaxy/alignment.py:66-120): out9 (device, row-major 3x3 double)
 * = sum over particles with sqrt(x^2+y^2+z^2) <= halfmass_radius (float32
 * test) of m (|r|^2 delta_ij - r_i r_j), fixed-order double reduction;
 * workspace rbx_inertia_workspace_size(n) bytes.
 *
 * rbx_rotate_rows replaces apply_init_rotation + apply_rotation
 * (alignment.py:123-137,208-225): rows (n, stride >= 3; columns past 2 are
 * copied) out = (in first) second with row-major 3x3 float matrices, each
 * product in float32 without fused multiply-adds. */
typedef struct rbx_rotation {
  float first[9];
  float second[9];
} rbx_rotation;

int rbx_filter_particles(const float* coords, int64_t n, int64_t stride,
                         float lo, float hi, const float* mass, const float* age,
                         const float* metallicity, float* mass_out, float* age_out,
                         float* metallicity_out, uint8_t* mask_out, void* stream);
int rbx_apply_mask(const uint8_t* mask, int64_t n, int64_t width,
                   const float* in, float* out, void* stream);
size_t rbx_inertia_workspace_size(int64_t n);
int rbx_inertia_tensor(const float* coords, const float* mass, int64_t n,
                       int64_t stride, float halfmass_radius, double* out9,
                       void* workspace, size_t workspace_bytes, void* stream);
int rbx_rotate_rows(const float* in, int64_t n, int64_t stride,
                    const rbx_rotation* rot, float* out, void* stream);

/* ---------------------------------------------------------------- noise
 * The apply_noise step after the LSF (rubix/core/noise.py:17-80 over
 * rubix/telescope/noise/noise.py:8-115), on a datacube of n_spaxel rows x
 * n_target bins (row-major, the reference's (n_x, n_y, n_wave) cube):
 *   flux[p] = sum_l cube[p, l] (double sum, rounded once);
 *   M       = median of flux if every flux[p] > 0, else 0 (jnp.median of
 *             where(flux > 0, flux, NaN) propagates the NaN; nan_to_num);
 *   s2n[p]  = flux[p] > 0 ? (sqrt(M) / signal_to_noise[p]) / sqrt(flux[p]) : 0;
 *   out     = cube + cube * (z * s2n[p]), each operation rounded in float32.
 * z[i] at flat index i: jax.random's Threefry-2x32 construction for
 * PRNGKey(seed) (the reference always uses seed 0): normal (sqrt(2) erfinv of
 * a uniform on (-1, 1)) or uniform on [0, 1).  The medians' ranks are
 * selected exactly (radix select); stats gets M and the spaxels at the two
 * midpoint ranks (stable order), used by the adjoint.  out may alias cube.
 *
 * rbx_noise_sample replaces sample_noise (noise.py:8-32): out[i] = z[i].
 * rbx_apply_noise_bwd: dcube = dout + dout * z * s2n + dL/dflux[p], with
 * dL/dflux from d s2n / d flux and d M / d flux at the median spaxels
 * (zero when M = 0); workspace rbx_noise_workspace_size(n_spaxel) bytes.
 * n_spaxel <= 2^31 - 1; the S2N map is one workgroup's work. */
#define RBX_NOISE_NORMAL 0
#define RBX_NOISE_UNIFORM 1

typedef struct rbx_noise_stats {
  float median;
  int32_t median_lo;      /* spaxel at rank floor((n-1)/2), -1 if M = 0 */
  int32_t median_hi;      /* spaxel at rank ceil((n-1)/2), -1 if M = 0 */
  int32_t any_nonpositive;
} rbx_noise_stats;

int rbx_noise_sample(int64_t n, uint64_t seed, int32_t distribution, float* out,
                     void* stream);
size_t rbx_noise_workspace_size(int32_t n_spaxel);
int rbx_apply_noise(const float* cube, int32_t n_spaxel, int32_t n_target,
                    const float* signal_to_noise, uint64_t seed, int32_t distribution,
                    float* flux, float* s2n, rbx_noise_stats* stats, float* out,
                    void* stream);
int rbx_apply_noise_bwd(const float* cube, const float* dout, int32_t n_spaxel,
                        int32_t n_target, const float* flux, const float* s2n,
                        const rbx_noise_stats* stats, uint64_t seed,
                        int32_t distribution, void* workspace, size_t workspace_bytes,
                        float* dcube, void* stream);

/* ---------------------------------------------------------------- dust
 * The calculate_extinction step of calc_dusty_ifu (rubix/core/dust.py:16-65
 * over rubix/spectra/dust/dust_extinction.py:163-358).
 *
 * rbx_extinction_curve: axav[l] = A(x)/A(V) of Cardelli89
 * (extinction_models.py:104-178) or Gordon23 (:255-445) at x = wave[l] / 1e4,
 * the argument the reference's step passes (microns; its Cardelli89 branches
 * treat the argument as a wavenumber).
 *
 * rbx_gas_extinction: per gas cell (metals rows of metals_stride >= 5
 * columns, H in column 0, O in column 4): log(O/H) + 12 = 12 +
 * log10(O / (16 H)), the Remy-Ruyer et al. (2014) dust-to-gas ratio
 * (dust_extinction.py:13-90), dust mass = mass * D/G and
 * ext = 3 pi dust C / (0.4 ln 10 * 5448e-8 * grain_density) / spaxel_area with
 * C = MSUN_TO_GRAMS / KPC_TO_CM^2 (:93-160, 269-287).
 *
 * rbx_column_av: A_V per star (dust_extinction.py:237-327): for the star's
 * spaxel s (stars with ids outside [0, n_spaxels) get 0), jnp.interp with
 * left = "extrapolate" at the star's z over every gas cell, keyed z for
 * cells of s and fl(z * 1e30) for the others, valued the inclusive
 * cumulative extinction of s's cells in z order (0 for the others), in the
 * stable order of a lexsort by (pixel, z).  Coordinates are rows of
 * *_stride >= 3 floats (z in column 2); gas pixel ids come from
 * rbx_spaxel_assign, sorted over n_keys buckets (ids >= n_keys after them).
 * If any gas extinction is not finite, every A_V is NaN.
 * Workspace: rbx_column_av_workspace_size(n_gas, n_keys) bytes.
 *
 * rbx_apply_extinction: out[i, l] = in[i, l] * 10^((-0.4 axav[l]) * av[i])
 * over (n, n_wave) rows; the same call is its own adjoint; out may alias
 * in. */
#define RBX_EXT_CARDELLI89 0
#define RBX_EXT_GORDON23 1
#define RBX_DTG_POWER_LAW_FREE 0  /* "power law slope free" */
#define RBX_DTG_BROKEN_POWER_LAW 1 /* "broken power law fit" */
#define RBX_XCO_MW 0
#define RBX_XCO_Z 1

int rbx_extinction_curve(int32_t model, float rv, const float* wave, int32_t n,
                         float* axav, void* stream);
int rbx_gas_extinction(const float* mass, const float* metals, int64_t n,
                       int64_t metals_stride, int32_t dtg_model, int32_t xco,
                       float grain_density, float spaxel_area, float* ext, void* stream);
size_t rbx_column_av_workspace_size(int64_t n_gas, int32_t n_keys);
int rbx_column_av(const float* gas_coords, int64_t gas_stride, const int32_t* gas_pix,
                  const float* gas_ext, int64_t n_gas, const float* star_coords,
                  int64_t star_stride, const int32_t* star_pix, int64_t n_star,
                  int32_t n_spaxels, int32_t n_keys, void* workspace,
                  size_t workspace_bytes, float* av, void* stream);
int rbx_apply_extinction(const float* in, const float* av, const float* axav, int64_t n,
                         int32_t n_wave, float* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* RUBIX_AMD_H */
