/*
 * lensim.h -- C ABI of the MI355X (gfx950) strong-lensing image-simulation
 * library (liblensim.so).
 *
 * One shared library, plain pointers and sizes, no torch types.  Every
 * buffer argument named d_* is DEVICE memory on the current HIP device;
 * every call is asynchronous on the given HIP stream (NULL = default stream)
 * and allocates no device memory: scratch memory comes in through `ws`
 * (query its size with lensim_workspace_size).  The graph entry points
 * create HIP graph objects.  All entry points return 0 on success or
 * a negative lensim_status; lensim_last_error() returns a message for the
 * last failure on the calling thread.
 *
 * Parameter blocks.  A model is an ordered list of profiles: n_lens
 * deflectors, then n_bulk further deflectors whose types come from a device
 * array (many-deflector models, e.g. substructure), then n_source
 * source-light profiles, then n_lens_light lens-light profiles.  Parameters
 * of a batch of n_img models are one float64 array of shape
 * [n_img][n_lens + n_bulk + n_source + n_lens_light][LENSIM_NPARAM]; each
 * profile slot holds that profile's parameters in the order of the
 * reference's `param_names` (unused trailing slots are ignored):
 *   EPL            theta_E gamma e1 e2 center_x center_y    (LensModel/Profiles/epl.py:50)
 *   SIE            theta_E e1 e2 center_x center_y           (LensModel/Profiles/sie.py:40)
 *   NIE            theta_E e1 e2 s_scale center_x center_y   (LensModel/Profiles/nie.py)
 *   NFW            Rs alpha_Rs center_x center_y             (LensModel/Profiles/nfw.py:48)
 *   TNFW           Rs alpha_Rs r_trunc center_x center_y     (LensModel/Profiles/tnfw.py:27)
 *   SHEAR          gamma1 gamma2 ra_0 dec_0                  (LensModel/Profiles/shear.py:16)
 *   CONVERGENCE    kappa ra_0 dec_0                          (LensModel/Profiles/convergence.py:15)
 *   SERSIC_ELLIPSE amp R_sersic n_sersic e1 e2 center_x center_y
 *                                                 (LightModel/Profiles/sersic_ellipse.py:25)
 *   SERSIC         amp R_sersic n_sersic center_x center_y  (LightModel/Profiles/sersic.py)
 *   GAUSSIAN       amp sigma center_x center_y               (LightModel/Profiles/gaussian.py)
 */
#ifndef LENSIM_H
#define LENSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LENSIM_ABI_VERSION 6
#define LENSIM_NPARAM 8        /* float64 slots per profile in a parameter block */
#define LENSIM_MAX_PROFILES 64 /* lens + source + lens-light profiles per model */
#define LENSIM_EPL_NMAX 500    /* largest series length of the EPL hypergeometric sum */

typedef enum {
  LENSIM_OK = 0,
  LENSIM_EINVAL = -1,        /* bad argument (shape, pointer, size) */
  LENSIM_EUNSUPPORTED = -2,  /* profile / option not supported */
  LENSIM_EDEVICE = -3,       /* HIP error */
  LENSIM_EWORKSPACE = -4     /* workspace too small */
} lensim_status;

typedef enum {
  LENSIM_FP64 = 0,  /* plain FP64 form: every step in float64 */
  LENSIM_MIXED = 1  /* float32 profiles and convolution, float64 coordinates and chi^2 sums */
} lensim_precision;

typedef enum {
  LENSIM_NULL = 0,
  /* deflectors */
  LENSIM_EPL = 1,
  LENSIM_SIE = 2,
  LENSIM_NIE = 3,
  LENSIM_NFW = 4,
  LENSIM_TNFW = 5,
  LENSIM_SHEAR = 6,
  LENSIM_CONVERGENCE = 7, /* mass sheet */
  /* light profiles */
  LENSIM_SERSIC = 32,
  LENSIM_SERSIC_ELLIPSE = 33,
  LENSIM_GAUSSIAN = 34
} lensim_profile;

typedef enum {
  LENSIM_PSF_NONE = 0,
  LENSIM_PSF_PIXEL = 1,    /* zero-padded 'same' convolution (ImSim/Numerics/convolution.py:74-86) */
  LENSIM_PSF_GAUSSIAN = 2  /* edge-padded 'same' convolution (ImSim/Numerics/convolution.py:217-236) */
} lensim_psf_type;

/* Flags of lensim_image */
#define LENSIM_IMAGE_SOURCE      1u  /* add the lensed source (source_add) */
#define LENSIM_IMAGE_LENS_LIGHT  2u  /* add the lens light (lens_light_add) */
#define LENSIM_IMAGE_UNCONVOLVED 4u  /* skip the PSF (unconvolved=True) */

/* Multi-plane ray tracing through redshift-ordered deflectors
 * (MultiPlaneBulk.ray_shooting, LensModel/MultiPlane/multi_plane_bulk.py:152-212):
 * starting from x = 0 with the ray angle alpha = theta, for every plane p
 *   x += alpha * T_ij[p];  alpha -= alpha_red_p(x / T_z[p]) * r2p[p]
 * where the last plane is the source (no deflector), and beta = x / T_z[last].
 * The arrays come from MultiPlaneBulk.prepare_ray_shooting_kwargs
 * (multi_plane_bulk.py:54-150). */
typedef struct {
  int32_t n_planes;          /* deflector planes + 1 (the source) */
  int32_t reserved;
  const int32_t *d_order;    /* [n_planes - 1] deflector index (inline lens, then bulk) of each plane */
  const double *d_T_ij;      /* [n_img][n_planes] transverse comoving distance from the previous plane */
  const double *d_T_z;       /* [n_img][n_planes] transverse comoving distance from z = 0 */
  const double *d_r2p;       /* [n_img][n_planes] reduced-to-physical deflection factor */
  int64_t stride;            /* elements between images in the three arrays (0: shared) */
} lensim_multiplane;

/* Model description: the profile lists of LensModel(lens_model_list),
 * LightModel(source list) and LightModel(lens light list)
 * (LensModel/lens_model.py:25-216, LightModel/light_model_base.py:63-228),
 * plus an optional bulk deflector list as packed by
 * SinglePlaneBulk.prepare_ray_shooting_kwargs (LensModel/single_plane_bulk.py:47-106):
 * bulk types may be SIE, NIE, NFW, TNFW, SHEAR, CONVERGENCE or LENSIM_NULL (padding). */
typedef struct {
  int32_t n_lens;
  int32_t n_source;
  int32_t n_lens_light;
  int32_t types[LENSIM_MAX_PROFILES]; /* lensim_profile, lens then source then lens light */
  double alpha_scaling;               /* SinglePlane alpha_scaling (LensModel/single_plane.py:139) */
  double flux_scaling;                /* ImageData flux_scaling (ImSim/image_model.py:60-63) */
  int32_t n_bulk;                     /* bulk deflectors after the n_lens inline ones */
  int32_t bulk_type;                  /* their common lensim_profile, or LENSIM_NULL if mixed */
  const int32_t *d_bulk_types;        /* [n_bulk] device array (read when bulk_type == LENSIM_NULL) */
  const lensim_multiplane *multiplane; /* NULL: single plane; else every deflector sits on its own plane */
} lensim_model;

/* Pixel grid + supersampling (ImSim/Numerics/grid.py:22-64): pixel (i, j)
 * (column i fastest) has centre ra = ra_at_xy_0 + M00 i + M01 j,
 * dec = dec_at_xy_0 + M10 i + M11 j; images are [ny][nx].
 * d_eval_mask (RegularGrid flux_evaluate_indexes, grid.py:47-58): when not
 * NULL, only pixels with a non-zero mask byte shoot rays; the others have
 * zero flux (before the convolution). */
typedef struct {
  int32_t nx;
  int32_t ny;
  int32_t supersampling;
  int32_t reserved;
  double ra_at_xy_0;
  double dec_at_xy_0;
  double pix2angle[4];         /* M00 M01 M10 M11 */
  const uint8_t *d_eval_mask;  /* [ny][nx] device bytes, or NULL for every pixel */
} lensim_grid;

/* PSF convolution (ImSim/Numerics/numerics.py:102-146).
 * supersampled = 0: d_kernel is sampled on the image pixels and convolves the
 *   s x s mean of the rays (PixelKernelConvolution / GaussianConvolution,
 *   ImSim/Numerics/convolution.py:74-96, :217-236).
 * supersampled = 1 (supersampling_convolution=True, numerics.py:103-116):
 *   d_kernel is sampled on the supersampled grid and convolves the
 *   (ny s) x (nx s) ray fluxes before the s x s mean
 *   (SubgridKernelConvolution.re_size_convolve, convolution.py:154-166, or
 *   GaussianConvolution.re_size_convolve, :238-252); an optional low-resolution
 *   kernel (kx_low x ky_low, zero padding) then convolves the s x s mean of the
 *   unconvolved rays and is added (the split kernel of convolution.py:117-131).
 *   Ignored when the grid's supersampling is 1. */
typedef struct {
  int32_t type;               /* lensim_psf_type */
  int32_t kx;                 /* odd kernel width */
  int32_t ky;                 /* odd kernel height */
  int32_t supersampled;       /* 0 or 1, see above */
  const double *d_kernel;     /* [ky][kx], float64 device memory */
  int32_t kx_low;             /* odd width of the low-resolution split kernel, or 0 for none */
  int32_t ky_low;             /* odd height, or 0 */
  const double *d_kernel_low; /* [ky_low][kx_low] float64 device memory, or NULL */
} lensim_psf;

/* Data and noise of one image for the likelihood
 * (Data/imaging_data.py:159-182, Data/image_noise.py:100-137). */
typedef struct {
  const double *d_data;         /* [ny][nx] observed image */
  const double *d_noise_map;    /* [ny][nx] 1-sigma noise, or NULL */
  const double *d_exposure_map; /* [ny][nx] exposure time, or NULL to use exposure_time */
  const double *d_mask;         /* [ny][nx] likelihood mask (0/1), or NULL for all ones */
  double exposure_time;         /* used when d_exposure_map is NULL */
  double background_rms;
  int64_t image_stride;         /* elements between the data (and noise map) of consecutive
                                   batch images; 0: one observation for the whole batch.
                                   The exposure map and mask are shared by the batch. */
} lensim_data;

/* Library version and a message for the last failure on this thread. */
int lensim_abi_version(void);
const char *lensim_last_error(void);

/* Bytes of scratch memory the calls below need for a batch of n_img models
 * (grid and psf may be NULL for lensim_raytrace / lensim_surface_brightness;
 * psf may be NULL for an unconvolved image).  Returns 0 on a bad argument. */
size_t lensim_workspace_size(const lensim_model *model, const lensim_grid *grid, const lensim_psf *psf,
                             int32_t n_img, int32_t precision);

/* Ray shooting: beta = theta - alpha_scaling * sum_k alpha_k(theta) for every
 * image of the batch at the same n_ray coordinates.
 * Replaces LensModel.ray_shooting (LensModel/lens_model.py:236-252) ->
 * SinglePlane.ray_shooting/alpha (LensModel/single_plane.py:38-56, :112-139).
 * d_x, d_y: [n_ray]; d_beta_x, d_beta_y: [n_img][n_ray] float64. */
int lensim_raytrace(const lensim_model *model, const double *d_params, int32_t n_img, const double *d_x,
                    const double *d_y, int64_t n_ray, int32_t precision, double *d_beta_x, double *d_beta_y,
                    void *ws, size_t ws_bytes, void *stream);

/* Surface brightness of the model's source-light profiles (which = 0) or
 * lens-light profiles (which = 1) at n_ray coordinates.
 * Replaces LightModel.surface_brightness (LightModel/light_model_base.py:230-252).
 * d_flux: [n_img][n_ray] float64. */
int lensim_surface_brightness(const lensim_model *model, const double *d_params, int32_t n_img,
                              int32_t which, const double *d_x, const double *d_y, int64_t n_ray,
                              int32_t precision, double *d_flux, void *ws, size_t ws_bytes, void *stream);

/* Model images of a batch: ray shooting + source light + lens light on the
 * supersampled grid, s x s mean, PSF convolution, x pixel_width^2 x flux_scaling.
 * Replaces ImageModel.image (ImSim/image_model.py:431-492) with its
 * Numerics.re_size_convolve (ImSim/Numerics/numerics.py:158-177).
 * d_out: [n_img][ny][nx], float64 for LENSIM_FP64 and float32 for LENSIM_MIXED. */
int lensim_image(const lensim_model *model, const lensim_grid *grid, const lensim_psf *psf,
                 const double *d_params, int32_t n_img, int32_t precision, uint32_t flags, void *d_out,
                 void *ws, size_t ws_bytes, void *stream);

/* log-likelihood of every model of a batch against one data image:
 * -1/2 sum mask (model - data)^2 / (C_D(model) + |error|).
 * Replaces ImageModel.likelihood_data_given_model (ImSim/image_model.py:162-221)
 * -> ImageData._log_likelihood (Data/imaging_data.py:159-182).
 * d_logl: [n_img] float64.  d_model_out: optional [n_img][ny][nx] model images
 * (dtype as lensim_image) or NULL. */
int lensim_loglike(const lensim_model *model, const lensim_grid *grid, const lensim_psf *psf,
                   const lensim_data *data, const double *d_params, int32_t n_img, int32_t precision,
                   double *d_logl, void *d_model_out, void *ws, size_t ws_bytes, void *stream);

/* log-likelihood of given model images against the data:
 * -1/2 sum mask (model - data)^2 / (C_D(model) + |error|), as lensim_loglike
 * computes it for its own models.
 * Replaces ImageData.log_likelihood (Data/imaging_data.py:159-182).
 * d_model: [n_img][ny][nx], float64 (precision LENSIM_FP64) or float32
 * (LENSIM_MIXED); the sums run in float64.  Needs ws of
 * lensim_data_loglike_workspace_size bytes. */
size_t lensim_data_loglike_workspace_size(int32_t nx, int32_t ny, int32_t n_img);
int lensim_data_loglike(const lensim_data *data, int32_t nx, int32_t ny, const void *d_model, int32_t n_img,
                        int32_t precision, double *d_logl, void *ws, size_t ws_bytes, void *stream);

/* Linear amplitude solve: the amplitude of every source and lens-light
 * profile (SERSIC, SERSIC_ELLIPSE, GAUSSIAN 'amp', in profile order) by
 * weighted least squares against the data, then the likelihood of the solved
 * model.  Replaces ImageLinearFit.image_linear_solve and
 * likelihood_data_given_model (ImSim/image_linear_solve.py:98-181, :229-340)
 * with linear_response_matrix (:362-455) and de_lens.get_param_WLS /
 * marginalization_new (ImSim/de_lens.py:10-77):
 *   A_l = the model image of light profile l alone at amp = 1 (x flux_scaling),
 *   w = mask / C_D(data) (the data-based variance, Data.C_D),
 *   M = A W A^T, R = A W d, amp = M^-1 R if cond(M) < 5 / eps else 0,
 *   model = sum_l amp_l A_l inside the mask, 0 outside,
 *   logL = -1/2 sum mask (model - d)^2 / C_D(model)
 *          [+ marginalisation constant of M^-1 when source_marg].
 * linear_prior <= 0 means none.  Outputs (device, per image of the batch):
 * d_amp [n_img][n_lin] or NULL, d_logl [n_img], d_model_out [n_img][ny][nx]
 * or NULL, d_cov [n_img][n_lin][n_lin] (M^-1, zero when unstable) or NULL,
 * d_response_out [n_lin][n_img][ny][nx] (the A_l) or NULL; images in the
 * compute type (float64 / float32).  n_lin = n_source + n_lens_light <= 8. */
size_t lensim_linear_workspace_size(const lensim_model *model, const lensim_grid *grid, const lensim_psf *psf,
                                    int32_t n_img, int32_t precision);
int lensim_linear_solve(const lensim_model *model, const lensim_grid *grid, const lensim_psf *psf,
                        const lensim_data *data, const double *d_params, int32_t n_img, int32_t precision,
                        int32_t source_marg, double linear_prior, double *d_amp, double *d_logl, void *d_model_out,
                        double *d_cov, void *d_response_out, void *ws, size_t ws_bytes, void *stream);

/* log-likelihood and its gradient with respect to every kwarg of every
 * profile, for a batch: d_grad is [n_img][n_profiles][LENSIM_NPARAM] in the
 * parameter-block layout (slots beyond a profile's kwargs are zero).
 * Replaces jax.grad of ImageModel.likelihood_data_given_model
 * (ImSim/image_model.py:162-221, as exercised by
 * test/test_ImSim/test_image_model.py:274-335; EPL's custom JVP epl.py:321-340).
 * Adjoint path: dlogL/dmodel (model-based variance included) -> correlation
 * with the PSF -> per-ray forward-mode derivatives of each profile.  Bulk
 * deflector lists are differentiated per deflector; a multi-plane model
 * (model->multiplane) by a reverse sweep over its planes
 * (LensModel/MultiPlane/multi_plane_bulk.py:152-212); an evaluation mask
 * (grid->d_eval_mask) restricts the rays as in lensim_image. */
size_t lensim_grad_workspace_size(const lensim_model *model, const lensim_grid *grid, const lensim_psf *psf,
                                  int32_t n_img, int32_t precision);
int lensim_loglike_grad(const lensim_model *model, const lensim_grid *grid, const lensim_psf *psf,
                        const lensim_data *data, const double *d_params, int32_t n_img, int32_t precision,
                        double *d_logl, double *d_grad, void *ws, size_t ws_bytes, void *stream);

/* lensim_loglike without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Its
 * scratch memory is a file-scope device array of lensim_noalloc_scratch_bytes()
 * (64 MiB) that every GPU the library runs on holds until the process ends;
 * a call whose lensim_workspace_size exceeds it returns LENSIM_EWORKSPACE
 * before any device work.  It calls no allocator.  Calls on one GPU share
 * the array: calls that overlap in time on one GPU (different streams or
 * threads) race and must be ordered by the caller.  Results equal
 * lensim_loglike's. */
size_t lensim_noalloc_scratch_bytes(void);
int lensim_loglike_noalloc(const lensim_model *model, const lensim_grid *grid, const lensim_psf *psf,
                           const lensim_data *data, const double *d_params, int32_t n_img, int32_t precision,
                           double *d_logl, void *d_model_out, void *stream);

/* Graph capture of a log-likelihood call, for callers that replay one batch
 * shape many times (a sampler's likelihood, a request server's lanes) and
 * want one launch per call.  Replaces, for those callers, the jit-compiled
 * and cached ImageModel.likelihood_data_given_model that the reference's
 * samplers call (Sampling/sampler.py:215-251, Sampling/likelihood.py:290-319).
 * lensim_graph_capture_loglike captures the launches lensim_loglike makes
 * with these arguments on `stream` (not the NULL stream) into a HIP graph and
 * instantiates it; nothing runs.  lensim_graph_launch replays it on a stream:
 * it reads the same device buffers (parameters, data, workspace) and writes
 * the same outputs, so a caller updates d_params (or the data) in place
 * between launches.  Launches of one graph must be ordered (they share the
 * workspace).  The model, grid, psf and data structs are read only at
 * capture.  lensim_graph_destroy frees the graph. */
typedef struct lensim_graph_s *lensim_graph;
int lensim_graph_capture_loglike(const lensim_model *model, const lensim_grid *grid, const lensim_psf *psf,
                                 const lensim_data *data, const double *d_params, int32_t n_img, int32_t precision,
                                 double *d_logl, void *d_model_out, void *ws, size_t ws_bytes, void *stream,
                                 lensim_graph *graph);
int lensim_graph_launch(lensim_graph graph, void *stream);
int lensim_graph_destroy(lensim_graph graph);

/* Kernel timing for benchmarks (no counterpart in the reference).  While
 * enabled, every lensim_image / lensim_loglike call records HIP events on its
 * own stream around its render kernel and its convolution kernel.
 * lensim_timing_read waits for the recorded events, returns the summed
 * milliseconds of each kernel and the number of calls since the last read, and
 * resets the sums.  Off by default; not thread-safe. */
int lensim_timing_enable(int on);
int lensim_timing_read(double *render_ms, double *conv_ms, int64_t *calls);

/* Bounds-checked debug build (no counterpart in the reference: JAX checks
 * nothing of the kind).  liblensim_checked.so is this library compiled with
 * -DLENSIM_BOUNDS_CHECK.  There every compute entry point registers the buffers
 * of its call (parameters, model arrays, mask, PSF, data, outputs, and the
 * workspace up to the size it needs), and every kernel access to global
 * memory is tested against them.  An access outside them is not made (a
 * store goes to a private sink, a load reads it); the entry point waits for
 * its stream and returns LENSIM_EDEVICE with the source line of the first such
 * access.  Calls under graph capture are not checked.  The registered ranges
 * are global: calls of the checked build must not overlap in time.
 * lensim_bounds_checked returns 1 in the checked build, 0 otherwise.
 * lensim_bounds_selftest registers [d_buf, d_buf + bytes) and launches a
 * kernel that stores bytes 0 .. bytes - 1 (byte i = i mod 256) and then
 * byte `bytes`, one past the range.  The checked build diverts that last
 * store and returns LENSIM_EDEVICE; the normal build launches nothing and
 * returns LENSIM_EUNSUPPORTED.  bytes is at most 2^20. */
int lensim_bounds_checked(void);
int lensim_bounds_selftest(void *d_buf, size_t bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* LENSIM_H */
