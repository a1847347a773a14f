/* Host-side checks of liblensim's C ABI that need no GPU: argument validation,
 * workspace layouts and error messages of every entry point, built together
 * with the library's host code under AddressSanitizer and
 * UndefinedBehaviorSanitizer (tests/test_sanitizers.py).  Every call here
 * returns before any device work; the program exits non-zero on the first
 * unexpected result, and the sanitizers abort on the first finding. */
#include <stdio.h>
#include <string.h>

#include "lensim.h"

static int n_fail = 0;

#define EXPECT(cond)                                                              \
  do {                                                                            \
    if (!(cond)) {                                                                \
      fprintf(stderr, "%s:%d: failed: %s (last error: %s)\n", __FILE__, __LINE__, \
              #cond, lensim_last_error());                                        \
      ++n_fail;                                                                   \
    }                                                                             \
  } while (0)

static lensim_model model_of(int n_lens, const int *lens, int n_src, const int *src, int n_light,
                             const int *light) {
  lensim_model m;
  memset(&m, 0, sizeof(m));
  m.n_lens = n_lens;
  m.n_source = n_src;
  m.n_lens_light = n_light;
  int k = 0;
  for (int i = 0; i < n_lens; ++i) m.types[k++] = lens[i];
  for (int i = 0; i < n_src; ++i) m.types[k++] = src[i];
  for (int i = 0; i < n_light; ++i) m.types[k++] = light[i];
  m.alpha_scaling = 1.0;
  m.flux_scaling = 1.0;
  return m;
}

static lensim_grid grid_of(int nx, int ny, int s) {
  lensim_grid g;
  memset(&g, 0, sizeof(g));
  g.nx = nx;
  g.ny = ny;
  g.supersampling = s;
  g.ra_at_xy_0 = -0.04 * (nx - 1) / 2.0;
  g.dec_at_xy_0 = -0.04 * (ny - 1) / 2.0;
  g.pix2angle[0] = 0.04;
  g.pix2angle[3] = 0.04;
  return g;
}

/* a non-NULL address standing in for device memory: never dereferenced here */
static double fake_device[64];

int main(void) {
  EXPECT(lensim_abi_version() == LENSIM_ABI_VERSION);
  const int lens_all[] = {LENSIM_EPL, LENSIM_SIE, LENSIM_NIE, LENSIM_NFW, LENSIM_TNFW, LENSIM_SHEAR};
  const int light_all[] = {LENSIM_SERSIC, LENSIM_SERSIC_ELLIPSE, LENSIM_GAUSSIAN};
  lensim_model m = model_of(6, lens_all, 3, light_all, 3, light_all);
  lensim_psf psf;
  memset(&psf, 0, sizeof(psf));
  psf.type = LENSIM_PSF_PIXEL;
  psf.d_kernel = fake_device;

  /* workspace sizes over shapes, PSFs, supersampling and both forms */
  const int sizes[][3] = {{1, 1, 1}, {7, 3, 1}, {60, 60, 2}, {120, 120, 2}, {181, 179, 3}, {240, 240, 5}};
  const int ks[] = {1, 3, 11, 21, 41, 61};
  for (unsigned a = 0; a < sizeof(sizes) / sizeof(sizes[0]); ++a) {
    lensim_grid g = grid_of(sizes[a][0], sizes[a][1], sizes[a][2]);
    for (unsigned b = 0; b < sizeof(ks) / sizeof(ks[0]); ++b) {
      for (int type = LENSIM_PSF_NONE; type <= LENSIM_PSF_GAUSSIAN; ++type) {
        for (int sup = 0; sup <= 1; ++sup) {
          psf.type = type;
          psf.kx = psf.ky = ks[b];
          psf.supersampled = sup;
          for (int prec = LENSIM_FP64; prec <= LENSIM_MIXED; ++prec) {
            size_t prev = 0;
            const int batches[] = {1, 2, 7, 64, 256, 1000};
            for (unsigned c = 0; c < sizeof(batches) / sizeof(batches[0]); ++c) {
              const size_t w = lensim_workspace_size(&m, &g, &psf, batches[c], prec);
              const size_t wg = lensim_grad_workspace_size(&m, &g, &psf, batches[c], prec);
              const size_t wl = lensim_linear_workspace_size(&m, &g, &psf, batches[c], prec);
              EXPECT(w > 0 && wg >= w && w >= prev && (w & 255) == 0);
              (void)wl;
              prev = w;
            }
          }
        }
      }
    }
  }
  psf.type = LENSIM_PSF_PIXEL;
  psf.kx = psf.ky = 21;
  psf.supersampled = 0;
  lensim_grid g = grid_of(120, 120, 2);
  EXPECT(lensim_data_loglike_workspace_size(120, 120, 64) > 0);
  EXPECT(lensim_data_loglike_workspace_size(0, 120, 64) == 0);

  /* bad models, grids and PSFs: size queries return 0 with a message */
  lensim_model bad = m;
  bad.types[0] = 99;
  EXPECT(lensim_workspace_size(&bad, &g, &psf, 4, 0) == 0 && strlen(lensim_last_error()) > 0);
  bad = m;
  bad.types[7] = LENSIM_EPL; /* a deflector type among the light profiles */
  EXPECT(lensim_workspace_size(&bad, &g, &psf, 4, 0) == 0);
  bad = m;
  bad.n_lens = 40;
  bad.n_source = 20;
  bad.n_lens_light = 20; /* 80 profiles */
  EXPECT(lensim_workspace_size(&bad, &g, &psf, 4, 0) == 0);
  bad = m;
  bad.n_lens = -1;
  EXPECT(lensim_workspace_size(&bad, &g, &psf, 4, 0) == 0);
  bad = m;
  bad.n_bulk = 10;
  bad.bulk_type = LENSIM_EPL; /* EPL is not a bulk type */
  EXPECT(lensim_workspace_size(&bad, &g, &psf, 4, 0) == 0);
  bad.bulk_type = LENSIM_NULL; /* mixed bulk types need the type array */
  EXPECT(lensim_workspace_size(&bad, &g, &psf, 4, 0) == 0);
  bad.n_bulk = (1 << 20) + 1;
  bad.bulk_type = LENSIM_TNFW;
  EXPECT(lensim_workspace_size(&bad, &g, &psf, 4, 0) == 0);
  lensim_model bulk = m;
  bulk.n_bulk = 1000;
  bulk.bulk_type = LENSIM_TNFW;
  EXPECT(lensim_workspace_size(&bulk, &g, &psf, 4, 0) > lensim_workspace_size(&m, &g, &psf, 4, 0));
  EXPECT(lensim_grad_workspace_size(&bulk, &g, &psf, 4, 1) > 0);
  lensim_multiplane mp;
  memset(&mp, 0, sizeof(mp));
  mp.n_planes = 3; /* must be n_defl + 1 = 7 */
  bad = m;
  bad.multiplane = &mp;
  EXPECT(lensim_workspace_size(&bad, &g, &psf, 4, 0) == 0);
  mp.n_planes = 7; /* arrays NULL */
  EXPECT(lensim_workspace_size(&bad, &g, &psf, 4, 0) == 0);
  lensim_grid bg = grid_of(0, 10, 1);
  EXPECT(lensim_workspace_size(&m, &bg, &psf, 4, 0) == 0);
  bg = grid_of(10, 10, 0);
  EXPECT(lensim_workspace_size(&m, &bg, &psf, 4, 0) == 0);
  bg = grid_of(1 << 16, 1 << 16, 1); /* 2^32 pixels */
  EXPECT(lensim_workspace_size(&m, &bg, &psf, 4, 0) == 0);
  lensim_psf bp = psf;
  bp.kx = 20; /* even */
  EXPECT(lensim_workspace_size(&m, &g, &bp, 4, 0) == 0);
  bp = psf;
  bp.type = 7;
  EXPECT(lensim_workspace_size(&m, &g, &bp, 4, 0) == 0);
  bp = psf;
  bp.d_kernel = NULL;
  EXPECT(lensim_workspace_size(&m, &g, &bp, 4, 0) == 0);
  bp = psf;
  bp.supersampled = 1;
  bp.kx_low = 4; /* even split kernel */
  bp.ky_low = 5;
  bp.d_kernel_low = fake_device;
  EXPECT(lensim_workspace_size(&m, &g, &bp, 4, 0) == 0);
  bp.kx_low = 5;
  EXPECT(lensim_workspace_size(&m, &g, &bp, 4, 0) > 0);
  bp.type = LENSIM_PSF_GAUSSIAN; /* a split kernel needs a PIXEL psf */
  EXPECT(lensim_workspace_size(&m, &g, &bp, 4, 0) == 0);

  /* entry points: bad arguments and short workspaces fail before device work */
  lensim_data d;
  memset(&d, 0, sizeof(d));
  d.d_data = fake_device;
  d.exposure_time = 100.0;
  d.background_rms = 0.01;
  double *P = fake_device, *out = fake_device + 8;
  EXPECT(lensim_loglike(NULL, &g, &psf, &d, P, 4, 0, out, NULL, NULL, 0, NULL) == LENSIM_EINVAL);
  EXPECT(lensim_loglike(&m, NULL, &psf, &d, P, 4, 0, out, NULL, NULL, 0, NULL) == LENSIM_EINVAL);
  EXPECT(lensim_loglike(&m, &g, &psf, NULL, P, 4, 0, out, NULL, NULL, 0, NULL) == LENSIM_EINVAL);
  EXPECT(lensim_loglike(&m, &g, &psf, &d, NULL, 4, 0, out, NULL, NULL, 0, NULL) == LENSIM_EINVAL);
  EXPECT(lensim_loglike(&m, &g, &psf, &d, P, -1, 0, out, NULL, NULL, 0, NULL) == LENSIM_EINVAL);
  EXPECT(lensim_loglike(&m, &g, &psf, &d, P, 0, 0, out, NULL, NULL, 0, NULL) == LENSIM_OK);
  EXPECT(lensim_loglike(&m, &g, &psf, &d, P, 70000, 0, out, NULL, NULL, 0, NULL) == LENSIM_EINVAL);
  EXPECT(lensim_loglike(&m, &g, &psf, &d, P, 4, 3, out, NULL, NULL, 0, NULL) == LENSIM_EINVAL);
  EXPECT(lensim_loglike(&m, &g, &psf, &d, P, 4, 0, out, NULL, NULL, 0, NULL) == LENSIM_EWORKSPACE);
  EXPECT(lensim_loglike(&m, &g, &psf, &d, P, 4, 0, out, NULL, NULL, 1 << 30, NULL) == LENSIM_EINVAL);
  EXPECT(lensim_loglike(&m, &g, &psf, &d, P, 4, 0, out, NULL, (char *)fake_device + 8, 1 << 30, NULL) ==
         LENSIM_EINVAL); /* misaligned workspace */
  lensim_data bd = d;
  bd.exposure_time = 0.0;
  EXPECT(lensim_loglike(&m, &g, &psf, &bd, P, 4, 0, out, NULL, NULL, 0, NULL) == LENSIM_EINVAL);
  bd = d;
  bd.image_stride = 100; /* smaller than one image */
  EXPECT(lensim_loglike(&m, &g, &psf, &bd, P, 4, 0, out, NULL, NULL, 0, NULL) == LENSIM_EINVAL);
  EXPECT(lensim_image(&m, &g, &psf, P, 4, 0, 3u, out, NULL, 0, NULL) == LENSIM_EWORKSPACE);
  EXPECT(lensim_image(&bad, &g, &psf, P, 4, 0, 3u, out, NULL, 0, NULL) < 0);
  EXPECT(lensim_raytrace(&m, P, 4, P, P, 100, 0, out, out, NULL, 0, NULL) == LENSIM_EWORKSPACE);
  EXPECT(lensim_raytrace(&m, P, 4, P, P, -5, 0, out, out, NULL, 0, NULL) == LENSIM_EINVAL);
  EXPECT(lensim_surface_brightness(&m, P, 4, 0, P, P, 100, 0, out, NULL, 0, NULL) < 0);
  EXPECT(lensim_surface_brightness(&m, P, 4, 5, P, P, 100, 0, out, NULL, 0, NULL) == LENSIM_EINVAL);
  EXPECT(lensim_loglike_grad(&m, &g, &psf, &d, P, 4, 0, out, out, NULL, 0, NULL) == LENSIM_EWORKSPACE);
  EXPECT(lensim_loglike_grad(&m, &g, &psf, &d, P, 4, 0, out, NULL, NULL, 0, NULL) == LENSIM_EINVAL);
  EXPECT(lensim_data_loglike(&d, 120, 120, P, 4, 0, out, NULL, 0, NULL) == LENSIM_EWORKSPACE);
  EXPECT(lensim_data_loglike(NULL, 120, 120, P, 4, 0, out, NULL, 0, NULL) == LENSIM_EINVAL);

  /* the allocation-free entry point: a batch beyond its scratch array is refused */
  EXPECT(lensim_noalloc_scratch_bytes() == (size_t)64 << 20);
  lensim_grid big = grid_of(180, 180, 3);
  EXPECT(lensim_loglike_noalloc(&m, &big, &psf, &d, P, 256, 0, out, NULL, NULL) == LENSIM_EWORKSPACE);
  EXPECT(strstr(lensim_last_error(), "scratch") != NULL);
  EXPECT(lensim_loglike_noalloc(&bad, &big, &psf, &d, P, 4, 0, out, NULL, NULL) < 0);

  /* graph capture needs a handle and a stream of its own */
  lensim_graph h = (lensim_graph)fake_device;
  EXPECT(lensim_graph_capture_loglike(&m, &g, &psf, &d, P, 4, 0, out, NULL, NULL, 0, NULL, &h) == LENSIM_EINVAL &&
         h == NULL);
  EXPECT(lensim_graph_capture_loglike(&m, &g, &psf, &d, P, 4, 0, out, NULL, NULL, 0, NULL, NULL) == LENSIM_EINVAL);
  EXPECT(lensim_graph_destroy(NULL) == LENSIM_OK);
  EXPECT(lensim_graph_launch(NULL, NULL) == LENSIM_EINVAL);

  /* timing with nothing recorded */
  double r = -1.0, c = -1.0;
  int64_t calls = -1;
  EXPECT(lensim_timing_enable(1) == LENSIM_OK);
  EXPECT(lensim_timing_read(&r, &c, &calls) == LENSIM_OK && r == 0.0 && c == 0.0 && calls == 0);
  EXPECT(lensim_timing_enable(0) == LENSIM_OK);

  if (n_fail) {
    fprintf(stderr, "%d checks failed\n", n_fail);
    return 1;
  }
  printf("host ABI checks passed\n");
  return 0;
}
