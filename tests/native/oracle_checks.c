/* Runs the C oracle (oracle/lensim_oracle.c) over scenes that reach every
 * profile, PSF mode, mask and noise branch -- odd and non-square grids,
 * kernels wider than the image, 1x1 images, batches, a long deflector list,
 * EPL slopes near the series' limits -- built under AddressSanitizer and
 * UndefinedBehaviorSanitizer with OpenMP on two threads
 * (tests/test_sanitizers.py).  Exits non-zero if a result is not finite; the
 * sanitizers abort on the first finding. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int oracle_image(int n_lens, int n_src, int n_light, const int *types, const double *params, int n_img, int nx,
                 int ny, int s, double ra0, double dec0, const double *M, int psf_type, int kx, int ky,
                 const double *kernel, double alpha_scaling, double flux_scaling, unsigned flags, double *out,
                 int n_threads);
int oracle_loglike(int n_lens, int n_src, int n_light, const int *types, const double *params, int n_img, int nx,
                   int ny, int s, double ra0, double dec0, const double *M, int psf_type, int kx, int ky,
                   const double *kernel, double alpha_scaling, double flux_scaling, const double *data,
                   const double *noise, const double *expo_map, double expo_time, double bkg, const double *mask,
                   double *logl, int n_threads);

enum { EPL = 1, SIE = 2, NIE = 3, NFW = 4, TNFW = 5, SHEAR = 6, SERSIC = 32, SERSIC_ELLIPSE = 33, GAUSSIAN = 34 };

static int n_bad = 0;

static void check_finite(const char *what, const double *v, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!isfinite(v[i])) {
      fprintf(stderr, "%s: element %zu is %g\n", what, i, v[i]);
      ++n_bad;
      return;
    }
}

static void set(double *p, const double *v, int n) { memcpy(p, v, sizeof(double) * n); }

/* one parameter block of n_prof profiles for each of n_img images, the
 * profiles' values varied a little per image */
static void fill(const int *types, int n_prof, int n_img, double *params, unsigned seed) {
  srand(seed);
  for (int b = 0; b < n_img; ++b)
    for (int k = 0; k < n_prof; ++k) {
      double *p = params + ((size_t)b * n_prof + k) * 8;
      const double j = 0.02 * (rand() % 100) / 100.0;
      memset(p, 0, 8 * sizeof(double));
      switch (types[k]) {
        case EPL: { const double v[] = {1.2 + j, 1.9 + j, 0.1, -0.05, 0.01, -0.02}; set(p, v, 6); break; }
        case SIE: { const double v[] = {0.5 + j, 0.05, 0.1, 0.3, -0.2}; set(p, v, 5); break; }
        case NIE: { const double v[] = {0.4 + j, -0.1, 0.05, 0.1, -0.3, 0.2}; set(p, v, 6); break; }
        case NFW: { const double v[] = {0.6 + j, 0.2, 0.5, 0.0}; set(p, v, 4); break; }
        case TNFW: {
          const double rs = 0.05 + j;
          const double v[] = {rs, 0.02, 5 * rs, 1.5 * cos(k), 1.5 * sin(k)};
          set(p, v, 5);
          break;
        }
        case SHEAR: { const double v[] = {0.03 + j, -0.02, 0.0, 0.0}; set(p, v, 4); break; }
        case SERSIC: { const double v[] = {10.0, 0.3 + j, 2.5, 0.05, 0.02}; set(p, v, 5); break; }
        case SERSIC_ELLIPSE: { const double v[] = {20.0, 0.2 + j, 1.5, 0.1, -0.1, 0.02, 0.0}; set(p, v, 7); break; }
        case GAUSSIAN: { const double v[] = {5.0, 0.2 + j, -0.1, 0.05}; set(p, v, 4); break; }
      }
    }
}

static void scene(const char *name, int n_lens, int n_src, int n_light, const int *types, int n_img, int nx, int ny,
                  int s, int psf_type, int kx, int ky, double epl_gamma) {
  const int n_prof = n_lens + n_src + n_light;
  double *params = malloc(sizeof(double) * 8 * n_prof * n_img);
  fill(types, n_prof, n_img, params, 7u + nx);
  if (epl_gamma > 0)
    for (int b = 0; b < n_img; ++b)
      for (int k = 0; k < n_lens; ++k)
        if (types[k] == EPL) params[((size_t)b * n_prof + k) * 8 + 1] = epl_gamma;
  const double M[4] = {0.05, 0.0, 0.0, 0.05};
  const double ra0 = -0.05 * (nx - 1) / 2.0, dec0 = -0.05 * (ny - 1) / 2.0;
  double *kernel = malloc(sizeof(double) * kx * ky);
  for (int i = 0; i < kx * ky; ++i) kernel[i] = 1.0 / (kx * ky) * (1.0 + 0.01 * (i % 7));
  const size_t npix = (size_t)nx * ny;
  double *img = malloc(sizeof(double) * npix * n_img);
  for (unsigned flags = 1; flags <= 7; ++flags) {
    if (oracle_image(n_lens, n_src, n_light, types, params, n_img, nx, ny, s, ra0, dec0, M, psf_type, kx, ky, kernel,
                     1.0, 1.0, flags, img, 2)) {
      fprintf(stderr, "%s: oracle_image failed\n", name);
      ++n_bad;
    }
    check_finite(name, img, npix * n_img);
  }
  double *data = malloc(sizeof(double) * npix), *noise = malloc(sizeof(double) * npix);
  double *expo = malloc(sizeof(double) * npix), *mask = malloc(sizeof(double) * npix);
  for (size_t p = 0; p < npix; ++p) {
    data[p] = img[p] + 0.01 * ((int)(p % 5) - 2);
    noise[p] = 0.02 + 0.001 * (p % 3);
    expo[p] = 100.0 + (double)(p % 11);
    mask[p] = (p % 4) ? 1.0 : 0.0;
  }
  double *logl = malloc(sizeof(double) * n_img);
  for (int v = 0; v < 4; ++v) {
    const int rc = oracle_loglike(n_lens, n_src, n_light, types, params, n_img, nx, ny, s, ra0, dec0, M, psf_type,
                                  kx, ky, kernel, 1.0, 1.0, data, v == 1 ? noise : NULL, v == 2 ? expo : NULL,
                                  100.0, 0.01, v == 3 ? mask : NULL, logl, 2);
    if (rc) {
      fprintf(stderr, "%s: oracle_loglike failed\n", name);
      ++n_bad;
    }
    check_finite(name, logl, n_img);
  }
  free(params);
  free(kernel);
  free(img);
  free(data);
  free(noise);
  free(expo);
  free(mask);
  free(logl);
}

int main(void) {
  const int all[] = {EPL, SIE, NIE, NFW, TNFW, SHEAR, SERSIC, SERSIC_ELLIPSE, GAUSSIAN, SERSIC_ELLIPSE};
  scene("all profiles, 17x13 s=3, 5x7 kernel", 6, 3, 1, all, 2, 17, 13, 3, 1, 5, 7, -1);
  scene("all profiles, edge-padded 9x3", 6, 3, 1, all, 1, 12, 10, 2, 2, 9, 3, -1);
  scene("1x1 image, 9x9 kernel", 6, 3, 1, all, 1, 1, 1, 1, 1, 9, 9, -1);
  scene("2x1 image, edge padding", 6, 3, 1, all, 3, 2, 1, 2, 2, 5, 5, -1);
  scene("no psf", 6, 3, 1, all, 2, 8, 9, 1, 0, 1, 1, -1);
  const int epl[] = {EPL, SHEAR, SERSIC_ELLIPSE, SERSIC_ELLIPSE};
  const double gammas[] = {1.05, 1.5, 2.0, 2.5, 2.95};
  for (int i = 0; i < 5; ++i) scene("EPL slope", 2, 1, 1, epl, 1, 24, 24, 2, 1, 11, 11, gammas[i]);
  int halos[124];
  halos[0] = EPL;
  halos[1] = SHEAR;
  for (int k = 2; k < 122; ++k) halos[k] = TNFW;
  halos[122] = SERSIC_ELLIPSE;
  halos[123] = SERSIC;
  scene("120 TNFW", 122, 1, 1, halos, 2, 30, 30, 2, 1, 7, 7, -1);
  if (n_bad) {
    fprintf(stderr, "%d oracle checks failed\n", n_bad);
    return 1;
  }
  printf("oracle checks passed\n");
  return 0;
}
