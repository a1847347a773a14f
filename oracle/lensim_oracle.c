/*
 * FP64 CPU oracle in C: a restatement of JAXtronomy 0.1.3's image-simulation
 * hot path (ray shooting -> Sersic light -> s x s pooling -> direct 'same'
 * convolution -> chi^2 log-likelihood), parallel over pixels with OpenMP.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as a second oracle (checked against
 * the numpy restatement oracle/lensim_oracle.py) and by bench.py as the timed
 * CPU baseline.  The product library never links or calls it.
 *
 * Per ray it follows the reference's algorithm literally, including the
 * Lopez-Temme series with its contiguous recurrence evaluated inside the
 * per-ray loop (Util/hyp2f1_util.py:136-176).  Parameter slots follow the
 * layout of include/lensim.h ([n_img][n_prof][8], reference param_names order).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { P_EPL = 1, P_SIE = 2, P_NIE = 3, P_NFW = 4, P_TNFW = 5, P_SHEAR = 6, P_CONVERGENCE = 7,
       P_SERSIC = 32, P_SERSIC_ELLIPSE = 33, P_GAUSSIAN = 34 };

static void rotate(double x, double y, double a, double *xo, double *yo) {
  /* Util/util.py:103-114 */
  *xo = x * cos(a) + y * sin(a);
  *yo = -x * sin(a) + y * cos(a);
}

static void ellipticity2phi_q(double e1, double e2, double *phi, double *q) {
  /* Util/param_util.py:88-100 */
  *phi = atan2(e2, e1) / 2;
  double c = sqrt(e1 * e1 + e2 * e2);
  if (c > 0.9999) c = 0.9999;
  *q = (1 - c) / (1 + c);
}

static int epl_nmax(double f) {
  /* LensModel/Profiles/epl.py:311-317 */
  if (f == 0) return 1;
  if (f < 0.12) return 10;
  if (f < 0.4) return 20;
  if (f < 0.6) return 35;
  if (f < 0.73) return 50;
  if (f < 0.86) return 100;
  if (f < 0.92) return 200;
  return 500;
}

static double complex hyp2f1_lt8(double a, double b, double c, double complex z, int nmax) {
  /* Util/hyp2f1_util.py:136-176 */
  double complex w = z / (z - 2.0);
  double complex pref = a * w;
  double fm2 = 1.0, fm1 = 1.0 - 2.0 * b / c;
  double complex sum = 1 + pref * fm1;
  for (int n = 2; n < nmax; ++n) {
    pref *= (a + n - 1.0) / n * w;
    double fn = ((n - 1.0) * fm2 - (2.0 * b - c) * fm1) / (c + n - 1.0);
    sum += pref * fn;
    fm2 = fm1;
    fm1 = fn;
  }
  return cpow(1.0 - z / 2.0, -a) * sum;
}

static double nan_to_1e10(double v) {
  if (isnan(v)) return 0.0;
  if (isinf(v)) return v > 0 ? 1e10 : -1e10;
  return v;
}

static void epl(const double *p, double x, double y, double *ax, double *ay) {
  /* EPL.derivatives epl.py:178-200, EPLMajorAxis.derivatives :364-394 */
  double phi, q;
  ellipticity2phi_q(p[2], p[3], &phi, &q);
  double t = p[1] - 1, b = p[0] * sqrt(q);
  double xr, yr;
  rotate(x - p[4], y - p[5], phi, &xr, &yr);
  double complex Z = q * xr + yr * I;
  double R = cabs(Z);
  if (R < 0.000000001) R = 0.000000001;
  double f = (1.0 - q) / (1.0 + q);
  double complex hyp;
  if (f == 0)
    hyp = 1.0;
  else
    hyp = hyp2f1_lt8(1.0, t / 2.0, 2.0 - t / 2.0, -f * Z / conj(Z), epl_nmax(f));
  double complex alpha = 2 / (1 + q) * pow(b / R, t) * (Z * hyp);
  double fr = nan_to_1e10(creal(alpha)), fi = nan_to_1e10(cimag(alpha));
  double ox, oy;
  rotate(fr, fi, -phi, &ox, &oy);
  *ax += ox;
  *ay += oy;
}

static void nie(double theta_E, double e1, double e2, double s_scale, double cx, double cy, double x, double y,
                double *ax, double *ay) {
  /* NIE.derivatives nie.py:86-107, _param_conv :182-200, NIEMajorAxis :280-295 */
  double phi, q;
  ellipticity2phi_q(e1, e2, &phi, &q);
  double theta_E_conv = theta_E / sqrt((1.0 + q * q) / (2.0 * q));
  double b = theta_E_conv * sqrt((1 + q * q) / 2);
  double s = s_scale / sqrt(q);
  double xr, yr;
  rotate(x - cx, y - cy, phi, &xr, &yr);
  if (q >= 1) q = 0.99999999;
  double psi = sqrt(q * q * (s * s + xr * xr) + yr * yr);
  double fx = b / sqrt(1.0 - q * q) * atan(sqrt(1.0 - q * q) * xr / (psi + s));
  double fy = b / sqrt(1.0 - q * q) * atanh(sqrt(1.0 - q * q) * yr / (psi + q * q * s));
  double ox, oy;
  rotate(fx, fy, -phi, &ox, &oy);
  *ax += ox;
  *ay += oy;
}

static double nfw_g(double X) {
  /* nfw.py:345-365 */
  double c = 0.000001;
  if (X < c) X = c;
  if (X < 1) return log(X / 2.0) + 1 / sqrt(1 - X * X) * acosh(1.0 / X);
  if (X == 1) return 1 + log(1.0 / 2.0);
  return log(X / 2) + 1 / sqrt(X * X - 1) * acos(1.0 / X);
}

static void nfw(const double *p, double x, double y, double *ax, double *ay) {
  /* nfw.py:86-105, nfw_alpha :257-279, alpha2rho0 :384-395 */
  double Rs = p[0];
  double rho0 = p[1] / (4.0 * Rs * Rs * (1.0 + log(1.0 / 2.0)));
  if (Rs < 0.0000001) Rs = 0.0000001;
  double dx = x - p[2], dy = y - p[3];
  double R = sqrt(dx * dx + dy * dy);
  if (R < 0.00000001) R = 0.00000001;
  double X = R / Rs;
  double a = 4 * rho0 * Rs * nfw_g(X) / (X * X);
  *ax += a * dx;
  *ay += a * dy;
}

static double tnfw_F(double x) {
  /* tnfw.py:288-304 */
  if (x < 0.001) x = 0.001;
  if (x < 1) return pow(1 - x * x, -0.5) * atanh(pow(1 - x * x, 0.5));
  if (x > 1) return pow(x * x - 1, -0.5) * atan(pow(x * x - 1, 0.5));
  return x;
}

static void tnfw(const double *p, double x, double y, double *ax, double *ay) {
  /* tnfw.py:67-87, :226-248, _g :329-349, _L :276-286 */
  double Rs = p[0], r_trunc = p[2];
  double rho0 = p[1] / (4.0 * Rs * Rs * (1.0 + log(1.0 / 2.0)));
  double dx = x - p[3], dy = y - p[4];
  double R = sqrt(dx * dx + dy * dy);
  if (R < 0.001 * Rs) R = 0.001 * Rs;
  double X = R / Rs;
  if (X < 0.001) X = 0.001;
  double tau = r_trunc / Rs, t2 = tau * tau;
  double L = log(X * pow(tau + sqrt(t2 + X * X), -1));
  double g = t2 / ((t2 + 1) * (t2 + 1)) *
             ((t2 + 1 + 2 * (X * X - 1)) * tnfw_F(X) + tau * M_PI + (t2 - 1) * log(tau) +
              sqrt(t2 + X * X) * (-M_PI + L * (t2 - 1) / tau));
  double a = 4 * rho0 * Rs * g / (X * X);
  *ax += a * dx;
  *ay += a * dy;
}

static void deflect(int type, const double *p, double x, double y, double *ax, double *ay) {
  switch (type) {
    case P_EPL: epl(p, x, y, ax, ay); break;
    case P_SIE: nie(p[0], p[1], p[2], 0.0000000001, p[3], p[4], x, y, ax, ay); break;
    case P_NIE: nie(p[0], p[1], p[2], p[3], p[4], p[5], x, y, ax, ay); break;
    case P_NFW: nfw(p, x, y, ax, ay); break;
    case P_TNFW: tnfw(p, x, y, ax, ay); break;
    case P_SHEAR: {
      /* shear.py:37-53 */
      double dx = x - p[2], dy = y - p[3];
      *ax += p[0] * dx + p[1] * dy;
      *ay += p[1] * dx - p[0] * dy;
      break;
    }
    case P_CONVERGENCE:
      /* convergence.py:34-46 */
      *ax += p[0] * (x - p[1]);
      *ay += p[0] * (y - p[2]);
      break;
    default: break;
  }
}

static double r_sersic(double R, double R_sersic, double n) {
  /* sersic_utils.py:239-267 */
  double R_ = R > 0.0001 ? R : 0.0001;
  double Rs_ = R_sersic > 0.0001 ? R_sersic : 0.0001;
  double bn = 1.9992 * n - 0.3271;
  if (bn < 0.00001) bn = 0.00001;
  double frac = R_ / Rs_;
  double v = frac <= 1000.0 ? exp(-bn * (pow(frac, 1.0 / n) - 1.0)) : 0.0;
  if (isnan(v)) v = 0.0;
  return v;
}

static double light(int type, const double *p, double x, double y) {
  switch (type) {
    case P_SERSIC_ELLIPSE: {
      /* sersic_ellipse.py:62-93, param_util.py:103-124 */
      double Rs = p[1] > 0 ? p[1] : 0;
      double e1 = p[3], e2 = p[4];
      double dx = x - p[5], dy = y - p[6];
      double norm = sqrt(fabs(1 - e1 * e1 - e2 * e2));
      if (norm < 0.000001) norm = 0.000001;
      double xe = ((1 - e1) * dx - e2 * dy) / norm;
      double ye = (-e2 * dx + (1 + e1) * dy) / norm;
      return p[0] * r_sersic(sqrt(xe * xe + ye * ye), Rs, p[2]);
    }
    case P_SERSIC: {
      double dx = x - p[3], dy = y - p[4];
      return p[0] * r_sersic(sqrt(dx * dx + dy * dy), p[1], p[2]);
    }
    case P_GAUSSIAN: {
      /* gaussian.py:27-42 */
      double s2 = p[1] * p[1];
      double dx = x - p[2], dy = y - p[3];
      return p[0] / (2 * M_PI * s2) * exp(-(dx * dx / s2 + dy * dy / s2) / 2.0);
    }
    default: return 0.0;
  }
}

/* Unconvolved, pooled image of one model: flux_scaling x mean over s x s. */
static void render(int n_lens, int n_src, int n_light, const int *types, const double *prm, int nx, int ny, int s,
                   double ra0, double dec0, const double *M, double alpha_scaling, double flux_scaling,
                   unsigned flags, double *low) {
  const int np_ = n_lens + n_src + n_light;
  (void)np_;
#pragma omp parallel for schedule(dynamic, 4)
  for (int j = 0; j < ny; ++j) {
    for (int i = 0; i < nx; ++i) {
      double ra_c = i * M[0] + j * M[1] + ra0, dec_c = i * M[2] + j * M[3] + dec0;
      double acc = 0.0;
      for (int sy = 0; sy < s; ++sy) {
        double oy = -1 / 2.0 + 1 / (2.0 * s) + sy / (double)s;
        for (int sx = 0; sx < s; ++sx) {
          double ox = -1 / 2.0 + 1 / (2.0 * s) + sx / (double)s;
          double x = ra_c + M[0] * ox + M[1] * oy, y = dec_c + M[2] * ox + M[3] * oy;
          if ((flags & 1u) && n_src > 0) {
            double ax = 0, ay = 0;
            for (int k = 0; k < n_lens; ++k) deflect(types[k], prm + 8 * k, x, y, &ax, &ay);
            double bx = x - ax * alpha_scaling, by = y - ay * alpha_scaling;
            double f = 0;
            for (int k = n_lens; k < n_lens + n_src; ++k) f += light(types[k], prm + 8 * k, bx, by);
            acc += f * flux_scaling;
          }
          if ((flags & 2u) && n_light > 0) {
            double f = 0;
            for (int k = n_lens + n_src; k < n_lens + n_src + n_light; ++k) f += light(types[k], prm + 8 * k, x, y);
            acc += f * flux_scaling;
          }
        }
      }
      low[j * nx + i] = acc / (double)(s * s);
    }
  }
}

/* 'same' convolution, zero (edge = 0) or replicated-edge (edge = 1) padding. */
static void convolve(const double *in, int nx, int ny, const double *k, int kx, int ky, int edge, double scale,
                     double *out) {
  const int cx = (kx - 1) / 2, cy = (ky - 1) / 2;
#pragma omp parallel for schedule(static)
  for (int y = 0; y < ny; ++y) {
    for (int x = 0; x < nx; ++x) {
      double acc = 0.0;
      for (int a = 0; a < ky; ++a) {
        int yy = y + cy - a;
        if (edge) yy = yy < 0 ? 0 : (yy >= ny ? ny - 1 : yy);
        else if (yy < 0 || yy >= ny) continue;
        for (int c = 0; c < kx; ++c) {
          int xx = x + cx - c;
          if (edge) xx = xx < 0 ? 0 : (xx >= nx ? nx - 1 : xx);
          else if (xx < 0 || xx >= nx) continue;
          acc += in[yy * nx + xx] * k[a * kx + c];
        }
      }
      out[y * nx + x] = acc * scale;
    }
  }
}

/* Model images [n_img][ny][nx]; psf_type 0 none, 1 zero-padded, 2 edge-padded. */
int oracle_image(int n_lens, int n_src, int n_light, const int *types, const double *params, int n_img, int nx,
                 int ny, int s, double ra0, double dec0, const double *M, int psf_type, int kx, int ky,
                 const double *kernel, double alpha_scaling, double flux_scaling, unsigned flags, double *out,
                 int n_threads) {
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
  const int np_ = n_lens + n_src + n_light;
  const double pw2 = fabs(M[0] * M[3] - M[1] * M[2]);
  double *low = (double *)malloc(sizeof(double) * (size_t)nx * ny);
  if (!low) return -1;
  for (int b = 0; b < n_img; ++b) {
    render(n_lens, n_src, n_light, types, params + (size_t)b * np_ * 8, nx, ny, s, ra0, dec0, M, alpha_scaling,
           flux_scaling, flags, low);
    double *o = out + (size_t)b * nx * ny;
    if (psf_type == 0 || (flags & 4u)) {
      for (size_t i = 0; i < (size_t)nx * ny; ++i) o[i] = low[i] * pw2;
    } else {
      convolve(low, nx, ny, kernel, kx, ky, psf_type == 2, pw2, o);
    }
  }
  free(low);
  return 0;
}

/* log-likelihoods [n_img] (Data/imaging_data.py:159-182, image_noise.py:100-137). */
int oracle_loglike(int n_lens, int n_src, int n_light, const int *types, const double *params, int n_img, int nx,
                   int ny, int s, double ra0, double dec0, const double *M, int psf_type, int kx, int ky,
                   const double *kernel, double alpha_scaling, double flux_scaling, const double *data,
                   const double *noise, const double *expo_map, double expo_time, double bkg, const double *mask,
                   double *logl, int n_threads) {
  const size_t npix = (size_t)nx * ny;
  double *model = (double *)malloc(sizeof(double) * npix * (size_t)n_img);
  if (!model) return -1;
  int rc = oracle_image(n_lens, n_src, n_light, types, params, n_img, nx, ny, s, ra0, dec0, M, psf_type, kx, ky,
                        kernel, alpha_scaling, flux_scaling, 3u, model, n_threads);
  if (rc) {
    free(model);
    return rc;
  }
  for (int b = 0; b < n_img; ++b) {
    const double *m = model + (size_t)b * npix;
    double sum = 0.0;
    for (size_t p = 0; p < npix; ++p) {
      double var;
      if (noise) {
        var = noise[p] * noise[p];
      } else {
        double t = expo_map ? expo_map[p] : expo_time;
        if (t <= 1e-10) t = 1e-10;
        var = (m[p] >= 0 ? m[p] : 0) / t + bkg * bkg;
      }
      double chi = (m[p] - data[p]) * (m[p] - data[p]) / var;
      if (mask) chi *= mask[p];
      sum += chi;
    }
    logl[b] = -sum / 2;
  }
  free(model);
  return 0;
}
