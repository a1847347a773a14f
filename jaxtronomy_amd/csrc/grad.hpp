// Profiles written once for plain scalars and forward-mode duals (gm::Dual):
// the gradient kernels evaluate one profile with its own parameters as dual
// seeds, so every derivative is taken with respect to the reference's kwargs
// (theta_E, gamma, e1, e2, ...), through the same parameter conversions and
// floors as the forward path (floors and branches are decided on values, as
// jnp.maximum / jnp.where are under JAX differentiation).
#pragma once

#include "dual.hpp"
#include "profiles.hpp"

namespace lensim {
namespace gm {

template <typename T, int N>
__device__ __forceinline__ Dual<T, N> lift(T x, const Dual<T, N> &) {
  return Dual<T, N>(x);
}
template <typename T>
__device__ __forceinline__ T lift(T x, T) {
  return x;
}

// c + dc * (dt/dseeds): a series coefficient F_n(t) as a function of the seeds
template <typename T, int N>
__device__ __forceinline__ Dual<T, N> coef_of_t(T c, T dc, const Dual<T, N> &t) {
  Dual<T, N> r;
  r.v = c;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = dc * t.d[i];
  return r;
}
template <typename T>
__device__ __forceinline__ T coef_of_t(T c, T, T) {
  return c;
}

template <typename S>
__device__ __forceinline__ bool is_nan(const S &x) {
  const auto v = val(x);
  return v != v;
}

// Util/param_util.py:88-100
template <typename T, typename S>
__device__ __forceinline__ void phi_q(const S &e1, const S &e2, S &phi, S &q) {
  phi = atan2(e2, e1) * T(0.5);
  S c = sqrt(e1 * e1 + e2 * e2);
  c = fmin(c, T(0.9999));
  q = (T(1) - c) / (T(1) + c);
}

// Sum_{n<N} c_n(t) w^n (quadratic-division form; see real_poly_at_complex)
template <typename T, typename S>
__device__ __forceinline__ void series_g(const T *__restrict__ c, const T *__restrict__ dc, int n, const S &t,
                                         const S &wr, const S &wi, S &pr, S &pi) {
  if (n == 1) {
    pr = coef_of_t(c[0], dc[0], t);
    pi = lift(T(0), t);
    return;
  }
  const S sw = T(2) * wr, tw = wr * wr + wi * wi;
  S b2 = lift(T(0), t), b1 = coef_of_t(c[n - 1], dc[n - 1], t);
  for (int k = n - 2; k >= 1; --k) {
    const S b0 = sw * b1 - tw * b2 + coef_of_t(c[k], dc[k], t);
    b2 = b1;
    b1 = b0;
  }
  pr = b1 * wr - tw * b2 + coef_of_t(c[0], dc[0], t);
  pi = b1 * wi;
}

// Clenshaw form of sum_{k<n} a(k) w^k for real coefficients a(k) at complex w
// (sw = 2 Re w, tw = |w|^2), as series_g's value.
template <typename T, typename A>
__device__ __forceinline__ void clenshaw_c(A a, int n, T wr, T wi, T sw, T tw, T &pr, T &pi) {
  if (n <= 0) {
    pr = pi = T(0);
    return;
  }
  if (n == 1) {
    pr = a(0);
    pi = T(0);
    return;
  }
  T b2 = T(0), b1 = a(n - 1);
  for (int k = n - 2; k >= 1; --k) {
    const T b0 = sw * b1 - tw * b2 + a(k);
    b2 = b1;
    b1 = b0;
  }
  pr = b1 * wr - tw * b2 + a(0);
  pi = b1 * wi;
}

// The series in duals without dual arithmetic inside the loop: the sum is
// holomorphic in w, so its tangents are S'(w) dw + U(w) dt with
// S'(w) = sum (k+1) c_{k+1} w^k and U(w) = sum dc_k w^k, three scalar
// recurrences instead of one recurrence on N + 1 components.
template <typename T, int N>
__device__ __forceinline__ void series_g(const T *__restrict__ c, const T *__restrict__ dc, int n,
                                         const Dual<T, N> &t, const Dual<T, N> &wr, const Dual<T, N> &wi,
                                         Dual<T, N> &pr, Dual<T, N> &pi) {
  const T wrv = wr.v, wiv = wi.v;
  const T sw = T(2) * wrv, tw = wrv * wrv + wiv * wiv;
  T sr, si, dr, di, ur, ui;
  clenshaw_c<T>([&](int k) { return c[k]; }, n, wrv, wiv, sw, tw, sr, si);
  clenshaw_c<T>([&](int k) { return T(k + 1) * c[k + 1]; }, n - 1, wrv, wiv, sw, tw, dr, di);
  clenshaw_c<T>([&](int k) { return dc[k]; }, n, wrv, wiv, sw, tw, ur, ui);
  pr.v = sr;
  pi.v = si;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    pr.d[i] = dr * wr.d[i] - di * wi.d[i] + ur * t.d[i];
    pi.d[i] = dr * wi.d[i] + di * wr.d[i] + ui * t.d[i];
  }
}

// EPL, LensModel/Profiles/epl.py:109-124, :178-200, :364-394.  p: theta_E,
// gamma, e1, e2, center_x, center_y.  nser: series length (from the value of f).
// q, cs, sn: the axis ratio and cos / sin of the orientation (from the record, or phi_q).
template <typename T, typename S>
__device__ __forceinline__ void epl_g_geo(const S *p, const S &q, const S &cs, const S &sn,
                                          const T *__restrict__ c, const T *__restrict__ dc, int nser, const S &x,
                                          const S &y, S &ax, S &ay) {
  const S t = p[1] - T(1);
  const S b = p[0] * sqrt(q);
  const S f = (T(1) - q) / (T(1) + q);
  const S dx = x - p[4], dy = y - p[5];
  const S xr = dx * cs + dy * sn;
  const S yr = dy * cs - dx * sn;
  const S Zr = q * xr, Zi = yr;
  const S R2 = Zr * Zr + Zi * Zi;
  if (val(R2) == T(0)) return;  // the reference's NaN at Z = 0 becomes 0 (epl.py:391-392)
  const S R = fmax(sqrt(R2), T(0.000000001));
  S Or, Oi;
  if (val(f) == T(0)) {
    Or = Zr;
    Oi = Zi;
  } else {
    const S zr = -f * (Zr * Zr - Zi * Zi) / R2;
    const S zi = -f * (T(2) * Zr * Zi) / R2;
    const S dr = zr - T(2);
    const S dd = T(1) / (dr * dr + zi * zi);
    const S wr = (zr * dr + zi * zi) * dd;
    const S wi = (zi * dr - zr * zi) * dd;
    S Sr, Si;
    series_g<T>(c, dc, nser, t, wr, wi, Sr, Si);
    const S er = T(1) - zr * T(0.5), ei = zi * T(-0.5);
    const S ed = T(1) / (er * er + ei * ei);
    const S hr = (Sr * er + Si * ei) * ed;
    const S hi = (Si * er - Sr * ei) * ed;
    Or = Zr * hr - Zi * hi;
    Oi = Zr * hi + Zi * hr;
  }
  const S scale = T(2) / (T(1) + q) * exp(t * log(b / R));
  S fr = scale * Or, fi = scale * Oi;
  if (is_nan(fr)) fr = lift(T(0), t);
  if (is_nan(fi)) fi = lift(T(0), t);
  ax = ax + (fr * cs - fi * sn);
  ay = ay + (fr * sn + fi * cs);
}

template <typename T, typename S>
__device__ void epl_g(const S *p, const T *__restrict__ c, const T *__restrict__ dc, int nser, const S &x,
                      const S &y, S &ax, S &ay) {
  S phi, q;
  phi_q<T>(p[2], p[3], phi, q);
  epl_g_geo<T>(p, q, cos(phi), sin(phi), c, dc, nser, x, y, ax, ay);
}

// NIE / SIE, LensModel/Profiles/nie.py:86-107, :182-200, :280-295, from the
// axis ratio q and the orientation's cos / sin (cs, sn).
template <typename T, typename S>
__device__ __forceinline__ void nie_g_geo(const S &theta_E, const S &q, const S &cs, const S &sn, const S &s_scale,
                                          const S &cx, const S &cy, const S &x, const S &y, S &ax, S &ay) {
  const S theta_E_conv = theta_E / sqrt((T(1) + q * q) / (T(2) * q));
  const S b = theta_E_conv * sqrt((T(1) + q * q) * T(0.5));
  const S s = s_scale / sqrt(q);
  const S dx = x - cx, dy = y - cy;
  const S xr = dx * cs + dy * sn;
  const S yr = dy * cs - dx * sn;
  const S qc = val(q) >= T(1) ? lift(T(0.99999999), q) : q;
  const S psi = sqrt(qc * qc * (s * s + xr * xr) + yr * yr);
  const S sq = sqrt(T(1) - qc * qc);
  const S fx = b / sq * atan(sq * xr / (psi + s));
  const S fy = b / sq * atanh(sq * yr / (psi + qc * qc * s));
  ax = ax + (fx * cs - fy * sn);
  ay = ay + (fx * sn + fy * cs);
}

// p: theta_E, e1, e2, s_scale, center_x, center_y (SIE: s_scale = 1e-10, centre at p[3..4]).
template <typename T, typename S>
__device__ void nie_g(const S &theta_E, const S &e1, const S &e2, const S &s_scale, const S &cx, const S &cy,
                      const S &x, const S &y, S &ax, S &ay) {
  S phi, q;
  phi_q<T>(e1, e2, phi, q);
  nie_g_geo<T>(theta_E, q, cos(phi), sin(phi), s_scale, cx, cy, x, y, ax, ay);
}

// F(X) of the NFW family: arctanh(sqrt(1 - X^2)) / sqrt(1 - X^2) = arccosh(1/X)
// / sqrt(1 - X^2) below X = 1, arctan(sqrt(X^2 - 1)) / sqrt(X^2 - 1) = arccos(1/X)
// / sqrt(X^2 - 1) above, 1 at X = 1 (nfw.py:345-365, tnfw.py:288-304).  Both
// sides are the one series sum_n (-u)^n / (2n + 1) in u = X^2 - 1, used for
// |u| < u0: there the closed forms' derivatives are differences of two terms
// that agree to O(u), so their duals lose (epsilon / |u|) relative precision
// (all of it in float within a few 1e-7 of X = 1); the series carries the
// derivative with no cancellation.  u0 = 0.1 (float) leaves a truncation of
// 0.1^9 / 19 after nine terms, u0 = 0.01 (double) 0.01^9 / 19.
template <typename T, typename S>
__device__ __forceinline__ S nfw_F_g(const S &X) {
  const S u = X * X - T(1);
  const T u0 = sizeof(T) == sizeof(float) ? T(0.1) : T(0.01);
  if (val(u) < u0 && val(u) > -u0) {
    S f = lift(T(1) / T(17), u);
#pragma unroll
    for (int n = 7; n >= 0; --n) f = T(1) / T(2 * n + 1) - u * f;
    return f;
  }
  if (val(u) < T(0)) {
    const S r = sqrt(-u);
    return atanh(r) / r;
  }
  const S r = sqrt(u);
  return atan(r) / r;
}

// NFW.g, nfw.py:345-365: log(X/2) + F(X)
template <typename T, typename S>
__device__ __forceinline__ S nfw_g_g(S X) {
  X = fmax(X, T(0.000001));
  return log(X * T(0.5)) + nfw_F_g<T>(X);
}

// NFW, nfw.py:86-105, :257-279, :384-395.  p: Rs, alpha_Rs, center_x, center_y.
template <typename T, typename S>
__device__ void nfw_g(const S *p, const S &x, const S &y, S &ax, S &ay) {
  const S rho0 = p[1] / (T(4) * p[0] * p[0] * (T(1) + T(::log(0.5))));
  const S Rs = fmax(p[0], T(0.0000001));
  const S dx = x - p[2], dy = y - p[3];
  const S R = fmax(sqrt(dx * dx + dy * dy), T(0.00000001));
  const S X = R / Rs;
  const S a = T(4) * rho0 * Rs * nfw_g_g<T>(X) / (X * X);
  ax = ax + a * dx;
  ay = ay + a * dy;
}

// TNFW, tnfw.py:67-87, :226-248, :276-349.  p: Rs, alpha_Rs, r_trunc, center_x, center_y.
template <typename T, typename S>
__device__ void tnfw_g(const S *p, const S &x, const S &y, S &ax, S &ay) {
  const S Rs = p[0];
  const S rho0 = p[1] / (T(4) * Rs * Rs * (T(1) + T(::log(0.5))));
  const S dx = x - p[3], dy = y - p[4];
  S R = sqrt(dx * dx + dy * dy);
  R = val(R) >= val(Rs) * T(0.001) ? R : Rs * T(0.001);
  const S X = fmax(R / Rs, T(0.001));
  const S tau = p[2] / Rs;
  const S t2 = tau * tau;
  const S F = nfw_F_g<T>(X);
  const S L = log(X / (tau + sqrt(t2 + X * X)));
  const S g = t2 / ((t2 + T(1)) * (t2 + T(1))) *
              ((t2 + T(1) + T(2) * (X * X - T(1))) * F + tau * T(M_PI) + (t2 - T(1)) * log(tau) +
               sqrt(t2 + X * X) * (L * (t2 - T(1)) / tau - T(M_PI)));
  const S a = T(4) * rho0 * Rs * g / (X * X);
  ax = ax + a * dx;
  ay = ay + a * dy;
}

// One deflector from its raw slot.
template <typename T, typename S>
__device__ __forceinline__ void defl_g(int type, const S *p, const T *c, const T *dc, int nser, const S &x,
                                       const S &y, S &ax, S &ay) {
  switch (type) {
    case LENSIM_EPL:
      epl_g<T>(p, c, dc, nser, x, y, ax, ay);
      break;
    case LENSIM_SIE:
      nie_g<T>(p[0], p[1], p[2], lift(T(0.0000000001), x), p[3], p[4], x, y, ax, ay);
      break;
    case LENSIM_NIE:
      nie_g<T>(p[0], p[1], p[2], p[3], p[4], p[5], x, y, ax, ay);
      break;
    case LENSIM_NFW:
      nfw_g<T>(p, x, y, ax, ay);
      break;
    case LENSIM_TNFW:
      tnfw_g<T>(p, x, y, ax, ay);
      break;
    case LENSIM_SHEAR: {
      const S dx = x - p[2], dy = y - p[3];
      ax = ax + (p[0] * dx + p[1] * dy);
      ay = ay + (p[1] * dx - p[0] * dy);
      break;
    }
    case LENSIM_CONVERGENCE:  // convergence.py:34-46; p: kappa, ra_0, dec_0
      ax = ax + p[0] * (x - p[1]);
      ay = ay + p[0] * (y - p[2]);
      break;
    default:
      break;
  }
}

// Sersic kernel, sersic_utils.py:239-267 (R_sersic already max(0, .) for the ellipse)
template <typename T, typename S>
__device__ __forceinline__ S sersic_g(const S &R, const S &R_sersic, const S &n) {
  const S R_ = fmax(R, T(0.0001));
  const S Rs_ = fmax(R_sersic, T(0.0001));
  const S bn = fmax(T(1.9992) * n - T(0.3271), T(0.00001));
  const S frac = R_ / Rs_;
  if (!(val(frac) <= T(1000))) return lift(T(0), R);
  return exp(-bn * (exp(log(frac) / n) - T(1)));
}

// SERSIC_ELLIPSE at (x, y) with its derivatives in closed form: returns the
// flux, writes d flux / d(amp, R_sersic, n_sersic, e1, e2, center_x,
// center_y) to gp and d flux / d(x, y) to gx, gy.  Same expression, floors
// and window as light_g (sersic_ellipse.py:62-93, param_util.py:103-124,
// sersic_utils.py:239-267); a floor that binds passes no derivative, as
// fmax does for duals.  With u = frac^(1/n) and E = exp(-b_n (u - 1)):
//   d flux / d ln(frac) = -amp E b_n u / n,
//   d flux / dn = -(d flux / d ln frac) ln(frac) / n - amp E (u - 1) db_n/dn.
template <typename T>
__device__ __forceinline__ T sersic_ellipse_grad(const double *p, T x, T y, T *gp, T &gx, T &gy) {
  const T amp = T(p[0]), p1 = T(p[1]), n = T(p[2]), e1 = T(p[3]), e2 = T(p[4]);
  const T dx = x - T(p[5]), dy = y - T(p[6]);
  const T Rs = p1 >= T(0) ? p1 : T(0);
  const T v = T(1) - e1 * e1 - e2 * e2;
  const T sq = sqrt(v >= T(0) ? v : -v);
  const bool nfree = sq >= T(0.000001);
  const T norm = nfree ? sq : T(0.000001);
  const T inorm = T(1) / norm;
  const T xe = ((T(1) - e1) * dx - e2 * dy) * inorm;
  const T ye = ((T(1) + e1) * dy - e2 * dx) * inorm;
  const T R = sqrt(xe * xe + ye * ye);
  const bool rfree = R >= T(0.0001), sfree = Rs >= T(0.0001);
  const T R_ = rfree ? R : T(0.0001);
  const T Rs_ = sfree ? Rs : T(0.0001);
  const T bn0 = T(1.9992) * n - T(0.3271);
  const bool bfree = bn0 >= T(0.00001);
  const T bn = bfree ? bn0 : T(0.00001);
  const T frac = R_ / Rs_;
#pragma unroll
  for (int i = 0; i < 7; ++i) gp[i] = T(0);
  gx = gy = T(0);
  if (!(frac <= T(1000))) return T(0);
  const T lf = log(frac);
  const T u = exp(lf / n);
  const T E = exp(-bn * (u - T(1)));
  const T c = -amp * E * bn * u / n;  // d flux / d ln(frac)
  gp[0] = E;
  gp[1] = (sfree && p1 >= T(0)) ? -c / Rs_ : T(0);
  gp[2] = -c * lf / n - (bfree ? amp * E * (u - T(1)) * T(1.9992) : T(0));
  // d flux / d(xe, ye) through R = |(xe, ye)|
  const T gR = rfree ? c / (R_ * R) : T(0);
  const T gxe = gR * xe, gye = gR * ye;
  // d norm / d e_i = -sign(v) e_i / sq where the floor does not bind
  const T sg = v >= T(0) ? T(1) : T(-1);
  const T dn1 = nfree ? -sg * e1 / sq : T(0), dn2 = nfree ? -sg * e2 / sq : T(0);
  gp[3] = (gxe * (-dx - xe * dn1) + gye * (dy - ye * dn1)) * inorm;
  gp[4] = (gxe * (-dy - xe * dn2) + gye * (-dx - ye * dn2)) * inorm;
  gx = (gxe * (T(1) - e1) - gye * e2) * inorm;
  gy = (gye * (T(1) + e1) - gxe * e2) * inorm;
  gp[5] = -gx;
  gp[6] = -gy;
  return amp * E;
}

// One light profile from its raw slot at (x, y).
template <typename T, typename S>
__device__ __forceinline__ S light_g(int type, const S *p, const S &x, const S &y) {
  switch (type) {
    case LENSIM_SERSIC_ELLIPSE: {
      // sersic_ellipse.py:62-93 with param_util.py:103-124 (product average)
      const S Rs = fmax(p[1], T(0));
      const S dx = x - p[5], dy = y - p[6];
      const S norm = fmax(sqrt(fabs(T(1) - p[3] * p[3] - p[4] * p[4])), T(0.000001));
      const S xe = ((T(1) - p[3]) * dx - p[4] * dy) / norm;
      const S ye = ((T(1) + p[3]) * dy - p[4] * dx) / norm;
      return p[0] * sersic_g<T>(sqrt(xe * xe + ye * ye), Rs, p[2]);
    }
    case LENSIM_SERSIC: {
      const S dx = x - p[3], dy = y - p[4];
      return p[0] * sersic_g<T>(sqrt(dx * dx + dy * dy), p[1], p[2]);
    }
    case LENSIM_GAUSSIAN: {
      // gaussian.py:27-42
      const S s2 = p[1] * p[1];
      const S dx = x - p[2], dy = y - p[3];
      return p[0] / (T(2 * M_PI) * s2) * exp(-(dx * dx / s2 + dy * dy / s2) * T(0.5));
    }
    default:
      return lift(T(0), x);
  }
}

}  // namespace gm
}  // namespace lensim
