// Device-side lens and light profiles for gfx950.
//
// Every profile has three pieces:
//   * a host/device "prepare" step (prepare_record) that turns the reference's
//     kwargs (one LENSIM_NPARAM slot of float64) into a record of derived
//     constants (rotation, axis ratio, normalisations, EPL series
//     coefficients), computed once per image and profile in float64;
//   * a per-ray evaluation templated on the compute type T (double for the
//     plain FP64 form, float for the mixed form).  Ray coordinates and the
//     deflection sums stay in double in both forms; only the per-profile
//     transcendental work runs in T.
//
// Formulas follow the reference profile files cited at each function; the
// numerical floors, branches and NaN policies are kept exactly in the FP64
// form.  The float form uses the same formulas except where float
// cancellation would lose accuracy (noted at each place).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/lensim.h"

namespace lensim {

constexpr int kRec = 24;  // derived constants per (image, profile)

// Sub-ray stepping of the supersampled grid, for the records' step constants:
// the sky offset from one sub-ray to the next along a pixel's first axis (sx),
// the offset from the end of a row of s sub-rays back to the start of the next
// row (rr = one step along the second axis minus s first-axis steps), and the
// deflection scaling alpha_scaling.  The render then updates every quantity
// that is affine in the ray position by one addition per sub-ray.
struct SubSteps {
  double sx[2], rr[2], asc;
};

// Record slot names shared by all profiles: centre first.
enum : int { R_CX = 0, R_CY = 1 };

// ---------------------------------------------------------------- maths ---

template <typename T>
struct Fn;

template <>
struct Fn<double> {
  static __device__ __forceinline__ double sqrt_(double x) { return ::sqrt(x); }
  static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
  static __device__ __forceinline__ double exp_(double x) { return ::exp(x); }
  static __device__ __forceinline__ double log_(double x) { return ::log(x); }
  static __device__ __forceinline__ double pow_(double a, double b) { return ::pow(a, b); }
  static __device__ __forceinline__ double atan_(double x) { return ::atan(x); }
  static __device__ __forceinline__ double atanh_(double x) { return ::atanh(x); }
  static __device__ __forceinline__ double acosh_(double x) { return ::acosh(x); }
  static __device__ __forceinline__ double acos_(double x) { return ::acos(x); }
};

template <>
struct Fn<float> {
  static __device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  static __device__ __forceinline__ float exp_(float x) { return __expf(x); }
  static __device__ __forceinline__ float log_(float x) { return __logf(x); }
  static __device__ __forceinline__ float pow_(float a, float b) { return __expf(b * __logf(a)); }
  static __device__ __forceinline__ float atan_(float x) { return ::atanf(x); }
  static __device__ __forceinline__ float atanh_(float x) { return ::atanhf(x); }
  static __device__ __forceinline__ float acosh_(float x) { return ::acoshf(x); }
  static __device__ __forceinline__ float acos_(float x) { return ::acosf(x); }
};

// 1/v to within two ulp: the hardware reciprocal refined by one Newton step
// (a correctly rounded float64 division issues about three times the
// instructions).  0, +-inf and NaN keep the hardware result (1/0 = inf,
// 1/inf = 0), where the Newton step itself would give NaN.
__device__ __forceinline__ double rcp_refined(double v) {
  const double r0 = __builtin_amdgcn_rcp(v);
  const double e = fma(-v, r0, 1.0);
  const double r = fma(e, r0, r0);
  return (e == e) ? r : r0;
}

// log2 x for finite x > 0 (the callers floor their arguments), to a few ulp
// in about a third of the library log2's instructions: x = m 2^e with m in
// [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s), s = (m - 1) / (m + 1), |s| <=
// 0.1716, and atanh(s) = s (1 + z / 3 + z^2 / 5 + ... + z^9 / 19), z = s^2 (the
// next term, s^21 / 21, is below 2.4e-17 s).  m + 1 lies in [1.7, 2.5], so the
// reciprocal's Newton step needs no guard.
__device__ __forceinline__ double log2_pos(double x) {
  int e;
  double m = ::frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
  if (m < 0.7071067811865476) {
    m += m;
    --e;
  }
  const double d = m + 1.0;
  double r = __builtin_amdgcn_rcp(d);
  r = fma(fma(-d, r, 1.0), r, r);
  const double sv = (m - 1.0) * r;
  const double z = sv * sv;
  double q = 0.05263157894736842;  // 1/19
  q = fma(q, z, 0.058823529411764705);
  q = fma(q, z, 0.06666666666666667);
  q = fma(q, z, 0.07692307692307693);
  q = fma(q, z, 0.09090909090909091);
  q = fma(q, z, 0.1111111111111111);
  q = fma(q, z, 0.14285714285714285);
  q = fma(q, z, 0.2);
  q = fma(q, z, 0.3333333333333333);
  const double at = fma(sv * z, q, sv);  // atanh(s)
  return fma(at, 2.8853900817779268, double(e));  // 2 atanh(s) / ln 2 + e
}

// atanh x for |x| < 1, within ~6 ulp in a fraction of the library atanh's
// instructions: the odd series x (1 + z / 3 + ... + z^9 / 19), z = x^2, for
// |x| <= 0.17 (the next term is below 1.2e-17 x); beyond, ln((1 + |x|) / (1 -
// |x|)) / 2 by log2_pos, whose argument is then at least 1.4 (no cancellation).
__device__ __forceinline__ double atanh_dbl(double x) {
  const double ax = fabs(x);
  if (ax <= 0.17) {
    const double z = x * x;
    double q = 0.05263157894736842;  // 1/19
    q = fma(q, z, 0.058823529411764705);
    q = fma(q, z, 0.06666666666666667);
    q = fma(q, z, 0.07692307692307693);
    q = fma(q, z, 0.09090909090909091);
    q = fma(q, z, 0.1111111111111111);
    q = fma(q, z, 0.14285714285714285);
    q = fma(q, z, 0.2);
    q = fma(q, z, 0.3333333333333333);
    return fma(x * z, q, x);
  }
  const double v = 0.34657359027997264 * log2_pos((1.0 + ax) * rcp_refined(1.0 - ax));  // ln 2 / 2
  return x < 0.0 ? -v : v;
}

// atan x for finite x, within ~2 ulp of the library's in a fraction of its
// instructions: |x| > 1 by pi/2 - atan(1/|x|); t > tan(pi/12) by pi/6 +
// atan(u), u = (t sqrt3 - 1) / (t + sqrt3), so |u| <= tan(pi/12) = 0.268; then
// the odd series u (1 - z / 3 + z^2 / 5 - ... + z^13 / 27), z = u^2 (the next
// term is below 3.5e-18 u).
__device__ __forceinline__ double atan_dbl(double x) {
  const double ax = fabs(x);
  const bool inv = ax > 1.0;
  const double t = inv ? rcp_refined(ax) : ax;
  const bool red = t > 0.2679491924311227;
  const double u = red ? fma(t, 1.7320508075688772, -1.0) * rcp_refined(t + 1.7320508075688772) : t;
  const double z = u * u;
  double q = 0.037037037037037035;  // 1/27
  q = fma(q, z, -0.04);
  q = fma(q, z, 0.043478260869565216);
  q = fma(q, z, -0.047619047619047616);
  q = fma(q, z, 0.05263157894736842);
  q = fma(q, z, -0.058823529411764705);
  q = fma(q, z, 0.06666666666666667);
  q = fma(q, z, -0.07692307692307693);
  q = fma(q, z, 0.09090909090909091);
  q = fma(q, z, -0.1111111111111111);
  q = fma(q, z, 0.14285714285714285);
  q = fma(q, z, -0.2);
  q = fma(q, z, 0.3333333333333333);
  double a = fma(u * z, -q, u);
  if (red) a += 0.5235987755982989;  // pi/6
  if (inv) a = 1.5707963267948966 - a;  // pi/2
  return x < 0.0 ? -a : a;
}

// 2^x for finite x, to a few ulp in half the library exp2's instructions:
// x = n + r, n = rint(x), |r| <= 1/2, 2^r = sum_k (r ln 2)^k / k! to k = 13
// (the next term is below 4.2e-18), times 2^n by ldexp; x is clamped to
// [-1100, 1100] first, so the result is 0 below 2^-1075 and inf above 2^1024.
__device__ __forceinline__ double exp2_fin(double x) {
  x = fmin(fmax(x, -1100.0), 1100.0);
  const double n = ::rint(x);
  const double r = x - n;
  double p = 1.3691488853904128e-12;
  p = fma(p, r, 2.5678435993488206e-11);
  p = fma(p, r, 4.4455382718708116e-10);
  p = fma(p, r, 7.054911620801123e-09);
  p = fma(p, r, 1.01780860092397e-07);
  p = fma(p, r, 1.321548679014431e-06);
  p = fma(p, r, 1.5252733804059841e-05);
  p = fma(p, r, 0.0001540353039338161);
  p = fma(p, r, 0.0013333558146428443);
  p = fma(p, r, 0.009618129107628477);
  p = fma(p, r, 0.05550410866482158);
  p = fma(p, r, 0.24022650695910072);
  p = fma(p, r, 0.6931471805599453);
  p = fma(p, r, 1.0);
  return ::ldexp(p, int(n));
}

// jnp.nan_to_num(v, posinf=1e10, neginf=-1e10) as used at LensModel/Profiles/epl.py:391-392
template <typename T>
__device__ __forceinline__ T nan_to_1e10(T v) {
  if (v != v) return T(0);
  if (v > T(1e10)) return (v == T(INFINITY)) ? T(1e10) : v;
  if (v < T(-1e10)) return (v == T(-INFINITY)) ? T(-1e10) : v;
  return v;
}

// ------------------------------------------------------ record layouts ----
// EPL     (epl.py:109-124): cos phi, sin phi, q, b, t, f, 2/(1+q), nmax
enum : int { EPL_C = 2, EPL_S, EPL_Q, EPL_B, EPL_T, EPL_F, EPL_PREF, EPL_NMAX,
              EPL_NZ,     // float form: terms of the z-series of delta (0: the w-series)
              EPL_NF,     // series length of the float form (truncation below 1e-8)
              EPL_F2,     // f^2 = |z|^2
              EPL_LNB,    // ln b
              EPL_TM1,    // t - 1
              EPL_PB,     // 2 b / (1 + q)
              EPL_ZSX,    // Z = (q xr, yr) steps: next sub-ray (real, imaginary) ...
              EPL_ZSXI,
              EPL_ZRR,    // ... and next row
              EPL_ZRRI,
              EPL_AC,     // alpha_scaling cos phi, alpha_scaling sin phi (the rotation back, scaled)
              EPL_AS,
              EPL_M2F,    // -2 f
              EPL_LB2 };  // log2 b
// In the float copy of an EPL record, EPL_NMAX, EPL_NZ and EPL_NF hold the
// int's bits (read with __float_as_int), so no per-ray conversion is needed.
static_assert(EPL_LB2 < kRec, "EPL record exceeds kRec");
// The float form's z-series coefficients g_1..g_NZ sit at this offset of an
// EPL's float coefficient array, past the w-series' F_n (nmax <= 50 whenever
// the z-series is chosen).
constexpr int kZMax = 24, kZOff = LENSIM_EPL_NMAX - kZMax - 4;
static_assert(kZOff >= 50, "z-series coefficients overlap the w-series");
// NIE/SIE (nie.py:182-200, :280-295): cos, sin, q (clamped), q^2, s, b/sqrt(1-q^2),
//          sqrt(1-q^2), q^2 s
//          then the steps of (xr, yr) and alpha_scaling (cos, sin) as for the EPL
enum : int { NIE_C = 2, NIE_S, NIE_Q, NIE_Q2, NIE_SS, NIE_BSQ, NIE_SQ, NIE_Q2S, NIE_XSX, NIE_YSX, NIE_XRR,
              NIE_YRR, NIE_AC, NIE_AS };
// NFW     (nfw.py:86-105, :257-279, :384-395): Rs (floored), 4 rho0 Rs, 1/Rs
enum : int { NFW_RS = 2, NFW_A, NFW_IRS };
// TNFW    (tnfw.py:67-87, :226-248, :329-349): Rs, 4 rho0 Rs, tau, tau^2,
//          tau^2/(tau^2+1)^2, tau pi + (tau^2-1) ln tau, (tau^2-1)/tau, 0.001 Rs, tau^2+1, 1/Rs
enum : int { TN_RS = 2, TN_A, TN_TAU, TN_TAU2, TN_PRE, TN_C0, TN_C1, TN_RMIN, TN_T2P1, TN_IRS,
              TN_K,      // pi (tau^2 + 1)^2 / 2
              TN_T2M1 }; // tau^2 - 1
// SHEAR   (shear.py:37-53): ra_0, dec_0 in the centre slots, gamma1, gamma2; the
//          steps of w = theta - alpha_scaling alpha_shear(theta) (affine in theta)
enum : int { SH_G1 = 2, SH_G2, SH_WSX, SH_WSY, SH_WRX, SH_WRY };
// CONVERGENCE (convergence.py:34-46): ra_0, dec_0 in the centre slots, kappa
enum : int { CV_KAPPA = 2 };
// SERSIC  (sersic_ellipse.py:62-93, sersic_utils.py:83-110, :239-267,
//          param_util.py:103-124): amp, the product-average matrix
//          ((1-e1), -e2; -e2, (1+e1))/norm, max(R_sersic, 1e-4), 1/n, b_n, 1/R_sersic_
enum : int { SE_AMP = 2, SE_M00, SE_M01, SE_M11, SE_RS, SE_INVN, SE_BN, SE_IRS,
              SE_C1,      // invn / 2            (frac^(1/n) = exp2(C1 log2(R^2 IRS2)))
              SE_IRS2,    // 1 / R_sersic_^2
              SE_R2MAX,   // (1000 R_sersic_)^2  (the 1000 R_sersic window on R^2)
              SE_LNRS,    // ln R_sersic_
              SE_K2LO,    // K2 - float(K2): the float form carries K2 as a head + tail pair
              SE_M00S,    // the matrix / R_sersic_: (xs, ys) = M (x - c) / R_sersic_, u = xs^2 + ys^2
              SE_M01S,
              SE_M11S,
              SE_UMIN,    // 1e-8 / R_sersic_^2 (the R^2 floor, on u)
              SE_XSX,     // steps of (xs, ys): next sub-ray, next row
              SE_YSX,
              SE_XRR,
              SE_YRR,
              SE_K2 };    // K2 = b_n log2(e): the profile is 2^(K2 (1 - frac^(1/n)))
static_assert(SE_K2 < kRec, "SERSIC record exceeds kRec");
// GAUSSIAN light (gaussian.py:27-42): amp/(2 pi sigma^2), 1/sigma^2
enum : int { GA_C = 2, GA_IS2 };

// ------------------------------------------------------------- prepare ---

// EPL hypergeometric series length as chosen at LensModel/Profiles/epl.py:311-317.
__host__ __device__ inline int epl_nmax(double f) {
  if (f == 0.0) return 1;
  if (f < 0.12) return 10;
  if (f < 0.4) return 20;
  if (f < 0.6) return 35;
  if (f < 0.73) return 50;
  if (f < 0.86) return 100;
  if (f < 0.92) return 200;
  return 500;
}

// Util/param_util.py:88-100
__host__ __device__ inline void ellipticity2phi_q(double e1, double e2, double &phi, double &q) {
  phi = atan2(e2, e1) / 2.0;
  double c = sqrt(e1 * e1 + e2 * e2);
  c = fmin(c, 0.9999);
  q = (1.0 - c) / (1.0 + c);
}

// Turns one kwargs slot p[LENSIM_NPARAM] into the derived record rec[kRec].
// For EPL also writes the series coefficients F_n = 2F1(-n, t/2; 2-t/2; 2),
// n < nmax, from Gauss' contiguous recurrence (Util/hyp2f1_util.py:156-166),
// their float copies (coef_f) and, for gradients, dF_n/dt (dcoef, dcoef_f).
// st: the grid's sub-ray steps (zero for calls without a grid).
__device__ inline void prepare_record(int type, const double *p, double *rec, double *coef, const SubSteps &st,
                                      double *dcoef = nullptr, float *coef_f = nullptr, float *dcoef_f = nullptr) {
  for (int i = 0; i < kRec; ++i) rec[i] = 0.0;
  switch (type) {
    case LENSIM_EPL: {
      const double theta_E = p[0], gamma = p[1], e1 = p[2], e2 = p[3];
      double phi, q;
      ellipticity2phi_q(e1, e2, phi, q);
      const double t = gamma - 1.0;
      const double b = theta_E * sqrt(q);
      const double f = (1.0 - q) / (1.0 + q);
      const int nmax = epl_nmax(f);
      rec[R_CX] = p[4];
      rec[R_CY] = p[5];
      rec[EPL_C] = cos(phi);
      rec[EPL_S] = sin(phi);
      rec[EPL_Q] = q;
      rec[EPL_B] = b;
      rec[EPL_T] = t;
      rec[EPL_F] = f;
      rec[EPL_PREF] = 2.0 / (1.0 + q);
      rec[EPL_NMAX] = (double)nmax;

      rec[EPL_F2] = f * f;
      rec[EPL_LNB] = log(b);
      rec[EPL_TM1] = t - 1.0;
      rec[EPL_PB] = 2.0 * b / (1.0 + q);
      rec[EPL_NF] = (double)nmax;
      {
        const double c = rec[EPL_C], sn = rec[EPL_S];
        rec[EPL_ZSX] = q * (st.sx[0] * c + st.sx[1] * sn);
        rec[EPL_ZSXI] = -st.sx[0] * sn + st.sx[1] * c;
        rec[EPL_ZRR] = q * (st.rr[0] * c + st.rr[1] * sn);
        rec[EPL_ZRRI] = -st.rr[0] * sn + st.rr[1] * c;
        rec[EPL_AC] = st.asc * c;
        rec[EPL_AS] = st.asc * sn;
      }
      rec[EPL_M2F] = -2.0 * f;
      rec[EPL_LB2] = log2(b);
      if (coef) {
        // F_n by the recurrence, with float copies and (for gradients) dF_n/dt
        // written as they are formed: nothing is read back from memory
        const double bb = t / 2.0, cc = 2.0 - bb, k2 = 2.0 * bb - cc;
        double fm2 = 1.0, fm1 = 1.0 - 2.0 * bb / cc;
        // d F_n / d t from the same recurrence: F_n = N_n / D_n with
        // N_n = (n-1) F_{n-2} - (1.5 t - 2) F_{n-1}, D_n = 1 + n - t/2
        double dm2 = 0.0, dm1 = -2.0 / (cc * cc);  // d/dt (1 - t / (2 - t/2))
        coef[0] = 1.0;
        if (coef_f) coef_f[0] = 1.f;
        if (dcoef) dcoef[0] = 0.0;
        if (dcoef_f) dcoef_f[0] = 0.f;
        if (nmax > 1) {
          coef[1] = fm1;
          if (coef_f) coef_f[1] = float(fm1);
          if (dcoef) dcoef[1] = dm1;
          if (dcoef_f) dcoef_f[1] = float(dm1);
        }
        double cmax = fmax(1.0, fabs(fm1));
        // the divisor 1/(cc + n - 1) does not depend on the recurrence: it is a
        // reciprocal refined to full precision by two Newton steps, which the
        // unrolled loop computes ahead, so the serial chain is two FMAs per n
#pragma unroll 4
        for (int n = 2; n < nmax; ++n) {
          const double d = cc + n - 1.0;
          double r = __builtin_amdgcn_rcp(d);
          r = fma(r, fma(-d, r, 1.0), r);
          r = fma(r, fma(-d, r, 1.0), r);
          const double fn = fma(n - 1.0, fm2, -k2 * fm1) * r;
          coef[n] = fn;
          if (coef_f) coef_f[n] = float(fn);
          if (dcoef) {
            const double dn = ((n - 1.0) * dm2 - 1.5 * fm1 - k2 * dm1) * r + 0.5 * fn * r;
            dcoef[n] = dn;
            if (dcoef_f) dcoef_f[n] = float(dn);
            dm2 = dm1;
            dm1 = dn;
          }
          cmax = fmax(cmax, fabs(fn));
          fm2 = fm1;
          fm1 = fn;
        }
        // |w| <= f / (2 - f) for |z| = f, so the terms n >= N add at most
        // max|F_n| wmax^N / (1 - wmax) to a sum of size ~1
        const double wmax = f / (2.0 - f);
        int nf = nmax;
        if (wmax > 0.0) {
          double tail = cmax / (1.0 - wmax);
          for (nf = 1; nf < nmax && tail >= 1e-9; ++nf) tail *= wmax;
        }
        rec[EPL_NF] = (double)nf;
        // float form, small f: delta = 2F1 - 1 = sum_{n>=1} g_n z^n with
        // g_n = prod_{j<n} (b + j)/(c + j) (b = t/2, c = 2 - t/2, so 0 < g_n < 1
        // for 0 < t < 2), |z| = f: the terms past N add at most f^(N+1)/(1-f)
        int nz = 0;
        if (coef_f && f > 0.0 && t > 0.0 && t < 2.0) {
          double tail = f * f / (1.0 - f);
          int n = 1;
          for (; n <= kZMax && tail >= 1e-9; ++n) tail *= f;
          if (tail < 1e-9) {  // the loop leaves tail = f^(n+1)/(1-f): n terms suffice
            nz = n;
            double g = 1.0;
            for (int j = 1; j <= nz; ++j) {
              g *= (bb + j - 1.0) / (cc + j - 1.0);
              coef_f[kZOff + j - 1] = float(g);
            }
          }
        }
        rec[EPL_NZ] = (double)nz;
      }
      break;
    }
    case LENSIM_SIE:
    case LENSIM_NIE: {
      const double theta_E = p[0], e1 = p[1], e2 = p[2];
      const double s_scale = (type == LENSIM_SIE) ? 0.0000000001 : p[3];
      const double cx = (type == LENSIM_SIE) ? p[3] : p[4];
      const double cy = (type == LENSIM_SIE) ? p[4] : p[5];
      double phi, q;
      ellipticity2phi_q(e1, e2, phi, q);
      const double theta_E_conv = theta_E / sqrt((1.0 + q * q) / (2.0 * q));
      const double b = theta_E_conv * sqrt((1.0 + q * q) / 2.0);
      const double s = s_scale / sqrt(q);
      const double qc = (q >= 1.0) ? 0.99999999 : q;
      const double sq = sqrt(1.0 - qc * qc);
      rec[R_CX] = cx;
      rec[R_CY] = cy;
      rec[NIE_C] = cos(phi);
      rec[NIE_S] = sin(phi);
      rec[NIE_Q] = qc;
      rec[NIE_Q2] = qc * qc;
      rec[NIE_SS] = s;
      rec[NIE_BSQ] = b / sq;
      rec[NIE_SQ] = sq;
      rec[NIE_Q2S] = qc * qc * s;
      {
        const double c = rec[NIE_C], sn = rec[NIE_S];
        rec[NIE_XSX] = st.sx[0] * c + st.sx[1] * sn;
        rec[NIE_YSX] = -st.sx[0] * sn + st.sx[1] * c;
        rec[NIE_XRR] = st.rr[0] * c + st.rr[1] * sn;
        rec[NIE_YRR] = -st.rr[0] * sn + st.rr[1] * c;
        rec[NIE_AC] = st.asc * c;
        rec[NIE_AS] = st.asc * sn;
      }
      break;
    }
    case LENSIM_NFW: {
      const double Rs0 = p[0], alpha_Rs = p[1];
      const double rho0 = alpha_Rs / (4.0 * Rs0 * Rs0 * (1.0 + log(1.0 / 2.0)));
      const double Rs = (Rs0 < 0.0000001) ? 0.0000001 : Rs0;
      rec[R_CX] = p[2];
      rec[R_CY] = p[3];
      rec[NFW_RS] = Rs;
      rec[NFW_A] = 4.0 * rho0 * Rs;
      rec[NFW_IRS] = 1.0 / Rs;
      break;
    }
    case LENSIM_TNFW: {
      const double Rs = p[0], alpha_Rs = p[1], r_trunc = p[2];
      const double rho0 = alpha_Rs / (4.0 * Rs * Rs * (1.0 + log(1.0 / 2.0)));
      const double tau = r_trunc / Rs, tau2 = tau * tau;
      rec[R_CX] = p[3];
      rec[R_CY] = p[4];
      rec[TN_RS] = Rs;
      rec[TN_A] = 4.0 * rho0 * Rs;
      rec[TN_TAU] = tau;
      rec[TN_TAU2] = tau2;
      rec[TN_PRE] = tau2 / ((tau2 + 1.0) * (tau2 + 1.0));
      rec[TN_C0] = tau * M_PI + (tau2 - 1.0) * log(tau);
      rec[TN_C1] = (tau2 - 1.0) / tau;
      rec[TN_RMIN] = 0.001 * Rs;
      rec[TN_T2P1] = tau2 + 1.0;
      rec[TN_IRS] = 1.0 / Rs;
      rec[TN_K] = 0.5 * M_PI * (tau2 + 1.0) * (tau2 + 1.0);
      rec[TN_T2M1] = tau2 - 1.0;
      break;
    }
    case LENSIM_SHEAR: {
      rec[R_CX] = p[2];
      rec[R_CY] = p[3];
      rec[SH_G1] = p[0];
      rec[SH_G2] = p[1];
      // w = theta - asc (g1 dx + g2 dy, g2 dx - g1 dy): its steps
      rec[SH_WSX] = st.sx[0] - st.asc * (p[0] * st.sx[0] + p[1] * st.sx[1]);
      rec[SH_WSY] = st.sx[1] - st.asc * (p[1] * st.sx[0] - p[0] * st.sx[1]);
      rec[SH_WRX] = st.rr[0] - st.asc * (p[0] * st.rr[0] + p[1] * st.rr[1]);
      rec[SH_WRY] = st.rr[1] - st.asc * (p[1] * st.rr[0] - p[0] * st.rr[1]);
      break;
    }
    case LENSIM_CONVERGENCE:
      rec[R_CX] = p[1];
      rec[R_CY] = p[2];
      rec[CV_KAPPA] = p[0];
      break;
    case LENSIM_SERSIC:
    case LENSIM_SERSIC_ELLIPSE: {
      const bool ell = (type == LENSIM_SERSIC_ELLIPSE);
      const double amp = p[0], n = p[2];
      double Rs = p[1];
      const double e1 = ell ? p[3] : 0.0, e2 = ell ? p[4] : 0.0;
      if (ell) Rs = fmax(0.0, Rs);
      const double norm = fmax(sqrt(fabs(1.0 - e1 * e1 - e2 * e2)), 0.000001);
      const double Rs_ = fmax(0.0001, Rs);
      rec[R_CX] = ell ? p[5] : p[3];
      rec[R_CY] = ell ? p[6] : p[4];
      rec[SE_AMP] = amp;
      rec[SE_M00] = (1.0 - e1) / norm;
      rec[SE_M01] = -e2 / norm;
      rec[SE_M11] = (1.0 + e1) / norm;
      rec[SE_RS] = Rs_;
      rec[SE_INVN] = 1.0 / n;
      rec[SE_BN] = fmax(1.9992 * n - 0.3271, 0.00001);
      rec[SE_IRS] = 1.0 / Rs_;
      rec[SE_C1] = 0.5 / n;
      rec[SE_IRS2] = 1.0 / (Rs_ * Rs_);
      rec[SE_R2MAX] = (1000.0 * Rs_) * (1000.0 * Rs_);
      rec[SE_LNRS] = log(Rs_);
      rec[SE_K2] = rec[SE_BN] * 1.4426950408889634;
      rec[SE_K2LO] = rec[SE_K2] - double(float(rec[SE_K2]));
      rec[SE_M00S] = rec[SE_M00] / Rs_;
      rec[SE_M01S] = rec[SE_M01] / Rs_;
      rec[SE_M11S] = rec[SE_M11] / Rs_;
      rec[SE_UMIN] = 1e-8 * rec[SE_IRS2];
      rec[SE_XSX] = rec[SE_M00S] * st.sx[0] + rec[SE_M01S] * st.sx[1];
      rec[SE_YSX] = rec[SE_M01S] * st.sx[0] + rec[SE_M11S] * st.sx[1];
      rec[SE_XRR] = rec[SE_M00S] * st.rr[0] + rec[SE_M01S] * st.rr[1];
      rec[SE_YRR] = rec[SE_M01S] * st.rr[0] + rec[SE_M11S] * st.rr[1];
      break;
    }
    case LENSIM_GAUSSIAN: {
      const double amp = p[0], sigma = p[1];
      rec[R_CX] = p[2];
      rec[R_CY] = p[3];
      rec[GA_C] = amp / (2.0 * M_PI * sigma * sigma);
      rec[GA_IS2] = 1.0 / (sigma * sigma);
      break;
    }
    default:
      break;
  }
}

// ------------------------------------------------------ per-ray lensing ---

// Record access: centres always in double; the rest in the compute type.
template <typename T>
struct Rec {
  const double *d;  // kRec doubles
  const T *t;       // kRec values of T (== d for T = double)
  __device__ __forceinline__ T operator[](int i) const { return t[i]; }
};

// max(u, floor) of a non-negative float by its bits (the order of non-negative
// floats is that of their bit patterns): one integer max, where fmax would
// first quiet its operands.
__device__ __forceinline__ float floor_nonneg_f32(float u, int floor_bits) {
  return __int_as_float(max(__float_as_int(u), floor_bits));
}

// Sum_{n < N} c_n w^n for real c_n at complex w, by dividing by the real
// quadratic (v - w)(v - conj w) = v^2 - s v + t (s = 2 Re w, t = |w|^2):
// b_k = c_k + s b_{k+1} - t b_{k+2}; the sum is b_1 w + c_0 - t b_2.
// Two FMAs per term instead of the four of a complex Horner step.
template <typename ST>
__device__ __forceinline__ void real_poly_at_complex(const ST *__restrict__ c, int N, ST wr, ST wi, ST &pr,
                                                     ST &pi) {
  if (N == 1) {
    pr = c[0];
    pi = ST(0);
    return;
  }
  const ST sw = ST(2) * wr, tw = wr * wr + wi * wi;
  ST b2 = ST(0), b1 = c[N - 1];
  int k = N - 2;
  // four terms per step: the coefficient loads of a step are issued together
  // (they do not depend on the recurrence) and the b1/b2 rotation unrolls away
  for (; k >= 4; k -= 4) {
    const ST c0 = c[k], c1 = c[k - 1], c2 = c[k - 2], c3 = c[k - 3];
    const ST a0 = fma(sw, b1, fma(-tw, b2, c0));
    const ST a1 = fma(sw, a0, fma(-tw, b1, c1));
    const ST a2 = fma(sw, a1, fma(-tw, a0, c2));
    const ST a3 = fma(sw, a2, fma(-tw, a1, c3));
    b2 = a2;
    b1 = a3;
  }
  // the last k < 4 steps, unrolled (no register rotation in a loop)
  if (k == 3) {
    const ST a0 = fma(sw, b1, fma(-tw, b2, c[3]));
    const ST a1 = fma(sw, a0, fma(-tw, b1, c[2]));
    const ST a2 = fma(sw, a1, fma(-tw, a0, c[1]));
    b2 = a1;
    b1 = a2;
  } else if (k == 2) {
    const ST a0 = fma(sw, b1, fma(-tw, b2, c[2]));
    const ST a1 = fma(sw, a0, fma(-tw, b1, c[1]));
    b2 = a0;
    b1 = a1;
  } else if (k == 1) {
    const ST a0 = fma(sw, b1, fma(-tw, b2, c[1]));
    b2 = b1;
    b1 = a0;
  }
  pr = fma(b1, wr, fma(-tw, b2, c[0]));
  pi = b1 * wi;
}

// EPL deflection, LensModel/Profiles/epl.py:178-200 and :364-394 (Tessore &
// Metcalf 2015 eqs. 22-23), FP64 form.  The Lopez-Temme series
// (Util/hyp2f1_util.py:136-176) in w = z/(z-2) with the reference's nmax terms
// over the per-image coefficients F_n; the reference's nan_to_num policy.
// The deflection in the major-axis frame (fr, fi) at Z = (Zr, Zi) = (q xr, yr),
// R2 = |Z|^2.
__device__ __forceinline__ void epl_major_f64(const Rec<double> &r, const double *__restrict__ coef, double Zr,
                                              double Zi, double R2, double &fr, double &fi) {
  const double t = r.d[EPL_T], f = r.d[EPL_F];
  // R = max(|Z|, 1e-9) enters only through ln R = ln(max(R^2, 1e-18)) / 2
  const double R2c = R2 > 1e-18 ? R2 : 1e-18;
  double Or, Oi;
  const int nmax = (int)r.d[EPL_NMAX];
  if (nmax == 1) {  // f == 0: the reference returns 2F1 = 1 (epl.py:294, :317)
    Or = Zr;
    Oi = Zi;
  } else {
    // z = -f Z / conj(Z) = -f Z^2 / |Z|^2 (NaN at Z = 0, as the reference);
    // the reciprocals refined to two ulp, not divided
    const double inv = rcp_refined(R2);
    const double zr = -f * (Zr * Zr - Zi * Zi) * inv;
    const double zi = -f * (2.0 * Zr * Zi) * inv;
    // w = z / (z - 2)
    const double dr = zr - 2.0;
    const double dd = rcp_refined(dr * dr + zi * zi);
    const double wr = (zr * dr + zi * zi) * dd;
    const double wi = (zi * dr - zr * zi) * dd;
    double Sr, Si;
    real_poly_at_complex<double>(coef, nmax, wr, wi, Sr, Si);
    // 2F1 = S / (1 - z/2)
    const double er = 1.0 - zr * 0.5, ei = -zi * 0.5;
    const double ed = rcp_refined(er * er + ei * ei);
    const double hr = (Sr * er + Si * ei) * ed;
    const double hi = (Si * er - Sr * ei) * ed;
    Or = Zr * hr - Zi * hi;
    Oi = Zr * hi + Zi * hr;
  }
  // (b/R)^t = 2^(t (log2 b - log2(R^2) / 2))
  const double scale = r.d[EPL_PREF] * exp2_fin(t * (r.d[EPL_LB2] - 0.5 * log2_pos(R2c)));
  fr = nan_to_1e10(scale * Or);
  fi = nan_to_1e10(scale * Oi);
}

__device__ __forceinline__ void epl_defl_f64(const Rec<double> &r, const double *__restrict__ coef, double x,
                                             double y, double &ax, double &ay) {
  const double dx = x - r.d[R_CX], dy = y - r.d[R_CY];
  const double c = r.d[EPL_C], s = r.d[EPL_S];
  const double xr = dx * c + dy * s;
  const double yr = -dx * s + dy * c;
  const double Zr = r.d[EPL_Q] * xr, Zi = yr;
  double fr, fi;
  epl_major_f64(r, coef, Zr, Zi, Zr * Zr + Zi * Zi, fr, fi);
  ax += fr * c - fi * s;
  ay += fr * s + fi * c;
}

// expm1(x) in float: for |x| < 0.5 the Taylor series to x^9 (truncation below
// 3e-10 |x|), else 2^(x log2 e) - 1 with the product split into head and
// rounding error (the subtraction then loses less than two bits).  Only the
// finite arguments of the EPL power law reach it.
__device__ __forceinline__ float expm1_f32(float x) {
  float p = fmaf(x, 2.7557319224e-6f, 2.4801587302e-5f);
  p = fmaf(x, p, 1.9841269841e-4f);
  p = fmaf(x, p, 1.3888888889e-3f);
  p = fmaf(x, p, 8.3333333333e-3f);
  p = fmaf(x, p, 4.1666666667e-2f);
  p = fmaf(x, p, 1.6666666667e-1f);
  p = fmaf(x, p, 0.5f);
  const float small = fmaf(x * x, p, x);
  // the EPL's argument (t - 1) ln(b/R) is below 0.5 in magnitude for most
  // rays: a wave whose lanes all are skips the exp2 form (a uniform branch)
  if (__all(fabsf(x) < 0.5f)) return small;
  asm volatile("");  // keeps the exp2 form behind the branch (else it is speculated and selected)
  const float L = 1.4426950408889634f;
  const float hi = x * L;
  const float lo = fmaf(x, L, -hi) + x * 1.925963033500011e-8f;  // log2(e) - L
  const float large = fmaf(__builtin_amdgcn_exp2f(hi), fmaf(lo, 0.6931471805599453f, 1.f), -1.f);
  return fabsf(x) < 0.5f ? small : large;
}

// EPL, mixed form.  The geometry runs in double (rotation, Z, 1/|Z| by a
// float rsqrt and one Newton step, the final products); float carries only
// the parts whose size bounds their error:
//  * delta = 2F1 - 1, of size ~ f t / (4 - t): for f < 0.44 the z-series
//    sum_{n>=1} g_n z^n (EPL_NZ terms, tail below 1e-9; z = -f (Z/|Z|)^2 from
//    the rsqrt, no division), otherwise the reference's w-series form
//    (S - 1 + z/2) / (1 - z/2), S - 1 = w sum_{n>=1} F_n w^(n-1), truncated at
//    EPL_NF terms;
//  * (b/R)^(t-1) = 1 + expm1((t - 1) ln(b/R)), |t - 1| small.
// Then alpha = 2b/(1+q) (b/R)^(t-1) (Z/R) (1 + delta), so its relative error
// is about eps (|delta| + |(t-1) ln(b/R)|) instead of several float ulps.
// Long series (nmax >= 100, f >= 0.73) form delta in double.
// The deflection in the major-axis frame (fr, fi) at Z = (Zrd, Zid) = (q xr, yr),
// R2d = |Z|^2.
__device__ __forceinline__ void epl_major_f32(const Rec<float> &r, const float *__restrict__ coef_f,
                                              const double *__restrict__ coef_d, double Zrd, double Zid, double R2d,
                                              double &fr, double &fi) {
  const float Zr = float(Zrd), Zi = float(Zid), R2 = float(R2d);
  float dr = 0.f, di = 0.f;  // delta (0 for f == 0, and at the centre where the reference's NaN becomes 0)
  const int nf = __float_as_int(r[EPL_NF]), nz = __float_as_int(r[EPL_NZ]);
  // 1/|Z|, unclamped: the direction Z/|Z| enters 2F1 however small |Z| is
  // (float R^2 holds |Z| down to 1e-19; 0 at the centre, where delta = 0)
  const float iRs = R2 > 0.f ? __builtin_amdgcn_rsqf(R2) : 0.f;
  if (nz > 0) {
    // z = -f (Z/|Z|)^2 from the rsqrt (no division); delta = z P(z) with the
    // real coefficients g_1..g_nz
    const float f = r[EPL_F];
    const float ur = Zr * iRs, ui = Zi * iRs;
    const float zr = -f * ((ur - ui) * (ur + ui));
    const float zi = r[EPL_M2F] * (ur * ui);
    float Pr, Pi;
    real_poly_at_complex<float>(coef_f + kZOff, nz, zr, zi, Pr, Pi);
    dr = fmaf(zr, Pr, -zi * Pi);
    di = fmaf(zr, Pi, zi * Pr);
  } else if (__float_as_int(r[EPL_NMAX]) >= 100) {
    // f >= 0.73: |w| approaches f/(2-f) ~ 1 and the long series amplifies any
    // error of w, so delta is formed in double throughout (uniform per image)
    if (R2d > 0.0) {
      const double f = r.d[EPL_F], f2 = f * f;
      const double inv = 1.0 / R2d;
      const double zr = -f * (Zrd * Zrd - Zid * Zid) * inv;
      const double zi = -f * (2.0 * Zrd * Zid) * inv;
      const double dd = 1.0 / fma(-4.0, zr, f2 + 4.0);
      const double wr = fma(-2.0, zr, f2) * dd;
      const double wi = -2.0 * zi * dd;
      double Pr, Pi;
      real_poly_at_complex<double>(coef_d + 1, (int)r.d[EPL_NMAX] - 1, wr, wi, Pr, Pi);
      const double nr = fma(wr, Pr, fma(-wi, Pi, 0.5 * zr));
      const double ni = fma(wr, Pi, fma(wi, Pr, 0.5 * zi));
      const double er = fma(-0.5, zr, 1.0), ei = -0.5 * zi;
      const double ed = 1.0 / fma(er, er, ei * ei);
      dr = float(fma(nr, er, ni * ei) * ed);
      di = float(fma(ni, er, -nr * ei) * ed);
    }
  } else if (nf > 1 && R2 > 0.f) {
    const float inv = __builtin_amdgcn_rcpf(R2);
    const float f = r[EPL_F];
    const float zr = -f * (Zr * Zr - Zi * Zi) * inv;
    const float zi = -f * (2.f * Zr * Zi) * inv;
    // w = z / (z - 2) with |z|^2 = f^2:  w = (f^2 - 2 z, -2 Im z) / (f^2 - 4 Re z + 4)
    const float f2 = r[EPL_F2];
    const float dd = __builtin_amdgcn_rcpf(fmaf(-4.f, zr, f2 + 4.f));
    const float wr = fmaf(-2.f, zr, f2) * dd;
    const float wi = -2.f * zi * dd;
    float Pr, Pi;
    real_poly_at_complex<float>(coef_f + 1, nf - 1, wr, wi, Pr, Pi);
    // numerator S - 1 + z/2 = w P + z/2, denominator 1 - z/2
    const float nr = fmaf(wr, Pr, fmaf(-wi, Pi, 0.5f * zr));
    const float ni = fmaf(wr, Pi, fmaf(wi, Pr, 0.5f * zi));
    const float er = fmaf(-0.5f, zr, 1.f), ei = -0.5f * zi;
    const float ed = __builtin_amdgcn_rcpf(fmaf(er, er, ei * ei));
    dr = fmaf(nr, er, ni * ei) * ed;
    di = fmaf(ni, er, -nr * ei) * ed;
  }
  // Omega = Z (1 + delta)
  const double Or = Zrd + double(fmaf(Zr, dr, -Zi * di));
  const double Oi = Zid + double(fmaf(Zr, di, Zi * dr));
  // 1/R with R = max(|Z|, 1e-9): one Newton step on the rsqrt, or 1e9 below the floor
  const bool above = R2d > 1e-18;
  double iR = double(iRs);
  iR = above ? iR * fma(-0.5 * R2d, iR * iR, 1.5) : 1e9;
  // (b/R)^(t-1)
  const float R2cf = floor_nonneg_f32(R2, 0x219392ef);  // max(R^2, 1e-18f); R2 >= 0
  const float lnbr = fmaf(-0.34657359f, __builtin_amdgcn_logf(R2cf), r[EPL_LNB]);  // ln b - ln(R^2)/2
  const float em = expm1_f32(r[EPL_TM1] * lnbr);
  const double pbr = r.d[EPL_PB] * iR;
  const double scale = fma(pbr, double(em), pbr);  // 2b/(1+q) / R (b/R)^(t-1)
  fr = scale * Or;
  fi = scale * Oi;
}

__device__ __forceinline__ void epl_defl_f32(const Rec<float> &r, const float *__restrict__ coef_f,
                                             const double *__restrict__ coef_d, double x, double y, double &ax,
                                             double &ay) {
  const double dx = x - r.d[R_CX], dy = y - r.d[R_CY];
  const double cd = r.d[EPL_C], sd = r.d[EPL_S];
  const double Zrd = r.d[EPL_Q] * fma(dx, cd, dy * sd);
  const double Zid = fma(-dx, sd, dy * cd);
  double fr, fi;
  epl_major_f32(r, coef_f, coef_d, Zrd, Zid, fma(Zrd, Zrd, Zid * Zid), fr, fi);
  ax += fr * cd - fi * sd;
  ay += fr * sd + fi * cd;
}

// NIE / SIE deflection, LensModel/Profiles/nie.py:86-107, :280-295, :333-342.
// DBL (every render and ray-tracing kernel of a model that holds an SIE / NIE
// or NFW deflector): the float form runs the FP64 form's arithmetic instead.
// Its float atan / atanh leave one-to-two-ulp random errors in alpha, which a
// compact cuspy source (Sersic n ~ 4, R_sersic 0.17") behind an NFW + SIE lens
// turned into a log-likelihood error of 1.8e-6 relative.  The kernels are
// instantiated both ways, so the double library atan / atanh cost registers
// only to the models that use them.  The gradient kernels keep the float form
// (their bar is 1e-4).
// The FP64 form's NIE in the major-axis frame at the rotated (xr, yr).
__device__ __forceinline__ void nie_major_dbl(const double *rd, double xr, double yr, double &fx, double &fy) {
  const double ss = rd[NIE_SS];
  const double psi = ::sqrt(rd[NIE_Q2] * (ss * ss + xr * xr) + yr * yr);
  const double sq = rd[NIE_SQ], bsq = rd[NIE_BSQ];
  fx = bsq * atan_dbl(sq * xr * rcp_refined(psi + ss));
  fy = bsq * atanh_dbl(sq * yr * rcp_refined(psi + rd[NIE_Q2S]));
}

template <typename T, bool DBL = false>
__device__ __forceinline__ void nie_defl(const Rec<T> &r, double x, double y, double &ax, double &ay) {
  const double dx = x - r.d[R_CX], dy = y - r.d[R_CY];
  const double cd = r.d[NIE_C], sd = r.d[NIE_S];
  if constexpr (DBL || sizeof(T) == sizeof(double)) {
    const double xr = dx * cd + dy * sd;
    const double yr = -dx * sd + dy * cd;
    double fx, fy;
    nie_major_dbl(r.d, xr, yr, fx, fy);
    ax += fx * cd - fy * sd;
    ay += fx * sd + fy * cd;
  } else {
    const T xr = T(dx * cd + dy * sd);
    const T yr = T(-dx * sd + dy * cd);
    const T ss = r[NIE_SS];
    const T psi = Fn<T>::sqrt_(r[NIE_Q2] * (ss * ss + xr * xr) + yr * yr);
    const T sq = r[NIE_SQ], bsq = r[NIE_BSQ];
    const T fx = bsq * Fn<T>::atan_(sq * xr / (psi + ss));
    const T fy = bsq * Fn<T>::atanh_(sq * yr / (psi + r[NIE_Q2S]));
    const T c = r[NIE_C], s = r[NIE_S];
    ax += double(fx * c - fy * s);
    ay += double(fx * s + fy * c);
  }
}

// atan(t) for 0 <= t <= 0.9 as t + t^3 P(t^2), P of degree 7 fitted to atan by
// iteratively reweighted least squares on the relative error: 1.3 ulp in float
// over the interval, with no range reduction, division or branch.
__device__ __forceinline__ float atan_unit_f32(float t) {
  const float s = t * t;
  float p = 0.00463163447f;
  p = fmaf(p, s, -0.0220576132f);
  p = fmaf(p, s, 0.0508678642f);
  p = fmaf(p, s, -0.0808524268f);
  p = fmaf(p, s, 0.108694514f);
  p = fmaf(p, s, -0.142524425f);
  p = fmaf(p, s, 0.199977465f);
  p = fmaf(p, s, -0.333332811f);
  return fmaf(t, p * s, t);
}

// asinh(z) for 0 <= z <= 0.5 as z + z^3 Q(z^2), Q of degree 4 fitted the same
// way: 1 ulp in float.
__device__ __forceinline__ float asinh_small_f32(float z) {
  const float s = z * z;
  float p = -0.0132715709f;
  p = fmaf(p, s, 0.0281910547f);
  p = fmaf(p, s, -0.0443894712f);
  p = fmaf(p, s, 0.0749869748f);
  p = fmaf(p, s, -0.166666449f);
  return fmaf(z, p * s, z);
}

// h(X) = acosh(1/X)/sqrt(1-X^2) (X<1), acos(1/X)/sqrt(X^2-1) (X>1), 1 at X=1.
// It equals sum_k v^k/(2k+1) with v = 1 - X^2; near X = 1 the float form uses
// that series because both closed branches cancel there.
__device__ __forceinline__ float nfw_h_f32(float X) {
  const float v = (1.0f - X) * (1.0f + X);
  if (fabsf(v) < 0.1f) {
    float s = 1.0f / 17.0f;
    s = s * v + 1.0f / 15.0f;
    s = s * v + 1.0f / 13.0f;
    s = s * v + 1.0f / 11.0f;
    s = s * v + 1.0f / 9.0f;
    s = s * v + 1.0f / 7.0f;
    s = s * v + 1.0f / 5.0f;
    s = s * v + 1.0f / 3.0f;
    s = s * v + 1.0f;
    return s;
  }
  if (v > 0.0f) {
    const float sv = __builtin_amdgcn_sqrtf(v);
    return __logf((1.0f + sv) / X) / sv;  // acosh(1/X) = ln((1 + sqrt(1-X^2))/X)
  }
  const float sv = __builtin_amdgcn_sqrtf(-v);
  if (X >= 1.5f) {
    // atan(sv) = pi/2 - atan(1/sv) with 1/sv <= 0.895 (no cancellation: the
    // result is at least 0.84)
    const float isv = __builtin_amdgcn_rcpf(sv);
    return (1.57079633f - atan_unit_f32(isv)) * isv;
  }
  return ::atanf(sv) / sv;  // acos(1/X) = atan(sqrt(X^2-1))
}

// NFW.g, LensModel/Profiles/nfw.py:345-365, exactly as the reference (FP64 form).
__device__ __forceinline__ double nfw_g(double X) {
  X = (X < 0.000001) ? 0.000001 : X;
  if (X < 1.0) return ::log(X / 2.0) + 1.0 / ::sqrt(1.0 - X * X) * ::acosh(1.0 / X);
  if (X == 1.0) return 1.0 + ::log(1.0 / 2.0);
  return ::log(X / 2.0) + 1.0 / ::sqrt(X * X - 1.0) * ::acos(1.0 / X);
}

// Float form of g.  For X < 0.8 the closed form ln(X/2) + h(X) cancels (both
// terms grow like ln(2/X) while g ~ X^2 ln(2/X) / 2), so it is rewritten with
// s = sqrt(1 - X^2), u = 1 - s = X^2 / (1 + s):
//   g = (u / s) ln(2 / X) + log1p(-u / 2) / s,
// which has no cancellation there.
__device__ __forceinline__ float nfw_g(float X) {
  X = (X < 0.000001f) ? 0.000001f : X;
  if (X < 0.8f) {
    const float s = __builtin_amdgcn_sqrtf((1.0f - X) * (1.0f + X));
    const float u = X * X / (1.0f + s);
    return (u * ::logf(2.0f / X) + ::log1pf(-0.5f * u)) / s;
  }
  return ::logf(0.5f * X) + nfw_h_f32(X);
}

// NFW deflection, LensModel/Profiles/nfw.py:86-105 with nfw_alpha :257-279.
// DBL (kernels of models that hold an NFW): the float form runs in double too.  The float
// g(X) = ln(X/2) + h(X) cancels around X ~ 1 (0.31 at X = 1 from terms of 0.7
// and 1), where an NFW main deflector's Einstein ring sits; with a compact
// cuspy source that moved the log-likelihood by 2.4e-6 relative.
// The FP64 form's NFW deflection at (dx, dy) from its centre: alpha = a (dx, dy).
__device__ __forceinline__ double nfw_factor_dbl(const double *rd, double dx, double dy) {
  double R = ::sqrt(dx * dx + dy * dy);
  R = R > 0.00000001 ? R : 0.00000001;
  // X = R / Rs by division: g(X) has an exact X == 1 branch and cancels
  // near it, so X must round as the reference's does.
  const double X = R / rd[NFW_RS];
  return rd[NFW_A] * nfw_g(X) / (X * X);
}

template <typename T, bool DBL = false>
__device__ __forceinline__ void nfw_defl(const Rec<T> &r, double x, double y, double &ax, double &ay) {
  const double dxd = x - r.d[R_CX], dyd = y - r.d[R_CY];
  if constexpr (DBL || sizeof(T) == sizeof(double)) {
    const double a = nfw_factor_dbl(r.d, dxd, dyd);
    ax += a * dxd;
    ay += a * dyd;
  } else {
    const T dx = T(dxd), dy = T(dyd);
    T R = Fn<T>::sqrt_(dx * dx + dy * dy);
    R = R > T(0.00000001) ? R : T(0.00000001);
    const T X = R * r[NFW_IRS];
    const T g = nfw_g(X);
    const T a = r[NFW_A] * g / (X * X);
    ax += double(a * dx);
    ay += double(a * dy);
  }
}

// TNFW.F, LensModel/Profiles/tnfw.py:288-304 (FP64 form; the division by
// sqrt(v) as a refined reciprocal, within two ulp of it).
__device__ __forceinline__ double tnfw_F(double x) {
  if (x < 1.0) {
    const double sv = ::sqrt(1.0 - x * x);
    return atanh_dbl(sv) * rcp_refined(sv);
  }
  if (x > 1.0) {
    const double sv = ::sqrt(x * x - 1.0);
    return atan_dbl(sv) * rcp_refined(sv);
  }
  return x;
}
__device__ __forceinline__ float tnfw_F(float x) { return nfw_h_f32(x); }

// TNFW deflection, LensModel/Profiles/tnfw.py:67-87, :226-248, _g :329-349,
// _L :276-286 (Baltz, Marshall & Oguri 2009).
template <typename T>
__device__ __forceinline__ void tnfw_defl(const Rec<T> &r, double x, double y, double &ax, double &ay) {
  const double dxd = x - r.d[R_CX], dyd = y - r.d[R_CY];
  const T dx = T(dxd), dy = T(dyd);
  T R = Fn<T>::sqrt_(dx * dx + dy * dy);
  const T rmin = r[TN_RMIN];
  R = R > rmin ? R : rmin;
  T X = R * r[TN_IRS];
  X = X > T(0.001) ? X : T(0.001);
  if (sizeof(T) == sizeof(float) && X < T(1)) {
    // Inside the scale radius the terms of g cancel to O(X^2 ln X).  With
    // s = sqrt(1 - X^2), w = sqrt(tau^2 + X^2) and F = atanh(s)/s, the bracket
    // of g equals X^2 C with
    //   C = (tau^2 - 1) [ ln X (1 - tau^2 - X^2) / ((w s + tau) tau s)
    //                     + ln(1 + s) / (s (1 + s)) - ln(tau + w) / (tau (tau + w))
    //                     - k log1p(-k X^2) / (-k X^2) ]  + 2 F - pi / (tau + w),
    //   k = (tau^2 + 1) / ((tau s + w)(tau + w)),
    // (from w s - tau = X^2 (1 - tau^2 - X^2) / (w s + tau), tau s - w =
    // -X^2 (tau^2 + 1) / (tau s + w), w - tau = X^2 / (w + tau)), so
    // alpha = 4 rho0 Rs P C (dx, dy) without cancellation or 1 / X^2.
    const float Xf = float(X), X2 = Xf * Xf;
    const float tau = float(r[TN_TAU]), tau2 = float(r[TN_TAU2]);
    const float sf = Fn<float>::sqrt_((1.0f - Xf) * (1.0f + Xf));
    const float w = Fn<float>::sqrt_(tau2 + X2);
    const float lnX = Xf > 0.5f ? ::log1pf(Xf - 1.0f) : ::logf(Xf);
    const float k = float(r[TN_T2P1]) * Fn<float>::rcp((tau * sf + w) * (tau + w));
    const float y = -k * X2;
    const float l1 = fabsf(y) < 1e-4f ? 1.0f + y * (-0.5f + y * (1.0f / 3.0f)) : ::log1pf(y) / y;
    const float itw = Fn<float>::rcp(tau + w);
    const float C = float(r[TN_T2M1]) * (lnX * (1.0f - tau2 - X2) * Fn<float>::rcp((w * sf + tau) * tau * sf) +
                                         ::log1pf(sf) * Fn<float>::rcp(sf * (1.0f + sf)) -
                                         ::logf(tau + w) * itw / tau - l1 * k) +
                    2.0f * nfw_h_f32(Xf) - float(M_PI) * itw;
    const float a = float(r[TN_A]) * float(r[TN_PRE]) * C;  // 4 rho0 Rs g / X^2
    ax += double(a * float(dx));
    ay += double(a * float(dy));
    return;
  }
  if (sizeof(T) == sizeof(float) && X >= T(1.5)) {
    // Outside the scale radius (2 X^2 + tau^2 - 1) F(X) and -pi sqrt(tau^2 + X^2)
    // both grow like pi X and cancel to O(1).  With v = sqrt(X^2 - 1),
    // w = sqrt(X^2 + tau^2) and F = (pi/2 - atan(1/v)) / v, their sum is
    //   pi (tau^2 + 1)^2 / (2 v (v + w)^2) - (2 X^2 + tau^2 - 1) atan(1/v) / v
    // (from (w - v)^2 = (tau^2 + 1)^2 / (v + w)^2), and
    // L = ln(X / (tau + w)) = -asinh(z), z = tau / X: a series for z <= 0.5
    // (where the logarithm's argument nears 1), else ln((tau + w) / X) with its
    // argument at least 1.6.  1/v <= 0.895 here, inside atan_unit_f32's range.
    const float Xf = float(X), tau2 = float(r[TN_TAU2]), tau = float(r[TN_TAU]);
    const float X2 = Xf * Xf;
    const float v = Fn<float>::sqrt_((Xf - 1.0f) * (Xf + 1.0f));
    const float w = Fn<float>::sqrt_(tau2 + X2);
    const float iv = Fn<float>::rcp(v);
    const float vw = v + w;
    const float A = float(r[TN_K]) * iv * Fn<float>::rcp(vw * vw) -
                    (2.0f * X2 + float(r[TN_T2M1])) * atan_unit_f32(iv) * iv;
    const float iX = Fn<float>::rcp(Xf);
    const float z = tau * iX;
    float ash = asinh_small_f32(z);
    // rays far from the subhalo (z <= 0.5) are the common case: a wave whose
    // lanes all are skips the logarithm (a uniform branch)
    if (!__all(z <= 0.5f)) {
      asm volatile("");  // keeps the logarithm behind the branch
      if (z > 0.5f) ash = 0.693147181f * __builtin_amdgcn_logf((tau + w) * iX);
    }
    const float g = float(r[TN_PRE]) * (A + float(r[TN_C0]) - w * ash * float(r[TN_C1]));
    const float a = float(r[TN_A]) * g * (iX * iX);
    ax += double(a * float(dx));
    ay += double(a * float(dy));
    return;
  }
  const T tau = r[TN_TAU], tau2 = r[TN_TAU2];
  const T X2 = X * X;
  const T sqtx = Fn<T>::sqrt_(tau2 + X2);
  // the FP64 form's divisions as refined reciprocals (within two ulp)
  constexpr bool kDbl = sizeof(T) == sizeof(double);
  // (the FP64 logarithm of the positive ratio by log2_pos)
  const T L = kDbl ? T(log2_pos(double(X) * rcp_refined(double(tau + sqtx))) * 0.6931471805599453)
                   : Fn<T>::log_(X / (tau + sqtx));
  const T g = r[TN_PRE] * ((r[TN_T2P1] + T(2) * (X2 - T(1))) * tnfw_F(X) + r[TN_C0] +
                           sqtx * (T(-M_PI) + L * r[TN_C1]));
  const T a = kDbl ? r[TN_A] * g * T(rcp_refined(double(X2))) : r[TN_A] * g / X2;
  ax += double(a * dx);
  ay += double(a * dy);
}

// External shear, LensModel/Profiles/shear.py:37-53 (linear: kept in double).
__device__ __forceinline__ void shear_defl(const double *rd, double x, double y, double &ax, double &ay) {
  const double dx = x - rd[R_CX], dy = y - rd[R_CY];
  const double g1 = rd[SH_G1], g2 = rd[SH_G2];
  ax += g1 * dx + g2 * dy;
  ay += g2 * dx - g1 * dy;
}

// Mass sheet, LensModel/Profiles/convergence.py:34-46 (linear: kept in double).
__device__ __forceinline__ void convergence_defl(const double *rd, double x, double y, double &ax, double &ay) {
  ax += rd[CV_KAPPA] * (x - rd[R_CX]);
  ay += rd[CV_KAPPA] * (y - rd[R_CY]);
}

// Adds one deflector's alpha at (x, y).  coef_d / coef_t: this profile's EPL
// series coefficients (NULL for other profiles).
// HALO: whether the NFW / TNFW branches are compiled in (kernels for models
// without them keep fewer registers).  DBL: SIE / NIE and NFW run the FP64
// form's arithmetic in the float form too (kernels for models that hold
// SIE / NIE or NFW, or halos on several planes; see nie_defl and nfw_defl);
// DBL_TNFW: TNFW too (the render and gradient kernels; the ray-shooting
// kernels keep the float TNFW, whose rays stay within 2e-6 of the FP64 form's).
template <typename T, bool HALO, bool DBL = false, bool DBL_TNFW = DBL>
__device__ __forceinline__ void add_deflection(int type, const Rec<T> &r, const double *__restrict__ coef_d,
                                               const T *__restrict__ coef_t, double x, double y, double &ax,
                                               double &ay) {
  switch (type) {
    case LENSIM_EPL:
      if constexpr (sizeof(T) == sizeof(float))
        epl_defl_f32(r, coef_t, coef_d, x, y, ax, ay);
      else
        epl_defl_f64(r, coef_d, x, y, ax, ay);
      break;
    case LENSIM_SIE:
    case LENSIM_NIE:
      nie_defl<T, DBL>(r, x, y, ax, ay);
      break;
    case LENSIM_NFW:
      if constexpr (HALO) nfw_defl<T, DBL>(r, x, y, ax, ay);
      break;
    case LENSIM_TNFW:
      if constexpr (HALO) {
        if constexpr (DBL_TNFW) tnfw_defl<double>(Rec<double>{r.d, r.d}, x, y, ax, ay);
        else tnfw_defl<T>(r, x, y, ax, ay);
      }
      break;
    case LENSIM_SHEAR:
      shear_defl(r.d, x, y, ax, ay);
      break;
    case LENSIM_CONVERGENCE:
      convergence_defl(r.d, x, y, ax, ay);
      break;
    default:
      break;
  }
}

// ---------------------------------------------------------- light --------

// Sersic kernel, LensModel/Profiles/sersic_utils.py:239-267 (FP64 form), on
// R^2: R_ = max(1e-4, R), frac = R_ / R_sersic_, frac^(1/n) = exp((ln R_ -
// ln R_sersic_) / n), zero beyond frac = 1000, NaN -> 0.
// In base 2 (log2_pos, exp2_fin): frac^(1/n) = 2^((log2(R^2) / 2 - log2 R_s) / n)
// and exp(b_n (1 - frac^(1/n))) = 2^(K2 (1 - frac^(1/n))).
__device__ __forceinline__ double sersic_kernel(double R2, const Rec<double> &r) {
  const double R2c = fmax(1e-8, R2);
  const double pw = exp2_fin((0.5 * log2_pos(R2c) - r.d[SE_LNRS] * 1.4426950408889634) * r.d[SE_INVN]);
  double v = (R2c <= r.d[SE_R2MAX]) ? exp2_fin(r.d[SE_K2] * (1.0 - pw)) : 0.0;
  if (v != v) v = 0.0;
  if (v == INFINITY) v = 1.7976931348623157e308;
  return v;
}

// The Sersic profile exp(b_n t) = 2^(K2 t), t = 1 - frac^(1/n), in float to
// ~2 ulp: K2 = b_n log2(e) is a float head + tail pair (rf[SE_K2],
// rf[SE_K2LO]), the head product's rounding error is kept, and the result is
// 2^hi (1 + lo ln 2), so the exponent carries no one-signed rounding error.
__device__ __forceinline__ float sersic_exp2_f32(const float *rf, float t) {
  const float hi = rf[SE_K2] * t;
  const float lo = fmaf(rf[SE_K2], t, -hi) + rf[SE_K2LO] * t;
  return __builtin_amdgcn_exp2f(hi) * fmaf(lo, 0.6931471805599453f, 1.f);
}

// Float form.  Avoided on purpose, because each would bias every value of a
// profile with one sign (the image sums many of them, so a bias of 1e-7 moves
// a log-likelihood of a poor fit by about 1e-6 relative):
//  * float copies of the ellipse matrix, 1/R_sersic^2 and C1: the elliptical
//    R^2, the ratio u = R^2 / R_sersic^2 and C1 log2 u are formed in double;
//  * v_log_f32 of a large |log2 u|: gfx950's v_log_f32 reads about 0.45 ulp
//    low on average, so the logarithm is split as exponent + log2 of the
//    mantissa in [0.5, 1), and only the small second term is approximated;
//  * a float b_n: the exponent b_n log2(e) (1 - frac^(1/n)) takes b_n log2(e)
//    as a head + tail pair with the head product's rounding error.
// frac^(1/n) = exp2(C1 log2 u) and exp of the exponent are then float
// (v_exp_f32 is unbiased to 0.03 ulp).  Window: R^2 <= (1000 R_sersic)^2.
__device__ __forceinline__ float sersic_value_f32(double dx, double dy, const Rec<float> &r) {
  const double xe = fma(r.d[SE_M00], dx, r.d[SE_M01] * dy);
  const double ye = fma(r.d[SE_M01], dx, r.d[SE_M11] * dy);
  const double R2 = fmax(1e-8, fma(xe, xe, ye * ye));
  const float u = float(R2 * r.d[SE_IRS2]);
  const float m = __builtin_amdgcn_frexp_mantf(u);  // u = m 2^ex, m in [0.5, 1)
  const int ex = __builtin_amdgcn_frexp_expf(u);
  const double L = double(ex) + double(__builtin_amdgcn_logf(m));
  const float pw = __builtin_amdgcn_exp2f(float(r.d[SE_C1] * L));
  return (R2 <= r.d[SE_R2MAX]) ? sersic_exp2_f32(r.t, 1.f - pw) : 0.f;
}

// The Sersic kernel (no amplitude) at the scaled elliptical coordinates
// (xs, ys) = M (x - c) / R_sersic_, i.e. at u = R^2 / R_sersic_^2 (the
// compile-time render's form: the matrix and 1/R_sersic_ are folded, and the
// lens light's (xs, ys) advance by a constant step per sub-ray).  frac^(1/n)
// = u^(1/(2n)); the floors and the window as sersic_kernel, on u.
// umin: rd[SE_UMIN], which a caller may hoist (canonicalised) out of its loop.
// In base 2: frac^(1/n) = 2^(C1 log2 u) and exp(b_n (1 - frac^(1/n))) =
// 2^(K2 (1 - frac^(1/n))), K2 = b_n log2(e), by log2_pos and exp2_fin.
__device__ __forceinline__ double sersic_u_f64(const double *rd, double xs, double ys, double umin) {
  const double u = fmax(fma(xs, xs, ys * ys), umin);
  const double pw = exp2_fin(log2_pos(u) * rd[SE_C1]);
  double v = (u <= 1e6) ? exp2_fin(rd[SE_K2] * (1.0 - pw)) : 0.0;
  if (v != v) v = 0.0;
  if (v == INFINITY) v = 1.7976931348623157e308;
  return v;
}

// Float form of sersic_u_f64 with sersic_value_f32's unbiased arithmetic.  u
// is formed in double and floored in float (umin_bits: the bits of
// float(rd[SE_UMIN]); the floor is monotone, so u below the double floor gets
// the float floor); the window u <= 1e6 is tested in float (1e6 is a float).
__device__ __forceinline__ float sersic_u_f32(const double *rd, const float *rf, double xs, double ys,
                                              int umin_bits) {
  const float uf = floor_nonneg_f32(float(fma(xs, xs, ys * ys)), umin_bits);
  const float m = __builtin_amdgcn_frexp_mantf(uf);  // u = m 2^ex, m in [0.5, 1)
  const int ex = __builtin_amdgcn_frexp_expf(uf);
  const double L = double(ex) + double(__builtin_amdgcn_logf(m));
  const float pw = __builtin_amdgcn_exp2f(float(rd[SE_C1] * L));
  return (uf <= 1e6f) ? sersic_exp2_f32(rf, 1.f - pw) : 0.f;
}

// One light profile at (x, y): SERSIC_ELLIPSE / SERSIC (sersic_ellipse.py:62-93,
// sersic.py:53-73) and GAUSSIAN (gaussian.py:27-42).
// SERSIC / SERSIC_ELLIPSE at (x, y) (the circular SERSIC's record has e1 = e2 = 0).
template <typename T>
__device__ __forceinline__ double sersic_light(const Rec<T> &r, double x, double y) {
  const double dxd = x - r.d[R_CX], dyd = y - r.d[R_CY];
  if constexpr (sizeof(T) == sizeof(float)) {
    // the amplitude multiplies in double (a float amp would bias the profile by its rounding)
    return r.d[SE_AMP] * double(sersic_value_f32(dxd, dyd, r));
  } else {
    const T dx = T(dxd), dy = T(dyd);
    const T xe = r[SE_M00] * dx + r[SE_M01] * dy;
    const T ye = r[SE_M01] * dx + r[SE_M11] * dy;
    return r.d[SE_AMP] * sersic_kernel(xe * xe + ye * ye, r);
  }
}

template <typename T>
__device__ __forceinline__ double light_value(int type, const Rec<T> &r, double x, double y) {
  const double dxd = x - r.d[R_CX], dyd = y - r.d[R_CY];
  const T dx = T(dxd), dy = T(dyd);
  switch (type) {
    case LENSIM_SERSIC:
    case LENSIM_SERSIC_ELLIPSE:
      return sersic_light<T>(r, x, y);
    case LENSIM_GAUSSIAN: {
      const T r2 = dx * dx * r[GA_IS2] + dy * dy * r[GA_IS2];
      return double(r[GA_C] * Fn<T>::exp_(-r2 / T(2)));
    }
    default:
      return 0.0;
  }
}

}  // namespace lensim
