// Exchange-correlation functionals by libxc id, for the device (and host):
//   1 lda_x, 7 lda_c_vwn (VWN5), 8 lda_c_vwn_rpa, 12 lda_c_pw (PW92 as
//   printed), 13 lda_c_pw_mod (libxc's more precise PW92 constants),
//   101 gga_x_pbe, 106 gga_x_b88, 130 gga_c_pbe, 131 gga_c_lyp,
//   202 mgga_x_tpss, 231 mgga_c_tpss, and the global hybrids
//   402 hyb_gga_xc_b3lyp (0.08 lda_x + 0.72 gga_x_b88 + 0.19 lda_c_vwn_rpa +
//   0.81 gga_c_lyp; 0.20 exact exchange) and 406 hyb_gga_xc_pbeh (PBE0:
//   0.75 gga_x_pbe + gga_c_pbe; 0.25 exact exchange), and the range-separated
//   433 hyb_gga_xc_cam_b3lyp and 470 hyb_gga_xc_camy_b3lyp (0.35 gga_x_b88 +
//   0.46 short-range B88 + 0.19 lda_c_vwn + 0.81 gga_c_lyp; 0.65 exact
//   exchange and -0.46 of its short-range part, erfc at omega 0.33 or Yukawa
//   at omega 0.34; short-range B88 with the erfc attenuation of Iikura et al.,
//   JCP 115, 3540 (2001), or the Yukawa one of Akinaga and Ten-no, CPL 462,
//   348 (2008))
// from the published formulas (Slater; Vosko-Wilk-Nusair, Can. J. Phys. 58,
// 1200 (1980), QMC and RPA fits; Perdew-Wang, PRB 45, 13244 (1992);
// Becke, PRA 38, 3098 (1988); Lee-Yang-Parr in the closed form of Miehlich
// et al., CPL 157, 200 (1989); Perdew-Burke-Ernzerhof, PRL 77, 3865 (1996);
// Tao-Perdew-Staroverov-Scuseria, PRL 91, 146401 (2003))
// with libxc's constants and thresholds: no output below the density
// threshold, 1 +- zeta floored at DBL_EPSILON, sigma_aa and sigma_bb floored
// at 1e-40, sigma_ab clamped to +-(sigma_aa + sigma_bb)/2; a polarized spin
// below the threshold adds no PBE or TPSS exchange; for TPSS tau floored at
// 1e-20 and sigma_ss capped at 8 rho_s tau_s (z <= 1).
//
// Each functional is one energy density per volume f(rho_a, rho_b, sigma_aa,
// sigma_ab, sigma_bb, tau_a, tau_b) written once over a scalar type; the
// potentials come from forward-mode differentiation (a value and one
// directional derivative, one pass per input), so the derivative formulas
// are never written by hand.
#pragma once
#include <cfloat>
#include <cmath>

namespace hfg {
namespace xcf {

// value and derivative along one input
struct Dl {
  double v, d;
};
__host__ __device__ inline Dl mk(double v, double d = 0.0) { return Dl{v, d}; }
__host__ __device__ inline Dl operator+(Dl a, Dl b) { return {a.v + b.v, a.d + b.d}; }
__host__ __device__ inline Dl operator-(Dl a, Dl b) { return {a.v - b.v, a.d - b.d}; }
__host__ __device__ inline Dl operator-(Dl a) { return {-a.v, -a.d}; }
__host__ __device__ inline Dl operator*(Dl a, Dl b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__host__ __device__ inline Dl operator/(Dl a, Dl b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__host__ __device__ inline Dl operator+(Dl a, double b) { return {a.v + b, a.d}; }
__host__ __device__ inline Dl operator+(double a, Dl b) { return {a + b.v, b.d}; }
__host__ __device__ inline Dl operator-(Dl a, double b) { return {a.v - b, a.d}; }
__host__ __device__ inline Dl operator-(double a, Dl b) { return {a - b.v, -b.d}; }
__host__ __device__ inline Dl operator*(Dl a, double b) { return {a.v * b, a.d * b}; }
__host__ __device__ inline Dl operator*(double a, Dl b) { return {a * b.v, a * b.d}; }
__host__ __device__ inline Dl operator/(Dl a, double b) { return {a.v / b, a.d / b}; }
__host__ __device__ inline Dl operator/(double a, Dl b) {
  const double q = a / b.v;
  return {q, -q * b.d / b.v};
}
__host__ __device__ inline Dl sqrt(Dl a) {
  const double s = ::sqrt(a.v);
  return {s, 0.5 * a.d / s};
}
__host__ __device__ inline Dl cbrt(Dl a) {
  const double c = ::cbrt(a.v);
  return {c, a.d / (3.0 * c * c)};
}
__host__ __device__ inline Dl log(Dl a) { return {::log(a.v), a.d / a.v}; }
__host__ __device__ inline Dl exp(Dl a) {
  const double e = ::exp(a.v);
  return {e, e * a.d};
}
__host__ __device__ inline Dl atan(Dl a) { return {::atan(a.v), a.d / (1.0 + a.v * a.v)}; }
__host__ __device__ inline Dl asinh(Dl a) { return {::asinh(a.v), a.d / ::sqrt(1.0 + a.v * a.v)}; }
__host__ __device__ inline Dl erf(Dl a) { return {::erf(a.v), a.d * 1.1283791670955126 * ::exp(-a.v * a.v)}; }
__host__ __device__ inline double val(Dl a) { return a.v; }
__host__ __device__ inline double val(double a) { return a; }

template <class T> __host__ __device__ inline T floor_at(T x, double lo) { return val(x) < lo ? T(x * 0.0 + lo) : x; }
__host__ __device__ inline Dl floor_at(Dl x, double lo) { return x.v < lo ? mk(lo) : x; }

constexpr double kPi = 3.14159265358979323846;

// x^(4/3) and x^(2/3)
template <class T> __host__ __device__ inline T p43(T x) { return x * cbrt(x); }
template <class T> __host__ __device__ inline T p23(T x) {
  const T c = cbrt(x);
  return c * c;
}

// VWN interpolation formula in x = sqrt(rs)
template <class T> __host__ __device__ inline T vwn_F(double A, double b, double c, double x0, T x) {
  const double Q = ::sqrt(4.0 * c - b * b), X0 = x0 * x0 + b * x0 + c;
  const double f1 = 2.0 * b / Q, f2 = b * x0 / X0, f3 = 2.0 * (2.0 * x0 + b) / Q;
  const T X = x * x + b * x + c;
  const T xm = x - x0;
  return A * (log(x * x / X) + (f1 - f2 * f3) * atan(Q / (2.0 * x + b)) - f2 * log(xm * xm / X));
}

// spin interpolation f(zeta) from 1 + zeta and 1 - zeta
template <class T> __host__ __device__ inline T fzeta(T opz, T omz) {
  return (p43(opz) + p43(omz) - 2.0) / (2.0 * ::cbrt(2.0) - 2.0);
}

// PW92: G(rs) per spin channel k (paramagnetic, ferromagnetic, -spin stiffness)
template <class T> __host__ __device__ inline T pw92_G(int k, bool mod, T rs, T srs) {
  const double A = mod ? (k == 0 ? 0.0310907 : k == 1 ? 0.01554535 : 0.0168869)
                       : (k == 0 ? 0.031091 : k == 1 ? 0.015545 : 0.016887);
  const double a1 = k == 0 ? 0.21370 : k == 1 ? 0.20548 : 0.11125;
  const double b1 = k == 0 ? 7.5957 : k == 1 ? 14.1189 : 10.357;
  const double b2 = k == 0 ? 3.5876 : k == 1 ? 6.1977 : 3.6231;
  const double b3 = k == 0 ? 1.6382 : k == 1 ? 3.3662 : 0.88026;
  const double b4 = k == 0 ? 0.49294 : k == 1 ? 0.62517 : 0.49671;
  const T den = srs * (b1 + b3 * rs) + rs * (b2 + b4 * rs);
  return -2.0 * A * (1.0 + a1 * rs) * log(1.0 + 1.0 / (2.0 * A * den));
}

// PW92 correlation energy per particle
template <class T> __host__ __device__ inline T pw92(bool mod, T rs, T z, T opz, T omz) {
  const double fz20 = mod ? 1.709920934161365617563962776245 : 1.709921;
  const T srs = sqrt(rs);
  const T g0 = pw92_G(0, mod, rs, srs), g1 = pw92_G(1, mod, rs, srs), g2 = pw92_G(2, mod, rs, srs);
  const T fz = fzeta(opz, omz), z2 = z * z;
  return g0 + z2 * z2 * fz * (g1 - g0 + g2 / fz20) - fz * g2 / fz20;
}

// PBE exchange energy per volume of an unpolarized density n, |grad n|^2 = s2n
template <class T> __host__ __device__ inline T pbe_x(T n, T s2n) {
  const double kappa = 0.8040, mu = 0.06672455060314922 * kPi * kPi / 3.0;
  const double c2 = 4.0 * ::cbrt(9.0 * kPi * kPi * kPi * kPi);  // 4 (3 pi^2)^(2/3)
  const T n43 = p43(n);
  const T s2 = s2n / (c2 * n43 * n43);
  return -0.75 * ::cbrt(3.0 / kPi) * n43 * (1.0 + kappa - kappa / (1.0 + (mu / kappa) * s2));
}

// PBE correlation per volume on PW92 (modified constants)
template <class T> __host__ __device__ inline T pbe_c(T ra, T rb, T saa, T sab, T sbb) {
  const double beta = 0.06672455060314922, gamma = (1.0 - 0.69314718055994530942) / (kPi * kPi);
  saa = floor_at(saa, 1e-40);
  sbb = floor_at(sbb, 1e-40);
  const T sav = 0.5 * (saa + sbb);
  if (val(sab) > val(sav)) sab = sav;
  if (val(sab) < -val(sav)) sab = -sav;
  const T rho = ra + rb;
  const T z = (ra - rb) / rho;
  const T opz = floor_at(1.0 + z, DBL_EPSILON), omz = floor_at(1.0 - z, DBL_EPSILON);
  const T rs = cbrt(3.0 / (4.0 * kPi * rho));
  const T ec = pw92(true, rs, z, opz, omz);
  const T phi = 0.5 * (p23(opz) + p23(omz)), phi3 = phi * phi * phi;
  const T ks2 = (4.0 / kPi) * cbrt(3.0 * kPi * kPi * rho);
  const T t2 = (saa + 2.0 * sab + sbb) / (4.0 * phi * phi * ks2 * rho * rho);
  const T A = (beta / gamma) / (exp(-ec / (gamma * phi3)) - 1.0);
  const T At2 = A * t2;
  const T H = gamma * phi3 * log(1.0 + (beta / gamma) * t2 * (1.0 + At2) / (1.0 + At2 + At2 * At2));
  return rho * (ec + H);
}

// TPSS exchange per volume of an unpolarized density n, |grad n|^2 = s2n,
// kinetic energy density tn
template <class T> __host__ __device__ inline T tpss_x(T n, T s2n, T tn) {
  const double kappa = 0.804, b = 0.40, c = 1.59096, e = 1.537, mu = 0.21951;
  const double ckf = ::cbrt(9.0 * kPi * kPi * kPi * kPi);  // (3 pi^2)^(2/3)
  tn = floor_at(tn, 1e-20);
  if (val(s2n) > val(8.0 * n * tn)) s2n = 8.0 * n * tn;
  const T n43 = p43(n);
  const T p = s2n / (4.0 * ckf * n43 * n43);
  const T tw = s2n / (8.0 * n);
  const T z = tw / tn;
  const T alpha = (tn - tw) / (0.3 * ckf * n43 * cbrt(n));
  const T qb = 0.45 * (alpha - 1.0) / sqrt(1.0 + b * alpha * (alpha - 1.0)) + (2.0 / 3.0) * p;
  const T z35 = 0.6 * z, zz = 1.0 + z * z, se = ::sqrt(e) * p + 1.0;
  const T x = ((10.0 / 81.0 + c * z * z / (zz * zz)) * p + (146.0 / 2025.0) * qb * qb -
               (73.0 / 405.0) * qb * sqrt(0.5 * z35 * z35 + 0.5 * p * p) + (100.0 / 6561.0 / kappa) * p * p +
               (20.0 / 81.0) * ::sqrt(e) * z35 * z35 + e * mu * p * p * p) /
              (se * se);
  return -0.75 * ::cbrt(3.0 / kPi) * n43 * (1.0 + kappa - kappa / (1.0 + x / kappa));
}

// TPSS correlation per volume: revPKZB built on PBE correlation
template <class T>
__host__ __device__ inline T tpss_c(T ra, T rb, T saa, T sab, T sbb, T ta, T tb, double thr) {
  const double d = 2.8;
  ta = floor_at(ta, 1e-20);
  tb = floor_at(tb, 1e-20);
  saa = floor_at(saa, 1e-40);
  sbb = floor_at(sbb, 1e-40);
  if (val(saa) > val(8.0 * ra * ta)) saa = 8.0 * ra * ta;
  if (val(sbb) > val(8.0 * rb * tb)) sbb = 8.0 * rb * tb;
  const T sav = 0.5 * (saa + sbb);
  if (val(sab) > val(sav)) sab = sav;
  if (val(sab) < -val(sav)) sab = -sav;
  const T zero = ra * 0.0;
  const T n = ra + rb;
  const T z = (saa + 2.0 * sab + sbb) / (8.0 * n * (ta + tb));
  const T zeta = (ra - rb) / n;
  const T opz = floor_at(1.0 + zeta, DBL_EPSILON), omz = floor_at(1.0 - zeta, DBL_EPSILON);
  const T eps = pbe_c(ra, rb, saa, sab, sbb) / n;
  T sum = zero;
  if (val(ra) >= thr) {
    T ea = pbe_c(ra, zero, saa, zero, zero) / ra;
    if (val(ea) < val(eps)) ea = eps;
    sum = sum + ra / n * ea;
  }
  if (val(rb) >= thr) {
    T eb = pbe_c(rb, zero, sbb, zero, zero) / rb;
    if (val(eb) < val(eps)) eb = eps;
    sum = sum + rb / n * eb;
  }
  const T gz2 = (omz * omz * saa - 2.0 * omz * opz * sab + opz * opz * sbb) / (n * n);
  const T xi2 = gz2 / (4.0 * p23(3.0 * kPi * kPi * n));
  const T z2 = zeta * zeta;
  const T C0 = 0.53 + z2 * (0.87 + z2 * (0.50 + 2.26 * z2));
  const T den = 1.0 + 0.5 * xi2 * (1.0 / p43(opz) + 1.0 / p43(omz));
  const T den2 = den * den;
  const T C = C0 / (den2 * den2);
  const T erev = eps * (1.0 + C * z * z) - (1.0 + C) * z * z * sum;
  return n * erev * (1.0 + d * erev * z * z * z);
}

// Slater exchange per volume, spin-scaled
template <class T> __host__ __device__ inline T slater_x(T ra, T rb) {
  T e = ra * 0.0;
  if (val(ra) > 0.0) e = e + p43(ra);
  if (val(rb) > 0.0) e = e + p43(rb);
  return -0.75 * ::cbrt(6.0 / kPi) * e;
}

// Becke 88 exchange per volume of one spin channel (density r, sigma s):
// -r^(4/3) [C_x + beta x^2 / (1 + 6 beta x asinh x)], x = sqrt(s) / r^(4/3)
template <class T> __host__ __device__ inline T b88_x(T r, T s) {
  const double beta = 0.0042, cx = 0.75 * ::cbrt(6.0 / kPi);
  const T r43 = p43(r), x = sqrt(s) / r43;
  return -1.0 * r43 * (cx + beta * x * x / (1.0 + 6.0 * beta * x * asinh(x)));
}

template <class T> __host__ __device__ inline T b88_pair(T ra, T rb, T saa, T sbb, double thr, bool pol) {
  T e = ra * 0.0;
  if (!pol || val(ra) >= thr) e = e + b88_x(ra, floor_at(saa, 1e-40));
  if (!pol || val(rb) >= thr) e = e + b88_x(rb, floor_at(sbb, 1e-40));
  return e;
}

// PBE exchange of both spin channels by spin scaling
template <class T> __host__ __device__ inline T pbe_x_pair(T ra, T rb, T saa, T sbb, double thr, bool pol) {
  T e = ra * 0.0;
  if (!pol || val(ra) >= thr) e = e + 0.5 * pbe_x(2.0 * ra, 4.0 * floor_at(saa, 1e-40));
  if (!pol || val(rb) >= thr) e = e + 0.5 * pbe_x(2.0 * rb, 4.0 * floor_at(sbb, 1e-40));
  return e;
}

// Short-range attenuation of the uniform-gas exchange of one spin channel,
// a = omega / (2 k_F): (8/3) int_0^1 w(x) (1 - 3x/2 + x^3/2) dx with w = 1 -
// exp(-x^2 / 4a^2) (erfc kernel) or x^2 / (x^2 + a^2) (Yukawa kernel
// exp(-omega r)/r).  Closed forms for small a; for large a, where they cancel,
// the series of the integral sum_n (-1)^(n+1) B_n c_n with
// B_n = 1/(2n+1) - 3/(4n+4) + 1/(4n+8), c_n = (4a^2)^-n / n! or a^-2n.
__host__ __device__ inline double series_B(int n) {
  return 1.0 / (2 * n + 1) - 3.0 / (4 * n + 4) + 1.0 / (4 * n + 8);
}
template <class T> __host__ __device__ inline T att_erfc(T a) {
  const T a2 = a * a;
  if (val(a) < 0.5)
    return 1.0 - (8.0 / 3.0) * a *
                     (1.7724538509055159 * erf(0.5 / a) + (2.0 * a - 4.0 * a2 * a) * exp(-0.25 / a2) - 3.0 * a +
                      4.0 * a2 * a);
  const T y = 0.25 / a2;
  T sum = a * 0.0, p = a * 0.0 + 1.0;
  double fact = 1.0;
  for (int n = 1; n <= 24; ++n) {
    p = p * y;
    fact *= n;
    sum = sum + ((n & 1) ? 1.0 : -1.0) * series_B(n) / fact * p;
  }
  return (8.0 / 3.0) * sum;
}
template <class T> __host__ __device__ inline T att_yukawa(T a) {
  const T a2 = a * a;
  if (val(a) < 2.0) return 1.0 - (8.0 / 3.0) * a * (atan(1.0 / a) + 0.25 * a * (1.0 - (a2 + 3.0) * log(1.0 + 1.0 / a2)));
  const T y = 1.0 / a2;
  T sum = a * 0.0, p = a * 0.0 + 1.0;
  for (int n = 1; n <= 40; ++n) {
    p = p * y;
    sum = sum + ((n & 1) ? 1.0 : -1.0) * series_B(n) * p;
  }
  return (8.0 / 3.0) * sum;
}

// Short-range B88 of one spin channel: the B88 enhancement F(x) times the
// attenuation at a = omega / (2 k), k = sqrt(9 pi / (2 C_x F)) r^(1/3)
template <class T> __host__ __device__ inline T sr_b88_x(T r, T s, double omega, bool yukawa) {
  const double beta = 0.0042, cx = 0.75 * ::cbrt(6.0 / kPi);
  const T r13 = cbrt(r), r43 = r * r13, x = sqrt(s) / r43;
  const T F = 1.0 + (beta / cx) * x * x / (1.0 + 6.0 * beta * x * asinh(x));
  const T k = sqrt(9.0 * kPi / (2.0 * cx * F)) * r13;
  const T a = omega / (2.0 * k);
  return -cx * r43 * F * (yukawa ? att_yukawa(a) : att_erfc(a));
}

template <class T> __host__ __device__ inline T sr_b88_pair(T ra, T rb, T saa, T sbb, double omega, bool yukawa,
                                                           double thr, bool pol) {
  T e = ra * 0.0;
  if (!pol || val(ra) >= thr) e = e + sr_b88_x(ra, floor_at(saa, 1e-40), omega, yukawa);
  if (!pol || val(rb) >= thr) e = e + sr_b88_x(rb, floor_at(sbb, 1e-40), omega, yukawa);
  return e;
}

// VWN5 per volume (QMC fit with the spin-stiffness interpolation)
template <class T> __host__ __device__ inline T vwn5_c(T rho, T rs, T z, T opz, T omz) {
  const T x = sqrt(rs);
  const T eP = vwn_F(0.0310907, 3.72744, 12.9352, -0.10498, x);
  const T eF = vwn_F(0.01554535, 7.06042, 18.0578, -0.32500, x);
  const T eA = vwn_F(-1.0 / (6.0 * kPi * kPi), 1.13107, 13.0045, -0.0047584, x);
  const double fpp = 4.0 / (9.0 * (::cbrt(2.0) - 1.0));
  const T fz = fzeta(opz, omz), z2 = z * z, z4 = z2 * z2;
  return rho * (eP + eA * fz * (1.0 - z4) / fpp + (eF - eP) * fz * z4);
}

// VWN RPA fit per volume: paramagnetic and ferromagnetic curves mixed by f(zeta)
template <class T> __host__ __device__ inline T vwn_rpa_c(T rho, T rs, T opz, T omz) {
  const T x = sqrt(rs);
  const T eP = vwn_F(0.0310907, 13.0720, 42.7198, -0.409286, x);
  const T eF = vwn_F(0.01554535, 20.1231, 101.578, -0.743294, x);
  return rho * (eP + (eF - eP) * fzeta(opz, omz));
}

// LYP correlation per volume (Miehlich et al. closed form) on the spin
// densities rebuilt from rho and the floored 1 +- zeta, sigmas floored and
// sigma_ab clamped as for every GGA
template <class T> __host__ __device__ inline T lyp_c(T rho, T opz, T omz, T saa, T sab, T sbb) {
  const double a = 0.04918, b = 0.132, c = 0.2533, d = 0.349;
  const double cf = 0.3 * ::cbrt(9.0 * kPi * kPi * kPi * kPi), c113 = 12.699208415745595;  // 2^(11/3)
  saa = floor_at(saa, 1e-40);
  sbb = floor_at(sbb, 1e-40);
  const T sav = 0.5 * (saa + sbb);
  if (val(sab) > val(sav)) sab = sav;
  if (val(sab) < -val(sav)) sab = -1.0 * sav;
  const T ra = 0.5 * rho * opz, rb = 0.5 * rho * omz;
  const T rm = 1.0 / cbrt(rho), den = 1.0 + d * rm;
  const T rm3 = rm * rm * rm;
  const T omega = exp(-c * rm) / den * (rm3 * rm3 * rm3 * rm * rm);  // ... rho^(-11/3)
  const T delta = c * rm + d * rm / den;
  const T sig = saa + 2.0 * sab + sbb;
  const T ra43 = p43(ra), rb43 = p43(rb);
  const T br = ra * rb *
               (c113 * cf * (ra43 * ra43 + rb43 * rb43) + (47.0 / 18.0 - 7.0 / 18.0 * delta) * sig -
                (2.5 - delta / 18.0) * (saa + sbb) - (delta - 11.0) / 9.0 * (ra * saa + rb * sbb) / rho);
  const T r2 = (2.0 / 3.0) * rho * rho;
  return -4.0 * a * ra * rb / (rho * den) -
         a * b * omega * (br - r2 * sig + (r2 - ra * ra) * sbb + (r2 - rb * rb) * saa);
}

// energy per volume of functional `id` (0: none)
template <class T>
__host__ __device__ inline T energy(int id, T ra, T rb, T saa, T sab, T sbb, T ta, T tb, double thr, bool pol) {
  if (id == 0) return ra * 0.0;
  if (id == 106) return b88_pair(ra, rb, saa, sbb, thr, pol);
  if (id == 202) {
    T e = ra * 0.0;
    if (!pol || val(ra) >= thr) e = e + 0.5 * tpss_x(2.0 * ra, 4.0 * floor_at(saa, 1e-40), 2.0 * ta);
    if (!pol || val(rb) >= thr) e = e + 0.5 * tpss_x(2.0 * rb, 4.0 * floor_at(sbb, 1e-40), 2.0 * tb);
    return e;
  }
  if (id == 231) return tpss_c(ra, rb, saa, sab, sbb, ta, tb, thr);
  if (id == 1) return slater_x(ra, rb);
  if (id == 101) return pbe_x_pair(ra, rb, saa, sbb, thr, pol);
  if (id == 406) return 0.75 * pbe_x_pair(ra, rb, saa, sbb, thr, pol) + pbe_c(ra, rb, saa, sab, sbb);
  const T rho = ra + rb;
  const T z = (ra - rb) / rho;
  const T opz = floor_at(1.0 + z, DBL_EPSILON), omz = floor_at(1.0 - z, DBL_EPSILON);
  const T rs = cbrt(3.0 / (4.0 * kPi * rho));
  if (id == 7) return vwn5_c(rho, rs, z, opz, omz);
  if (id == 12 || id == 13) return rho * pw92(id == 13, rs, z, opz, omz);
  if (id == 8) return vwn_rpa_c(rho, rs, opz, omz);
  if (id == 131) return lyp_c(rho, opz, omz, saa, sab, sbb);
  if (id == 402)
    return 0.08 * slater_x(ra, rb) + 0.72 * b88_pair(ra, rb, saa, sbb, thr, pol) +
           0.19 * vwn_rpa_c(rho, rs, opz, omz) + 0.81 * lyp_c(rho, opz, omz, saa, sab, sbb);
  if (id == 433 || id == 470)
    return 0.35 * b88_pair(ra, rb, saa, sbb, thr, pol) +
           0.46 * sr_b88_pair(ra, rb, saa, sbb, id == 433 ? 0.33 : 0.34, id == 470, thr, pol) +
           0.19 * vwn5_c(rho, rs, z, opz, omz) + 0.81 * lyp_c(rho, opz, omz, saa, sab, sbb);
  // 130: PBE correlation
  return pbe_c(ra, rb, saa, sab, sbb);
}

__host__ __device__ inline bool is_gga(int id) {
  return id == 101 || id == 106 || id == 130 || id == 131 || id == 202 || id == 231 || id == 402 || id == 406 ||
         id == 433 || id == 470;
}
__host__ __device__ inline bool is_mgga(int id) { return id == 202 || id == 231; }
__host__ __device__ inline bool exchange_ok(int id) {
  return id == 0 || id == 1 || id == 101 || id == 106 || id == 202 || id == 402 || id == 406 || id == 433 ||
         id == 470;
}
__host__ __device__ inline bool correlation_ok(int id) {
  return id == 0 || id == 7 || id == 8 || id == 12 || id == 13 || id == 130 || id == 131 || id == 231;
}

// exact-exchange fractions of a hybrid (full range, short range), its
// range-separation parameter and short-range kernel (1 erfc, 2 Yukawa); zero
// for the pure functionals
__host__ __device__ inline void exact_exchange(int id, double &kfrac, double &kshort, double &omega, int &kernel) {
  const bool cam = id == 433 || id == 470;
  kfrac = id == 402 ? 0.20 : id == 406 ? 0.25 : cam ? 0.65 : 0.0;
  kshort = cam ? -0.46 : 0.0;
  omega = id == 433 ? 0.33 : id == 470 ? 0.34 : 0.0;
  kernel = id == 433 ? 1 : id == 470 ? 2 : 0;
}

template <class T>
__host__ __device__ inline T total(int xf, int cf, T ra, T rb, T saa, T sab, T sbb, T ta, T tb, double thr,
                                   bool pol) {
  return energy(xf, ra, rb, saa, sab, sbb, ta, tb, thr, pol) + energy(cf, ra, rb, saa, sab, sbb, ta, tb, thr, pol);
}

// Unpolarized: x = (rho, sigma = |grad rho|^2, tau) -> energy per volume f and
// v[k] = df/dx_k (sigma only with gga, tau only with mgga).
__host__ __device__ inline void eval_unpol(int xf, int cf, const double x[3], double thr, bool gga, bool mgga, double &f,
                                           double v[3]) {
  f = v[0] = v[1] = v[2] = 0.0;
  if (!(x[0] >= thr)) return;
  const int nk = mgga ? 3 : gga ? 2 : 1;
  for (int k = 0; k < nk; ++k) {
    const Dl r = mk(x[0], k == 0 ? 1.0 : 0.0), s = mk(x[1], k == 1 ? 1.0 : 0.0), t = mk(x[2], k == 2 ? 1.0 : 0.0);
    const Dl e = total(xf, cf, 0.5 * r, 0.5 * r, 0.25 * s, 0.25 * s, 0.25 * s, 0.5 * t, 0.5 * t, thr, false);
    f = e.v;
    v[k] = e.d;
  }
}

// Polarized: x = (rho_a, rho_b, sigma_aa, sigma_ab, sigma_bb, tau_a, tau_b)
// -> f and v[k] = df/dx_k (the sigma derivatives only with gga, tau with mgga).
__host__ __device__ inline void eval_pol(int xf, int cf, const double x[7], double thr, bool gga, bool mgga,
                                         double &f, double v[7]) {
  f = 0.0;
  for (int k = 0; k < 7; ++k) v[k] = 0.0;
  if (!(x[0] + x[1] >= thr)) return;
  const int nk = mgga ? 7 : gga ? 5 : 2;
  for (int k = 0; k < nk; ++k) {
    Dl y[7];
    for (int j = 0; j < 7; ++j) y[j] = mk(x[j], j == k ? 1.0 : 0.0);
    const Dl e = total(xf, cf, y[0], y[1], y[2], y[3], y[4], y[5], y[6], thr, true);
    f = e.v;
    v[k] = e.d;
  }
}

}  // namespace xcf
}  // namespace hfg
