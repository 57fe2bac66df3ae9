// Oracle (test infrastructure only): exchange-correlation functionals by
// their published formulas, for the functional ids HelFEM passes to libxc
// (src/general/dftfuncs.cpp; libxc is not available here):
//   1   lda_x        Slater exchange, spin-scaled
//   7   lda_c_vwn    VWN5 (Vosko, Wilk, Nusair, Can. J. Phys. 58, 1200 (1980))
//   12  lda_c_pw     Perdew-Wang 92 (Phys. Rev. B 45, 13244 (1992)), as printed
//   13  lda_c_pw_mod the same with libxc's more precise A and f''(0)
//   101 gga_x_pbe    PBE exchange (Perdew, Burke, Ernzerhof, PRL 77, 3865 (1996)),
//                    kappa = 0.804, mu = beta pi^2 / 3
//   130 gga_c_pbe    PBE correlation on lda_c_pw_mod, beta = 0.06672455060314922,
//                    gamma = (1 - ln 2)/pi^2
//   202 mgga_x_tpss  TPSS exchange (Tao, Perdew, Staroverov, Scuseria, PRL 91,
//                    146401 (2003)): kappa 0.804, b 0.40, c 1.59096, e 1.537,
//                    mu 0.21951, by spin scaling
//   231 mgga_c_tpss  TPSS correlation (same paper): revPKZB on PBE correlation,
//                    C(zeta, xi) with 0.53, 0.87, 0.50, 2.26, d = 2.8
//   8   lda_c_vwn_rpa VWN's fit to the RPA correlation energy (same VWN paper),
//                    interpolated in f(zeta) between its paramagnetic and
//                    ferromagnetic curves
//   106 gga_x_b88    Becke 88 exchange (Phys. Rev. A 38, 3098 (1988)), beta 0.0042,
//                    by spin channel
//   131 gga_c_lyp    Lee-Yang-Parr correlation in the closed form of Miehlich,
//                    Savin, Stoll, Preuss (Chem. Phys. Lett. 157, 200 (1989)),
//                    a 0.04918, b 0.132, c 0.2533, d 0.349
//   402 hyb_gga_xc_b3lyp  0.08 lda_x + 0.72 gga_x_b88 + 0.19 lda_c_vwn_rpa +
//                    0.81 gga_c_lyp, with 0.20 exact exchange
//   406 hyb_gga_xc_pbeh   (PBE0) 0.75 gga_x_pbe + gga_c_pbe, with 0.25 exact exchange
//   433 hyb_gga_xc_cam_b3lyp   0.35 gga_x_b88 + 0.46 short-range B88 (erfc
//                    attenuation of Iikura, Tsuneda, Yanai, Hirao, JCP 115, 3540
//                    (2001), omega 0.33) + 0.19 lda_c_vwn + 0.81 gga_c_lyp, with
//                    0.65 exact exchange and -0.46 of its erfc short-range part
//   470 hyb_gga_xc_camy_b3lyp  the same with the Yukawa attenuation (Savin-Flad
//                    form; Akinaga, Ten-no, CPL 462, 348 (2008)), omega 0.34
// Each functional is one energy density per volume, f(rho_a, rho_b, sigma_aa,
// sigma_ab, sigma_bb, tau_a, tau_b), written once as a template; the
// potentials (vrho = df/drho, vsigma = df/dsigma, vtau = df/dtau) are its
// complex-step derivatives, Im f(x + i h)/h, exact to rounding for these
// analytic formulas.  Meta-GGAs also follow libxc's tau floor (1e-20) and
// its Fermi-hole constraint sigma_ss <= 8 rho_s tau_s (z <= 1).
// Thresholds follow libxc's: a point whose total density is below thr gives
// zero; 1 +- zeta is floored at DBL_EPSILON; sigma_ab is clamped to
// +-(sigma_aa + sigma_bb)/2 and the sigmas floored at 1e-40; in a polarized
// evaluation a spin below thr adds no PBE exchange (libxc floors it to thr
// instead: a difference of at most |e_x(thr)| ~ 1e-16 per point).
#include "dftgrid.hpp"
#include <cfloat>
#include <complex>

namespace orc {

namespace {
using cd = std::complex<double>;
const double kPi = 3.14159265358979323846264338327950288;

double re(double x) { return x; }
double re(const cd &x) { return x.real(); }
double cbrt_(double x) { return std::cbrt(x); }
cd cbrt_(const cd &x) { return std::pow(x, 1.0 / 3.0); }
double p43(double x) { return x * std::cbrt(x); }
cd p43(const cd &x) { return x * cbrt_(x); }
double p23(double x) { const double c = std::cbrt(x); return c * c; }
cd p23(const cd &x) { const cd c = cbrt_(x); return c * c; }

template <class T> T floor_at(const T &x, double lo) { return re(x) < lo ? T(lo) : x; }

// VWN interpolation in x = sqrt(rs) (libxc maple/vwn.mpl f_aux)
template <class T> T vwn_F(double A, double b, double c, double x0, const T &x) {
  const double Q = std::sqrt(4.0 * c - b * b), X0 = x0 * x0 + b * x0 + c;
  const double f1 = 2.0 * b / Q, f2 = b * x0 / X0, f3 = 2.0 * (2.0 * x0 + b) / Q;
  const T X = x * x + b * x + c;
  return A * (log(x * x / X) + (f1 - f2 * f3) * atan(Q / (2.0 * x + b)) - f2 * log((x - x0) * (x - x0) / X));
}

template <class T> T fzeta(const T &opz, const T &omz) {
  return (p43(opz) + p43(omz) - 2.0) / (2.0 * std::cbrt(2.0) - 2.0);
}

// Perdew-Wang 92 correlation energy per particle
template <class T> T pw92(const T &rs, const T &z, const T &opz, const T &omz, bool mod) {
  static const double A[2][3] = {{0.031091, 0.015545, 0.016887}, {0.0310907, 0.01554535, 0.0168869}};
  static const double a1[3] = {0.21370, 0.20548, 0.11125}, b1[3] = {7.5957, 14.1189, 10.357},
                      b2[3] = {3.5876, 6.1977, 3.6231}, b3[3] = {1.6382, 3.3662, 0.88026},
                      b4[3] = {0.49294, 0.62517, 0.49671};
  const double fz20 = mod ? 1.709920934161365617563962776245 : 1.709921;
  const T srs = sqrt(rs);
  T g[3];
  for (int k = 0; k < 3; ++k) {
    const double a = A[mod ? 1 : 0][k];
    const T aux = b1[k] * srs + b2[k] * rs + b3[k] * rs * srs + b4[k] * rs * rs;
    g[k] = -2.0 * a * (1.0 + a1[k] * rs) * log(1.0 + 1.0 / (2.0 * a * aux));
  }
  const T fz = fzeta(opz, omz), z4 = z * z * z * z;
  return g[0] + z4 * fz * (g[1] - g[0] + g[2] / fz20) - fz * g[2] / fz20;
}

// PBE exchange of an unpolarized density n with gradient invariant s2n = |grad n|^2
template <class T> T pbe_x(const T &n, const T &s2n) {
  const double kappa = 0.8040, mu = 0.06672455060314922 * kPi * kPi / 3.0;
  const T s2 = s2n / (4.0 * std::pow(3.0 * kPi * kPi, 2.0 / 3.0) * p43(n) * p43(n));
  const T F = 1.0 + kappa - kappa / (1.0 + mu * s2 / kappa);
  return -0.75 * std::cbrt(3.0 / kPi) * p43(n) * F;
}

// TPSS exchange of an unpolarized density n, |grad n|^2 = s2n, kinetic energy
// density tn (per volume)
template <class T> T tpss_x(const T &n, T s2n, T tn) {
  const double kappa = 0.804, b = 0.40, c = 1.59096, e = 1.537, mu = 0.21951;
  const double ckf = std::pow(3.0 * kPi * kPi, 2.0 / 3.0);
  tn = floor_at(tn, 1e-20);
  if (re(s2n) > re(8.0 * n * tn)) s2n = 8.0 * n * tn;
  const T n83 = p43(n) * p43(n);
  const T p = s2n / (4.0 * ckf * n83);
  const T tw = s2n / (8.0 * n);
  const T z = tw / tn;
  const T tunif = 0.3 * ckf * p43(n) * cbrt_(n);
  const T alpha = (tn - tw) / tunif;
  const T qb = 0.45 * (alpha - 1.0) / sqrt(1.0 + b * alpha * (alpha - 1.0)) + 2.0 * p / 3.0;
  const T z35 = 0.6 * z;
  const T x = ((10.0 / 81.0 + c * z * z / ((1.0 + z * z) * (1.0 + z * z))) * p + (146.0 / 2025.0) * qb * qb -
               (73.0 / 405.0) * qb * sqrt(0.5 * z35 * z35 + 0.5 * p * p) + (1.0 / kappa) * (100.0 / 6561.0) * p * p +
               2.0 * std::sqrt(e) * (10.0 / 81.0) * z35 * z35 + e * mu * p * p * p) /
              ((1.0 + std::sqrt(e) * p) * (1.0 + std::sqrt(e) * p));
  return -0.75 * std::cbrt(3.0 / kPi) * p43(n) * (1.0 + kappa - kappa / (1.0 + x / kappa));
}

template <class T> T energy(int id, T ra, T rb, T saa, T sab, T sbb, T ta, T tb, double thr, bool pol);

double asinh_(double x) { return std::asinh(x); }
cd asinh_(const cd &x) { return std::log(x + std::sqrt(x * x + 1.0)); }

// Becke 88 exchange per volume of one spin channel (density r, sigma s):
// -r^(4/3) [C_x + beta x^2 / (1 + 6 beta x asinh x)], x = sqrt(s) / r^(4/3),
// C_x = (3/4)(6/pi)^(1/3) the spin-scaled Slater constant
template <class T> T b88_x(const T &r, const T &s) {
  const double beta = 0.0042, cx = 0.75 * std::cbrt(6.0 / kPi);
  const T r43 = p43(r), x = sqrt(s) / r43;
  return -r43 * (cx + beta * x * x / (1.0 + 6.0 * beta * x * asinh_(x)));
}

double erf_(double x) { return std::erf(x); }
cd erf_(const cd &z) {  // first order in the (complex-step) imaginary part
  const double x = z.real();
  return cd(std::erf(x), z.imag() * 2.0 / std::sqrt(kPi) * std::exp(-x * x));
}
double atan_(double x) { return std::atan(x); }
cd atan_(const cd &z) { return std::atan(z); }

// Short-range attenuation of the uniform-gas exchange of one spin channel,
// a = omega / (2 k_F): (8/3) int_0^1 w(x) (1 - 3x/2 + x^3/2) dx with w = 1 -
// exp(-x^2 / 4a^2) for the erfc kernel and x^2 / (x^2 + a^2) for the Yukawa
// kernel exp(-omega r)/r.  Closed forms for small a; for large a (where they
// cancel) the series of the integral, sum_n (-1)^(n+1) B_n c_n,
// B_n = 1/(2n+1) - 3/(4n+4) + 1/(4n+8), c_n = (4a^2)^-n / n! (erfc), a^-2n (Yukawa).
double series_B(int n) { return 1.0 / (2 * n + 1) - 3.0 / (4 * n + 4) + 1.0 / (4 * n + 8); }
template <class T> T att_erfc(const T &a) {
  const T a2 = a * a;
  if (re(a) < 0.5)
    return 1.0 - (8.0 / 3.0) * a *
                     (std::sqrt(kPi) * erf_(0.5 / a) + (2.0 * a - 4.0 * a2 * a) * exp(-0.25 / a2) - 3.0 * a +
                      4.0 * a2 * a);
  const T y = 0.25 / a2;
  T s = T(0.0), p = T(1.0);
  double fact = 1.0;
  for (int n = 1; n <= 24; ++n) {
    p = p * y;
    fact *= n;
    s = s + ((n & 1) ? 1.0 : -1.0) * series_B(n) / fact * p;
  }
  return (8.0 / 3.0) * s;
}
template <class T> T att_yukawa(const T &a) {
  const T a2 = a * a;
  if (re(a) < 2.0) return 1.0 - (8.0 / 3.0) * a * (atan_(1.0 / a) + 0.25 * a * (1.0 - (a2 + 3.0) * log(1.0 + 1.0 / a2)));
  const T y = 1.0 / a2;
  T s = T(0.0), p = T(1.0);
  for (int n = 1; n <= 40; ++n) {
    p = p * y;
    s = s + ((n & 1) ? 1.0 : -1.0) * series_B(n) * p;
  }
  return (8.0 / 3.0) * s;
}

// Short-range B88 of one spin channel: the B88 enhancement F(x) times the
// attenuation at a = omega / (2 k), k = sqrt(9 pi / (2 C_x F)) r^(1/3) the
// GGA Fermi momentum of Iikura et al. (ITYH; with the Yukawa attenuation, SFAT)
template <class T> T sr_b88_x(const T &r, const T &s, double omega, bool yukawa) {
  const double beta = 0.0042, cx = 0.75 * std::cbrt(6.0 / kPi);
  const T r13 = cbrt_(r), r43 = r * r13, x = sqrt(s) / r43;
  const T F = 1.0 + (beta / cx) * x * x / (1.0 + 6.0 * beta * x * asinh_(x));
  const T k = sqrt(9.0 * kPi / (2.0 * cx * F)) * r13;
  const T a = omega / (2.0 * k);
  return -cx * r43 * F * (yukawa ? att_yukawa(a) : att_erfc(a));
}

// LYP correlation per volume (Miehlich et al. closed form) on the spin
// densities rebuilt from rho and the floored 1 +- zeta, as libxc evaluates it
template <class T> T lyp_c(const T &rho, const T &opz, const T &omz, const T &saa, const T &sab, const T &sbb) {
  const double a = 0.04918, b = 0.132, c = 0.2533, d = 0.349;
  const double cf = 0.3 * std::cbrt(9.0 * kPi * kPi * kPi * kPi), c113 = std::pow(2.0, 11.0 / 3.0);
  const T ra = 0.5 * rho * opz, rb = 0.5 * rho * omz;
  const T rm = 1.0 / cbrt_(rho), den = 1.0 + d * rm;
  const T rm3 = rm * rm * rm;
  const T omega = exp(-c * rm) / den * (rm3 * rm3 * rm3 * rm * rm);  // ... rho^(-11/3)
  const T delta = c * rm + d * rm / den;
  const T sig = saa + 2.0 * sab + sbb;
  const T ra43 = p43(ra), rb43 = p43(rb);
  const T br = ra * rb *
               (c113 * cf * (ra43 * ra43 + rb43 * rb43) + (47.0 / 18.0 - 7.0 / 18.0 * delta) * sig -
                (2.5 - delta / 18.0) * (saa + sbb) - (delta - 11.0) / 9.0 * (ra * saa + rb * sbb) / rho);
  const T r2 = (2.0 / 3.0) * rho * rho;
  return -4.0 * a * ra * rb / (rho * den) - a * b * omega * (br - r2 * sig + (r2 - ra * ra) * sbb + (r2 - rb * rb) * saa);
}

// PBE correlation energy per particle
template <class T> T pbe_c_eps(const T &ra, const T &rb, const T &saa, const T &sab, const T &sbb, double thr) {
  return energy(130, ra, rb, saa, sab, sbb, T(0.0), T(0.0), thr, true) / (ra + rb);
}

// TPSS correlation energy per volume
template <class T> T tpss_c(T ra, T rb, T saa, T sab, T sbb, T ta, T tb, double thr) {
  const double d = 2.8;
  ta = floor_at(ta, 1e-20);
  tb = floor_at(tb, 1e-20);
  saa = floor_at(saa, 1e-40);
  sbb = floor_at(sbb, 1e-40);
  if (re(saa) > re(8.0 * ra * ta)) saa = 8.0 * ra * ta;
  if (re(sbb) > re(8.0 * rb * tb)) sbb = 8.0 * rb * tb;
  const T sav = 0.5 * (saa + sbb);
  if (re(sab) > re(sav)) sab = sav;
  if (re(sab) < -re(sav)) sab = -sav;
  const T n = ra + rb, sig = saa + 2.0 * sab + sbb, tau = ta + tb;
  const T z = sig / (8.0 * n * tau);
  const T zeta = (ra - rb) / n;
  const T opz = floor_at(T(1.0) + zeta, DBL_EPSILON), omz = floor_at(T(1.0) - zeta, DBL_EPSILON);
  const T eps = pbe_c_eps(ra, rb, saa, sab, sbb, thr);
  // epsilon~_s = max(eps_PBE(n_s, 0, grad n_s, 0), eps_PBE(n_up, n_dn, ...)); a spin
  // below the density threshold carries no weight
  T sum = T(0.0);
  if (re(ra) >= thr) {
    T ea = pbe_c_eps(ra, T(0.0), saa, T(0.0), T(0.0), thr);
    if (re(ea) < re(eps)) ea = eps;
    sum = sum + ra / n * ea;
  }
  if (re(rb) >= thr) {
    T eb = pbe_c_eps(rb, T(0.0), sbb, T(0.0), T(0.0), thr);
    if (re(eb) < re(eps)) eb = eps;
    sum = sum + rb / n * eb;
  }
  const T gz2 = (omz * omz * saa - 2.0 * omz * opz * sab + opz * opz * sbb) / (n * n);
  const T xi2 = gz2 / (4.0 * p23(3.0 * kPi * kPi * n));
  const T z2 = zeta * zeta;
  const T C0 = 0.53 + 0.87 * z2 + 0.50 * z2 * z2 + 2.26 * z2 * z2 * z2;
  const T den = 1.0 + 0.5 * xi2 * (1.0 / p43(opz) + 1.0 / p43(omz));
  const T C = C0 / (den * den * den * den);
  const T erev = eps * (1.0 + C * z * z) - (1.0 + C) * z * z * sum;
  return n * erev * (1.0 + d * erev * z * z * z);
}

// f(rho_a, rho_b, sigma_aa, sigma_ab, sigma_bb, tau_a, tau_b) per volume of functional id
template <class T>
T energy(int id, T ra, T rb, T saa, T sab, T sbb, T ta, T tb, double thr, bool pol) {
  const T rho = ra + rb;
  if (id == 402)
    return 0.08 * energy(1, ra, rb, saa, sab, sbb, ta, tb, thr, pol) +
           0.72 * energy(106, ra, rb, saa, sab, sbb, ta, tb, thr, pol) +
           0.19 * energy(8, ra, rb, saa, sab, sbb, ta, tb, thr, pol) +
           0.81 * energy(131, ra, rb, saa, sab, sbb, ta, tb, thr, pol);
  if (id == 406)
    return 0.75 * energy(101, ra, rb, saa, sab, sbb, ta, tb, thr, pol) +
           energy(130, ra, rb, saa, sab, sbb, ta, tb, thr, pol);
  if (id == 433 || id == 470) {
    const double omega = id == 433 ? 0.33 : 0.34;
    T sr = T(0.0);
    const T fa = floor_at(saa, 1e-40), fb = floor_at(sbb, 1e-40);
    if (!pol || re(ra) >= thr) sr = sr + sr_b88_x(ra, fa, omega, id == 470);
    if (!pol || re(rb) >= thr) sr = sr + sr_b88_x(rb, fb, omega, id == 470);
    return 0.35 * energy(106, ra, rb, saa, sab, sbb, ta, tb, thr, pol) + 0.46 * sr +
           0.19 * energy(7, ra, rb, saa, sab, sbb, ta, tb, thr, pol) +
           0.81 * energy(131, ra, rb, saa, sab, sbb, ta, tb, thr, pol);
  }
  if (id == 106) {
    saa = floor_at(saa, 1e-40);
    sbb = floor_at(sbb, 1e-40);
    T e = T(0.0);
    if (!pol || re(ra) >= thr) e = e + b88_x(ra, saa);
    if (!pol || re(rb) >= thr) e = e + b88_x(rb, sbb);
    return e;
  }
  if (id == 202) {
    saa = floor_at(saa, 1e-40);
    sbb = floor_at(sbb, 1e-40);
    T e = T(0.0);
    if (!pol || re(ra) >= thr) e = e + 0.5 * tpss_x(2.0 * ra, 4.0 * saa, 2.0 * ta);
    if (!pol || re(rb) >= thr) e = e + 0.5 * tpss_x(2.0 * rb, 4.0 * sbb, 2.0 * tb);
    return e;
  }
  if (id == 231) return tpss_c(ra, rb, saa, sab, sbb, ta, tb, thr);
  if (id == 1) {
    T e = T(0.0);
    if (re(ra) > 0.0) e = e + p43(ra);
    if (re(rb) > 0.0) e = e + p43(rb);
    return -0.75 * std::cbrt(6.0 / kPi) * e;
  }
  if (id == 101) {
    saa = floor_at(saa, 1e-40);
    sbb = floor_at(sbb, 1e-40);
    T e = T(0.0);
    if (!pol || re(ra) >= thr) e = e + 0.5 * pbe_x(2.0 * ra, 4.0 * saa);
    if (!pol || re(rb) >= thr) e = e + 0.5 * pbe_x(2.0 * rb, 4.0 * sbb);
    return e;
  }
  const T z = (ra - rb) / rho;
  const T opz = floor_at(T(1.0) + z, DBL_EPSILON), omz = floor_at(T(1.0) - z, DBL_EPSILON);
  const T rs = cbrt_(3.0 / (4.0 * kPi * rho));
  if (id == 7) {
    const T x = sqrt(rs);
    const T eP = vwn_F(0.0310907, 3.72744, 12.9352, -0.10498, x);
    const T eF = vwn_F(0.01554535, 7.06042, 18.0578, -0.32500, x);
    const T eA = vwn_F(-1.0 / (6.0 * kPi * kPi), 1.13107, 13.0045, -0.0047584, x);
    const double fpp = 4.0 / (9.0 * (std::cbrt(2.0) - 1.0));
    const T fz = fzeta(opz, omz), z4 = z * z * z * z;
    return rho * (eP + eA * fz * (1.0 - z4) / fpp + (eF - eP) * fz * z4);
  }
  if (id == 8) {
    const T x = sqrt(rs);
    const T eP = vwn_F(0.0310907, 13.0720, 42.7198, -0.409286, x);
    const T eF = vwn_F(0.01554535, 20.1231, 101.578, -0.743294, x);
    return rho * (eP + (eF - eP) * fzeta(opz, omz));
  }
  if (id == 12 || id == 13) return rho * pw92(rs, z, opz, omz, id == 13);
  if (id == 131) {
    saa = floor_at(saa, 1e-40);
    sbb = floor_at(sbb, 1e-40);
    const T sav = 0.5 * (saa + sbb);
    if (re(sab) > re(sav)) sab = sav;
    if (re(sab) < -re(sav)) sab = -sav;
    return lyp_c(rho, opz, omz, saa, sab, sbb);
  }
  if (id == 130) {
    const double beta = 0.06672455060314922, gamma = (1.0 - std::log(2.0)) / (kPi * kPi);
    saa = floor_at(saa, 1e-40);
    sbb = floor_at(sbb, 1e-40);
    const T sav = 0.5 * (saa + sbb);
    if (re(sab) > re(sav)) sab = sav;
    if (re(sab) < -re(sav)) sab = -sav;
    const T ec = pw92(rs, z, opz, omz, true);
    const T phi = 0.5 * (p23(opz) + p23(omz)), phi3 = phi * phi * phi;
    const T kf = cbrt_(3.0 * kPi * kPi * rho), ks2 = 4.0 * kf / kPi;
    const T t2 = (saa + 2.0 * sab + sbb) / (4.0 * phi * phi * ks2 * rho * rho);
    const T A = (beta / gamma) / (exp(-ec / (gamma * phi3)) - 1.0);
    const T At2 = A * t2;
    const T H = gamma * phi3 * log(1.0 + (beta / gamma) * t2 * (1.0 + At2) / (1.0 + At2 + At2 * At2));
    return rho * (ec + H);
  }
  return T(0.0);
}

template <class T>
T total(int xf, int cf, const T &ra, const T &rb, const T &saa, const T &sab, const T &sbb, const T &ta, const T &tb,
        double thr, bool pol) {
  T f = T(0.0);
  if (xf) f = f + energy(xf, ra, rb, saa, sab, sbb, ta, tb, thr, pol);
  if (cf) f = f + energy(cf, ra, rb, saa, sab, sbb, ta, tb, thr, pol);
  return f;
}

double step(double x) { return std::fabs(x) > 1e-250 ? 1e-20 * std::fabs(x) : 1e-270; }
}  // namespace

bool xc_is_gga(int id) {
  return id == 101 || id == 106 || id == 130 || id == 131 || id == 202 || id == 231 || id == 402 || id == 406 ||
         id == 433 || id == 470;
}
bool xc_is_mgga(int id) { return id == 202 || id == 231; }

void xc_check(int xf, int cf) {
  if (!(xf == 0 || xf == 1 || xf == 101 || xf == 106 || xf == 202 || xf == 402 || xf == 406 || xf == 433 ||
        xf == 470))
    throw std::invalid_argument("unsupported exchange functional");
  if (!(cf == 0 || cf == 7 || cf == 8 || cf == 12 || cf == 13 || cf == 130 || cf == 131 || cf == 231))
    throw std::invalid_argument("unsupported correlation functional");
}

void xc_exact_exchange(int xf, double &kfrac, double &kshort, double &omega, int &kernel) {
  kfrac = xf == 402 ? 0.20 : xf == 406 ? 0.25 : (xf == 433 || xf == 470) ? 0.65 : 0.0;
  kshort = (xf == 433 || xf == 470) ? -0.46 : 0.0;
  omega = xf == 433 ? 0.33 : xf == 470 ? 0.34 : 0.0;
  kernel = xf == 433 ? 1 : xf == 470 ? 2 : 0;
}

void xc_unpolarized(int xf, int cf, double rho, double sigma, double tau, double thr, double &exc, double &vrho,
                    double &vsigma, double &vtau) {
  xc_check(xf, cf);
  exc = vrho = vsigma = vtau = 0.0;
  if (!(rho >= thr)) return;
  auto f = [&](auto r, auto s, auto t) {
    return total(xf, cf, 0.5 * r, 0.5 * r, 0.25 * s, 0.25 * s, 0.25 * s, 0.5 * t, 0.5 * t, thr, false);
  };
  exc = f(rho, sigma, tau) / rho;
  const double hr = step(rho), hs = step(sigma), ht = step(tau);
  vrho = f(cd(rho, hr), cd(sigma, 0.0), cd(tau, 0.0)).imag() / hr;
  vsigma = f(cd(rho, 0.0), cd(sigma, hs), cd(tau, 0.0)).imag() / hs;
  vtau = f(cd(rho, 0.0), cd(sigma, 0.0), cd(tau, ht)).imag() / ht;
}

void xc_polarized(int xf, int cf, const double rho[2], const double sigma[3], const double tau[2], double thr,
                  double &exc, double vrho[2], double vsigma[3], double vtau[2]) {
  xc_check(xf, cf);
  exc = 0.0;
  vrho[0] = vrho[1] = vsigma[0] = vsigma[1] = vsigma[2] = vtau[0] = vtau[1] = 0.0;
  const double rt = rho[0] + rho[1];
  if (!(rt >= thr)) return;
  double x[7] = {rho[0], rho[1], sigma[0], sigma[1], sigma[2], tau[0], tau[1]};
  exc = total(xf, cf, x[0], x[1], x[2], x[3], x[4], x[5], x[6], thr, true) / rt;
  double *out[7] = {&vrho[0], &vrho[1], &vsigma[0], &vsigma[1], &vsigma[2], &vtau[0], &vtau[1]};
  for (int k = 0; k < 7; ++k) {
    cd y[7];
    for (int j = 0; j < 7; ++j) y[j] = cd(x[j], 0.0);
    const double h = step(x[k]);
    y[k] = cd(x[k], h);
    *out[k] = total(xf, cf, y[0], y[1], y[2], y[3], y[4], y[5], y[6], thr, true).imag() / h;
  }
}

}  // namespace orc
