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
// Each functional is one energy density per volume, f(rho_a, rho_b, sigma_aa,
// sigma_ab, sigma_bb), written once as a template; the potentials
// (vrho = df/drho, vsigma = df/dsigma) are its complex-step derivatives,
// Im f(x + i h)/h, exact to rounding for these analytic formulas.
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

// f(rho_a, rho_b, sigma_aa, sigma_ab, sigma_bb) per volume of functional id
template <class T>
T energy(int id, T ra, T rb, T saa, T sab, T sbb, double thr, bool pol) {
  const T rho = ra + rb;
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
  if (id == 12 || id == 13) return rho * pw92(rs, z, opz, omz, id == 13);
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
T total(int xf, int cf, const T &ra, const T &rb, const T &saa, const T &sab, const T &sbb, double thr, bool pol) {
  T f = T(0.0);
  if (xf) f = f + energy(xf, ra, rb, saa, sab, sbb, thr, pol);
  if (cf) f = f + energy(cf, ra, rb, saa, sab, sbb, thr, pol);
  return f;
}

double step(double x) { return std::fabs(x) > 1e-250 ? 1e-20 * std::fabs(x) : 1e-270; }
}  // namespace

bool xc_is_gga(int id) { return id == 101 || id == 130; }

void xc_check(int xf, int cf) {
  if (!(xf == 0 || xf == 1 || xf == 101)) throw std::invalid_argument("unsupported exchange functional");
  if (!(cf == 0 || cf == 7 || cf == 12 || cf == 13 || cf == 130))
    throw std::invalid_argument("unsupported correlation functional");
}

void xc_unpolarized(int xf, int cf, double rho, double sigma, double thr, double &exc, double &vrho,
                    double &vsigma) {
  xc_check(xf, cf);
  exc = vrho = vsigma = 0.0;
  if (!(rho >= thr)) return;
  auto f = [&](auto r, auto s) { return total(xf, cf, 0.5 * r, 0.5 * r, 0.25 * s, 0.25 * s, 0.25 * s, thr, false); };
  exc = f(rho, sigma) / rho;
  const double hr = step(rho), hs = step(sigma);
  vrho = f(cd(rho, hr), cd(sigma, 0.0)).imag() / hr;
  vsigma = f(cd(rho, 0.0), cd(sigma, hs)).imag() / hs;
}

void xc_polarized(int xf, int cf, const double rho[2], const double sigma[3], double thr, double &exc,
                  double vrho[2], double vsigma[3]) {
  xc_check(xf, cf);
  exc = 0.0;
  vrho[0] = vrho[1] = vsigma[0] = vsigma[1] = vsigma[2] = 0.0;
  const double rt = rho[0] + rho[1];
  if (!(rt >= thr)) return;
  double x[5] = {rho[0], rho[1], sigma[0], sigma[1], sigma[2]};
  exc = total(xf, cf, x[0], x[1], x[2], x[3], x[4], thr, true) / rt;
  double *out[5] = {&vrho[0], &vrho[1], &vsigma[0], &vsigma[1], &vsigma[2]};
  for (int k = 0; k < 5; ++k) {
    cd y[5];
    for (int j = 0; j < 5; ++j) y[j] = cd(x[j], 0.0);
    const double h = step(x[k]);
    y[k] = cd(x[k], h);
    *out[k] = total(xf, cf, y[0], y[1], y[2], y[3], y[4], thr, true).imag() / h;
  }
}

}  // namespace orc
