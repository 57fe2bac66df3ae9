// Oracle (test infrastructure only): LDA exchange-correlation on HelFEM's
// atomic DFT grid, restating src/atomic/dftgrid.cpp (DFTGrid::eval_Fxc,
// :591-653; DFTGridWorker::compute_bf :470-576, update_density :51-128,
// eval_Fxc :304-359), src/general/angular.cpp (angular_chebyshev),
// src/general/dftgrid_common.{h,cpp} (eval_Exc, compute_Nel,
// increment_lda_split) and the libxc functionals lda_x (id 1) and
// lda_c_vwn (id 7, VWN5) through their published formulas (libxc is not
// available here).  The element loop is the reference's dense one: complex
// basis values on every grid point of the element, density by P * bf,
// potential matrix by (a.*wv) a' + (b.*wv) b'.
#include "orc.hpp"
#include "dftgrid.hpp"
#include "diatomic.hpp"
#include <algorithm>
#include <complex>

namespace orc {

namespace {
const double kPi = 3.14159265358979323846264338327950288;

// libxc maple/vwn.mpl: f_aux(A, b, c, x0, rs) and its x = sqrt(rs) derivative
void vwn_aux(double A, double b, double c, double x0, double x, double &F, double &dFdx) {
  const double Q = std::sqrt(4.0 * c - b * b);
  const double X = x * x + b * x + c, Xp = 2.0 * x + b;
  const double X0 = x0 * x0 + b * x0 + c;
  const double f1 = 2.0 * b / Q, f2 = b * x0 / X0, f3 = 2.0 * (2.0 * x0 + b) / Q;
  const double at = std::atan(Q / (2.0 * x + b));
  F = A * (std::log(x * x / X) + (f1 - f2 * f3) * at - f2 * std::log((x - x0) * (x - x0) / X));
  // d/dx atan(Q/(2x+b)) = -2Q/((2x+b)^2 + Q^2) = -Q/(2X)
  dFdx = A * (2.0 / x - Xp / X + (f1 - f2 * f3) * (-Q / (2.0 * X)) - f2 * (2.0 / (x - x0) - Xp / X));
}
}  // namespace

// Spin-unpolarized LDA energy density per particle and potential of the
// sum of exchange functional xf and correlation functional cf (0 = none);
// densities below thr give zero, as libxc's density threshold does.
void lda_unpolarized(int xf, int cf, double rho, double thr, double &exc, double &vrho) {
  exc = 0.0;
  vrho = 0.0;
  if (!(rho >= thr)) return;
  if (xf == 1) {
    // Slater: e_x = -(3/4)(3/pi)^(1/3) rho^(1/3), v_x = (4/3) e_x
    const double ex = -0.75 * std::cbrt(3.0 / kPi) * std::cbrt(rho);
    exc += ex;
    vrho += 4.0 / 3.0 * ex;
  } else if (xf != 0) {
    throw std::invalid_argument("unsupported exchange functional");
  }
  if (cf == 7) {
    // VWN5 paramagnetic: A = 0.0310907, b = 3.72744, c = 12.9352, x0 = -0.10498
    const double rs = std::cbrt(3.0 / (4.0 * kPi * rho));
    const double x = std::sqrt(rs);
    double F, dF;
    vwn_aux(0.0310907, 3.72744, 12.9352, -0.10498, x, F, dF);
    exc += F;
    vrho += F - x / 6.0 * dF;  // e_c - (rs/3) de_c/drs
  } else if (cf != 0) {
    throw std::invalid_argument("unsupported correlation functional");
  }
}

// Spin-polarized form (libxc's XC_POLARIZED): e per particle of the total
// density and the two potentials.  Slater exchange by spin scaling,
// e_x rho = -(3/4)(6/pi)^(1/3) (rho_a^(4/3) + rho_b^(4/3)), v_xs = -(6/pi)^(1/3) rho_s^(1/3);
// VWN5 correlation e_c = e_P + alpha_c f(z)(1 - z^4)/f''(0) + (e_F - e_P) f(z) z^4
// (libxc maple/lda_exc/lda_c_vwn.mpl), v_cs = e_c - (rs/3) de/drs - (z - s) de/dz.
void lda_polarized(int xf, int cf, double ra, double rb, double thr, double &exc, double &va, double &vb) {
  exc = 0.0;
  va = 0.0;
  vb = 0.0;
  const double rho = ra + rb;
  if (!(rho >= thr)) return;
  if (xf == 1) {
    const double c6 = std::cbrt(6.0 / kPi);
    const double a3 = ra > 0.0 ? std::cbrt(ra) : 0.0, b3 = rb > 0.0 ? std::cbrt(rb) : 0.0;
    exc += -0.75 * c6 * (ra * a3 + rb * b3) / rho;
    va += -c6 * a3;
    vb += -c6 * b3;
  } else if (xf != 0) {
    throw std::invalid_argument("unsupported exchange functional");
  }
  if (cf == 7) {
    const double rs = std::cbrt(3.0 / (4.0 * kPi * rho)), x = std::sqrt(rs);
    double z = (ra - rb) / rho;
    z = std::max(-1.0, std::min(1.0, z));
    double eP, dP, eF, dF, eA, dA;
    vwn_aux(0.0310907, 3.72744, 12.9352, -0.10498, x, eP, dP);
    vwn_aux(0.01554535, 7.06042, 18.0578, -0.32500, x, eF, dF);
    vwn_aux(-1.0 / (6.0 * kPi * kPi), 1.13107, 13.0045, -0.0047584, x, eA, dA);
    const double c43 = std::cbrt(2.0) * 2.0 - 2.0;  // 2^(4/3) - 2
    const double opz = 1.0 + z, omz = 1.0 - z;
    const double opz13 = std::cbrt(opz), omz13 = std::cbrt(omz);
    const double fz = (opz * opz13 + omz * omz13 - 2.0) / c43;
    const double dfz = 4.0 / 3.0 * (opz13 - omz13) / c43;
    const double fpp = 4.0 / (9.0 * (std::cbrt(2.0) - 1.0));
    const double z3 = z * z * z, z4 = z3 * z;
    const double e = eP + eA * fz * (1.0 - z4) / fpp + (eF - eP) * fz * z4;
    const double dedx = dP + dA * fz * (1.0 - z4) / fpp + (dF - dP) * fz * z4;
    const double dedz = eA * (dfz * (1.0 - z4) - 4.0 * z3 * fz) / fpp + (eF - eP) * (dfz * z4 + 4.0 * z3 * fz);
    const double base = e - x / 6.0 * dedx;
    exc += e;
    va += base - (z - 1.0) * dedz;
    vb += base - (z + 1.0) * dedz;
  } else if (cf != 0) {
    throw std::invalid_argument("unsupported correlation functional");
  }
}

// angular_chebyshev (src/general/angular.cpp:20-72): Chebyshev rule in
// cos(theta) times an even phi grid, w = w_cth * 2pi/nphi.
void angular_chebyshev(int l, int m, std::vector<double> &cth, std::vector<double> &phi, std::vector<double> &w) {
  std::vector<double> x, wx;
  chebyshev(l, x, wx);
  const double dphi = 2.0 * kPi / m;
  cth.assign((size_t)l * m, 0.0);
  phi = cth;
  w = cth;
  for (int i = 0; i < l; ++i)
    for (int j = 0; j < m; ++j) {
      const size_t q = (size_t)i * m + j;
      cth[q] = x[i];
      phi[q] = j * dphi;
      w[q] = wx[i] * dphi;
    }
}

// Y_l^m(cos theta, phi) with the Condon-Shortley phase
// (src/general/spherical_harmonics.cpp:19-35)
std::complex<double> spherical_harmonic(int l, int m, double cth, double phi) {
  if (m < 0) return std::conj(std::pow(-1.0, m) * spherical_harmonic(l, -m, cth, phi));
  std::complex<double> y = std::exp(std::complex<double>(0.0, m * phi));
  y *= std::sph_legendre((unsigned)l, (unsigned)m, std::acos(cth));
  return y;
}

AtomicXC::AtomicXC(const AtomicBasis &b_, int lang_, int mang_) : b(b_), lang(lang_), mang(mang_) {
  angular_chebyshev(lang, mang, cth, phi, wang);
}

// DFTGrid::eval_Fxc (restricted), src/atomic/dftgrid.cpp:591-653
void AtomicXC::eval_Fxc(int xf, int cf, const Mat &P, Mat &H, double &Exc, double &Nel, double thr) const {
  Mat Hb;
  eval(xf, cf, P, nullptr, H, Hb, Exc, Nel, thr);
}

// DFTGrid::eval_Fxc (spin-polarized), src/atomic/dftgrid.cpp:655-712
void AtomicXC::eval_Fxc(int xf, int cf, const Mat &Pa, const Mat &Pb, Mat &Ha, Mat &Hb, double &Exc, double &Nel,
                        double thr) const {
  eval(xf, cf, Pa, &Pb, Ha, Hb, Exc, Nel, thr);
}

// Shared element loop; Pb == nullptr selects the restricted form with P the total density.
// GGA functionals add the density gradient in spherical components
// (update_density, dftgrid.cpp:75-86, 166-180: g_c = 2 Re sum P (d_c chi) chi* /
// scale_c with scale = 1, r, r sin(theta)) from the basis derivatives of
// eval_df (TwoDBasis.cpp:1157-1200: d/dr, d/dtheta, d/dphi of R Y), and the
// potential term H += X + X', X = Re(sum_c gn_c d_c chi) chi*', gn_c = w (2
// vsigma g_c) / scale_c restricted, w (2 vsigma_ss g_s,c + vsigma_ab g_s',c) /
// scale_c polarized (eval_Fxc, dftgrid.cpp:320-334, 382-405;
// increment_gga_split, dftgrid_common.h:188-216).
void AtomicXC::eval(int xf, int cf, const Mat &P, const Mat *Pbp, Mat &H, Mat &Hb, double &Exc, double &Nel,
                    double thr) const {
  xc_check(xf, cf);
  const bool pol = Pbp != nullptr;
  const bool gga = xc_is_gga(xf) || xc_is_gga(cf);
  const bool mgga = xc_is_mgga(xf) || xc_is_mgga(cf);
  ekin = 0.0;
  const bool plain_lda = !gga && (xf == 0 || xf == 1) && (cf == 0 || cf == 7);
  const RadialBasis &rad = b.radial;
  const int Na = b.Nang(), Nr = rad.Nbf();
  H = Mat(P.r, P.c);
  if (pol) Hb = Mat(P.r, P.c);
  Exc = 0.0;
  Nel = 0.0;
  const int nang = (int)wang.size();
  for (int iel = 0; iel < rad.Nel(); ++iel) {
    // compute_bf, dftgrid.cpp:470-576
    const Mat R = rad.get_bf(rad.xq, iel);  // (nrad x Ni)
    const Mat dR = gga ? rad.get_df(rad.xq, iel) : Mat();
    const std::vector<double> r = rad.eval_coord(rad.xq, iel);
    const double h = rad.scaling_factor(iel);
    const int nrad = (int)rad.xq.size(), Ni = R.c, nbf = Na * Ni, npt = nrad * nang;
    std::vector<int> idx(nbf);
    for (int a = 0; a < Na; ++a)
      for (int j = 0; j < Ni; ++j) idx[a * Ni + j] = a * Nr + rad.first[iel] + j;
    std::vector<double> wtot(npt), sc_th(npt), sc_ph(npt);
    for (int ia = 0; ia < nang; ++ia) {
      const double sth = std::sqrt(1.0 - cth[ia] * cth[ia]);
      for (int ir = 0; ir < nrad; ++ir) {
        wtot[ia * nrad + ir] = wang[ia] * (h * rad.wq[ir]) * std::pow(r[ir], 2);
        sc_th[ia * nrad + ir] = r[ir];
        sc_ph[ia * nrad + ir] = r[ir] * sth;
      }
    }
    Mat are(nbf, npt), aim(nbf, npt);  // real and imaginary parts of conj(chi)
    Mat dre[3], dim[3];                // the same of conj(d_c chi), c = r, theta, phi
    if (gga)
      for (int c = 0; c < 3; ++c) dre[c] = dim[c] = Mat(nbf, npt);
    for (int ia = 0; ia < nang; ++ia)
      for (int a = 0; a < Na; ++a) {
        const int l = b.lval[a], m = b.mval[a];
        const std::complex<double> Y = spherical_harmonic(l, m, cth[ia], phi[ia]);
        const std::complex<double> y = std::conj(Y);
        std::complex<double> yc[3];
        if (gga) {
          const double sinth = std::sqrt(std::max((1.0 - cth[ia]) * (1.0 + cth[ia]), 0.0));
          const double cotth = sinth > 0.0 ? cth[ia] / sinth : 0.0;
          std::complex<double> angfac = double(m) * cotth * Y;
          if (m < l)
            angfac += std::sqrt(double((l - m) * (l + m + 1))) * std::exp(std::complex<double>(0.0, -phi[ia])) *
                      spherical_harmonic(l, m + 1, cth[ia], phi[ia]);
          yc[0] = y;
          yc[1] = std::conj(angfac);
          yc[2] = std::conj(std::complex<double>(0.0, m) * Y);
        }
        for (int j = 0; j < Ni; ++j)
          for (int ir = 0; ir < nrad; ++ir) {
            const int u = a * Ni + j, p = ia * nrad + ir;
            are(u, p) = y.real() * R(ir, j);
            aim(u, p) = y.imag() * R(ir, j);
            if (gga) {
              dre[0](u, p) = yc[0].real() * dR(ir, j);
              dim[0](u, p) = yc[0].imag() * dR(ir, j);
              for (int c = 1; c < 3; ++c) {
                dre[c](u, p) = yc[c].real() * R(ir, j);
                dim[c](u, p) = yc[c].imag() * R(ir, j);
              }
            }
          }
      }
    // update_density: rho = colwise (P a).a + (P b).b, per spin when polarized;
    // with GGA also g_c = 2 ((P a).da_c + (P b).db_c) / scale_c
    auto density = [&](const Mat &Pm, std::vector<double> &rho, std::vector<double> g[3], std::vector<double> &tau) {
      Mat Pe(nbf, nbf);
      for (int v = 0; v < nbf; ++v)
        for (int u = 0; u < nbf; ++u) Pe(u, v) = Pm(idx[u], idx[v]);
      const Mat PA = matmul(Pe, are), PB = matmul(Pe, aim);
      rho.assign(npt, 0.0);
      for (int p = 0; p < npt; ++p) {
        double s = 0.0, t = 0.0;
        for (int u = 0; u < nbf; ++u) {
          s += PA(u, p) * are(u, p);
          t += PB(u, p) * aim(u, p);
        }
        rho[p] = s + t;
      }
      if (!gga) return;
      for (int c = 0; c < 3; ++c) {
        g[c].assign(npt, 0.0);
        for (int p = 0; p < npt; ++p) {
          double s = 0.0, t = 0.0;
          for (int u = 0; u < nbf; ++u) {
            s += PA(u, p) * dre[c](u, p);
            t += PB(u, p) * dim[c](u, p);
          }
          const double sc = c == 0 ? 1.0 : (c == 1 ? sc_th[p] : sc_ph[p]);
          g[c][p] = 2.0 * (s + t) / sc;
        }
      }
      // kinetic energy density tau = 1/2 sum_c ((P da_c).da_c + (P db_c).db_c) / scale_c^2
      // (update_density, dftgrid.cpp:89-107, 186-214)
      if (!mgga) return;
      tau.assign(npt, 0.0);
      for (int c = 0; c < 3; ++c) {
        const Mat QA = matmul(Pe, dre[c]), QB = matmul(Pe, dim[c]);
        for (int p = 0; p < npt; ++p) {
          double s = 0.0;
          for (int u = 0; u < nbf; ++u) s += QA(u, p) * dre[c](u, p) + QB(u, p) * dim[c](u, p);
          const double sc = c == 0 ? 1.0 : (c == 1 ? sc_th[p] : sc_ph[p]);
          tau[p] += 0.5 * s / (sc * sc);
        }
      }
    };
    std::vector<double> ra, rb, ga[3], gb[3], ta, tb;
    density(P, ra, ga, ta);
    if (pol) density(*Pbp, rb, gb, tb);
    std::vector<double> va(npt), vb(npt), gna[3], gnb[3], vta(npt, 0.0), vtb(npt, 0.0);
    if (gga)
      for (int c = 0; c < 3; ++c) {
        gna[c].assign(npt, 0.0);
        if (pol) gnb[c].assign(npt, 0.0);
      }
    for (int p = 0; p < npt; ++p) {
      double exc;
      const double rt = pol ? ra[p] + rb[p] : ra[p];
      if (plain_lda) {
        if (pol)
          lda_polarized(xf, cf, ra[p], rb[p], thr, exc, va[p], vb[p]);
        else
          lda_unpolarized(xf, cf, ra[p], thr, exc, va[p]);
      } else if (pol) {
        const double rho[2] = {ra[p], rb[p]};
        double sig[3] = {0.0, 0.0, 0.0}, vr[2], vs[3];
        if (gga)
          for (int c = 0; c < 3; ++c) {
            sig[0] += ga[c][p] * ga[c][p];
            sig[1] += ga[c][p] * gb[c][p];
            sig[2] += gb[c][p] * gb[c][p];
          }
        const double tt[2] = {mgga ? ta[p] : 0.0, mgga ? tb[p] : 0.0};
        double vt[2];
        xc_polarized(xf, cf, rho, sig, tt, thr, exc, vr, vs, vt);
        va[p] = vr[0];
        vb[p] = vr[1];
        vta[p] = vt[0];
        vtb[p] = vt[1];
        if (gga)
          for (int c = 0; c < 3; ++c) {
            const double sc = c == 0 ? 1.0 : (c == 1 ? sc_th[p] : sc_ph[p]);
            gna[c][p] = wtot[p] * (2.0 * vs[0] * ga[c][p] + vs[1] * gb[c][p]) / sc;
            gnb[c][p] = wtot[p] * (2.0 * vs[2] * gb[c][p] + vs[1] * ga[c][p]) / sc;
          }
      } else {
        double sig = 0.0, vs;
        if (gga)
          for (int c = 0; c < 3; ++c) sig += ga[c][p] * ga[c][p];
        xc_unpolarized(xf, cf, ra[p], sig, mgga ? ta[p] : 0.0, thr, exc, va[p], vs, vta[p]);
        if (gga)
          for (int c = 0; c < 3; ++c) {
            const double sc = c == 0 ? 1.0 : (c == 1 ? sc_th[p] : sc_ph[p]);
            gna[c][p] = 2.0 * wtot[p] * vs * ga[c][p] / sc;
          }
      }
      Nel += wtot[p] * rt;
      Exc += wtot[p] * exc * rt;
      if (mgga) ekin += wtot[p] * (pol ? ta[p] + tb[p] : ta[p]);
    }
    // eval_Fxc + increment_lda_split: H += (a .* w v) a' + (b .* w v) b';
    // increment_gga_split: X = (sum_c gn_c .* da_c) a' + (sum_c gn_c .* db_c) b', H += X + X'
    auto potential = [&](const std::vector<double> &vr, const std::vector<double> *gn, const std::vector<double> &vt,
                         Mat &Hm) {
      Mat av = are, bv = aim;
      for (int p = 0; p < npt; ++p) {
        const double f = vr[p] * wtot[p];
        for (int u = 0; u < nbf; ++u) {
          av(u, p) *= f;
          bv(u, p) *= f;
        }
      }
      Mat Hs = matmul_nt(av, are);
      const Mat H2 = matmul_nt(bv, aim);
      for (size_t q = 0; q < Hs.a.size(); ++q) Hs.a[q] += H2.a[q];
      if (gga) {
        Mat gre(nbf, npt), gim(nbf, npt);
        for (int c = 0; c < 3; ++c)
          for (int p = 0; p < npt; ++p)
            for (int u = 0; u < nbf; ++u) {
              gre(u, p) += gn[c][p] * dre[c](u, p);
              gim(u, p) += gn[c][p] * dim[c](u, p);
            }
        Mat X = matmul_nt(gre, are);
        const Mat X2 = matmul_nt(gim, aim);
        for (size_t q = 0; q < X.a.size(); ++q) X.a[q] += X2.a[q];
        for (int v = 0; v < nbf; ++v)
          for (int u = 0; u < nbf; ++u) Hs(u, v) += X(u, v) + X(v, u);
      }
      // meta-GGA: H += sum_c (da_c .* w vtau/2 / scale_c^2) da_c' + (db_c ...) db_c'
      // (eval_Fxc, dftgrid.cpp:336-349, increment_lda_split per component)
      if (mgga)
        for (int c = 0; c < 3; ++c) {
          Mat dv = dre[c], ev = dim[c];
          for (int p = 0; p < npt; ++p) {
            const double sc = c == 0 ? 1.0 : (c == 1 ? sc_th[p] : sc_ph[p]);
            const double f = 0.5 * vt[p] * wtot[p] / (sc * sc);
            for (int u = 0; u < nbf; ++u) {
              dv(u, p) *= f;
              ev(u, p) *= f;
            }
          }
          const Mat T1 = matmul_nt(dv, dre[c]), T2 = matmul_nt(ev, dim[c]);
          for (size_t q = 0; q < Hs.a.size(); ++q) Hs.a[q] += T1.a[q] + T2.a[q];
        }
      for (int v = 0; v < nbf; ++v)
        for (int u = 0; u < nbf; ++u) Hm(idx[u], idx[v]) += Hs(u, v);
    };
    potential(va, gna, vta, H);
    if (pol) potential(vb, gnb, vtb, Hb);
  }
}

DiatomicXC::DiatomicXC(const DiatomicBasis &b_, int lang_, int mang_) : b(b_), lang(lang_), mang(mang_) {
  chebyshev(lang, cth, wang);
  if (mang > 0) angular_chebyshev(lang, mang, cth3, phi3, wang3);
}

// PureMDFTGrid::eval_Fxc (restricted), src/diatomic/dftgrid_purem.cpp:675-703,
// with the worker's compute_bf (:87-199), update_density (:212-260) and
// eval_Fxc (:468-500): per element and radial point, real per-m basis values
// B_j(mu) Y_l^m(nu, 0) on the nu rule, the phi integral done analytically.
void DiatomicXC::eval_Fxc(int xf, int cf, const Mat &Pin, Mat &Hout, double &Exc, double &Nel, double thr) const {
  Mat Hb;
  eval(xf, cf, Pin, nullptr, Hout, Hb, Exc, Nel, thr);
}

// PureMDFTGrid::eval_Fxc (spin-polarized), dftgrid_purem.cpp:705-740
void DiatomicXC::eval_Fxc(int xf, int cf, const Mat &Pa, const Mat &Pb, Mat &Ha, Mat &Hb, double &Exc,
                          double &Nel, double thr) const {
  eval(xf, cf, Pa, &Pb, Ha, Hb, Exc, Nel, thr);
}

// The general 3D grid (DFTGrid::eval_Fxc, src/diatomic/dftgrid.cpp:533-620):
// per element and radial point, the complex basis values B_j(mu) Y_l^m(nu, phi)
// (eval_bf) on the compound (cos nu, phi) rule, total weight
// wang * w_mu * Rh^3 sinh(mu) (sinh^2 mu + sin^2 nu) (compute_bf, :414-460),
// density and potential as in the atomic grid (update_density, increment_lda).
void DiatomicXC::eval3d(int xf, int cf, const Mat &Pin, const Mat *Pbin, Mat &Hout, Mat &Hbout, double &Exc,
                        double &Nel, double thr) const {
  xc_check(xf, cf);
  const bool pol = Pbin != nullptr;
  const bool gga = xc_is_gga(xf) || xc_is_gga(cf);
  const bool mgga = xc_is_mgga(xf) || xc_is_mgga(cf);
  ekin = 0.0;
  const bool plain_lda = !gga && (xf == 0 || xf == 1) && (cf == 0 || cf == 7);
  const RadialBasis &rad = b.radial;
  const Mat P = b.expand_boundaries(Pin);
  const Mat Pbx = pol ? b.expand_boundaries(*Pbin) : Mat();
  const int Na = b.Nang(), Nr = rad.Nbf(), nang = (int)wang3.size();
  Mat H(b.Ndummy(), b.Ndummy()), H2(pol ? b.Ndummy() : 0, pol ? b.Ndummy() : 0);
  Exc = 0.0;
  Nel = 0.0;
  const double Rh3 = std::pow(b.Rhalf, 3);
  for (int iel = 0; iel < rad.Nel(); ++iel) {
    const Mat Bv = rad.eval_dnf(rad.xq, 0, iel);  // (nq x Ni)
    const Mat Bd = gga ? rad.eval_dnf(rad.xq, 1, iel) : Mat();
    const std::vector<double> mu = rad.eval_coord(rad.xq, iel);
    const double h = rad.scaling_factor(iel);
    const int nrad = (int)mu.size(), Ni = Bv.c, nbf = Na * Ni, npt = nrad * nang;
    std::vector<int> idx(nbf);
    for (int a = 0; a < Na; ++a)
      for (int j = 0; j < Ni; ++j) idx[a * Ni + j] = a * Nr + rad.first[iel] + j;
    // weights and the scale factors h_mu = h_nu = Rh sqrt(sinh^2 mu + sin^2 nu),
    // h_phi = Rh sinh(mu) sin(nu) (compute_bf, dftgrid.cpp:414-450)
    std::vector<double> wtot(npt), sc[3];
    for (int c = 0; c < 3; ++c) sc[c].assign(npt, 1.0);
    for (int ia = 0; ia < nang; ++ia) {
      const double sth = std::sqrt(std::max((1.0 - cth3[ia]) * (1.0 + cth3[ia]), 0.0));
      for (int ir = 0; ir < nrad; ++ir) {
        const double shmu = std::sinh(mu[ir]);
        const int p = ia * nrad + ir;
        wtot[p] = wang3[ia] * (h * rad.wq[ir]) * Rh3 * shmu * (shmu * shmu + sth * sth);
        sc[0][p] = sc[1][p] = b.Rhalf * std::sqrt(shmu * shmu + sth * sth);
        sc[2][p] = b.Rhalf * shmu * sth;
      }
    }
    Mat are(nbf, npt), aim(nbf, npt);  // real and imaginary parts of conj(chi)
    Mat dre[3], dim[3];                // the same of conj(d_c chi), c = mu, nu, phi (eval_df, basis.cpp:2204-2243)
    if (gga)
      for (int c = 0; c < 3; ++c) dre[c] = dim[c] = Mat(nbf, npt);
    for (int ia = 0; ia < nang; ++ia)
      for (int a = 0; a < Na; ++a) {
        const int l = b.lval[a], m = b.mval[a];
        const std::complex<double> Y = spherical_harmonic(l, m, cth3[ia], phi3[ia]);
        const std::complex<double> y = std::conj(Y);
        std::complex<double> yc[3];
        if (gga) {
          const double sinth = std::sqrt(std::max((1.0 - cth3[ia]) * (1.0 + cth3[ia]), 0.0));
          const double cotth = sinth > 0.0 ? cth3[ia] / sinth : 0.0;
          std::complex<double> angfac = double(m) * cotth * Y;
          if (m < l)
            angfac += std::sqrt(double((l - m) * (l + m + 1))) * std::exp(std::complex<double>(0.0, -phi3[ia])) *
                      spherical_harmonic(l, m + 1, cth3[ia], phi3[ia]);
          yc[0] = y;
          yc[1] = std::conj(angfac);
          yc[2] = std::conj(std::complex<double>(0.0, m) * Y);
        }
        for (int j = 0; j < Ni; ++j)
          for (int ir = 0; ir < nrad; ++ir) {
            const int u = a * Ni + j, p = ia * nrad + ir;
            are(u, p) = y.real() * Bv(ir, j);
            aim(u, p) = y.imag() * Bv(ir, j);
            if (gga) {
              dre[0](u, p) = yc[0].real() * Bd(ir, j);
              dim[0](u, p) = yc[0].imag() * Bd(ir, j);
              for (int c = 1; c < 3; ++c) {
                dre[c](u, p) = yc[c].real() * Bv(ir, j);
                dim[c](u, p) = yc[c].imag() * Bv(ir, j);
              }
            }
          }
      }
    auto density = [&](const Mat &Pm, std::vector<double> &rho, std::vector<double> g[3], std::vector<double> &tau) {
      Mat Pe(nbf, nbf);
      for (int v = 0; v < nbf; ++v)
        for (int u = 0; u < nbf; ++u) Pe(u, v) = Pm(idx[u], idx[v]);
      const Mat PA = matmul(Pe, are), PB = matmul(Pe, aim);
      rho.assign(npt, 0.0);
      for (int p = 0; p < npt; ++p) {
        double s = 0.0, t = 0.0;
        for (int u = 0; u < nbf; ++u) {
          s += PA(u, p) * are(u, p);
          t += PB(u, p) * aim(u, p);
        }
        rho[p] = s + t;
      }
      if (!gga) return;
      for (int c = 0; c < 3; ++c) {
        g[c].assign(npt, 0.0);
        for (int p = 0; p < npt; ++p) {
          double s = 0.0, t = 0.0;
          for (int u = 0; u < nbf; ++u) {
            s += PA(u, p) * dre[c](u, p);
            t += PB(u, p) * dim[c](u, p);
          }
          g[c][p] = 2.0 * (s + t) / sc[c][p];
        }
      }
      // kinetic energy density tau = 1/2 sum_c ((P da_c).da_c + (P db_c).db_c) / scale_c^2
      // (update_density, dftgrid.cpp:89-107, 186-214)
      if (!mgga) return;
      tau.assign(npt, 0.0);
      for (int c = 0; c < 3; ++c) {
        const Mat QA = matmul(Pe, dre[c]), QB = matmul(Pe, dim[c]);
        for (int p = 0; p < npt; ++p) {
          double s = 0.0;
          for (int u = 0; u < nbf; ++u) s += QA(u, p) * dre[c](u, p) + QB(u, p) * dim[c](u, p);
          const double hc = sc[c][p];
          tau[p] += 0.5 * s / (hc * hc);
        }
      }
    };
    std::vector<double> ra, rb, ga[3], gb[3], ta, tb;
    density(P, ra, ga, ta);
    if (pol) density(Pbx, rb, gb, tb);
    std::vector<double> va(npt), vb(npt), gna[3], gnb[3], vta(npt, 0.0), vtb(npt, 0.0);
    if (gga)
      for (int c = 0; c < 3; ++c) {
        gna[c].assign(npt, 0.0);
        if (pol) gnb[c].assign(npt, 0.0);
      }
    for (int p = 0; p < npt; ++p) {
      double exc;
      const double rt = pol ? ra[p] + rb[p] : ra[p];
      if (plain_lda) {
        if (pol)
          lda_polarized(xf, cf, ra[p], rb[p], thr, exc, va[p], vb[p]);
        else
          lda_unpolarized(xf, cf, ra[p], thr, exc, va[p]);
      } else if (pol) {
        const double rho[2] = {ra[p], rb[p]};
        double sig[3] = {0.0, 0.0, 0.0}, vr[2], vs[3];
        if (gga)
          for (int c = 0; c < 3; ++c) {
            sig[0] += ga[c][p] * ga[c][p];
            sig[1] += ga[c][p] * gb[c][p];
            sig[2] += gb[c][p] * gb[c][p];
          }
        const double tt[2] = {mgga ? ta[p] : 0.0, mgga ? tb[p] : 0.0};
        double vt[2];
        xc_polarized(xf, cf, rho, sig, tt, thr, exc, vr, vs, vt);
        va[p] = vr[0];
        vb[p] = vr[1];
        vta[p] = vt[0];
        vtb[p] = vt[1];
        if (gga)
          for (int c = 0; c < 3; ++c) {
            gna[c][p] = wtot[p] * (2.0 * vs[0] * ga[c][p] + vs[1] * gb[c][p]) / sc[c][p];
            gnb[c][p] = wtot[p] * (2.0 * vs[2] * gb[c][p] + vs[1] * ga[c][p]) / sc[c][p];
          }
      } else {
        double sig = 0.0, vs;
        if (gga)
          for (int c = 0; c < 3; ++c) sig += ga[c][p] * ga[c][p];
        xc_unpolarized(xf, cf, ra[p], sig, mgga ? ta[p] : 0.0, thr, exc, va[p], vs, vta[p]);
        if (gga)
          for (int c = 0; c < 3; ++c) gna[c][p] = 2.0 * wtot[p] * vs * ga[c][p] / sc[c][p];
      }
      Nel += wtot[p] * rt;
      Exc += wtot[p] * exc * rt;
      if (mgga) ekin += wtot[p] * (pol ? ta[p] + tb[p] : ta[p]);
    }
    auto potential = [&](const std::vector<double> &vr, const std::vector<double> *gn, const std::vector<double> &vt,
                         Mat &Hm) {
      Mat av = are, bv = aim;
      for (int p = 0; p < npt; ++p) {
        const double f = vr[p] * wtot[p];
        for (int u = 0; u < nbf; ++u) {
          av(u, p) *= f;
          bv(u, p) *= f;
        }
      }
      Mat Hs = matmul_nt(av, are);
      const Mat H2s = matmul_nt(bv, aim);
      for (size_t q = 0; q < Hs.a.size(); ++q) Hs.a[q] += H2s.a[q];
      if (gga) {
        Mat gre(nbf, npt), gim(nbf, npt);
        for (int c = 0; c < 3; ++c)
          for (int p = 0; p < npt; ++p)
            for (int u = 0; u < nbf; ++u) {
              gre(u, p) += gn[c][p] * dre[c](u, p);
              gim(u, p) += gn[c][p] * dim[c](u, p);
            }
        Mat X = matmul_nt(gre, are);
        const Mat X2 = matmul_nt(gim, aim);
        for (size_t q = 0; q < X.a.size(); ++q) X.a[q] += X2.a[q];
        for (int v = 0; v < nbf; ++v)
          for (int u = 0; u < nbf; ++u) Hs(u, v) += X(u, v) + X(v, u);
      }
      // meta-GGA: H += sum_c (da_c .* w vtau/2 / scale_c^2) da_c' + (db_c ...) db_c'
      // (eval_Fxc, dftgrid.cpp:336-349, increment_lda_split per component)
      if (mgga)
        for (int c = 0; c < 3; ++c) {
          Mat dv = dre[c], ev = dim[c];
          for (int p = 0; p < npt; ++p) {
            const double hc = sc[c][p];
            const double f = 0.5 * vt[p] * wtot[p] / (hc * hc);
            for (int u = 0; u < nbf; ++u) {
              dv(u, p) *= f;
              ev(u, p) *= f;
            }
          }
          const Mat T1 = matmul_nt(dv, dre[c]), T2 = matmul_nt(ev, dim[c]);
          for (size_t q = 0; q < Hs.a.size(); ++q) Hs.a[q] += T1.a[q] + T2.a[q];
        }
      for (int v = 0; v < nbf; ++v)
        for (int u = 0; u < nbf; ++u) Hm(idx[u], idx[v]) += Hs(u, v);
    };
    potential(va, gna, vta, H);
    if (pol) potential(vb, gnb, vtb, H2);
  }
  Hout = b.remove_boundaries(H);
  if (pol) Hbout = b.remove_boundaries(H2);
}

// GGA on the pure-m grid (dftgrid_purem.cpp:87-199, 248-260, 382-395,
// 468-500): per m block the mu and nu derivatives of B_j(mu) Y_lm(nu, 0)
// (dr_m = Y B', dth_m = dY/dnu B), gradient g_c = 2 sum (P f) d_c f / h with
// h = h_mu = h_nu = Rh sqrt(sinh^2 mu + sin^2 nu) (no phi component: the
// density is phi-independent), potential X = (sum_c gn_c d_c) f', H += X + X'.
void DiatomicXC::eval(int xf, int cf, const Mat &Pin, const Mat *Pbin, Mat &Hout, Mat &Hbout, double &Exc,
                      double &Nel, double thr) const {
  if (mang > 0) return eval3d(xf, cf, Pin, Pbin, Hout, Hbout, Exc, Nel, thr);
  xc_check(xf, cf);
  const bool pol = Pbin != nullptr;
  const bool gga = xc_is_gga(xf) || xc_is_gga(cf);
  const bool mgga = xc_is_mgga(xf) || xc_is_mgga(cf);
  ekin = 0.0;
  const bool plain_lda = !gga && (xf == 0 || xf == 1) && (cf == 0 || cf == 7);
  const RadialBasis &rad = b.radial;
  const Mat P = b.expand_boundaries(Pin);
  const Mat Pbx = pol ? b.expand_boundaries(*Pbin) : Mat();
  const int Na = b.Nang(), Nr = rad.Nbf(), nang = (int)cth.size();
  Mat H(b.Ndummy(), b.Ndummy()), H2(pol ? b.Ndummy() : 0, pol ? b.Ndummy() : 0);
  Exc = 0.0;
  Nel = 0.0;
  std::vector<int> mlist(b.mval.begin(), b.mval.end());
  std::sort(mlist.begin(), mlist.end());
  mlist.erase(std::unique(mlist.begin(), mlist.end()), mlist.end());
  const double Rh3 = std::pow(b.Rhalf, 3);
  for (int iel = 0; iel < rad.Nel(); ++iel) {
    const Mat Bv = rad.eval_dnf(rad.xq, 0, iel);  // (nq x Ni)
    const Mat Bd = gga ? rad.eval_dnf(rad.xq, 1, iel) : Mat();
    const std::vector<double> mu = rad.eval_coord(rad.xq, iel);
    const double h = rad.scaling_factor(iel);
    const int Ni = Bv.c;
    for (int ir = 0; ir < (int)mu.size(); ++ir) {
      const double wr = h * rad.wq[ir], shmu = std::sinh(mu[ir]);
      std::vector<double> wtot(nang), hs(nang), hphi(nang);
      for (int ia = 0; ia < nang; ++ia) {
        const double sth = std::sqrt(std::max((1.0 - cth[ia]) * (1.0 + cth[ia]), 0.0));
        wtot[ia] = 2.0 * kPi * wang[ia] * wr * Rh3 * shmu * (shmu * shmu + sth * sth);
        hs[ia] = b.Rhalf * std::sqrt(shmu * shmu + sth * sth);
        hphi[ia] = b.Rhalf * shmu * sth;
      }
      std::vector<std::vector<int>> idx(mlist.size());
      std::vector<Mat> f(mlist.size()), fr(mlist.size()), ft(mlist.size());
      std::vector<double> rho(nang, 0.0), rhob(nang, 0.0), taua(nang, 0.0), taub(nang, 0.0);
      std::vector<double> ga[2] = {std::vector<double>(nang, 0.0), std::vector<double>(nang, 0.0)};
      std::vector<double> gb[2] = {std::vector<double>(nang, 0.0), std::vector<double>(nang, 0.0)};
      for (size_t im = 0; im < mlist.size(); ++im) {
        for (int a = 0; a < Na; ++a)
          if (b.mval[a] == mlist[im])
            for (int j = 0; j < Ni; ++j) idx[im].push_back(a * Nr + rad.first[iel] + j);
        const int nbf = (int)idx[im].size();
        f[im] = Mat(nbf, nang);
        if (gga) fr[im] = ft[im] = Mat(nbf, nang);
        int row = 0;
        for (int a = 0; a < Na; ++a) {
          if (b.mval[a] != mlist[im]) continue;
          const int l = b.lval[a], m = b.mval[a];
          for (int ia = 0; ia < nang; ++ia) {
            const double Y = std::real(spherical_harmonic(l, m, cth[ia], 0.0));
            double dY = 0.0;
            if (gga) {
              const double sinth = std::sqrt(std::max((1.0 - cth[ia]) * (1.0 + cth[ia]), 0.0));
              const double cotth = sinth > 0.0 ? cth[ia] / sinth : 0.0;
              dY = m * cotth * Y;
              if (m < l) dY += std::sqrt(double((l - m) * (l + m + 1))) * std::real(spherical_harmonic(l, m + 1, cth[ia], 0.0));
            }
            for (int j = 0; j < Ni; ++j) {
              f[im](row + j, ia) = Y * Bv(ir, j);
              if (gga) {
                fr[im](row + j, ia) = Y * Bd(ir, j);
                ft[im](row + j, ia) = dY * Bv(ir, j);
              }
            }
          }
          row += Ni;
        }
        auto block_density = [&](const Mat &Pm, std::vector<double> &r, std::vector<double> *g,
                                 std::vector<double> &tau) {
          Mat Pblk(nbf, nbf);
          for (int v = 0; v < nbf; ++v)
            for (int u = 0; u < nbf; ++u) Pblk(u, v) = Pm(idx[im][u], idx[im][v]);
          const Mat Pv = matmul(Pblk, f[im]);
          for (int ia = 0; ia < nang; ++ia) {
            double s = 0.0, s0 = 0.0, s1 = 0.0;
            for (int u = 0; u < nbf; ++u) {
              s += Pv(u, ia) * f[im](u, ia);
              if (gga) {
                s0 += Pv(u, ia) * fr[im](u, ia);
                s1 += Pv(u, ia) * ft[im](u, ia);
              }
            }
            r[ia] += s;
            if (gga) {
              g[0][ia] += 2.0 * s0;
              g[1][ia] += 2.0 * s1;
            }
          }
          // tau = 1/2 [((P dr).dr + (P dth).dth) / h^2 + m^2 rho_m / h_phi^2]
          // (update_density, dftgrid_purem.cpp:287-318)
          if (mgga) {
            const Mat Qr = matmul(Pblk, fr[im]), Qt = matmul(Pblk, ft[im]);
            const double m2 = double(mlist[im]) * mlist[im];
            for (int ia = 0; ia < nang; ++ia) {
              double kr = 0.0, kt = 0.0, rm = 0.0;
              for (int u = 0; u < nbf; ++u) {
                kr += Qr(u, ia) * fr[im](u, ia);
                kt += Qt(u, ia) * ft[im](u, ia);
                rm += Pv(u, ia) * f[im](u, ia);
              }
              const double kp = hphi[ia] > 0.0 ? m2 * rm / (hphi[ia] * hphi[ia]) : 0.0;
              tau[ia] += 0.5 * ((kr + kt) / (hs[ia] * hs[ia]) + kp);
            }
          }
        };
        block_density(P, rho, ga, taua);
        if (pol) block_density(Pbx, rhob, gb, taub);
      }
      std::vector<double> wv(nang), wvb(nang), gna[2], gnb[2], vta(nang, 0.0), vtb(nang, 0.0);
      for (int c = 0; c < 2; ++c) {
        gna[c].assign(nang, 0.0);
        gnb[c].assign(nang, 0.0);
      }
      for (int ia = 0; ia < nang; ++ia) {
        double exc, v, vb2 = 0.0;
        for (int c = 0; c < 2 && gga; ++c) {
          ga[c][ia] /= hs[ia];
          gb[c][ia] /= hs[ia];
        }
        if (plain_lda) {
          if (pol)
            lda_polarized(xf, cf, rho[ia], rhob[ia], thr, exc, v, vb2);
          else
            lda_unpolarized(xf, cf, rho[ia], thr, exc, v);
        } else if (pol) {
          const double rr[2] = {rho[ia], rhob[ia]};
          const double sig[3] = {ga[0][ia] * ga[0][ia] + ga[1][ia] * ga[1][ia],
                                 ga[0][ia] * gb[0][ia] + ga[1][ia] * gb[1][ia],
                                 gb[0][ia] * gb[0][ia] + gb[1][ia] * gb[1][ia]};
          double vr[2], vs[3], vt[2];
          const double tt[2] = {taua[ia], taub[ia]};
          xc_polarized(xf, cf, rr, sig, tt, thr, exc, vr, vs, vt);
          v = vr[0];
          vb2 = vr[1];
          vta[ia] = vt[0];
          vtb[ia] = vt[1];
          for (int c = 0; c < 2 && gga; ++c) {
            gna[c][ia] = wtot[ia] * (2.0 * vs[0] * ga[c][ia] + vs[1] * gb[c][ia]) / hs[ia];
            gnb[c][ia] = wtot[ia] * (2.0 * vs[2] * gb[c][ia] + vs[1] * ga[c][ia]) / hs[ia];
          }
        } else {
          double vs;
          xc_unpolarized(xf, cf, rho[ia], ga[0][ia] * ga[0][ia] + ga[1][ia] * ga[1][ia], taua[ia], thr, exc, v, vs,
                         vta[ia]);
          for (int c = 0; c < 2 && gga; ++c) gna[c][ia] = 2.0 * wtot[ia] * vs * ga[c][ia] / hs[ia];
        }
        const double rt = pol ? rho[ia] + rhob[ia] : rho[ia];
        Nel += wtot[ia] * rt;
        Exc += wtot[ia] * exc * rt;
        if (mgga) ekin += wtot[ia] * (pol ? taua[ia] + taub[ia] : taua[ia]);
        wv[ia] = v * wtot[ia];
        wvb[ia] = vb2 * wtot[ia];
      }
      auto potential = [&](const std::vector<double> &w, const std::vector<double> *gn, const std::vector<double> &vt,
                           Mat &Hm) {
        for (size_t im = 0; im < mlist.size(); ++im) {
          Mat fw = f[im];
          for (int ia = 0; ia < nang; ++ia)
            for (int u = 0; u < fw.r; ++u) fw(u, ia) *= w[ia];
          Mat Hblk = matmul_nt(fw, f[im]);
          if (gga) {
            Mat gr(fw.r, nang);
            for (int ia = 0; ia < nang; ++ia)
              for (int u = 0; u < fw.r; ++u) gr(u, ia) = gn[0][ia] * fr[im](u, ia) + gn[1][ia] * ft[im](u, ia);
            const Mat X = matmul_nt(gr, f[im]);
            for (int v = 0; v < X.c; ++v)
              for (int u = 0; u < X.r; ++u) Hblk(u, v) += X(u, v) + X(v, u);
          }
          // meta-GGA: w vtau/2 / h^2 on dr and dth, m^2 w vtau/2 / h_phi^2 on f
          // (eval_Fxc, dftgrid_purem.cpp:482-494)
          if (mgga) {
            const double m2 = double(mlist[im]) * mlist[im];
            Mat a = fr[im], c = ft[im], e = f[im];
            for (int ia = 0; ia < nang; ++ia) {
              const double vtl = 0.5 * vt[ia] * wtot[ia];
              const double wr = vtl / (hs[ia] * hs[ia]);
              const double wp = hphi[ia] > 0.0 ? m2 * vtl / (hphi[ia] * hphi[ia]) : 0.0;
              for (int u = 0; u < a.r; ++u) {
                a(u, ia) *= wr;
                c(u, ia) *= wr;
                e(u, ia) *= wp;
              }
            }
            const Mat T1 = matmul_nt(a, fr[im]), T2 = matmul_nt(c, ft[im]), T3 = matmul_nt(e, f[im]);
            for (size_t q = 0; q < Hblk.a.size(); ++q) Hblk.a[q] += T1.a[q] + T2.a[q] + T3.a[q];
          }
          for (int v = 0; v < Hblk.c; ++v)
            for (int u = 0; u < Hblk.r; ++u) Hm(idx[im][u], idx[im][v]) += Hblk(u, v);
        }
      };
      potential(wv, gna, vta, H);
      if (pol) potential(wvb, gnb, vtb, H2);
    }
  }
  Hout = b.remove_boundaries(H);
  if (pol) Hbout = b.remove_boundaries(H2);
}

}  // namespace orc
