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
  const RadialBasis &rad = b.radial;
  const int Na = b.Nang(), Nr = rad.Nbf();
  H = Mat(P.r, P.c);
  Exc = 0.0;
  Nel = 0.0;
  const int nang = (int)wang.size();
  for (int iel = 0; iel < rad.Nel(); ++iel) {
    // compute_bf, dftgrid.cpp:470-576
    const Mat R = rad.get_bf(rad.xq, iel);  // (nrad x Ni)
    const std::vector<double> r = rad.eval_coord(rad.xq, iel);
    const double h = rad.scaling_factor(iel);
    const int nrad = (int)rad.xq.size(), Ni = R.c, nbf = Na * Ni, npt = nrad * nang;
    std::vector<int> idx(nbf);
    for (int a = 0; a < Na; ++a)
      for (int j = 0; j < Ni; ++j) idx[a * Ni + j] = a * Nr + rad.first[iel] + j;
    std::vector<double> wtot(npt);
    for (int ia = 0; ia < nang; ++ia)
      for (int ir = 0; ir < nrad; ++ir) wtot[ia * nrad + ir] = wang[ia] * (h * rad.wq[ir]) * std::pow(r[ir], 2);
    Mat are(nbf, npt), aim(nbf, npt);  // real and imaginary parts of conj(chi)
    for (int ia = 0; ia < nang; ++ia)
      for (int a = 0; a < Na; ++a) {
        const std::complex<double> y = std::conj(spherical_harmonic(b.lval[a], b.mval[a], cth[ia], phi[ia]));
        for (int j = 0; j < Ni; ++j)
          for (int ir = 0; ir < nrad; ++ir) {
            are(a * Ni + j, ia * nrad + ir) = y.real() * R(ir, j);
            aim(a * Ni + j, ia * nrad + ir) = y.imag() * R(ir, j);
          }
      }
    // update_density: rho = colwise (P a).a + (P b).b
    Mat Pe(nbf, nbf);
    for (int v = 0; v < nbf; ++v)
      for (int u = 0; u < nbf; ++u) Pe(u, v) = P(idx[u], idx[v]);
    const Mat PA = matmul(Pe, are), PB = matmul(Pe, aim);
    std::vector<double> rho(npt), exc(npt), vrho(npt);
    for (int p = 0; p < npt; ++p) {
      double s = 0.0, t = 0.0;
      for (int u = 0; u < nbf; ++u) {
        s += PA(u, p) * are(u, p);
        t += PB(u, p) * aim(u, p);
      }
      rho[p] = s + t;
    }
    for (int p = 0; p < npt; ++p) {
      Nel += wtot[p] * rho[p];
      lda_unpolarized(xf, cf, rho[p], thr, exc[p], vrho[p]);
      Exc += wtot[p] * exc[p] * rho[p];
    }
    // eval_Fxc + increment_lda_split: H += (a .* w v) a' + (b .* w v) b'
    Mat av = are, bv = aim;
    for (int p = 0; p < npt; ++p) {
      const double f = vrho[p] * wtot[p];
      for (int u = 0; u < nbf; ++u) {
        av(u, p) *= f;
        bv(u, p) *= f;
      }
    }
    const Mat Ha = matmul_nt(av, are), Hb = matmul_nt(bv, aim);
    for (int v = 0; v < nbf; ++v)
      for (int u = 0; u < nbf; ++u) H(idx[u], idx[v]) += Ha(u, v) + Hb(u, v);
  }
}

DiatomicXC::DiatomicXC(const DiatomicBasis &b_, int lang_) : b(b_), lang(lang_) { chebyshev(lang, cth, wang); }

// PureMDFTGrid::eval_Fxc (restricted), src/diatomic/dftgrid_purem.cpp:675-703,
// with the worker's compute_bf (:87-199), update_density (:212-260) and
// eval_Fxc (:468-500): per element and radial point, real per-m basis values
// B_j(mu) Y_l^m(nu, 0) on the nu rule, the phi integral done analytically.
void DiatomicXC::eval_Fxc(int xf, int cf, const Mat &Pin, Mat &Hout, double &Exc, double &Nel, double thr) const {
  const RadialBasis &rad = b.radial;
  const Mat P = b.expand_boundaries(Pin);
  const int Na = b.Nang(), Nr = rad.Nbf(), nang = (int)cth.size();
  Mat H(b.Ndummy(), b.Ndummy());
  Exc = 0.0;
  Nel = 0.0;
  std::vector<int> mlist(b.mval.begin(), b.mval.end());
  std::sort(mlist.begin(), mlist.end());
  mlist.erase(std::unique(mlist.begin(), mlist.end()), mlist.end());
  const double Rh3 = std::pow(b.Rhalf, 3);
  for (int iel = 0; iel < rad.Nel(); ++iel) {
    const Mat Bv = rad.eval_dnf(rad.xq, 0, iel);  // (nq x Ni)
    const std::vector<double> mu = rad.eval_coord(rad.xq, iel);
    const double h = rad.scaling_factor(iel);
    const int Ni = Bv.c;
    for (int ir = 0; ir < (int)mu.size(); ++ir) {
      const double wr = h * rad.wq[ir], shmu = std::sinh(mu[ir]);
      std::vector<double> wtot(nang);
      for (int ia = 0; ia < nang; ++ia) {
        const double sth = std::sqrt(std::max((1.0 - cth[ia]) * (1.0 + cth[ia]), 0.0));
        wtot[ia] = 2.0 * kPi * wang[ia] * wr * Rh3 * shmu * (shmu * shmu + sth * sth);
      }
      std::vector<std::vector<int>> idx(mlist.size());
      std::vector<Mat> f(mlist.size());
      std::vector<double> rho(nang, 0.0);
      for (size_t im = 0; im < mlist.size(); ++im) {
        for (int a = 0; a < Na; ++a)
          if (b.mval[a] == mlist[im])
            for (int j = 0; j < Ni; ++j) idx[im].push_back(a * Nr + rad.first[iel] + j);
        const int nbf = (int)idx[im].size();
        f[im] = Mat(nbf, nang);
        int row = 0;
        for (int a = 0; a < Na; ++a) {
          if (b.mval[a] != mlist[im]) continue;
          for (int ia = 0; ia < nang; ++ia) {
            const double Y = std::real(spherical_harmonic(b.lval[a], b.mval[a], cth[ia], 0.0));
            for (int j = 0; j < Ni; ++j) f[im](row + j, ia) = Y * Bv(ir, j);
          }
          row += Ni;
        }
        Mat Pb(nbf, nbf);
        for (int v = 0; v < nbf; ++v)
          for (int u = 0; u < nbf; ++u) Pb(u, v) = P(idx[im][u], idx[im][v]);
        const Mat Pv = matmul(Pb, f[im]);
        for (int ia = 0; ia < nang; ++ia) {
          double s = 0.0;
          for (int u = 0; u < nbf; ++u) s += Pv(u, ia) * f[im](u, ia);
          rho[ia] += s;
        }
      }
      std::vector<double> wv(nang);
      for (int ia = 0; ia < nang; ++ia) {
        double exc, v;
        lda_unpolarized(xf, cf, rho[ia], thr, exc, v);
        Nel += wtot[ia] * rho[ia];
        Exc += wtot[ia] * exc * rho[ia];
        wv[ia] = v * wtot[ia];
      }
      for (size_t im = 0; im < mlist.size(); ++im) {
        Mat fw = f[im];
        for (int ia = 0; ia < nang; ++ia)
          for (int u = 0; u < fw.r; ++u) fw(u, ia) *= wv[ia];
        const Mat Hm = matmul_nt(fw, f[im]);
        for (int v = 0; v < Hm.c; ++v)
          for (int u = 0; u < Hm.r; ++u) H(idx[im][u], idx[im][v]) += Hm(u, v);
      }
    }
  }
  Hout = b.remove_boundaries(H);
}

}  // namespace orc
