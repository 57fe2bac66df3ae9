// Oracle (test infrastructure only): LDA exchange-correlation declarations.
#pragma once
#include "orc.hpp"
#include <complex>

namespace orc {

// lda_x (1) + lda_c_vwn (7, VWN5), spin-unpolarized; thr as libxc's density threshold
void lda_unpolarized(int xf, int cf, double rho, double thr, double &exc, double &vrho);
void lda_polarized(int xf, int cf, double ra, double rb, double thr, double &exc, double &va, double &vb);
// Functionals by id (lda_x 1, lda_c_vwn 7, lda_c_vwn_rpa 8, lda_c_pw 12, lda_c_pw_mod 13,
// gga_x_pbe 101, gga_x_b88 106, gga_c_pbe 130, gga_c_lyp 131, and the hybrids
// hyb_gga_xc_b3lyp 402, hyb_gga_xc_pbeh 406, hyb_gga_xc_cam_b3lyp 433 and
// hyb_gga_xc_camy_b3lyp 470; xcfunc.cpp): energy per particle exc,
// vrho = d(rho exc)/drho and vsigma = d(rho exc)/dsigma; polarized sigma is
// (aa, ab, bb) and vsigma likewise.
// mgga_x_tpss 202 and mgga_c_tpss 231 also take the kinetic energy density
// tau (per spin when polarized) and give vtau = d(rho exc)/dtau.
bool xc_is_gga(int id);   // needs the density gradient (GGA or meta-GGA)
bool xc_is_mgga(int id);  // needs tau
void xc_check(int xf, int cf);
// exact-exchange fractions of a hybrid: full-range kfrac, short-range kshort
// at range-separation omega with the kernel (1 erfc, 2 Yukawa) of its
// short-range part (dftfuncs.cpp range_separation, is_range_separated); zero otherwise
void xc_exact_exchange(int xf, double &kfrac, double &kshort, double &omega, int &kernel);
void xc_unpolarized(int xf, int cf, double rho, double sigma, double tau, double thr, double &exc, double &vrho,
                    double &vsigma, double &vtau);
void xc_polarized(int xf, int cf, const double rho[2], const double sigma[3], const double tau[2], double thr,
                  double &exc, double vrho[2], double vsigma[3], double vtau[2]);
void angular_chebyshev(int l, int m, std::vector<double> &cth, std::vector<double> &phi, std::vector<double> &w);
std::complex<double> spherical_harmonic(int l, int m, double cth, double phi);

// atomic::dftgrid::DFTGrid (src/atomic/dftgrid.h:140-173), LDA restricted
struct AtomicXC {
  const AtomicBasis &b;
  int lang, mang;
  mutable double ekin = 0.0;  // integral of tau over the grid, meta-GGA (compute_Ekin)
  std::vector<double> cth, phi, wang;
  AtomicXC(const AtomicBasis &b, int lang, int mang);
  void eval_Fxc(int xf, int cf, const Mat &P, Mat &H, double &Exc, double &Nel, double thr) const;
  void eval_Fxc(int xf, int cf, const Mat &Pa, const Mat &Pb, Mat &Ha, Mat &Hb, double &Exc, double &Nel,
                double thr) const;
  void eval(int xf, int cf, const Mat &P, const Mat *Pb, Mat &H, Mat &Hb, double &Exc, double &Nel,
            double thr) const;
};

struct DiatomicBasis;
// diatomic::dftgrid_purem::PureMDFTGrid (src/diatomic/dftgrid_purem.h) when
// mang == 0; the general 3D diatomic::dftgrid::DFTGrid (src/diatomic/dftgrid.h,
// dftgrid.cpp:533-620) on the lang x mang (cos nu, phi) rule when mang > 0.
struct DiatomicXC {
  const DiatomicBasis &b;
  int lang, mang;
  mutable double ekin = 0.0;  // integral of tau over the grid, meta-GGA (compute_Ekin)
  std::vector<double> cth, wang, phi3, cth3, wang3;
  DiatomicXC(const DiatomicBasis &b, int lang, int mang = 0);
  void eval_Fxc(int xf, int cf, const Mat &P, Mat &H, double &Exc, double &Nel, double thr) const;
  void eval_Fxc(int xf, int cf, const Mat &Pa, const Mat &Pb, Mat &Ha, Mat &Hb, double &Exc, double &Nel,
                double thr) const;
  void eval(int xf, int cf, const Mat &P, const Mat *Pb, Mat &H, Mat &Hb, double &Exc, double &Nel,
            double thr) const;
  void eval3d(int xf, int cf, const Mat &P, const Mat *Pb, Mat &H, Mat &Hb, double &Exc, double &Nel,
              double thr) const;
};

}  // namespace orc
