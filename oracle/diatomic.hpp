// Oracle (test infrastructure only): diatomic TwoDBasis declarations.
#pragma once
#include "orc.hpp"

namespace orc {

// Gaunt helpers with the semantics of gaunt.h:53-132 (coeff / mod_coeff)
// and gaunt.cpp:275-300 (cosine couplings).
double gaunt_coeff(int L, int M, int l, int m, int lp);
double gaunt_mod_coeff(int lj, int mj, int L, int M, int li, int mi);
double gaunt_cosine(int lj, int mj, int li, int mi);
double gaunt_cosine2(int lj, int mj, int li, int mi);
void diatomic_angular(const std::vector<int> &lmmax, std::vector<int> &lval, std::vector<int> &mval);

// Element data of the in-element two-electron integral
// (quadrature::TwoElectronElement, src/diatomic/quadrature.h:113-137).
struct TwoEl {
  std::vector<double> x, w, mu, chmu, shmu;
  double mulen = 0.0;
  Mat bf, bfprod;
  struct Sub {
    double mulen = 0.0;
    std::vector<double> chmu, shmu;
    Mat bf;
  };
  std::vector<Sub> sub;
};

struct DiatomicBasis {
  int Z1 = 0, Z2 = 0;
  double Rhalf = 0.0;
  int nquad = 0;
  RadialBasis radial;
  std::vector<int> lval, mval;
  std::vector<std::pair<int, int>> lm_map;  // (L, |M|), |M|-major
  std::vector<std::pair<int, int>> LM_map;  // (L, M), L-major
  std::vector<int> pure_idx;
  std::vector<Mat> P0, P2, Q0, Q2, cd_B;    // [ilm*Nel + iel]
  std::vector<std::vector<double>> cd_sigma;
  DiatomicBasis(int Z1, int Z2, double Rhalf, int nnodes, int nquad, const std::vector<double> &bval,
                const std::vector<int> &lval, const std::vector<int> &mval);
  int Nrad() const { return radial.Nbf(); }
  int Nang() const { return (int)lval.size(); }
  int Ndummy() const { return Nrad() * Nang(); }
  int Nbf() const { return (int)pure_idx.size(); }
  int lmind(int L, int M) const;
  int LMind(int L, int M) const;
  double LMfac(int L, int M) const;
  Mat radial_integral(int m, int n) const;
  Mat radial_kinetic() const;
  Mat PQ_integral(bool Q, int k, int iel, int L, LegendreColumn &leg) const;
  TwoEl twoe_element(int iel, int n) const;
  Mat twoe_integral(const TwoEl &el, int k, int l, int L, LegendreColumn &leg) const;
  Mat kernel_W(const TwoEl &el, int L, LegendreColumn &leg) const;
  int converged_twoe_order(int iel) const;
  void compute_tei();
  Mat expand_boundaries(const Mat &P) const;
  Mat remove_boundaries(const Mat &F) const;
  void add_sub(Mat &M, int i, int j, const Mat &S, double f) const;
  Mat overlap() const;
  Mat kinetic() const;
  Mat nuclear() const;
  Mat coulomb(const Mat &P) const;
  Mat exchange(const Mat &P) const;
  Mat exchange_blocks(const Mat &P, const std::vector<int> &jk) const;
};

}  // namespace orc
