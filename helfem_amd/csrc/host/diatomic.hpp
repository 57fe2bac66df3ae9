// Host-side setup of the diatomic (prolate spheroidal) two-dimensional basis:
// radial FE in mu times Y_lm(nu, phi), src/diatomic/basis.{h,cpp}.
//
// Builds what the GPU Coulomb/exchange kernels consume (compute_tei,
// src/diatomic/basis.cpp:1382-1547): the disjoint-element P/Q integrals
// P0, P2, Q0, Q2 per ((L,|M|), element) and the sign-aware Cholesky factor
// W = B diag(sigma) B' of the two-channel in-element kernel, plus the
// one-electron matrices in the pure basis.  FP64, OpenMP over elements.
#pragma once
#include "fem.hpp"
#include <unordered_map>
#include <utility>

namespace hfg {

// P_l^m(x), Q_l^m(x) for l <= lmax, m <= mmax; out[m*(lmax+1) + l]
// (src/legendre/Legendre.h:335-372).
void legendre_p(double *P, int lmax, int mmax, double x);
void legendre_q(double *Q, int lmax, int mmax, double x);

// One |M| column of P_L^|M|(x), Q_L^|M|(x), L <= Lmax, memoised per abscissa;
// non-normal nonzero values flush to zero and Q is 0 at x = 1
// (MLegendreCache, src/diatomic/quadrature.h:47-84).
struct LegendreCache {
  int Mabs = 0, Lmax = 0;
  std::unordered_map<double, std::vector<double>> Ptab, Qtab;
  LegendreCache() {}
  LegendreCache(int Mabs, int Lmax);
  const std::vector<double> &column(bool q, double x);
  double P(int L, double x) { return column(false, x)[L]; }
  double Q(int L, double x) { return x == 1.0 ? 0.0 : column(true, x)[L]; }
};

// Gaunt couplings of the diatomic basis (src/general/gaunt.h:53-132,
// src/general/gaunt.cpp:275-300); zero outside the triangle conditions.
double gaunt3(int L, int M, int l, int m, int lp);
double gaunt_mod(int lj, int mj, int L, int M, int li, int mi);
double gaunt_cos(int lj, int mj, int li, int mi);
double gaunt_cos2(int lj, int mj, int li, int mi);

// Shell list (l, m) of lm_to_l_m (src/diatomic/basis.cpp:500-516):
// |m|-major, l ascending, +m before -m.
void diatomic_shells(const std::vector<int> &lmax_per_m, std::vector<int> &lval, std::vector<int> &mval);

struct DiatomicSetup {
  int Z1 = 0, Z2 = 0;
  double Rh = 0.0;  // half bond length
  int nquad = 0;
  RadialFE rad;
  std::vector<int> lval, mval;
  std::vector<std::pair<int, int>> lm_map;  // (L, |M|), |M|-major then L
  std::vector<std::pair<int, int>> LM_map;  // (L, M), L-major then M
  std::vector<int64_t> poff;                // pure index of (shell a, radial s_a)
  std::vector<Blk> P0, P2, Q0, Q2, cdB;     // index ilm*Nel + iel
  std::vector<std::vector<double>> cdS;     // signs sigma_p
  bool have_tei = false;

  void build(int Z1, int Z2, double Rh, int nnodes, int nquad, const std::vector<double> &bval,
             const std::vector<int> &lval, const std::vector<int> &mval);
  int Nrad() const { return rad.Nrad(); }
  int Nang() const { return (int)lval.size(); }
  int Nel() const { return rad.Nel(); }
  int64_t Ndummy() const { return (int64_t)Nrad() * Nang(); }
  int64_t Nbf() const { return poff.empty() ? 0 : poff.back(); }
  int skip(int a) const { return mval[a] != 0 ? 1 : 0; }  // radial fn 0 removed for m != 0
  int lmind(int L, int M) const;
  int LMind(int L, int M) const;
  double LMfac(int L, int M) const;
  void compute_tei();
  // the disjoint P/Q integrals alone (the first half of compute_tei) and the
  // sizing of the in-element factor caches
  void compute_pq();
  // Q_L^|M|(cosh mu_i) at the n outer Chebyshev nodes of element iel, the
  // values the in-element kernel's outer contraction weights with, as
  // [i * (Lmax + 1) + L] (the device setup takes them from here)
  std::vector<double> outer_legendre_q(int iel, int n, int Mabs, int Lmax) const;
  // dense pure-basis one-electron matrices, column-major Nbf x Nbf
  void overlap(double *out, int64_t ld) const;
  void kinetic(double *out, int64_t ld) const;
  void nuclear(double *out, int64_t ld) const;
  // the same as their nonzero shell-pair blocks: which 0 overlap, 1 kinetic, 2 nuclear
  void one_electron_blocks(int which, BlockList &out) const;

  Blk radial_integral(int m, int n) const;  // int B_i B_j sinh^m cosh^n
  Blk radial_kinetic() const;               // int B_i' B_j' sinh
  Blk pq_integral(bool q, int k, int iel, int L, LegendreCache &leg) const;
  int twoe_order(int iel) const;
  std::vector<int> twoe_orders() const;  // twoe_order of every element (OpenMP)
};

}  // namespace hfg
