// Oracle: a plain CPU restatement of the HelFEM Fock-build path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library (helfem_amd/)
// may include, link or call anything in this directory.  The oracle is the
// checker the GPU path is compared against, and the CPU baseline that
// bench.py times beside it.  It follows the reference's loop structure and
// arithmetic order; each function cites the reference file:line it restates.
//
// Parity pinning (what checks each part):
//   lobatto/chebyshev/grid ....... closed forms + polynomial exactness (tests/test_oracle_fem.py)
//   LIP basis .................... Lagrange property, partition of unity, derivative vs finite difference
//   radial_integral .............. exact polynomial moments; H-atom 1e energy -0.5
//   twoe_integral ................ brute-force 2D quadrature split at r=r' (mpmath-free, numpy);
//                                  symmetry + PSD; H atom Coulomb 0.3125005143 (refs/ci.json)
//   pivoted Cholesky ............. reconstruction |LL'-T| <= tol, rank 27-29
//   gaunt ........................ 550 golden doubles of src/general/gaunt_test.cpp + sympy exact
//   atomic coulomb/exchange ...... He/Be/Ne HF energies of tests/refs/ci.json through oracle/scf.py;
//                                  Exx = -Coulomb/2 (He closed shell), J symmetric, K symmetric
//   PBE (xcfunc.cpp) + GGA grid .. uniform-gas and gradient-expansion limits, Lieb-Oxford saturation,
//                                  finite-difference potentials and dExc/dP, He PBE energies
//                                  (restricted and unrestricted) of tests/refs/ci.json, Ne (diatomic)
//   TPSS (xcfunc.cpp) + tau ...... uniform-gas limits, vanishing one-electron correlation,
//                                  finite-difference potentials and dExc/dP, He and Ne TPSS energies
#pragma once
#include <vector>
#include <cmath>
#include <cstdint>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <algorithm>
#include <functional>
#include <limits>
#include <unordered_map>

namespace orc {

// Column-major dense matrix (the reference's helfem::Mat<double> is an
// Eigen column-major matrix, libhelfem/include/Matrix.h:61).
struct Mat {
  int r = 0, c = 0;
  std::vector<double> a;
  Mat() {}
  Mat(int r_, int c_) : r(r_), c(c_), a((size_t)r_ * c_, 0.0) {}
  double &operator()(int i, int j) { return a[(size_t)i + (size_t)j * r]; }
  double operator()(int i, int j) const { return a[(size_t)i + (size_t)j * r]; }
  double *data() { return a.data(); }
  const double *data() const { return a.data(); }
  size_t size() const { return a.size(); }
};

Mat matmul(const Mat &A, const Mat &B);           // A*B
Mat matmul_tn(const Mat &A, const Mat &B);        // A'*B
Mat matmul_nt(const Mat &A, const Mat &B);        // A*B'
double max_abs(const Mat &A);
double max_abs_diff(const Mat &A, const Mat &B);

// ---- quadrature rules -----------------------------------------------------
// Gauss-Lobatto: libhelfem/include/lobatto.h:35-108
void lobatto(int n, std::vector<double> &x, std::vector<double> &w);
// Modified Gauss-Chebyshev of the 2nd kind: libhelfem/include/chebyshev.h:31-55
void chebyshev(int n, std::vector<double> &x, std::vector<double> &w);
// Element boundary grids: libhelfem/include/grid.h:37-101
std::vector<double> get_grid(double rmax, int num_el, int igrid, double zexp);

// ---- LIP polynomial basis (primbas=4): libhelfem/include/LIPBasis.h:27-113
struct LIPBasis {
  std::vector<double> x0;      // control nodes (Gauss-Lobatto)
  std::vector<int> enabled;    // enabled primitive indices
  LIPBasis() {}
  explicit LIPBasis(int nnodes);
  int nbf() const { return (int)enabled.size(); }
  int nprim() const { return (int)x0.size(); }
  void drop_first() { enabled.erase(enabled.begin()); }
  void drop_last() { enabled.pop_back(); }
  // d^n/dx^n of primitives at points x; (npts x nprim) column-major.
  // libhelfem/include/LIPBasis_eval.h:29-  (n = 0,1,2)
  static Mat eval_prim(const std::vector<double> &x, const std::vector<double> &nodes, int n);
  // PolynomialBasisT::eval_dnf: prim(:, enabled) / element_length^n (PolynomialBasisT.h:88-94)
  Mat eval_dnf(const std::vector<double> &x, int n, double element_length) const;
  // Analytic B/r on the first element (LIPBasis.h:83-111)
  Mat eval_over_r(const std::vector<double> &x, int n, double element_length) const;
};

// ---- finite-element radial basis -----------------------------------------
// FiniteElementBasisT (libhelfem/src/FiniteElementBasis.cpp) + FEMRadialBasisT
// (libhelfem/src/RadialBasis.cpp) for zero_func_left = zero_func_right = true.
struct RadialBasis {
  LIPBasis poly;
  std::vector<double> bval;
  std::vector<int> first, last;  // FiniteElementBasis.cpp:51-64
  std::vector<double> xq, wq;    // Chebyshev rule, RadialBasis.cpp:179-188
  bool dirichlet_left = true;     // drop the r=0 (mu=0) function
  RadialBasis() {}
  RadialBasis(int nnodes, const std::vector<double> &bval, int nquad, bool dirichlet_left = true);
  // fixed-rule element matrix (FiniteElementBasis.cpp:460-491 on [-1,1])
  Mat matel_rule(int iel, int lkind, int rkind, const std::vector<double> &x, const std::vector<double> &w,
                 const std::function<double(double)> *f) const;
  int Nel() const { return (int)bval.size() - 1; }
  int Nbf() const { return last.back() + 1; }
  int Nprim(int iel) const { return last[iel] - first[iel] + 1; }
  LIPBasis get_basis(int iel) const;  // FiniteElementBasis.cpp:284-293
  double scaling_factor(int iel) const { return 0.5 * (bval[iel + 1] - bval[iel]); }
  double element_midpoint(int iel) const { return 0.5 * (bval[iel] + bval[iel + 1]); }
  std::vector<double> eval_coord(const std::vector<double> &x, int iel) const;
  Mat eval_dnf(const std::vector<double> &x, int n, int iel) const;
  Mat get_bf(const std::vector<double> &x, int iel) const;   // RadialBasis.cpp:881-890 (R = B/r)
  Mat get_df(const std::vector<double> &x, int iel) const;   // RadialBasis.cpp:890-905 (d/dr of B/r)
  // kind: 0 = B0, 1 = B1, 10 = R0
  Mat eval_kind(int kind, const std::vector<double> &x, int iel) const;
  // matrix_element_auto with no breakpoints (FiniteElementBasis.cpp:556-604)
  Mat matrix_element_auto(int iel, int lkind, int rkind, const std::function<double(double)> *f) const;
  Mat radial_integral(int n, int iel) const;   // RadialBasis.cpp:246-257
  Mat kinetic(int iel) const;                  // RadialBasis.cpp:485-487
  Mat kinetic_l(int iel) const;                // RadialBasis.cpp:489-491
  Mat twoe_integral(int L, int iel) const;     // RadialBasis.cpp:644-663
  Mat twoe_integral_cholesky(int L, int iel, double tol) const;  // RadialBasis.cpp:712-716
  // auto-converged in-element tensor for a kernel fsb(r<, R>), fbig(R)
  Mat twoe_integral_conv(int iel, const std::function<double(double, double)> &fsb,
                         const std::function<double(double)> &fbig) const;
  Mat bessel_il_integral(int L, double lambda, int iel) const;  // RadialBasis.cpp:321-324
  Mat bessel_kl_integral(int L, double lambda, int iel) const;  // RadialBasis.cpp:327-330
  Mat yukawa_integral(int L, double lambda, int iel) const;     // RadialBasis.cpp:727-741
  Mat erfc_integral(int L, double mu, int iel, int kel) const;  // RadialBasis.cpp:743-812
};

// Range-separated kernels (libhelfem/include/math.h, erfc_expn.h)
double bessel_il(double x, int L);
double bessel_kl(double x, int L);
void erfc_phi(int Lmax, double x, double y, double *out);  // Phi_0..Lmax(x, y)
Mat exchange_order(const Mat &T, int Ni, int Nk);

// quadrature::twoe_integral, libhelfem/src/quadrature.cpp:136-161
Mat twoe_integral_rule(double rmin, double rmax, const std::vector<double> &x,
                       const std::vector<double> &w, const LIPBasis &poly, int L);
// RadialBasis.cpp:671-709
Mat pivoted_cholesky(const Mat &A, double tol);

// ---- Gaunt coefficients ---------------------------------------------------
// Exact-rational evaluation of int Y_L^M* Y_l^m Y_lp^mp (gaunt.cpp:87-93).
double gaunt_coefficient(int L, int M, int l, int m, int lp, int mp);
double modified_gaunt_coefficient(int lj, int mj, int L, int M, int li, int mi);
// Table: coeff(L,M,l,m,lp) with mp = M - m (gaunt.h:53-132 semantics)
struct Gaunt {
  int Lmax = 0, lmax = 0, lpmax = 0;
  std::vector<double> tab;  // dense (L,M,l,m,lp)
  Gaunt() {}
  Gaunt(int Lmax, int lmax, int lpmax);
  double coeff(int L, int M, int l, int m, int lp) const;
  double mod_coeff(int lj, int mj, int L, int M, int li, int mi) const;
};

// ---- atomic two-dimensional basis (src/atomic/TwoDBasis.{h,cpp}) -----------
struct AtomicBasis {
  int Z = 0;
  RadialBasis radial;
  std::vector<int> lval, mval;
  std::vector<Mat> disjoint_L, disjoint_m1L, prim_chol;  // TwoDBasis.h:83-104
  double chol_tol = 1e-12;                               // TwoDBasis.h:118-119
  // range-separated caches: Yukawa (disjoint_iL/kL, rs_chol) or erfc (rs_ktei)
  bool yukawa = false;
  double lambda = 0.0;
  std::vector<Mat> disjoint_iL, disjoint_kL, rs_chol, rs_ktei;
  AtomicBasis() {}
  AtomicBasis(int Z, int nnodes, int nquad, const std::vector<double> &bval,
              const std::vector<int> &lval, const std::vector<int> &mval);
  int Nrad() const { return radial.Nbf(); }
  int Nang() const { return (int)lval.size(); }
  int Nbf() const { return Nrad() * Nang(); }
  int N_L() const { return 2 * (*std::max_element(lval.begin(), lval.end())) + 1; }
  void compute_tei();                 // TwoDBasis.cpp:709-735
  Mat overlap() const;                // TwoDBasis.cpp:321-333
  Mat kinetic() const;                // TwoDBasis.cpp:354-374
  Mat nuclear() const;                // TwoDBasis.cpp:377-392 (point nucleus, Zl=Zr=0)
  Mat coulomb(const Mat &P) const;    // TwoDBasis.cpp:774-877
  Mat exchange(const Mat &P) const;   // TwoDBasis.cpp:880-999 (returns -K)
  // Same, restricted to the listed output blocks j*Nang+k (CPU-baseline sample).
  Mat exchange_blocks(const Mat &P, const std::vector<int> &jk) const;
  void compute_yukawa(double lambda);  // TwoDBasis.cpp:738-760
  void compute_erfc(double mu);        // TwoDBasis.cpp:762-771
  Mat rs_exchange(const Mat &P) const; // TwoDBasis.cpp:1002-1131 (returns -K)
};
// angular_basis, src/atomic/basis.cpp:179-203
void angular_basis(int lmax, int mmax, std::vector<int> &lval, std::vector<int> &mval);

// ---- Legendre functions (src/legendre/Legendre.h) --------------------------
// out[m*(lmax+1) + l] = P_l^m(x) / Q_l^m(x), 0 <= m <= mmax
void legendre_plm(double *out, int lmax, int mmax, double x);
void legendre_qlm(double *out, int lmax, int mmax, double x);
// One |M| column cached per argument (src/diatomic/quadrature.h:47-84)
struct LegendreColumn {
  int Mabs, Lmax;
  std::unordered_map<double, std::vector<double>> Ptab, Qtab;
  LegendreColumn(int Mabs, int Lmax);
  const std::vector<double> &fetch(bool wantP, double x);
  double P(int L, double x);
  double Q(int L, double x);
};

Mat assemble_K_pairwise(const RadialBasis &radial, const std::vector<Mat> &ktei, size_t off, const Mat &P);
// FE assemblers, libhelfemqc/include/CoulombExchangeFE.h:432-530
Mat assemble_J_chol(const RadialBasis &radial, const std::vector<Mat> &rs, const std::vector<Mat> &rb,
                    const std::vector<Mat> &chol, int off, const Mat &P);
Mat assemble_K_chol(const RadialBasis &radial, const std::vector<Mat> &rs, const std::vector<Mat> &rb,
                    const std::vector<Mat> &chol, int off, const Mat &P);

}  // namespace orc
