// Host-side finite-element setup for the HelFEM Fock-build engine.
//
// This is the part of TwoDBasis that builds the per-(L, element) factor
// caches the GPU kernels consume (compute_tei, src/atomic/TwoDBasis.cpp:709-735),
// plus the one-electron matrices an SCF driver needs.  It runs once per
// basis in FP64 on the host with OpenMP; the per-iteration Fock build runs
// on the GPU.
#pragma once
#include <vector>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <algorithm>
#include <functional>
#include <limits>

namespace hfg {

// Column-major dense block, the layout of helfem::Mat<double>
// (libhelfem/include/Matrix.h:61).
struct Blk {
  int r = 0, c = 0;
  std::vector<double> v;
  Blk() {}
  Blk(int r_, int c_) : r(r_), c(c_), v((size_t)r_ * c_, 0.0) {}
  double &operator()(int i, int j) { return v[(size_t)i + (size_t)j * r]; }
  double operator()(int i, int j) const { return v[(size_t)i + (size_t)j * r]; }
};

Blk blk_tn(const Blk &A, const Blk &B);  // A' * B

// Gauss-Lobatto rule (libhelfem/include/lobatto.h:35-108)
void gauss_lobatto(int n, std::vector<double> &x, std::vector<double> &w);
// Modified Gauss-Chebyshev rule of the 2nd kind (libhelfem/include/chebyshev.h:31-55)
void gauss_chebyshev2(int n, std::vector<double> &x, std::vector<double> &w);
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w);
// Element boundaries (libhelfem/include/grid.h:37-101), igrid 1..4
std::vector<double> element_grid(double rmax, int num_el, int igrid, double zexp);

// Lagrange interpolating polynomials on Gauss-Lobatto nodes (primbas=4,
// libhelfem/include/LIPBasis.h).  `lo`/`hi` select the enabled primitive
// range [lo, hi) after the Dirichlet drops at r=0 and r=rmax.
struct LipShape {
  std::vector<double> nodes;
  int lo = 0, hi = 0;
  int nbf() const { return hi - lo; }
  // values (deriv 0/1/2) at x, scaled by 1/h^deriv; (npts x nbf)
  Blk eval(const std::vector<double> &x, int deriv, double h) const;
  // d^deriv/dr^deriv of B(x)/r on the first element, where B(-1) = 0
  // (LIPBasis.h:83-111)
  Blk eval_over_r(const std::vector<double> &x, double h, int deriv = 0) const;
};
Blk lagrange_prim(const std::vector<double> &x, const std::vector<double> &nodes, int deriv);

// Finite-element radial basis with zero value at both ends
// (FiniteElementBasisT + FEMRadialBasisT).
// Dense matrix given by its nonzero blocks: per block (row0, col0, rows,
// cols, offset into vals), values column-major.  to_dense zeroes the rest.
struct BlockList {
  std::vector<long long> meta;
  std::vector<double> vals;
  void add(long long row0, long long col0, int rows, int cols, const double *v) {
    meta.insert(meta.end(), {row0, col0, (long long)rows, (long long)cols, (long long)vals.size()});
    vals.insert(vals.end(), v, v + (size_t)rows * cols);
  }
  size_t count() const { return meta.size() / 5; }
  void to_dense(double *out, int64_t ld, int64_t n) const;
};

struct RadialFE {
  int nnodes = 0, nquad = 0;
  bool drop_left = true;       // zero value at the left end (r = 0)
  std::vector<double> nodes, bval;
  std::vector<int> first, ni;  // first global index and size of each element
  RadialFE() {}
  RadialFE(int nnodes, const std::vector<double> &bval, int nquad, bool drop_left = true);
  // fixed-rule element matrix <lkind| f |rkind> on [-1,1] (kinds 0, 1, 10)
  Blk matel_rule(int iel, int lk, int rk, const std::vector<double> &x, const std::vector<double> &w,
                 const std::function<double(double)> *f) const;
  int Nel() const { return (int)bval.size() - 1; }
  int Nrad() const { return first.back() + ni.back(); }
  LipShape shape(int iel) const;
  double half_len(int iel) const { return 0.5 * (bval[iel + 1] - bval[iel]); }
  double mid(int iel) const { return 0.5 * (bval[iel] + bval[iel + 1]); }
  // radial function values R = B/r at element coordinates x
  Blk rvalues(const std::vector<double> &x, int iel) const;
  // their r derivatives d/dr (B/r) (RadialBasis.cpp:890-905)
  Blk rderiv(const std::vector<double> &x, int iel) const;
  // kind 0: B, 1: dB/dr, 10: R = B/r
  Blk values(int kind, const std::vector<double> &x, int iel) const;
  // auto-converged <lkind| f |rkind> over element iel
  Blk matel(int iel, int lkind, int rkind, const std::function<double(double)> *f) const;
  Blk rpow(int n, int iel) const;  // <R| r^n |R> = int R R r^(n+2) dr
  Blk kin(int iel) const;          // 1/2 int B' B'
  Blk kin_l(int iel) const;        // 1/2 int R R
  Blk twoe(int L, int iel) const;  // in-element Slater tensor, (ni^2 x ni^2)
  // range-separated element integrals (RadialBasis.cpp:321-330, 719-812)
  Blk bessel_i(int L, double lambda, int iel) const;  // int B B i_L(lambda r)
  Blk bessel_k(int L, double lambda, int iel) const;  // int B B k_L(lambda r)
  Blk yukawa(int L, double lambda, int iel) const;    // in-element Yukawa tensor
  // erfc tensor of an element pair on the fixed Chebyshev rule, Coulomb
  // ordered: rows a*ni + c (element iel), cols b*nk + d (element kel)
  Blk erfc_pair(int L, double mu, int iel, int kel) const;
};

Blk twoe_fixed_rule(double rmin, double rmax, const std::vector<double> &x, const std::vector<double> &w,
                    const LipShape &shape, int L);
Blk twoe_fixed_rule_k(double rmin, double rmax, const std::vector<double> &x, const std::vector<double> &w,
                      const LipShape &shape, const std::function<double(double, double)> &fsb,
                      const std::function<double(double)> &fbig);
// Modified spherical Bessel functions i_L = sqrt(pi/2x) I_{L+1/2},
// k_L = sqrt(2/(pi x)) K_{L+1/2} (libhelfem/include/math.h:52-115), and the
// multipole kernel of erfc(mu d)/d = mu sum_L Phi_L(mu r, mu r') P_L(cos g).
double sph_bessel_i(double x, int L);
double sph_bessel_k(double x, int L);
void erfc_multipoles(int Lmax, double x, double y, double *out);
Blk pivoted_cholesky(const Blk &A, double tol);
// Lobatto order doubling until the block changes by <= 8 eps (or stalls at
// the roundoff floor), from nstart up to nmax points.
Blk refine(const std::function<Blk(int)> &eval, int nstart, int nmax);

// Gaunt coefficient int Y_L^M* Y_l^m Y_lp^mp, exact rational rounded once.
double gaunt_coeff(int L, int M, int l, int m, int lp, int mp);

// Atomic two-dimensional basis (src/atomic/TwoDBasis.h:52-283): radial FE x Y_lm.
struct AtomicSetup {
  int Z = 0;
  RadialFE rad;
  std::vector<int> lval, mval;
  int N_L = 0;
  std::vector<Blk> rsmall, rbig, chol;  // index L*Nel + iel
  double chol_tol = 1e-12;
  bool have_tei = false;
  // range separation (TwoDBasis.cpp:738-771): kind 1 Yukawa (rs_small/big =
  // i_L/k_L element integrals, rs_chol), kind 2 erfc (rs_pair, index
  // L*Nel*Nel + iel*Nel + kel, Coulomb-ordered element-pair tensors)
  int rs_kind = 0;
  double rs_param = 0.0;
  std::vector<Blk> rs_small, rs_big, rs_chol, rs_pair;
  void compute_yukawa(double lambda);
  void compute_erfc(double mu);
  void build(int Z, int nnodes, int nquad, const std::vector<double> &bval, const std::vector<int> &lval,
             const std::vector<int> &mval);
  int Nrad() const { return rad.Nrad(); }
  int Nang() const { return (int)lval.size(); }
  int64_t Nbf() const { return (int64_t)Nrad() * Nang(); }
  void compute_tei();
  // the r^L and r^-L-1 element matrices alone (the part of compute_tei that
  // stays on the host when the Slater tensors are built on the device)
  void compute_radial_factors();
  // dense one-electron matrices, column-major Nbf x Nbf with leading dim ld
  void overlap(double *out, int64_t ld) const;
  void kinetic(double *out, int64_t ld) const;
  void nuclear(double *out, int64_t ld) const;
  // the same as their nonzero blocks: which 0 overlap, 1 kinetic, 2 nuclear
  void one_electron_blocks(int which, BlockList &out) const;
  // the radial factors of those blocks: shell a's block is f M + l_a (l_a + 1) X
  // with f = 1 (overlap, kinetic) or -Z (nuclear); X only for the kinetic energy
  void radial_one_electron(int which, std::vector<double> &M, std::vector<double> &X) const;
};

}  // namespace hfg
