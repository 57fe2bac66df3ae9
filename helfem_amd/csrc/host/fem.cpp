// Host-side FE setup: quadrature rules, LIP shape functions, auto-converged
// element matrices, the in-element two-electron tensor and its pivoted
// Cholesky factor.  Reference behaviour cited per function.
#include "fem.hpp"
#ifdef _OPENMP
#include <omp.h>
#endif

namespace hfg {

Blk blk_tn(const Blk &A, const Blk &B) {
  Blk C(A.c, B.c);
  for (int j = 0; j < B.c; ++j)
    for (int i = 0; i < A.c; ++i) {
      const double *a = &A.v[(size_t)i * A.r];
      const double *b = &B.v[(size_t)j * B.r];
      double s = 0.0;
      for (int k = 0; k < A.r; ++k) s += a[k] * b[k];
      C(i, j) = s;
    }
  return C;
}

namespace {
double blk_maxabs(const Blk &A) {
  double m = 0.0;
  for (double x : A.v) m = std::max(m, std::abs(x));
  return m;
}
double blk_maxdiff(const Blk &A, const Blk &B) {
  double m = 0.0;
  for (size_t i = 0; i < A.v.size(); ++i) m = std::max(m, std::abs(A.v[i] - B.v[i]));
  return m;
}

}  // namespace

// Order refinement shared by the one- and two-electron element integrals:
// double the rule until the block changes by <= 8 eps, or until the change
// stalls at the roundoff floor (FiniteElementBasis.cpp:500-553,
// RadialBasis.cpp:114-162).
Blk refine(const std::function<Blk(int)> &eval, int nstart, int nmax) {
  const double eps = std::numeric_limits<double>::epsilon();
  const double tol = 8.0 * eps, sq = std::sqrt(eps);
  Blk prev, cur;
  bool have = false;
  double d1 = -1.0, d2 = -1.0;
  int n = std::max(nstart, 2);
  for (;;) {
    cur = eval(n);
    if (have) {
      const double diff = blk_maxdiff(cur, prev), scale = blk_maxabs(cur);
      if (diff <= tol * (scale + tol)) return cur;
      if (d1 >= 0.0 && diff <= sq * (scale + tol) && diff > 0.5 * d1) return cur;
      if (d2 >= 0.0 && diff <= sq * (scale + tol) && diff > 0.125 * d2) return cur;
      d2 = d1;
      d1 = diff;
    }
    prev = cur;
    have = true;
    if (n >= nmax) return cur;
    n = std::min(2 * n, nmax);
  }
}


void gauss_lobatto(int n, std::vector<double> &x, std::vector<double> &w) {
  if (n < 2) throw std::runtime_error("Gauss-Lobatto rule needs n >= 2");
  x.resize(n);
  w.resize(n);
  const double pi = std::acos(-1.0);
  const double tolerance = 100.0 * std::numeric_limits<double>::epsilon();
  for (int i = 0; i < n; ++i) x[i] = std::cos(pi * double(i) / double(n - 1));
  // Legendre table p[j][i] = P_j(x_i); Newton on P'_{n-1} via
  // (x P_{n-1} - P_{n-2}) / (n P_{n-1}).
  std::vector<double> xo(n), pj((size_t)n * n);
  for (;;) {
    xo = x;
    for (int i = 0; i < n; ++i) {
      pj[i] = 1.0;
      pj[(size_t)n + i] = x[i];
    }
    for (int j = 2; j < n; ++j) {
      const double a = double(2 * j - 1), b = double(j - 1), c = 1.0 / double(j);
      for (int i = 0; i < n; ++i)
        pj[(size_t)j * n + i] = (a * x[i] * pj[(size_t)(j - 1) * n + i] - b * pj[(size_t)(j - 2) * n + i]) * c;
    }
    double err = 0.0;
    for (int i = 0; i < n; ++i) {
      const double pn1 = pj[(size_t)(n - 1) * n + i], pn2 = pj[(size_t)(n - 2) * n + i];
      x[i] = xo[i] - (x[i] * pn1 - pn2) / (double(n) * pn1);
      err = std::max(err, std::abs(x[i] - xo[i]));
    }
    if (err <= tolerance) break;
  }
  std::reverse(x.begin(), x.end());
  for (int i = 0; i < n; ++i) {
    double p = x[i], pm = 1.0;
    for (int j = 2; j < n; ++j) {
      const double pmm = pm;
      pm = p;
      p = (double(2 * j - 1) * x[i] * pm - double(j - 1) * pmm) / double(j);
    }
    w[i] = 2.0 / (double(n - 1) * double(n) * p * p);
  }
}

// Gauss-Legendre rule on [-1, 1] (exact for polynomials of degree 2n - 1):
// Newton on P_n from the Chebyshev guesses cos(pi (i - 1/4) / (n + 1/2)),
// weights 2 / ((1 - x^2) P_n'(x)^2); nodes increasing.
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
  if (n < 1) throw std::runtime_error("Gauss-Legendre rule needs n >= 1");
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const double pi = std::acos(-1.0);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(pi * (double(i) + 0.75) / (double(n) + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1.0, p0 = 0.0;
      for (int j = 1; j <= n; ++j) {
        const double p2 = p0;
        p0 = p1;
        p1 = ((2.0 * j - 1.0) * z * p0 - (j - 1.0) * p2) / double(j);
      }
      dp = double(n) * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::abs(dz) <= 4.0 * std::numeric_limits<double>::epsilon()) break;
    }
    {  // derivative at the converged node
      double p1 = 1.0, p0 = 0.0;
      for (int j = 1; j <= n; ++j) {
        const double p2 = p0;
        p0 = p1;
        p1 = ((2.0 * j - 1.0) * z * p0 - (j - 1.0) * p2) / double(j);
      }
      dp = double(n) * (z * p1 - p0) / (z * z - 1.0);
    }
    x[n - 1 - i] = z;
    w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

void gauss_chebyshev2(int n, std::vector<double> &x, std::vector<double> &w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const double pi = std::acos(-1.0);
  const double inv = 1.0 / double(n + 1);
  for (int i = 1; i <= n; ++i) {
    const double t = double(i) * pi * inv;
    const double s = std::sin(t), s2 = s * s, c = std::cos(t);
    // stored reversed so the nodes increase
    w[n - i] = 16.0 / 3.0 / double(n + 1) * s2 * s2;
    x[n - i] = 1.0 - 2.0 * double(i) * inv + (2.0 / pi) * (1.0 + 2.0 / 3.0 * s2) * c * s;
  }
}

std::vector<double> element_grid(double rmax, int num_el, int igrid, double zexp) {
  std::vector<double> b(num_el + 1);
  auto lin = [num_el](double hi, int i) { return i == num_el ? hi : double(i) * (hi / double(num_el)); };
  for (int i = 0; i <= num_el; ++i) {
    switch (igrid) {
      case 1: b[i] = lin(rmax, i); break;
      case 2: b[i] = double(i) * double(i) * rmax * (1.0 / (double(num_el) * double(num_el))); break;
      case 3: b[i] = rmax * std::pow(double(i) * (1.0 / double(num_el)), zexp); break;
      case 4: b[i] = std::exp(std::pow(lin(std::pow(std::log(rmax + 1.0), 1.0 / zexp), i), zexp)) - 1.0; break;
      default: throw std::logic_error("unsupported element grid type");
    }
  }
  b[0] = 0.0;
  b[num_el] = rmax;
  return b;
}

// d^k/dx^k of the Lagrange polynomials through `nodes` (k <= 2), evaluated
// as the explicit node products of LIPBasis_eval.h:29-.
Blk lagrange_prim(const std::vector<double> &x, const std::vector<double> &nodes, int deriv) {
  const int N = (int)nodes.size(), np = (int)x.size();
  Blk out(np, N);
  for (int ix = 0; ix < np; ++ix) {
    const double xv = x[ix];
    for (int f = 0; f < N; ++f) {
      const double xf = nodes[f];
      auto prod_skip = [&](int s1, int s2) {
        double v = 1.0;
        for (int p = 0; p < N; ++p) {
          if (p == f || p == s1 || p == s2) continue;
          v *= (xv - nodes[p]) / (xf - nodes[p]);
        }
        return v;
      };
      double val;
      if (deriv == 0) {
        val = prod_skip(-1, -1);
      } else if (deriv == 1) {
        double s = 0.0;
        for (int d = 0; d < N; ++d) {
          if (d == f) continue;
          double v = prod_skip(d, -1);
          v /= (xf - nodes[d]);
          s += v;
        }
        val = s;
      } else if (deriv == 2) {
        double s = 0.0;
        for (int d = 0; d < N; ++d) {
          if (d == f) continue;
          for (int e = 0; e < d; ++e) {
            if (e == f) continue;
            double v = prod_skip(d, e);
            v /= (xf - nodes[d]) * (xf - nodes[e]);
            s += v;
          }
        }
        val = 2.0 * s;
      } else {
        throw std::logic_error("LIP derivative order > 2");
      }
      out(ix, f) = val;
    }
  }
  return out;
}

Blk LipShape::eval(const std::vector<double> &x, int deriv, double h) const {
  const Blk prim = lagrange_prim(x, nodes, deriv);
  const double s = std::pow(h, deriv);
  Blk out((int)x.size(), nbf());
  for (int k = 0; k < nbf(); ++k)
    for (int i = 0; i < (int)x.size(); ++i) out(i, k) = prim(i, lo + k) / s;
  return out;
}

Blk LipShape::eval_over_r(const std::vector<double> &x, double h, int deriv) const {
  if (lo < 1) throw std::logic_error("B/r needs the r=0 function dropped");
  // B_k(x) = (x + 1)/(x_k + 1) L_k(x) with L_k the Lagrange polynomial on the
  // nodes without x = -1, and r = h (x + 1) on an element that starts at 0
  const std::vector<double> red(nodes.begin() + 1, nodes.end());
  const Blk p = lagrange_prim(x, red, deriv);
  const double scale = 1.0 / std::pow(h, deriv + 1);
  Blk out((int)x.size(), nbf());
  for (int k = 0; k < nbf(); ++k) {
    const int full = lo + k;
    const double den = nodes[full] + 1.0;
    for (int i = 0; i < (int)x.size(); ++i) out(i, k) = scale * p(i, full - 1) / den;
  }
  return out;
}

RadialFE::RadialFE(int nn, const std::vector<double> &bv, int nq, bool dl) : nnodes(nn), nquad(nq), drop_left(dl), bval(bv) {
  std::vector<double> w;
  gauss_lobatto(nnodes, nodes, w);
  std::sort(nodes.begin(), nodes.end());
  const int nel = Nel();
  if (nel < 1) throw std::logic_error("need at least one radial element");
  first.assign(nel, 0);
  ni.assign(nel, 0);
  for (int e = 0; e < nel; ++e) {
    ni[e] = shape(e).nbf();
    first[e] = e == 0 ? 0 : first[e - 1] + ni[e - 1] - 1;
  }
}

LipShape RadialFE::shape(int iel) const {
  LipShape s;
  s.nodes = nodes;
  s.lo = (iel == 0 && drop_left) ? 1 : 0;
  s.hi = iel == Nel() - 1 ? nnodes - 1 : nnodes;
  return s;
}

Blk RadialFE::rvalues(const std::vector<double> &x, int iel) const {
  const LipShape s = shape(iel);
  if (iel == 0) return s.eval_over_r(x, half_len(0));
  Blk v = s.eval(x, 0, half_len(iel));
  const double m = mid(iel), h = half_len(iel);
  for (int f = 0; f < v.c; ++f)
    for (int i = 0; i < v.r; ++i) v(i, f) /= (m + h * x[i]);
  return v;
}

Blk RadialFE::rderiv(const std::vector<double> &x, int iel) const {
  const LipShape s = shape(iel);
  const double h = half_len(iel), m = mid(iel);
  if (iel == 0 && drop_left) return s.eval_over_r(x, h, 1);
  const Blk f = s.eval(x, 0, h), d = s.eval(x, 1, h);
  Blk v(f.r, f.c);
  for (int c = 0; c < f.c; ++c)
    for (int i = 0; i < f.r; ++i) {
      const double ir = 1.0 / (m + h * x[i]);
      v(i, c) = (d(i, c) - f(i, c) * ir) * ir;
    }
  return v;
}

Blk RadialFE::values(int kind, const std::vector<double> &x, int iel) const {
  if (kind == 10) return rvalues(x, iel);
  return shape(iel).eval(x, kind, half_len(iel));
}

// Fixed-rule element matrix, FiniteElementBasis.cpp:460-491 on [-1,1]; a
// non-finite weight at a node counts as zero (the basis vanishes there).
Blk RadialFE::matel_rule(int iel, int lk, int rk, const std::vector<double> &x, const std::vector<double> &w,
                         const std::function<double(double)> *f) const {
  const double h = half_len(iel), m = mid(iel);
  const int n = (int)x.size();
  Blk lh = values(lk, x, iel);
  const Blk rh = values(rk, x, iel);
  for (int i = 0; i < n; ++i) {
    double wp = w[i] * h * 1.0;
    if (f) {
      const double fv = (*f)(m + h * x[i]);
      wp = std::isfinite(fv) ? wp * fv : 0.0;
    }
    for (int j = 0; j < lh.c; ++j) lh(i, j) *= wp;
  }
  return blk_tn(lh, rh);
}

// FiniteElementBasis.cpp:556-604 (one panel, start order for a degree
// nnodes-1 basis, Lobatto doubling to 512).
Blk RadialFE::matel(int iel, int lk, int rk, const std::function<double(double)> *f) const {
  const int nstart = std::min(std::max((2 * (nnodes - 1) + 4) / 2 + 2, 5), 512);
  return refine(
      [&](int n) {
        std::vector<double> x, w;
        gauss_lobatto(n, x, w);
        return matel_rule(iel, lk, rk, x, w, f);
      },
      nstart, 512);
}

Blk RadialFE::rpow(int n, int iel) const {
  const std::function<double(double)> f = [n](double r) { return std::pow(r, double(n + 2)); };
  return matel(iel, 10, 10, &f);
}

Blk RadialFE::kin(int iel) const {
  Blk b = matel(iel, 1, 1, nullptr);
  for (auto &x : b.v) x *= 0.5;
  return b;
}

Blk RadialFE::kin_l(int iel) const {
  Blk b = matel(iel, 10, 10, nullptr);
  for (auto &x : b.v) x *= 0.5;
  return b;
}

// quadrature::twoe_integral (libhelfem/src/quadrature.cpp:36-161): the
// cumulative inner integral phi_kl(r_p) = r_p^(-L-1) int_rmin^{r_p} B_k B_l r'^L
// on the same n-point rule mapped to each [r_{p-1}, r_p], then the outer
// contraction with B_i B_j, symmetrised.
Blk twoe_fixed_rule(double rmin, double rmax, const std::vector<double> &x, const std::vector<double> &w,
                    const LipShape &shape, int L) {
  const std::function<double(double, double)> fsb = [L](double r, double R) {
    return std::pow(r / R, double(L)) / R;
  };
  const std::function<double(double)> fbig = [L](double r) { return std::pow(r, double(-L - 1)); };
  return twoe_fixed_rule_k(rmin, rmax, x, w, shape, fsb, fbig);
}

// The same rule for a kernel f(r<, r>) = fsb(r<, r>) with the r>-dependence
// fbig(r>) used to carry the cumulative inner integral between nodes
// (r^L / r^(L+1) for Coulomb, i_L(lambda r) k_L(lambda R) for Yukawa,
// quadrature.cpp:164-198).
Blk twoe_fixed_rule_k(double rmin, double rmax, const std::vector<double> &x, const std::vector<double> &w,
                      const LipShape &shape, const std::function<double(double, double)> &fsb,
                      const std::function<double(double)> &fbig) {
  const int n = (int)x.size(), nb = shape.nbf(), nb2 = nb * nb;
  const double rmid = 0.5 * (rmax + rmin), rlen = 0.5 * (rmax - rmin);
  std::vector<double> r(n);
  for (int i = 0; i < n; ++i) r[i] = rmid + rlen * x[i];
  Blk inner(n, nb2);
  auto segment = [&](double a, double b, int row) {
    const double sm = 0.5 * (b + a), sl = 0.5 * (b - a);
    std::vector<double> rr(n), wp(n), xp(n);
    for (int i = 0; i < n; ++i) {
      rr[i] = sm + sl * x[i];
      wp[i] = w[i] * fsb(rr[i], b) * sl;
      xp[i] = (rr[i] - rmid) / rlen;
    }
    const Blk bf = shape.eval(xp, 0, rlen);
    Blk wbf = bf;
    for (int k = 0; k < nb; ++k)
      for (int i = 0; i < n; ++i) wbf(i, k) *= wp[i];
    const Blk M = blk_tn(wbf, bf);
    for (int k = 0; k < nb2; ++k) inner(row, k) = M.v[k];
  };
  const bool empty_first = (r[0] == rmin);
  if (!empty_first) segment(rmin, r[0], 0);
  for (int p = 1; p < n; ++p) segment(r[p - 1], r[p], p);
  for (int p = 1; p < n; ++p) {
    if (p == 1 && empty_first) continue;
    const double fac = fbig(r[p]) / fbig(r[p - 1]);
    for (int k = 0; k < nb2; ++k) inner(p, k) += inner(p - 1, k) * fac;
  }
  const Blk bf = shape.eval(x, 0, rlen);
  Blk prod(n, nb2);
  for (int a = 0; a < nb; ++a)
    for (int b = 0; b < nb; ++b)
      for (int i = 0; i < n; ++i) prod(i, a * nb + b) = bf(i, a) * bf(i, b) * (w[i] * rlen);
  const Blk ints = blk_tn(prod, inner);
  Blk out(nb2, nb2);
  for (int j = 0; j < nb2; ++j)
    for (int i = 0; i < nb2; ++i) out(i, j) = ints(i, j) + ints(j, i);
  return out;
}

// RadialBasis.cpp:644-663: Lobatto order refinement seeded at nquad.
Blk RadialFE::twoe(int L, int iel) const {
  const LipShape s = shape(iel);
  const double a = bval[iel], b = bval[iel + 1];
  return refine(
      [&](int n) {
        std::vector<double> x, w;
        gauss_lobatto(n, x, w);
        return twoe_fixed_rule(a, b, x, w, s, L);
      },
      std::min(std::max(nquad, 5), 512), 512);
}

// Greedy largest-diagonal pivoted Cholesky, absolute tolerance on the
// residual diagonal (RadialBasis.cpp:671-709).  Column-major (n x r).
Blk pivoted_cholesky(const Blk &A, double tol) {
  const int n = A.r;
  std::vector<double> D(n);
  for (int i = 0; i < n; ++i) D[i] = A(i, i);
  std::vector<char> used(n, 0);
  std::vector<double> cols;  // packed columns
  int rank = 0;
  for (int k = 0; k < n; ++k) {
    int piv = n;
    double pv = tol;
    for (int i = 0; i < n; ++i)
      if (!used[i] && D[i] > pv) {
        piv = i;
        pv = D[i];
      }
    if (piv == n) break;
    used[piv] = 1;
    const double sq = std::sqrt(pv);
    cols.resize((size_t)(rank + 1) * n);
    double *col = &cols[(size_t)rank * n];
    for (int i = 0; i < n; ++i) {
      if (used[i] && i != piv) {
        col[i] = 0.0;
        continue;
      }
      double s = A(i, piv);
      for (int j = 0; j < rank; ++j) s -= cols[(size_t)j * n + i] * cols[(size_t)j * n + piv];
      col[i] = s / sq;
    }
    col[piv] = sq;
    for (int i = 0; i < n; ++i)
      if (!used[i]) D[i] -= col[i] * col[i];
    ++rank;
  }
  Blk L(n, rank);
  L.v.assign(cols.begin(), cols.begin() + (size_t)rank * n);
  return L;
}

void AtomicSetup::build(int Z_, int nnodes, int nquad, const std::vector<double> &bval,
                        const std::vector<int> &lv, const std::vector<int> &mv) {
  if (lv.size() != mv.size() || lv.empty()) throw std::logic_error("angular basis: lval/mval mismatch");
  for (size_t i = 0; i < lv.size(); ++i)
    if (lv[i] < 0 || std::abs(mv[i]) > lv[i]) throw std::logic_error("angular basis: invalid (l,m)");
  Z = Z_;
  rad = RadialFE(nnodes, bval, nquad);
  lval = lv;
  mval = mv;
  N_L = 2 * (*std::max_element(lval.begin(), lval.end())) + 1;
  have_tei = false;
}

void AtomicSetup::compute_tei() {
  const int nel = rad.Nel();
  rsmall.assign((size_t)N_L * nel, Blk());
  rbig.assign((size_t)N_L * nel, Blk());
  chol.assign((size_t)N_L * nel, Blk());
  // The r^-L-1 factor on the innermost element diverges and is never used
  // (CoulombExchangeFE.h:193-213): it stays empty.
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int L = 0; L < N_L; ++L)
    for (int e = 0; e < nel; ++e) {
      const size_t i = (size_t)L * nel + e;
      rsmall[i] = rad.rpow(L, e);
      if (e > 0) rbig[i] = rad.rpow(-L - 1, e);
      chol[i] = pivoted_cholesky(rad.twoe(L, e), chol_tol);
    }
  have_tei = true;
}

void AtomicSetup::compute_radial_factors() {
  const int nel = rad.Nel();
  rsmall.assign((size_t)N_L * nel, Blk());
  rbig.assign((size_t)N_L * nel, Blk());
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int L = 0; L < N_L; ++L)
    for (int e = 0; e < nel; ++e) {
      const size_t i = (size_t)L * nel + e;
      rsmall[i] = rad.rpow(L, e);
      if (e > 0) rbig[i] = rad.rpow(-L - 1, e);
    }
}

namespace {
void radial_diag(const RadialFE &rad, const std::function<Blk(int)> &k, std::vector<double> &M) {
  const int n = rad.Nrad();
  M.assign((size_t)n * n, 0.0);
  for (int e = 0; e < rad.Nel(); ++e) {
    const Blk b = k(e);
    const int f = rad.first[e];
    for (int j = 0; j < b.c; ++j)
      for (int i = 0; i < b.r; ++i) M[(size_t)(f + i) + (size_t)(f + j) * n] += b(i, j);
  }
}
}  // namespace

void BlockList::to_dense(double *out, int64_t ld, int64_t n) const {
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < n; ++i) out[i + j * ld] = 0.0;
#pragma omp parallel for schedule(dynamic)
  for (size_t b = 0; b < count(); ++b) {
    const long long *m = &meta[5 * b];
    const double *v = &vals[(size_t)m[4]];
    for (long long c = 0; c < m[3]; ++c)
      for (long long r = 0; r < m[2]; ++r) out[(m[0] + r) + (m[1] + c) * ld] = v[r + c * m[2]];
  }
}

// TwoDBasis.cpp:321-392: every one-electron matrix is block diagonal over the
// angular shells; overlap S, kinetic T + l(l+1) T_l, nuclear -Z <1/r> (point
// nucleus at the origin) on each diagonal block.
void AtomicSetup::radial_one_electron(int which, std::vector<double> &M, std::vector<double> &X) const {
  if (which < 0 || which > 2) throw std::invalid_argument("which must be 0, 1 or 2");
  X.clear();
  if (which == 0)
    radial_diag(rad, [&](int e) { return rad.rpow(0, e); }, M);
  else if (which == 1) {
    radial_diag(rad, [&](int e) { return rad.kin(e); }, M);
    radial_diag(rad, [&](int e) { return rad.kin_l(e); }, X);
  } else
    radial_diag(rad, [&](int e) { return rad.rpow(-1, e); }, M);
}

void AtomicSetup::one_electron_blocks(int which, BlockList &out) const {
  std::vector<double> M, X;
  radial_one_electron(which, M, X);
  const int n = Nrad();
  std::vector<double> blk((size_t)n * n);
  for (int a = 0; a < Nang(); ++a) {
    const double f = which == 2 ? -double(Z) : 1.0;
    const double g = which == 1 ? double(lval[a] * (lval[a] + 1)) : 0.0;
    for (size_t q = 0; q < blk.size(); ++q) {
      double v = f * M[q];
      if (g != 0.0) v += g * X[q];
      blk[q] = v;
    }
    out.add((long long)a * n, (long long)a * n, n, n, blk.data());
  }
}

void AtomicSetup::overlap(double *out, int64_t ld) const {
  BlockList b;
  one_electron_blocks(0, b);
  b.to_dense(out, ld, Nbf());
}

void AtomicSetup::kinetic(double *out, int64_t ld) const {
  BlockList b;
  one_electron_blocks(1, b);
  b.to_dense(out, ld, Nbf());
}

void AtomicSetup::nuclear(double *out, int64_t ld) const {
  BlockList b;
  one_electron_blocks(2, b);
  b.to_dense(out, ld, Nbf());
}

}  // namespace hfg
