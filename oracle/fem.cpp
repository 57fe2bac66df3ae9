// Oracle (test infrastructure only): finite-element radial core.
// Restates libhelfem's Lobatto/Chebyshev rules, grids, LIP basis, the
// auto-converging one-electron matrix elements, the in-element radial
// two-electron tensor and its pivoted Cholesky factorisation.
#include "orc.hpp"
#include <cstdio>

namespace orc {

Mat matmul(const Mat &A, const Mat &B) {
  if (A.c != B.r) throw std::logic_error("matmul: shape mismatch");
  Mat C(A.r, B.c);
  for (int j = 0; j < B.c; ++j)
    for (int k = 0; k < A.c; ++k) {
      const double b = B(k, j);
      if (b == 0.0) continue;
      const double *a = &A.a[(size_t)k * A.r];
      double *c = &C.a[(size_t)j * C.r];
      for (int i = 0; i < A.r; ++i) c[i] += a[i] * b;
    }
  return C;
}

// Every entry is a dot product summed in index order; four columns of B are
// processed together only to let the loop vectorise (same sums, same order).
Mat matmul_tn(const Mat &A, const Mat &B) {
  if (A.r != B.r) throw std::logic_error("matmul_tn: shape mismatch");
  Mat C(A.c, B.c);
  const int K = A.r;
  int j = 0;
  for (; j + 4 <= B.c; j += 4)
    for (int i = 0; i < A.c; ++i) {
      const double *a = &A.a[(size_t)i * K];
      const double *b = &B.a[(size_t)j * K];
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      for (int k = 0; k < K; ++k) {
        s0 += a[k] * b[k];
        s1 += a[k] * b[k + K];
        s2 += a[k] * b[k + 2 * (size_t)K];
        s3 += a[k] * b[k + 3 * (size_t)K];
      }
      C(i, j) = s0;
      C(i, j + 1) = s1;
      C(i, j + 2) = s2;
      C(i, j + 3) = s3;
    }
  for (; j < B.c; ++j)
    for (int i = 0; i < A.c; ++i) {
      const double *a = &A.a[(size_t)i * K];
      const double *b = &B.a[(size_t)j * K];
      double s = 0.0;
      for (int k = 0; k < K; ++k) s += a[k] * b[k];
      C(i, j) = s;
    }
  return C;
}

Mat matmul_nt(const Mat &A, const Mat &B) {
  if (A.c != B.c) throw std::logic_error("matmul_nt: shape mismatch");
  Mat C(A.r, B.r);
  for (int k = 0; k < A.c; ++k)
    for (int j = 0; j < B.r; ++j) {
      const double b = B(j, k);
      const double *a = &A.a[(size_t)k * A.r];
      double *c = &C.a[(size_t)j * C.r];
      for (int i = 0; i < A.r; ++i) c[i] += a[i] * b;
    }
  return C;
}

double max_abs(const Mat &A) {
  double m = 0.0;
  for (double v : A.a) m = std::max(m, std::abs(v));
  return m;
}

double max_abs_diff(const Mat &A, const Mat &B) {
  double m = 0.0;
  for (size_t i = 0; i < A.a.size(); ++i) m = std::max(m, std::abs(A.a[i] - B.a[i]));
  return m;
}

// libhelfem/include/lobatto.h:35-108: Newton iteration for the roots of
// P'_{n-1}, started from the Chebyshev-Gauss-Lobatto points.
void lobatto(int n, std::vector<double> &x, std::vector<double> &w) {
  if (n < 2) throw std::runtime_error("Lobatto needs n >= 2");
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const double pi = std::acos(-1.0);
  const double tolerance = 100.0 * std::numeric_limits<double>::epsilon();
  for (int i = 0; i < n; ++i) x[i] = std::cos(pi * double(i) / double(n - 1));
  std::vector<double> xold(n), p((size_t)n * n);
  for (;;) {
    xold = x;
    for (int i = 0; i < n; ++i) {
      p[i] = 1.0;
      p[i + (size_t)n] = x[i];
    }
    for (int j = 2; j <= n - 1; ++j) {
      const double tj = double(2 * j - 1), jm1 = double(j - 1), invj = 1.0 / double(j);
      for (int i = 0; i < n; ++i)
        p[i + (size_t)j * n] = (tj * x[i] * p[i + (size_t)(j - 1) * n] - jm1 * p[i + (size_t)(j - 2) * n]) * invj;
    }
    for (int i = 0; i < n; ++i)
      x[i] = xold[i] - (x[i] * p[i + (size_t)(n - 1) * n] - p[i + (size_t)(n - 2) * n]) /
                           (double(n) * p[i + (size_t)(n - 1) * n]);
    double err = 0.0;
    for (int i = 0; i < n; ++i) err = std::max(err, std::abs(x[i] - xold[i]));
    if (err <= tolerance) break;
  }
  std::reverse(x.begin(), x.end());
  for (int i = 0; i < n; ++i) {
    double pn = x[i], pnm1 = 1.0, pnm2;
    for (int j = 2; j <= n - 1; ++j) {
      pnm2 = pnm1;
      pnm1 = pn;
      pn = (double(2 * j - 1) * x[i] * pnm1 - double(j - 1) * pnm2) / double(j);
    }
    w[i] = 2.0 / (double(n - 1) * double(n) * pn * pn);
  }
}

// libhelfem/include/chebyshev.h:31-55
void chebyshev(int n, std::vector<double> &x, std::vector<double> &w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const double pi = std::acos(-1.0);
  const double oonpp = 1.0 / double(n + 1);
  const double two_over_pi = 2.0 / pi;
  for (int i = 1; i <= n; ++i) {
    const double angle = double(i) * pi * oonpp;
    const double s = std::sin(angle), s2 = s * s, c = std::cos(angle);
    w[i - 1] = 16.0 / 3.0 / double(n + 1) * s2 * s2;
    x[i - 1] = 1.0 - 2.0 * double(i) * oonpp + two_over_pi * (1.0 + 2.0 / 3.0 * s2) * c * s;
  }
  std::reverse(x.begin(), x.end());
  std::reverse(w.begin(), w.end());
}

// libhelfem/include/grid.h:37-101 (igrid 1..4)
std::vector<double> get_grid(double rmax, int num_el, int igrid, double zexp) {
  std::vector<double> b(num_el + 1, 0.0);
  auto linspaced = [](int size, double lo, double hi, int i) {
    // Eigen LinSpaced semantics for |hi| >= |lo|
    if (i == size - 1) return hi;
    return lo + double(i) * ((hi - lo) / double(size - 1));
  };
  switch (igrid) {
    case 1:
      for (int i = 0; i <= num_el; ++i) b[i] = linspaced(num_el + 1, 0.0, rmax, i);
      break;
    case 2: {
      const double inv = 1.0 / (double(num_el) * double(num_el));
      for (int i = 0; i <= num_el; ++i) b[i] = double(i) * double(i) * rmax * inv;
      break;
    }
    case 3: {
      const double inv = 1.0 / double(num_el);
      for (int i = 0; i <= num_el; ++i) b[i] = rmax * std::pow(double(i) * inv, zexp);
      break;
    }
    case 4: {
      const double upper = std::pow(std::log(rmax + 1.0), 1.0 / zexp);
      for (int i = 0; i <= num_el; ++i)
        b[i] = std::exp(std::pow(linspaced(num_el + 1, 0.0, upper, i), zexp)) - 1.0;
      break;
    }
    default:
      throw std::logic_error("Invalid choice for grid");
  }
  b[0] = 0.0;
  b[num_el] = rmax;
  return b;
}

LIPBasis::LIPBasis(int nnodes) {
  std::vector<double> w;
  lobatto(nnodes, x0, w);
  std::sort(x0.begin(), x0.end());
  enabled.resize(nnodes);
  for (int i = 0; i < nnodes; ++i) enabled[i] = i;
}

// libhelfem/include/LIPBasis_eval.h:29- : explicit Lagrange products.
Mat LIPBasis::eval_prim(const std::vector<double> &x, const std::vector<double> &x0, int n) {
  const int N = (int)x0.size();
  Mat d((int)x.size(), N);
  for (size_t ix = 0; ix < x.size(); ++ix)
    for (int fi = 0; fi < N; ++fi) {
      double val = 0.0;
      if (n == 0) {
        double v = 1.0;
        for (int ip = 0; ip < N; ++ip) {
          if (ip == fi) continue;
          v *= (x[ix] - x0[ip]) / (x0[fi] - x0[ip]);
        }
        val = v;
      } else if (n == 1) {
        double el = 0.0;
        for (int d1 = 0; d1 < N; ++d1) {
          if (d1 == fi) continue;
          double v = 1.0;
          for (int ip = 0; ip < N; ++ip) {
            if (ip == d1 || ip == fi) continue;
            v *= (x[ix] - x0[ip]) / (x0[fi] - x0[ip]);
          }
          v /= (x0[fi] - x0[d1]);
          el += v;
        }
        val = el;
      } else if (n == 2) {
        double el = 0.0;
        for (int d1 = 0; d1 < N; ++d1) {
          if (d1 == fi) continue;
          for (int d2 = 0; d2 < d1; ++d2) {
            if (d2 == fi) continue;
            double v = 1.0;
            for (int ip = 0; ip < N; ++ip) {
              if (ip == d1 || ip == d2 || ip == fi) continue;
              v *= (x[ix] - x0[ip]) / (x0[fi] - x0[ip]);
            }
            v /= (x0[fi] - x0[d1]) * (x0[fi] - x0[d2]);
            el += v;
          }
        }
        val = 2.0 * el;
      } else {
        throw std::logic_error("LIP derivative order not implemented");
      }
      d((int)ix, fi) = val;
    }
  return d;
}

Mat LIPBasis::eval_dnf(const std::vector<double> &x, int n, double element_length) const {
  Mat prim = eval_prim(x, x0, n);
  Mat out((int)x.size(), nbf());
  const double s = std::pow(element_length, n);
  for (int k = 0; k < nbf(); ++k)
    for (int i = 0; i < (int)x.size(); ++i) out(i, k) = prim(i, enabled[k]) / s;
  return out;
}

Mat LIPBasis::eval_over_r(const std::vector<double> &x, int n, double element_length) const {
  if (enabled.empty() || enabled[0] == 0) throw std::logic_error("eval_over_r needs drop_first");
  std::vector<double> xr(x0.begin() + 1, x0.end());
  Mat red = eval_prim(x, xr, n);
  const double scale = 1.0 / std::pow(element_length, n + 1);
  Mat out((int)x.size(), nbf());
  for (int k = 0; k < nbf(); ++k) {
    const int ifull = enabled[k], ired = ifull - 1;
    const double denom = x0[ifull] + 1.0;
    for (int i = 0; i < (int)x.size(); ++i) out(i, k) = scale * red(i, ired) / denom;
  }
  return out;
}

RadialBasis::RadialBasis(int nnodes, const std::vector<double> &bv, int nquad, bool dl)
    : poly(nnodes), bval(bv), dirichlet_left(dl) {
  const int nel = Nel();
  first.assign(nel, 0);
  last.assign(nel, 0);
  for (int iel = 0; iel < nel; ++iel) {
    first[iel] = (iel == 0) ? 0 : last[iel - 1];  // noverlap = 1
    last[iel] = first[iel] + get_basis(iel).nbf() - 1;
  }
  chebyshev(nquad, xq, wq);
}

LIPBasis RadialBasis::get_basis(int iel) const {
  LIPBasis p = poly;
  if (iel == 0 && dirichlet_left) p.drop_first();
  if (iel == Nel() - 1) p.drop_last();
  return p;
}

std::vector<double> RadialBasis::eval_coord(const std::vector<double> &x, int iel) const {
  std::vector<double> r(x.size());
  const double mid = element_midpoint(iel), sf = scaling_factor(iel);
  for (size_t i = 0; i < x.size(); ++i) r[i] = mid + sf * x[i];
  return r;
}

Mat RadialBasis::eval_dnf(const std::vector<double> &x, int n, int iel) const {
  return get_basis(iel).eval_dnf(x, n, scaling_factor(iel));
}

Mat RadialBasis::get_bf(const std::vector<double> &x, int iel) const {
  if (iel == 0) return get_basis(0).eval_over_r(x, 0, scaling_factor(0));
  Mat v = eval_dnf(x, 0, iel);
  const std::vector<double> r = eval_coord(x, iel);
  for (int f = 0; f < v.c; ++f)
    for (int i = 0; i < v.r; ++i) v(i, f) /= r[i];
  return v;
}

// d/dr (B/r): the analytic deflated form on the first element, else
// (B' - B/r)/r (RadialBasis.cpp:890-905)
Mat RadialBasis::get_df(const std::vector<double> &x, int iel) const {
  if (iel == 0) return get_basis(0).eval_over_r(x, 1, scaling_factor(0));
  const Mat f = eval_dnf(x, 0, iel), d = eval_dnf(x, 1, iel);
  const std::vector<double> r = eval_coord(x, iel);
  Mat out(f.r, f.c);
  for (int j = 0; j < f.c; ++j)
    for (int i = 0; i < f.r; ++i) {
      const double ir = 1.0 / r[i];
      out(i, j) = (-f(i, j) * ir + d(i, j)) * ir;
    }
  return out;
}

Mat RadialBasis::eval_kind(int kind, const std::vector<double> &x, int iel) const {
  if (kind == 0) return eval_dnf(x, 0, iel);
  if (kind == 1) return eval_dnf(x, 1, iel);
  if (kind == 10) return get_bf(x, iel);
  throw std::logic_error("eval_kind");
}

namespace {
// FiniteElementBasis.cpp:460-491 on the full reference range [-1,1].
Mat matel_fixed(const RadialBasis &rb, int iel, int lk, int rk, const std::vector<double> &xq,
                const std::vector<double> &wq, const std::function<double(double)> *f) {
  const std::vector<double> r = rb.eval_coord(xq, iel);
  std::vector<double> wp(wq.size());
  const double sf = rb.scaling_factor(iel);
  for (size_t i = 0; i < wq.size(); ++i) {
    wp[i] = wq[i] * sf * 1.0;
    if (f) {
      const double fv = (*f)(r[i]);
      wp[i] = std::isfinite(fv) ? wp[i] * fv : 0.0;
    }
  }
  Mat lh = rb.eval_kind(lk, xq, iel);
  const Mat rh = rb.eval_kind(rk, xq, iel);
  for (int j = 0; j < lh.c; ++j)
    for (int i = 0; i < lh.r; ++i) lh(i, j) *= wp[i];
  return matmul_tn(lh, rh);
}
}  // namespace

Mat RadialBasis::matel_rule(int iel, int lk, int rk, const std::vector<double> &x, const std::vector<double> &w,
                            const std::function<double(double)> *f) const {
  return matel_fixed(*this, iel, lk, rk, x, w, f);
}

// converge_panel, FiniteElementBasis.cpp:500-553, with the start order of
// matrix_element_auto (:556-604): basis degree nnodes-1, one panel [-1,1].
Mat RadialBasis::matrix_element_auto(int iel, int lk, int rk, const std::function<double(double)> *f) const {
  const int basisdeg = std::max(0, 1 * poly.nprim() - 1);
  const int deg = 2 * basisdeg + 0;
  const int nmax = 512;
  int nstart = (deg + 4) / 2 + 2;
  nstart = std::min(std::max(nstart, 5), nmax);
  const double eps = std::numeric_limits<double>::epsilon();
  const double tol = 8.0 * eps, sqrteps = std::sqrt(eps);
  Mat prev, cur;
  bool have = false;
  double prevdiff = -1.0, prevprevdiff = -1.0;
  int n = std::max(nstart, 2);
  std::vector<double> x, w;
  for (;;) {
    lobatto(n, x, w);
    cur = matel_fixed(*this, iel, lk, rk, x, w, f);
    if (have) {
      const double diff = max_abs_diff(cur, prev), scale = max_abs(cur);
      if (diff <= tol * (scale + tol)) return cur;
      if (prevdiff >= 0.0 && diff <= sqrteps * (scale + tol) && diff > 0.5 * prevdiff) return cur;
      if (prevprevdiff >= 0.0 && diff <= sqrteps * (scale + tol) && diff > 0.125 * prevprevdiff) return cur;
      prevprevdiff = prevdiff;
      prevdiff = diff;
    }
    prev = cur;
    have = true;
    if (n >= nmax) return cur;
    n = std::min(2 * n, nmax);
  }
}

Mat RadialBasis::radial_integral(int Rexp, int iel) const {
  const std::function<double(double)> f = [Rexp](double r) { return std::pow(r, double(Rexp + 2)); };
  return matrix_element_auto(iel, 10, 10, &f);
}

Mat RadialBasis::kinetic(int iel) const {
  Mat m = matrix_element_auto(iel, 1, 1, nullptr);
  for (auto &v : m.a) v *= 0.5;
  return m;
}

Mat RadialBasis::kinetic_l(int iel) const {
  Mat m = matrix_element_auto(iel, 10, 10, nullptr);
  for (auto &v : m.a) v *= 0.5;
  return m;
}

namespace {
// quadrature.cpp:36-74; fsb(r, R) is the kernel r<-part times R>-part
// (r^L / R^(L+1) for Coulomb, i_L(lambda r) k_L(lambda R) for Yukawa)
std::vector<double> inner_wrk(double rmin, double rmax, double rmin0, double rmax0,
                              const std::vector<double> &x, const std::vector<double> &wx,
                              const LIPBasis &poly, const std::function<double(double, double)> &fsbf) {
  const double rmid = 0.5 * (rmax + rmin), rlen = 0.5 * (rmax - rmin);
  const double rmid0 = 0.5 * (rmax0 + rmin0), rlen0 = 0.5 * (rmax0 - rmin0);
  const size_t n = x.size();
  std::vector<double> r(n), wp(n), xp(n);
  for (size_t i = 0; i < n; ++i) {
    r[i] = rmid + rlen * x[i];
    const double fsb = fsbf(r[i], rmax);
    wp[i] = wx[i] * fsb * rlen;
    xp[i] = (r[i] - rmid0) / rlen0;
  }
  const Mat bf = poly.eval_dnf(xp, 0, rlen0);
  Mat wbf = bf;
  for (int k = 0; k < wbf.c; ++k)
    for (int i = 0; i < wbf.r; ++i) wbf(i, k) *= wp[i];
  const Mat M = matmul_tn(wbf, bf);
  return M.a;
}
}  // namespace

// quadrature.cpp:76-161 with the kernel pieces fsb(r, R) and fbig(R)
Mat twoe_integral_rule_k(double rmin, double rmax, const std::vector<double> &x, const std::vector<double> &wx,
                         const LIPBasis &poly, const std::function<double(double, double)> &fsb,
                         const std::function<double(double)> &fbig) {
  const double rmid = 0.5 * (rmax + rmin), rlen = 0.5 * (rmax - rmin);
  const int n = (int)x.size(), nbf = poly.nbf(), nb2 = nbf * nbf;
  std::vector<double> r(n);
  for (int i = 0; i < n; ++i) r[i] = rmid + rlen * x[i];
  Mat inner(n, nb2);
  const bool empty_first = (r[0] == rmin);
  if (!empty_first) {
    const std::vector<double> v = inner_wrk(rmin, r[0], rmin, rmax, x, wx, poly, fsb);
    for (int k = 0; k < nb2; ++k) inner(0, k) = v[k];
  }
  for (int ip = 1; ip < n; ++ip) {
    const std::vector<double> v = inner_wrk(r[ip - 1], r[ip], rmin, rmax, x, wx, poly, fsb);
    for (int k = 0; k < nb2; ++k) inner(ip, k) = v[k];
  }
  for (int ip = 1; ip < n; ++ip) {
    if (ip == 1 && empty_first) continue;
    const double fac = fbig(r[ip]) / fbig(r[ip - 1]);
    for (int k = 0; k < nb2; ++k) inner(ip, k) += inner(ip - 1, k) * fac;
  }
  const Mat bf = poly.eval_dnf(x, 0, rlen);
  Mat bfprod(n, nb2);
  for (int fi = 0; fi < nbf; ++fi)
    for (int fj = 0; fj < nbf; ++fj)
      for (int i = 0; i < n; ++i) bfprod(i, fi * nbf + fj) = bf(i, fi) * bf(i, fj) * (wx[i] * rlen);
  Mat ints = matmul_tn(bfprod, inner);
  Mat out(nb2, nb2);
  for (int j = 0; j < nb2; ++j)
    for (int i = 0; i < nb2; ++i) out(i, j) = ints(i, j) + ints(j, i);
  return out;
}

Mat twoe_integral_rule(double rmin, double rmax, const std::vector<double> &x, const std::vector<double> &wx,
                       const LIPBasis &poly, int L) {
  const std::function<double(double, double)> fsb = [L](double r, double R) {
    return std::pow(r / R, double(L)) / R;
  };
  const std::function<double(double)> fbig = [L](double r) { return std::pow(r, double(-L - 1)); };
  return twoe_integral_rule_k(rmin, rmax, x, wx, poly, fsb, fbig);
}

Mat RadialBasis::twoe_integral(int L, int iel) const {
  const std::function<double(double, double)> fsb = [L](double r, double R) {
    return std::pow(r / R, double(L)) / R;
  };
  const std::function<double(double)> fbig = [L](double r) { return std::pow(r, double(-L - 1)); };
  return twoe_integral_conv(iel, fsb, fbig);
}

// converge_rule, RadialBasis.cpp:114-162, seeded at nquad (twoe_nstart :165-173)
Mat RadialBasis::twoe_integral_conv(int iel, const std::function<double(double, double)> &fsb,
                                    const std::function<double(double)> &fbig) const {
  const double Rmin = bval[iel], Rmax = bval[iel + 1];
  const LIPBasis p = get_basis(iel);
  const int nmax = 512;
  int nstart = std::min(std::max((int)xq.size(), 5), nmax);
  const double eps = std::numeric_limits<double>::epsilon();
  const double tol = 8.0 * eps, sqrteps = std::sqrt(eps);
  Mat prev, cur;
  bool have = false;
  double prevdiff = -1.0, prevprevdiff = -1.0;
  int n = std::max(nstart, 2);
  std::vector<double> x, w;
  for (;;) {
    lobatto(n, x, w);
    cur = twoe_integral_rule_k(Rmin, Rmax, x, w, p, fsb, fbig);
    if (have) {
      const double diff = max_abs_diff(cur, prev), scale = max_abs(cur);
      if (diff <= tol * (scale + tol)) return cur;
      if (prevdiff >= 0.0 && diff <= sqrteps * (scale + tol) && diff > 0.5 * prevdiff) return cur;
      if (prevprevdiff >= 0.0 && diff <= sqrteps * (scale + tol) && diff > 0.125 * prevprevdiff) return cur;
      prevprevdiff = prevdiff;
      prevdiff = diff;
    }
    prev = cur;
    have = true;
    if (n >= nmax) return cur;
    n = std::min(2 * n, nmax);
  }
}

// RadialBasis.cpp:671-709: greedy largest-diagonal pivoted Cholesky with an
// absolute stopping tolerance on the residual diagonal.
Mat pivoted_cholesky(const Mat &A, double tol) {
  const int n = A.r;
  std::vector<double> D(n);
  for (int i = 0; i < n; ++i) D[i] = A(i, i);
  std::vector<unsigned char> done(n, 0);
  std::vector<std::vector<double>> cols;
  for (int k = 0; k < n; ++k) {
    int pivot = n;
    double pv = tol;
    for (int i = 0; i < n; ++i)
      if (!done[i] && D[i] > pv) {
        pivot = i;
        pv = D[i];
      }
    if (pivot == n) break;
    done[pivot] = 1;
    const double sq = std::sqrt(pv);
    std::vector<double> col(n);
    for (int i = 0; i < n; ++i) {
      if (done[i] && i != pivot) {
        col[i] = 0.0;
        continue;
      }
      double s = A(i, pivot);
      for (size_t j = 0; j < cols.size(); ++j) s -= cols[j][i] * cols[j][pivot];
      col[i] = s / sq;
    }
    col[pivot] = sq;
    cols.push_back(col);
    for (int i = 0; i < n; ++i)
      if (!done[i]) D[i] -= col[i] * col[i];
  }
  Mat L(n, (int)cols.size());
  for (size_t j = 0; j < cols.size(); ++j)
    for (int i = 0; i < n; ++i) L(i, (int)j) = cols[j][i];
  return L;
}

Mat RadialBasis::twoe_integral_cholesky(int L, int iel, double tol) const {
  return pivoted_cholesky(twoe_integral(L, iel), tol);
}

}  // namespace orc
