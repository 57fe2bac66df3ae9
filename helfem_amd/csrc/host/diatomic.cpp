// Diatomic basis setup: angular maps, disjoint P/Q integrals, the two-channel
// in-element kernel and its sign-aware Cholesky factor, one-electron
// matrices.  Reference behaviour cited per function.
#include "diatomic.hpp"
#include <set>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace hfg {

namespace {

constexpr int kMaxRule = 512;

double maxabs(const Blk &A) {
  double m = 0.0;
  for (double x : A.v) m = std::max(m, std::abs(x));
  return m;
}

double maxdiff(const Blk &A, const Blk &B) {
  double m = 0.0;
  for (size_t i = 0; i < A.v.size(); ++i) m = std::max(m, std::abs(A.v[i] - B.v[i]));
  return m;
}

// Rule doubling with the diatomic acceptance test (src/diatomic/basis.cpp:126-200):
// accept at <= 8 eps relative change, or below sqrt(eps) once the change
// reaches the 256 eps floor or stops halving.  At the 512-point cap either
// the last result or, with `first_on_cap`, the starting-rule result is kept.
Blk converge(const std::function<Blk(int)> &eval, int nstart, int *nused, bool first_on_cap) {
  const double eps = std::numeric_limits<double>::epsilon();
  const double tol = 8.0 * eps, sq = std::sqrt(eps), floor_rel = 256.0 * eps;
  const int n0 = std::max(nstart, 2);
  Blk first, prev, cur;
  double lastdiff = -1.0;
  int n = n0;
  for (bool have = false;; have = true) {
    cur = eval(n);
    if (!have) first = cur;
    if (have) {
      const double diff = maxdiff(cur, prev), scale = maxabs(cur) + tol;
      const bool done = diff <= tol * scale ||
                        (diff <= sq * scale && (diff <= floor_rel * scale || (lastdiff >= 0.0 && diff > 0.5 * lastdiff)));
      if (done) {
        if (nused) *nused = n;
        return cur;
      }
      lastdiff = diff;
    }
    if (n >= kMaxRule) {
      if (nused) *nused = first_on_cap ? n0 : n;
      return first_on_cap ? first : cur;
    }
    prev = cur;
    n = std::min(2 * n, kMaxRule);
  }
}

double ipow(double x, int n) { return std::pow(x, double(n)); }

// x^n by repeated squaring, as Eigen's array pow with an int exponent
// evaluates the two-electron quadrature weights (quadrature.cpp:216, :241).
double int_pow(double x, int n) {
  double r = 1.0, b = x;
  for (unsigned e = (unsigned)n; e;) {
    if (e & 1u) r *= b;
    e >>= 1;
    if (e) b *= b;
  }
  return r;
}

// C = A' B with each entry summed in index order, eight columns of C at a
// time so the inner loop vectorises without reordering any sum.
Blk tn_product(const Blk &A, const Blk &B) {
  Blk C(A.c, B.c);
  const int K = A.r;
  for (int j0 = 0; j0 < B.c; j0 += 8) {
    const int nj = std::min(8, B.c - j0);
    for (int i = 0; i < A.c; ++i) {
      const double *a = &A.v[(size_t)i * K];
      double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (nj == 8) {
        const double *b = &B.v[(size_t)j0 * K];
        for (int k = 0; k < K; ++k) {
          const double ak = a[k];
#pragma GCC unroll 8
          for (int jj = 0; jj < 8; ++jj) s[jj] += ak * b[k + (size_t)jj * K];
        }
      } else {
        for (int jj = 0; jj < nj; ++jj) {
          const double *b = &B.v[(size_t)(j0 + jj) * K];
          double t = 0.0;
          for (int k = 0; k < K; ++k) t += a[k] * b[k];
          s[jj] = t;
        }
      }
      for (int jj = 0; jj < nj; ++jj) C(i, j0 + jj) = s[jj];
    }
  }
  return C;
}

// Data of one element for the in-element two-electron integral at an n-point
// Chebyshev rule (quadrature::twoe_element, src/diatomic/quadrature.cpp:162-207):
// outer nodes mu_i, and for each outer node the same rule mapped onto
// [mu_{i-1}, mu_i] for the cumulative inner integral.
struct ElementRule {
  std::vector<double> x, w, ch, sh;
  double len = 0.0;
  Blk bf, bfprod;
  std::vector<double> slen;
  std::vector<std::vector<double>> sch, ssh;
  std::vector<Blk> sbf;
};

ElementRule element_rule(const RadialFE &rad, int iel, int n) {
  ElementRule el;
  gauss_chebyshev2(n, el.x, el.w);
  const LipShape shape = rad.shape(iel);
  const double lo = rad.bval[iel], hi = rad.bval[iel + 1];
  const double mid = 0.5 * (hi + lo);
  el.len = 0.5 * (hi - lo);
  const int nb = shape.nbf();
  std::vector<double> mu(n);
  el.ch.resize(n);
  el.sh.resize(n);
  for (int i = 0; i < n; ++i) {
    mu[i] = mid + el.len * el.x[i];
    el.ch[i] = std::cosh(mu[i]);
    el.sh[i] = std::sinh(mu[i]);
  }
  el.bf = shape.eval(el.x, 0, el.len);
  el.bfprod = Blk(n, nb * nb);
  for (int a = 0; a < nb; ++a)
    for (int b = 0; b < nb; ++b)
      for (int i = 0; i < n; ++i) el.bfprod(i, a * nb + b) = el.bf(i, a) * el.bf(i, b);
  el.slen.resize(n);
  el.sch.assign(n, std::vector<double>(n));
  el.ssh.assign(n, std::vector<double>(n));
  el.sbf.resize(n);
  std::vector<double> xs(n);
  for (int p = 0; p < n; ++p) {
    const double a = p == 0 ? lo : mu[p - 1], b = mu[p];
    const double smid = 0.5 * (b + a);
    el.slen[p] = 0.5 * (b - a);
    for (int i = 0; i < n; ++i) {
      const double m = smid + el.slen[p] * el.x[i];
      el.sch[p][i] = std::cosh(m);
      el.ssh[p][i] = std::sinh(m);
      xs[i] = (m - mid) / el.len;
    }
    el.sbf[p] = shape.eval(xs, 0, el.len);
  }
  return el;
}

// Cumulative inner integrals I_p(ab) = int_{mu_lo}^{mu_p} B_a B_b sinh cosh^l P_L,
// rows p (quadrature.cpp:209-240, inner part).
Blk inner_cumulative(const ElementRule &el, int l, int L, LegendreCache &leg) {
  const int n = (int)el.x.size(), nb = el.bf.c, nb2 = nb * nb;
  Blk inner(n, nb2);
  std::vector<double> acc(nb2, 0.0), wp(n);
  for (int p = 0; p < n; ++p) {
    for (int i = 0; i < n; ++i) {
      double v = el.slen[p] * el.w[i];
      v *= el.ssh[p][i];
      if (l) v *= int_pow(el.sch[p][i], l);
      v *= leg.P(L, el.sch[p][i]);
      wp[i] = v;
    }
    Blk wbf = el.sbf[p];
    for (int c = 0; c < nb; ++c)
      for (int i = 0; i < n; ++i) wbf(i, c) *= wp[i];
    const Blk prod = tn_product(wbf, el.sbf[p]);
    for (int q = 0; q < nb2; ++q) {
      acc[q] += prod.v[q];
      inner(p, q) = acc[q];
    }
  }
  return inner;
}

// Outer contraction with B_c B_d sinh cosh^k Q_L (quadrature.cpp:241-258).
Blk outer_contract(const ElementRule &el, int k, int L, LegendreCache &leg, const Blk &inner) {
  const int n = (int)el.x.size(), nb2 = el.bfprod.c;
  std::vector<double> wp(n);
  for (int i = 0; i < n; ++i) {
    double v = el.len * el.w[i];
    v *= el.sh[i];
    if (k) v *= int_pow(el.ch[i], k);
    v *= leg.Q(L, el.ch[i]);
    wp[i] = v;
  }
  Blk wbp = el.bfprod;
  for (int c = 0; c < nb2; ++c)
    for (int i = 0; i < n; ++i) wbp(i, c) *= wp[i];
  return tn_product(wbp, inner);
}

// Symmetrised two-channel kernel W = [[T00, -T02], [-T02', T22]] with
// T_kl = O(k,l) + O(l,k)' and O(k,l) the outer(cosh^k) x inner(cosh^l) product
// (quadrature::twoe_integral, quadrature.cpp:209-258; basis.cpp:1470-1490).
Blk kernel_w(const ElementRule &el, int L, LegendreCache &leg) {
  const Blk in0 = inner_cumulative(el, 0, L, leg), in2 = inner_cumulative(el, 2, L, leg);
  const Blk o00 = outer_contract(el, 0, L, leg, in0), o02 = outer_contract(el, 0, L, leg, in2);
  const Blk o20 = outer_contract(el, 2, L, leg, in0), o22 = outer_contract(el, 2, L, leg, in2);
  const int n = o00.r;
  Blk T00(n, n), T02(n, n), T22(n, n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      T00(i, j) = o00(i, j) + o00(j, i);
      T02(i, j) = o02(i, j) + o20(j, i);
      T22(i, j) = o22(i, j) + o22(j, i);
    }
  const int N = 2 * n;
  Blk W(N, N);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      W(i, j) = T00(i, j);
      W(i, n + j) = -T02(i, j);
      W(n + i, j) = -T02(j, i);
      W(n + i, n + j) = T22(i, j);
    }
  Blk S(N, N);
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i) S(i, j) = 0.5 * (W(i, j) + W(j, i));
  return S;
}

// Sign-aware diagonal-pivoted Cholesky W = B diag(sigma) B', stop when the
// largest residual |diagonal| drops to 1e-12 of the initial largest
// (basis.cpp:1492-1537).
void signed_cholesky(const Blk &W, Blk &B, std::vector<double> &sigma) {
  const int N = W.r;
  std::vector<double> d(N);
  double dmax0 = 0.0;
  for (int i = 0; i < N; ++i) {
    d[i] = W(i, i);
    dmax0 = std::max(dmax0, std::abs(d[i]));
  }
  std::vector<double> cols;
  sigma.clear();
  int rank = 0;
  for (int p = 0; p < N; ++p) {
    int piv = 0;
    double best = std::abs(d[0]);
    for (int i = 1; i < N; ++i)
      if (std::abs(d[i]) > best) {
        best = std::abs(d[i]);
        piv = i;
      }
    if (dmax0 <= 0.0 || best <= 1e-12 * dmax0) break;
    const double dp = d[piv], s = dp >= 0.0 ? 1.0 : -1.0;
    cols.resize((size_t)(rank + 1) * N);
    double *col = &cols[(size_t)rank * N];
    for (int i = 0; i < N; ++i) col[i] = W(i, piv);
    for (int q = 0; q < rank; ++q) {
      const double *cq = &cols[(size_t)q * N];
      const double f = sigma[q] * cq[piv];
      for (int i = 0; i < N; ++i) col[i] -= f * cq[i];
    }
    const double sq = std::sqrt(std::abs(dp));
    for (int i = 0; i < N; ++i) col[i] /= sq;
    for (int i = 0; i < N; ++i) d[i] -= s * col[i] * col[i];
    d[piv] = 0.0;
    sigma.push_back(s);
    ++rank;
  }
  B = Blk(N, rank);
  B.v.assign(cols.begin(), cols.begin() + (size_t)rank * N);
}

}  // namespace

double gaunt3(int L, int M, int l, int m, int lp) {
  if (L < 0 || l < 0 || lp < 0) return 0.0;
  if (std::abs(M) > L || std::abs(m) > l || std::abs(M - m) > lp) return 0.0;
  return gaunt_coeff(L, M, l, m, lp, M - m);
}

// Coupling of the cosh^0 channel: Y_L^M expanded against cos^2 nu
// (gaunt.h:88-132, modified_coeff).
double gaunt_mod(int lj, int mj, int L, int M, int li, int mi) {
  if (mj != M + mi) return 0.0;
  static const double c0 = 2.0 / 3.0 * std::sqrt(M_PI);
  static const double c2 = 4.0 / 15.0 * std::sqrt(5.0 * M_PI);
  const double g0 = gaunt3(L, M, 0, 0, L) * gaunt3(lj, mj, li, mi, L);
  double g2 = 0.0;
  for (int Lp = std::max(std::max(L - 2, 0), std::abs(M)); Lp <= L + 2; ++Lp)
    g2 += gaunt3(Lp, M, 2, 0, L) * gaunt3(lj, mj, li, mi, Lp);
  return c0 * g0 + c2 * g2;
}

double gaunt_cos(int lj, int mj, int li, int mi) {
  if (mi != mj) return 0.0;
  static const double c1 = 2.0 * std::sqrt(M_PI / 3.0);
  return c1 * gaunt3(lj, mj, 1, 0, li);
}

double gaunt_cos2(int lj, int mj, int li, int mi) {
  if (mi != mj) return 0.0;
  static const double c0 = 2.0 / 3.0 * std::sqrt(M_PI);
  static const double c2 = 4.0 / 15.0 * std::sqrt(5.0 * M_PI);
  return c0 * gaunt3(lj, mj, 0, 0, li) + c2 * gaunt3(lj, mj, 2, 0, li);
}

void diatomic_shells(const std::vector<int> &lmax_per_m, std::vector<int> &lval, std::vector<int> &mval) {
  lval.clear();
  mval.clear();
  for (int ma = 0; ma < (int)lmax_per_m.size(); ++ma)
    for (int l = ma; l <= lmax_per_m[ma]; ++l) {
      lval.push_back(l);
      mval.push_back(ma);
      if (ma > 0) {
        lval.push_back(l);
        mval.push_back(-ma);
      }
    }
}

// TwoDBasis constructor maps (basis.cpp:525-647) and pure indices (:717-737).
void DiatomicSetup::build(int Z1_, int Z2_, double Rh_, int nnodes, int nquad_, const std::vector<double> &bval,
                          const std::vector<int> &lv, const std::vector<int> &mv) {
  if (lv.size() != mv.size() || lv.empty()) throw std::logic_error("angular basis: lval/mval mismatch");
  for (size_t i = 0; i < lv.size(); ++i)
    if (lv[i] < 0 || std::abs(mv[i]) > lv[i]) throw std::logic_error("angular basis: invalid (l,m)");
  if (!(Rh_ > 0.0)) throw std::logic_error("half bond length must be positive");
  Z1 = Z1_;
  Z2 = Z2_;
  Rh = Rh_;
  nquad = nquad_;
  rad = RadialFE(nnodes, bval, nquad_, false);
  lval = lv;
  mval = mv;
  std::set<std::pair<int, int>> absm_L, L_M;
  for (int i = 0; i < Nang(); ++i)
    for (int j = 0; j < Nang(); ++j) {
      const int M = mval[j] - mval[i];
      for (int L = std::max(std::abs(lval[j] - lval[i]) - 2, std::abs(M)); L <= lval[j] + lval[i] + 2; ++L) {
        absm_L.insert({std::abs(M), L});
        L_M.insert({L, M});
      }
    }
  lm_map.clear();
  for (auto &p : absm_L) lm_map.push_back({p.second, p.first});
  LM_map.assign(L_M.begin(), L_M.end());
  poff.assign(Nang() + 1, 0);
  for (int a = 0; a < Nang(); ++a) poff[a + 1] = poff[a] + (Nrad() - skip(a));
  P0.clear();
  P2.clear();
  Q0.clear();
  Q2.clear();
  cdB.clear();
  cdS.clear();
  have_tei = false;
}

int DiatomicSetup::lmind(int L, int M) const {
  // lm_map holds (L, |M|) ordered by |M| first
  auto less = [](const std::pair<int, int> &a, const std::pair<int, int> &b) {
    return a.second != b.second ? a.second < b.second : a.first < b.first;
  };
  const std::pair<int, int> key{L, std::abs(M)};
  auto it = std::lower_bound(lm_map.begin(), lm_map.end(), key, less);
  if (it == lm_map.end() || *it != key) throw std::logic_error("(L,|M|) not in the map");
  return (int)(it - lm_map.begin());
}

int DiatomicSetup::LMind(int L, int M) const {
  auto it = std::lower_bound(LM_map.begin(), LM_map.end(), std::make_pair(L, M));
  if (it == LM_map.end() || *it != std::make_pair(L, M)) throw std::logic_error("(L,M) not in the map");
  return (int)(it - LM_map.begin());
}

// (-1)^M 4 pi Rh^5 (L-|M|)!/(L+|M|)!, basis.cpp:1549-1560
double DiatomicSetup::LMfac(int L, int M) const {
  const int Ma = std::abs(M);
  double fr = 1.0;
  for (int p = L + Ma; p > L - Ma; --p) fr *= p;
  const double v = 4.0 * M_PI * std::pow(Rh, 5) / fr;
  return (M & 1) ? -v : v;
}

// RadialBasis::radial_integral(m, n), basis.cpp:289-310: whole-matrix
// convergence over Lobatto rules seeded at nquad.
Blk DiatomicSetup::radial_integral(int m, int n) const {
  std::function<double(double)> g;
  if (m != 0 && n != 0)
    g = [m, n](double mu) { return ipow(std::sinh(mu), m) * ipow(std::cosh(mu), n); };
  else if (m != 0)
    g = [m](double mu) { return ipow(std::sinh(mu), m); };
  else if (n != 0)
    g = [n](double mu) { return ipow(std::cosh(mu), n); };
  const std::function<double(double)> *f = g ? &g : nullptr;
  return converge(
      [&](int nq) {
        std::vector<double> x, w;
        gauss_lobatto(nq, x, w);
        Blk out(Nrad(), Nrad());
        for (int e = 0; e < Nel(); ++e) {
          const Blk b = rad.matel_rule(e, 0, 0, x, w, f);
          for (int jj = 0; jj < b.c; ++jj)
            for (int ii = 0; ii < b.r; ++ii) out(rad.first[e] + ii, rad.first[e] + jj) += b(ii, jj);
        }
        return out;
      },
      std::max(nquad, 5), nullptr, false);
}

// RadialBasis::kinetic, basis.cpp:396-409
Blk DiatomicSetup::radial_kinetic() const {
  const std::function<double(double)> sh = [](double mu) { return std::sinh(mu); };
  return converge(
      [&](int nq) {
        std::vector<double> x, w;
        gauss_lobatto(nq, x, w);
        Blk out(Nrad(), Nrad());
        for (int e = 0; e < Nel(); ++e) {
          const Blk b = rad.matel_rule(e, 1, 1, x, w, &sh);
          for (int jj = 0; jj < b.c; ++jj)
            for (int ii = 0; ii < b.r; ++ii) out(rad.first[e] + ii, rad.first[e] + jj) += b(ii, jj);
        }
        return out;
      },
      std::max(nquad, 5), nullptr, false);
}

// Plm_integral / Qlm_integral, basis.cpp:352-394: int B_i B_j sinh cosh^k P_L^|M|
// (or Q) over one element, Chebyshev rules seeded at nquad.
Blk DiatomicSetup::pq_integral(bool q, int k, int iel, int L, LegendreCache &leg) const {
  std::function<double(double)> f;
  if (q)
    f = [&leg, k, L](double mu) {
      return k ? std::sinh(mu) * ipow(std::cosh(mu), k) * leg.Q(L, std::cosh(mu)) : std::sinh(mu) * leg.Q(L, std::cosh(mu));
    };
  else
    f = [&leg, k, L](double mu) {
      return k ? std::sinh(mu) * ipow(std::cosh(mu), k) * leg.P(L, std::cosh(mu)) : std::sinh(mu) * leg.P(L, std::cosh(mu));
    };
  return converge(
      [&](int n) {
        std::vector<double> x, w;
        gauss_chebyshev2(n, x, w);
        return rad.matel_rule(iel, 0, 0, x, w, &f);
      },
      std::max(nquad, 5), nullptr, true);
}

// converged_twoe_order, basis.cpp:1334-1380: the rule at which the kernel of
// the largest (L, |M|) converges sets the rule for every (L, |M|) of the element.
int DiatomicSetup::twoe_order(int iel) const {
  const int nstart = std::min(std::max(nquad, 5), kMaxRule);
  if (lm_map.empty()) return nstart;
  int L = lm_map[0].first, M = lm_map[0].second;
  for (auto &e : lm_map)
    if (e.first > L || (e.first == L && e.second > M)) {
      L = e.first;
      M = e.second;
    }
  int nconv = nstart;
  converge(
      [&](int n) {
        const ElementRule el = element_rule(rad, iel, n);
        LegendreCache leg(M, L);
        return kernel_w(el, L, leg);
      },
      nstart, &nconv, false);
  return nconv;
}

namespace {
// runs of equal |M| in lm_map, with the largest L of each run
void m_runs(const std::vector<std::pair<int, int>> &lm_map, std::vector<int> &run_lo, std::vector<int> &run_hi,
            std::vector<int> &run_Lhi) {
  const int nlm = (int)lm_map.size();
  for (int i = 0; i < nlm;) {
    int j = i, Lhi = 0;
    while (j < nlm && lm_map[j].second == lm_map[i].second) Lhi = std::max(Lhi, lm_map[j++].first);
    run_lo.push_back(i);
    run_hi.push_back(j);
    run_Lhi.push_back(Lhi);
    i = j;
  }
}
}  // namespace

std::vector<double> DiatomicSetup::outer_legendre_q(int iel, int n, int Mabs, int Lmax) const {
  // the outer nodes of element_rule
  std::vector<double> x, w;
  gauss_chebyshev2(n, x, w);
  const double lo = rad.bval[iel], hi = rad.bval[iel + 1];
  const double mid = 0.5 * (hi + lo), len = 0.5 * (hi - lo);
  LegendreCache leg(Mabs, Lmax);
  std::vector<double> out((size_t)n * (Lmax + 1));
  for (int i = 0; i < n; ++i) {
    const double ch = std::cosh(mid + len * x[i]);
    for (int L = 0; L <= Lmax; ++L) out[(size_t)i * (Lmax + 1) + L] = leg.Q(L, ch);
  }
  return out;
}

void DiatomicSetup::compute_pq() {
  const int ne = Nel(), nlm = (int)lm_map.size();
  P0.assign((size_t)ne * nlm, Blk());
  P2 = P0;
  Q0 = P0;
  Q2 = P0;
  cdB = P0;
  cdS.assign((size_t)ne * nlm, std::vector<double>());
  std::vector<int> run_lo, run_hi, run_Lhi;
  m_runs(lm_map, run_lo, run_hi, run_Lhi);
  const int nrun = (int)run_lo.size();
#pragma omp parallel for schedule(dynamic)
  for (int t = 0; t < nrun * ne; ++t) {
    const int g = t / ne, e = t - g * ne;
    LegendreCache leg(lm_map[run_lo[g]].second, run_Lhi[g]);
    for (int q = run_lo[g]; q < run_hi[g]; ++q) {
      const int L = lm_map[q].first;
      const size_t i = (size_t)q * ne + e;
      P0[i] = pq_integral(false, 0, e, L, leg);
      P2[i] = pq_integral(false, 2, e, L, leg);
      Q0[i] = pq_integral(true, 0, e, L, leg);
      Q2[i] = pq_integral(true, 2, e, L, leg);
    }
  }
}

std::vector<int> DiatomicSetup::twoe_orders() const {
  std::vector<int> order(Nel());
#pragma omp parallel for schedule(dynamic)
  for (int e = 0; e < Nel(); ++e) order[e] = twoe_order(e);
  return order;
}

// compute_tei, basis.cpp:1382-1547
void DiatomicSetup::compute_tei() {
  compute_pq();
  const int ne = Nel();
  std::vector<int> run_lo, run_hi, run_Lhi;
  m_runs(lm_map, run_lo, run_hi, run_Lhi);
  const int nrun = (int)run_lo.size();
  std::vector<int> order(ne);
#pragma omp parallel for schedule(dynamic)
  for (int e = 0; e < ne; ++e) order[e] = twoe_order(e);
#pragma omp parallel for schedule(dynamic)
  for (int t = 0; t < nrun * ne; ++t) {
    const int g = t / ne, e = t - g * ne;
    const ElementRule el = element_rule(rad, e, order[e]);
    LegendreCache leg(lm_map[run_lo[g]].second, run_Lhi[g]);
    for (int q = run_lo[g]; q < run_hi[g]; ++q) {
      const Blk W = kernel_w(el, lm_map[q].first, leg);
      signed_cholesky(W, cdB[(size_t)q * ne + e], cdS[(size_t)q * ne + e]);
    }
  }
  have_tei = true;
}

// overlap, kinetic, nuclear in the pure basis (basis.cpp:1032-1156): each
// (shell, shell) block is the Rh-scaled radial combination, restricted to the
// radial functions that survive remove_boundaries.
namespace {
template <class F>
void collect_pure(const DiatomicSetup &s, BlockList &out, F block_value) {
  const int N = s.Nrad(), na = s.Nang();
  std::vector<std::pair<int, int>> pairs;
  for (int j = 0; j < na; ++j)
    for (int i = 0; i < na; ++i)
      if (s.mval[i] == s.mval[j]) pairs.emplace_back(i, j);
  std::vector<std::vector<double>> blk(pairs.size());
#pragma omp parallel for schedule(dynamic)
  for (size_t p = 0; p < pairs.size(); ++p) {
    const int i = pairs[p].first, j = pairs[p].second, si = s.skip(i), sj = s.skip(j);
    std::vector<double> v((size_t)(N - si) * (N - sj));
    bool any = false;
    for (int b = sj; b < N; ++b)
      for (int a = si; a < N; ++a) {
        const double x = block_value(i, j, a, b);
        v[(size_t)(a - si) + (size_t)(b - sj) * (N - si)] = x;
        any = any || x != 0.0;
      }
    if (any) blk[p] = std::move(v);
  }
  for (size_t p = 0; p < pairs.size(); ++p)
    if (!blk[p].empty()) {
      const int i = pairs[p].first, j = pairs[p].second;
      out.add(s.poff[i], s.poff[j], N - s.skip(i), N - s.skip(j), blk[p].data());
    }
}

// per shell pair coupling coefficient, evaluated once per block
template <class G>
std::vector<double> pair_table(const DiatomicSetup &s, G coef) {
  const int na = s.Nang();
  std::vector<double> t((size_t)na * na);
  for (int j = 0; j < na; ++j)
    for (int i = 0; i < na; ++i) t[(size_t)i * na + j] = coef(i, j);
  return t;
}
}  // namespace

void DiatomicSetup::one_electron_blocks(int which, BlockList &out) const {
  if (which == 0) {
    const Blk I10 = radial_integral(1, 0), I12 = radial_integral(1, 2);
    const double f = std::pow(Rh, 3);
    const std::vector<double> cc =
        pair_table(*this, [&](int i, int j) { return gaunt_cos2(lval[j], mval[j], lval[i], mval[i]); });
    collect_pure(*this, out, [&](int i, int j, int a, int b) {
      double v = 0.0;
      if (lval[i] == lval[j]) v += 1.0 * I12(a, b);
      const double c = cc[(size_t)i * Nang() + j];
      if (c != 0.0) v += -c * I10(a, b);
      return v * f;
    });
  } else if (which == 1) {
    const Blk Tr = radial_kinetic(), Ip1 = radial_integral(1, 0), Im1 = radial_integral(-1, 0);
    const double f = Rh / 2.0;
    collect_pure(*this, out, [&](int i, int j, int a, int b) {
      if (i != j) return 0.0;
      double v = 0.0;
      v += 1.0 * Tr(a, b);
      if (lval[i] != 0) v += double(lval[i] * (lval[i] + 1)) * Ip1(a, b);
      if (mval[i] != 0) v += double(mval[i] * mval[i]) * Im1(a, b);
      return v * f;
    });
  } else if (which == 2) {
    const Blk I10 = radial_integral(1, 0), I11 = radial_integral(1, 1);
    const double f = -std::pow(Rh, 2);
    const std::vector<double> cc =
        pair_table(*this, [&](int i, int j) { return gaunt_cos(lval[j], mval[j], lval[i], mval[i]); });
    collect_pure(*this, out, [&](int i, int j, int a, int b) {
      double v = 0.0;
      if (lval[i] == lval[j]) v += double(Z1 + Z2) * I11(a, b);
      if (Z1 != Z2) {
        const double c = cc[(size_t)i * Nang() + j];
        if (c != 0.0) v += (double(Z2 - Z1) * c) * I10(a, b);
      }
      return v * f;
    });
  } else {
    throw std::invalid_argument("which must be 0, 1 or 2");
  }
}

void DiatomicSetup::overlap(double *out, int64_t ld) const {
  BlockList b;
  one_electron_blocks(0, b);
  b.to_dense(out, ld, Nbf());
}

void DiatomicSetup::kinetic(double *out, int64_t ld) const {
  BlockList b;
  one_electron_blocks(1, b);
  b.to_dense(out, ld, Nbf());
}

void DiatomicSetup::nuclear(double *out, int64_t ld) const {
  BlockList b;
  one_electron_blocks(2, b);
  b.to_dense(out, ld, Nbf());
}

}  // namespace hfg
