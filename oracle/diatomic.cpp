// Oracle (test infrastructure only): diatomic (prolate spheroidal) TwoDBasis
// Fock-build path, restating src/diatomic/basis.cpp and
// src/diatomic/quadrature.cpp.  Coordinates: mu (radial FE), nu/phi (Y_lm).
#include "orc.hpp"
#include "diatomic.hpp"
#include <set>
#include <cstdio>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

namespace {
const int kTwoeMax = 512;

// converge_block, src/diatomic/basis.cpp:126-200
Mat converge_block(const std::function<Mat(int)> &probe, int nstart, int *nconv, bool seed_fallback) {
  const double eps = std::numeric_limits<double>::epsilon();
  const double tol = 8.0 * eps, sqrteps = std::sqrt(eps), floor_rel = 256.0 * eps;
  Mat prev, cur, seed;
  bool have = false;
  double prevdiff = -1.0;
  int n = std::max(nstart, 2);
  for (;;) {
    cur = probe(n);
    if (!have) seed = cur;
    if (have) {
      const double diff = max_abs_diff(cur, prev), scale = max_abs(cur);
      if (diff <= tol * (scale + tol)) {
        if (nconv) *nconv = n;
        return cur;
      }
      if (diff <= sqrteps * (scale + tol) &&
          (diff <= floor_rel * (scale + tol) || (prevdiff >= 0.0 && diff > 0.5 * prevdiff))) {
        if (nconv) *nconv = n;
        return cur;
      }
      prevdiff = diff;
    }
    prev = cur;
    have = true;
    if (n >= kTwoeMax) {
      if (seed_fallback) {
        if (nconv) *nconv = std::max(nstart, 2);
        return seed;
      }
      if (nconv) *nconv = n;
      return cur;
    }
    n = std::min(2 * n, kTwoeMax);
  }
}

double ipow(double x, int n) { return std::pow(x, double(n)); }

// Integer power by repeated squaring: the weights of the two-electron
// quadrature use Eigen's array pow with an int exponent (quadrature.cpp:216,
// :241), which multiplies rather than calling pow().
double ipow_sq(double x, int n) {
  double r = 1.0, b = x;
  for (unsigned e = (unsigned)n; e; ) {
    if (e & 1u) r *= b;
    e >>= 1;
    if (e) b *= b;
  }
  return r;
}
}  // namespace

// gaunt.h:53-132 semantics without storage bounds (the reference sizes its
// table to cover every index the diatomic basis asks for, basis.cpp:612-625)
double gaunt_coeff(int L, int M, int l, int m, int lp) {
  if (L < 0 || l < 0 || lp < 0) return 0.0;
  if (std::abs(M) > L || std::abs(m) > l || std::abs(M - m) > lp) return 0.0;
  return gaunt_coefficient(L, M, l, m, lp, M - m);
}

double gaunt_mod_coeff(int lj, int mj, int L, int M, int li, int mi) {
  if (mj != M + mi) return 0.0;
  static const double const0 = 2.0 / 3.0 * std::sqrt(M_PI);
  static const double const2 = 4.0 / 15.0 * std::sqrt(5.0 * M_PI);
  const double cpl0 = gaunt_coeff(L, M, 0, 0, L) * gaunt_coeff(lj, mj, li, mi, L);
  double cpl2 = 0.0;
  for (int Lp = std::max(std::max(L - 2, 0), std::abs(M)); Lp <= L + 2; ++Lp)
    cpl2 += gaunt_coeff(Lp, M, 2, 0, L) * gaunt_coeff(lj, mj, li, mi, Lp);
  return const0 * cpl0 + const2 * cpl2;
}

// gaunt.cpp:275-300
double gaunt_cosine(int lj, int mj, int li, int mi) {
  if (mi != mj) return 0.0;
  static const double c1 = 2.0 * std::sqrt(M_PI / 3.0);
  return c1 * gaunt_coeff(lj, mj, 1, 0, li);
}
double gaunt_cosine2(int lj, int mj, int li, int mi) {
  if (mi != mj) return 0.0;
  static const double c0 = 2.0 / 3.0 * std::sqrt(M_PI);
  static const double c2 = 4.0 / 15.0 * std::sqrt(5.0 * M_PI);
  return c0 * gaunt_coeff(lj, mj, 0, 0, li) + c2 * gaunt_coeff(lj, mj, 2, 0, li);
}

// lm_to_l_m, basis.cpp:500-516
void diatomic_angular(const std::vector<int> &lmmax, std::vector<int> &lval, std::vector<int> &mval) {
  lval.clear();
  mval.clear();
  for (int mabs = 0; mabs < (int)lmmax.size(); ++mabs)
    for (int l = mabs; l <= lmmax[mabs]; ++l) {
      lval.push_back(l);
      mval.push_back(mabs);
      if (mabs > 0) {
        lval.push_back(l);
        mval.push_back(-mabs);
      }
    }
}

// TwoDBasis ctor, basis.cpp:525-647 (maps) + pure_indices :717-737
DiatomicBasis::DiatomicBasis(int Z1_, int Z2_, double Rhalf_, int nnodes, int nquad_, const std::vector<double> &bval,
                             const std::vector<int> &lv, const std::vector<int> &mv)
    : Z1(Z1_), Z2(Z2_), Rhalf(Rhalf_), nquad(nquad_), radial(nnodes, bval, nquad_, false), lval(lv), mval(mv) {
  std::set<std::pair<int, int>> lms, LMs;
  for (int i = 0; i < Nang(); ++i)
    for (int j = 0; j < Nang(); ++j) {
      const int M = mval[j] - mval[i];
      for (int L = std::max(std::abs(lval[j] - lval[i]) - 2, std::abs(M)); L <= lval[j] + lval[i] + 2; ++L) {
        lms.insert({std::abs(M), L});  // |M|-major, L-minor (lm_less)
        LMs.insert({L, M});
      }
    }
  for (auto &p : lms) lm_map.push_back({p.second, p.first});
  LM_map.assign(LMs.begin(), LMs.end());
  for (int a = 0; a < Nang(); ++a) {
    const int s = mval[a] == 0 ? 0 : 1;
    for (int r = s; r < Nrad(); ++r) pure_idx.push_back(a * Nrad() + r);
  }
}

int DiatomicBasis::lmind(int L, int M) const {
  const std::pair<int, int> key{L, std::abs(M)};
  for (size_t i = 0; i < lm_map.size(); ++i)
    if (lm_map[i] == key) return (int)i;
  throw std::logic_error("(L,|M|) not in the map");
}

int DiatomicBasis::LMind(int L, int M) const {
  auto it = std::lower_bound(LM_map.begin(), LM_map.end(), std::make_pair(L, M));
  if (it == LM_map.end() || *it != std::make_pair(L, M)) throw std::logic_error("(L,M) not in the map");
  return (int)(it - LM_map.begin());
}

// RadialBasis::radial_integral(m, n), basis.cpp:289-310
Mat DiatomicBasis::radial_integral(int m, int n) const {
  std::function<double(double)> chsh;
  if (m != 0 && n != 0)
    chsh = [m, n](double mu) { return ipow(std::sinh(mu), m) * ipow(std::cosh(mu), n); };
  else if (m != 0)
    chsh = [m](double mu) { return ipow(std::sinh(mu), m); };
  else if (n != 0)
    chsh = [n](double mu) { return ipow(std::cosh(mu), n); };
  const std::function<double(double)> *f = chsh ? &chsh : nullptr;
  return converge_block(
      [&](int nq) {
        std::vector<double> x, w;
        lobatto(nq, x, w);
        Mat M(Nrad(), Nrad());
        for (int e = 0; e < radial.Nel(); ++e) {
          const Mat b = radial.matel_rule(e, 0, 0, x, w, f);
          const int fe = radial.first[e];
          for (int jj = 0; jj < b.c; ++jj)
            for (int ii = 0; ii < b.r; ++ii) M(fe + ii, fe + jj) += b(ii, jj);
        }
        return M;
      },
      std::max(nquad, 5), nullptr, false);
}

// RadialBasis::kinetic, basis.cpp:396-409
Mat DiatomicBasis::radial_kinetic() const {
  const std::function<double(double)> sh = [](double mu) { return std::sinh(mu); };
  return converge_block(
      [&](int nq) {
        std::vector<double> x, w;
        lobatto(nq, x, w);
        Mat M(Nrad(), Nrad());
        for (int e = 0; e < radial.Nel(); ++e) {
          const Mat b = radial.matel_rule(e, 1, 1, x, w, &sh);
          const int fe = radial.first[e];
          for (int jj = 0; jj < b.c; ++jj)
            for (int ii = 0; ii < b.r; ++ii) M(fe + ii, fe + jj) += b(ii, jj);
        }
        return M;
      },
      std::max(nquad, 5), nullptr, false);
}

// Plm_integral / Qlm_integral, basis.cpp:352-394
Mat DiatomicBasis::PQ_integral(bool Q, int k, int iel, int L, LegendreColumn &leg) const {
  std::function<double(double)> f;
  if (Q)
    f = [&leg, k, L](double mu) {
      return k ? std::sinh(mu) * ipow(std::cosh(mu), k) * leg.Q(L, std::cosh(mu)) : std::sinh(mu) * leg.Q(L, std::cosh(mu));
    };
  else
    f = [&leg, k, L](double mu) {
      return k ? std::sinh(mu) * ipow(std::cosh(mu), k) * leg.P(L, std::cosh(mu)) : std::sinh(mu) * leg.P(L, std::cosh(mu));
    };
  return converge_block(
      [&](int n) {
        std::vector<double> x, w;
        chebyshev(n, x, w);
        return radial.matel_rule(iel, 0, 0, x, w, &f);
      },
      std::max(nquad, 5), nullptr, true);
}

// quadrature::twoe_element, quadrature.cpp:162-207
TwoEl DiatomicBasis::twoe_element(int iel, int n) const {
  TwoEl el;
  chebyshev(n, el.x, el.w);
  const LIPBasis p = radial.get_basis(iel);
  const double mumin = radial.bval[iel], mumax = radial.bval[iel + 1];
  const double mumid = 0.5 * (mumax + mumin);
  el.mulen = 0.5 * (mumax - mumin);
  const int nq = n, nb = p.nbf();
  el.mu.resize(nq);
  el.chmu.resize(nq);
  el.shmu.resize(nq);
  for (int i = 0; i < nq; ++i) {
    el.mu[i] = mumid + el.mulen * el.x[i];
    el.chmu[i] = std::cosh(el.mu[i]);
    el.shmu[i] = std::sinh(el.mu[i]);
  }
  el.bf = p.eval_dnf(el.x, 0, el.mulen);
  el.bfprod = Mat(nq, nb * nb);
  for (int fi = 0; fi < nb; ++fi)
    for (int fj = 0; fj < nb; ++fj)
      for (int i = 0; i < nq; ++i) el.bfprod(i, fi * nb + fj) = el.bf(i, fi) * el.bf(i, fj);
  el.sub.resize(nq);
  for (int ip = 0; ip < nq; ++ip) {
    const double smin = ip == 0 ? mumin : el.mu[ip - 1], smax = el.mu[ip];
    const double smid = 0.5 * (smax + smin), slen = 0.5 * (smax - smin);
    TwoEl::Sub &s = el.sub[ip];
    s.mulen = slen;
    s.chmu.resize(nq);
    s.shmu.resize(nq);
    std::vector<double> xp(nq);
    for (int i = 0; i < nq; ++i) {
      const double mu = smid + slen * el.x[i];
      s.chmu[i] = std::cosh(mu);
      s.shmu[i] = std::sinh(mu);
      xp[i] = (mu - mumid) / el.mulen;
    }
    s.bf = p.eval_dnf(xp, 0, el.mulen);
  }
  return el;
}

// quadrature::twoe_integral(el, k, l, L, M), quadrature.cpp:209-258
Mat DiatomicBasis::twoe_integral(const TwoEl &el, int k, int l, int L, LegendreColumn &leg) const {
  const int nq = (int)el.x.size(), nb = el.bf.c, nb2 = nb * nb;
  auto outer = [&](int kv, int lv) {
    Mat inner(nq, nb2);
    std::vector<double> acc(nb2, 0.0), wp(nq);
    for (int ip = 0; ip < nq; ++ip) {
      const TwoEl::Sub &s = el.sub[ip];
      for (int i = 0; i < nq; ++i) {
        double v = s.mulen * el.w[i];
        v *= s.shmu[i];
        if (lv) v *= ipow_sq(s.chmu[i], lv);
        v *= leg.P(L, s.chmu[i]);
        wp[i] = v;
      }
      Mat wbf = s.bf;
      for (int c = 0; c < nb; ++c)
        for (int i = 0; i < nq; ++i) wbf(i, c) *= wp[i];
      const Mat prod = matmul_tn(wbf, s.bf);
      for (int q = 0; q < nb2; ++q) {
        acc[q] += prod.a[q];
        inner(ip, q) = acc[q];
      }
    }
    for (int i = 0; i < nq; ++i) {
      double v = el.mulen * el.w[i];
      v *= el.shmu[i];
      if (kv) v *= ipow_sq(el.chmu[i], kv);
      v *= leg.Q(L, el.chmu[i]);
      wp[i] = v;
    }
    Mat wbp = el.bfprod;
    for (int c = 0; c < nb2; ++c)
      for (int i = 0; i < nq; ++i) wbp(i, c) *= wp[i];
    return matmul_tn(wbp, inner);
  };
  const Mat t1 = outer(k, l), t2 = k == l ? t1 : outer(l, k);
  Mat out(nb2, nb2);
  for (int j = 0; j < nb2; ++j)
    for (int i = 0; i < nb2; ++i) out(i, j) = t1(i, j) + t2(j, i);
  return out;
}

Mat DiatomicBasis::kernel_W(const TwoEl &el, int L, LegendreColumn &leg) const {
  const Mat T00 = twoe_integral(el, 0, 0, L, leg), T02 = twoe_integral(el, 0, 2, L, leg),
            T22 = twoe_integral(el, 2, 2, L, leg);
  const int n = T00.r;
  Mat W(2 * n, 2 * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      W(i, j) = T00(i, j);
      W(i, n + j) = -T02(i, j);
      W(n + i, j) = -T02(j, i);
      W(n + i, n + j) = T22(i, j);
    }
  Mat S(2 * n, 2 * n);
  for (int j = 0; j < 2 * n; ++j)
    for (int i = 0; i < 2 * n; ++i) S(i, j) = 0.5 * (W(i, j) + W(j, i));
  return S;
}

// converged_twoe_order, basis.cpp:1334-1380
int DiatomicBasis::converged_twoe_order(int iel) const {
  int nstart = std::min(std::max(nquad, 5), kTwoeMax);
  if (lm_map.empty()) return nstart;
  int L = lm_map[0].first, M = lm_map[0].second;
  for (auto &e : lm_map)
    if (e.first > L || (e.first == L && e.second > M)) {
      L = e.first;
      M = e.second;
    }
  int nconv = nstart;
  converge_block(
      [&](int n) {
        const TwoEl el = twoe_element(iel, n);
        LegendreColumn leg(M, L);
        return kernel_W(el, L, leg);
      },
      nstart, &nconv, false);
  return nconv;
}

// compute_tei, basis.cpp:1382-1547
void DiatomicBasis::compute_tei() {
  const int Nel = radial.Nel(), nlm = (int)lm_map.size();
  P0.assign((size_t)Nel * nlm, Mat());
  P2 = P0;
  Q0 = P0;
  Q2 = P0;
  std::vector<std::pair<int, int>> runs;
  for (int i = 0; i < nlm;) {
    int j = i;
    while (j < nlm && lm_map[j].second == lm_map[i].second) ++j;
    runs.push_back({i, j});
    i = j;
  }
  std::vector<std::pair<int, int>> tasks;
  for (size_t g = 0; g < runs.size(); ++g)
    for (int e = 0; e < Nel; ++e) tasks.push_back({(int)g, e});
#pragma omp parallel for schedule(dynamic)
  for (int t = 0; t < (int)tasks.size(); ++t) {
    const int g = tasks[t].first, e = tasks[t].second;
    const int lo = runs[g].first, hi = runs[g].second, Mabs = lm_map[lo].second;
    int Lhi = 0;
    for (int q = lo; q < hi; ++q) Lhi = std::max(Lhi, lm_map[q].first);
    LegendreColumn leg(Mabs, Lhi);
    for (int q = lo; q < hi; ++q) {
      const int L = lm_map[q].first;
      const size_t i = (size_t)q * Nel + e;
      P0[i] = PQ_integral(false, 0, e, L, leg);
      P2[i] = PQ_integral(false, 2, e, L, leg);
      Q0[i] = PQ_integral(true, 0, e, L, leg);
      Q2[i] = PQ_integral(true, 2, e, L, leg);
    }
  }
  cd_B.assign((size_t)Nel * nlm, Mat());
  cd_sigma.assign((size_t)Nel * nlm, std::vector<double>());
#pragma omp parallel for schedule(dynamic)
  for (int e = 0; e < Nel; ++e) {
    const int nconv = converged_twoe_order(e);
    const TwoEl el = twoe_element(e, nconv);
    LegendreColumn leg(0, 0);
    int curM = -1;
    for (int q = 0; q < nlm; ++q) {
      const int L = lm_map[q].first, M = lm_map[q].second;
      if (M != curM) {
        curM = M;
        int Lhi = 0;
        for (int i = q; i < nlm && lm_map[i].second == M; ++i) Lhi = std::max(Lhi, lm_map[i].first);
        leg = LegendreColumn(M, Lhi);
      }
      const Mat W = kernel_W(el, L, leg);
      // sign-aware diagonal-pivoted Cholesky, relative threshold 1e-12 (:1492-1537)
      const int N = W.r;
      std::vector<double> d(N);
      double dmax0 = 0.0;
      for (int i = 0; i < N; ++i) {
        d[i] = W(i, i);
        dmax0 = std::max(dmax0, std::abs(d[i]));
      }
      std::vector<std::vector<double>> cols;
      std::vector<double> sg;
      for (int p = 0; p < N; ++p) {
        int piv = 0;
        double dp_abs = std::abs(d[0]);
        for (int i = 1; i < N; ++i)
          if (std::abs(d[i]) > dp_abs) {
            dp_abs = std::abs(d[i]);
            piv = i;
          }
        if (dmax0 <= 0.0 || dp_abs <= 1e-12 * dmax0) break;
        const double dp = d[piv], s = dp >= 0.0 ? 1.0 : -1.0;
        std::vector<double> col(N);
        for (int i = 0; i < N; ++i) col[i] = W(i, piv);
        for (size_t q2 = 0; q2 < cols.size(); ++q2) {
          const double f = sg[q2] * cols[q2][piv];
          for (int i = 0; i < N; ++i) col[i] -= f * cols[q2][i];
        }
        const double sq = std::sqrt(std::abs(dp));
        for (int i = 0; i < N; ++i) col[i] /= sq;
        for (int i = 0; i < N; ++i) d[i] -= s * col[i] * col[i];
        d[piv] = 0.0;
        cols.push_back(col);
        sg.push_back(s);
      }
      Mat B(N, (int)cols.size());
      for (size_t p = 0; p < cols.size(); ++p)
        for (int i = 0; i < N; ++i) B(i, (int)p) = cols[p][i];
      cd_B[(size_t)q * Nel + e] = B;
      cd_sigma[(size_t)q * Nel + e] = sg;
    }
  }
}

Mat DiatomicBasis::remove_boundaries(const Mat &F) const {
  const int n = (int)pure_idx.size();
  Mat O(n, n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) O(i, j) = F(pure_idx[i], pure_idx[j]);
  return O;
}

Mat DiatomicBasis::expand_boundaries(const Mat &P) const {
  const int n = (int)pure_idx.size();
  if (P.r != n || P.c != n) throw std::logic_error("expand_boundaries: size mismatch");
  Mat O(Ndummy(), Ndummy());
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) O(pure_idx[i], pure_idx[j]) = P(i, j);
  return O;
}

void DiatomicBasis::add_sub(Mat &M, int i, int j, const Mat &S, double f) const {
  const int N = Nrad();
  for (int b = 0; b < N; ++b)
    for (int a = 0; a < N; ++a) M(i * N + a, j * N + b) += f * S(a, b);
}

// overlap, kinetic, nuclear: basis.cpp:1032-1156
Mat DiatomicBasis::overlap() const {
  const Mat I10 = radial_integral(1, 0), I12 = radial_integral(1, 2);
  Mat S(Ndummy(), Ndummy());
  for (int i = 0; i < Nang(); ++i)
    for (int j = 0; j < Nang(); ++j) {
      if (mval[i] != mval[j]) continue;
      if (lval[i] == lval[j]) add_sub(S, i, j, I12, 1.0);
      const double c = gaunt_cosine2(lval[j], mval[j], lval[i], mval[i]);
      if (c != 0.0) add_sub(S, i, j, I10, -c);
    }
  const double f = std::pow(Rhalf, 3);
  for (auto &v : S.a) v *= f;
  return remove_boundaries(S);
}

Mat DiatomicBasis::kinetic() const {
  const Mat Tr = radial_kinetic(), Ip1 = radial_integral(1, 0), Im1 = radial_integral(-1, 0);
  Mat T(Ndummy(), Ndummy());
  for (int a = 0; a < Nang(); ++a) {
    add_sub(T, a, a, Tr, 1.0);
    if (lval[a] != 0) add_sub(T, a, a, Ip1, double(lval[a] * (lval[a] + 1)));
    if (mval[a] != 0) add_sub(T, a, a, Im1, double(mval[a] * mval[a]));
  }
  for (auto &v : T.a) v *= Rhalf / 2.0;
  return remove_boundaries(T);
}

Mat DiatomicBasis::nuclear() const {
  const Mat I10 = radial_integral(1, 0), I11 = radial_integral(1, 1);
  Mat V(Ndummy(), Ndummy());
  for (int i = 0; i < Nang(); ++i)
    for (int j = 0; j < Nang(); ++j) {
      if (mval[i] != mval[j]) continue;
      if (lval[i] == lval[j]) add_sub(V, i, j, I11, double(Z1 + Z2));
      if (Z1 != Z2) {
        const double c = gaunt_cosine(lval[j], mval[j], lval[i], mval[i]);
        if (c != 0.0) add_sub(V, i, j, I10, double(Z2 - Z1) * c);
      }
    }
  const double f = -std::pow(Rhalf, 2);
  for (auto &v : V.a) v *= f;
  return remove_boundaries(V);
}

// build_LMfac_abs, basis.cpp:1549-1560
double DiatomicBasis::LMfac(int L, int M) const {
  const int Ma = std::abs(M);
  double fr = 1.0;
  for (int p = L + Ma; p > L - Ma; --p) fr *= p;
  const double v = 4.0 * M_PI * std::pow(Rhalf, 5) / fr;
  return (M & 1) ? -v : v;
}

// coulomb, basis.cpp:1627-1816
Mat DiatomicBasis::coulomb(const Mat &Pin) const {
  if (cd_B.empty()) throw std::logic_error("Primitive teis have not been computed!");
  const Mat P = expand_boundaries(Pin);
  const int Nel = radial.Nel(), N = Nrad(), nLM = (int)LM_map.size();
  std::vector<Mat> Pa0(nLM, Mat(N, N)), Pa2(nLM, Mat(N, N));
  for (int k = 0; k < Nang(); ++k)
    for (int l = 0; l < Nang(); ++l) {
      const int lk = lval[k], mk = mval[k], ll = lval[l], ml = mval[l], M = mk - ml;
      for (int L = std::max(std::abs(lk - ll) - 2, std::abs(M)); L <= lk + ll + 2; ++L) {
        const int q = LMind(L, M);
        const double c0 = gaunt_mod_coeff(lk, mk, L, M, ll, ml), c2 = gaunt_coeff(lk, mk, L, M, ll);
        for (int b = 0; b < N; ++b)
          for (int a = 0; a < N; ++a) {
            const double v = P(k * N + a, l * N + b);
            if (c0 != 0.0) Pa0[q](a, b) += c0 * v;
            if (c2 != 0.0) Pa2[q](a, b) += c2 * v;
          }
      }
    }
  std::vector<Mat> Ja0(nLM, Mat(N, N)), Ja2(nLM, Mat(N, N));
  for (int q = 0; q < nLM; ++q) {
    const int L = LM_map[q].first, M = LM_map[q].second;
    const int ilm = lmind(L, M);
    const double fac = LMfac(L, M);
    for (int jel = 0; jel < Nel; ++jel) {
      const int jf = radial.first[jel], Nj = radial.Nprim(jel);
      Mat s0(Nj, Nj), s2(Nj, Nj);
      for (int b = 0; b < Nj; ++b)
        for (int a = 0; a < Nj; ++a) {
          s0(a, b) = Pa0[q](jf + a, jf + b);
          s2(a, b) = Pa2[q](jf + a, jf + b);
        }
      auto tr = [&](const Mat &A, const Mat &X) {
        double t = 0.0;
        for (int a = 0; a < Nj; ++a)
          for (int b = 0; b < Nj; ++b) t += A(a, b) * X(b, a);
        return t;
      };
      const size_t jj = (size_t)ilm * Nel + jel;
      const double js0 = fac * tr(P0[jj], s0), jb0 = fac * tr(Q0[jj], s0);
      const double js2 = fac * tr(P2[jj], s2), jb2 = fac * tr(Q2[jj], s2);
      for (int iel = 0; iel < Nel; ++iel) {
        if (iel == jel) continue;
        const int fi = radial.first[iel], Ni = radial.Nprim(iel);
        const size_t ii = (size_t)ilm * Nel + iel;
        const bool lower = iel < jel;
        const double f0 = lower ? (jb0 - jb2) : (js0 - js2), f2 = lower ? (-jb0 + jb2) : (-js0 + js2);
        const Mat &m0 = lower ? P0[ii] : Q0[ii];
        const Mat &m2 = lower ? P2[ii] : Q2[ii];
        for (int b = 0; b < Ni; ++b)
          for (int a = 0; a < Ni; ++a) {
            Ja0[q](fi + a, fi + b) += f0 * m0(a, b);
            Ja2[q](fi + a, fi + b) += f2 * m2(a, b);
          }
      }
      const Mat &B = cd_B[jj];
      const std::vector<double> &sg = cd_sigma[jj];
      const int nn = Nj * Nj;
      std::vector<double> p2(2 * nn);
      for (int u = 0; u < nn; ++u) {
        p2[u] = s0.a[u];
        p2[nn + u] = s2.a[u];
      }
      std::vector<double> c(B.c, 0.0);
      for (int p = 0; p < B.c; ++p) {
        double t = 0.0;
        for (int u = 0; u < 2 * nn; ++u) t += B(u, p) * p2[u];
        c[p] = t * sg[p];
      }
      for (int u = 0; u < 2 * nn; ++u) {
        double t = 0.0;
        for (int p = 0; p < B.c; ++p) t += B(u, p) * c[p];
        t *= fac;
        const int uu = u % nn, a = uu % Nj, b = uu / Nj;
        if (u < nn)
          Ja0[q](jf + a, jf + b) += t;
        else
          Ja2[q](jf + a, jf + b) += t;
      }
    }
  }
  Mat J(Ndummy(), Ndummy());
  for (int i = 0; i < Nang(); ++i)
    for (int j = 0; j < Nang(); ++j) {
      const int li = lval[i], mi = mval[i], lj = lval[j], mj = mval[j], M = mj - mi;
      for (int L = std::max(std::abs(lj - li) - 2, std::abs(M)); L <= lj + li + 2; ++L) {
        const int q = LMind(L, M);
        const double c0 = gaunt_mod_coeff(lj, mj, L, M, li, mi);
        if (c0 != 0.0) add_sub(J, i, j, Ja0[q], c0);
        const double c2 = gaunt_coeff(lj, mj, L, M, li);
        if (c2 != 0.0) add_sub(J, i, j, Ja2[q], c2);
      }
    }
  return remove_boundaries(J);
}

// exchange, basis.cpp:1818-2089 (absm_symmetric = false)
Mat DiatomicBasis::exchange(const Mat &Pin) const {
  std::vector<int> all((size_t)Nang() * Nang());
  for (size_t q = 0; q < all.size(); ++q) all[q] = (int)q;
  return exchange_blocks(Pin, all);
}

// The same restricted to the output blocks j*Nang+k listed in jks.
Mat DiatomicBasis::exchange_blocks(const Mat &Pin, const std::vector<int> &jks) const {
  if (cd_B.empty()) throw std::logic_error("Primitive teis have not been computed!");
  const Mat P = expand_boundaries(Pin);
  const int Nel = radial.Nel(), N = Nrad(), Na = Nang(), nlm = (int)lm_map.size();
  int mabs = 0;
  for (int m : mval) mabs = std::max(mabs, std::abs(m));
  const int dmoff = 2 * mabs;
  std::vector<std::vector<std::pair<int, int>>> bydm(4 * mabs + 1);
  for (int i = 0; i < Na; ++i)
    for (int l = 0; l < Na; ++l) {
      double s = 0.0;
      for (int b = 0; b < N; ++b)
        for (int a = 0; a < N; ++a) {
          const double v = P(i * N + a, l * N + b);
          s += v * v;
        }
      if (std::sqrt(s) < 10 * std::numeric_limits<double>::epsilon()) continue;
      const int dm = mval[i] - mval[l];
      if (std::abs(dm) > 2 * mabs) continue;
      bydm[dm + dmoff].push_back({i, l});
    }
  Mat K(Ndummy(), Ndummy());
#pragma omp parallel for schedule(dynamic)
  for (size_t t = 0; t < jks.size(); ++t) {
      const int j = jks[t] / Na, k = jks[t] % Na;
      const int lj = lval[j], mj = mval[j], lk = lval[k], mk = mval[k];
      const int dmjk = mj - mk;
      if (std::abs(dmjk) > 2 * mabs) continue;
      const auto &cand = bydm[dmjk + dmoff];
      if (cand.empty()) continue;
      std::vector<Mat> R00(nlm), R02(nlm), R20(nlm), R22(nlm);
      std::vector<bool> couple(nlm, false);
      bool any = false;
      for (auto &pr : cand) {
        const int i = pr.first, l = pr.second;
        const int li = lval[i], mi = mval[i], ll = lval[l], ml = mval[l];
        const int M = mj - mi;
        const int Lmin = std::max(std::max(std::abs(li - lj), std::abs(lk - ll)) - 2, std::abs(M));
        const int Lmax = std::min(li + lj, lk + ll) + 2;
        for (int L = Lmin; L <= Lmax; ++L) {
          const double mji = gaunt_mod_coeff(lj, mj, L, M, li, mi), mkl = gaunt_mod_coeff(lk, mk, L, M, ll, ml);
          const double gji = gaunt_coeff(lj, mj, L, M, li), gkl = gaunt_coeff(lk, mk, L, M, ll);
          const double c00 = mji * mkl, c02 = -mji * gkl, c20 = -gji * mkl, c22 = gji * gkl;
          if (c00 == 0.0 && c02 == 0.0 && c20 == 0.0 && c22 == 0.0) continue;
          const int ilm = lmind(L, M);
          const double fac = LMfac(L, M);
          if (!couple[ilm]) {
            R00[ilm] = Mat(N, N);
            R02[ilm] = Mat(N, N);
            R20[ilm] = Mat(N, N);
            R22[ilm] = Mat(N, N);
            couple[ilm] = true;
            any = true;
          }
          const double f00 = fac * c00, f02 = fac * c02, f20 = fac * c20, f22 = fac * c22;
          for (int b = 0; b < N; ++b)
            for (int a = 0; a < N; ++a) {
              const double v = P(i * N + a, l * N + b);
              R00[ilm](a, b) += f00 * v;
              R02[ilm](a, b) += f02 * v;
              R20[ilm](a, b) += f20 * v;
              R22[ilm](a, b) += f22 * v;
            }
        }
      }
      if (!any) continue;
      for (int iel = 0; iel < Nel; ++iel) {
        const int fi = radial.first[iel], Ni = radial.Nprim(iel);
        for (int jel = 0; jel < Nel; ++jel) {
          const int fj = radial.first[jel], Nj = radial.Nprim(jel);
          auto sub = [&](const Mat &R) {
            Mat X(Ni, Nj);
            for (int b = 0; b < Nj; ++b)
              for (int a = 0; a < Ni; ++a) X(a, b) = R(fi + a, fj + b);
            return X;
          };
          Mat Ks(Ni, Nj);
          for (int ilm = 0; ilm < nlm; ++ilm) {
            if (!couple[ilm]) continue;
            const Mat r00 = sub(R00[ilm]), r02 = sub(R02[ilm]), r20 = sub(R20[ilm]), r22 = sub(R22[ilm]);
            if (iel == jel) {
              const Mat &B = cd_B[(size_t)ilm * Nel + iel];
              const std::vector<double> &sg = cd_sigma[(size_t)ilm * Nel + iel];
              const int nn = Ni * Ni;
              for (int p = 0; p < B.c; ++p) {
                Mat M0t(Ni, Ni), M2t(Ni, Ni);  // transposed channel blocks
                for (int b = 0; b < Ni; ++b)
                  for (int a = 0; a < Ni; ++a) {
                    M0t(b, a) = B(a + b * Ni, p);
                    M2t(b, a) = B(nn + a + b * Ni, p);
                  }
                const Mat t1 = matmul(M0t, matmul(r00, M0t)), t2 = matmul(M0t, matmul(r02, M2t)),
                          t3 = matmul(M2t, matmul(r20, M0t)), t4 = matmul(M2t, matmul(r22, M2t));
                for (size_t u = 0; u < Ks.a.size(); ++u) Ks.a[u] += sg[p] * (t1.a[u] - t2.a[u] - t3.a[u] + t4.a[u]);
              }
            } else {
              const size_t ii = (size_t)ilm * Nel + iel, jj = (size_t)ilm * Nel + jel;
              const bool up = iel > jel;
              const Mat &i0 = up ? Q0[ii] : P0[ii], &i2 = up ? Q2[ii] : P2[ii];
              const Mat &j0 = up ? P0[jj] : Q0[jj], &j2 = up ? P2[jj] : Q2[jj];
              const Mat Ta = matmul_nt(r00, j0), Tb = matmul_nt(r02, j2);
              Mat T(Ni, Nj);
              for (size_t u = 0; u < T.a.size(); ++u) T.a[u] = Ta.a[u] + Tb.a[u];
              const Mat X1 = matmul(i0, T);
              const Mat Tc = matmul_nt(r20, j0), Td = matmul_nt(r22, j2);
              for (size_t u = 0; u < T.a.size(); ++u) T.a[u] = Tc.a[u] + Td.a[u];
              const Mat X2 = matmul(i2, T);
              for (size_t u = 0; u < Ks.a.size(); ++u) Ks.a[u] -= X1.a[u] + X2.a[u];
            }
          }
          // reference: in-element K.block -= Ksubm ; off-element K.block += Ksub
          for (int b = 0; b < Nj; ++b)
            for (int a = 0; a < Ni; ++a) {
              const double v = (iel == jel) ? -Ks(a, b) : Ks(a, b);
              K(j * N + fi + a, k * N + fj + b) += v;
            }
        }
      }
    }
  return remove_boundaries(K);
}

}  // namespace orc
