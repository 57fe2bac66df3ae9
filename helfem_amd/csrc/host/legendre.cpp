// Associated Legendre functions P_l^m(x), Q_l^m(x) for x >= 1 (and on the
// cut), HelFEM's convention (Hobson form off the cut; src/legendre/Legendre.h).
// P: upward recurrence in l from the diagonal seeds.  Q: closed forms for
// l <= 1, then Christoffel's decomposition Q_l = P_l Q_0 - W_l near x = 1,
// Miller's downward recurrence with a Lentz continued-fraction start above
// cosh(ln(1e3)/(2 lmax)), and the upward m recurrence for m >= 2.
#include "diatomic.hpp"

namespace hfg {
namespace {

inline size_t at(int l, int m, int lmax) { return (size_t)m * (size_t)(lmax + 1) + (size_t)l; }
inline double side(double x) { return std::abs(x) <= 1.0 ? 1.0 : -1.0; }
inline double root0(double y) { return std::sqrt(std::max(y, 0.0)); }
inline double qzero(double x) { return 0.5 * (std::log(std::abs(x + 1.0)) - std::log(std::abs(x - 1.0))); }

void q_m_upward(double *Q, int lmax, int mmax, double x) {
  if (mmax < 2) return;
  const double s = side(x), w = root0(s * (1.0 - x * x));
  if (w == 0.0) throw std::domain_error("Legendre Q: singular m recurrence at |x| = 1");
  for (int l = 0; l <= lmax; ++l)
    for (int m = 1; m < mmax; ++m)
      Q[at(l, m + 1, lmax)] = -double(2 * m) * x / w * Q[at(l, m, lmax)] -
                              s * double(l + m) * double(l - m + 1) * Q[at(l, m - 1, lmax)];
}

double q_ratio_cf(double x, int lmax, int m) {
  const double tiny = std::numeric_limits<double>::min() * 1e4;
  const double tol = 8.0 * std::numeric_limits<double>::epsilon();
  double f = tiny, C = f, D = 0.0, a = 1.0;
  int n = lmax;
  for (int it = 0; it < 1000000; ++it, ++n) {
    const double b = double(2 * n + 1) * x / double(n + m);
    D = b + a * D;
    if (D == 0.0) D = tiny;
    C = b + a / C;
    if (C == 0.0) C = tiny;
    D = 1.0 / D;
    const double delta = C * D;
    f *= delta;
    if (std::abs(delta - 1.0) < tol) return f;
    a = -double(n - m + 1) / double(n + m);
  }
  throw std::runtime_error("Legendre Q ratio continued fraction failed to converge");
}

void q_downward(double *Q, int lmax, int m, double x) {
  for (int l = lmax - 1; l >= 1; --l)
    Q[at(l - 1, m, lmax)] =
        (double(2 * l + 1) * x * Q[at(l, m, lmax)] - double(l + 1 - m) * Q[at(l + 1, m, lmax)]) / double(l + m);
}

}  // namespace

void legendre_p(double *P, int lmax, int mmax, double x) {
  std::fill_n(P, (size_t)(lmax + 1) * (mmax + 1), 0.0);
  const double s = side(x), w = root0(s * (1.0 - x * x));
  P[at(0, 0, lmax)] = 1.0;
  for (int m = 1; m <= std::min(mmax, lmax); ++m) P[at(m, m, lmax)] = -s * double(2 * m - 1) * w * P[at(m - 1, m - 1, lmax)];
  for (int m = 0; m <= std::min(mmax, lmax - 1); ++m) P[at(m + 1, m, lmax)] = double(2 * m + 1) * x * P[at(m, m, lmax)];
  for (int m = 0; m <= std::min(mmax, lmax); ++m)
    for (int l = m + 1; l < lmax; ++l)
      P[at(l + 1, m, lmax)] =
          (double(2 * l + 1) * x * P[at(l, m, lmax)] - double(l + m) * P[at(l - 1, m, lmax)]) / double(l + 1 - m);
}

void legendre_q(double *Q, int lmax, int mmax, double x) {
  if (std::abs(x) == 1.0) throw std::domain_error("Legendre Q is singular at |x| = 1");
  std::fill_n(Q, (size_t)(lmax + 1) * (mmax + 1), 0.0);
  const double s = side(x), w = root0(s * (1.0 - x * x)), q0 = qzero(x);
  Q[at(0, 0, lmax)] = q0;
  if (lmax >= 1) Q[at(1, 0, lmax)] = x * q0 - 1.0;
  if (mmax >= 1) {
    Q[at(0, 1, lmax)] = -1.0 / w;
    if (lmax >= 1) Q[at(1, 1, lmax)] = -s * w * (q0 + x / (1.0 - x * x));
  }
  if (std::abs(x) <= 1.0) {
    for (int m = 0; m <= std::min(mmax, 1); ++m)
      if (lmax >= 2)
        for (int l = 1; l < lmax; ++l)
          Q[at(l + 1, m, lmax)] =
              (double(2 * l + 1) * x * Q[at(l, m, lmax)] - double(l + m) * Q[at(l - 1, m, lmax)]) / double(l + 1 - m);
    q_m_upward(Q, lmax, mmax, x);
    return;
  }
  const double xsplit = lmax > 0 ? std::cosh(std::log(1e3) / double(2 * lmax)) : std::numeric_limits<double>::infinity();
  if (std::abs(x) < xsplit) {
    // Christoffel: Q_l = P_l Q_0 - W_l, Q_l^1 = w P_l' Q_0 - P_l/w - w W_l'
    std::vector<double> p(lmax + 1, 0.0), W(lmax + 1, 0.0), dp(lmax + 1, 0.0), dW(lmax + 1, 0.0);
    p[0] = 1.0;
    if (lmax >= 1) {
      p[1] = x;
      W[1] = 1.0;
      dp[1] = 1.0;
    }
    for (int n = 1; n < lmax; ++n) {
      const double inv = 1.0 / double(n + 1), a = double(2 * n + 1), b = double(n);
      p[n + 1] = (a * x * p[n] - b * p[n - 1]) * inv;
      W[n + 1] = (a * x * W[n] - b * W[n - 1]) * inv;
      dp[n + 1] = (a * (p[n] + x * dp[n]) - b * dp[n - 1]) * inv;
      dW[n + 1] = (a * (W[n] + x * dW[n]) - b * dW[n - 1]) * inv;
    }
    for (int l = 0; l <= lmax; ++l) Q[at(l, 0, lmax)] = p[l] * q0 - W[l];
    if (mmax >= 1) {
      const double v = std::sqrt(std::max(x * x - 1.0, 0.0));
      Q[at(0, 1, lmax)] = -1.0 / v;
      for (int l = 1; l <= lmax; ++l) Q[at(l, 1, lmax)] = v * dp[l] * q0 - p[l] / v - v * dW[l];
    }
  } else {
    const double seed = std::numeric_limits<double>::min() * 1e4;
    if (lmax >= 1) {
      const double r0 = q_ratio_cf(x, lmax, 0);
      Q[at(lmax - 1, 0, lmax)] = seed;
      Q[at(lmax, 0, lmax)] = seed * r0;
      q_downward(Q, lmax, 0, x);
      const double sc = q0 / Q[at(0, 0, lmax)];
      for (int l = 0; l <= lmax; ++l) Q[at(l, 0, lmax)] *= sc;
    }
    if (mmax >= 1 && lmax >= 1) {
      const double r1 = q_ratio_cf(x, lmax, 1);
      Q[at(lmax - 1, 1, lmax)] = seed;
      Q[at(lmax, 1, lmax)] = seed * r1;
      q_downward(Q, lmax, 1, x);
      const double sc = (-1.0 / w) / Q[at(0, 1, lmax)];
      for (int l = 0; l <= lmax; ++l) Q[at(l, 1, lmax)] *= sc;
    }
  }
  q_m_upward(Q, lmax, mmax, x);
}

LegendreCache::LegendreCache(int Mabs_, int Lmax_) : Mabs(Mabs_), Lmax(Lmax_) {}

const std::vector<double> &LegendreCache::column(bool q, double x) {
  auto &tab = q ? Qtab : Ptab;
  auto it = tab.find(x);
  if (it != tab.end()) return it->second;
  std::vector<double> full((size_t)(Lmax + 1) * (Mabs + 1));
  if (q)
    legendre_q(full.data(), Lmax, Mabs, x);
  else
    legendre_p(full.data(), Lmax, Mabs, x);
  std::vector<double> col(Lmax + 1);
  for (int L = 0; L <= Lmax; ++L) {
    const double v = full[(size_t)L + (size_t)Mabs * (Lmax + 1)];
    col[L] = (v != 0.0 && !std::isnormal(v)) ? 0.0 : v;
  }
  return tab.emplace(x, std::move(col)).first->second;
}

}  // namespace hfg
