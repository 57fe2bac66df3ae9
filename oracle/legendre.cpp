// Oracle (test infrastructure only): associated Legendre functions of the
// first and second kinds, P_l^m(x) and Q_l^m(x), in HelFEM's convention
// (Hobson form outside the cut), restating src/legendre/Legendre.h.
// Pinned by tests/golden/legendre_reference.json (Maple values from
// src/legendre/legendre_reference_data.h) at the reference test's tolerance.
#include "orc.hpp"
#include <sstream>

namespace orc {
namespace {

inline size_t ix(int l, int m, int lmax) { return (size_t)m * (size_t)(lmax + 1) + (size_t)l; }
inline double sgnf(double x) { return std::abs(x) <= 1.0 ? 1.0 : -1.0; }
inline double sqclamp(double y) { return std::sqrt(std::max(y, 0.0)); }
inline double q00(double x) { return 0.5 * (std::log(std::abs(x + 1.0)) - std::log(std::abs(x - 1.0))); }

// Legendre.h:94-108 + :111-121
void plm_impl(double *P, int lmax, int mmax, double x) {
  std::fill_n(P, (size_t)(lmax + 1) * (mmax + 1), 0.0);
  const double s = sgnf(x), w = sqclamp(s * (1.0 - x * x));
  P[ix(0, 0, lmax)] = 1.0;
  for (int m = 1; m <= std::min(mmax, lmax); ++m) P[ix(m, m, lmax)] = -s * double(2 * m - 1) * w * P[ix(m - 1, m - 1, lmax)];
  for (int m = 0; m <= std::min(mmax, lmax - 1); ++m) P[ix(m + 1, m, lmax)] = double(2 * m + 1) * x * P[ix(m, m, lmax)];
  for (int m = 0; m <= std::min(mmax, lmax); ++m)
    for (int l = m + 1; l < lmax; ++l)
      P[ix(l + 1, m, lmax)] =
          (double(2 * l + 1) * x * P[ix(l, m, lmax)] - double(l + m) * P[ix(l - 1, m, lmax)]) / double(l + 1 - m);
}

// Legendre.h:134-157
void seed_q(double *Q, int lmax, int mmax, double x) {
  const double s = sgnf(x), w = sqclamp(s * (1.0 - x * x)), Q0 = q00(x);
  Q[ix(0, 0, lmax)] = Q0;
  if (lmax >= 1) Q[ix(1, 0, lmax)] = x * Q0 - 1.0;
  if (mmax >= 1) {
    if (w == 0.0) {
      Q[ix(0, 1, lmax)] = std::numeric_limits<double>::infinity();
      if (lmax >= 1) Q[ix(1, 1, lmax)] = std::numeric_limits<double>::infinity();
    } else {
      Q[ix(0, 1, lmax)] = -1.0 / w;
      if (lmax >= 1) Q[ix(1, 1, lmax)] = -s * w * (Q0 + x / (1.0 - x * x));
    }
  }
}

// Lentz-Thompson continued fraction for Q_lmax / Q_(lmax-1), Legendre.h:161-191
double cf_ratio(double x, int lmax, int m) {
  const double tiny = std::numeric_limits<double>::min() * 1e4;
  const double tol = 8.0 * std::numeric_limits<double>::epsilon();
  double f = tiny, C = f, D = 0.0, a = 1.0;
  int n = lmax;
  for (int it = 0; it < 1000000; ++it) {
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
    ++n;
  }
  throw std::runtime_error("Legendre Q continued fraction did not converge");
}

void q_up_l(double *Q, int lmax, int m, double x) {
  for (int l = 1; l < lmax; ++l)
    Q[ix(l + 1, m, lmax)] =
        (double(2 * l + 1) * x * Q[ix(l, m, lmax)] - double(l + m) * Q[ix(l - 1, m, lmax)]) / double(l + 1 - m);
}

void q_down_l(double *Q, int lmax, int m, double x) {
  for (int l = lmax - 1; l >= 1; --l)
    Q[ix(l - 1, m, lmax)] =
        (double(2 * l + 1) * x * Q[ix(l, m, lmax)] - double(l + 1 - m) * Q[ix(l + 1, m, lmax)]) / double(l + m);
}

void q_up_m(double *Q, int lmax, int mmax, double x) {
  if (mmax < 2) return;
  const double s = sgnf(x), w = sqclamp(s * (1.0 - x * x));
  if (w == 0.0) throw std::domain_error("Legendre Q: w = 0 in the m recurrence");
  for (int l = 0; l <= lmax; ++l)
    for (int m = 1; m < mmax; ++m)
      Q[ix(l, m + 1, lmax)] = -double(2 * m) * x / w * Q[ix(l, m, lmax)] -
                              s * double(l + m) * double(l - m + 1) * Q[ix(l, m - 1, lmax)];
}

// Christoffel/Bonnet decomposition Q_l = P_l Q_0 - W_l, Legendre.h:244-291
void q_christoffel(double *Q, int lmax, int mmax, double x) {
  std::vector<double> P(lmax + 1, 0.0), W(lmax + 1, 0.0), Pp(lmax + 1, 0.0), Wp(lmax + 1, 0.0);
  P[0] = 1.0;
  if (lmax >= 1) {
    P[1] = x;
    W[1] = 1.0;
    Pp[1] = 1.0;
  }
  for (int n = 1; n < lmax; ++n) {
    const double inv = 1.0 / double(n + 1), a = double(2 * n + 1), b = double(n);
    P[n + 1] = (a * x * P[n] - b * P[n - 1]) * inv;
    W[n + 1] = (a * x * W[n] - b * W[n - 1]) * inv;
    Pp[n + 1] = (a * (P[n] + x * Pp[n]) - b * Pp[n - 1]) * inv;
    Wp[n + 1] = (a * (W[n] + x * Wp[n]) - b * Wp[n - 1]) * inv;
  }
  const double Q0 = q00(x);
  for (int l = 0; l <= lmax; ++l) Q[ix(l, 0, lmax)] = P[l] * Q0 - W[l];
  if (mmax >= 1) {
    const double w = std::sqrt(std::max(x * x - 1.0, 0.0));
    Q[ix(0, 1, lmax)] = -1.0 / w;
    for (int l = 1; l <= lmax; ++l) Q[ix(l, 1, lmax)] = w * Pp[l] * Q0 - P[l] / w - w * Wp[l];
  }
  q_up_m(Q, lmax, mmax, x);
}

// Miller's downward recurrence normalised to the closed forms, Legendre.h:295-320
void q_miller(double *Q, int lmax, int mmax, double x) {
  const double seed = std::numeric_limits<double>::min() * 1e4;
  if (lmax >= 1) {
    const double r0 = cf_ratio(x, lmax, 0);
    Q[ix(lmax - 1, 0, lmax)] = seed;
    Q[ix(lmax, 0, lmax)] = seed * r0;
    q_down_l(Q, lmax, 0, x);
    const double sc = q00(x) / Q[ix(0, 0, lmax)];
    for (int l = 0; l <= lmax; ++l) Q[ix(l, 0, lmax)] *= sc;
  }
  if (mmax >= 1 && lmax >= 1) {
    const double r1 = cf_ratio(x, lmax, 1);
    Q[ix(lmax - 1, 1, lmax)] = seed;
    Q[ix(lmax, 1, lmax)] = seed * r1;
    q_down_l(Q, lmax, 1, x);
    const double w = sqclamp(sgnf(x) * (1.0 - x * x));
    const double sc = (-1.0 / w) / Q[ix(0, 1, lmax)];
    for (int l = 0; l <= lmax; ++l) Q[ix(l, 1, lmax)] *= sc;
  }
  q_up_m(Q, lmax, mmax, x);
}

}  // namespace

void legendre_plm(double *out, int lmax, int mmax, double x) {
  if (lmax < 0 || mmax < 0) throw std::invalid_argument("plm: negative order");
  plm_impl(out, lmax, mmax, x);
}

// Legendre.h:343-372
void legendre_qlm(double *out, int lmax, int mmax, double x) {
  if (lmax < 0 || mmax < 0) throw std::invalid_argument("qlm: negative order");
  if (std::abs(x) == 1.0) throw std::domain_error("qlm: singular at |x| = 1");
  std::fill_n(out, (size_t)(lmax + 1) * (mmax + 1), 0.0);
  seed_q(out, lmax, mmax, x);
  if (std::abs(x) <= 1.0) {
    if (lmax >= 2) q_up_l(out, lmax, 0, x);
    if (mmax >= 1 && lmax >= 2) q_up_l(out, lmax, 1, x);
    q_up_m(out, lmax, mmax, x);
    return;
  }
  const double xmax = lmax > 0 ? std::cosh(std::log(1e3) / double(2 * lmax)) : std::numeric_limits<double>::infinity();
  if (std::abs(x) < xmax)
    q_christoffel(out, lmax, mmax, x);
  else
    q_miller(out, lmax, mmax, x);
}

// MLegendreCache (src/diatomic/quadrature.h:47-84): one |M| column of a
// recurrence run to Lmax, with non-normal nonzero values flushed to zero.
LegendreColumn::LegendreColumn(int Mabs_, int Lmax_) : Mabs(Mabs_), Lmax(Lmax_) {}

const std::vector<double> &LegendreColumn::fetch(bool wantP, double x) {
  auto &tab = wantP ? Ptab : Qtab;
  auto it = tab.find(x);
  if (it != tab.end()) return it->second;
  std::vector<double> full((size_t)(Lmax + 1) * (Mabs + 1), 0.0);
  if (wantP)
    legendre_plm(full.data(), Lmax, Mabs, x);
  else
    legendre_qlm(full.data(), Lmax, Mabs, x);
  std::vector<double> col(Lmax + 1);
  for (int L = 0; L <= Lmax; ++L) {
    double v = full[(size_t)L + (size_t)Mabs * (Lmax + 1)];
    if (v != 0.0 && !std::isnormal(v)) v = 0.0;
    col[L] = v;
  }
  return tab.emplace(x, std::move(col)).first->second;
}

double LegendreColumn::P(int L, double x) { return fetch(true, x)[L]; }
double LegendreColumn::Q(int L, double x) { return x == 1.0 ? 0.0 : fetch(false, x)[L]; }

}  // namespace orc
