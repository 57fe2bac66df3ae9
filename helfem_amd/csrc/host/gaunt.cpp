// Gaunt coefficients for the angular couplings of J and K.
//
// int Y_L^M* Y_l^m Y_lp^mp dOmega (src/general/gaunt.cpp:87-93).  The value is
// (-1)^M sqrt((2L+1)(2l+1)(2lp+1)/4pi) (L l lp;0 0 0)(L l lp;-M m mp), with
// each 3j symbol taken from the Racah sum evaluated exactly: the terms share
// a common denominator expressed by prime exponents, their numerators are
// summed as exact integers, and the result is rounded once from binary128.
// This is the same exact-rational approach as libwignernj (the reference's
// dependency, v0.8.0), so the doubles agree with it to the last bit in
// practice.
#include "fem.hpp"
#include <quadmath.h>
#include <mutex>
#include <map>
#include <array>

namespace hfg {
namespace {

// Unsigned magnitude in base 2^32 plus a sign flag.
struct Wide {
  bool neg = false;
  std::vector<uint32_t> w;
  bool is_zero() const { return w.empty(); }
  void norm() {
    while (!w.empty() && !w.back()) w.pop_back();
    if (w.empty()) neg = false;
  }
  void mul(uint32_t m) {
    uint64_t c = 0;
    for (auto &x : w) {
      uint64_t t = (uint64_t)x * m + c;
      x = (uint32_t)t;
      c = t >> 32;
    }
    if (c) w.push_back((uint32_t)c);
  }
  static int cmpmag(const Wide &a, const Wide &b) {
    if (a.w.size() != b.w.size()) return a.w.size() < b.w.size() ? -1 : 1;
    for (size_t i = a.w.size(); i-- > 0;)
      if (a.w[i] != b.w[i]) return a.w[i] < b.w[i] ? -1 : 1;
    return 0;
  }
  // |a| - |b| with |a| >= |b|
  static std::vector<uint32_t> sub(const std::vector<uint32_t> &a, const std::vector<uint32_t> &b) {
    std::vector<uint32_t> r(a.size());
    int64_t borrow = 0;
    for (size_t i = 0; i < a.size(); ++i) {
      int64_t t = (int64_t)a[i] - borrow - (i < b.size() ? (int64_t)b[i] : 0);
      borrow = t < 0 ? 1 : 0;
      r[i] = (uint32_t)(t + (borrow << 32));
    }
    return r;
  }
  static std::vector<uint32_t> add(const std::vector<uint32_t> &a, const std::vector<uint32_t> &b) {
    std::vector<uint32_t> r(std::max(a.size(), b.size()) + 1, 0);
    uint64_t c = 0;
    for (size_t i = 0; i + 1 < r.size(); ++i) {
      uint64_t t = c + (i < a.size() ? a[i] : 0) + (i < b.size() ? b[i] : 0);
      r[i] = (uint32_t)t;
      c = t >> 32;
    }
    r.back() = (uint32_t)c;
    return r;
  }
  void accumulate(const Wide &o) {
    if (o.is_zero()) return;
    if (is_zero()) {
      *this = o;
      return;
    }
    if (neg == o.neg) {
      w = add(w, o.w);
    } else {
      const int c = cmpmag(*this, o);
      if (c >= 0) {
        w = sub(w, o.w);
      } else {
        w = sub(o.w, w);
        neg = o.neg;
      }
    }
    norm();
  }
  __float128 value() const {
    __float128 v = 0;
    for (size_t i = w.size(); i-- > 0;) v = v * (__float128)4294967296.0 + (__float128)w[i];
    return neg ? -v : v;
  }
};

constexpr int kMaxFact = 1024;

struct PrimeTable {
  std::vector<int> p;
  std::vector<std::vector<int>> fe;  // fe[n][i]: exponent of p[i] in n!
  PrimeTable() {
    std::vector<char> comp(kMaxFact + 1, 0);
    for (int i = 2; i <= kMaxFact; ++i)
      if (!comp[i]) {
        p.push_back(i);
        for (int j = 2 * i; j <= kMaxFact; j += i) comp[j] = 1;
      }
    fe.assign(kMaxFact + 1, std::vector<int>(p.size(), 0));
    for (int n = 2; n <= kMaxFact; ++n) {
      fe[n] = fe[n - 1];
      int v = n;
      for (size_t i = 0; i < p.size() && v > 1; ++i)
        while (v % p[i] == 0) {
          ++fe[n][i];
          v /= p[i];
        }
    }
  }
};

const PrimeTable &primes() {
  static const PrimeTable t;
  return t;
}

// 3j symbol = S * prod_i p_i^(e2[i]/2), with S an exact signed integer.
bool racah_3j(int j1, int j2, int j3, int m1, int m2, int m3, Wide &S, std::vector<int> &e2) {
  const PrimeTable &T = primes();
  const size_t np = T.p.size();
  e2.assign(np, 0);
  if (m1 + m2 + m3 != 0 || std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return false;
  if (j3 < std::abs(j1 - j2) || j3 > j1 + j2) return false;
  if (j1 + j2 + j3 + 1 > kMaxFact) throw std::logic_error("Gaunt: angular momentum too large");
  const int sq_pos[9] = {j1 + j2 - j3, j1 - j2 + j3, -j1 + j2 + j3, j1 + m1, j1 - m1, j2 + m2, j2 - m2, j3 + m3, j3 - m3};
  for (int n : sq_pos)
    for (size_t i = 0; i < np; ++i) e2[i] += T.fe[n][i];
  for (size_t i = 0; i < np; ++i) e2[i] -= T.fe[j1 + j2 + j3 + 1][i];
  const int k0 = std::max({0, j2 - j3 - m1, j1 - j3 + m2});
  const int k1 = std::min({j1 + j2 - j3, j1 - m1, j2 + m2});
  if (k0 > k1) return false;
  std::vector<std::vector<int>> den;
  std::vector<int> lcm(np, 0);
  for (int k = k0; k <= k1; ++k) {
    std::vector<int> d(np, 0);
    const int a[6] = {k, j3 - j2 + k + m1, j3 - j1 + k - m2, j1 + j2 - j3 - k, j1 - k - m1, j2 - k + m2};
    for (int n : a)
      for (size_t i = 0; i < np; ++i) d[i] += T.fe[n][i];
    for (size_t i = 0; i < np; ++i) lcm[i] = std::max(lcm[i], d[i]);
    den.push_back(std::move(d));
  }
  S = Wide();
  for (int k = k0; k <= k1; ++k) {
    Wide t;
    t.w = {1u};
    const auto &d = den[k - k0];
    for (size_t i = 0; i < np; ++i)
      for (int q = d[i]; q < lcm[i]; ++q) t.mul((uint32_t)T.p[i]);
    t.neg = (k & 1) != 0;
    S.accumulate(t);
  }
  if (S.is_zero()) return false;
  if ((j1 - j2 - m3) & 1) S.neg = !S.neg;
  for (size_t i = 0; i < np; ++i) e2[i] -= 2 * lcm[i];
  return true;
}

double gaunt_triple(int l1, int m1, int l2, int m2, int l3, int m3) {
  if (m1 + m2 + m3 != 0 || ((l1 + l2 + l3) & 1)) return 0.0;
  Wide s0, s1;
  std::vector<int> e0, e1;
  if (!racah_3j(l1, l2, l3, 0, 0, 0, s0, e0)) return 0.0;
  if (!racah_3j(l1, l2, l3, m1, m2, m3, s1, e1)) return 0.0;
  const PrimeTable &T = primes();
  for (size_t i = 0; i < T.p.size(); ++i) e0[i] += e1[i];
  for (int f : {2 * l1 + 1, 2 * l2 + 1, 2 * l3 + 1})
    for (size_t i = 0; i < T.p.size() && f > 1; ++i)
      while (f % T.p[i] == 0) {
        ++e0[i];
        f /= T.p[i];
      }
  __float128 num = 1, den = 1;
  for (size_t i = 0; i < T.p.size(); ++i) {
    int e = e0[i];
    if (!e) continue;
    const __float128 pr = T.p[i];
    const int whole = e / 2, odd = e % 2;
    for (int q = 0; q < std::abs(whole); ++q) (whole > 0 ? num : den) *= pr;
    if (odd) (odd > 0 ? num : den) *= sqrtq(pr);
  }
  const __float128 v = s0.value() * s1.value() * num / den / sqrtq(4 * acosq((__float128)-1));
  return (double)v;
}

}  // namespace

double gaunt_coeff(int L, int M, int l, int m, int lp, int mp) {
  static std::mutex mtx;
  static std::map<std::array<int, 6>, double> memo;
  const std::array<int, 6> key = {L, M, l, m, lp, mp};
  {
    std::lock_guard<std::mutex> g(mtx);
    auto it = memo.find(key);
    if (it != memo.end()) return it->second;
  }
  const double v = ((M & 1) ? -1.0 : 1.0) * gaunt_triple(L, -M, l, m, lp, mp);
  std::lock_guard<std::mutex> g(mtx);
  memo[key] = v;
  return v;
}

}  // namespace hfg
