// Oracle (test infrastructure only): Gaunt coefficients.
//
// The reference takes them from libwignernj v0.8.0 (fetched, absent here;
// CMakeLists.txt:288-306), whose published method is an exact
// prime-factorised rational rounded once at the end (gaunt.cpp:28-34).  This
// restatement evaluates the Racah closed form of the two 3j symbols with
// exact big-integer sums over a common prime-factorised denominator, and
// rounds once from binary128.  Pinned by the 550 golden doubles in
// src/general/gaunt_test.cpp (tests/golden/gaunt_reference.json).
#include "orc.hpp"
#include <quadmath.h>
#include <map>
#include <mutex>
#include <tuple>
#include <cstring>

namespace orc {
namespace {

struct BigInt {  // sign-magnitude, base 2^32
  bool neg = false;
  std::vector<uint32_t> d;
  bool zero() const { return d.empty(); }
  void trim() {
    while (!d.empty() && d.back() == 0) d.pop_back();
    if (d.empty()) neg = false;
  }
  static BigInt from(uint64_t v) {
    BigInt b;
    while (v) {
      b.d.push_back((uint32_t)v);
      v >>= 32;
    }
    return b;
  }
  void mul_small(uint32_t m) {
    uint64_t carry = 0;
    for (auto &w : d) {
      const uint64_t t = (uint64_t)w * m + carry;
      w = (uint32_t)t;
      carry = t >> 32;
    }
    if (carry) d.push_back((uint32_t)carry);
    trim();
  }
  static int cmp_mag(const BigInt &a, const BigInt &b) {
    if (a.d.size() != b.d.size()) return a.d.size() < b.d.size() ? -1 : 1;
    for (size_t i = a.d.size(); i-- > 0;)
      if (a.d[i] != b.d[i]) return a.d[i] < b.d[i] ? -1 : 1;
    return 0;
  }
  static BigInt add_mag(const BigInt &a, const BigInt &b) {
    BigInt r;
    const size_t n = std::max(a.d.size(), b.d.size());
    r.d.resize(n);
    uint64_t c = 0;
    for (size_t i = 0; i < n; ++i) {
      const uint64_t t = (uint64_t)(i < a.d.size() ? a.d[i] : 0) + (i < b.d.size() ? b.d[i] : 0) + c;
      r.d[i] = (uint32_t)t;
      c = t >> 32;
    }
    if (c) r.d.push_back((uint32_t)c);
    return r;
  }
  static BigInt sub_mag(const BigInt &a, const BigInt &b) {  // |a| >= |b|
    BigInt r;
    r.d.resize(a.d.size());
    int64_t br = 0;
    for (size_t i = 0; i < a.d.size(); ++i) {
      int64_t t = (int64_t)a.d[i] - (i < b.d.size() ? (int64_t)b.d[i] : 0) - br;
      br = t < 0;
      if (t < 0) t += ((int64_t)1 << 32);
      r.d[i] = (uint32_t)t;
    }
    r.trim();
    return r;
  }
  void add(const BigInt &o) {
    if (o.zero()) return;
    if (zero()) {
      *this = o;
      return;
    }
    if (neg == o.neg) {
      const bool s = neg;
      *this = add_mag(*this, o);
      neg = s;
    } else {
      const int c = cmp_mag(*this, o);
      if (c == 0) {
        d.clear();
        neg = false;
      } else if (c > 0) {
        const bool s = neg;
        *this = sub_mag(*this, o);
        neg = s;
      } else {
        const bool s = o.neg;
        *this = sub_mag(o, *this);
        neg = s;
      }
    }
    trim();
  }
  __float128 to_q() const {
    __float128 v = 0;
    for (size_t i = d.size(); i-- > 0;) v = v * (__float128)4294967296.0 + (__float128)d[i];
    return neg ? -v : v;
  }
};

std::vector<int> primes_upto(int n) {
  std::vector<int> p;
  std::vector<char> s(n + 1, 1);
  for (int i = 2; i <= n; ++i)
    if (s[i]) {
      p.push_back(i);
      for (long j = (long)i * i; j <= n; j += i) s[j] = 0;
    }
  return p;
}

const int NMAXF = 1200;
const std::vector<int> &primes() {
  static const std::vector<int> p = primes_upto(NMAXF);
  return p;
}

// exponent of each prime in n!  (Legendre's formula)
std::vector<int> fact_exp(int n) {
  const auto &P = primes();
  std::vector<int> e(P.size(), 0);
  for (size_t i = 0; i < P.size() && P[i] <= n; ++i) {
    long pk = P[i];
    while (pk <= n) {
      e[i] += n / pk;
      pk *= P[i];
    }
  }
  return e;
}

const std::vector<int> &fexp(int n) {
  static std::vector<std::vector<int>> cache;
  static std::mutex mtx;
  std::lock_guard<std::mutex> lk(mtx);
  if (n < 0 || n > NMAXF) throw std::logic_error("factorial out of table range");
  if (cache.empty()) cache.resize(NMAXF + 1);
  if (cache[n].empty()) cache[n] = fact_exp(n);
  return cache[n];
}

// One 3j symbol as (sign, integer part N, exponent vector E with 2*E
// integers):  value = N * prod p^(E/2).  Racah formula.
struct Exact3j {
  bool zero = true;
  BigInt N;
  std::vector<int> twoE;  // twice the prime exponents
};

Exact3j threej_exact(int j1, int j2, int j3, int m1, int m2, int m3) {
  Exact3j r;
  const auto &P = primes();
  r.twoE.assign(P.size(), 0);
  if (m1 + m2 + m3 != 0) return r;
  if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return r;
  if (j3 < std::abs(j1 - j2) || j3 > j1 + j2) return r;
  // sqrt( Delta * prod (j+-m)! ):  Delta = (j1+j2-j3)!(j1-j2+j3)!(-j1+j2+j3)!/(j1+j2+j3+1)!
  auto addf = [&](int n, int s) {
    const auto &e = fexp(n);
    for (size_t i = 0; i < P.size(); ++i) r.twoE[i] += s * e[i];
  };
  addf(j1 + j2 - j3, 1);
  addf(j1 - j2 + j3, 1);
  addf(-j1 + j2 + j3, 1);
  addf(j1 + j2 + j3 + 1, -1);
  addf(j1 + m1, 1);
  addf(j1 - m1, 1);
  addf(j2 + m2, 1);
  addf(j2 - m2, 1);
  addf(j3 + m3, 1);
  addf(j3 - m3, 1);
  // The sum: S = sum_k (-1)^k / D_k.
  const int kmin = std::max(0, std::max(j2 - j3 - m1, j1 - j3 + m2));
  const int kmax = std::min(j1 + j2 - j3, std::min(j1 - m1, j2 + m2));
  if (kmin > kmax) return r;
  std::vector<std::vector<int>> Dk;
  std::vector<int> Dmax(P.size(), 0);
  for (int k = kmin; k <= kmax; ++k) {
    std::vector<int> e(P.size(), 0);
    const int args[6] = {k, j3 - j2 + k + m1, j3 - j1 + k - m2, j1 + j2 - j3 - k, j1 - k - m1, j2 - k + m2};
    for (int a : args) {
      const auto &f = fexp(a);
      for (size_t i = 0; i < P.size(); ++i) e[i] += f[i];
    }
    for (size_t i = 0; i < P.size(); ++i) Dmax[i] = std::max(Dmax[i], e[i]);
    Dk.push_back(e);
  }
  BigInt S;
  for (int k = kmin; k <= kmax; ++k) {
    const auto &e = Dk[k - kmin];
    BigInt t = BigInt::from(1);
    for (size_t i = 0; i < P.size(); ++i)
      for (int q = 0; q < Dmax[i] - e[i]; ++q) t.mul_small((uint32_t)P[i]);
    if (k & 1) t.neg = true;
    S.add(t);
  }
  if (S.zero()) return r;
  // overall phase (-1)^(j1-j2-m3)
  const int ph = j1 - j2 - m3;
  if (ph & 1) S.neg = !S.neg;
  for (size_t i = 0; i < P.size(); ++i) r.twoE[i] -= 2 * Dmax[i];
  r.N = S;
  r.zero = false;
  return r;
}

// int Y_l1^m1 Y_l2^m2 Y_l3^m3 dOmega, rounded once from binary128.
double gaunt3(int l1, int m1, int l2, int m2, int l3, int m3) {
  if (m1 + m2 + m3 != 0) return 0.0;
  if ((l1 + l2 + l3) & 1) return 0.0;
  const Exact3j a = threej_exact(l1, l2, l3, 0, 0, 0);
  if (a.zero) return 0.0;
  const Exact3j b = threej_exact(l1, l2, l3, m1, m2, m3);
  if (b.zero) return 0.0;
  const auto &P = primes();
  std::vector<int> te(P.size());
  for (size_t i = 0; i < P.size(); ++i) te[i] = a.twoE[i] + b.twoE[i];
  // (2l1+1)(2l2+1)(2l3+1) under the square root
  for (int f : {2 * l1 + 1, 2 * l2 + 1, 2 * l3 + 1}) {
    int v = f;
    for (size_t i = 0; i < P.size() && v > 1; ++i)
      while (v % P[i] == 0) {
        te[i] += 1;
        v /= P[i];
      }
  }
  // value = Na*Nb * prod p^(te/2) / sqrt(4 pi)
  __float128 num = 1, den = 1;
  for (size_t i = 0; i < P.size(); ++i) {
    const int full = te[i] / 2;  // C++ truncation toward zero
    const int half = te[i] - 2 * full;
    __float128 pq = P[i];
    if (full > 0)
      for (int q = 0; q < full; ++q) num *= pq;
    else
      for (int q = 0; q < -full; ++q) den *= pq;
    if (half > 0) num *= sqrtq(pq);
    if (half < 0) den *= sqrtq(pq);
  }
  __float128 v = a.N.to_q() * b.N.to_q() * num / den;
  v /= sqrtq(4 * acosq((__float128)-1));
  return (double)v;
}

}  // namespace

// gaunt.cpp:87-93: int Y_L^M* Y_l^m Y_lp^mp = (-1)^M gaunt(L,-M,l,m,lp,mp).
// Exact values are memoised: the reference rebuilds its table on every
// coulomb/exchange call (TwoDBasis.cpp:787-788) from a fast library; here the
// exact evaluation is paid once per distinct argument tuple.
double gaunt_coefficient(int L, int M, int l, int m, int lp, int mp) {
  if (M - m - mp != 0 || ((L + l + lp) & 1) || lp < std::abs(L - l) || lp > L + l) return 0.0;
  static std::mutex mtx;
  static std::map<std::tuple<int, int, int, int, int, int>, double> memo;
  const auto key = std::make_tuple(L, M, l, m, lp, mp);
  {
    std::lock_guard<std::mutex> g(mtx);
    auto it = memo.find(key);
    if (it != memo.end()) return it->second;
  }
  const double sign = (M & 1) ? -1.0 : 1.0;
  const double v = sign * gaunt3(L, -M, l, m, lp, mp);
  std::lock_guard<std::mutex> g(mtx);
  memo[key] = v;
  return v;
}

// gaunt.cpp:99-112
double modified_gaunt_coefficient(int lj, int mj, int L, int M, int li, int mi) {
  const double const0 = 2.0 / 3.0 * std::sqrt(M_PI);
  const double const2 = 4.0 / 15.0 * std::sqrt(5.0 * M_PI);
  const double cpl0 = gaunt_coefficient(L, M, 0, 0, L, M) * gaunt_coefficient(lj, mj, li, mi, L, M);
  double cpl2 = 0.0;
  for (int Lp = std::max(std::max(L - 2, 0), std::abs(M)); Lp <= L + 2; ++Lp)
    cpl2 += gaunt_coefficient(Lp, M, 2, 0, L, M) * gaunt_coefficient(lj, mj, li, mi, Lp, M);
  return const0 * cpl0 + const2 * cpl2;
}

Gaunt::Gaunt(int Lmax_, int lmax_, int lpmax_) : Lmax(Lmax_), lmax(lmax_), lpmax(lpmax_) {
  const size_t nL = Lmax + 1, nM = 2 * Lmax + 1, nl = lmax + 1, nm = 2 * lmax + 1, nlp = lpmax + 1;
  tab.assign(nL * nM * nl * nm * nlp, 0.0);
  for (int L = 0; L <= Lmax; ++L)
    for (int M = -L; M <= L; ++M)
      for (int l = 0; l <= lmax; ++l)
        for (int m = -l; m <= l; ++m)
          for (int lp = 0; lp <= lpmax; ++lp) {
            const int mp = M - m;
            if (std::abs(mp) > lp) continue;
            const size_t idx = ((((size_t)L * nM + (M + Lmax)) * nl + l) * nm + (m + lmax)) * nlp + lp;
            tab[idx] = gaunt_coefficient(L, M, l, m, lp, mp);
          }
}

// gaunt.cpp:202-250 semantics (out-of-range probes are zero)
double Gaunt::coeff(int L, int M, int l, int m, int lp) const {
  if (L < 0 || L > Lmax || l < 0 || l > lmax || lp < 0 || lp > lpmax) return 0.0;
  if (std::abs(M) > L || std::abs(m) > l) return 0.0;
  if (std::abs(M - m) > lp) return 0.0;
  const size_t nM = 2 * Lmax + 1, nl = lmax + 1, nm = 2 * lmax + 1, nlp = lpmax + 1;
  return tab[((((size_t)L * nM + (M + Lmax)) * nl + l) * nm + (m + lmax)) * nlp + lp];
}

// gaunt.cpp:254-270
double Gaunt::mod_coeff(int lj, int mj, int L, int M, int li, int mi) const {
  if (mj != M + mi) return 0.0;
  const double const0 = 2.0 / 3.0 * std::sqrt(M_PI);
  const double const2 = 4.0 / 15.0 * std::sqrt(5.0 * M_PI);
  const double cpl0 = coeff(L, M, 0, 0, L) * coeff(lj, mj, li, mi, L);
  double cpl2 = 0.0;
  for (int Lp = std::max(std::max(L - 2, 0), std::abs(M)); Lp <= L + 2; ++Lp)
    cpl2 += coeff(Lp, M, 2, 0, L) * coeff(lj, mj, li, mi, Lp);
  return const0 * cpl0 + const2 * cpl2;
}

}  // namespace orc
