// Host setup of the range-separated exchange caches of the atomic basis
// (TwoDBasisT::compute_yukawa / compute_erfc, src/atomic/TwoDBasis.cpp:738-771).
//
// Yukawa, exp(-lambda d)/d = 4 pi lambda sum_L i_L(lambda r<) k_L(lambda r>)
// sum_M Y Y*: element integrals of i_L and k_L replace r^L and r^-L-1, and
// the in-element tensor uses the same cumulative rule as the Coulomb one
// with the kernel i_L k_L; all auto-converged like their Coulomb
// counterparts.
// erfc, erfc(mu d)/d = mu sum_L Phi_L(mu r, mu r') P_L(cos g): no
// factorisation across elements, so every element pair gets a tensor from
// the basis' fixed Chebyshev rule (RadialBasis.cpp:743-812).  Phi_L comes
// from the Gaussian representation of erf(mu d)/d and
// int_-1^1 exp(a t) P_L(t) dt = 2 i_L(a):
//   Phi_L(x, y) = (2L+1) 2/sqrt(pi) int_1^inf exp(-u^2 (x^2+y^2)) i_L(2 u^2 x y) du,
// evaluated for all L at once by adaptive Gauss-Kronrod quadrature, in the
// variable t = 1/u for large arguments and as the Coulomb term minus the
// complementary integral over u in [0, 1] for small ones.
#include "fem.hpp"
#include <cmath>
#include <limits>
#include <stdexcept>

namespace hfg {

double sph_bessel_i(double x, int L) {
  const double ax = std::fabs(x);
  double v;
  if (ax < 0.5) {
    // x^L / (2L+1)!! sum_k (x^2/2)^k / (k! prod_{j<=k} (2L+2j+1))
    double term = 1.0;
    for (int j = 1; j <= L; ++j) term *= ax / double(2 * j + 1);
    v = term;
    const double h = 0.5 * ax * ax;
    for (int k = 1; k < 256; ++k) {
      term *= h / (double(k) * double(2 * L + 2 * k + 1));
      v += term;
      if (std::fabs(term) <= 0.01 * std::numeric_limits<double>::epsilon() * std::fabs(v)) break;
    }
  } else {
    v = std::cyl_bessel_i(L + 0.5, ax) * std::sqrt(M_PI / (2.0 * ax));
  }
  return (x < 0.0 && (L & 1)) ? -v : v;
}

double sph_bessel_k(double x, int L) { return std::cyl_bessel_k(L + 0.5, x) * std::sqrt(2.0 / (M_PI * x)); }

namespace {

// exp(-z) i_L(z) without overflow (terminating expansion for large z).
double scaled_i(double z, int L) {
  if (z < 600.0) return std::exp(-z) * sph_bessel_i(z, L);
  double plus = 0.0, alt = 0.0, a = 1.0;
  for (int k = 0; k <= L; ++k) {
    if (k > 0) a *= double((L + k) * (L - k + 1)) / (double(k) * 2.0 * z);
    alt += (k & 1) ? -a : a;
    plus += a;
  }
  return (alt + ((L & 1) ? 1.0 : -1.0) * std::exp(-2.0 * z) * plus) / (2.0 * z);
}

struct GK15 {
  double x[8], wk[8], wg[4];
};
const GK15 gk = {{0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                  0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                  0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                  0.207784955007898467600689403773245, 0.0},
                 {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                  0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                  0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                  0.204432940075298892414161999234649, 0.209482141084727828012999174891714},
                 {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                  0.381830050505118944950369775488975, 0.417959183673469387755102040816327}};

template <typename F>
void adapt(const F &f, int n, double a, double b, double *acc, double atol, int depth) {
  double kr[64] = {0}, gr[64] = {0}, v[64];
  const double c = 0.5 * (a + b), h = 0.5 * (b - a);
  for (int j = 0; j < 8; ++j)
    for (int side = 0; side < (j == 7 ? 1 : 2); ++side) {
      f(side == 0 ? c - h * gk.x[j] : c + h * gk.x[j], v);
      for (int i = 0; i < n; ++i) {
        kr[i] += gk.wk[j] * v[i];
        if (j & 1) gr[i] += gk.wg[j / 2] * v[i];
      }
    }
  double err = 0.0, mag = 0.0;
  for (int i = 0; i < n; ++i) {
    err = std::max(err, std::fabs(h * (kr[i] - gr[i])));
    mag = std::max(mag, std::fabs(h * kr[i]));
  }
  if (err <= std::max(atol, 1e-15 * mag) || depth >= 40) {
    for (int i = 0; i < n; ++i) acc[i] += h * kr[i];
    return;
  }
  adapt(f, n, a, c, acc, 0.5 * atol, depth + 1);
  adapt(f, n, c, b, acc, 0.5 * atol, depth + 1);
}

}  // namespace

void erfc_multipoles(int Lmax, double x, double y, double *out) {
  const int n = Lmax + 1;
  if (n > 64) throw std::invalid_argument("erfc multipoles: L above 63");
  double acc[64] = {0};
  const double c2 = 2.0 / std::sqrt(M_PI);
  if (std::max(x, y) <= 1.0) {
    const double s2 = x * x + y * y, p = 2.0 * x * y;
    adapt([&](double u, double *v) {
            const double e = std::exp(-u * u * s2), z = u * u * p;
            for (int L = 0; L < n; ++L) v[L] = e * sph_bessel_i(z, L);
          },
          n, 0.0, 1.0, acc, 1e-17, 0);
    const double lo = std::min(x, y), hi = std::max(x, y);
    for (int L = 0; L < n; ++L)
      out[L] = (hi > 0.0 ? std::pow(lo, L) / std::pow(hi, L + 1) : 0.0) - (2 * L + 1) * c2 * acc[L];
    return;
  }
  const double d2 = (x - y) * (x - y), p = 2.0 * x * y;
  adapt([&](double t, double *v) {
          if (t <= 0.0) {
            for (int L = 0; L < n; ++L) v[L] = d2 > 0.0 ? 0.0 : 0.5 / p;
            return;
          }
          const double it2 = 1.0 / (t * t), e = std::exp(-d2 * it2), z = p * it2;
          for (int L = 0; L < n; ++L) v[L] = (p > 0.0 ? e * scaled_i(z, L) : (L == 0 ? e : 0.0)) * it2;
        },
        n, 0.0, 1.0, acc, 1e-17, 0);
  for (int L = 0; L < n; ++L) out[L] = (2 * L + 1) * c2 * acc[L];
}

Blk RadialFE::bessel_i(int L, double lambda, int iel) const {
  const std::function<double(double)> f = [L, lambda](double r) { return sph_bessel_i(r * lambda, L); };
  return matel(iel, 0, 0, &f);
}

Blk RadialFE::bessel_k(int L, double lambda, int iel) const {
  const std::function<double(double)> f = [L, lambda](double r) { return sph_bessel_k(r * lambda, L); };
  return matel(iel, 0, 0, &f);
}

Blk RadialFE::yukawa(int L, double lambda, int iel) const {
  const LipShape s = shape(iel);
  const double a = bval[iel], b = bval[iel + 1];
  const std::function<double(double, double)> fsb = [L, lambda](double r, double R) {
    return sph_bessel_i(r * lambda, L) * sph_bessel_k(R * lambda, L);
  };
  const std::function<double(double)> fbig = [L, lambda](double r) { return sph_bessel_k(r * lambda, L); };
  return refine(
      [&](int n) {
        std::vector<double> x, w;
        gauss_lobatto(n, x, w);
        return twoe_fixed_rule_k(a, b, x, w, s, fsb, fbig);
      },
      std::min(std::max(nquad, 5), 512), 512);
}

Blk RadialFE::erfc_pair(int L, double mu, int iel, int kel) const {
  std::vector<double> xq, wq;
  gauss_chebyshev2(nquad, xq, wq);
  const int nq = nquad, nint = iel == kel ? nq : 1;
  std::vector<double> xk((size_t)nq * nint), wk((size_t)nq * nint);
  for (int s = 0; s < nint; ++s) {
    const double st = s * 2.0 / nint - 1.0, en = (s + 1) * 2.0 / nint - 1.0;
    const double m = 0.5 * (en + st), l = 0.5 * (en - st);
    for (int i = 0; i < nq; ++i) {
      xk[(size_t)s * nq + i] = m + xq[i] * l;
      wk[(size_t)s * nq + i] = wq[i] * l;
    }
  }
  const Blk bi = values(0, xq, iel), bk = values(0, xk, kel);
  const double hi = half_len(iel), hk = half_len(kel), mi = mid(iel), mk = mid(kel);
  const int ni = bi.c, nk = bk.c, pi = (int)xq.size(), pk = (int)xk.size();
  Blk F(pi, pk);
  std::vector<double> tmp(L + 1);
#pragma omp parallel for schedule(dynamic) firstprivate(tmp)
  for (int i = 0; i < pi; ++i)
    for (int k = 0; k < pk; ++k) {
      erfc_multipoles(L, mu * (mi + hi * xq[i]), mu * (mk + hk * xk[k]), tmp.data());
      F(i, k) = tmp[L];
    }
  Blk ui(pi, ni * ni), uk(pk, nk * nk);
  for (int a = 0; a < ni; ++a)
    for (int b = 0; b < ni; ++b)
      for (int i = 0; i < pi; ++i) ui(i, a * ni + b) = bi(i, a) * bi(i, b) * (wq[i] * hi);
  for (int a = 0; a < nk; ++a)
    for (int b = 0; b < nk; ++b)
      for (int k = 0; k < pk; ++k) uk(k, a * nk + b) = bk(k, a) * bk(k, b) * (wk[k] * hk);
  const Blk t = blk_tn(ui, F);  // (ni^2 x pk)
  Blk out(ni * ni, nk * nk);
  for (int q = 0; q < nk * nk; ++q)
    for (int p = 0; p < ni * ni; ++p) {
      double sum = 0.0;
      for (int k = 0; k < pk; ++k) sum += t(p, k) * uk(k, q);
      out(p, q) = sum;
    }
  if (iel == kel) {
    Blk sym(out.r, out.c);
    for (int q = 0; q < out.c; ++q)
      for (int p = 0; p < out.r; ++p) sym(p, q) = 0.5 * (out(p, q) + out(q, p));
    out = sym;
  }
  return out;
}

void AtomicSetup::compute_yukawa(double lambda) {
  if (!(lambda > 0.0)) throw std::invalid_argument("Yukawa range parameter must be positive");
  const int nel = rad.Nel();
  rs_small.assign((size_t)N_L * nel, Blk());
  rs_big.assign((size_t)N_L * nel, Blk());
  rs_chol.assign((size_t)N_L * nel, Blk());
  rs_pair.clear();
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int L = 0; L < N_L; ++L)
    for (int e = 0; e < nel; ++e) {
      const size_t i = (size_t)L * nel + e;
      rs_small[i] = rad.bessel_i(L, lambda, e);
      if (e > 0) rs_big[i] = rad.bessel_k(L, lambda, e);
      rs_chol[i] = pivoted_cholesky(rad.yukawa(L, lambda, e), chol_tol);
    }
  rs_kind = 1;
  rs_param = lambda;
}

void AtomicSetup::compute_erfc(double mu) {
  if (!(mu > 0.0)) throw std::invalid_argument("erfc range parameter must be positive");
  const int nel = rad.Nel();
  rs_small.clear();
  rs_big.clear();
  rs_chol.clear();
  rs_pair.assign((size_t)N_L * nel * nel, Blk());
  for (int L = 0; L < N_L; ++L)
    for (int e = 0; e < nel; ++e)
      for (int f = 0; f < nel; ++f) rs_pair[((size_t)L * nel + e) * nel + f] = rad.erfc_pair(L, mu, e, f);
  rs_kind = 2;
  rs_param = mu;
}

}  // namespace hfg
