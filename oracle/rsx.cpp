// Oracle (test infrastructure only): range-separated exchange of the atomic
// basis, TwoDBasisT::rs_exchange (src/atomic/TwoDBasis.cpp:1002-1131), with
// the Yukawa caches of compute_yukawa (:738-760) and the erfc caches of
// compute_erfc (:762-771).
//
// Yukawa, exp(-lambda r12)/r12 = 4 pi lambda sum_L i_L(lambda r<) k_L(lambda r>)
//   sum_M Y Y*: the bare-Coulomb machinery with r^L -> i_L(lambda r),
//   r^-L-1 -> k_L(lambda r) (RadialBasis.cpp:321-330, quadrature.cpp:164-198)
//   and the L factor 4 pi lambda.
// erfc, erfc(mu r12)/r12 = mu sum_L Phi_L(mu r, mu r') P_L(cos gamma): no
//   factorisation, so every element pair keeps a dense exchange-ordered
//   tensor from a fixed Chebyshev rule (RadialBasis.cpp:743-812,
//   quadrature.cpp:203-245); L factor 4 pi mu / (2L + 1).
//
// Phi_L is evaluated from the Gaussian representation
//   erf(mu d)/d = 2/sqrt(pi) int_0^mu exp(-s^2 d^2) ds,
//   int_-1^1 exp(a t) P_L(t) dt = 2 i_L(a),
// which gives, with x = mu r, y = mu r',
//   Phi_L(x, y) = x<^L / x>^(L+1) - (2L+1) 2/sqrt(pi) int_0^1 exp(-u^2 (x^2+y^2)) i_L(2 u^2 x y) du
//               = (2L+1) 2/sqrt(pi) int_0^1 exp(-(x-y)^2/t^2) [exp(-z) i_L(z)] dt / t^2,  z = 2xy/t^2.
// The first form is used for small arguments (no cancellation there), the
// second (no cancellation at all) otherwise; both integrands are smooth and
// are integrated by adaptive Gauss-Kronrod (7/15) for all L at once.  The
// reference evaluates the same function by the Angyan-Gerber-Marsman series
// (erfc_expn.cpp); this restatement is pinned against mpmath.
#include "orc.hpp"
#include <cmath>
#include <functional>
#include <limits>
#include <stdexcept>

namespace orc {

// Modified spherical Bessel functions, normalised as
//   i_L(x) = sqrt(pi/(2x)) I_{L+1/2}(x),  k_L(x) = sqrt(2/(pi x)) K_{L+1/2}(x)
// (libhelfem/include/math.h:52-115): Taylor series below |x| = 0.5, the
// standard library's cylinder functions above.
double bessel_il(double x, int L) {
  const double ax = std::fabs(x);
  double v;
  if (ax < 0.5) {
    double df = 1.0;
    for (int j = 3; j <= 2 * L + 1; j += 2) df *= j;
    double term = std::pow(ax, L) / df;
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

double bessel_kl(double x, int L) { return std::cyl_bessel_k(L + 0.5, x) * std::sqrt(2.0 / (M_PI * x)); }

namespace {

// exp(-z) i_L(z) for z >= 0 without overflow: direct below z = 600, the
// terminating expansion i_L(z) = [e^z S(-z) - (-1)^L e^-z S(z)] / (2z),
// S(z) = sum_k (L+k)! / (k! (L-k)! (2z)^k) ... with alternating signs, above
// (where z >> L^2 and nothing cancels).
double scaled_il(double z, int L) {
  if (z < 600.0) return std::exp(-z) * bessel_il(z, L);
  double sp = 0.0, sm = 0.0, a = 1.0;  // a_k = (L+k)! / (k! (L-k)! (2z)^k)
  for (int k = 0; k <= L; ++k) {
    if (k > 0) a *= double((L + k) * (L - k + 1)) / (double(k) * 2.0 * z);
    sp += (k & 1) ? -a : a;
    sm += a;
  }
  const double sgn = (L & 1) ? 1.0 : -1.0;  // (-1)^(L+1)
  return (sp + sgn * std::exp(-2.0 * z) * sm) / (2.0 * z);
}

// Gauss-Kronrod 7/15 nodes and weights on [-1, 1].
const double gk_x[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                        0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                        0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                        0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
const double gk_wk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                         0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                         0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                         0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
const double gk_wg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                         0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

// Adaptive GK15 of a vector integrand f(t, out[0..n)) over [a, b].
void gk_vec(const std::function<void(double, double *)> &f, int n, double a, double b, double *res, double atol,
            int depth) {
  std::vector<double> kr(n, 0.0), gr(n, 0.0), v(n);
  const double c = 0.5 * (a + b), h = 0.5 * (b - a);
  for (int j = 0; j < 8; ++j) {
    const int npts = j == 7 ? 1 : 2;
    for (int side = 0; side < npts; ++side) {
      const double t = side == 0 ? c - h * gk_x[j] : c + h * gk_x[j];
      f(t, v.data());
      for (int i = 0; i < n; ++i) {
        kr[i] += gk_wk[j] * v[i];
        if (j & 1) gr[i] += gk_wg[j / 2] * v[i];
      }
    }
  }
  double err = 0.0, mag = 0.0;
  for (int i = 0; i < n; ++i) {
    kr[i] *= h;
    gr[i] *= h;
    err = std::max(err, std::fabs(kr[i] - gr[i]));
    mag = std::max(mag, std::fabs(kr[i]));
  }
  if (err <= std::max(atol, 1e-15 * mag) || depth >= 40) {
    for (int i = 0; i < n; ++i) res[i] += kr[i];
    return;
  }
  gk_vec(f, n, a, c, res, 0.5 * atol, depth + 1);
  gk_vec(f, n, c, b, res, 0.5 * atol, depth + 1);
}

}  // namespace

// Phi_L(x, y) for L = 0..Lmax into out[0..Lmax].
void erfc_phi(int Lmax, double x, double y, double *out) {
  const int n = Lmax + 1;
  std::vector<double> acc(n, 0.0);
  const double c = 2.0 / std::sqrt(M_PI);
  if (std::max(x, y) <= 1.0) {
    const double s2 = x * x + y * y, p = 2.0 * x * y;
    gk_vec([&](double u, double *v) {
             const double e = std::exp(-u * u * s2), z = u * u * p;
             for (int L = 0; L < n; ++L) v[L] = e * bessel_il(z, L);
           },
           n, 0.0, 1.0, acc.data(), 1e-17, 0);
    const double lo = std::min(x, y), hi = std::max(x, y);
    for (int L = 0; L < n; ++L) {
      const double coul = hi > 0.0 ? std::pow(lo, L) / std::pow(hi, L + 1) : 0.0;
      out[L] = coul - (2 * L + 1) * c * acc[L];
    }
  } else {
    const double d2 = (x - y) * (x - y), p = 2.0 * x * y;
    gk_vec([&](double t, double *v) {
             if (t <= 0.0) {
               // limit t -> 0: exp(-d^2/t^2) -> 0 unless x = y, where
               // exp(-z) i_L(z) / t^2 -> 1 / (2 p)
               for (int L = 0; L < n; ++L) v[L] = d2 > 0.0 ? 0.0 : 0.5 / p;
               return;
             }
             const double it2 = 1.0 / (t * t), e = std::exp(-d2 * it2), z = p * it2;
             for (int L = 0; L < n; ++L) v[L] = (p > 0.0 ? e * scaled_il(z, L) : (L == 0 ? e : 0.0)) * it2;
           },
           n, 0.0, 1.0, acc.data(), 1e-17, 0);
    for (int L = 0; L < n; ++L) out[L] = (2 * L + 1) * c * acc[L];
  }
}

Mat RadialBasis::bessel_il_integral(int L, double lambda, int iel) const {
  const std::function<double(double)> f = [L, lambda](double r) { return bessel_il(r * lambda, L); };
  return matrix_element_auto(iel, 0, 0, &f);
}

Mat RadialBasis::bessel_kl_integral(int L, double lambda, int iel) const {
  const std::function<double(double)> f = [L, lambda](double r) { return bessel_kl(r * lambda, L); };
  return matrix_element_auto(iel, 0, 0, &f);
}

// RadialBasis.cpp:727-741 + quadrature.cpp:164-198: the twoe rule with the
// kernel i_L(lambda r) k_L(lambda R), auto-converged like twoe_integral.
Mat RadialBasis::yukawa_integral(int L, double lambda, int iel) const {
  const std::function<double(double, double)> fsb = [L, lambda](double r, double R) {
    return bessel_il(r * lambda, L) * bessel_kl(R * lambda, L);
  };
  const std::function<double(double)> fbig = [L, lambda](double r) { return bessel_kl(r * lambda, L); };
  return twoe_integral_conv(iel, fsb, fbig);
}

// RadialBasis.cpp:743-812 + quadrature.cpp:203-245: the fixed Chebyshev rule
// of the basis (Nq points) on element iel, and on element kel Nq points in
// each of Nint = Nq (kel == iel) or 1 equal sub-intervals; Coulomb-ordered
// (ab, cd) tensor, symmetrised when iel == kel.
Mat RadialBasis::erfc_integral(int L, double mu, int iel, int kel) const {
  const int Nq = (int)xq.size(), Nint = iel == kel ? Nq : 1;
  std::vector<double> xk((size_t)Nq * Nint), wk((size_t)Nq * Nint);
  for (int s = 0; s < Nint; ++s) {
    const double st = s * 2.0 / Nint - 1.0, en = (s + 1) * 2.0 / Nint - 1.0;
    const double mid = 0.5 * (en + st), len = 0.5 * (en - st);
    for (int i = 0; i < Nq; ++i) {
      xk[(size_t)s * Nq + i] = mid + xq[i] * len;
      wk[(size_t)s * Nq + i] = wq[i] * len;
    }
  }
  const Mat bi = eval_dnf(xq, 0, iel), bk = eval_dnf(xk, 0, kel);
  const std::vector<double> ri = eval_coord(xq, iel), rk = eval_coord(xk, kel);
  const double hi = scaling_factor(iel), hk = scaling_factor(kel);
  const int ni = bi.c, nk = bk.c, npi = (int)ri.size(), npk = (int)rk.size();
  Mat Fn(npi, npk);
  std::vector<double> tmp(L + 1);
  for (int i = 0; i < npi; ++i)
    for (int k = 0; k < npk; ++k) {
      erfc_phi(L, mu * ri[i], mu * rk[k], tmp.data());
      Fn(i, k) = tmp[L];
    }
  Mat pi_(npi, ni * ni), pk(npk, nk * nk);
  for (int a = 0; a < ni; ++a)
    for (int b = 0; b < ni; ++b)
      for (int i = 0; i < npi; ++i) pi_(i, a * ni + b) = bi(i, a) * bi(i, b) * (wq[i] * hi);
  for (int a = 0; a < nk; ++a)
    for (int b = 0; b < nk; ++b)
      for (int k = 0; k < npk; ++k) pk(k, a * nk + b) = bk(k, a) * bk(k, b) * (wk[k] * hk);
  Mat out = matmul(matmul_tn(pi_, Fn), pk);
  if (iel == kel) {
    Mat s(out.r, out.c);
    for (int j = 0; j < out.c; ++j)
      for (int i = 0; i < out.r; ++i) s(i, j) = 0.5 * (out(i, j) + out(j, i));
    out = s;
  }
  return out;
}

// utils::exchange_tei: (ij|kl) -> (jk|il) ordering, i.e. the exchange
// operator K(a c, b d) = T(a b, c d) applied to vec(P) with P(b, d).
Mat exchange_order(const Mat &T, int Ni, int Nk) {
  // T rows a*Ni + b (element i pair), cols c*Nk + d (element k pair);
  // out rows a + Ni*c (K sub-block entry), cols b + Ni*d (P sub-block entry)
  Mat out(Ni * Nk, Ni * Nk);
  for (int a = 0; a < Ni; ++a)
    for (int b = 0; b < Ni; ++b)
      for (int c = 0; c < Nk; ++c)
        for (int d = 0; d < Nk; ++d) out(a + Ni * c, b + Ni * d) = T(a * Ni + b, c * Nk + d);
  return out;
}

void AtomicBasis::compute_yukawa(double lambda_) {
  lambda = lambda_;
  yukawa = true;
  const int nL = N_L(), Nel = radial.Nel();
  disjoint_iL.assign((size_t)nL * Nel, Mat());
  disjoint_kL.assign((size_t)nL * Nel, Mat());
  rs_chol.assign((size_t)nL * Nel, Mat());
  rs_ktei.clear();
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int L = 0; L < nL; ++L)
    for (int iel = 0; iel < Nel; ++iel) {
      disjoint_iL[(size_t)L * Nel + iel] = radial.bessel_il_integral(L, lambda, iel);
      if (iel > 0) disjoint_kL[(size_t)L * Nel + iel] = radial.bessel_kl_integral(L, lambda, iel);
      rs_chol[(size_t)L * Nel + iel] = pivoted_cholesky(radial.yukawa_integral(L, lambda, iel), chol_tol);
    }
}

void AtomicBasis::compute_erfc(double mu) {
  lambda = mu;
  yukawa = false;
  disjoint_iL.clear();
  disjoint_kL.clear();
  rs_chol.clear();
  const int nL = N_L(), Nel = radial.Nel();
  rs_ktei.assign((size_t)nL * Nel * Nel, Mat());
#pragma omp parallel for collapse(3) schedule(dynamic)
  for (int L = 0; L < nL; ++L)
    for (int iel = 0; iel < Nel; ++iel)
      for (int jel = 0; jel < Nel; ++jel)
        rs_ktei[(size_t)L * Nel * Nel + (size_t)iel * Nel + jel] =
            exchange_order(radial.erfc_integral(L, mu, iel, jel), radial.Nprim(iel), radial.Nprim(jel));
}

// CoulombExchangeFE.h:412-440: K_ij += reshape(ktei(i,j) vec(P_ij)).
Mat assemble_K_pairwise(const RadialBasis &radial, const std::vector<Mat> &ktei, size_t off, const Mat &P) {
  const int Nel = radial.Nel(), Nrad = radial.Nbf();
  Mat K(Nrad, Nrad);
  for (int iel = 0; iel < Nel; ++iel) {
    const int fi = radial.first[iel], Ni = radial.Nprim(iel);
    for (int jel = 0; jel < Nel; ++jel) {
      const int fj = radial.first[jel], Nj = radial.Nprim(jel);
      const Mat &T = ktei[off + (size_t)iel * Nel + jel];
      for (int b = 0; b < Nj; ++b)
        for (int a = 0; a < Ni; ++a) {
          double s = 0.0;
          for (int d = 0; d < Nj; ++d)
            for (int c = 0; c < Ni; ++c) s += T(a + Ni * b, c + Ni * d) * P(fi + c, fj + d);
          K(fi + a, fj + b) += s;
        }
    }
  }
  return K;
}

// TwoDBasis.cpp:1002-1131 (returns -K_sr)
Mat AtomicBasis::rs_exchange(const Mat &P) const {
  if (rs_chol.empty() && rs_ktei.empty()) throw std::logic_error("Primitive teis have not been computed!");
  const int Nel = radial.Nel(), Nrad = radial.Nbf(), Na = Nang();
  const int lmx = *std::max_element(lval.begin(), lval.end());
  const int gmax = std::max(lmx, *std::max_element(mval.begin(), mval.end()));
  const Gaunt gaunt(gmax, 2 * gmax, gmax);
  const double bdens_thr = 10.0 * std::numeric_limits<double>::epsilon();
  const int nL = 2 * lmx + 1;
  std::vector<double> bnorm((size_t)Na * Na);
  for (int i = 0; i < Na; ++i)
    for (int l = 0; l < Na; ++l) {
      double s = 0.0;
      for (int b = 0; b < Nrad; ++b)
        for (int a = 0; a < Nrad; ++a) {
          const double v = P(i * Nrad + a, l * Nrad + b);
          s += v * v;
        }
      bnorm[(size_t)i * Na + l] = std::sqrt(s);
    }
  Mat K(Nbf(), Nbf());
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int j = 0; j < Na; ++j)
    for (int k = 0; k < Na; ++k) {
      const int lj = lval[j], mj = mval[j], lk = lval[k], mk = mval[k];
      std::vector<Mat> R(nL);
      std::vector<bool> couple(nL, false);
      for (int L = 0; L < nL; ++L) R[L] = Mat(Nrad, Nrad);
      for (int i = 0; i < Na; ++i) {
        const int li = lval[i], mi = mval[i];
        for (int l = 0; l < Na; ++l) {
          const int ll = lval[l], ml = mval[l];
          const int M = mj - mi;
          if (M != mk - ml) continue;
          if (bnorm[(size_t)i * Na + l] < bdens_thr) continue;
          const int Lmin = std::max(std::max(std::abs(li - lj), std::abs(lk - ll)), std::abs(M));
          const int Lmax = std::min(li + lj, lk + ll);
          for (int L = Lmin; L <= Lmax; ++L) {
            const double cpl = gaunt.coeff(lj, mj, L, M, li) * gaunt.coeff(lk, mk, L, M, ll);
            if (cpl == 0.0) continue;
            const double Lfac = yukawa ? 4.0 * M_PI * lambda : 4.0 * M_PI * lambda / double(2 * L + 1);
            const double f = Lfac * cpl;
            Mat &X = R[L];
            for (int b = 0; b < Nrad; ++b)
              for (int a = 0; a < Nrad; ++a) X(a, b) += f * P(i * Nrad + a, l * Nrad + b);
            couple[L] = true;
          }
        }
      }
      Mat Kb(Nrad, Nrad);
      for (int L = 0; L < nL; ++L) {
        if (!couple[L]) continue;
        const Mat t = yukawa ? assemble_K_chol(radial, disjoint_iL, disjoint_kL, rs_chol, L * Nel, R[L])
                             : assemble_K_pairwise(radial, rs_ktei, (size_t)L * Nel * Nel, R[L]);
        for (size_t u = 0; u < t.a.size(); ++u) Kb.a[u] += t.a[u];
      }
      for (int b = 0; b < Nrad; ++b)
        for (int a = 0; a < Nrad; ++a) K(j * Nrad + a, k * Nrad + b) -= Kb(a, b);
    }
  return K;
}

}  // namespace orc
