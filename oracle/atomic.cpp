// Oracle (test infrastructure only): atomic TwoDBasis Fock-build path.
#include "orc.hpp"
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// src/atomic/basis.cpp:179-203
void angular_basis(int lmax, int mmax, std::vector<int> &lval, std::vector<int> &mval) {
  lval.clear();
  mval.clear();
  for (int mabs = 0; mabs <= mmax; ++mabs)
    for (int l = mabs; l <= lmax; ++l) {
      lval.push_back(l);
      mval.push_back(mabs);
      if (mabs > 0) {
        lval.push_back(l);
        mval.push_back(-mabs);
      }
    }
}

AtomicBasis::AtomicBasis(int Z_, int nnodes, int nquad, const std::vector<double> &bval,
                         const std::vector<int> &lv, const std::vector<int> &mv)
    : Z(Z_), radial(nnodes, bval, nquad), lval(lv), mval(mv) {}

// TwoDBasis.cpp:709-735 with CoulombExchangeFE.h:180-217
void AtomicBasis::compute_tei() {
  const int nL = N_L(), Nel = radial.Nel();
  disjoint_L.assign((size_t)nL * Nel, Mat());
  disjoint_m1L.assign((size_t)nL * Nel, Mat());
  prim_chol.assign((size_t)nL * Nel, Mat());
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int L = 0; L < nL; ++L)
    for (int iel = 0; iel < Nel; ++iel) {
      disjoint_L[(size_t)L * Nel + iel] = radial.radial_integral(L, iel);
      if (iel > 0) disjoint_m1L[(size_t)L * Nel + iel] = radial.radial_integral(-L - 1, iel);
      prim_chol[(size_t)L * Nel + iel] = radial.twoe_integral_cholesky(L, iel, chol_tol);
    }
}

namespace {
Mat assemble_diag(const RadialBasis &rb, const std::function<Mat(int)> &k) {
  Mat M(rb.Nbf(), rb.Nbf());
  for (int iel = 0; iel < rb.Nel(); ++iel) {
    const Mat b = k(iel);
    const int f = rb.first[iel];
    for (int j = 0; j < b.c; ++j)
      for (int i = 0; i < b.r; ++i) M(f + i, f + j) += b(i, j);
  }
  return M;
}
void set_sub(Mat &M, int Nrad, int iang, int jang, const Mat &S, double fac = 1.0, bool add = false) {
  for (int j = 0; j < Nrad; ++j)
    for (int i = 0; i < Nrad; ++i) {
      double &t = M(iang * Nrad + i, jang * Nrad + j);
      t = add ? t + fac * S(i, j) : fac * S(i, j);
    }
}
}  // namespace

Mat AtomicBasis::overlap() const {
  const Mat O = assemble_diag(radial, [&](int iel) { return radial.radial_integral(0, iel); });
  Mat S(Nbf(), Nbf());
  for (int a = 0; a < Nang(); ++a) set_sub(S, Nrad(), a, a, O);
  return S;
}

Mat AtomicBasis::kinetic() const {
  const Mat T0 = assemble_diag(radial, [&](int iel) { return radial.kinetic(iel); });
  const Mat Tl = assemble_diag(radial, [&](int iel) { return radial.kinetic_l(iel); });
  Mat T(Nbf(), Nbf());
  for (int a = 0; a < Nang(); ++a) {
    set_sub(T, Nrad(), a, a, T0);
    if (lval[a] > 0) set_sub(T, Nrad(), a, a, Tl, double(lval[a] * (lval[a] + 1)), true);
  }
  return T;
}

Mat AtomicBasis::nuclear() const {
  Mat V(Nbf(), Nbf());
  if (Z == 0) return V;
  const Mat Vr = assemble_diag(radial, [&](int iel) { return radial.radial_integral(-1, iel); });
  for (int a = 0; a < Nang(); ++a) set_sub(V, Nrad(), a, a, Vr, double(-Z));
  return V;
}

// CoulombExchangeFE.h:432-482
Mat assemble_J_chol(const RadialBasis &radial, const std::vector<Mat> &rs, const std::vector<Mat> &rb,
                    const std::vector<Mat> &chol, int off, const Mat &P) {
  const int Nel = radial.Nel(), Nrad = radial.Nbf();
  Mat J(Nrad, Nrad);
  for (int jel = 0; jel < Nel; ++jel) {
    const int jf = radial.first[jel], Nj = radial.Nprim(jel);
    Mat Psub(Nj, Nj);
    for (int b = 0; b < Nj; ++b)
      for (int a = 0; a < Nj; ++a) Psub(a, b) = P(jf + a, jf + b);
    const Mat &rsj = rs[off + jel];
    double jsmall = 0.0, jbig = 0.0;
    for (int a = 0; a < Nj; ++a)
      for (int b = 0; b < Nj; ++b) jsmall += rsj(a, b) * Psub(b, a);
    if (jel > 0) {
      const Mat &rbj = rb[off + jel];
      for (int a = 0; a < Nj; ++a)
        for (int b = 0; b < Nj; ++b) jbig += rbj(a, b) * Psub(b, a);
    }
    for (int iel = 0; iel < jel; ++iel) {
      const int f = radial.first[iel], Ni = radial.Nprim(iel);
      const Mat &m = rs[off + iel];
      for (int b = 0; b < Ni; ++b)
        for (int a = 0; a < Ni; ++a) J(f + a, f + b) += jbig * m(a, b);
    }
    for (int iel = jel + 1; iel < Nel; ++iel) {
      const int f = radial.first[iel], Ni = radial.Nprim(iel);
      const Mat &m = rb[off + iel];
      for (int b = 0; b < Ni; ++b)
        for (int a = 0; a < Ni; ++a) J(f + a, f + b) += jsmall * m(a, b);
    }
    const Mat &Lc = chol[off + jel];
    std::vector<double> s(Lc.c, 0.0);
    for (int p = 0; p < Lc.c; ++p) {
      double t = 0.0;
      for (int k = 0; k < Nj * Nj; ++k) t += Lc(k, p) * Psub.a[k];
      s[p] = t;
    }
    for (int k = 0; k < Nj * Nj; ++k) {
      double t = 0.0;
      for (int p = 0; p < Lc.c; ++p) t += Lc(k, p) * s[p];
      J(jf + k % Nj, jf + k / Nj) += t;
    }
  }
  return J;
}

// CoulombExchangeFE.h:484-530
Mat assemble_K_chol(const RadialBasis &radial, const std::vector<Mat> &rs, const std::vector<Mat> &rb,
                    const std::vector<Mat> &chol, int off, const Mat &P) {
  const int Nel = radial.Nel(), Nrad = radial.Nbf();
  Mat K(Nrad, Nrad);
  for (int iel = 0; iel < Nel; ++iel) {
    const int fi = radial.first[iel], Ni = radial.Nprim(iel);
    for (int jel = 0; jel < Nel; ++jel) {
      const int fj = radial.first[jel], Nj = radial.Nprim(jel);
      Mat Psub(Ni, Nj);
      for (int b = 0; b < Nj; ++b)
        for (int a = 0; a < Ni; ++a) Psub(a, b) = P(fi + a, fj + b);
      Mat Ksub(Ni, Nj);
      if (iel == jel) {
        const Mat &Lc = chol[off + iel];
        for (int p = 0; p < Lc.c; ++p) {
          Mat Mp(Ni, Ni);
          for (int k = 0; k < Ni * Ni; ++k) Mp.a[k] = Lc(k, p);
          const Mat t = matmul(Mp, matmul_nt(Psub, Mp));
          for (size_t k = 0; k < t.a.size(); ++k) Ksub.a[k] += t.a[k];
        }
      } else {
        const Mat &ii = (iel > jel) ? rb[off + iel] : rs[off + iel];
        const Mat &jj = (iel > jel) ? rs[off + jel] : rb[off + jel];
        Ksub = matmul(ii, matmul_nt(Psub, jj));
      }
      for (int b = 0; b < Nj; ++b)
        for (int a = 0; a < Ni; ++a) K(fi + a, fj + b) += Ksub(a, b);
    }
  }
  return K;
}

// TwoDBasis.cpp:774-877
Mat AtomicBasis::coulomb(const Mat &P) const {
  if (prim_chol.empty()) throw std::logic_error("Primitive teis have not been computed!");
  const int Nel = radial.Nel(), Nrad = radial.Nbf(), Na = Nang();
  const int lmx = *std::max_element(lval.begin(), lval.end());
  const int gmax = std::max(lmx, *std::max_element(mval.begin(), mval.end()));
  const Gaunt gaunt(gmax, 2 * gmax, gmax);
  const int Mmax = *std::max_element(mval.begin(), mval.end()) - *std::min_element(mval.begin(), mval.end());
  const int nL = 2 * lmx + 1;
  std::vector<std::vector<Mat>> Paux(nL, std::vector<Mat>(2 * Mmax + 1));
  for (int L = 0; L < nL; ++L)
    for (int M = -std::min(L, Mmax); M <= std::min(L, Mmax); ++M) Paux[L][M + Mmax] = Mat(Nrad, Nrad);
  for (int k = 0; k < Na; ++k)
    for (int l = 0; l < Na; ++l) {
      const int lk = lval[k], mk = mval[k], ll = lval[l], ml = mval[l];
      const int M = mk - ml;
      const int Lmin = std::max(std::abs(lk - ll), std::abs(M)), Lmax = lk + ll;
      for (int L = Lmin; L <= Lmax; ++L) {
        const double cpl = gaunt.coeff(lk, mk, L, M, ll);
        Mat &X = Paux[L][M + Mmax];
        for (int j = 0; j < Nrad; ++j)
          for (int i = 0; i < Nrad; ++i) X(i, j) += cpl * P(k * Nrad + i, l * Nrad + j);
      }
    }
  std::vector<std::vector<Mat>> Jaux(nL, std::vector<Mat>(2 * Mmax + 1));
  for (int L = 0; L < nL; ++L) {
    const double Lfac = 4.0 * M_PI / double(2 * L + 1);
    for (int M = -std::min(L, Mmax); M <= std::min(L, Mmax); ++M) {
      Mat j = assemble_J_chol(radial, disjoint_L, disjoint_m1L, prim_chol, L * Nel, Paux[L][M + Mmax]);
      for (auto &v : j.a) v *= Lfac;
      Jaux[L][M + Mmax] = j;
    }
  }
  Mat J(Nbf(), Nbf());
  for (int i = 0; i < Na; ++i)
    for (int j = 0; j < Na; ++j) {
      const int li = lval[i], mi = mval[i], lj = lval[j], mj = mval[j];
      const int M = mj - mi;
      const int Lmin = std::max(std::abs(lj - li), std::abs(M)), Lmax = lj + li;
      for (int L = Lmin; L <= Lmax; ++L) {
        const double cpl = gaunt.coeff(lj, mj, L, M, li);
        if (cpl != 0.0) set_sub(J, Nrad, i, j, Jaux[L][M + Mmax], cpl, true);
      }
    }
  return J;
}

// TwoDBasis.cpp:880-999
Mat AtomicBasis::exchange(const Mat &P) const {
  std::vector<int> all;
  for (int q = 0; q < Nang() * Nang(); ++q) all.push_back(q);
  return exchange_blocks(P, all);
}

Mat AtomicBasis::exchange_blocks(const Mat &P, const std::vector<int> &jks) const {
  if (prim_chol.empty()) throw std::logic_error("Primitive teis have not been computed!");
  const int Nel = radial.Nel(), Nrad = radial.Nbf(), Na = Nang();
  const int lmx = *std::max_element(lval.begin(), lval.end());
  const int gmax = std::max(lmx, *std::max_element(mval.begin(), mval.end()));
  const Gaunt gaunt(gmax, 2 * gmax, gmax);
  const double bdens_thr = 10.0 * std::numeric_limits<double>::epsilon();
  const int nL = 2 * lmx + 1;
  // Frobenius norms of every radial block (the reference recomputes them
  // inside the (jang,kang) loop, TwoDBasis.cpp:950; the values are equal).
  std::vector<double> bnorm((size_t)Na * Na);
#pragma omp parallel for collapse(2) schedule(static)
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
  const int njk = (int)jks.size();
#pragma omp parallel for schedule(dynamic)
  for (int q = 0; q < njk; ++q) {
    {
      const int j = jks[q] / Na, k = jks[q] % Na;
      const int lj = lval[j], mj = mval[j], lk = lval[k], mk = mval[k];
      std::vector<Mat> R(nL);
      std::vector<bool> couple(nL, false);
      for (int L = 0; L < nL; ++L) R[L] = Mat(Nrad, Nrad);
      for (int i = 0; i < Na; ++i) {
        const int li = lval[i], mi = mval[i];
        for (int l = 0; l < Na; ++l) {
          const int ll = lval[l], ml = mval[l];
          const int M = mj - mi, Mp = mk - ml;
          if (M != Mp) continue;
          if (bnorm[(size_t)i * Na + l] < bdens_thr) continue;
          const int Lmin = std::max(std::max(std::abs(li - lj), std::abs(lk - ll)), std::abs(M));
          const int Lmax = std::min(li + lj, lk + ll);
          for (int L = Lmin; L <= Lmax; ++L) {
            const double cpl = gaunt.coeff(lj, mj, L, M, li) * gaunt.coeff(lk, mk, L, M, ll);
            if (cpl == 0.0) continue;
            const double Lfac = 4.0 * M_PI / double(2 * L + 1);
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
        const Mat t = assemble_K_chol(radial, disjoint_L, disjoint_m1L, prim_chol, L * Nel, R[L]);
        for (size_t u = 0; u < t.a.size(); ++u) Kb.a[u] += t.a[u];
      }
      for (int b = 0; b < Nrad; ++b)
        for (int a = 0; a < Nrad; ++a) K(j * Nrad + a, k * Nrad + b) -= Kb(a, b);
    }
  }
  return K;
}

}  // namespace orc
