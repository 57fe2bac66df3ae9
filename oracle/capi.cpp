// Oracle (test infrastructure only): C entry points for the Python tests.
// Loaded by tests/ and by bench.py's cpu_baseline leg through ctypes.
#include "orc.hpp"
#include "diatomic.hpp"
#include "dftgrid.hpp"
#include <cstring>
#include <omp.h>
#include <stdexcept>
#include <string>

using namespace orc;

namespace {
thread_local std::string g_err;
template <typename F>
int guard(F f) {
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}
void put(const Mat &m, double *out) { std::memcpy(out, m.data(), sizeof(double) * m.size()); }
Mat get(const double *p, int r, int c) {
  Mat m(r, c);
  std::memcpy(m.data(), p, sizeof(double) * m.size());
  return m;
}
}  // namespace

extern "C" {

const char *orc_last_error() { return g_err.c_str(); }

int orc_lobatto(int n, double *x, double *w) {
  return guard([&] {
    std::vector<double> xx, ww;
    lobatto(n, xx, ww);
    std::memcpy(x, xx.data(), sizeof(double) * n);
    std::memcpy(w, ww.data(), sizeof(double) * n);
  });
}

int orc_chebyshev(int n, double *x, double *w) {
  return guard([&] {
    std::vector<double> xx, ww;
    chebyshev(n, xx, ww);
    std::memcpy(x, xx.data(), sizeof(double) * n);
    std::memcpy(w, ww.data(), sizeof(double) * n);
  });
}

int orc_grid(double rmax, int nel, int igrid, double zexp, double *out) {
  return guard([&] {
    const auto b = get_grid(rmax, nel, igrid, zexp);
    std::memcpy(out, b.data(), sizeof(double) * b.size());
  });
}

double orc_gaunt(int L, int M, int l, int m, int lp, int mp) { return gaunt_coefficient(L, M, l, m, lp, mp); }
double orc_mod_gaunt(int lj, int mj, int L, int M, int li, int mi) {
  return modified_gaunt_coefficient(lj, mj, L, M, li, mi);
}

// LIP primitives on Gauss-Lobatto nodes: out is (npts x nnodes)
int orc_lip_eval(int nnodes, int npts, const double *x, int n, double *out) {
  return guard([&] {
    LIPBasis p(nnodes);
    const Mat m = LIPBasis::eval_prim(std::vector<double>(x, x + npts), p.x0, n);
    put(m, out);
  });
}

// In-element radial 2e tensor at a fixed Lobatto order (no convergence loop)
int orc_twoe_rule(double rmin, double rmax, int nquad, int nnodes, int drop_first, int drop_last, int L,
                  double *out) {
  return guard([&] {
    LIPBasis p(nnodes);
    if (drop_first) p.drop_first();
    if (drop_last) p.drop_last();
    std::vector<double> x, w;
    lobatto(nquad, x, w);
    put(twoe_integral_rule(rmin, rmax, x, w, p, L), out);
  });
}

void *orc_atomic_create(int Z, int nnodes, int nquad, int nbval, const double *bval, int nang, const int *lval,
                        const int *mval) {
  try {
    return new AtomicBasis(Z, nnodes, nquad, std::vector<double>(bval, bval + nbval),
                           std::vector<int>(lval, lval + nang), std::vector<int>(mval, mval + nang));
  } catch (const std::exception &e) {
    g_err = e.what();
    return nullptr;
  }
}

void orc_atomic_destroy(void *h) { delete (AtomicBasis *)h; }

// out: Nbf, Nrad, Nel, N_L, Nang
int orc_atomic_sizes(void *h, long *out) {
  AtomicBasis *b = (AtomicBasis *)h;
  out[0] = b->Nbf();
  out[1] = b->Nrad();
  out[2] = b->radial.Nel();
  out[3] = b->N_L();
  out[4] = b->Nang();
  return 0;
}

// first index and size of every element
int orc_atomic_elements(void *h, int *first, int *nprim) {
  AtomicBasis *b = (AtomicBasis *)h;
  for (int e = 0; e < b->radial.Nel(); ++e) {
    first[e] = b->radial.first[e];
    nprim[e] = b->radial.Nprim(e);
  }
  return 0;
}

int orc_atomic_compute_tei(void *h) { return guard([&] { ((AtomicBasis *)h)->compute_tei(); }); }

// which: 0 overlap, 1 kinetic, 2 nuclear
int orc_atomic_matrix(void *h, int which, double *out) {
  return guard([&] {
    AtomicBasis *b = (AtomicBasis *)h;
    Mat m = which == 0 ? b->overlap() : which == 1 ? b->kinetic() : b->nuclear();
    put(m, out);
  });
}

int orc_atomic_coulomb(void *h, const double *P, double *J) {
  return guard([&] {
    AtomicBasis *b = (AtomicBasis *)h;
    put(b->coulomb(get(P, b->Nbf(), b->Nbf())), J);
  });
}

int orc_atomic_exchange(void *h, const double *P, double *K) {
  return guard([&] {
    AtomicBasis *b = (AtomicBasis *)h;
    put(b->exchange(get(P, b->Nbf(), b->Nbf())), K);
  });
}

// Range-separated exchange: caches, then -K_sr (TwoDBasis.cpp:738-771, 1002-1131)
int orc_atomic_compute_yukawa(void *h, double lambda) {
  return guard([&] { ((AtomicBasis *)h)->compute_yukawa(lambda); });
}
int orc_atomic_compute_erfc(void *h, double mu) {
  return guard([&] { ((AtomicBasis *)h)->compute_erfc(mu); });
}
int orc_atomic_rs_exchange(void *h, const double *P, double *K) {
  return guard([&] {
    AtomicBasis *b = (AtomicBasis *)h;
    put(b->rs_exchange(get(P, b->Nbf(), b->Nbf())), K);
  });
}
// Range-separated cache read-back: kind 0 i_L element integral, 1 k_L, 2
// Yukawa Cholesky factor (L*Nel + iel), 3 erfc exchange-ordered tensor
// (L*Nel*Nel + iel*Nel + jel).  dims = (rows, cols); data may be NULL.
int orc_atomic_rs_cache(void *h, int kind, int index, int *dims, double *data) {
  return guard([&] {
    AtomicBasis *b = (AtomicBasis *)h;
    const std::vector<Mat> *v = kind == 0 ? &b->disjoint_iL : kind == 1 ? &b->disjoint_kL
                                : kind == 2 ? &b->rs_chol : &b->rs_ktei;
    if (index < 0 || index >= (int)v->size()) throw std::invalid_argument("cache index out of range");
    const Mat &M = (*v)[index];
    dims[0] = M.r;
    dims[1] = M.c;
    if (data) put(M, data);
  });
}
// Special functions: i_L, k_L and the erfc multipole kernel Phi_0..Lmax
double orc_bessel_il(double x, int L) { return bessel_il(x, L); }
double orc_bessel_kl(double x, int L) { return bessel_kl(x, L); }
int orc_erfc_phi(int Lmax, double x, double y, double *out) {
  return guard([&] { erfc_phi(Lmax, x, y, out); });
}

int orc_atomic_exchange_blocks(void *h, const double *P, double *K, int nblocks, const int *jk) {
  return guard([&] {
    AtomicBasis *b = (AtomicBasis *)h;
    put(b->exchange_blocks(get(P, b->Nbf(), b->Nbf()), std::vector<int>(jk, jk + nblocks)), K);
  });
}

int orc_legendre(int lmax, int mmax, double x, double *P, double *Q) {
  return guard([&] {
    legendre_plm(P, lmax, mmax, x);
    legendre_qlm(Q, lmax, mmax, x);
  });
}

// Cached factor export.  kind: 0 disjoint_L (r^L), 1 disjoint_m1L (r^-L-1),
// 2 prim_chol.  dims receives (rows, cols); out may be null to query dims.
int orc_atomic_cache(void *h, int kind, int L, int iel, int *dims, double *out) {
  return guard([&] {
    AtomicBasis *b = (AtomicBasis *)h;
    const size_t idx = (size_t)L * b->radial.Nel() + iel;
    const std::vector<Mat> &v = kind == 0 ? b->disjoint_L : kind == 1 ? b->disjoint_m1L : b->prim_chol;
    if (idx >= v.size()) throw std::logic_error("cache index out of range");
    dims[0] = v[idx].r;
    dims[1] = v[idx].c;
    if (out) put(v[idx], out);
  });
}

int orc_atomic_twoe(void *h, int L, int iel, double *out) {
  return guard([&] { put(((AtomicBasis *)h)->radial.twoe_integral(L, iel), out); });
}

int orc_atomic_radial_integral(void *h, int n, int iel, double *out) {
  return guard([&] { put(((AtomicBasis *)h)->radial.radial_integral(n, iel), out); });
}

// Radial basis values R = B/r at arbitrary element coordinates x in [-1,1]
int orc_atomic_radial_bf(void *h, int iel, int npts, const double *x, double *out) {
  return guard([&] {
    AtomicBasis *b = (AtomicBasis *)h;
    put(b->radial.get_bf(std::vector<double>(x, x + npts), iel), out);
  });
}

// ---- diatomic ----
void *orc_diatomic_create(int Z1, int Z2, double Rhalf, int nnodes, int nquad, int nbval, const double *bval,
                          int nang, const int *lval, const int *mval) {
  try {
    return new DiatomicBasis(Z1, Z2, Rhalf, nnodes, nquad, std::vector<double>(bval, bval + nbval),
                             std::vector<int>(lval, lval + nang), std::vector<int>(mval, mval + nang));
  } catch (const std::exception &e) {
    g_err = e.what();
    return nullptr;
  }
}
void orc_diatomic_destroy(void *h) { delete (DiatomicBasis *)h; }
// out: Nbf, Ndummy, Nrad, Nang, Nel, n_lm, n_LM
int orc_diatomic_sizes(void *h, long *out) {
  DiatomicBasis *b = (DiatomicBasis *)h;
  out[0] = b->Nbf();
  out[1] = b->Ndummy();
  out[2] = b->Nrad();
  out[3] = b->Nang();
  out[4] = b->radial.Nel();
  out[5] = (long)b->lm_map.size();
  out[6] = (long)b->LM_map.size();
  return 0;
}
int orc_diatomic_pure_idx(void *h, int *out) {
  DiatomicBasis *b = (DiatomicBasis *)h;
  for (size_t i = 0; i < b->pure_idx.size(); ++i) out[i] = b->pure_idx[i];
  return 0;
}
int orc_diatomic_compute_tei(void *h) { return guard([&] { ((DiatomicBasis *)h)->compute_tei(); }); }
int orc_diatomic_matrix(void *h, int which, double *out) {
  return guard([&] {
    DiatomicBasis *b = (DiatomicBasis *)h;
    put(which == 0 ? b->overlap() : which == 1 ? b->kinetic() : b->nuclear(), out);
  });
}
int orc_diatomic_coulomb(void *h, const double *P, double *J) {
  return guard([&] {
    DiatomicBasis *b = (DiatomicBasis *)h;
    put(b->coulomb(get(P, b->Nbf(), b->Nbf())), J);
  });
}
int orc_diatomic_exchange(void *h, const double *P, double *K) {
  return guard([&] {
    DiatomicBasis *b = (DiatomicBasis *)h;
    put(b->exchange(get(P, b->Nbf(), b->Nbf())), K);
  });
}
int orc_diatomic_exchange_blocks(void *h, const double *P, double *K, int nblocks, const int *jk) {
  return guard([&] {
    DiatomicBasis *b = (DiatomicBasis *)h;
    put(b->exchange_blocks(get(P, b->Nbf(), b->Nbf()), std::vector<int>(jk, jk + nblocks)), K);
  });
}
// LDA functional values at n densities
int orc_lda(int xf, int cf, int n, const double *rho, double thr, double *exc, double *vrho) {
  return guard([&] {
    for (int i = 0; i < n; ++i) lda_unpolarized(xf, cf, rho[i], thr, exc[i], vrho[i]);
  });
}
int orc_lda_pol(int xf, int cf, int n, const double *ra, const double *rb, double thr, double *exc, double *va,
                double *vb) {
  return guard([&] {
    for (int i = 0; i < n; ++i) lda_polarized(xf, cf, ra[i], rb[i], thr, exc[i], va[i], vb[i]);
  });
}
// Any supported functional pair at n points: unpolarized (rho, sigma, tau) ->
// (exc, vrho, vsigma, vtau) when pol = 0; polarized rho (2n: a, b), sigma
// (3n: aa, ab, bb), tau (2n) -> exc, vrho (2n), vsigma (3n), vtau (2n) when
// pol = 1; tau and vtau may be NULL (zero kinetic energy density)
int orc_xc_exact_exchange(int xf, double *out) {
  return guard([&] {
    int kernel = 0;
    xc_exact_exchange(xf, out[0], out[1], out[2], kernel);
    out[3] = kernel;
  });
}

int orc_xc_eval(int xf, int cf, int pol, int n, const double *rho, const double *sigma, const double *tau, double thr,
                double *exc, double *vrho, double *vsigma, double *vtau) {
  return guard([&] {
    for (int i = 0; i < n; ++i) {
      if (pol) {
        const double r[2] = {rho[i], rho[n + i]}, s[3] = {sigma[i], sigma[n + i], sigma[2 * n + i]};
        const double t[2] = {tau ? tau[i] : 0.0, tau ? tau[n + i] : 0.0};
        double vr[2], vs[3], vt[2];
        xc_polarized(xf, cf, r, s, t, thr, exc[i], vr, vs, vt);
        if (vtau) {
          vtau[i] = vt[0];
          vtau[n + i] = vt[1];
        }
        vrho[i] = vr[0];
        vrho[n + i] = vr[1];
        for (int k = 0; k < 3; ++k) vsigma[k * n + i] = vs[k];
      } else {
        double vt;
        xc_unpolarized(xf, cf, rho[i], sigma[i], tau ? tau[i] : 0.0, thr, exc[i], vrho[i], vsigma[i], vt);
        if (vtau) vtau[i] = vt;
      }
    }
  });
}
// Atomic DFT grid: H (Nbf x Nbf) and out = (Exc, Nel, Ekin = integral of tau, meta-GGA)
int orc_atomic_xc(void *h, int lang, int mang, int xf, int cf, const double *P, double thr, double *H, double *out) {
  return guard([&] {
    AtomicBasis *b = (AtomicBasis *)h;
    AtomicXC g(*b, lang, mang);
    Mat Hm;
    g.eval_Fxc(xf, cf, get(P, b->Nbf(), b->Nbf()), Hm, out[0], out[1], thr);
    out[2] = g.ekin;
    put(Hm, H);
  });
}
// Diatomic pure-m DFT grid: H (Nbf x Nbf, pure) and out = (Exc, Nel)
int orc_diatomic_xc(void *h, int lang, int xf, int cf, const double *P, double thr, double *H, double *out) {
  return guard([&] {
    DiatomicBasis *b = (DiatomicBasis *)h;
    DiatomicXC g(*b, lang);
    Mat Hm;
    g.eval_Fxc(xf, cf, get(P, b->Nbf(), b->Nbf()), Hm, out[0], out[1], thr);
    out[2] = g.ekin;
    put(Hm, H);
  });
}
// Spin-polarized grids: Ha, Hb and out = (Exc, Nel)
int orc_atomic_xc_pol(void *h, int lang, int mang, int xf, int cf, const double *Pa, const double *Pb, double thr,
                      double *Ha, double *Hb, double *out) {
  return guard([&] {
    AtomicBasis *b = (AtomicBasis *)h;
    AtomicXC g(*b, lang, mang);
    Mat A, B;
    g.eval_Fxc(xf, cf, get(Pa, b->Nbf(), b->Nbf()), get(Pb, b->Nbf(), b->Nbf()), A, B, out[0], out[1], thr);
    out[2] = g.ekin;
    put(A, Ha);
    put(B, Hb);
  });
}
int orc_diatomic_xc_pol(void *h, int lang, int xf, int cf, const double *Pa, const double *Pb, double thr, double *Ha,
                        double *Hb, double *out) {
  return guard([&] {
    DiatomicBasis *b = (DiatomicBasis *)h;
    DiatomicXC g(*b, lang);
    Mat A, B;
    g.eval_Fxc(xf, cf, get(Pa, b->Nbf(), b->Nbf()), get(Pb, b->Nbf(), b->Nbf()), A, B, out[0], out[1], thr);
    out[2] = g.ekin;
    put(A, Ha);
    put(B, Hb);
  });
}
// General 3D diatomic grid (lang x mang): restricted and spin-polarized
int orc_diatomic_xc3d(void *h, int lang, int mang, int xf, int cf, const double *P, double thr, double *H,
                      double *out) {
  return guard([&] {
    DiatomicBasis *b = (DiatomicBasis *)h;
    if (mang < 1) throw std::invalid_argument("mang must be positive");
    DiatomicXC g(*b, lang, mang);
    Mat Hm;
    g.eval_Fxc(xf, cf, get(P, b->Nbf(), b->Nbf()), Hm, out[0], out[1], thr);
    out[2] = g.ekin;
    put(Hm, H);
  });
}
int orc_diatomic_xc3d_pol(void *h, int lang, int mang, int xf, int cf, const double *Pa, const double *Pb,
                          double thr, double *Ha, double *Hb, double *out) {
  return guard([&] {
    DiatomicBasis *b = (DiatomicBasis *)h;
    if (mang < 1) throw std::invalid_argument("mang must be positive");
    DiatomicXC g(*b, lang, mang);
    Mat A, B;
    g.eval_Fxc(xf, cf, get(Pa, b->Nbf(), b->Nbf()), get(Pb, b->Nbf(), b->Nbf()), A, B, out[0], out[1], thr);
    out[2] = g.ekin;
    put(A, Ha);
    put(B, Hb);
  });
}
// lm map entry i: out = (L, |M|); LM map similarly (L, M)
int orc_diatomic_maps(void *h, int *lm, int *LM) {
  DiatomicBasis *b = (DiatomicBasis *)h;
  for (size_t i = 0; i < b->lm_map.size(); ++i) {
    lm[2 * i] = b->lm_map[i].first;
    lm[2 * i + 1] = b->lm_map[i].second;
  }
  for (size_t i = 0; i < b->LM_map.size(); ++i) {
    LM[2 * i] = b->LM_map[i].first;
    LM[2 * i + 1] = b->LM_map[i].second;
  }
  return 0;
}
// kind 0 P0, 1 P2, 2 Q0, 3 Q2, 4 cd_B, 5 cd_sigma (as a column)
int orc_diatomic_cache(void *h, int kind, int ilm, int iel, int *dims, double *out) {
  return guard([&] {
    DiatomicBasis *b = (DiatomicBasis *)h;
    const size_t i = (size_t)ilm * b->radial.Nel() + iel;
    if (i >= b->cd_B.size()) throw std::logic_error("cache index out of range");
    if (kind == 5) {
      dims[0] = (int)b->cd_sigma[i].size();
      dims[1] = 1;
      if (out) std::memcpy(out, b->cd_sigma[i].data(), sizeof(double) * b->cd_sigma[i].size());
      return;
    }
    const Mat &m = kind == 0 ? b->P0[i] : kind == 1 ? b->P2[i] : kind == 2 ? b->Q0[i] : kind == 3 ? b->Q2[i] : b->cd_B[i];
    dims[0] = m.r;
    dims[1] = m.c;
    if (out) put(m, out);
  });
}

// Thread count of the oracle's OpenMP loops (the CPU baseline is timed at
// all host threads and at one).
int orc_set_threads(int n) {
  if (n < 1) return -1;
  omp_set_num_threads(n);
  return 0;
}

}  // extern "C"
