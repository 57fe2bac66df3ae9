// C-ABI layer (include/hfgpu.h): handles, table construction, error
// translation.  Host code only; the kernels live in atomic_jk.hip.
#include "capi_common.hpp"
#include <cstring>
#include <cmath>
#include <memory>

using namespace hfg;

namespace hfg {
thread_local std::string g_err;
}

namespace {

// Pack per-(L, element) factors as zero-padded 16x16 tiles: the "small"
// (r^L or i_L) and "big" (r^-L-1 or k_L, empty on element 0) element
// matrices, and the in-element Cholesky vectors M_p(a,b) = L(a + b n, p).
void upload_factor_tiles(const AtomicSetup &S, const std::vector<Blk> &small, const std::vector<Blk> &big,
                         const std::vector<Blk> &chol, DevArray<double> &d_rs, DevArray<double> &d_rb,
                         DevArray<double> &d_chol, DevArray<int> &d_off, DevArray<int> &d_rank, int &maxr,
                         hipStream_t s) {
  const int Nel = S.rad.Nel(), nL = S.N_L;
  std::vector<double> rs((size_t)nL * Nel * TILE2, 0.0), rb((size_t)nL * Nel * TILE2, 0.0);
  std::vector<double> ch;
  std::vector<int> choff((size_t)nL * Nel), chrank((size_t)nL * Nel);
  maxr = 0;
  for (int L = 0; L < nL; ++L)
    for (int e = 0; e < Nel; ++e) {
      const size_t q = (size_t)L * Nel + e;
      const int n = S.rad.ni[e];
      const Blk &a = small[q];
      need(a.r == n && a.c == n, "small-r factor cache has the wrong shape");
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) rs[q * TILE2 + i + TILE * j] = a(i, j);
      if (e > 0) {
        const Blk &b = big[q];
        need(b.r == n && b.c == n, "large-r factor cache has the wrong shape");
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i) rb[q * TILE2 + i + TILE * j] = b(i, j);
      }
      const Blk &c = chol[q];
      need(c.r == n * n, "Cholesky cache has the wrong number of rows");
      choff[q] = (int)ch.size();
      chrank[q] = c.c;
      maxr = std::max(maxr, c.c);
      need(c.c <= TILE2, "Cholesky rank above 256 is not supported");
      const size_t base = ch.size();
      ch.resize(base + (size_t)c.c * TILE2, 0.0);
      for (int p = 0; p < c.c; ++p)
        for (int bb = 0; bb < n; ++bb)
          for (int aa = 0; aa < n; ++aa) ch[base + (size_t)p * TILE2 + aa + TILE * bb] = c(aa + bb * n, p);
    }
  d_rs.upload(rs, s);
  d_rb.upload(rb, s);
  if (ch.empty()) ch.push_back(0.0);
  d_chol.upload(ch, s);
  d_off.upload(choff, s);
  d_rank.upload(chrank, s);
}

// Range-separated caches on the device: Yukawa factor tiles with the L
// scale lambda (2L+1) that turns the stored 4 pi/(2L+1) Gaunt weights into
// 4 pi lambda; erfc element-pair operators, T[k = c + 16 d][t = a + 16 b] =
// (a c | b d) for output (a, b) and density (c, d) entries, with the scale mu
// (4 pi mu/(2L+1), TwoDBasis.cpp:1082-1083).
void atomic_upload_rs(hf_atomic *h) {
  const AtomicSetup &S = h->host;
  AtomicDevice &d = h->dev;
  hipStream_t s = h->ctx->cur;
  const int Nel = S.rad.Nel(), nL = S.N_L;
  std::vector<double> lscale(nL);
  if (S.rs_kind == 1) {
    int maxr = 0;
    upload_factor_tiles(S, S.rs_small, S.rs_big, S.rs_chol, d.y_rs, d.y_rb, d.y_chol, d.y_chol_off,
                        d.y_chol_rank, maxr, s);
    for (int L = 0; L < nL; ++L) lscale[L] = S.rs_param * (2 * L + 1);
    d.e_T.release();
  } else if (S.rs_kind == 2) {
    const size_t per = (size_t)TILE2 * TILE2;
    std::vector<double> T(per * nL * Nel * Nel, 0.0);
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int L = 0; L < nL; ++L)
      for (int e = 0; e < Nel; ++e)
        for (int f = 0; f < Nel; ++f) {
          const size_t x = ((size_t)L * Nel + e) * Nel + f;
          const Blk &E = S.rs_pair[x];
          const int ni = S.rad.ni[e], nj = S.rad.ni[f];
          need(E.r == ni * ni && E.c == nj * nj, "erfc pair cache has the wrong shape");
          double *t = &T[x * per];
          for (int a = 0; a < ni; ++a)
            for (int c = 0; c < ni; ++c)
              for (int b = 0; b < nj; ++b)
                for (int dd = 0; dd < nj; ++dd)
                  t[(size_t)(c + TILE * dd) * TILE2 + a + TILE * b] = E(a * ni + c, b * nj + dd);
        }
    d.e_T.upload(T, s);
    for (int L = 0; L < nL; ++L) lscale[L] = S.rs_param;
  } else {
    throw Error(HF_ESTATE, "range-separated integrals have not been computed");
  }
  d.rs_Lscale.upload(lscale, s);
  d.rs_kind = S.rs_kind;
  hip_check(hipStreamSynchronize(s), "range-separated upload");
}

// Build the Gaunt coupling tables and upload caches for the atomic kernels.
void atomic_upload(hf_atomic *h) {
  const AtomicSetup &S = h->host;
  AtomicDevice &d = h->dev;
  hipStream_t s = h->ctx->cur;
  const int Nel = S.rad.Nel(), Nrad = S.Nrad(), Na = S.Nang(), nL = S.N_L;
  d.Nel = Nel;
  d.Nrad = Nrad;
  d.Nang = Na;
  d.N_L = nL;
  d.first.upload(S.rad.first, s);
  d.ni.upload(S.rad.ni, s);
  for (int e = 0; e < Nel; ++e)
    if (S.rad.ni[e] > TILE) throw Error(HF_EINVAL, "elements with more than 16 functions are not supported");

  // radial factor tiles
  upload_factor_tiles(S, S.rsmall, S.rbig, S.chol, d.rs, d.rb, d.chol, d.chol_off, d.chol_rank, d.max_rank, s);

  // Coulomb tables (TwoDBasis.cpp:790-822, :852-874)
  const auto &lv = S.lval;
  const auto &mv = S.mval;
  const int Mmax = *std::max_element(mv.begin(), mv.end()) - *std::min_element(mv.begin(), mv.end());
  std::vector<int> lmL, lmIndex((size_t)nL * (2 * Mmax + 1), -1);
  for (int L = 0; L < nL; ++L)
    for (int M = -std::min(L, Mmax); M <= std::min(L, Mmax); ++M) {
      lmIndex[(size_t)L * (2 * Mmax + 1) + (M + Mmax)] = (int)lmL.size();
      lmL.push_back(L);
    }
  const int nlm = (int)lmL.size();
  std::vector<std::vector<std::pair<int, double>>> pairs(nlm);
  for (int k = 0; k < Na; ++k)
    for (int l = 0; l < Na; ++l) {
      const int M = mv[k] - mv[l];
      for (int L = std::max(std::abs(lv[k] - lv[l]), std::abs(M)); L <= lv[k] + lv[l]; ++L) {
        const double c = gaunt_coeff(lv[k], mv[k], L, M, lv[l], mv[k] - M);
        if (c == 0.0) continue;
        pairs[lmIndex[(size_t)L * (2 * Mmax + 1) + (M + Mmax)]].push_back({k * Na + l, c});
      }
    }
  std::vector<int> pstart(nlm + 1, 0), pkl;
  std::vector<double> pc;
  for (int q = 0; q < nlm; ++q) {
    for (auto &pr : pairs[q]) {
      pkl.push_back(pr.first);
      pc.push_back(pr.second);
    }
    pstart[q + 1] = (int)pkl.size();
  }
  std::vector<int> ostart((size_t)Na * Na + 1, 0), olm;
  std::vector<double> oc;
  for (int i = 0; i < Na; ++i)
    for (int j = 0; j < Na; ++j) {
      const int M = mv[j] - mv[i];
      for (int L = std::max(std::abs(lv[j] - lv[i]), std::abs(M)); L <= lv[j] + lv[i]; ++L) {
        const double c = gaunt_coeff(lv[j], mv[j], L, M, lv[i], mv[j] - M);
        if (c == 0.0) continue;
        olm.push_back(lmIndex[(size_t)L * (2 * Mmax + 1) + (M + Mmax)]);
        oc.push_back(c * (kFourPi / double(2 * L + 1)));
      }
      ostart[(size_t)i * Na + j + 1] = (int)olm.size();
    }
  d.n_lm = nlm;
  d.n_pairs = (int64_t)pkl.size();
  d.lm_L.upload(lmL, s);
  d.lm_pstart.upload(pstart, s);
  if (pkl.empty()) {
    pkl.push_back(0);
    pc.push_back(0.0);
  }
  d.pair_kl.upload(pkl, s);
  d.pair_c.upload(pc, s);
  d.out_start.upload(ostart, s);
  if (olm.empty()) {
    olm.push_back(0);
    oc.push_back(0.0);
  }
  d.out_lm.upload(olm, s);
  d.out_c.upload(oc, s);

  // Exchange tables (TwoDBasis.cpp:918-968)
  std::vector<int> act, tstart(1, 0), til;
  std::vector<double> tc;
  for (int j = 0; j < Na; ++j)
    for (int k = 0; k < Na; ++k) {
      std::vector<std::vector<std::pair<int, double>>> perL(nL);
      bool anyt = false;
      for (int i = 0; i < Na; ++i)
        for (int l = 0; l < Na; ++l) {
          const int M = mv[j] - mv[i];
          if (M != mv[k] - mv[l]) continue;
          const int Lmin = std::max(std::max(std::abs(lv[i] - lv[j]), std::abs(lv[k] - lv[l])), std::abs(M));
          const int Lmax = std::min(lv[i] + lv[j], lv[k] + lv[l]);
          for (int L = Lmin; L <= Lmax; ++L) {
            const double c = gaunt_coeff(lv[j], mv[j], L, M, lv[i], mv[i]) *
                             gaunt_coeff(lv[k], mv[k], L, M, lv[l], mv[l]);
            if (c == 0.0) continue;
            perL[L].push_back({i * Na + l, (kFourPi / double(2 * L + 1)) * c});
            anyt = true;
          }
        }
      if (!anyt) continue;
      act.push_back(j * Na + k);
      for (int L = 0; L < nL; ++L) {
        for (auto &pr : perL[L]) {
          til.push_back(pr.first);
          tc.push_back(pr.second);
        }
        tstart.push_back((int)til.size());
      }
    }
  // Radial RI-K flops per coupled (output block, L): 4 r Ni^3 inside each
  // element and 2 Ni Nj (Ni + Nj) per element pair.
  d.h_radL.assign(nL, 0.0);
  d.h_radL_in.assign(nL, 0.0);
  d.h_radL_low.assign(nL, 0.0);
  d.h_ni = S.rad.ni;
  {
    std::vector<int> efi, efo, efl;
    for (int e = 0; e < Nel; ++e)
      for (int f = 0; f < Nel; ++f) {
        (e == f ? efi : efo).push_back(e * Nel + f);
        if (e > f) efl.push_back(e * Nel + f);
      }
    if (efo.empty()) efo.push_back(0);
    if (efl.empty()) efl.push_back(0);
    d.ef_in.upload(efi, s);
    d.ef_off.upload(efo, s);
    d.ef_low.upload(efl, s);
  }
  for (int L = 0; L < nL; ++L) {
    double f = 0.0;
    for (int e = 0; e < Nel; ++e) {
      const double n = S.rad.ni[e];
      f += 4.0 * S.chol[(size_t)L * Nel + e].c * n * n * n;
      d.h_radL_in[L] += 4.0 * S.chol[(size_t)L * Nel + e].c * n * n * n;
      for (int g = 0; g < Nel; ++g)
        if (g != e) {
          const double m = S.rad.ni[g];
          f += 2.0 * n * m * (n + m);
          if (e > g) d.h_radL_low[L] += 2.0 * n * m * (n + m);
        }
    }
    d.h_radL[L] = f;
  }
  d.h_act = act;
  d.h_tstart = tstart;
  d.h_til = til;
  d.h_tc = tc;
  d.screen_mask.clear();
  d.n_act = 0;
  d.n_terms = 0;
  d.h_bnorm.alloc((size_t)Na * Na);
  d.h_stats.alloc((size_t)Na * Na * 4 * block_stats_chunks(d.Nrad));
  d.stats.alloc((size_t)Na * Na * 4 * block_stats_chunks(d.Nrad));
  d.Jin.alloc((size_t)std::max(nlm, 1) * Nel * TILE2);
  d.jsm.alloc((size_t)std::max(nlm, 1) * Nel);
  d.jbg.alloc((size_t)std::max(nlm, 1) * Nel);
  d.preA.alloc((size_t)std::max(nlm, 1) * Nel);
  d.preB.alloc((size_t)std::max(nlm, 1) * Nel);

  hip_check(hipStreamSynchronize(s), "upload sync");
  atomic_build_operator(S, d, s);
  h->uploaded = true;
}

void ensure_ready(hf_atomic *h) {
  if (!h->ctx) throw Error(HF_ESTATE, "handle was created without a device context");
  if (!h->host.have_tei) throw Error(HF_ESTATE, "Primitive teis have not been computed!");
  if (!h->uploaded) atomic_upload(h);
}

}  // namespace

extern "C" {

const char *hf_last_error(void) { return g_err.c_str(); }
const char *hf_version(void) { return "hfgpu 0.1 (gfx950)"; }

int hf_element_grid(double rmax, int nel, int igrid, double zexp, double *out) {
  return guarded([&] {
    need(out != nullptr, "out is NULL");
    need(nel >= 1, "need at least one element");
    need(rmax > 0.0, "rmax must be positive");
    need(igrid >= 1 && igrid <= 4, "igrid must be 1..4");
    const std::vector<double> b = element_grid(rmax, nel, igrid, zexp);
    std::copy(b.begin(), b.end(), out);
  });
}

int hf_ctx_create(int device, unsigned flags, hf_ctx **out) {
  return guarded([&] {
    need(out != nullptr, "out is NULL");
    need(flags == 0, "flags must be 0");
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    need(device >= 0 && device < n, "device index out of range");
    std::unique_ptr<hf_ctx> c(new hf_ctx);
    c->device = device;
    DeviceGuard g(device);
    hip_check(hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking), "hipStreamCreate");
    c->cur = c->own;
    noalloc_scratch_init(device);
    *out = c.release();
  });
}

int hf_ctx_destroy(hf_ctx *ctx) {
  return guarded([&] {
    if (!ctx) return;
    DeviceGuard g(ctx->device);
    (void)hipStreamSynchronize(ctx->cur);
    hf_ctx_release_blas(ctx);
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
    delete ctx;
  });
}

int hf_ctx_set_stream(hf_ctx *ctx, void *stream) {
  return guarded([&] {
    need(ctx != nullptr, "ctx is NULL");
    ctx->cur = (hipStream_t)stream;
  });
}

int hf_ctx_get_stream(hf_ctx *ctx, void **stream) {
  return guarded([&] {
    need(ctx && stream, "NULL argument");
    *stream = (void *)ctx->cur;
  });
}

int hf_ctx_set_timing(hf_ctx *ctx, int on) {
  return guarded([&] {
    need(ctx != nullptr, "ctx is NULL");
    ctx->timer.collect();
    ctx->timer.on = on != 0;
    if (!on) ctx->timer.acc.clear();
  });
}

int hf_ctx_kernel_stats(hf_ctx *ctx, const char *name, double out[3]) {
  return guarded([&] {
    need(ctx && name && out, "NULL argument");
    DeviceGuard g(ctx->device);
    ctx->timer.collect();
    auto it = ctx->timer.acc.find(name);
    if (it == ctx->timer.acc.end()) {
      out[0] = out[1] = out[2] = 0.0;
      return;
    }
    out[0] = (double)it->second.count;
    out[1] = it->second.ms;
    out[2] = it->second.work;
  });
}

int hf_ctx_synchronize(hf_ctx *ctx) {
  return guarded([&] {
    need(ctx != nullptr, "ctx is NULL");
    DeviceGuard g(ctx->device);
    hip_check(hipStreamSynchronize(ctx->cur), "hipStreamSynchronize");
  });
}

int hf_atomic_create(hf_ctx *ctx, int Z, int nnodes, int nquad, const double *bval, int nbval, const int *lval,
                     const int *mval, int nang, hf_atomic **out) {
  return guarded([&] {
    need(bval && lval && mval && out, "NULL argument");
    need(nbval >= 2, "need at least one element (nbval >= 2)");
    need(nang >= 1, "need at least one angular shell");
    need(nnodes >= 2 && nnodes <= TILE, "nnodes must be in [2, 16]");
    need(Z >= 0, "Z must be non-negative");
    for (int i = 0; i + 1 < nbval; ++i) need(bval[i + 1] > bval[i], "element boundaries must increase");
    need(bval[0] == 0.0, "the first element boundary must be r=0");
    std::unique_ptr<hf_atomic> h(new hf_atomic);
    h->ctx = ctx;
    h->host.build(Z, nnodes, nquad > 0 ? nquad : 5 * nnodes, std::vector<double>(bval, bval + nbval),
                  std::vector<int>(lval, lval + nang), std::vector<int>(mval, mval + nang));
    *out = h.release();
  });
}

int hf_atomic_clone(const hf_atomic *h, hf_atomic **out) {
  return guarded([&] {
    need(h != nullptr && out != nullptr, "NULL argument");
    std::unique_ptr<hf_atomic> c(new hf_atomic);
    c->ctx = h->ctx;
    c->host = h->host;
    *out = c.release();
  });
}

int hf_atomic_destroy(hf_atomic *h) {
  return guarded([&] {
    if (!h) return;
    if (h->ctx) {
      DeviceGuard g(h->ctx->device);
      (void)hipStreamSynchronize(h->ctx->cur);
      delete h;
    } else {
      delete h;
    }
  });
}

int hf_atomic_sizes(const hf_atomic *h, int64_t out[5]) {
  return guarded([&] {
    need(h && out, "NULL argument");
    out[0] = h->host.Nbf();
    out[1] = h->host.Nrad();
    out[2] = h->host.Nang();
    out[3] = h->host.rad.Nel();
    out[4] = h->host.N_L;
  });
}

int hf_atomic_work(hf_atomic *h, double out[4]) {
  return guarded([&] {
    need(h && out, "NULL argument");
    if (!h->host.have_tei) throw Error(HF_ESTATE, "Primitive teis have not been computed!");
    if (!h->ctx) throw Error(HF_ESTATE, "handle was created without a device context");
    DeviceGuard g(h->ctx->device);
    if (!h->uploaded) atomic_upload(h);
    out[0] = h->dev.flops_k_angular;
    out[1] = h->dev.flops_k_radial;
    out[2] = (double)h->dev.n_terms;
    out[3] = (double)h->dev.n_act;
  });
}

int hf_atomic_elements(const hf_atomic *h, int *first, int *nprim) {
  return guarded([&] {
    need(h && first && nprim, "NULL argument");
    for (int e = 0; e < h->host.rad.Nel(); ++e) {
      first[e] = h->host.rad.first[e];
      nprim[e] = h->host.rad.ni[e];
    }
  });
}

int hf_atomic_compute_tei(hf_atomic *h, int exchange) {
  (void)exchange;
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    h->host.compute_tei();
    h->uploaded = false;
  });
}

int hf_atomic_compute_tei_device(hf_atomic *h, double *seconds) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    need(h->ctx != nullptr, "handle was created without a device context");
    DeviceGuard g(h->ctx->device);
    atomic_tei_device(h->host, h->ctx->cur, seconds);
    h->uploaded = false;
  });
}

int hf_atomic_compute_yukawa(hf_atomic *h, double lambda) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    need(lambda > 0.0 && std::isfinite(lambda), "Yukawa range parameter must be positive");
    h->host.compute_yukawa(lambda);
    h->rs_uploaded = false;
  });
}

int hf_atomic_compute_erfc(hf_atomic *h, double mu) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    need(mu > 0.0 && std::isfinite(mu), "erfc range parameter must be positive");
    h->host.compute_erfc(mu);
    h->rs_uploaded = false;
  });
}

int hf_atomic_set_cache(hf_atomic *h, int kind, int L, int iel, int rows, int cols, const double *data) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    AtomicSetup &S = h->host;
    const int Nel = S.rad.Nel();
    need(kind >= 0 && kind <= 2, "kind must be 0, 1 or 2");
    need(L >= 0 && L < S.N_L && iel >= 0 && iel < Nel, "(L, iel) out of range");
    need(rows >= 0 && cols >= 0 && (data || rows * cols == 0), "bad dimensions");
    const size_t n = (size_t)S.N_L * Nel;
    if (S.rsmall.size() != n) {
      S.rsmall.assign(n, Blk());
      S.rbig.assign(n, Blk());
      S.chol.assign(n, Blk());
    }
    Blk b(rows, cols);
    if (rows * cols) std::memcpy(b.v.data(), data, sizeof(double) * (size_t)rows * cols);
    std::vector<Blk> &v = kind == 0 ? S.rsmall : kind == 1 ? S.rbig : S.chol;
    v[(size_t)L * Nel + iel] = std::move(b);
    S.have_tei = false;
    h->uploaded = false;
  });
}

int hf_atomic_commit_tei(hf_atomic *h) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    AtomicSetup &S = h->host;
    const int Nel = S.rad.Nel();
    need(S.chol.size() == (size_t)S.N_L * Nel, "no caches have been set");
    for (int L = 0; L < S.N_L; ++L)
      for (int e = 0; e < Nel; ++e) {
        const size_t q = (size_t)L * Nel + e;
        const int n = S.rad.ni[e];
        need(S.rsmall[q].r == n && S.rsmall[q].c == n, "r^L cache missing or of the wrong shape");
        need(e == 0 || (S.rbig[q].r == n && S.rbig[q].c == n), "r^-L-1 cache missing or of the wrong shape");
        need(S.chol[q].r == n * n, "Cholesky cache missing or of the wrong shape");
      }
    S.have_tei = true;
    h->uploaded = false;
  });
}

int hf_atomic_get_cache(const hf_atomic *h, int kind, int L, int iel, int dims[2], double *data) {
  return guarded([&] {
    need(h && dims, "NULL argument");
    const AtomicSetup &S = h->host;
    const int Nel = S.rad.Nel();
    need(kind >= 0 && kind <= 2, "kind must be 0, 1 or 2");
    need(L >= 0 && L < S.N_L && iel >= 0 && iel < Nel, "(L, iel) out of range");
    if (!S.have_tei) throw Error(HF_ESTATE, "Primitive teis have not been computed!");
    const std::vector<Blk> &v = kind == 0 ? S.rsmall : kind == 1 ? S.rbig : S.chol;
    const Blk &b = v[(size_t)L * Nel + iel];
    dims[0] = b.r;
    dims[1] = b.c;
    if (data && !b.v.empty()) std::memcpy(data, b.v.data(), sizeof(double) * b.v.size());
  });
}

int hf_atomic_one_electron(const hf_atomic *h, int which, double *out, int64_t ld) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    check_matrix(out, ld, h->host.Nbf(), "out");
    if (which == 0)
      h->host.overlap(out, ld);
    else if (which == 1)
      h->host.kinetic(out, ld);
    else if (which == 2)
      h->host.nuclear(out, ld);
    else
      throw Error(HF_EINVAL, "which must be 0 (overlap), 1 (kinetic) or 2 (nuclear)");
  });
}

namespace {
// the radial one-electron matrices of kind `which` (and the shells' l
// values) on the device, computed and uploaded on first use
void prepare_one_electron(const hf_atomic *hc, int which) {
  hf_atomic *h = const_cast<hf_atomic *>(hc);
  if (h->lval_dev.p && h->one_M[which].p) return;
  // the host copies stay alive until the stream has consumed them
  std::vector<int> l(h->host.lval.begin(), h->host.lval.end());
  std::vector<double> M, X;
  if (!h->lval_dev.p) h->lval_dev.upload(l, h->ctx->cur);
  if (!h->one_M[which].p) {
    h->host.radial_one_electron(which, M, X);
    h->one_M[which].upload(M, h->ctx->cur);
    if (which == 1) h->one_X.upload(X, h->ctx->cur);
  }
  hip_check(hipStreamSynchronize(h->ctx->cur), "one-electron upload");
}
}  // namespace

int hf_atomic_one_electron_dev(const hf_atomic *h, int which, double *dOut, int64_t ld) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    need(h->ctx != nullptr, "handle was created without a device context");
    need(which >= 0 && which <= 2, "which must be 0 (overlap), 1 (kinetic) or 2 (nuclear)");
    DeviceGuard g(h->ctx->device);
    check_matrix(dOut, ld, h->host.Nbf(), "out");
    prepare_one_electron(h, which);
    shell_diag_launch(dOut, ld, h->host.Nbf(), h->host.Nrad(), h->one_M[which].p,
                      which == 1 ? h->one_X.p : nullptr, h->lval_dev.p, which == 2 ? -double(h->host.Z) : 1.0,
                      h->ctx->cur);
  });
}

int hf_atomic_one_electron_prepare(hf_atomic *h) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    need(h->ctx != nullptr, "handle was created without a device context");
    DeviceGuard g(h->ctx->device);
    for (int w = 0; w < 3; ++w) prepare_one_electron(h, w);
  });
}

int hf_atomic_coulomb_dev(hf_atomic *h, const double *dP, int64_t ldp, double *dJ, int64_t ldj) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    need(h->ctx != nullptr, "handle was created without a device context");
    DeviceGuard g(h->ctx->device);
    ensure_ready(h);
    check_matrix(dP, ldp, h->host.Nbf(), "P");
    check_matrix(dJ, ldj, h->host.Nbf(), "J");
    atomic_coulomb_launch(h->dev, dP, ldp, dJ, ldj, h->ctx->cur, &h->ctx->timer);
  });
}

int hf_atomic_exchange_dev(hf_atomic *h, const double *dP, int64_t ldp, double *dK, int64_t ldk) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    need(h->ctx != nullptr, "handle was created without a device context");
    DeviceGuard g(h->ctx->device);
    ensure_ready(h);
    check_matrix(dP, ldp, h->host.Nbf(), "P");
    check_matrix(dK, ldk, h->host.Nbf(), "K");
    atomic_exchange_launch(h->dev, dP, ldp, dK, ldk, h->ctx->cur, &h->ctx->timer);
  });
}

int hf_atomic_exchange_fp32_dev(hf_atomic *h, const double *dP, int64_t ldp, double *dK, int64_t ldk) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    need(h->ctx != nullptr, "handle was created without a device context");
    DeviceGuard g(h->ctx->device);
    ensure_ready(h);
    check_matrix(dP, ldp, h->host.Nbf(), "P");
    check_matrix(dK, ldk, h->host.Nbf(), "K");
    atomic_exchange_launch(h->dev, dP, ldp, dK, ldk, h->ctx->cur, &h->ctx->timer, true);
  });
}

namespace {
void rs_ready(hf_atomic *h) {
  ensure_ready(h);
  if (h->host.rs_kind == 0) throw Error(HF_ESTATE, "Primitive teis have not been computed!");
  if (!h->rs_uploaded) {
    atomic_upload_rs(h);
    h->rs_uploaded = true;
  }
}
}  // namespace

int hf_atomic_rs_exchange_dev(hf_atomic *h, const double *dP, int64_t ldp, double *dK, int64_t ldk) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    need(h->ctx != nullptr, "handle was created without a device context");
    DeviceGuard g(h->ctx->device);
    rs_ready(h);
    check_matrix(dP, ldp, h->host.Nbf(), "P");
    check_matrix(dK, ldk, h->host.Nbf(), "K");
    atomic_rs_exchange_launch(h->dev, dP, ldp, dK, ldk, h->ctx->cur, &h->ctx->timer);
  });
}

int hf_atomic_rs_exchange(hf_atomic *h, const double *P, int64_t ldp, double *K, int64_t ldk) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    need(h->ctx != nullptr, "handle was created without a device context");
    DeviceGuard g(h->ctx->device);
    rs_ready(h);
    host_roundtrip(h->ctx->cur, h->host.Nbf(), h->dev.hostP, h->dev.hostOut, P, ldp, K, ldk,
                   [&](const double *dp, int64_t lp, double *dout, int64_t lo) {
                     atomic_rs_exchange_launch(h->dev, dp, lp, dout, lo, h->ctx->cur, &h->ctx->timer);
                   });
  });
}

int hf_atomic_coulomb(hf_atomic *h, const double *P, int64_t ldp, double *J, int64_t ldj) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    need(h->ctx != nullptr, "handle was created without a device context");
    DeviceGuard g(h->ctx->device);
    ensure_ready(h);
    host_roundtrip(h->ctx->cur, h->host.Nbf(), h->dev.hostP, h->dev.hostOut, P, ldp, J, ldj, [&](const double *dp, int64_t lp, double *dout, int64_t lo) {
      atomic_coulomb_launch(h->dev, dp, lp, dout, lo, h->ctx->cur, &h->ctx->timer);
    });
  });
}

int hf_atomic_exchange(hf_atomic *h, const double *P, int64_t ldp, double *K, int64_t ldk) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    need(h->ctx != nullptr, "handle was created without a device context");
    DeviceGuard g(h->ctx->device);
    ensure_ready(h);
    host_roundtrip(h->ctx->cur, h->host.Nbf(), h->dev.hostP, h->dev.hostOut, P, ldp, K, ldk, [&](const double *dp, int64_t lp, double *dout, int64_t lo) {
      atomic_exchange_launch(h->dev, dp, lp, dout, lo, h->ctx->cur, &h->ctx->timer);
    });
  });
}

}  // extern "C"

// ------------------------------------------------------------ atomic batch

struct hf_atomic_batch {
  hf_ctx *ctx = nullptr;
  std::vector<hf_atomic *> handles;
  hfg::AtomicBatch b;
};

extern "C" {

int hf_atomic_batch_create(hf_ctx *ctx, hf_atomic *const *handles, int n, hf_atomic_batch **out) {
  return guarded([&] {
    need(ctx != nullptr && out != nullptr, "NULL context or output");
    need(n >= 1 && handles != nullptr, "a batch needs at least one handle");
    DeviceGuard g(ctx->device);
    std::unique_ptr<hf_atomic_batch> b(new hf_atomic_batch);
    b->ctx = ctx;
    for (int i = 0; i < n; ++i) {
      hf_atomic *h = handles[i];
      need(h != nullptr, "NULL handle in the batch");
      need(h->ctx == ctx, "every handle of a batch must use the batch's context");
      ensure_ready(h);
      need(!h->dev.dense_in, "the dense in-element operator is not supported in a batch");
      b->handles.push_back(h);
      b->b.sys.push_back(&h->dev);
    }
    *out = b.release();
  });
}

int hf_atomic_batch_destroy(hf_atomic_batch *b) {
  return guarded([&] {
    if (!b) return;
    DeviceGuard g(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->cur);
    delete b;
  });
}

int hf_atomic_batch_set_options(hf_atomic_batch *b, int use_graphs, int autotune) {
  return guarded([&] {
    need(b != nullptr, "NULL batch");
    b->b.use_graphs = use_graphs != 0;
    b->b.autotune = autotune != 0;
  });
}

int hf_atomic_batch_tuning(const hf_atomic_batch *b, int nw[8], double ms[24]) {
  return guarded([&] {
    need(b != nullptr, "NULL batch");
    for (int c = 0; c < 4; ++c)
      for (int p = 0; p < 2; ++p) {
        if (nw) nw[c * 2 + p] = b->b.nw[c][p];
        for (int q = 0; q < 3; ++q)
          if (ms) ms[(c * 2 + p) * 3 + q] = b->b.tune_ms[c][p][q];
      }
  });
}

namespace {
void check_batch_mats(hf_atomic_batch *b, const double *const *P, const int64_t *ldp, double *const *O,
                      const int64_t *ldo) {
  need(P != nullptr && ldp != nullptr && O != nullptr && ldo != nullptr, "NULL pointer array");
  for (size_t i = 0; i < b->handles.size(); ++i) {
    check_matrix(P[i], ldp[i], b->handles[i]->host.Nbf(), "P");
    check_matrix(O[i], ldo[i], b->handles[i]->host.Nbf(), "output");
  }
}
}  // namespace

int hf_atomic_batch_coulomb_dev(hf_atomic_batch *b, const double *const *dP, const int64_t *ldp, double *const *dJ,
                                const int64_t *ldj) {
  return guarded([&] {
    need(b != nullptr, "NULL batch");
    DeviceGuard g(b->ctx->device);
    check_batch_mats(b, dP, ldp, dJ, ldj);
    atomic_batch_coulomb(b->b, dP, ldp, dJ, ldj, b->ctx->cur, &b->ctx->timer);
  });
}

int hf_atomic_batch_exchange_dev(hf_atomic_batch *b, const double *const *dP, const int64_t *ldp, double *const *dK,
                                 const int64_t *ldk, const int *fp32) {
  return guarded([&] {
    need(b != nullptr, "NULL batch");
    DeviceGuard g(b->ctx->device);
    check_batch_mats(b, dP, ldp, dK, ldk);
    std::vector<int> f(b->handles.size(), 0);
    if (fp32)
      for (size_t i = 0; i < f.size(); ++i) f[i] = fp32[i] != 0;
    atomic_batch_exchange(b->b, dP, ldp, dK, ldk, f.data(), b->ctx->cur, &b->ctx->timer);
  });
}

}  // extern "C"
