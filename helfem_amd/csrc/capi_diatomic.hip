// C-ABI for the diatomic basis (include/hfgpu.h): handle, device tables and
// the coupling lists of the Coulomb and exchange kernels (diatomic_jk.hip).
#include "capi_common.hpp"
#include <cstring>
#include <cstdlib>
#include <cmath>
#include <memory>
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace hfg;

namespace hfg {

// In-element exchange operator per ((L,|M|), element), from the sign-aware
// Cholesky factor: W~(i, j) = sum_p sigma_p B(i, p) B(j, p) over the two
// channel blocks (basis.cpp:1992-2033 contracts the same factor as
// sum_p sigma_p [M0' R00 M0' - M0' R02 M2' - M2' R20 M0' + M2' R22 M2']).
// Built only when the environment variable HFGPU_EXCHANGE_OPERATOR_GIB grants
// a budget (in GiB) that holds the tables; otherwise (dense_in = false) the
// Cholesky-form in-element kernel runs, which on gfx950 is the faster of the
// two (64 vs 216 ms for N2 at lmax 16), so the operator path is opt-in.
void diatomic_build_operator(const DiatomicSetup &S, DiatomicDevice &d, hipStream_t s) {
  const int Nel = S.Nel(), nlm = (int)S.lm_map.size();
  const size_t per = (size_t)4 * KC * TILE2;
  const double bytes = 8.0 * per * nlm * Nel;
  double budget = 0.0;
  if (const char *env = std::getenv("HFGPU_EXCHANGE_OPERATOR_GIB")) budget = std::atof(env) * (1ull << 30);
  d.dense_in = false;
  d.Top.release();
  if (nlm == 0 || bytes > budget) return;
  std::vector<double> T(per * nlm * Nel, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (int x = 0; x < nlm * Nel; ++x) {
    const int e = x % Nel;
    const int n = S.rad.ni[e], nn = n * n, N = 2 * nn;
    const Blk &B = S.cdB[x];
    const std::vector<double> &sg = S.cdS[x];
    std::vector<double> W((size_t)N * N, 0.0);
    for (int p = 0; p < B.c; ++p) {
      const double *bp = &B.v[(size_t)p * N];
      for (int j = 0; j < N; ++j) {
        const double f = sg[p] * bp[j];
        double *wj = &W[(size_t)j * N];
        for (int i = 0; i < N; ++i) wj[i] += f * bp[i];
      }
    }
    double *Tx = &T[per * x];
    const int xo[4] = {0, 0, nn, nn}, yo[4] = {0, nn, 0, nn};
    const double sgn[4] = {1.0, -1.0, -1.0, 1.0};
    for (int ch = 0; ch < 4; ++ch)
      for (int dd = 0; dd < n; ++dd)
        for (int cc = 0; cc < n; ++cc) {
          double *row = Tx + ((size_t)ch * KC + cc + 15 * dd) * TILE2;
          for (int b = 0; b < n; ++b)
            for (int a = 0; a < n; ++a)
              row[a + TILE * b] = sgn[ch] * W[(size_t)(xo[ch] + cc + a * n) + (size_t)(yo[ch] + b + dd * n) * N];
        }
  }
  d.Top.upload(T, s);
  hip_check(hipStreamSynchronize(s), "exchange operator upload");
  d.dense_in = true;
}

}  // namespace hfg

namespace {

// Per (j, i, L) Gaunt couplings with M = mj - mi: g0 = mod_coeff (cosh^0
// channel), g2 = coeff (cosh^2 channel), basis.cpp:1650-1665, 1908-1916.
void gaunt_tables(const DiatomicSetup &S, DiatomicDevice &d) {
  const int Na = S.Nang();
  int lmax = 0;
  for (int l : S.lval) lmax = std::max(lmax, l);
  d.gL = 2 * lmax + 3;
  d.h_g0.assign((size_t)Na * Na * d.gL, 0.0);
  d.h_g2.assign((size_t)Na * Na * d.gL, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (int j = 0; j < Na; ++j)
    for (int i = 0; i < Na; ++i) {
      const int lj = S.lval[j], mj = S.mval[j], li = S.lval[i], mi = S.mval[i], M = mj - mi;
      for (int L = std::max(std::abs(lj - li) - 2, std::abs(M)); L <= lj + li + 2; ++L) {
        const size_t x = ((size_t)j * Na + i) * d.gL + L;
        d.h_g0[x] = gaunt_mod(lj, mj, L, M, li, mi);
        d.h_g2[x] = gaunt3(lj, mj, L, M, li);
      }
    }
}

void diatomic_upload(hf_diatomic *h) {
  const DiatomicSetup &S = h->host;
  DiatomicDevice &d = h->dev;
  hipStream_t s = h->ctx->cur;
  const int Nel = S.Nel(), Na = S.Nang(), nlm = (int)S.lm_map.size(), nLM = (int)S.LM_map.size();
  d.Nel = Nel;
  d.Nrad = S.Nrad();
  d.Nang = Na;
  d.n_lm = nlm;
  d.n_LM = nLM;
  d.Nbf = S.Nbf();
  for (int e = 0; e < Nel; ++e)
    if (S.rad.ni[e] > TILE) throw Error(HF_EINVAL, "elements with more than 16 functions are not supported");
  d.first.upload(S.rad.first, s);
  d.ni.upload(S.rad.ni, s);
  std::vector<int> skip(Na);
  for (int a = 0; a < Na; ++a) skip[a] = S.skip(a);
  d.skip.upload(skip, s);
  std::vector<long long> poff(S.poff.begin(), S.poff.end());
  d.poff.upload(poff, s);

  // P/Q tiles and Cholesky channel tiles
  std::vector<double> pq((size_t)std::max(nlm, 1) * Nel * 4 * TILE2, 0.0), ch, sg;
  std::vector<int> choff((size_t)std::max(nlm, 1) * Nel, 0), chrank((size_t)std::max(nlm, 1) * Nel, 0);
  int maxr = 0;
  for (int q = 0; q < nlm; ++q)
    for (int e = 0; e < Nel; ++e) {
      const size_t x = (size_t)q * Nel + e;
      const int n = S.rad.ni[e];
      const Blk *src[4] = {&S.P0[x], &S.P2[x], &S.Q0[x], &S.Q2[x]};
      for (int kind = 0; kind < 4; ++kind) {
        need(src[kind]->r == n && src[kind]->c == n, "P/Q cache has the wrong shape");
        for (int b = 0; b < n; ++b)
          for (int a = 0; a < n; ++a) pq[(x * 4 + kind) * TILE2 + a + TILE * b] = (*src[kind])(a, b);
      }
      const Blk &B = S.cdB[x];
      const int nn = n * n;
      need(B.r == 2 * nn, "Cholesky cache has the wrong number of rows");
      need((int)S.cdS[x].size() == B.c, "Cholesky signs have the wrong length");
      need(B.c <= 2 * TILE2, "Cholesky rank above 512 is not supported");
      choff[x] = (int)sg.size();
      chrank[x] = B.c;
      maxr = std::max(maxr, B.c);
      const size_t base = ch.size();
      ch.resize(base + (size_t)B.c * 2 * TILE2, 0.0);
      for (int p = 0; p < B.c; ++p) {
        double *m0 = &ch[base + (size_t)p * 2 * TILE2], *m2 = m0 + TILE2;
        for (int b = 0; b < n; ++b)
          for (int a = 0; a < n; ++a) {
            m0[a + TILE * b] = B(a + b * n, p);
            m2[a + TILE * b] = B(nn + a + b * n, p);
          }
        sg.push_back(S.cdS[x][p]);
      }
    }
  d.max_rank = maxr;
  if (ch.empty()) ch.push_back(0.0);
  if (sg.empty()) sg.push_back(0.0);
  d.pq.upload(pq, s);
  // log scale per (L,|M|, element) that brings the P and Q tiles to the same
  // magnitude, sqrt(max|P| max|Q|): P_L(cosh mu) reaches 1e75 and Q_L 1e-56
  // at L = 34, outside FP32, while their products across elements are O(1)
  {
    std::vector<double> lsig((size_t)std::max(nlm, 1) * Nel, 0.0);
    for (size_t x = 0; x < (size_t)nlm * Nel; ++x) {
      double mp = 0.0, mq = 0.0;
      for (int kind = 0; kind < 4; ++kind)
        for (int u = 0; u < TILE2; ++u) {
          const double v = std::fabs(pq[(x * 4 + kind) * TILE2 + u]);
          if (kind < 2)
            mp = std::max(mp, v);
          else
            mq = std::max(mq, v);
        }
      if (mp > 0.0 && mq > 0.0) lsig[x] = 0.5 * (std::log(mp) - std::log(mq));
    }
    d.lsig.upload(lsig, s);
  }
  d.chol.upload(ch, s);
  d.sigma.upload(sg, s);
  d.chol_off.upload(choff, s);
  d.chol_rank.upload(chrank, s);

  // Coulomb tables (basis.cpp:1633-1660, 1790-1810)
  gaunt_tables(S, d);
  const auto &lv = S.lval;
  const auto &mv = S.mval;
  std::vector<int> LMilm(std::max(nLM, 1), 0);
  std::vector<double> LMfac(std::max(nLM, 1), 0.0);
  for (int q = 0; q < nLM; ++q) {
    LMilm[q] = S.lmind(S.LM_map[q].first, S.LM_map[q].second);
    LMfac[q] = S.LMfac(S.LM_map[q].first, S.LM_map[q].second);
  }
  std::vector<std::vector<std::pair<int, std::pair<double, double>>>> pairs(nLM);
  for (int k = 0; k < Na; ++k)
    for (int l = 0; l < Na; ++l) {
      const int M = mv[k] - mv[l];
      for (int L = std::max(std::abs(lv[k] - lv[l]) - 2, std::abs(M)); L <= lv[k] + lv[l] + 2; ++L) {
        const size_t x = ((size_t)k * Na + l) * d.gL + L;
        const double c0 = d.h_g0[x], c2 = d.h_g2[x];
        if (c0 == 0.0 && c2 == 0.0) continue;
        pairs[S.LMind(L, M)].push_back({k * Na + l, {c0, c2}});
      }
    }
  std::vector<int> pstart(nLM + 1, 0), pkl;
  std::vector<double> pc;
  for (int q = 0; q < nLM; ++q) {
    for (auto &pr : pairs[q]) {
      pkl.push_back(pr.first);
      pc.push_back(pr.second.first);
      pc.push_back(pr.second.second);
    }
    pstart[q + 1] = (int)pkl.size();
  }
  std::vector<int> ostart((size_t)Na * Na + 1, 0), oq;
  std::vector<double> oc;
  for (int i = 0; i < Na; ++i)
    for (int j = 0; j < Na; ++j) {
      const int M = mv[j] - mv[i];
      for (int L = std::max(std::abs(lv[j] - lv[i]) - 2, std::abs(M)); L <= lv[j] + lv[i] + 2; ++L) {
        const size_t x = ((size_t)j * Na + i) * d.gL + L;
        const double c0 = d.h_g0[x], c2 = d.h_g2[x];
        if (c0 == 0.0 && c2 == 0.0) continue;
        oq.push_back(S.LMind(L, M));
        oc.push_back(c0);
        oc.push_back(c2);
      }
      ostart[(size_t)i * Na + j + 1] = (int)oq.size();
    }
  d.n_pairs = (int64_t)pkl.size();
  if (pkl.empty()) {
    pkl.push_back(0);
    pc.assign(2, 0.0);
  }
  if (oq.empty()) {
    oq.push_back(0);
    oc.assign(2, 0.0);
  }
  d.LM_ilm.upload(LMilm, s);
  d.LM_fac.upload(LMfac, s);
  d.pstart.upload(pstart, s);
  d.pair_kl.upload(pkl, s);
  d.pair_c.upload(pc, s);
  d.out_start.upload(ostart, s);
  d.out_q.upload(oq, s);
  d.out_c.upload(oc, s);

  // radial RI-K flops per coupled (output block, (L,|M|)): 12 r n^3 inside
  // each element, 2 (4 ni nj^2 + 2 ni^2 nj) per element pair
  d.h_rad_in.assign(std::max(nlm, 1), 0.0);
  d.h_rad_off.assign(std::max(nlm, 1), 0.0);
  d.h_rad_low.assign(std::max(nlm, 1), 0.0);
  for (int q = 0; q < nlm; ++q) {
    double fi = 0.0, fo = 0.0, fl = 0.0;
    for (int e = 0; e < Nel; ++e) {
      const double n = S.rad.ni[e];
      fi += 12.0 * S.cdB[(size_t)q * Nel + e].c * n * n * n;
      for (int g = 0; g < Nel; ++g)
        if (g != e) {
          const double m = S.rad.ni[g], x = 2.0 * (4.0 * n * m * m + 2.0 * n * n * m);
          fo += x;
          if (e > g) fl += x;
        }
    }
    d.h_rad_in[q] = fi;
    d.h_rad_off[q] = fo;
    d.h_rad_low[q] = fl;
  }
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
  // absm mirror pairs: K(j,k) = K(jm,km) for mj, mk < 0 (basis.cpp:2067-2086)
  std::vector<int> mir;
  if (h->absm) {
    std::vector<int> partner(Na, -1);
    for (int a = 0; a < Na; ++a)
      for (int b = 0; b < Na; ++b)
        if (lv[b] == lv[a] && mv[b] == -mv[a]) {
          partner[a] = b;
          break;
        }
    for (int j = 0; j < Na; ++j) {
      if (mv[j] >= 0 || partner[j] < 0) continue;
      for (int k = 0; k < Na; ++k) {
        if (mv[k] >= 0 || partner[k] < 0) continue;
        mir.push_back(j * Na + k);
        mir.push_back(partner[j] * Na + partner[k]);
      }
    }
  }
  d.absm = h->absm;
  d.n_mirror = (int)mir.size() / 2;
  if (mir.empty()) mir.assign(2, 0);
  d.mirror.upload(mir, s);
  d.screen_mask.clear();
  d.n_act = 0;
  d.h_bnorm.alloc((size_t)Na * Na);

  d.h_stats.alloc((size_t)Na * Na * 4 * block_stats_chunks(d.Nrad));
  d.stats.alloc((size_t)Na * Na * 4 * block_stats_chunks(d.Nrad));
  d.Jin.alloc((size_t)std::max(nLM, 1) * Nel * 2 * TILE2);
  d.ds.alloc((size_t)std::max(nLM, 1) * Nel);
  d.db.alloc((size_t)std::max(nLM, 1) * Nel);
  d.preA.alloc((size_t)std::max(nLM, 1) * Nel);
  d.preB.alloc((size_t)std::max(nLM, 1) * Nel);
  hip_check(hipStreamSynchronize(s), "upload sync");
  diatomic_build_operator(S, d, s);
  diatomic_op_build(d, s);
  h->uploaded = true;
}

// Lists of the two-step cross-element exchange (k_dia_k2_off).  The terms of
// output block (j,k) and (L,M) factor as fac(L,M) u(j,i) v(k,l) with u = (mod
// Gaunt, plain Gaunt) of (j,i) and v the same of (k,l) (basis.cpp:1905-1960),
// so R(j,k) = fac sum_i u(j,i) (x) V(i) with V_b(i) = sum_l v_b(k,l) P(i,l):
// the V tiles of one (k, (L,|M|)) group serve every row j.  l_j + L + l_i is
// even for every coupling, so a group splits by the parity of l_i into two
// with disjoint rows.  A slot is one (M, i) of a group; its lt entries are the
// (tile of (i,l), v0, v1) that pass the screen; a row's jt entries are (slot,
// fac u0, fac u1).  Leaves d.k2 off when a column has more than K2_JMAX rows
// or a group more than K2_SLOTS slots.
void build_twostep(hf_diatomic *h, const std::vector<int> &act, const std::vector<int> &tile_of,
                   const std::vector<unsigned char> &mask, const std::vector<int> &ilm_of,
                   const std::vector<double> &fac_of, int Mext, int dmoff, hipStream_t s) {
  const DiatomicSetup &S = h->host;
  DiatomicDevice &d = h->dev;
  const int Na = S.Nang(), Nel = S.Nel();
  const auto &lv = S.lval;
  const auto &mv = S.mval;
  d.k2 = false;
  d.n_wg_k2 = 0;
  if (!d.tiled || Nel < 2 || act.empty()) return;
  // rows of every column shell k
  std::vector<std::vector<int>> rows(Na);
  for (int jk : act) rows[jk % Na].push_back(jk / Na);
  std::vector<int> ks;
  for (int k = 0; k < Na; ++k) {
    if (rows[k].empty()) continue;
    if ((int)rows[k].size() > K2_JMAX) return;
    ks.push_back(k);
  }
  struct Group {
    int ilm, par;
    std::vector<std::pair<int, int>> slots;               // (M, i)
    std::vector<std::vector<std::pair<int, std::pair<double, double>>>> lt;  // per slot: (tile, v)
    std::vector<std::vector<std::pair<int, std::pair<double, double>>>> jt;  // per row: (slot, c)
  };
  const int nlm = d.n_lm, gL = d.gL;
  std::vector<std::vector<int>> sector(2 * dmoff + 1);  // shells by m + dmoff
  for (int a = 0; a < Na; ++a) sector[mv[a] + dmoff].push_back(a);
  std::vector<std::vector<Group>> groups(ks.size());
  bool fits = true;
#pragma omp parallel for schedule(dynamic)
  for (size_t kq = 0; kq < ks.size(); ++kq) {
    const int k = ks[kq], mk = mv[k];
    const std::vector<int> &js = rows[k];
    std::vector<Group> &G = groups[kq];
    std::vector<int> gidx((size_t)2 * nlm, -1);
    std::vector<int> Lk;
    for (int L = 0; L < gL; ++L)
      for (int M = -std::min(L, dmoff); M <= std::min(L, dmoff); ++M) {
        const int ml = mk - M;
        if (std::abs(ml) > dmoff / 2) continue;
        Lk.clear();
        for (int l : sector[ml + dmoff]) {
          const size_t x = ((size_t)k * Na + l) * gL + L;
          if (d.h_g0[x] != 0.0 || d.h_g2[x] != 0.0) Lk.push_back(l);
        }
        if (Lk.empty()) continue;
        const size_t lm = (size_t)L * Mext + M + dmoff;
        const int ilm = ilm_of[lm];
        const double fac = fac_of[lm];
        for (size_t q = 0; q < js.size(); ++q) {
          const int j = js[q], mi = mv[j] - M;
          if (std::abs(mi) > dmoff / 2) continue;
          for (int i : sector[mi + dmoff]) {
            const size_t xu = ((size_t)j * Na + i) * gL + L;
            const double u0 = d.h_g0[xu], u1 = d.h_g2[xu];
            if (u0 == 0.0 && u1 == 0.0) continue;
            const int par = lv[i] & 1;
            int &g = gidx[(size_t)2 * ilm + par];
            int sl = -1;
            if (g >= 0)
              for (size_t z = 0; z < G[g].slots.size(); ++z)
                if (G[g].slots[z].first == M && G[g].slots[z].second == i) {
                  sl = (int)z;
                  break;
                }
            if (sl < 0) {
              // the slot's density blocks (i,l) that pass the screen; a group
              // is opened only with its first non-empty slot
              std::vector<std::pair<int, std::pair<double, double>>> ltv;
              for (int l : Lk) {
                const int tl = mask[(size_t)i * Na + l] ? tile_of[(size_t)i * Na + l] : -1;
                if (tl < 0) continue;
                const size_t x = ((size_t)k * Na + l) * gL + L;
                ltv.push_back({tl, {d.h_g0[x], d.h_g2[x]}});
              }
              if (ltv.empty()) continue;
              if (g < 0) {
                g = (int)G.size();
                G.push_back(Group{ilm, par, {}, {},
                                  std::vector<std::vector<std::pair<int, std::pair<double, double>>>>(js.size())});
              }
              sl = (int)G[g].slots.size();
              G[g].slots.push_back({M, i});
              G[g].lt.push_back(std::move(ltv));
            }
            Group &gr = G[g];
            gr.jt[q].push_back({sl, {fac * u0, fac * u1}});
          }
        }
      }
    std::sort(G.begin(), G.end(),
              [](const Group &a, const Group &b) { return a.ilm < b.ilm || (a.ilm == b.ilm && a.par < b.par); });
    for (const Group &gr : G)
      if ((int)gr.slots.size() > K2_SLOTS) {
#pragma omp critical
        fits = false;
      }
  }
  if (!fits) return;
  // flatten
  std::vector<int> kshell, kj(1, 0), jshell, kg(1, 0), gilm, gslot(1, 0), gjt, sl(1, 0), lt, jto, jts;
  std::vector<double> ltv, jtc;
  std::vector<double> kcost(ks.size(), 0.0);
  double n_lt = 0.0, n_jt = 0.0, frad = 0.0;
  for (size_t kq = 0; kq < ks.size(); ++kq) {
    kshell.push_back(ks[kq]);
    for (int j : rows[ks[kq]]) jshell.push_back(j);
    kj.push_back((int)jshell.size());
    for (const Group &gr : groups[kq]) {
      gilm.push_back(gr.ilm);
      gjt.push_back((int)jto.size());
      for (size_t x = 0; x < gr.slots.size(); ++x) {
        for (auto &e : gr.lt[x]) {
          lt.push_back(e.first);
          ltv.push_back(e.second.first);
          ltv.push_back(e.second.second);
        }
        sl.push_back((int)lt.size());
        n_lt += (double)gr.lt[x].size();
        kcost[kq] += 2.0 * (double)gr.lt[x].size();
      }
      gslot.push_back((int)sl.size() - 1);
      jto.push_back((int)jts.size());
      for (auto &rowv : gr.jt) {
        for (auto &e : rowv) {
          jts.push_back(e.first);
          jtc.push_back(e.second.first);
          jtc.push_back(e.second.second);
        }
        jto.push_back((int)jts.size());
        n_jt += (double)rowv.size();
        if (!rowv.empty()) {
          frad += d.sym ? d.h_rad_low[gr.ilm] : d.h_rad_off[gr.ilm];
          kcost[kq] += (double)rowv.size() + 24.0;
        }
      }
    }
    kg.push_back((int)gilm.size());
  }
  // workgroups: element pairs of the launch dealt to XCDs (pair index % 8),
  // columns in decreasing cost within each pair
  std::vector<int> efs;
  for (int e = 0; e < Nel; ++e)
    for (int f = 0; f < Nel; ++f)
      if (d.sym ? e > f : e != f) efs.push_back(e * Nel + f);
  std::vector<int> korder(ks.size());
  for (size_t q = 0; q < ks.size(); ++q) korder[q] = (int)q;
  std::sort(korder.begin(), korder.end(), [&](int a, int b) { return kcost[a] > kcost[b]; });
  std::vector<std::vector<int>> per_xcd(8);
  for (size_t p = 0; p < efs.size(); ++p)
    for (int kq : korder) {
      auto &v = per_xcd[p % 8];
      v.push_back(kq);
      v.push_back(efs[p]);
    }
  size_t mx = 0;
  for (auto &v : per_xcd) mx = std::max(mx, v.size() / 2);
  std::vector<int> wg(mx * 8 * 2, -1);
  for (int x = 0; x < 8; ++x)
    for (size_t q = 0; q < per_xcd[x].size() / 2; ++q) {
      wg[(q * 8 + x) * 2] = per_xcd[x][2 * q];
      wg[(q * 8 + x) * 2 + 1] = per_xcd[x][2 * q + 1];
    }
  if (lt.empty()) {
    lt.push_back(0);
    ltv.assign(2, 0.0);
  }
  if (jts.empty()) {
    jts.push_back(0);
    jtc.assign(2, 0.0);
  }
  if (gilm.empty()) {
    gilm.push_back(0);
    gjt.push_back(0);
  }
  d.k2_wg.upload(wg, s);
  d.k2_kshell.upload(kshell, s);
  d.k2_kj.upload(kj, s);
  d.k2_jshell.upload(jshell, s);
  d.k2_kg.upload(kg, s);
  d.k2_gilm.upload(gilm, s);
  d.k2_gslot.upload(gslot, s);
  d.k2_gjt.upload(gjt, s);
  d.k2_sl.upload(sl, s);
  d.k2_lt.upload(lt, s);
  d.k2_ltv.upload(ltv, s);
  d.k2_jto.upload(jto, s);
  d.k2_jts.upload(jts, s);
  d.k2_jtc.upload(jtc, s);
  d.n_wg_k2 = (int)(mx * 8);
  // per element pair: 4 flops per lt entry (two channels) and 8 per jt entry
  // (four channels) on every tile entry, plus the radial products
  double sum_off = 0.0;
  for (int e = 0; e < Nel; ++e)
    for (int f = 0; f < Nel; ++f)
      if (d.sym ? e > f : e != f) sum_off += (double)S.rad.ni[e] * S.rad.ni[f];
  d.flops_off_k2 = frad + (4.0 * n_lt + 8.0 * n_jt) * sum_off;
  d.k2 = std::getenv("HFGPU_EXCHANGE_TWOSTEP") == nullptr || std::atoi(std::getenv("HFGPU_EXCHANGE_TWOSTEP")) != 0;
}

// Exchange coupling lists for the density blocks that pass the norm screen
// (|P_il|_F >= 10 eps, basis.cpp:1858-1875).  For each output block (j,k):
// runs of terms per (L,|M|), each term a density block (i,l) with
// mi - ml = mj - mk and the four channel weights LMfac * {mji mkl, -mji gkl,
// -gji mkl, gji gkl} (basis.cpp:1905-1960).  Rebuilt only when the screening
// pattern changes.
void build_exchange_lists(hf_diatomic *h, hipStream_t s) {
  const DiatomicSetup &S = h->host;
  DiatomicDevice &d = h->dev;
  const int Na = S.Nang(), nlm = d.n_lm;
  const auto &lv = S.lval;
  const auto &mv = S.mval;
  std::vector<unsigned char> mask((size_t)Na * Na);
  const double thr = 10.0 * 2.220446049250313e-16;
  for (size_t q = 0; q < mask.size(); ++q) mask[q] = d.h_bnorm.p[q] >= thr ? 1 : 0;
  // the symmetric exchange needs the mirror of every coupled block coupled
  if (d.sym)
    for (int i = 0; i < Na; ++i)
      for (int l = 0; l < i; ++l) {
        const unsigned char m = mask[(size_t)i * Na + l] | mask[(size_t)l * Na + i];
        mask[(size_t)i * Na + l] = mask[(size_t)l * Na + i] = m;
      }
  if (!d.screen_mask.empty() && mask == d.screen_mask && d.sym == d.list_sym) return;
  int mabs = 0;
  for (int m : mv) mabs = std::max(mabs, std::abs(m));
  const int dmoff = 2 * mabs;
  std::vector<std::vector<int>> bydm(4 * mabs + 1);
  for (int i = 0; i < Na; ++i)
    for (int l = 0; l < Na; ++l)
      if (mask[(size_t)i * Na + l]) bydm[mv[i] - mv[l] + dmoff].push_back(i * Na + l);
  // (L, M) -> ilm and LMfac lookups
  const int Lext = d.gL, Mext = 2 * dmoff + 1;
  std::vector<int> ilm_of((size_t)Lext * Mext, -1);
  std::vector<double> fac_of((size_t)Lext * Mext, 0.0);
  for (auto &e : S.LM_map)
    if (e.first < Lext && std::abs(e.second) <= dmoff) {
      ilm_of[(size_t)e.first * Mext + e.second + dmoff] = S.lmind(e.first, e.second);
      fac_of[(size_t)e.first * Mext + e.second + dmoff] = S.LMfac(e.first, e.second);
    }
  struct Term {
    int ilm, il;
    double f[4];
  };
  const int njk = Na * Na;
  std::vector<std::vector<Term>> per(njk);
#pragma omp parallel for schedule(dynamic, 4)
  for (int jk = 0; jk < njk; ++jk) {
    const int j = jk / Na, k = jk - j * Na;
    const int lj = lv[j], mj = mv[j], lk = lv[k], mk = mv[k];
    if (h->absm && (mj < 0 || mk < 0)) continue;
    const int dm = mj - mk;
    if (std::abs(dm) > 2 * mabs) continue;
    std::vector<Term> &out = per[jk];
    for (int il : bydm[dm + dmoff]) {
      const int i = il / Na, l = il - i * Na;
      const int li = lv[i], mi = mv[i], ll = lv[l], ml = mv[l], M = mj - mi;
      const int Lmin = std::max(std::max(std::abs(li - lj), std::abs(lk - ll)) - 2, std::abs(M));
      const int Lmax = std::min(li + lj, lk + ll) + 2;
      const size_t gji = ((size_t)j * Na + i) * Lext, gkl = ((size_t)k * Na + l) * Lext;
      for (int L = Lmin; L <= Lmax; ++L) {
        const double mji = d.h_g0[gji + L], g_ji = d.h_g2[gji + L];
        const double mkl = d.h_g0[gkl + L], g_kl = d.h_g2[gkl + L];
        const double c00 = mji * mkl, c02 = -mji * g_kl, c20 = -g_ji * mkl, c22 = g_ji * g_kl;
        if (c00 == 0.0 && c02 == 0.0 && c20 == 0.0 && c22 == 0.0) continue;
        const size_t lm = (size_t)L * Mext + M + dmoff;
        const double fac = fac_of[lm];
        out.push_back({ilm_of[lm], il, {fac * c00, fac * c02, fac * c20, fac * c22}});
      }
    }
    std::stable_sort(out.begin(), out.end(), [](const Term &a, const Term &b) { return a.ilm < b.ilm; });
  }
  std::vector<int> act, lstart(1, 0), rilm, rstart(1, 0), til;
  std::vector<double> tf;
  // flops as executed: in the symmetric mode the cross-element launch covers
  // the pairs e > f and the in-element one the blocks j <= k
  double frad_in = 0.0, frad_off = 0.0;
  int64_t nterm = 0, nterm_in = 0;
  for (int jk = 0; jk < njk; ++jk) {
    const std::vector<Term> &v = per[jk];
    if (v.empty()) continue;
    const bool in_done = !d.sym || jk / Na <= jk % Na;
    act.push_back(jk);
    for (size_t a = 0; a < v.size();) {
      size_t b = a;
      while (b < v.size() && v[b].ilm == v[a].ilm) {
        til.push_back(v[b].il);
        tf.insert(tf.end(), v[b].f, v[b].f + 4);
        ++b;
      }
      rilm.push_back(v[a].ilm);
      rstart.push_back((int)til.size());
      if (in_done) frad_in += d.h_rad_in[v[a].ilm];
      frad_off += d.sym ? d.h_rad_low[v[a].ilm] : d.h_rad_off[v[a].ilm];
      a = b;
    }
    lstart.push_back((int)rilm.size());
    nterm += (int64_t)v.size();
    if (in_done) nterm_in += (int64_t)v.size();
  }
  d.n_act = (int)act.size();
  d.n_runs = (int64_t)rilm.size();
  d.n_terms = nterm;
  // Gaunt-weighted density sums: 4 channels x 2 flops per term and tile entry
  double sum_in = 0.0, sum_off = 0.0;
  for (int e = 0; e < S.Nel(); ++e) {
    sum_in += (double)S.rad.ni[e] * S.rad.ni[e];
    for (int f = 0; f < S.Nel(); ++f)
      if (d.sym ? e > f : e != f) sum_off += (double)S.rad.ni[e] * S.rad.ni[f];
  }
  d.flops_in = frad_in + 8.0 * (double)nterm_in * sum_in;
  d.flops_off = frad_off + 8.0 * (double)nterm * sum_off;
  d.flops_k_angular = 8.0 * ((double)nterm_in * sum_in + (double)nterm * sum_off);
  d.flops_k_radial = frad_in + frad_off;
  // dense in-element path: runs grouped by (L,|M|); one workgroup per
  // (ilm, element, 32 runs), groups dealt round-robin to the 8 XCDs so the
  // workgroups sharing one operator table run on the same L2
  if (d.dense_in && d.n_runs > 0) {
    const int Nel = S.Nel();
    std::vector<int> rs(nlm + 1, 0), runs(rilm.size());
    for (int r : rilm) ++rs[r + 1];
    for (int q = 0; q < nlm; ++q) rs[q + 1] += rs[q];
    std::vector<int> fill(rs.begin(), rs.end() - 1);
    for (size_t r = 0; r < rilm.size(); ++r) runs[fill[rilm[r]]++] = (int)r;
    std::vector<std::vector<int>> per_xcd(8);
    int grp = 0;
    for (int q = 0; q < nlm; ++q) {
      const int nr = rs[q + 1] - rs[q];
      if (!nr) continue;
      for (int e = 0; e < Nel; ++e, ++grp)
        for (int c0 = 0; c0 < nr; c0 += 32) {
          auto &v = per_xcd[grp % 8];
          v.push_back(q);
          v.push_back(e);
          v.push_back(c0);
        }
    }
    size_t mx = 0;
    for (auto &v : per_xcd) mx = std::max(mx, v.size() / 3);
    std::vector<int> wg(mx * 8 * 3, -1);
    for (int x = 0; x < 8; ++x)
      for (size_t j = 0; j < per_xcd[x].size() / 3; ++j)
        for (int u = 0; u < 3; ++u) wg[(j * 8 + x) * 3 + u] = per_xcd[x][3 * j + u];
    d.n_wg_in = (int)(mx * 8);
    d.ilm_rstart.upload(rs, s);
    d.ilm_runs.upload(runs, s);
    d.wg_in.upload(wg, s);
    d.Ypart.alloc((size_t)d.n_runs * Nel * TILE2);
    // the dense operator replaces 12 r n^3 by 8 n^4 per (run, element)
    double fr = 0.0;
    for (int e = 0; e < Nel; ++e) fr += 8.0 * std::pow((double)S.rad.ni[e], 4);
    d.flops_in = (double)d.n_runs * fr + 8.0 * (double)nterm * sum_in;
  }
  // in-element workgroups of the LDS-resident kernel: runs grouped by ilm,
  // batches of KL_BATCH per (ilm, element), element e's batches on XCD e % 8
  d.n_wg_kl = 0;
  d.n_wg_op = 0;
  if (!d.dense_in && d.max_rank <= KL_RANK_MAX && !rilm.empty()) {
    const int Nel = S.Nel(), nr = (int)rilm.size();
    std::vector<int> run_act(nr), cnt(nlm + 1, 0), runs;
    std::vector<unsigned char> keep(nr, 1);
    for (size_t a = 0; a < act.size(); ++a)
      for (int r = lstart[a]; r < lstart[a + 1]; ++r) {
        run_act[r] = (int)a;
        if (d.sym && act[a] / Na > act[a] % Na) keep[r] = 0;  // the mirror of block (k, j) covers it
      }
    for (int r = 0; r < nr; ++r)
      if (keep[r]) ++cnt[rilm[r] + 1];
    for (int q = 0; q < nlm; ++q) cnt[q + 1] += cnt[q];
    runs.resize(cnt[nlm]);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int r = 0; r < nr; ++r)
      if (keep[r]) runs[fill[rilm[r]]++] = r;
    if (runs.empty()) runs.push_back(0);
    std::vector<std::vector<int>> per_xcd(8);
    for (int q = 0; q < nlm; ++q)
      for (int e = 0; e < Nel; ++e)
        for (int b = cnt[q]; b < cnt[q + 1]; b += KL_BATCH) {
          auto &v = per_xcd[e % 8];
          v.insert(v.end(), {q, e, b, std::min(KL_BATCH, cnt[q + 1] - b)});
        }
    size_t mx = 0;
    for (auto &v : per_xcd) mx = std::max(mx, v.size() / 4);
    std::vector<int> wg(mx * 8 * 4, -1);
    for (int x = 0; x < 8; ++x)
      for (size_t j = 0; j < per_xcd[x].size() / 4; ++j)
        for (int u = 0; u < 4; ++u) wg[(j * 8 + x) * 4 + u] = per_xcd[x][4 * j + u];
    d.wg_kl.upload(wg, s);
    d.kl_runs.upload(runs, s);
    d.run_act.upload(run_act, s);
    d.n_wg_kl = (int)(mx * 8);
    // the matrix-core operator form: OP_RUNS runs per workgroup, the groups
    // of one (ilm, element) dealt to one XCD so they share its operator in L2
    if (d.op_in) {
      std::vector<std::vector<int>> px(8);
      int grp = 0;
      for (int q = 0; q < nlm; ++q)
        for (int e = 0; e < Nel; ++e) {
          if (cnt[q + 1] == cnt[q]) continue;
          auto &v = px[grp++ % 8];
          for (int b = cnt[q]; b < cnt[q + 1]; b += OP_RUNS)
            v.insert(v.end(), {q, e, b, std::min(OP_RUNS, cnt[q + 1] - b)});
        }
      size_t mo = 0;
      for (auto &v : px) mo = std::max(mo, v.size() / 4);
      std::vector<int> wo(mo * 8 * 4, -1);
      for (int xc = 0; xc < 8; ++xc)
        for (size_t j = 0; j < px[xc].size() / 4; ++j)
          for (int u = 0; u < 4; ++u) wo[(j * 8 + xc) * 4 + u] = px[xc][4 * j + u];
      d.wg_op.upload(wo, s);
      d.n_wg_op = (int)(mo * 8);
    }
  }
  if (act.empty()) act.push_back(0);
  if (rilm.empty()) rilm.push_back(0);
  if (til.empty()) {
    til.push_back(0);
    tf.assign(4, 0.0);
  }
  d.act_jk.upload(act, s);
  d.act_lstart.upload(lstart, s);
  d.lrun_ilm.upload(rilm, s);
  d.lrun_tstart.upload(rstart, s);
  d.term_il.upload(til, s);
  {
    std::vector<int4> tpos(til.size());
    for (size_t q = 0; q < til.size(); ++q) {
      const int i = til[q] / Na, l = til[q] % Na;
      tpos[q] = make_int4((int)S.poff[i], (int)(S.poff[l] - S.skip(l)), S.skip(i), S.skip(l));
    }
    d.term_pos.upload(tpos, s);
  }
  {
    // source blocks referenced by the terms, and each term's tile among them
    std::vector<int> id((size_t)Na * Na, -1), src, tt(til.size());
    for (size_t q = 0; q < til.size(); ++q) {
      int &x = id[til[q]];
      if (x < 0) {
        x = (int)src.size();
        src.push_back(til[q]);
      }
      tt[q] = x;
    }
    const double bytes = (double)src.size() * S.Nel() * S.Nel() * TILE2 * sizeof(double);
    d.tiled = bytes <= exchange_tile_budget(d.ptile.n * sizeof(double) + d.ptile32.n * sizeof(float));
    d.n_src = (int)src.size();
    d.src_blk.upload(src, s);
    d.term_tile.upload(tt, s);
    build_twostep(h, act, id, mask, ilm_of, fac_of, Mext, dmoff, s);
  }
  d.term_f.upload(tf, s);
  hip_check(hipStreamSynchronize(s), "exchange list upload");
  d.screen_mask = mask;
  d.list_sym = d.sym;
  ++d.list_gen;
}

void ensure_ready(hf_diatomic *h) {
  if (!h->ctx) throw Error(HF_ESTATE, "handle was created without a device context");
  if (!h->host.have_tei) throw Error(HF_ESTATE, "Primitive teis have not been computed!");
  if (!h->uploaded) diatomic_upload(h);
}

void exchange_launch(hf_diatomic *h, const double *P, int64_t ldp, double *K, int64_t ldk, bool fp32 = false) {
  DiatomicDevice &d = h->dev;
  hipStream_t s = h->ctx->cur;
  block_stats_launch(P, ldp, d.Nrad, d.Nang, d.skip.p, d.poff.p, d.stats.p, d.h_stats.p, s);
  hip_check(hipStreamSynchronize(s), "block statistics sync");
  const bool symmetric = block_stats_reduce(d.h_stats.p, d.Nrad, d.Nang, d.h_bnorm.p);
  d.sym = !d.dense_in && symmetric_exchange_allowed() && symmetric;
  build_exchange_lists(h, s);
  diatomic_exchange_kernels(d, P, ldp, K, ldk, s, &h->ctx->timer, fp32);
}

hf_diatomic *checked(hf_diatomic *h) {
  need(h != nullptr, "NULL handle");
  need(h->ctx != nullptr, "handle was created without a device context");
  return h;
}

}  // namespace

extern "C" {

int hf_diatomic_create(hf_ctx *ctx, int Z1, int Z2, double Rhalf, int nnodes, int nquad, const double *bval,
                       int nbval, const int *lval, const int *mval, int nang, hf_diatomic **out) {
  return guarded([&] {
    need(bval && lval && mval && out, "NULL argument");
    need(nbval >= 2, "need at least one element (nbval >= 2)");
    need(nang >= 1, "need at least one angular shell");
    need(nnodes >= 2 && nnodes <= TILE, "nnodes must be in [2, 16]");
    need(Z1 >= 0 && Z2 >= 0, "nuclear charges must be non-negative");
    need(Rhalf > 0.0, "Rhalf must be positive");
    for (int i = 0; i + 1 < nbval; ++i) need(bval[i + 1] > bval[i], "element boundaries must increase");
    need(bval[0] == 0.0, "the first element boundary must be mu=0");
    std::unique_ptr<hf_diatomic> h(new hf_diatomic);
    h->ctx = ctx;
    h->host.build(Z1, Z2, Rhalf, nnodes, nquad > 0 ? nquad : 5 * nnodes, std::vector<double>(bval, bval + nbval),
                  std::vector<int>(lval, lval + nang), std::vector<int>(mval, mval + nang));
    *out = h.release();
  });
}

int hf_diatomic_destroy(hf_diatomic *h) {
  return guarded([&] {
    if (!h) return;
    if (h->ctx) {
      DeviceGuard g(h->ctx->device);
      (void)hipStreamSynchronize(h->ctx->cur);
    }
    delete h;
  });
}

int hf_diatomic_sizes(const hf_diatomic *h, int64_t out[7]) {
  return guarded([&] {
    need(h && out, "NULL argument");
    out[0] = h->host.Nbf();
    out[1] = h->host.Ndummy();
    out[2] = h->host.Nrad();
    out[3] = h->host.Nang();
    out[4] = h->host.Nel();
    out[5] = (int64_t)h->host.lm_map.size();
    out[6] = (int64_t)h->host.LM_map.size();
  });
}

int hf_diatomic_pure_index(const hf_diatomic *h, int64_t *idx) {
  return guarded([&] {
    need(h && idx, "NULL argument");
    const DiatomicSetup &S = h->host;
    int64_t q = 0;
    for (int a = 0; a < S.Nang(); ++a)
      for (int r = S.skip(a); r < S.Nrad(); ++r) idx[q++] = (int64_t)a * S.Nrad() + r;
  });
}

int hf_diatomic_maps(const hf_diatomic *h, int *lm, int *LM) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    const DiatomicSetup &S = h->host;
    if (lm)
      for (size_t i = 0; i < S.lm_map.size(); ++i) {
        lm[2 * i] = S.lm_map[i].first;
        lm[2 * i + 1] = S.lm_map[i].second;
      }
    if (LM)
      for (size_t i = 0; i < S.LM_map.size(); ++i) {
        LM[2 * i] = S.LM_map[i].first;
        LM[2 * i + 1] = S.LM_map[i].second;
      }
  });
}

int hf_diatomic_set_absm_symmetric(hf_diatomic *h, int sym) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    if (h->absm != (sym != 0)) {
      h->absm = sym != 0;
      h->uploaded = false;
    }
  });
}

int hf_diatomic_compute_tei(hf_diatomic *h, int exchange) {
  (void)exchange;
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    h->host.compute_tei();
    h->uploaded = false;
  });
}

int hf_diatomic_compute_tei_device(hf_diatomic *h, double *seconds) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    need(h->ctx != nullptr, "handle was created without a device context");
    DeviceGuard g(h->ctx->device);
    diatomic_tei_device(h->host, h->ctx->cur, seconds);
    h->uploaded = false;
  });
}

int hf_diatomic_get_cache(const hf_diatomic *h, int kind, int ilm, int iel, int dims[2], double *data) {
  return guarded([&] {
    need(h && dims, "NULL argument");
    const DiatomicSetup &S = h->host;
    need(kind >= 0 && kind <= 5, "kind must be in 0..5");
    need(ilm >= 0 && ilm < (int)S.lm_map.size() && iel >= 0 && iel < S.Nel(), "(ilm, iel) out of range");
    if (!S.have_tei) throw Error(HF_ESTATE, "Primitive teis have not been computed!");
    const size_t x = (size_t)ilm * S.Nel() + iel;
    if (kind == 5) {
      dims[0] = (int)S.cdS[x].size();
      dims[1] = 1;
      if (data && dims[0]) std::memcpy(data, S.cdS[x].data(), sizeof(double) * S.cdS[x].size());
      return;
    }
    const Blk &b = kind == 0 ? S.P0[x] : kind == 1 ? S.P2[x] : kind == 2 ? S.Q0[x] : kind == 3 ? S.Q2[x] : S.cdB[x];
    dims[0] = b.r;
    dims[1] = b.c;
    if (data && !b.v.empty()) std::memcpy(data, b.v.data(), sizeof(double) * b.v.size());
  });
}

int hf_diatomic_one_electron(const hf_diatomic *h, int which, double *out, int64_t ld) {
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

int hf_diatomic_work(hf_diatomic *h, double out[4]) {
  return guarded([&] {
    checked(h);
    out[0] = h->dev.flops_k_angular;
    out[1] = h->dev.flops_k_radial;
    out[2] = (double)h->dev.n_terms;
    out[3] = (double)h->dev.n_act;
  });
}

int hf_diatomic_one_electron_dev(const hf_diatomic *h, int which, double *dOut, int64_t ld) {
  return guarded([&] {
    need(h != nullptr, "NULL handle");
    need(h->ctx != nullptr, "handle was created without a device context");
    need(which >= 0 && which <= 2, "which must be 0 (overlap), 1 (kinetic) or 2 (nuclear)");
    DeviceGuard g(h->ctx->device);
    check_matrix(dOut, ld, h->host.Nbf(), "out");
    BlockList b;
    h->host.one_electron_blocks(which, b);
    scatter_blocks_launch(dOut, ld, h->host.Nbf(), b.meta, b.vals, h->ctx->cur);
  });
}

int hf_diatomic_coulomb_dev(hf_diatomic *h, const double *dP, int64_t ldp, double *dJ, int64_t ldj) {
  return guarded([&] {
    checked(h);
    DeviceGuard g(h->ctx->device);
    ensure_ready(h);
    check_matrix(dP, ldp, h->host.Nbf(), "P");
    check_matrix(dJ, ldj, h->host.Nbf(), "J");
    diatomic_coulomb_launch(h->dev, dP, ldp, dJ, ldj, h->ctx->cur, &h->ctx->timer);
  });
}

int hf_diatomic_exchange_dev(hf_diatomic *h, const double *dP, int64_t ldp, double *dK, int64_t ldk) {
  return guarded([&] {
    checked(h);
    DeviceGuard g(h->ctx->device);
    ensure_ready(h);
    check_matrix(dP, ldp, h->host.Nbf(), "P");
    check_matrix(dK, ldk, h->host.Nbf(), "K");
    exchange_launch(h, dP, ldp, dK, ldk);
  });
}

int hf_diatomic_coulomb(hf_diatomic *h, const double *P, int64_t ldp, double *J, int64_t ldj) {
  return guarded([&] {
    checked(h);
    DeviceGuard g(h->ctx->device);
    ensure_ready(h);
    host_roundtrip(h->ctx->cur, h->host.Nbf(), h->dev.hostP, h->dev.hostOut, P, ldp, J, ldj,
                   [&](const double *dp, int64_t lp, double *dout, int64_t lo) {
                     diatomic_coulomb_launch(h->dev, dp, lp, dout, lo, h->ctx->cur, &h->ctx->timer);
                   });
  });
}

int hf_diatomic_exchange_fp32_dev(hf_diatomic *h, const double *dP, int64_t ldp, double *dK, int64_t ldk) {
  return guarded([&] {
    checked(h);
    DeviceGuard g(h->ctx->device);
    ensure_ready(h);
    check_matrix(dP, ldp, h->host.Nbf(), "P");
    check_matrix(dK, ldk, h->host.Nbf(), "K");
    exchange_launch(h, dP, ldp, dK, ldk, true);
  });
}

int hf_diatomic_exchange(hf_diatomic *h, const double *P, int64_t ldp, double *K, int64_t ldk) {
  return guarded([&] {
    checked(h);
    DeviceGuard g(h->ctx->device);
    ensure_ready(h);
    host_roundtrip(h->ctx->cur, h->host.Nbf(), h->dev.hostP, h->dev.hostOut, P, ldp, K, ldk,
                   [&](const double *dp, int64_t lp, double *dout, int64_t lo) { exchange_launch(h, dp, lp, dout, lo); });
  });
}

}  // extern "C"
