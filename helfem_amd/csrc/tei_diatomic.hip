// Diatomic two-electron setup on the device: the two-channel in-element
// kernel W of every ((L,|M|), element) and its sign-aware pivoted Cholesky
// factor (TwoDBasis::compute_tei, src/diatomic/basis.cpp:1382-1547; the
// kernel is quadrature::twoe_integral, src/diatomic/quadrature.cpp:162-258,
// symmetrised as basis.cpp:1470-1490; the factorisation basis.cpp:1492-1537).
//
// The host keeps the cheap parts: the disjoint P/Q integrals, the rule size
// of each element (converged on the largest (L,|M|), basis.cpp:1334-1380)
// and the Q_L^|M| values at the outer nodes.  Per group of tasks ((L,|M|),
// element) sharing a rule size n:
//   k_dw_inner   WG per (task, segment p): for l = 0, 2 the segment integrals
//                sum_i slen w_i sinh cosh^l P_L^|M|(cosh) B_a B_b over the rule
//                mapped to [mu_{p-1}, mu_p], P by its upward recurrence per point;
//   k_dw_cumul   running sums over the segments (the cumulative inner integral);
//   k_dw_bfprod  B_a B_b at the outer nodes, per element;
//   k_dw_outer   WG per (task, 16x16 tile (u,v) of the (ab) x (cd) products): the
//                eight outer contractions o_kl(u,v), o_kl(v,u) (k, l = 0, 2) with
//                weights len w_i sinh cosh^k Q, and the kernel written directly
//                in its symmetrised two-channel form
//                W = [[T00, -T02], [-T02', T22]], T_kl = o_kl + o_lk';
//   k_dw_chol    one workgroup per task (a thread per row, N = 2 nb^2 <= 512):
//                largest |residual diagonal| pivot, the sign of the pivot kept as
//                sigma_p, stop at 1e-12 of the largest initial |diagonal|.
// The factors and signs come back into the host caches.  Floating-point
// contraction is off in this file, as in the host setup, so the recurrences
// and sums round as there; the sums run in the same index order.
#include "capi_common.hpp"
#include <chrono>
#include <map>

#pragma clang fp contract(off)

using namespace hfg;

namespace {

constexpr int DW_NODES = 16;
constexpr int DW_NB2S = 256;     // stride of a (ab) pair index
constexpr int DW_NMAX = 512;     // 2 nb^2 rows of a kernel
constexpr size_t DW_WSTRIDE = (size_t)DW_NMAX * DW_NMAX;
constexpr int DW_CHUNK = 256;

struct DwElem {
  double lo, hi;
  int flo, nb;  // first shape function, number of functions
};

struct DwArgs {
  const double *nodes;
  int nnodes;
  const DwElem *el;
  const int4 *task;  // (L, |M|, element, task slot)
  int ntask;
  const double *x, *w;  // Chebyshev rule of this group
  int n;
  const double *qtab;   // [task slot][n]: Q_L^|M|(cosh mu_i)
  double *inner;        // [task slot][2][n][DW_NB2S]
  double *bfp;          // [element][n][DW_NB2S]
  double *W;            // [task slot][DW_WSTRIDE]
};

__device__ inline double lagrange(double xv, const double *nodes, int N, int f) {
  const double xf = nodes[f];
  double v = 1.0;
  for (int p = 0; p < N; ++p) {
    if (p == f) continue;
    v *= (xv - nodes[p]) / (xf - nodes[p]);
  }
  return v;
}

// P_L^M(x) for x > 1 by the host recurrence (diagonal seeds, then upward in
// L); values that are not normal numbers are flushed to zero as the host
// cache does.
__device__ inline double legendre_p(int L, int M, double x) {
  if (L < M) return 0.0;
  const double s = fabs(x) <= 1.0 ? 1.0 : -1.0;
  const double w = sqrt(fmax(s * (1.0 - x * x), 0.0));
  double pmm = 1.0;
  for (int m = 1; m <= M; ++m) pmm = -s * double(2 * m - 1) * w * pmm;
  double v;
  if (L == M) {
    v = pmm;
  } else {
    double p0 = pmm, p1 = double(2 * M + 1) * x * pmm;
    for (int l = M + 1; l < L; ++l) {
      const double p2 = (double(2 * l + 1) * x * p1 - double(l + M) * p0) / double(l + 1 - M);
      p0 = p1;
      p1 = p2;
    }
    v = p1;
  }
  return (v != 0.0 && !isnormal(v)) ? 0.0 : v;
}

__global__ void __launch_bounds__(256) k_dw_inner(DwArgs A) {
  const int tk = blockIdx.x / A.n, p = blockIdx.x - tk * A.n;
  const int4 T = A.task[tk];
  const DwElem E = A.el[T.z];
  const int t = threadIdx.x, nb = E.nb, nb2 = nb * nb;
  const double mid = 0.5 * (E.hi + E.lo), len = 0.5 * (E.hi - E.lo);
  const double a = p == 0 ? E.lo : mid + len * A.x[p - 1], b = mid + len * A.x[p];
  const double smid = 0.5 * (b + a), slen = 0.5 * (b - a);
  __shared__ double nodes[DW_NODES];
  __shared__ double bf[DW_CHUNK][DW_NODES];
  __shared__ double wp[2][DW_CHUNK];
  if (t < A.nnodes) nodes[t] = A.nodes[t];
  const int k = t % nb, l = t / nb;
  const bool valid = t < nb2;
  double s0 = 0.0, s2 = 0.0;
  for (int i0 = 0; i0 < A.n; i0 += DW_CHUNK) {
    __syncthreads();
    const int i = i0 + t;
    if (i < A.n) {
      const double m = smid + slen * A.x[i];
      const double ch = cosh(m), sh = sinh(m);
      const double P = legendre_p(T.x, T.y, ch);
      double v0 = slen * A.w[i];
      v0 *= sh;
      double v2 = v0;
      v0 *= P;
      v2 *= ch * ch;  // cosh^2 by repeated squaring is one product
      v2 *= P;
      wp[0][t] = v0;
      wp[1][t] = v2;
      const double xs = (m - mid) / len;
      for (int f = 0; f < nb; ++f) bf[t][f] = lagrange(xs, nodes, A.nnodes, E.flo + f);
    }
    __syncthreads();
    const int cnt = min(DW_CHUNK, A.n - i0);
    if (valid)
      for (int ii = 0; ii < cnt; ++ii) {
        s0 += (bf[ii][k] * wp[0][ii]) * bf[ii][l];
        s2 += (bf[ii][k] * wp[1][ii]) * bf[ii][l];
      }
  }
  if (valid) {
    double *o = A.inner + (((size_t)tk * 2) * A.n + p) * DW_NB2S;
    o[k + nb * l] = s0;
    o[(size_t)A.n * DW_NB2S + k + nb * l] = s2;
  }
}

__global__ void __launch_bounds__(256) k_dw_cumul(DwArgs A) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int tk = (int)(g / (2 * DW_NB2S)), r = (int)(g - (int64_t)tk * 2 * DW_NB2S);
  if (tk >= A.ntask) return;
  const int ch = r / DW_NB2S, q = r - ch * DW_NB2S;
  const DwElem E = A.el[A.task[tk].z];
  if (q >= E.nb * E.nb) return;
  double *col = A.inner + (((size_t)tk * 2 + ch) * A.n) * DW_NB2S + q;
  double acc = 0.0;
  for (int p = 0; p < A.n; ++p) {
    acc += col[(size_t)p * DW_NB2S];
    col[(size_t)p * DW_NB2S] = acc;
  }
}

__global__ void __launch_bounds__(256) k_dw_bfprod(DwArgs A, int e0) {
  const int e = e0 + blockIdx.x / A.n, i = blockIdx.x % A.n;
  const DwElem E = A.el[e];
  const int t = threadIdx.x, nb = E.nb;
  __shared__ double nodes[DW_NODES], bv[DW_NODES];
  if (t < A.nnodes) nodes[t] = A.nodes[t];
  __syncthreads();
  if (t < nb) bv[t] = lagrange(A.x[i], nodes, A.nnodes, E.flo + t);
  __syncthreads();
  if (t < nb * nb) {
    const int a = t / nb, b = t - a * nb;
    A.bfp[((size_t)e * A.n + i) * DW_NB2S + t] = bv[a] * bv[b];
  }
}

__global__ void __launch_bounds__(256) k_dw_outer(DwArgs A) {
  constexpr int NT16 = 16;
  const int tk = blockIdx.x / (NT16 * NT16), tt = blockIdx.x - tk * NT16 * NT16;
  const int tu = tt % NT16, tv = tt / NT16;
  const int4 T = A.task[tk];
  const DwElem E = A.el[T.z];
  const int nb2 = E.nb * E.nb, N = 2 * nb2;
  if (tu * 16 >= nb2 || tv * 16 >= nb2) return;
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const double mid = 0.5 * (E.hi + E.lo), len = 0.5 * (E.hi - E.lo);
  const double *bp = A.bfp + (size_t)T.z * A.n * DW_NB2S;
  const double *in0 = A.inner + ((size_t)tk * 2) * A.n * DW_NB2S, *in2 = in0 + (size_t)A.n * DW_NB2S;
  const double *Q = A.qtab + (size_t)tk * A.n;
  __shared__ double bu[32][17], bv[32][17], iu0[32][17], iv0[32][17], iu2[32][17], iv2[32][17];
  __shared__ double wq[2][32];
  double o00uv = 0, o00vu = 0, o02uv = 0, o02vu = 0, o20uv = 0, o20vu = 0, o22uv = 0, o22vu = 0;
  for (int i0 = 0; i0 < A.n; i0 += 32) {
    __syncthreads();
    for (int q = t; q < 512; q += 256) {
      const int ii = q >> 4, cc = q & 15, i = i0 + ii;
      const int cu = tu * 16 + cc, cv = tv * 16 + cc;
      const bool ok = i < A.n;
      const size_t ro = (size_t)i * DW_NB2S;
      bu[ii][cc] = (ok && cu < nb2) ? bp[ro + cu] : 0.0;
      bv[ii][cc] = (ok && cv < nb2) ? bp[ro + cv] : 0.0;
      iu0[ii][cc] = (ok && cu < nb2) ? in0[ro + cu] : 0.0;
      iv0[ii][cc] = (ok && cv < nb2) ? in0[ro + cv] : 0.0;
      iu2[ii][cc] = (ok && cu < nb2) ? in2[ro + cu] : 0.0;
      iv2[ii][cc] = (ok && cv < nb2) ? in2[ro + cv] : 0.0;
    }
    if (t < 32) {
      const int i = i0 + t;
      double v0 = 0.0, v2 = 0.0;
      if (i < A.n) {
        const double mu = mid + len * A.x[i];
        const double ch = cosh(mu), sh = sinh(mu);
        v0 = len * A.w[i];
        v0 *= sh;
        v2 = v0;
        v2 *= ch * ch;
        v0 *= Q[i];
        v2 *= Q[i];
      }
      wq[0][t] = v0;
      wq[1][t] = v2;
    }
    __syncthreads();
    const int cnt = min(32, A.n - i0);
    for (int ii = 0; ii < cnt; ++ii) {
      const double pu0 = bu[ii][tx] * wq[0][ii], pv0 = bv[ii][ty] * wq[0][ii];
      const double pu2 = bu[ii][tx] * wq[1][ii], pv2 = bv[ii][ty] * wq[1][ii];
      o00uv += pu0 * iv0[ii][ty];
      o00vu += pv0 * iu0[ii][tx];
      o02uv += pu0 * iv2[ii][ty];
      o02vu += pv0 * iu2[ii][tx];
      o20uv += pu2 * iv0[ii][ty];
      o20vu += pv2 * iu0[ii][tx];
      o22uv += pu2 * iv2[ii][ty];
      o22vu += pv2 * iu2[ii][tx];
    }
  }
  const int u = tu * 16 + tx, v = tv * 16 + ty;
  if (u < nb2 && v < nb2) {
    double *Wt = A.W + (size_t)T.w * DW_WSTRIDE;
    Wt[u + (size_t)N * v] = o00uv + o00vu;
    Wt[u + (size_t)N * (nb2 + v)] = -(o02uv + o20vu);
    Wt[(nb2 + u) + (size_t)N * v] = -(o02vu + o20uv);
    Wt[(nb2 + u) + (size_t)N * (nb2 + v)] = o22uv + o22vu;
  }
}

struct DwCholArgs {
  const DwElem *el;
  const int4 *task;
  const double *W;
  double *B;      // [task slot][DW_WSTRIDE]: column p at p * DW_NMAX
  double *sigma;  // [task slot][DW_NMAX]
  int *rank;      // [task slot]
};

// (|d|, index, d): the larger |d| wins, ties to the lower index (the host
// scan keeps the first maximum).
__device__ inline void pick(double &a, int &ia, double &sa, double b, int ib, double sb) {
  if (b > a || (b == a && ib < ia)) {
    a = b;
    ia = ib;
    sa = sb;
  }
}

__global__ void __launch_bounds__(512) k_dw_chol(DwCholArgs C) {
  const int4 T = C.task[blockIdx.x];
  const DwElem E = C.el[T.z];
  const int N = 2 * E.nb * E.nb, i = threadIdx.x, lane = i & 63, wv = i >> 6;
  const bool valid = i < N;
  const double *Wt = C.W + (size_t)T.w * DW_WSTRIDE;
  double *Bt = C.B + (size_t)T.w * DW_WSTRIDE;
  double *sg = C.sigma + (size_t)T.w * DW_NMAX;
  __shared__ double rv[8], rs[8];
  __shared__ int ri[8];
  __shared__ double s_dmax0;
  double d = valid ? Wt[i + (size_t)N * i] : 0.0;
  {
    double a = valid ? fabs(d) : 0.0;
    for (int off = 32; off > 0; off >>= 1) a = fmax(a, __shfl_down(a, off, 64));
    if (lane == 0) rv[wv] = a;
    __syncthreads();
    if (i == 0) {
      double m = rv[0];
      for (int q = 1; q < 8; ++q) m = fmax(m, rv[q]);
      s_dmax0 = m;
    }
    __syncthreads();
  }
  const double dmax0 = s_dmax0;
  int rank = 0;
  for (int p = 0; p < N; ++p) {
    double a = valid ? fabs(d) : -1.0, sa = d;
    int ia = valid ? i : 1 << 30;
    for (int off = 32; off > 0; off >>= 1) {
      const double b = __shfl_down(a, off, 64), sb = __shfl_down(sa, off, 64);
      const int ib = __shfl_down(ia, off, 64);
      pick(a, ia, sa, b, ib, sb);
    }
    if (lane == 0) {
      rv[wv] = a;
      ri[wv] = ia;
      rs[wv] = sa;
    }
    __syncthreads();
    double best = rv[0], dp = rs[0];
    int piv = ri[0];
    for (int q = 1; q < 8; ++q) pick(best, piv, dp, rv[q], ri[q], rs[q]);
    __syncthreads();
    if (dmax0 <= 0.0 || best <= 1e-12 * dmax0) break;
    const double s = dp >= 0.0 ? 1.0 : -1.0;
    if (valid) {
      double col = Wt[i + (size_t)N * piv];
      for (int q = 0; q < rank; ++q) {
        const double f = sg[q] * Bt[(size_t)q * DW_NMAX + piv];
        col -= f * Bt[(size_t)q * DW_NMAX + i];
      }
      col /= sqrt(fabs(dp));
      d -= s * col * col;
      if (i == piv) d = 0.0;
      Bt[(size_t)rank * DW_NMAX + i] = col;
    }
    if (i == 0) sg[rank] = s;
    ++rank;
    __syncthreads();  // column rank - 1 and its sign complete before the next step reads them
  }
  if (i == 0) C.rank[T.w] = rank;
}

// Packs each task's rank columns (N rows each) contiguously at off[task].
__global__ void __launch_bounds__(256) k_dw_pack(const double *B, const int4 *task, const DwElem *el,
                                                 const int *rank, const long long *off, double *dst) {
  const int j = blockIdx.x;
  const int4 T = task[j];
  const int N = 2 * el[T.z].nb * el[T.z].nb, r = rank[j];
  const double *src = B + (size_t)T.w * DW_WSTRIDE;
  for (int q = threadIdx.x; q < r * N; q += 256) {
    const int p = q / N, a = q - p * N;
    dst[off[j] + q] = src[(size_t)p * DW_NMAX + a];
  }
}

}  // namespace

namespace hfg {

void diatomic_tei_device(DiatomicSetup &S, hipStream_t s, double *seconds) {
  const auto t0 = std::chrono::steady_clock::now();
  const RadialFE &R = S.rad;
  const int ne = S.Nel(), nlm = (int)S.lm_map.size();
  if ((int)R.nodes.size() > DW_NODES) throw Error(HF_EINVAL, "device setup handles at most 16 nodes per element");
  S.compute_pq();
  const std::vector<int> order = S.twoe_orders();
  // largest L of each |M|: the host caches evaluate Q over L = 0..Lhi of
  // the task's |M| (the same values whatever Lhi)
  std::map<int, int> Lhi;
  for (auto &lm : S.lm_map) Lhi[lm.second] = std::max(Lhi[lm.second], lm.first);
  std::vector<DwElem> el(ne);
  for (int e = 0; e < ne; ++e) {
    const LipShape sh = R.shape(e);
    el[e] = DwElem{R.bval[e], R.bval[e + 1], sh.lo, sh.nbf()};
    if (2 * sh.nbf() * sh.nbf() > DW_NMAX) throw Error(HF_EINVAL, "element too large for the device setup");
  }
  DevArray<double> d_nodes, d_x, d_w, d_q, inner, bfp, Wb, Bb, sg;
  DevArray<DwElem> d_el;
  DevArray<int4> d_task;
  DevArray<int> d_rank;
  DevArray<long long> d_off;
  DevArray<double> packed;
  PinnedArray<double> hpacked, hsig;
  d_nodes.upload(R.nodes, s);
  d_el.upload(el, s);
  // tasks grouped by rule size, in chunks that bound the work memory
  std::map<int, std::vector<std::pair<int, int>>> by_n;  // n -> (ilm, element)
  for (int q = 0; q < nlm; ++q)
    for (int e = 0; e < ne; ++e) by_n[order[e]].push_back({q, e});
  constexpr int CHUNK = 128;
  for (auto &grp : by_n) {
    const int n = grp.first;
    std::vector<double> x, w;
    gauss_chebyshev2(n, x, w);
    d_x.upload(x, s);
    d_w.upload(w, s);
    bfp.alloc((size_t)ne * n * DW_NB2S);
    // Q tables per (|M|, element) at this rule size
    std::map<std::pair<int, int>, std::vector<double>> qcols;
    {
      std::vector<std::pair<int, int>> keys;
      for (auto &te : grp.second) keys.push_back({S.lm_map[te.first].second, te.second});
      std::sort(keys.begin(), keys.end());
      keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
      std::vector<std::vector<double>> vals(keys.size());
      for (int k = 0; k < (int)keys.size(); ++k)
        vals[k] = S.outer_legendre_q(keys[k].second, n, keys[k].first, Lhi[keys[k].first]);
      for (size_t k = 0; k < keys.size(); ++k) qcols[keys[k]] = std::move(vals[k]);
    }
    DwArgs A{d_nodes.p, (int)R.nodes.size(), d_el.p, nullptr, 0, d_x.p, d_w.p, n, nullptr, nullptr, bfp.p, nullptr};
    hipLaunchKernelGGL(k_dw_bfprod, dim3(ne * n), dim3(256), 0, s, A, 0);
    const auto &tl = grp.second;
    for (size_t c0 = 0; c0 < tl.size(); c0 += CHUNK) {
      const int nt = (int)std::min<size_t>(CHUNK, tl.size() - c0);
      std::vector<int4> task(nt);
      std::vector<double> qt((size_t)nt * n);
      for (int j = 0; j < nt; ++j) {
        const int q = tl[c0 + j].first, e = tl[c0 + j].second;
        const int L = S.lm_map[q].first, M = S.lm_map[q].second;
        task[j] = int4{L, M, e, j};
        const std::vector<double> &col = qcols[{M, e}];
        const int ld = Lhi[M] + 1;
        for (int i = 0; i < n; ++i) qt[(size_t)j * n + i] = col[(size_t)i * ld + L];
      }
      d_task.upload(task, s);
      d_q.upload(qt, s);
      inner.alloc((size_t)nt * 2 * n * DW_NB2S);
      Wb.alloc((size_t)nt * DW_WSTRIDE);
      Bb.alloc((size_t)nt * DW_WSTRIDE);
      sg.alloc((size_t)nt * DW_NMAX);
      d_rank.alloc(nt);
      A.task = d_task.p;
      A.ntask = nt;
      A.qtab = d_q.p;
      A.inner = inner.p;
      A.W = Wb.p;
      hipLaunchKernelGGL(k_dw_inner, dim3(nt * n), dim3(256), 0, s, A);
      hipLaunchKernelGGL(k_dw_cumul, dim3((unsigned)(((int64_t)nt * 2 * DW_NB2S + 255) / 256)), dim3(256), 0, s, A);
      hipLaunchKernelGGL(k_dw_outer, dim3(nt * 256), dim3(256), 0, s, A);
      DwCholArgs C{d_el.p, d_task.p, Wb.p, Bb.p, sg.p, d_rank.p};
      hipLaunchKernelGGL(k_dw_chol, dim3(nt), dim3(512), 0, s, C);
      hip_check(hipGetLastError(), "device diatomic two-electron setup");
      // ranks first, then only the factors' rank columns come back (packed)
      std::vector<int> rk(nt);
      hip_check(hipMemcpyAsync(rk.data(), d_rank.p, nt * sizeof(int), hipMemcpyDeviceToHost, s), "ranks");
      hip_check(hipStreamSynchronize(s), "device diatomic two-electron setup");
      std::vector<long long> off(nt + 1, 0);
      for (int j = 0; j < nt; ++j) off[j + 1] = off[j] + (long long)rk[j] * 2 * el[tl[c0 + j].second].nb *
                                                             el[tl[c0 + j].second].nb;
      d_off.upload(off, s);
      packed.alloc((size_t)std::max<long long>(off[nt], 1));
      hpacked.alloc((size_t)std::max<long long>(off[nt], 1));
      hsig.alloc((size_t)nt * DW_NMAX);
      hipLaunchKernelGGL(k_dw_pack, dim3(nt), dim3(256), 0, s, Bb.p, d_task.p, d_el.p, d_rank.p, d_off.p, packed.p);
      hip_check(hipGetLastError(), "device diatomic factor pack");
      hip_check(hipMemcpyAsync(hpacked.p, packed.p, off[nt] * sizeof(double), hipMemcpyDeviceToHost, s), "factors");
      hip_check(hipMemcpyAsync(hsig.p, sg.p, (size_t)nt * DW_NMAX * sizeof(double), hipMemcpyDeviceToHost, s),
                "signs");
      hip_check(hipStreamSynchronize(s), "device diatomic two-electron setup");
      for (int j = 0; j < nt; ++j) {
        const int q = tl[c0 + j].first, e = tl[c0 + j].second, N = 2 * el[e].nb * el[e].nb, r = rk[j];
        Blk B(N, r);
        std::copy(hpacked.p + off[j], hpacked.p + off[j + 1], B.v.begin());
        S.cdB[(size_t)q * ne + e] = std::move(B);
        S.cdS[(size_t)q * ne + e].assign(hsig.p + (size_t)j * DW_NMAX, hsig.p + (size_t)j * DW_NMAX + r);
      }
    }
  }
  S.have_tei = true;
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace hfg
