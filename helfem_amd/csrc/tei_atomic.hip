// Atomic two-electron setup on the device: the in-element Slater tensor of
// every (L, element) by the split-at-r=r' quadrature with Lobatto order
// refinement, and its pivoted Cholesky factor (the expensive half of
// TwoDBasisT::compute_tei, TwoDBasis.cpp:709-735; the tensor follows
// quadrature::twoe_integral, libhelfem/src/quadrature.cpp:36-161, and the
// refinement and factorisation RadialBasis.cpp:644-716).
//
// Per refinement level n (all tasks share the rule sizes nstart, 2 nstart,
// ..., 512) one launch of each kernel covers every (L, element) still
// refining:
//   k_tei_inner   WG per (task, segment p): the segment integrals
//                 M_p(k,l) = sum_i w_i (r_i/r_p)^L / r_p (dr/2) B_k B_l over the
//                 rule mapped to [r_{p-1}, r_p]; shape values and weights of 256
//                 points at a time in LDS;
//   k_tei_cumul   the cumulative inner integral across segments, scaled by
//                 (r_p / r_{p-1})^(-L-1) (one thread per (task, (k,l)));
//   k_tei_prod    the outer weights B_a B_b w_p (dr/2) per (element, point);
//   k_tei_gemm    T(u,v) = sum_p prod(p,u) inner(p,v) + (u <-> v), 16x16 output
//                 tiles through LDS;
//   k_tei_metrics max |T_n - T_{n/2}| and max |T_n| per task, which the host
//                 turns into the refinement decision of the reference (change
//                 below 8 eps, or stalled at the roundoff floor).
// k_tei_chol then factors each converged tensor in one workgroup (one thread
// per row, greedy largest-diagonal pivot, absolute tolerance on the residual
// diagonal).  The factors come back to the host caches, so every later path
// (upload, clones, cache reads) is the one of the host setup.  Results agree
// with the host setup to rounding (the sums run in another order); the
// pivot order, and so the individual Cholesky vectors, can differ where two
// residual diagonals tie to rounding, the reconstructed tensor cannot.
#include "capi_common.hpp"
#include <chrono>
#include <limits>

using namespace hfg;

namespace {

constexpr int TEI_NODES = 16;          // LIP nodes per element (at most)
constexpr int NB2S = 256;              // stride of a (k,l) pair index (nb^2 <= 225)
constexpr size_t TSTRIDE = 65536;      // stride of one task's tensor (nb^2 x nb^2)
constexpr int CHUNK = 256;             // rule points per LDS chunk

struct TeiElem {
  double a, b;      // element bounds
  int lo, nb;       // first shape function, number of functions
  int empty_first;  // the first rule point is the element start: segment 0 is empty
  int pad;
};

struct TeiArgs {
  const double *nodes;
  int nnodes;
  const TeiElem *el;
  const int4 *task;  // (L, element, task id, -)
  int ntask;
  const double *x, *w;  // the level's Lobatto rule on [-1, 1]
  int n;
  double *inner;        // [active task][p][NB2S]
  double *prod;         // [element][p][NB2S]
  double *cur, *prev;   // [task id][TSTRIDE]
  double *metric;       // [active task][2]
};

// Lagrange polynomial f through the nodes at xv, as the node products of the
// host shape functions (same order of factors).
__device__ inline double lagrange(double xv, const double *nodes, int N, int f) {
  const double xf = nodes[f];
  double v = 1.0;
  for (int p = 0; p < N; ++p) {
    if (p == f) continue;
    v *= (xv - nodes[p]) / (xf - nodes[p]);
  }
  return v;
}

__global__ void __launch_bounds__(256) k_tei_inner(TeiArgs A) {
  const int tk = blockIdx.x / A.n, p = blockIdx.x - tk * A.n;
  const int4 T = A.task[tk];
  const TeiElem E = A.el[T.y];
  const int t = threadIdx.x, nb = E.nb, nb2 = nb * nb;
  double *out = A.inner + ((size_t)tk * A.n + p) * NB2S;
  const double rmid = 0.5 * (E.b + E.a), rlen = 0.5 * (E.b - E.a);
  if (p == 0 && E.empty_first) {
    if (t < nb2) out[t] = 0.0;
    return;
  }
  const double sa = p == 0 ? E.a : rmid + rlen * A.x[p - 1];
  const double sb = rmid + rlen * A.x[p];
  const double sm = 0.5 * (sb + sa), sl = 0.5 * (sb - sa);
  __shared__ double nodes[TEI_NODES];
  __shared__ double bf[CHUNK][TEI_NODES];
  __shared__ double wp[CHUNK];
  if (t < A.nnodes) nodes[t] = A.nodes[t];
  const int k = t % nb, l = t / nb;
  const bool valid = t < nb2;
  double acc = 0.0;
  const double Ld = (double)T.x;
  for (int i0 = 0; i0 < A.n; i0 += CHUNK) {
    __syncthreads();
    const int i = i0 + t;
    if (i < A.n) {
      const double rr = sm + sl * A.x[i];
      wp[t] = A.w[i] * (pow(rr / sb, Ld) / sb) * sl;
      const double xp = (rr - rmid) / rlen;
      for (int f = 0; f < nb; ++f) bf[t][f] = lagrange(xp, nodes, A.nnodes, E.lo + f);
    }
    __syncthreads();
    const int m = min(CHUNK, A.n - i0);
    if (valid)
      for (int ii = 0; ii < m; ++ii) acc += (bf[ii][k] * wp[ii]) * bf[ii][l];
  }
  if (valid) out[k + nb * l] = acc;
}

__global__ void __launch_bounds__(256) k_tei_cumul(TeiArgs A) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int tk = (int)(g / NB2S), kl = (int)(g - (int64_t)tk * NB2S);
  if (tk >= A.ntask) return;
  const int4 T = A.task[tk];
  const TeiElem E = A.el[T.y];
  if (kl >= E.nb * E.nb) return;
  const double rmid = 0.5 * (E.b + E.a), rlen = 0.5 * (E.b - E.a);
  const double e1 = (double)(-T.x - 1);
  double *col = A.inner + (size_t)tk * A.n * NB2S + kl;
  double prev = col[0];
  for (int p = 1; p < A.n; ++p) {
    double v = col[(size_t)p * NB2S];
    if (!(p == 1 && E.empty_first)) {
      const double r1 = rmid + rlen * A.x[p], r0 = rmid + rlen * A.x[p - 1];
      v += prev * (pow(r1, e1) / pow(r0, e1));
      col[(size_t)p * NB2S] = v;
    }
    prev = v;
  }
}

// prod(p, a*nb + b) = B_a(x_p) B_b(x_p) (w_p dr/2), one workgroup per
// (element, point).
__global__ void __launch_bounds__(256) k_tei_prod(TeiArgs A, int nel) {
  const int e = blockIdx.x / A.n, p = blockIdx.x - e * A.n;
  const TeiElem E = A.el[e];
  const int t = threadIdx.x, nb = E.nb;
  __shared__ double nodes[TEI_NODES];
  __shared__ double bv[TEI_NODES];
  if (t < A.nnodes) nodes[t] = A.nodes[t];
  __syncthreads();
  if (t < nb) bv[t] = lagrange(A.x[p], nodes, A.nnodes, E.lo + t);
  __syncthreads();
  if (t < nb * nb) {
    const int a = t / nb, b = t - a * nb;
    const double rlen = 0.5 * (E.b - E.a);
    A.prod[((size_t)e * A.n + p) * NB2S + t] = bv[a] * bv[b] * (A.w[p] * rlen);
  }
}

// One workgroup per (task, 16x16 output tile): T(u,v) = ints(u,v) + ints(v,u),
// ints(u,v) = sum_p prod(p,u) inner(p,v).
__global__ void __launch_bounds__(256) k_tei_gemm(TeiArgs A) {
  constexpr int NT16 = 16;  // 16-wide tiles of nb^2 <= 256
  const int tk = blockIdx.x / (NT16 * NT16), tt = blockIdx.x - tk * NT16 * NT16;
  const int tu = tt % NT16, tv = tt / NT16;
  const int4 T = A.task[tk];
  const TeiElem E = A.el[T.y];
  const int nb2 = E.nb * E.nb;
  if (tu * 16 >= nb2 || tv * 16 >= nb2) return;
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const double *P = A.prod + (size_t)T.y * A.n * NB2S;
  const double *I = A.inner + (size_t)tk * A.n * NB2S;
  __shared__ double pu[32][17], pv[32][17], iu[32][17], iv[32][17];
  double s1 = 0.0, s2 = 0.0;
  for (int p0 = 0; p0 < A.n; p0 += 32) {
    __syncthreads();
    for (int q = t; q < 512; q += 256) {
      const int pp = q >> 4, cc = q & 15, p = p0 + pp;
      const int cu = tu * 16 + cc, cv = tv * 16 + cc;
      const bool okp = p < A.n;
      pu[pp][cc] = (okp && cu < nb2) ? P[(size_t)p * NB2S + cu] : 0.0;
      iu[pp][cc] = (okp && cu < nb2) ? I[(size_t)p * NB2S + cu] : 0.0;
      pv[pp][cc] = (okp && cv < nb2) ? P[(size_t)p * NB2S + cv] : 0.0;
      iv[pp][cc] = (okp && cv < nb2) ? I[(size_t)p * NB2S + cv] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int pp = 0; pp < 32; ++pp) {
      s1 += pu[pp][tx] * iv[pp][ty];
      s2 += pv[pp][ty] * iu[pp][tx];
    }
  }
  const int u = tu * 16 + tx, v = tv * 16 + ty;
  if (u < nb2 && v < nb2) A.cur[(size_t)T.z * TSTRIDE + u + (size_t)nb2 * v] = s1 + s2;
}

// max |cur - prev| and max |cur| of each active task.
__global__ void __launch_bounds__(256) k_tei_metrics(TeiArgs A) {
  const int tk = blockIdx.x;
  const int4 T = A.task[tk];
  const TeiElem E = A.el[T.y];
  const int nb2 = E.nb * E.nb, n = nb2 * nb2;
  const double *c = A.cur + (size_t)T.z * TSTRIDE, *pr = A.prev + (size_t)T.z * TSTRIDE;
  double md = 0.0, ma = 0.0;
  for (int q = threadIdx.x; q < n; q += 256) {
    md = fmax(md, fabs(c[q] - pr[q]));
    ma = fmax(ma, fabs(c[q]));
  }
  __shared__ double r[2][256];
  r[0][threadIdx.x] = md;
  r[1][threadIdx.x] = ma;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      r[0][threadIdx.x] = fmax(r[0][threadIdx.x], r[0][threadIdx.x + s]);
      r[1][threadIdx.x] = fmax(r[1][threadIdx.x], r[1][threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    A.metric[2 * tk] = r[0][0];
    A.metric[2 * tk + 1] = r[1][0];
  }
}

struct CholArgs {
  const TeiElem *el;
  const int4 *task;       // all tasks (L, element, id, buffer 0/1 holding the tensor)
  const double *buf0, *buf1;
  double *fac;            // [task id][TSTRIDE]: column j of the factor at j*NB2S
  int *rank;
  double tol;
};

// Greedy pivoted Cholesky of one task's tensor: thread i owns row i.  The
// pivot is the largest residual diagonal above tol, the lowest index among
// equal ones (the host scan's choice).
__global__ void __launch_bounds__(256) k_tei_chol(CholArgs C) {
  const int4 T = C.task[blockIdx.x];
  const TeiElem E = C.el[T.y];
  const int n2 = E.nb * E.nb, i = threadIdx.x, lane = i & 63, wv = i >> 6;
  const bool valid = i < n2;
  const double *Am = (T.w ? C.buf1 : C.buf0) + (size_t)T.z * TSTRIDE;
  double *F = C.fac + (size_t)T.z * TSTRIDE;
  __shared__ double sv[4];
  __shared__ int si[4];
  double D = valid ? Am[i + (size_t)n2 * i] : 0.0;
  bool used = false;
  int rank = 0;
  for (int k = 0; k < n2; ++k) {
    double v = (valid && !used && D > C.tol) ? D : -1.0;
    int ix = valid ? i : 1 << 30;
    for (int off = 32; off > 0; off >>= 1) {
      const double v2 = __shfl_down(v, off, 64);
      const int i2 = __shfl_down(ix, off, 64);
      if (v2 > v || (v2 == v && i2 < ix)) {
        v = v2;
        ix = i2;
      }
    }
    if (lane == 0) {
      sv[wv] = v;
      si[wv] = ix;
    }
    __syncthreads();
    double bv = sv[0];
    int piv = si[0];
    for (int q = 1; q < 4; ++q)
      if (sv[q] > bv || (sv[q] == bv && si[q] < piv)) {
        bv = sv[q];
        piv = si[q];
      }
    __syncthreads();
    if (bv <= C.tol) break;  // no residual diagonal above the tolerance
    if (i == piv) used = true;
    const double sq = sqrt(bv);
    if (valid) {
      double col;
      if (used && i != piv) {
        col = 0.0;
      } else {
        double s = Am[i + (size_t)n2 * piv];
        for (int j = 0; j < rank; ++j) s -= F[(size_t)j * NB2S + i] * F[(size_t)j * NB2S + piv];
        col = s / sq;
      }
      if (i == piv) col = sq;
      F[(size_t)rank * NB2S + i] = col;
      if (!used) D -= col * col;
    }
    ++rank;
    __syncthreads();  // column `rank - 1` complete before the next step reads its pivot entry
  }
  if (i == 0) C.rank[T.z] = rank;
}

// Packs each task's rank columns (n2 rows each) contiguously at off[task].
__global__ void __launch_bounds__(256) k_tei_pack(const double *F, const TeiElem *el, const int4 *task,
                                                  const int *rank, const long long *off, double *dst) {
  const int4 T = task[blockIdx.x];
  const int n2 = el[T.y].nb * el[T.y].nb, r = rank[T.z];
  const double *src = F + (size_t)T.z * TSTRIDE;
  for (int q = threadIdx.x; q < r * n2; q += 256) {
    const int p = q / n2, a = q - p * n2;
    dst[off[T.z] + q] = src[(size_t)p * NB2S + a];
  }
}

}  // namespace

namespace hfg {

// The refinement decision of the host setup (fem.cpp refine()): returns true
// when the level's tensor is final.
static bool refine_done(double diff, double scale, double &d1, double &d2) {
  const double eps = std::numeric_limits<double>::epsilon();
  const double tol = 8.0 * eps, sq = std::sqrt(eps);
  if (diff <= tol * (scale + tol)) return true;
  if (d1 >= 0.0 && diff <= sq * (scale + tol) && diff > 0.5 * d1) return true;
  if (d2 >= 0.0 && diff <= sq * (scale + tol) && diff > 0.125 * d2) return true;
  d2 = d1;
  d1 = diff;
  return false;
}

void atomic_tei_device(AtomicSetup &S, hipStream_t s, double *seconds) {
  const auto t0 = std::chrono::steady_clock::now();
  const RadialFE &R = S.rad;
  const int nel = R.Nel(), NL = S.N_L, ntask = NL * nel;
  if ((int)R.nodes.size() > TEI_NODES) throw Error(HF_EINVAL, "device setup handles at most 16 nodes per element");
  std::vector<TeiElem> el(nel);
  for (int e = 0; e < nel; ++e) {
    const LipShape sh = R.shape(e);
    const double a = R.bval[e], b = R.bval[e + 1];
    const double rmid = 0.5 * (b + a), rlen = 0.5 * (b - a);
    el[e] = TeiElem{a, b, sh.lo, sh.nbf(), (rmid + rlen * -1.0) == a ? 1 : 0, 0};
  }
  DevArray<double> d_nodes, d_x, d_w, inner, prod, buf[2], metric, fac;
  DevArray<TeiElem> d_el;
  DevArray<int4> d_task, d_all;
  DevArray<int> d_rank;
  d_nodes.upload(R.nodes, s);
  d_el.upload(el, s);
  buf[0].alloc((size_t)ntask * TSTRIDE);
  buf[1].alloc((size_t)ntask * TSTRIDE);
  struct St {
    double d1 = -1.0, d2 = -1.0;
    bool have = false, done = false;
    int fin = 0;
  };
  std::vector<St> st(ntask);
  const int nmax = 512;
  int n = std::max(std::min(std::max(R.nquad, 5), 512), 2);
  std::vector<double> hm;
  for (int level = 0;; ++level) {
    std::vector<int4> act;
    for (int L = 0; L < NL; ++L)
      for (int e = 0; e < nel; ++e)
        if (!st[L * nel + e].done) act.push_back(int4{L, e, L * nel + e, 0});
    if (act.empty()) break;
    std::vector<double> x, w;
    gauss_lobatto(n, x, w);
    d_x.upload(x, s);
    d_w.upload(w, s);
    d_task.upload(act, s);
    const int na = (int)act.size();
    inner.alloc((size_t)na * n * NB2S);
    prod.alloc((size_t)nel * n * NB2S);
    metric.alloc((size_t)2 * na);
    TeiArgs A{d_nodes.p, (int)R.nodes.size(), d_el.p, d_task.p, na, d_x.p, d_w.p, n, inner.p, prod.p,
              buf[level & 1].p, buf[(level + 1) & 1].p, metric.p};
    hipLaunchKernelGGL(k_tei_inner, dim3(na * n), dim3(256), 0, s, A);
    hipLaunchKernelGGL(k_tei_cumul, dim3((unsigned)(((int64_t)na * NB2S + 255) / 256)), dim3(256), 0, s, A);
    hipLaunchKernelGGL(k_tei_prod, dim3(nel * n), dim3(256), 0, s, A, nel);
    hipLaunchKernelGGL(k_tei_gemm, dim3(na * 256), dim3(256), 0, s, A);
    const bool compare = st[act[0].z].have;  // every active task reached the same level
    if (compare) hipLaunchKernelGGL(k_tei_metrics, dim3(na), dim3(256), 0, s, A);
    hip_check(hipGetLastError(), "device two-electron setup");
    hm.assign((size_t)2 * na, 0.0);
    if (compare)
      hip_check(hipMemcpyAsync(hm.data(), metric.p, hm.size() * sizeof(double), hipMemcpyDeviceToHost, s),
                "setup metrics");
    hip_check(hipStreamSynchronize(s), "device two-electron setup");
    for (int q = 0; q < na; ++q) {
      St &z = st[act[q].z];
      if (z.have && refine_done(hm[2 * q], hm[2 * q + 1], z.d1, z.d2)) {
        z.done = true;
        z.fin = level & 1;
        continue;
      }
      z.have = true;
      if (n >= nmax) {
        z.done = true;
        z.fin = level & 1;
      }
    }
    n = std::min(2 * n, nmax);
  }
  inner.release();
  prod.release();
  // factor every tensor
  std::vector<int4> all(ntask);
  for (int L = 0; L < NL; ++L)
    for (int e = 0; e < nel; ++e) all[L * nel + e] = int4{L, e, L * nel + e, st[L * nel + e].fin};
  d_all.upload(all, s);
  fac.alloc((size_t)ntask * TSTRIDE);
  d_rank.alloc(ntask);
  CholArgs C{d_el.p, d_all.p, buf[0].p, buf[1].p, fac.p, d_rank.p, S.chol_tol};
  hipLaunchKernelGGL(k_tei_chol, dim3(ntask), dim3(256), 0, s, C);
  hip_check(hipGetLastError(), "device pivoted Cholesky");
  // ranks first, then only the factors' rank columns come back (packed)
  std::vector<int> rk(ntask);
  hip_check(hipMemcpyAsync(rk.data(), d_rank.p, ntask * sizeof(int), hipMemcpyDeviceToHost, s), "ranks");
  hip_check(hipStreamSynchronize(s), "device two-electron setup");
  std::vector<long long> off(ntask + 1, 0);
  for (int i = 0; i < ntask; ++i) off[i + 1] = off[i] + (long long)rk[i] * el[i % nel].nb * el[i % nel].nb;
  DevArray<long long> d_off;
  DevArray<double> packed;
  PinnedArray<double> hp;
  d_off.upload(off, s);
  packed.alloc((size_t)std::max<long long>(off[ntask], 1));
  hp.alloc((size_t)std::max<long long>(off[ntask], 1));
  hipLaunchKernelGGL(k_tei_pack, dim3(ntask), dim3(256), 0, s, fac.p, d_el.p, d_all.p, d_rank.p, d_off.p, packed.p);
  hip_check(hipGetLastError(), "device factor pack");
  hip_check(hipMemcpyAsync(hp.p, packed.p, off[ntask] * sizeof(double), hipMemcpyDeviceToHost, s), "factors");
  hip_check(hipStreamSynchronize(s), "device two-electron setup");
  S.compute_radial_factors();
  S.chol.assign((size_t)NL * nel, Blk());
  for (int i = 0; i < ntask; ++i) {
    const int e = i % nel, n2 = el[e].nb * el[e].nb, r = rk[i];
    Blk Lf(n2, r);
    std::copy(hp.p + off[i], hp.p + off[i + 1], Lf.v.begin());
    S.chol[i] = std::move(Lf);
  }
  S.have_tei = true;
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace hfg
