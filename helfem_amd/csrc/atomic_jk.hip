// Atomic Coulomb and exchange kernels (FP64) for gfx950.
//
// Data layout (device):
//   P, J, K     dense Nbf x Nbf column-major, angular-major blocks:
//               index (ang*Nrad + rad) as in TwoDBasis::add_sub
//               (src/atomic/TwoDBasis.cpp:310-318).
//   rs, rb      per (L, element) 16x16 zero-padded tiles of the r^L and
//               r^-L-1 element matrices (CoulombExchangeFE.h:180-217).
//   chol        per (L, element) the Cholesky vectors reshaped to 16x16
//               tiles M_p(a,b) = L(a + b*Ni, p) (CoulombExchangeFE.h:503-510).
//
// Coulomb (TwoDBasis.cpp:774-877 + CoulombExchangeFE.h:432-482):
//   only the element-diagonal Ni x Ni bands of each P block enter J, and J
//   is nonzero only on those bands.  Kernel 1 gathers the Gaunt-weighted
//   band of every (L,M) into LDS, reduces the two multipole traces and the
//   Cholesky matvec; kernel 2 prefix-sums the cross-element traces; kernel 3
//   expands the (L,M) bands back into the angular blocks of J.
// Exchange (TwoDBasis.cpp:880-999 + CoulombExchangeFE.h:484-530):
//   one workgroup per (output block, element pair (e,f)): the Gaunt-weighted
//   density sub-block R_L(e,f) is formed in registers for each coupled L and
//   immediately contracted with the radial factors (A R B' across elements,
//   sum_p M_p R M_p' inside an element), so the Nrad x Nrad Rmat of the
//   reference never exists.
#include "hfgpu_internal.hpp"
#include "device_common.hpp"
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <cmath>
#include <array>
#include <map>
#include <mutex>

#ifndef HFG_K_INFLIGHT
#define HFG_K_INFLIGHT 4
#endif
#ifndef HFG_K_OFF_WAVES
#define HFG_K_OFF_WAVES 6
#endif

namespace hfg {
namespace {

using namespace dev;

// Zero a rows x cols matrix; a contiguous one (ld == rows) as 16-byte stores.
// Workgroup blk of nblk zeroes its grid-stride share.
__device__ inline void fill_zero_body(double *A, int64_t rows, int64_t cols, int64_t ld, int64_t blk, int64_t nblk) {
  const int64_t n = rows * cols, stride = nblk * NT;
  if (ld == rows && (reinterpret_cast<uintptr_t>(A) & 15) == 0) {
    double2 *A2 = reinterpret_cast<double2 *>(A);
    const int64_t n2 = n / 2;
    for (int64_t q = blk * NT + threadIdx.x; q < n2; q += stride) A2[q] = make_double2(0.0, 0.0);
    if (blk == 0 && threadIdx.x == 0 && (n & 1)) A[n - 1] = 0.0;
    return;
  }
  for (int64_t q = blk * NT + threadIdx.x; q < n; q += stride) {
    const int64_t j = q / rows, i = q - j * rows;
    A[i + j * ld] = 0.0;
  }
}

__global__ void __launch_bounds__(NT) k_fill_zero(double *A, int64_t rows, int64_t cols, int64_t ld) {
  fill_zero_body(A, rows, cols, ld, blockIdx.x, gridDim.x);
}

// --- Coulomb -----------------------------------------------------------------

struct JArgs {
  const double *P;
  int64_t ldp;
  int Nel, Nrad, Nang;
  const int *first, *ni;
  const double *rs, *rb, *chol;
  const int *chol_off, *chol_rank;
  const int *lm_L, *lm_pstart, *pair_kl;
  const double *pair_c;
  double *Jin, *jsm, *jbg;
};

// One workgroup per ((L,M), element): blk = lm * Nel + e.
__device__ inline void j_gather_body(const JArgs &A, int blk) {
  const int lm = blk / A.Nel, e = blk - lm * A.Nel;
  const int t = threadIdx.x, a = t & 15, b = t >> 4;
  const int ni = A.ni[e], f0 = A.first[e], L = A.lm_L[lm];
  __shared__ double sP[TILE2];
  __shared__ double sS[TILE2];
  __shared__ double red[8];
  double v = 0.0;
  if (a < ni && b < ni) {
    const int p0 = A.lm_pstart[lm], p1 = A.lm_pstart[lm + 1];
    for (int p = p0; p < p1; ++p) {
      const int kl = A.pair_kl[p];
      const int k = kl / A.Nang, l = kl - k * A.Nang;
      v += A.pair_c[p] * A.P[(int64_t)(k * A.Nrad + f0 + a) + (int64_t)(l * A.Nrad + f0 + b) * A.ldp];
    }
  }
  sP[t] = v;
  __syncthreads();
  // tr(r_small P) and tr(r_big P): sum_ab r(a,b) P(b,a)
  const size_t tile = ((size_t)L * A.Nel + e) * TILE2;
  const double pt = sP[b + TILE * a];
  double ts = A.rs[tile + t] * pt;
  double tb = e > 0 ? A.rb[tile + t] * pt : 0.0;
  block_sum2(ts, tb, red);
  if (t == 0) {
    A.jsm[(size_t)lm * A.Nel + e] = ts;
    A.jbg[(size_t)lm * A.Nel + e] = tb;
  }
  // s_p = <M_p, P>, wave w handles p = w, w+4, ...
  const int rank = A.chol_rank[L * A.Nel + e];
  const double *M = A.chol + A.chol_off[L * A.Nel + e];
  const int lane = t & 63, w = t >> 6;
  for (int p = w; p < rank; p += 4) {
    const double *Mp = M + (size_t)p * TILE2;
    double s = Mp[lane] * sP[lane] + Mp[lane + 64] * sP[lane + 64] + Mp[lane + 128] * sP[lane + 128] +
               Mp[lane + 192] * sP[lane + 192];
    s = wave_sum(s);
    if (lane == 0) sS[p] = s;
  }
  __syncthreads();
  double jin = 0.0;
  for (int p = 0; p < rank; ++p) jin += M[(size_t)p * TILE2 + t] * sS[p];
  A.Jin[((size_t)lm * A.Nel + e) * TILE2 + t] = jin;
}

__global__ void __launch_bounds__(NT) k_atomic_j_gather(JArgs A) { j_gather_body(A, blockIdx.x); }

// Cross-element weights: pre_a(lm,e) = sum_{j>e} jbig(j), pre_b = sum_{j<e} jsmall(j)
__device__ inline void j_prefix_body(int lm, int n_lm, int Nel, const double *jsm, const double *jbg,
                                     double *preA, double *preB) {
  if (lm >= n_lm) return;
  const double *s = jsm + (size_t)lm * Nel;
  const double *g = jbg + (size_t)lm * Nel;
  double acc = 0.0;
  for (int e = 0; e < Nel; ++e) {
    preB[(size_t)lm * Nel + e] = acc;
    acc += s[e];
  }
  acc = 0.0;
  for (int e = Nel - 1; e >= 0; --e) {
    preA[(size_t)lm * Nel + e] = acc;
    acc += g[e];
  }
}

__global__ void k_atomic_j_prefix(int n_lm, int Nel, const double *jsm, const double *jbg, double *preA,
                                  double *preB) {
  j_prefix_body(blockIdx.x * blockDim.x + threadIdx.x, n_lm, Nel, jsm, jbg, preA, preB);
}

struct JOutArgs {
  double *J;
  int64_t ldj;
  int Nel, Nrad, Nang;
  const int *first, *ni, *lm_L;
  const double *rs, *rb, *Jin, *preA, *preB;
  const int *out_start, *out_lm;
  const double *out_c;
};

__device__ inline double jband(const JOutArgs &A, int lm, int e, int t) {
  const int L = A.lm_L[lm];
  const size_t tile = ((size_t)L * A.Nel + e) * TILE2;
  const size_t q = (size_t)lm * A.Nel + e;
  double v = A.Jin[q * TILE2 + t] + A.preA[q] * A.rs[tile + t];
  if (e > 0) v += A.preB[q] * A.rb[tile + t];
  return v;
}

// One workgroup per (angular output block (i,j), element): blk = ij * Nel + e.
__device__ inline void j_expand_body(const JOutArgs &A, int blk) {
  const int ij = blk / A.Nel, e = blk - ij * A.Nel;
  const int q0 = A.out_start[ij], q1 = A.out_start[ij + 1];
  if (q0 == q1) return;
  const int t = threadIdx.x, a = t & 15, b = t >> 4;
  const int ni = A.ni[e], f0 = A.first[e];
  if (a >= ni || b >= ni) return;
  if (e > 0 && a == 0 && b == 0) return;  // shared corner, written by e-1
  const bool corner = (a == ni - 1 && b == ni - 1 && e + 1 < A.Nel);
  double v = 0.0;
  for (int q = q0; q < q1; ++q) {
    const int lm = A.out_lm[q];
    double x = jband(A, lm, e, t);
    if (corner) x += jband(A, lm, e + 1, 0);
    v += A.out_c[q] * x;
  }
  const int i = ij / A.Nang, j = ij - i * A.Nang;
  A.J[(int64_t)(i * A.Nrad + f0 + a) + (int64_t)(j * A.Nrad + f0 + b) * A.ldj] = v;
}

__global__ void __launch_bounds__(NT) k_atomic_j_expand(JOutArgs A) { j_expand_body(A, blockIdx.x); }

// --- Exchange ----------------------------------------------------------------


template <typename T>
struct KArgs {
  const T *P;
  int64_t ldp;
  double *K;
  int64_t ldk;
  int Nel, Nrad, Nang, N_L, n_act;
  const int *first, *ni;
  const T *rs, *rb, *chol;
  const int *chol_off, *chol_rank;
  const int *act_jk, *term_start;
  const int2 *term_off;  // per term (i*Nrad, l*Nrad) of its density block
  const T *term_c;
  const int *efl;  // element pairs e*Nel+f of this launch
  int n_ef;
  const T *Lscale;  // optional per-L factor on the Gaunt weights (range separation)
  const T *pt;      // tiled density (k_atomic_ptile) or null: gather from P
  const int *ttile; // per term: its tile in an element pair's slice
  int n_src;
  int sym;          // symmetric density: e > f pairs (e == f: j <= k) also written to the mirror
};

// Density tiles in gather order: tile (ef = e*Nel+f, src) holds the 16x16
// block of P between element e's radial functions of shell i and element f's
// of shell l (src = i*Nang+l from the source list), lane-major in the
// matrix-core operand order (entry lane*4 + s is P(c, row(g,s)) for lane
// (g,c)), zero outside the elements.  One wave per tile; wg is the
// workgroup index (four tiles per workgroup).
template <typename T>
__device__ inline void ptile_body(const double *P, int64_t ldp, int Nel, int Nrad, const int *first, const int *ni,
                                  int Nang, const int *src_blk, int n_src, T *out, int64_t wg) {
  const int64_t tile = wg * (NT / 64) + (threadIdx.x >> 6);
  if (tile >= (int64_t)Nel * Nel * n_src) return;
  const int ef = (int)(tile / n_src), src = (int)(tile - (int64_t)ef * n_src);
  const int e = ef / Nel, f = ef - e * Nel;
  const int blk = src_blk[src], i = blk / Nang, l = blk - i * Nang;
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const bool rv = c < ni[e];
  const int64_t row = (int64_t)i * Nrad + first[e] + c, col0 = (int64_t)l * Nrad + first[f];
  typename Mma<T>::acc_t v;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int kk = Mma<T>::row(g, s);
    v[s] = (rv && kk < ni[f]) ? (T)P[row + (col0 + kk) * ldp] : T(0);
  }
  *reinterpret_cast<typename Mma<T>::acc_t *>(out + tile * TILE2 + lane * 4) = v;
}

template <typename T>
__global__ void __launch_bounds__(NT) k_atomic_ptile(const double *P, int64_t ldp, int Nel, int Nrad,
                                                     const int *first, const int *ni, int Nang, const int *src_blk,
                                                     int n_src, T *out) {
  ptile_body<T>(P, ldp, Nel, Nrad, first, ni, Nang, src_blk, n_src, out, blockIdx.x);
}

// One workgroup per (coupled output block (j,k), element pair (e,f)); the
// four waves split the multipoles L.  For each L a wave forms the
// Gaunt-weighted density tile R_L(e,f) = sum_terms c * P_il(e,f) directly in
// matrix-operand layout (lane (g,c) holds R(c, row(g,s))), then contracts it
// with the radial factors on the matrix cores:
//   e != f:  K += A R B'       (A, B the r^L / r^-L-1 element matrices)
//   e == f:  K += sum_p M_p R M_p'   (M_p the Cholesky vectors as 16x16)
// T = double is the reference's arithmetic; T = float runs the gathers and
// products in FP32 (half the bytes, the FP32 matrix rate) and sums the four
// waves' tiles in FP64.  Workgroups sharing (e,f) run on one XCD.  The
// density tiles come from the tiled copy (k_atomic_ptile: one aligned 2 KB
// tile per (element pair, density block)) when it fits in memory, else
// straight from P.  NW (waves per workgroup, 1, 2 or 4) is the workgroup
// shape: fewer waves suit bases with few multipoles (N_L < 4 leaves waves of
// a four-wave group idle); w is the workgroup's index within its system's
// grid (a multiple-of-8 offset of the launch's, so w & 7 is its XCD).
template <bool INSIDE, bool TILED, typename T, int NW = 4>
__device__ inline void atomic_k_body(const KArgs<T> &A, int w) {
  typedef Mma<T> MM;
  typedef typename MM::acc_t acc_t;
  const int xcd = w & 7, loc = w >> 3;
  const int efi = (loc / A.n_act) * 8 + xcd;
  const int act = loc - (loc / A.n_act) * A.n_act;
  if (efi >= A.n_ef) return;
  const int ef = A.efl[efi], e = ef / A.Nel, f = ef - e * A.Nel;
  const int jk = A.act_jk[act], j = jk / A.Nang, k = jk - j * A.Nang;
  if (INSIDE && A.sym && j > k) return;  // the mirror of block (k, j) covers it
  const int t = threadIdx.x, lane = t & 63, g = lane >> 4, c = lane & 15;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int ni = A.ni[e], nj = A.ni[f], fi = A.first[e], fj = A.first[f];
  __shared__ double red[NW * TILE2];
  acc_t acc = {0, 0, 0, 0}, acc2 = {0, 0, 0, 0};
  // Lanes outside the element pair read a clamped entry of the same tile and
  // are masked once per tile, so the term loop has no branches and its loads
  // can all be in flight together.
  const int rloc = fi + min(c, ni - 1);
  int cloc[4], kk[4];
  T cmask[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    kk[s] = MM::row(g, s);
    cloc[s] = fj + min(kk[s], nj - 1);
    cmask[s] = (c < ni && kk[s] < nj) ? T(1) : T(0);
  }
  typedef typename MM::acc_t vec;
  const T *pt = TILED ? A.pt + (size_t)ef * A.n_src * TILE2 + lane * 4 : nullptr;
  for (int L = wv; L < A.N_L; L += NW) {
    const int q0 = A.term_start[act * A.N_L + L], q1 = A.term_start[act * A.N_L + L + 1];
    if (q0 == q1) continue;
    T r[4] = {0, 0, 0, 0};
    int q = q0;
    if constexpr (TILED) {
      // one aligned 4-vector per lane and term, masks already in the tiles;
      // HFG_K_INFLIGHT terms' loads issued together
      constexpr int KU = HFG_K_INFLIGHT;
      for (; q + KU <= q1; q += KU) {
        vec x[KU];
        T cq[KU];
#pragma unroll
        for (int u = 0; u < KU; ++u) {
          x[u] = *reinterpret_cast<const vec *>(pt + (size_t)A.ttile[q + u] * TILE2);
          cq[u] = A.term_c[q + u];
        }
#pragma unroll
        for (int u = 0; u < KU; ++u)
#pragma unroll
          for (int s = 0; s < 4; ++s) r[s] += cq[u] * x[u][s];
      }
      for (; q < q1; ++q) {
        const vec x = *reinterpret_cast<const vec *>(pt + (size_t)A.ttile[q] * TILE2);
        const T cq = A.term_c[q];
#pragma unroll
        for (int s = 0; s < 4; ++s) r[s] += cq * x[s];
      }
    } else {
      for (; q + 4 <= q1; q += 4) {
        T x[4][4], cq[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int2 to = A.term_off[q + u];
          cq[u] = A.term_c[q + u];
          const T *Pr = A.P + (int64_t)(to.x + rloc);
#pragma unroll
          for (int s = 0; s < 4; ++s) x[u][s] = Pr[(int64_t)(to.y + cloc[s]) * A.ldp];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int s = 0; s < 4; ++s) r[s] += cq[u] * x[u][s];
      }
      for (; q < q1; ++q) {
        const int2 to = A.term_off[q];
        const T cq = A.term_c[q];
        const T *Pr = A.P + (int64_t)(to.x + rloc);
#pragma unroll
        for (int s = 0; s < 4; ++s) r[s] += cq * Pr[(int64_t)(to.y + cloc[s]) * A.ldp];
      }
    }
    const T msc = A.Lscale ? A.Lscale[L] : T(1);
#pragma unroll
    for (int s = 0; s < 4; ++s) r[s] *= cmask[s] * msc;
    const size_t te = ((size_t)L * A.Nel + e) * TILE2, tf = ((size_t)L * A.Nel + f) * TILE2;
    if (!INSIDE) {
      const T *At = (e > f) ? A.rb + te : A.rs + te;
      const T *Bt = (e > f) ? A.rs + tf : A.rb + tf;
      T as[4], bs[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        as[s] = At[c + TILE * kk[s]];
        bs[s] = Bt[c + TILE * kk[s]];
      }
      acc_t y = {0, 0, 0, 0};
      y = mma16<T>(r, bs, y);    // Y = R B'
      acc = mma16<T>(as, y, acc);  // K += A Y
    } else {
      const int rank = A.chol_rank[L * A.Nel + e];
      const T *M = A.chol + A.chol_off[L * A.Nel + e];
      // two Cholesky vectors per step, so two independent chains overlap;
      // M_p(c, k) serves as both the Y operand of R M_p' and the X operand of M_p Y
      for (int p = 0; p < rank; p += 2) {
        const T *Ma = M + (size_t)p * TILE2;
        const T *Mb = M + (size_t)(p + 1 < rank ? p + 1 : p) * TILE2;
        const T wb = p + 1 < rank ? T(1) : T(0);
        T ma[4], mb[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          ma[s] = Ma[c + TILE * kk[s]];
          mb[s] = wb * Mb[c + TILE * kk[s]];
        }
        acc_t ya = {0, 0, 0, 0}, yb = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          ya = MM::step(r[s], ma[s], ya);  // R M_a'
          yb = MM::step(r[s], mb[s], yb);  // R M_b'
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          acc = MM::step(ma[s], ya[s], acc);    // M_a Y_a
          acc2 = MM::step(mb[s], yb[s], acc2);  // M_b Y_b
        }
      }
    }
  }
  // combine the waves' partial tiles (lane holds K(row(g,r), c)) in FP64
  acc += acc2;
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) red[wv * TILE2 + kk[rr] + TILE * c] = (double)acc[rr];
  __syncthreads();
  for (int tt = t; tt < TILE2; tt += 64 * NW) {
    const int a = tt & 15, b = tt >> 4;
    if (a >= ni || b >= nj) continue;
    double v = red[tt];
#pragma unroll
    for (int q = 1; q < NW; ++q) v += red[q * TILE2 + tt];
    const int64_t row = (int64_t)j * A.Nrad + fi + a, col = (int64_t)k * A.Nrad + fj + b;
    const bool srow = (a == 0 && e > 0) || (a == ni - 1 && e + 1 < A.Nel);
    const bool scol = (b == 0 && f > 0) || (b == nj - 1 && f + 1 < A.Nel);
    double *dst = A.K + row + col * A.ldk;
    if (srow || scol)
      atomicAdd(dst, -v);
    else
      *dst = -v;
    if (A.sym && (!INSIDE || j != k)) {
      // mirror tile (k, j, f, e) = this tile transposed
      double *mdst = A.K + col + row * A.ldk;
      if (srow || scol)
        atomicAdd(mdst, -v);
      else
        *mdst = -v;
    }
  }
}

// Launch wrappers.  The cross-element body fits 6 waves per SIMD in FP64
// (75 registers) and 8 in FP32 without spills; the in-element body keeps the
// compiler's register choice.
template <bool TILED>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(HFG_K_OFF_WAVES))) k_atomic_k_off_f64(KArgs<double> A) {
  atomic_k_body<false, TILED, double>(A, blockIdx.x);
}
template <bool TILED>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8))) k_atomic_k_off_f32(KArgs<float> A) {
  atomic_k_body<false, TILED, float>(A, blockIdx.x);
}
template <bool TILED>
__global__ void __launch_bounds__(NT) k_atomic_k_in_f64(KArgs<double> A) {
  atomic_k_body<true, TILED, double>(A, blockIdx.x);
}
template <bool TILED>
__global__ void __launch_bounds__(NT) k_atomic_k_in_f32(KArgs<float> A) {
  atomic_k_body<true, TILED, float>(A, blockIdx.x);
}

void launch_atomic_k(int part, bool tiled, int64_t grid, hipStream_t s, const KArgs<double> &a) {
  const dim3 gd((unsigned)grid), bd(NT);
  if (part == 0 && tiled)
    hipLaunchKernelGGL(k_atomic_k_off_f64<true>, gd, bd, 0, s, a);
  else if (part == 0)
    hipLaunchKernelGGL(k_atomic_k_off_f64<false>, gd, bd, 0, s, a);
  else if (tiled)
    hipLaunchKernelGGL(k_atomic_k_in_f64<true>, gd, bd, 0, s, a);
  else
    hipLaunchKernelGGL(k_atomic_k_in_f64<false>, gd, bd, 0, s, a);
}
void launch_atomic_k(int part, bool tiled, int64_t grid, hipStream_t s, const KArgs<float> &a) {
  const dim3 gd((unsigned)grid), bd(NT);
  if (part == 0 && tiled)
    hipLaunchKernelGGL(k_atomic_k_off_f32<true>, gd, bd, 0, s, a);
  else if (part == 0)
    hipLaunchKernelGGL(k_atomic_k_off_f32<false>, gd, bd, 0, s, a);
  else if (tiled)
    hipLaunchKernelGGL(k_atomic_k_in_f32<true>, gd, bd, 0, s, a);
  else
    hipLaunchKernelGGL(k_atomic_k_in_f32<false>, gd, bd, 0, s, a);
}

// --- Range-separated exchange with non-separable (erfc) element pairs ------------
//
// One workgroup per (coupled output block, element pair (e,f)), all pairs
// including e == f, waves splitting L.  The Gaunt-weighted density tile
// R_L(e,f) is gathered as in k_atomic_k, staged in LDS, and multiplied by
// the pair's dense operator: K(a,b) += sum_cd T_L,ef[c + 16d][a + 16b]
// R(c,d) (CoulombExchangeFE.h:412-440, which applies the exchange-ordered
// element-pair tensor to vec(P_ij)).  The operator rows are read coalesced.

struct KPairArgs {
  const double *P;
  int64_t ldp;
  double *K;
  int64_t ldk;
  int Nel, Nrad, Nang, N_L, n_act;
  const int *first, *ni;
  const double *T, *Lscale;
  const int *act_jk, *term_start, *term_il;
  const double *term_c;
};

__global__ void __launch_bounds__(NT) k_atomic_k_pair(KPairArgs A) {
  const int w = blockIdx.x;
  const int xcd = w & 7, loc = w >> 3;
  const int n_ef = A.Nel * A.Nel;
  const int efi = (loc / A.n_act) * 8 + xcd;
  const int act = loc - (loc / A.n_act) * A.n_act;
  if (efi >= n_ef) return;
  const int e = efi / A.Nel, f = efi - e * A.Nel;
  const int jk = A.act_jk[act], j = jk / A.Nang, k = jk - j * A.Nang;
  const int t = threadIdx.x, lane = t & 63, g = lane >> 4, c = lane & 15;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int ni = A.ni[e], nj = A.ni[f], fi = A.first[e], fj = A.first[f];
  __shared__ double Rl[4 * TILE2];
  __shared__ double red[4 * TILE2];
  const int rloc = fi + min(c, ni - 1);
  int cloc[4], kk[4];
  double cmask[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    kk[s] = g + 4 * s;
    cloc[s] = fj + min(kk[s], nj - 1);
    cmask[s] = (c < ni && kk[s] < nj) ? 1.0 : 0.0;
  }
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  double *R = Rl + wv * TILE2;
  for (int L = wv; L < A.N_L; L += 4) {
    const int q0 = A.term_start[act * A.N_L + L], q1 = A.term_start[act * A.N_L + L + 1];
    if (q0 == q1) continue;
    double r[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = q0; q < q1; ++q) {
      const int il = A.term_il[q], i = il / A.Nang, l = il - i * A.Nang;
      const double cq = A.term_c[q];
      const double *Pr = A.P + (int64_t)(i * A.Nrad + rloc);
#pragma unroll
      for (int s = 0; s < 4; ++s) r[s] += cq * Pr[(int64_t)(l * A.Nrad + cloc[s]) * A.ldp];
    }
    const double msc = A.Lscale[L];
#pragma unroll
    for (int s = 0; s < 4; ++s) R[c + TILE * kk[s]] = r[s] * cmask[s] * msc;
    // the wave's own LDS slice: a wave-wide barrier orders the writes
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const double *Tp = A.T + (((size_t)L * A.Nel + e) * A.Nel + f) * TILE2 * TILE2 + lane;
    for (int kx = 0; kx < TILE2; ++kx) {
      const double rv = R[kx];
      const double *Tk = Tp + (size_t)kx * TILE2;
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u] += Tk[64 * u] * rv;
    }
    __builtin_amdgcn_wave_barrier();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) red[wv * TILE2 + lane + 64 * u] = acc[u];
  __syncthreads();
  const int a = t & 15, b = t >> 4;
  if (a >= ni || b >= nj) return;
  const double v = red[t] + red[TILE2 + t] + red[2 * TILE2 + t] + red[3 * TILE2 + t];
  const int64_t row = (int64_t)j * A.Nrad + fi + a, col = (int64_t)k * A.Nrad + fj + b;
  const bool srow = (a == 0 && e > 0) || (a == ni - 1 && e + 1 < A.Nel);
  const bool scol = (b == 0 && f > 0) || (b == nj - 1 && f + 1 < A.Nel);
  double *dst = A.K + row + col * A.ldk;
  if (srow || scol)
    atomicAdd(dst, -v);
  else
    *dst = -v;
}

// --- In-element exchange as a dense operator ----------------------------------
//
// For one (L, element e) the in-element exchange of every output block coupled
// to L is Y(a,b) = sum_cd T((a,b),(c,d)) R(c,d) with T(a + c n, b + d n) =
// sum_p M_p(a,c) M_p(b,d): a GEMM Y (256 x runs) = T (256 x KC) X (KC x runs)
// on the f64 matrix cores, X the Gaunt-weighted density tiles formed in LDS.

constexpr int NCOL = 32;

struct KInArgs {
  const double *P;
  int64_t ldp;
  int Nel, Nrad, Nang;
  const int *first, *ni;
  const double *T;
  const int *wg, *L_rstart, *L_runs, *run_q, *term_il;
  const double *term_c;
  double *Y;
};

__global__ void __launch_bounds__(NT) k_atomic_kin_gemm(KInArgs A) {
  const int w = blockIdx.x;
  const int L = A.wg[3 * w], e = A.wg[3 * w + 1], c0 = A.wg[3 * w + 2];
  if (L < 0) return;
  const int r0 = A.L_rstart[L] + c0, r1 = min(A.L_rstart[L + 1], r0 + NCOL);
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, g = lane >> 4, c = lane & 15;
  const int n = A.ni[e], fe = A.first[e];
  __shared__ double X[KC * NCOL];
  __shared__ int srun[NCOL];
  if (t < NCOL) srun[t] = r0 + t < r1 ? A.L_runs[r0 + t] : -1;
  __syncthreads();
  const int cc = t % 15, dd = t / 15;
  const bool cdok = t < 225 && cc < n && dd < n;
  if (t < KC)
    for (int j = 0; j < NCOL; ++j) {
      const int run = srun[j];
      double v = 0.0;
      if (run >= 0 && cdok)
        for (int q = A.run_q[run]; q < A.run_q[run + 1]; ++q) {
          const int il = A.term_il[q], i = il / A.Nang, l = il - i * A.Nang;
          v += A.term_c[q] * A.P[(int64_t)(i * A.Nrad + fe + cc) + (int64_t)(l * A.Nrad + fe + dd) * A.ldp];
        }
      X[t * NCOL + j] = v;
    }
  __syncthreads();
  v4d acc[4][2];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) acc[rt][0] = acc[rt][1] = v4d{0.0, 0.0, 0.0, 0.0};
  const double *Tc = A.T + (size_t)(L * A.Nel + e) * KC * TILE2 + 64 * wv + c;
  for (int ks = 0; ks < KC / 4; ++ks) {
    const int k = 4 * ks + g;
    const double b0 = X[k * NCOL + c], b1 = X[k * NCOL + 16 + c];
    const double *Tk = Tc + (size_t)k * TILE2;
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      const double a = Tk[16 * rt];
      acc[rt][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b0, acc[rt][0], 0, 0, 0);
      acc[rt][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b1, acc[rt][1], 0, 0, 0);
    }
  }
#pragma unroll
  for (int ct = 0; ct < 2; ++ct) {
    const int run = srun[16 * ct + c];
    if (run < 0) continue;
    double *out = A.Y + ((size_t)run * A.Nel + e) * TILE2 + 64 * wv + g;
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r) out[16 * rt + 4 * r] = acc[rt][ct][r];
  }
}

struct KRedArgs {
  double *K;
  int64_t ldk;
  int Nel, Nrad, Nang;
  const int *first, *ni, *act_jk, *act_rstart;
  const double *Y;
};

// One workgroup per (active output block, element): K tile = -sum of the
// block's partial tiles; shared boundary rows/cols accumulated atomically.
__global__ void __launch_bounds__(NT) k_atomic_kin_reduce(KRedArgs A) {
  const int act = blockIdx.x / A.Nel, e = blockIdx.x - act * A.Nel;
  const int t = threadIdx.x, a = t & 15, b = t >> 4;
  const int n = A.ni[e], fe = A.first[e];
  if (a >= n || b >= n) return;
  const int jk = A.act_jk[act], j = jk / A.Nang, k = jk - j * A.Nang;
  double v = 0.0;
  for (int run = A.act_rstart[act]; run < A.act_rstart[act + 1]; ++run)
    v += A.Y[((size_t)run * A.Nel + e) * TILE2 + t];
  const bool srow = (a == 0 && e > 0) || (a == n - 1 && e + 1 < A.Nel);
  const bool scol = (b == 0 && e > 0) || (b == n - 1 && e + 1 < A.Nel);
  double *dst = A.K + (int64_t)(j * A.Nrad + fe + a) + (int64_t)(k * A.Nrad + fe + b) * A.ldk;
  if (srow || scol)
    atomicAdd(dst, -v);
  else
    *dst = -v;
}

}  // namespace

// Dense in-element operators T = L L' per (L, element) from the Cholesky
// factors (CoulombExchangeFE.h:503-510 contracts the same factor as
// sum_p M_p P M_p').  Built only when the environment variable
// HFGPU_EXCHANGE_OPERATOR_GIB grants a budget (in GiB) that holds them: on
// gfx950 the Cholesky-form in-element kernel is the faster of the two
// (0.85 vs 1.16 ms for Xe at lmax 6), so the operator path is opt-in.
void atomic_build_operator(const AtomicSetup &S, AtomicDevice &d, hipStream_t s) {
  const int Nel = S.rad.Nel(), nL = S.N_L;
  const size_t per = (size_t)KC * TILE2;
  double budget = 0.0;
  if (const char *env = std::getenv("HFGPU_EXCHANGE_OPERATOR_GIB")) budget = std::atof(env) * (1ull << 30);
  d.dense_in = false;
  d.Top.release();
  if (nL == 0 || 8.0 * per * nL * Nel > budget) return;
  std::vector<double> T(per * nL * Nel, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (int x = 0; x < nL * Nel; ++x) {
    const int e = x % Nel, n = S.rad.ni[e], nn = n * n;
    const Blk &C = S.chol[x];
    std::vector<double> W((size_t)nn * nn, 0.0);
    for (int p = 0; p < C.c; ++p) {
      const double *cp = &C.v[(size_t)p * nn];
      for (int j = 0; j < nn; ++j) {
        const double f = cp[j];
        double *wj = &W[(size_t)j * nn];
        for (int i = 0; i < nn; ++i) wj[i] += f * cp[i];
      }
    }
    double *Tx = &T[per * x];
    for (int dd = 0; dd < n; ++dd)
      for (int cc = 0; cc < n; ++cc) {
        double *row = Tx + (size_t)(cc + 15 * dd) * TILE2;
        for (int b = 0; b < n; ++b)
          for (int a = 0; a < n; ++a) row[a + TILE * b] = W[(size_t)(a + cc * n) + (size_t)(b + dd * n) * nn];
      }
  }
  d.Top.upload(T, s);
  hip_check(hipStreamSynchronize(s), "exchange operator upload");
  d.dense_in = true;
}

namespace {
__global__ void k_to_f32(const double *src, float *dst, size_t n) {
  for (size_t q = (size_t)blockIdx.x * NT + threadIdx.x; q < n; q += (size_t)gridDim.x * NT) dst[q] = (float)src[q];
}
__global__ void k_mat_to_f32(const double *A, int64_t lda, float *B, int64_t n) {
  const int64_t col = blockIdx.y;
  for (int64_t r = (int64_t)blockIdx.x * NT + threadIdx.x; r < n; r += (int64_t)gridDim.x * NT)
    B[r + col * n] = (float)A[r + col * lda];
}
}  // namespace

void to_f32_launch(const double *src, float *dst, size_t n, hipStream_t s) {
  if (n == 0) return;
  const unsigned blocks = (unsigned)std::min<size_t>((n + NT - 1) / NT, 65536);
  hipLaunchKernelGGL(k_to_f32, dim3(blocks), dim3(NT), 0, s, src, dst, n);
  hip_check(hipGetLastError(), "to_f32");
}

void mat_to_f32_launch(const double *A, int64_t lda, float *B, int64_t n, hipStream_t s) {
  if (n == 0) return;
  if (n > 65535) throw Error(-1, "matrix too large for the FP32 conversion");
  const unsigned bx = (unsigned)std::min<int64_t>((n + NT - 1) / NT, 64);
  hipLaunchKernelGGL(k_mat_to_f32, dim3(bx, (unsigned)n), dim3(NT), 0, s, A, lda, B, n);
  hip_check(hipGetLastError(), "mat_to_f32");
}

namespace {
// One workgroup per block: out(row0 + r, col0 + c) = vals[off + r + c rows].
__global__ void __launch_bounds__(NT) k_scatter_blocks(double *out, int64_t ld, const long long *meta,
                                                       const double *vals) {
  const long long *m = meta + 5 * (size_t)blockIdx.x;
  const long long rows = m[2], n = m[2] * m[3];
  const double *v = vals + m[4];
  for (long long q = threadIdx.x; q < n; q += NT) {
    const long long c = q / rows, r = q - c * rows;
    out[(m[0] + r) + (m[1] + c) * ld] = v[q];
  }
}
}  // namespace

namespace {
// Block-diagonal one-electron matrix of an atomic basis: shell a's diagonal
// block is f M + l_a (l_a + 1) X (X null: no l term), zero elsewhere; the
// products and the sum are rounded separately, as the host assembly does.
__global__ void __launch_bounds__(NT) k_shell_diag_expand(double *out, int64_t ld, int64_t nbf, int n,
                                                          const double *M, const double *X, const int *lval,
                                                          double f) {
#pragma clang fp contract(off)  // two roundings, as the host assembly: no fused multiply-add
  const int64_t total = nbf * nbf;
  for (int64_t q = (int64_t)blockIdx.x * NT + threadIdx.x; q < total; q += (int64_t)gridDim.x * NT) {
    const int64_t j = q / nbf, i = q - j * nbf;
    const int64_t a = i / n;
    double v = 0.0;
    if (a == j / n) {
      const int64_t m = (i - a * n) + (j - a * n) * n;
      v = f * M[m];
      if (X) {
        const int l = lval[a];
        if (l != 0) v += double(l * (l + 1)) * X[m];
      }
    }
    out[i + j * ld] = v;
  }
}
}  // namespace

void shell_diag_launch(double *out, int64_t ld, int64_t nbf, int n, const double *M, const double *X,
                       const int *lval, double f, hipStream_t s) {
  const int64_t total = nbf * nbf;
  const unsigned blocks = (unsigned)std::min<int64_t>((total + NT - 1) / NT, 8192);
  hipLaunchKernelGGL(k_shell_diag_expand, dim3(std::max(1u, blocks)), dim3(NT), 0, s, out, ld, nbf, n, M, X, lval, f);
  hip_check(hipGetLastError(), "shell_diag_expand");
}

void scatter_blocks_launch(double *out, int64_t ld, int64_t n, const std::vector<long long> &meta,
                           const std::vector<double> &vals, hipStream_t s) {
  fill_zero_launch(out, n, n, ld, s);
  const size_t nblk = meta.size() / 5;
  if (nblk == 0) return;
  for (size_t b = 0; b < nblk; ++b) {
    const long long *m = &meta[5 * b];
    if (m[0] < 0 || m[1] < 0 || m[0] + m[2] > n || m[1] + m[3] > n || m[4] + m[2] * m[3] > (long long)vals.size())
      throw Error(-1, "one-electron block outside the matrix");
  }
  if (nblk > 0x7fffffff) throw Error(-1, "too many one-electron blocks");
  DevArray<long long> dm;
  DevArray<double> dv;
  dm.upload(meta, s);
  dv.upload(vals, s);
  hipLaunchKernelGGL(k_scatter_blocks, dim3((unsigned)nblk), dim3(NT), 0, s, out, ld, dm.p, dv.p);
  hip_check(hipGetLastError(), "scatter_blocks");
  hip_check(hipStreamSynchronize(s), "scatter_blocks sync");  // before the staging buffers go
}

double exchange_tile_budget(size_t held_bytes) {
  if (const char *env = std::getenv("HFGPU_EXCHANGE_TILE_GIB")) return std::atof(env) * (double)(1ull << 30);
  size_t free_b = 0, total_b = 0;
  hip_check(hipMemGetInfo(&free_b, &total_b), "device memory info");
  return 0.5 * (double)(free_b + held_bytes);
}

void fill_zero_launch(double *A, int64_t rows, int64_t cols, int64_t ld, hipStream_t s) {
  const int64_t n = rows * cols;
  if (n == 0) return;
  int64_t blocks = (n + NT - 1) / NT;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(NT), 0, s, A, rows, cols, ld);
  hip_check(hipGetLastError(), "fill_zero");
}

size_t atomic_coulomb_work_doubles(const AtomicDevice &d) {
  return (size_t)std::max(d.n_lm, 1) * d.Nel * (TILE2 + 4);
}

void atomic_coulomb_launch(const AtomicDevice &d, const double *P, int64_t ldp, double *J, int64_t ldj,
                           hipStream_t s, KernelTimer *timer, double *work) {
  const int64_t nbf = (int64_t)d.Nrad * d.Nang;
  fill_zero_launch(J, nbf, nbf, ldj, s);
  if (d.n_lm == 0) return;
  // work space: the handle's buffers, or one caller-given block laid out as
  // Jin [n_lm][Nel][TILE2], then jsm, jbg, preA, preB [n_lm][Nel] each
  const size_t nq = (size_t)d.n_lm * d.Nel;
  double *Jin = work ? work : d.Jin.p;
  double *jsm = work ? work + nq * TILE2 : d.jsm.p;
  double *jbg = work ? jsm + nq : d.jbg.p;
  double *preA = work ? jbg + nq : d.preA.p;
  double *preB = work ? preA + nq : d.preB.p;
  JArgs a{P, ldp, d.Nel, d.Nrad, d.Nang, d.first.p, d.ni.p, d.rs.p, d.rb.p, d.chol.p, d.chol_off.p,
          d.chol_rank.p, d.lm_L.p, d.lm_pstart.p, d.pair_kl.p, d.pair_c.p, Jin, jsm, jbg};
  {
    TimedLaunch tl(timer, "atomic_j_gather", s);
    hipLaunchKernelGGL(k_atomic_j_gather, dim3(d.n_lm * d.Nel), dim3(NT), 0, s, a);
    hip_check(hipGetLastError(), "atomic_j_gather");
  }
  hipLaunchKernelGGL(k_atomic_j_prefix, dim3((d.n_lm + 63) / 64), dim3(64), 0, s, d.n_lm, d.Nel, jsm, jbg,
                     preA, preB);
  hip_check(hipGetLastError(), "atomic_j_prefix");
  JOutArgs o{J, ldj, d.Nel, d.Nrad, d.Nang, d.first.p, d.ni.p, d.lm_L.p, d.rs.p, d.rb.p, Jin,
             preA, preB, d.out_start.p, d.out_lm.p, d.out_c.p};
  {
    TimedLaunch tl(timer, "atomic_j_expand", s);
    hipLaunchKernelGGL(k_atomic_j_expand, dim3(d.Nang * d.Nang * d.Nel), dim3(NT), 0, s, o);
    hip_check(hipGetLastError(), "atomic_j_expand");
  }
}

// Compact the static coupling lists to the density blocks that pass the
// norm screen (|P_il|_F >= 10 eps, TwoDBasis.cpp:949-952).  The result only
// changes when the screening pattern does, so it is cached by pattern.
static void compact_exchange_lists(AtomicDevice &d, hipStream_t s) {
  const int Na = d.Nang, nL = d.N_L;
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
  if (mask == d.screen_mask && !d.screen_mask.empty() && d.sym == d.list_sym) return;
  std::vector<int> act, tstart(1, 0), til;
  std::vector<double> tc;
  // flops as executed: in the symmetric mode the cross-element launch covers
  // the pairs e > f and the in-element one the blocks j <= k
  double frad_off = 0.0, frad_in = 0.0;
  int64_t nterm_in = 0;
  for (size_t a = 0; a < d.h_act.size(); ++a) {
    const size_t n0 = til.size();
    const bool in_done = !d.sym || d.h_act[a] / Na <= d.h_act[a] % Na;
    std::vector<int> starts;
    for (int L = 0; L < nL; ++L) {
      const int q0 = d.h_tstart[a * nL + L], q1 = d.h_tstart[a * nL + L + 1];
      const size_t before = til.size();
      for (int q = q0; q < q1; ++q)
        if (mask[d.h_til[q]]) {
          til.push_back(d.h_til[q]);
          tc.push_back(d.h_tc[q]);
        }
      if (til.size() > before) {
        frad_off += d.sym ? d.h_radL_low[L] : d.h_radL[L] - d.h_radL_in[L];
        if (in_done) frad_in += d.h_radL_in[L];
      }
      starts.push_back((int)til.size());
    }
    if (til.size() == n0) continue;  // nothing reaches this output block
    if (in_done) nterm_in += (int64_t)(til.size() - n0);
    act.push_back(d.h_act[a]);
    tstart.insert(tstart.end(), starts.begin(), starts.end());
  }
  d.n_act = (int)act.size();
  d.n_terms = (int64_t)til.size();
  // angular sums: 2 flops per term and tile entry of the element pairs covered
  double area_in = 0.0, area_off = 0.0;
  for (int e = 0; e < d.Nel; ++e) {
    area_in += (double)d.h_ni[e] * d.h_ni[e];
    for (int f = 0; f < d.Nel; ++f)
      if (d.sym ? e > f : e != f) area_off += (double)d.h_ni[e] * d.h_ni[f];
  }
  d.flops_in = frad_in + 2.0 * (double)nterm_in * area_in;
  d.flops_off = frad_off + 2.0 * (double)d.n_terms * area_off;
  d.flops_k_radial = frad_in + frad_off;
  d.flops_k_angular = d.flops_in + d.flops_off - d.flops_k_radial;
  if (d.dense_in && !act.empty()) {
    // runs (act, L) with terms, act-major; their term ranges are contiguous,
    // rq[r] .. rq[r+1]; grouped by L for the GEMM
    std::vector<int> rq(1, 0), rL, ars(1, 0);
    for (size_t a = 0; a < act.size(); ++a) {
      for (int L = 0; L < nL; ++L) {
        const int q0 = tstart[a * nL + L], q1 = tstart[a * nL + L + 1];
        if (q1 == q0) continue;
        rq.back() = q0;
        rL.push_back(L);
        rq.push_back(q1);
      }
      ars.push_back((int)rL.size());
    }
    const int nr = (int)rL.size(), Nel = d.Nel;
    std::vector<int> Ls(nL + 1, 0), Lruns(nr);
    for (int L : rL) ++Ls[L + 1];
    for (int L = 0; L < nL; ++L) Ls[L + 1] += Ls[L];
    std::vector<int> fill(Ls.begin(), Ls.end() - 1);
    for (int r = 0; r < nr; ++r) Lruns[fill[rL[r]]++] = r;
    std::vector<std::vector<int>> per_xcd(8);
    int grp = 0;
    for (int L = 0; L < nL; ++L) {
      const int cnt = Ls[L + 1] - Ls[L];
      if (!cnt) continue;
      for (int e = 0; e < Nel; ++e, ++grp)
        for (int c0 = 0; c0 < cnt; c0 += 32) {
          auto &v = per_xcd[grp % 8];
          v.push_back(L);
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
    d.n_runs = nr;
    d.run_q.upload(rq, s);
    d.act_rstart.upload(ars, s);
    d.L_rstart.upload(Ls, s);
    d.L_runs.upload(Lruns, s);
    d.wg_in.upload(wg, s);
    d.Ypart.alloc((size_t)nr * Nel * TILE2);
    double fr = 0.0;
    for (int e = 0; e < Nel; ++e) fr += 2.0 * std::pow((double)d.h_ni[e], 4);
    d.flops_in = (double)nr * fr + 2.0 * (double)nterm_in * area_in;
  }
  if (act.empty()) act.push_back(0);
  if (til.empty()) {
    til.push_back(0);
    tc.push_back(0.0);
  }
  d.act_jk.upload(act, s);
  d.term_start.upload(tstart, s);
  d.term_il.upload(til, s);
  {
    std::vector<int2> toff(til.size());
    for (size_t q = 0; q < til.size(); ++q)
      toff[q] = make_int2((til[q] / d.Nang) * d.Nrad, (til[q] % d.Nang) * d.Nrad);
    d.term_off.upload(toff, s);
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
    const double bytes = (double)src.size() * d.Nel * d.Nel * TILE2 * sizeof(double);
    d.tiled = bytes <= exchange_tile_budget(d.ptile.n * sizeof(double) + d.ptile32.n * sizeof(float));
    d.n_src = (int)src.size();
    d.src_blk.upload(src, s);
    d.term_tile.upload(tt, s);
  }
  d.term_c.upload(tc, s);
  hip_check(hipStreamSynchronize(s), "exchange list upload");
  d.screen_mask = mask;
  d.list_sym = d.sym;
  ++d.list_gen;
}

// Block-norm screen of P, compacted coupling lists, zeroed K; false when no
// output block is coupled (K stays zero).
static bool exchange_prologue(AtomicDevice &d, const double *P, int64_t ldp, double *K, int64_t ldk,
                              hipStream_t s, bool allow_sym) {
  const int64_t nbf = (int64_t)d.Nrad * d.Nang;
  if (d.h_act.empty()) {
    fill_zero_launch(K, nbf, nbf, ldk, s);
    return false;
  }
  block_stats_launch(P, ldp, d.Nrad, d.Nang, nullptr, nullptr, d.stats.p, d.h_stats.p, s);
  hip_check(hipStreamSynchronize(s), "block statistics sync");
  const bool symmetric = block_stats_reduce(d.h_stats.p, d.Nrad, d.Nang, d.h_bnorm.p);
  d.sym = allow_sym && !d.dense_in && symmetric_exchange_allowed() && symmetric;
  compact_exchange_lists(d, s);
  fill_zero_launch(K, nbf, nbf, ldk, s);
  return d.n_act > 0;
}

// The tiled density of this call (FP64 or FP32), when the tiles are in use.
static void build_density_tiles(AtomicDevice &d, const double *P, int64_t ldp, bool fp32, hipStream_t s,
                                KernelTimer *timer) {
  if (!d.tiled) return;
  const int64_t n_tiles = (int64_t)d.Nel * d.Nel * d.n_src, n_wg = (n_tiles + NT / 64 - 1) / (NT / 64);
  if (n_wg > 0x7fffffff) throw Error(-1, "density tile grid too large");
  TimedLaunch tl(timer, fp32 ? "atomic_ptile_f32" : "atomic_ptile", s, 0.0);
  if (fp32) {
    d.ptile32.alloc((size_t)n_tiles * TILE2);
    hipLaunchKernelGGL(k_atomic_ptile<float>, dim3((unsigned)n_wg), dim3(NT), 0, s, P, ldp, d.Nel, d.Nrad,
                       d.first.p, d.ni.p, d.Nang, d.src_blk.p, d.n_src, d.ptile32.p);
  } else {
    d.ptile.alloc((size_t)n_tiles * TILE2);
    hipLaunchKernelGGL(k_atomic_ptile<double>, dim3((unsigned)n_wg), dim3(NT), 0, s, P, ldp, d.Nel, d.Nrad,
                       d.first.p, d.ni.p, d.Nang, d.src_blk.p, d.n_src, d.ptile.p);
  }
  hip_check(hipGetLastError(), "atomic_ptile");
}

void atomic_rs_exchange_launch(AtomicDevice &d, const double *P, int64_t ldp, double *K, int64_t ldk,
                               hipStream_t s, KernelTimer *timer) {
  // the erfc pair kernel covers every element pair itself
  if (!exchange_prologue(d, P, ldp, K, ldk, s, d.rs_kind == 1)) return;
  if (d.rs_kind == 1) {
    build_density_tiles(d, P, ldp, false, s, timer);
    for (int part = 0; part < 2; ++part) {
      const int n_ef = part == 0 ? d.Nel * (d.Nel - 1) / (d.sym ? 2 : 1) : d.Nel;
      if (n_ef == 0) continue;
      const int64_t grid = (int64_t)(((n_ef + 7) / 8) * 8) * d.n_act;
      if (grid > 0x7fffffff) throw Error(-1, "exchange grid too large");
      const int *efl = part == 0 ? (d.sym ? d.ef_low.p : d.ef_off.p) : d.ef_in.p;
      KArgs<double> a{P, ldp, K, ldk, d.Nel, d.Nrad, d.Nang, d.N_L, d.n_act, d.first.p, d.ni.p, d.y_rs.p,
                      d.y_rb.p, d.y_chol.p, d.y_chol_off.p, d.y_chol_rank.p, d.act_jk.p, d.term_start.p,
                      d.term_off.p, d.term_c.p, efl, n_ef, d.rs_Lscale.p,
                      d.tiled ? d.ptile.p : nullptr, d.term_tile.p, d.n_src, (int)d.sym};
      TimedLaunch tl(timer, "atomic_rs_k_yukawa", s);
      launch_atomic_k(part, d.tiled, grid, s, a);
      hip_check(hipGetLastError(), "atomic_rs_k");
    }
  } else if (d.rs_kind == 2) {
    const int n_ef = d.Nel * d.Nel;
    const int64_t grid = (int64_t)(((n_ef + 7) / 8) * 8) * d.n_act;
    if (grid > 0x7fffffff) throw Error(-1, "exchange grid too large");
    KPairArgs a{P, ldp, K, ldk, d.Nel, d.Nrad, d.Nang, d.N_L, d.n_act, d.first.p, d.ni.p, d.e_T.p,
                d.rs_Lscale.p, d.act_jk.p, d.term_start.p, d.term_il.p, d.term_c.p};
    TimedLaunch tl(timer, "atomic_rs_k_erfc", s);
    hipLaunchKernelGGL(k_atomic_k_pair, dim3((unsigned)grid), dim3(NT), 0, s, a);
    hip_check(hipGetLastError(), "atomic_rs_k_pair");
  } else {
    throw Error(-1, "range-separated tables are not on the device");
  }
}

void atomic_exchange_launch(AtomicDevice &d, const double *P, int64_t ldp, double *K, int64_t ldk,
                            hipStream_t s, KernelTimer *timer, bool fp32) {
  const int64_t nbf = (int64_t)d.Nrad * d.Nang;
  if (!exchange_prologue(d, P, ldp, K, ldk, s, true)) return;
  if (fp32) {
    if (!d.f32_tables) {
      d.rs32.alloc(d.rs.n);
      d.rb32.alloc(d.rb.n);
      d.chol32.alloc(d.chol.n);
      to_f32_launch(d.rs.p, d.rs32.p, d.rs.n, s);
      to_f32_launch(d.rb.p, d.rb32.p, d.rb.n, s);
      to_f32_launch(d.chol.p, d.chol32.p, d.chol.n, s);
      d.f32_tables = true;
    }
    if (d.term_c32_gen != d.list_gen) {
      d.term_c32.alloc(d.term_c.n);
      to_f32_launch(d.term_c.p, d.term_c32.p, d.term_c.n, s);
      d.term_c32_gen = d.list_gen;
    }
    if (!d.tiled) {
      d.P32.alloc((size_t)nbf * nbf);
      mat_to_f32_launch(P, ldp, d.P32.p, nbf, s);
    }
  }
  build_density_tiles(d, P, ldp, fp32, s, timer);
  // two launches: element pairs e != f and the in-element tiles; boundary
  // rows/cols shared by neighbouring tiles are accumulated atomically in both
  for (int part = 0; part < 2; ++part) {
    const int n_ef = part == 0 ? d.Nel * (d.Nel - 1) / (d.sym ? 2 : 1) : d.Nel;
    if (n_ef == 0) continue;
    const char *name = part == 0 ? (fp32 ? "atomic_k_off_f32" : "atomic_k_off")
                                 : (fp32 ? "atomic_k_in_f32" : "atomic_k_in");
    const double flops = part == 0 ? d.flops_off : d.flops_in;
    if (part == 1 && d.dense_in && !fp32) {
      KInArgs g{P, ldp, d.Nel, d.Nrad, d.Nang, d.first.p, d.ni.p, d.Top.p, d.wg_in.p, d.L_rstart.p,
                d.L_runs.p, d.run_q.p, d.term_il.p, d.term_c.p, d.Ypart.p};
      KRedArgs r{K, ldk, d.Nel, d.Nrad, d.Nang, d.first.p, d.ni.p, d.act_jk.p, d.act_rstart.p, d.Ypart.p};
      TimedLaunch tl(timer, name, s, flops);
      hipLaunchKernelGGL(k_atomic_kin_gemm, dim3(d.n_wg_in), dim3(NT), 0, s, g);
      hip_check(hipGetLastError(), "atomic_kin_gemm");
      hipLaunchKernelGGL(k_atomic_kin_reduce, dim3(d.n_act * d.Nel), dim3(NT), 0, s, r);
      hip_check(hipGetLastError(), "atomic_kin_reduce");
      continue;
    }
    const int n_ef_pad = ((n_ef + 7) / 8) * 8;
    const int64_t grid = (int64_t)n_ef_pad * d.n_act;
    if (grid > 0x7fffffff) throw Error(-1, "exchange grid too large");
    const int *efl = part == 0 ? (d.sym ? d.ef_low.p : d.ef_off.p) : d.ef_in.p;
    TimedLaunch tl(timer, name, s, flops);
    if (fp32) {
      KArgs<float> a{d.P32.p, nbf, K, ldk, d.Nel, d.Nrad, d.Nang, d.N_L, d.n_act, d.first.p, d.ni.p,
                     d.rs32.p, d.rb32.p, d.chol32.p, d.chol_off.p, d.chol_rank.p, d.act_jk.p, d.term_start.p,
                     d.term_off.p, d.term_c32.p, efl, n_ef, nullptr, d.tiled ? d.ptile32.p : nullptr,
                     d.term_tile.p, d.n_src, (int)d.sym};
      launch_atomic_k(part, d.tiled, grid, s, a);
    } else {
      KArgs<double> a{P, ldp, K, ldk, d.Nel, d.Nrad, d.Nang, d.N_L, d.n_act, d.first.p, d.ni.p, d.rs.p, d.rb.p,
                      d.chol.p, d.chol_off.p, d.chol_rank.p, d.act_jk.p, d.term_start.p, d.term_off.p,
                      d.term_c.p, efl, n_ef, nullptr, d.tiled ? d.ptile.p : nullptr, d.term_tile.p, d.n_src,
                      (int)d.sym};
      launch_atomic_k(part, d.tiled, grid, s, a);
    }
    hip_check(hipGetLastError(), "atomic_k");
  }
}


// --- Batched launches over many atomic systems ------------------------------
//
// Each kernel of the Coulomb and exchange builds runs once for all systems of
// an AtomicBatch: the grid is the concatenation of the systems' grids, a
// device table holds every system's argument record and the first workgroup
// of each system, and a workgroup finds its system by binary search
// (batch_find).  Every system's grid is padded to a multiple of 8 workgroups
// where the kernel maps workgroups to XCDs, so the XCD mapping inside a
// system is the one of its own launch.

namespace {

struct FillArgs {
  double *A;
  int64_t rows, cols, ld;
  int nblk;
};

struct PrefixArgs {
  int n_lm, Nel;
  const double *jsm, *jbg;
  double *preA, *preB;
};

template <typename T>
struct PtileArgs {
  const double *P;
  int64_t ldp;
  int Nel, Nrad, Nang, n_src;
  const int *first, *ni, *src_blk;
  T *out;
};

__global__ void __launch_bounds__(NT) k_fill_zero_batch(const FillArgs *args, const int *start, int n) {
  const int s = batch_find(start, n, blockIdx.x);
  const FillArgs &a = args[s];
  fill_zero_body(a.A, a.rows, a.cols, a.ld, blockIdx.x - start[s], a.nblk);
}

__global__ void __launch_bounds__(NT) k_atomic_j_gather_batch(const JArgs *args, const int *start, int n) {
  const int s = batch_find(start, n, blockIdx.x);
  j_gather_body(args[s], blockIdx.x - start[s]);
}

__global__ void __launch_bounds__(64) k_atomic_j_prefix_batch(const PrefixArgs *args, const int *start, int n) {
  const int s = batch_find(start, n, blockIdx.x);
  const PrefixArgs &a = args[s];
  j_prefix_body((blockIdx.x - start[s]) * 64 + threadIdx.x, a.n_lm, a.Nel, a.jsm, a.jbg, a.preA, a.preB);
}

__global__ void __launch_bounds__(NT) k_atomic_j_expand_batch(const JOutArgs *args, const int *start, int n) {
  const int s = batch_find(start, n, blockIdx.x);
  j_expand_body(args[s], blockIdx.x - start[s]);
}

template <typename T>
__global__ void __launch_bounds__(NT) k_atomic_ptile_batch(const PtileArgs<T> *args, const int *start, int n) {
  const int s = batch_find(start, n, blockIdx.x);
  const PtileArgs<T> &a = args[s];
  ptile_body<T>(a.P, a.ldp, a.Nel, a.Nrad, a.first, a.ni, a.Nang, a.src_blk, a.n_src, a.out, blockIdx.x - start[s]);
}

// The exchange kernels in three workgroup shapes; the cross-element ones keep
// the occupancy hints of the single-system launches.
#define HFG_K_BATCH(NAME, INSIDE, T, ATTR)                                                              \
  template <bool TILED, int NW>                                                                         \
  __global__ void __launch_bounds__(64 * NW) ATTR NAME(const KArgs<T> *args, const int *start, int n) { \
    const int s = batch_find(start, n, blockIdx.x);                                                     \
    atomic_k_body<INSIDE, TILED, T, NW>(args[s], blockIdx.x - start[s]);                                 \
  }
// one wave per SIMD fewer than the single-system kernels: the per-system
// argument struct read from memory needs the registers (no spills)
HFG_K_BATCH(k_atomic_k_off_f64_batch, false, double, __attribute__((amdgpu_waves_per_eu(5))))
HFG_K_BATCH(k_atomic_k_off_f32_batch, false, float, __attribute__((amdgpu_waves_per_eu(7))))
HFG_K_BATCH(k_atomic_k_in_f64_batch, true, double, )
HFG_K_BATCH(k_atomic_k_in_f32_batch, true, float, )
#undef HFG_K_BATCH

template <typename T, bool INSIDE, bool TILED, int NW>
void launch_k_batch_nw(const void *args, const int *start, int n, int grid, hipStream_t s) {
  const KArgs<T> *a = static_cast<const KArgs<T> *>(args);
  if constexpr (std::is_same<T, double>::value) {
    if constexpr (INSIDE)
      hipLaunchKernelGGL((k_atomic_k_in_f64_batch<TILED, NW>), dim3(grid), dim3(64 * NW), 0, s, a, start, n);
    else
      hipLaunchKernelGGL((k_atomic_k_off_f64_batch<TILED, NW>), dim3(grid), dim3(64 * NW), 0, s, a, start, n);
  } else {
    if constexpr (INSIDE)
      hipLaunchKernelGGL((k_atomic_k_in_f32_batch<TILED, NW>), dim3(grid), dim3(64 * NW), 0, s, a, start, n);
    else
      hipLaunchKernelGGL((k_atomic_k_off_f32_batch<TILED, NW>), dim3(grid), dim3(64 * NW), 0, s, a, start, n);
  }
}

// cls = fp32 * 2 + tiled; part 0 cross-element, 1 in-element; nw in {1, 2, 4}
void launch_k_batch(int cls, int part, int nw, const void *args, const int *start, int n, int grid, hipStream_t s) {
  if (grid <= 0) return;
#define HFG_KB(T, IN, TL)                                                     \
  do {                                                                        \
    if (nw == 1)                                                              \
      launch_k_batch_nw<T, IN, TL, 1>(args, start, n, grid, s);               \
    else if (nw == 2)                                                         \
      launch_k_batch_nw<T, IN, TL, 2>(args, start, n, grid, s);               \
    else                                                                      \
      launch_k_batch_nw<T, IN, TL, 4>(args, start, n, grid, s);               \
  } while (0)
  const bool f32 = cls >= 2, tiled = cls & 1;
  if (!f32 && !part && !tiled) HFG_KB(double, false, false);
  else if (!f32 && !part && tiled) HFG_KB(double, false, true);
  else if (!f32 && part && !tiled) HFG_KB(double, true, false);
  else if (!f32 && part && tiled) HFG_KB(double, true, true);
  else if (f32 && !part && !tiled) HFG_KB(float, false, false);
  else if (f32 && !part && tiled) HFG_KB(float, false, true);
  else if (f32 && part && !tiled) HFG_KB(float, true, false);
  else HFG_KB(float, true, true);
#undef HFG_KB
  hip_check(hipGetLastError(), "atomic_k_batch");
}

// Host image of the device argument tables: sections aligned to 256 bytes.
struct TabImage {
  std::vector<char> buf;
  template <typename X>
  size_t add(const std::vector<X> &v) {
    const size_t off = (buf.size() + 255) & ~(size_t)255;
    buf.resize(off + std::max<size_t>(v.size(), 1) * sizeof(X), 0);
    if (!v.empty()) std::memcpy(buf.data() + off, v.data(), v.size() * sizeof(X));
    return off;
  }
};

// Copy a table image to the device through the pinned staging buffer (after
// the previous copy out of it has finished); returns the device base.
char *upload_table(const TabImage &img, DevArray<char> &dtab, PinnedArray<char> &htab, hipEvent_t &ev,
                   hipStream_t s) {
  if (ev) hip_check(hipEventSynchronize(ev), "table staging wait");
  else hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
  htab.alloc(img.buf.size());
  dtab.alloc(img.buf.size());
  std::memcpy(htab.p, img.buf.data(), img.buf.size());
  hip_check(hipMemcpyAsync(dtab.p, htab.p, img.buf.size(), hipMemcpyHostToDevice, s), "table upload");
  hip_check(hipEventRecord(ev, s), "hipEventRecord");
  return dtab.p;
}

int fill_blocks(int64_t rows, int64_t cols) {
  const int64_t n = rows * cols;
  return (int)std::max<int64_t>(1, std::min<int64_t>(4096, (n + 8 * NT - 1) / (8 * NT)));
}

// Autotuned workgroup shapes, per (device, class, part, size bucket), shared
// by every batch of the process.
std::mutex g_tune_mu;
std::map<std::array<int, 4>, int> g_tuned;

int tuned_shape(int dev, int c, int p, int bucket) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  const auto it = g_tuned.find({dev, c, p, bucket});
  return it == g_tuned.end() ? 0 : it->second;
}

void remember_shape(int dev, int c, int p, int bucket, int nw) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned[{dev, c, p, bucket}] = nw;
}

// Run `launches` as a graph captured for `sig` (re-captured when sig
// changes), or directly when graphs are off or the stream cannot capture
// (the legacy null stream).
template <typename F>
void run_graphed(bool use, hipGraphExec_t &gexec, std::vector<long long> &cur, const std::vector<long long> &sig,
                 hipStream_t s, F launches) {
  if (!use || s == nullptr) {
    launches();
    return;
  }
  if (!gexec || cur != sig) {
    if (gexec) {
      (void)hipGraphExecDestroy(gexec);
      gexec = nullptr;
    }
    hipGraph_t g = nullptr;
    hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      launches();
    } catch (...) {
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    hip_check(hipStreamEndCapture(s, &g), "hipStreamEndCapture");
    const hipError_t e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    hip_check(e, "hipGraphInstantiate");
    cur = sig;
  }
  hip_check(hipGraphLaunch(gexec, s), "hipGraphLaunch");
}

}  // namespace

AtomicBatch::~AtomicBatch() {
  if (gexec_j) (void)hipGraphExecDestroy(gexec_j);
  if (gexec_k) (void)hipGraphExecDestroy(gexec_k);
  for (hipEvent_t e : {ev_j, ev_k, ev_s})
    if (e) (void)hipEventDestroy(e);
}

void atomic_batch_coulomb(AtomicBatch &b, const double *const *P, const int64_t *ldp, double *const *J,
                          const int64_t *ldj, hipStream_t s, KernelTimer *timer) {
  const int n = (int)b.sys.size();
  std::vector<FillArgs> fills(n);
  std::vector<JArgs> gat(n);
  std::vector<PrefixArgs> pre(n);
  std::vector<JOutArgs> outs(n);
  std::vector<int> sF(n + 1, 0), sG(n + 1, 0), sP(n + 1, 0), sE(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    AtomicDevice &d = *b.sys[i];
    const int64_t nbf = (int64_t)d.Nrad * d.Nang;
    fills[i] = FillArgs{J[i], nbf, nbf, ldj[i], fill_blocks(nbf, nbf)};
    sF[i + 1] = sF[i] + fills[i].nblk;
    gat[i] = JArgs{P[i], ldp[i], d.Nel, d.Nrad, d.Nang, d.first.p, d.ni.p, d.rs.p, d.rb.p, d.chol.p,
                   d.chol_off.p, d.chol_rank.p, d.lm_L.p, d.lm_pstart.p, d.pair_kl.p, d.pair_c.p,
                   d.Jin.p, d.jsm.p, d.jbg.p};
    pre[i] = PrefixArgs{d.n_lm, d.Nel, d.jsm.p, d.jbg.p, d.preA.p, d.preB.p};
    outs[i] = JOutArgs{J[i], ldj[i], d.Nel, d.Nrad, d.Nang, d.first.p, d.ni.p, d.lm_L.p, d.rs.p, d.rb.p,
                       d.Jin.p, d.preA.p, d.preB.p, d.out_start.p, d.out_lm.p, d.out_c.p};
    const bool any = d.n_lm > 0;
    sG[i + 1] = sG[i] + (any ? d.n_lm * d.Nel : 0);
    sP[i + 1] = sP[i] + (any ? (d.n_lm + 63) / 64 : 0);
    sE[i + 1] = sE[i] + (any ? d.Nang * d.Nang * d.Nel : 0);
  }
  TabImage img;
  const size_t oF = img.add(fills), oG = img.add(gat), oP = img.add(pre), oE = img.add(outs);
  const size_t osF = img.add(sF), osG = img.add(sG), osP = img.add(sP), osE = img.add(sE);
  char *base = upload_table(img, b.dtab_j, b.htab_j, b.ev_j, s);
  const std::vector<long long> sig{sF[n], sG[n], sP[n], sE[n]};
  auto launches = [&](KernelTimer *t) {
    const int *dsF = (const int *)(base + osF), *dsG = (const int *)(base + osG);
    const int *dsP = (const int *)(base + osP), *dsE = (const int *)(base + osE);
    hipLaunchKernelGGL(k_fill_zero_batch, dim3(sF[n]), dim3(NT), 0, s, (const FillArgs *)(base + oF), dsF, n);
    hip_check(hipGetLastError(), "fill_zero_batch");
    if (sG[n] == 0) return;
    {
      TimedLaunch tl(t, "batch_j_gather", s);
      hipLaunchKernelGGL(k_atomic_j_gather_batch, dim3(sG[n]), dim3(NT), 0, s, (const JArgs *)(base + oG), dsG, n);
      hip_check(hipGetLastError(), "atomic_j_gather_batch");
    }
    hipLaunchKernelGGL(k_atomic_j_prefix_batch, dim3(sP[n]), dim3(64), 0, s, (const PrefixArgs *)(base + oP), dsP,
                       n);
    hip_check(hipGetLastError(), "atomic_j_prefix_batch");
    {
      TimedLaunch tl(t, "batch_j_expand", s);
      hipLaunchKernelGGL(k_atomic_j_expand_batch, dim3(sE[n]), dim3(NT), 0, s, (const JOutArgs *)(base + oE), dsE,
                         n);
      hip_check(hipGetLastError(), "atomic_j_expand_batch");
    }
  };
  if (timer && timer->on)
    launches(timer);
  else
    run_graphed(b.use_graphs, b.gexec_j, b.sig_j, sig, s, [&] { launches(nullptr); });
}

void atomic_batch_exchange(AtomicBatch &b, const double *const *P, const int64_t *ldp, double *const *K,
                           const int64_t *ldk, const int *fp32, hipStream_t s, KernelTimer *timer) {
  const int n = (int)b.sys.size();
  // 1. block statistics of every density in one launch, one copy, one sync
  {
    std::vector<StatsArgs> st(n);
    std::vector<int> sS(n + 1, 0);
    b.stats_off.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
      AtomicDevice &d = *b.sys[i];
      const int nch = block_stats_chunks(d.Nrad);
      const size_t cnt = d.h_act.empty() ? 0 : (size_t)d.Nang * d.Nang * 4 * nch;
      b.stats_off[i + 1] = b.stats_off[i] + cnt;
      st[i] = StatsArgs{P[i], ldp[i], d.Nrad, d.Nang, nch, nullptr, nullptr, nullptr};
      sS[i + 1] = sS[i] + (cnt ? d.Nang * d.Nang * nch : 0);
    }
    b.stats.alloc(std::max<size_t>(b.stats_off[n], 1));
    b.h_stats.alloc(std::max<size_t>(b.stats_off[n], 1));
    for (int i = 0; i < n; ++i) st[i].part = b.stats.p + b.stats_off[i];
    TabImage img;
    const size_t oA = img.add(st), oS = img.add(sS);
    char *base = upload_table(img, b.dtab_s, b.htab_s, b.ev_s, s);
    {
      TimedLaunch tl(timer, "batch_stats", s);
      block_stats_batch_launch((const StatsArgs *)(base + oA), (const int *)(base + oS), n, sS[n], s);
    }
    if (b.stats_off[n])
      hip_check(hipMemcpyAsync(b.h_stats.p, b.stats.p, sizeof(double) * b.stats_off[n], hipMemcpyDeviceToHost, s),
                "copy block statistics");
    hip_check(hipStreamSynchronize(s), "block statistics sync");
  }
  // 2. per system: screen, symmetry mode, compacted lists, reduced-precision tables
  for (int i = 0; i < n; ++i) {
    AtomicDevice &d = *b.sys[i];
    if (d.h_act.empty()) continue;
    const bool symmetric = block_stats_reduce(b.h_stats.p + b.stats_off[i], d.Nrad, d.Nang, d.h_bnorm.p);
    d.sym = !d.dense_in && symmetric_exchange_allowed() && symmetric;
    compact_exchange_lists(d, s);
    if (fp32[i] && d.n_act > 0) {
      if (!d.f32_tables) {
        d.rs32.alloc(d.rs.n);
        d.rb32.alloc(d.rb.n);
        d.chol32.alloc(d.chol.n);
        to_f32_launch(d.rs.p, d.rs32.p, d.rs.n, s);
        to_f32_launch(d.rb.p, d.rb32.p, d.rb.n, s);
        to_f32_launch(d.chol.p, d.chol32.p, d.chol.n, s);
        d.f32_tables = true;
      }
      if (d.term_c32_gen != d.list_gen) {
        d.term_c32.alloc(d.term_c.n);
        to_f32_launch(d.term_c.p, d.term_c32.p, d.term_c.n, s);
        d.term_c32_gen = d.list_gen;
      }
      const int64_t nbf = (int64_t)d.Nrad * d.Nang;
      if (!d.tiled) {
        d.P32.alloc((size_t)nbf * nbf);
        mat_to_f32_launch(P[i], ldp[i], d.P32.p, nbf, s);
      }
    }
    if (d.tiled && d.n_act > 0) {
      const size_t nt = (size_t)d.Nel * d.Nel * d.n_src * TILE2;
      if (fp32[i])
        d.ptile32.alloc(nt);
      else
        d.ptile.alloc(nt);
    }
  }
  // 3. argument tables per class (fp32 * 2 + tiled)
  std::vector<FillArgs> fills(n);
  std::vector<int> sF(n + 1, 0);
  std::vector<KArgs<double>> kd[2][2];  // [tiled][part]
  std::vector<KArgs<float>> kf[2][2];
  std::vector<int> sk[4][2];
  std::vector<PtileArgs<double>> pd;
  std::vector<PtileArgs<float>> pf;
  std::vector<int> spd(1, 0), spf(1, 0);
  for (int c = 0; c < 4; ++c)
    for (int p = 0; p < 2; ++p) sk[c][p].assign(1, 0);
  double flops[4][2] = {};
  for (int i = 0; i < n; ++i) {
    AtomicDevice &d = *b.sys[i];
    const int64_t nbf = (int64_t)d.Nrad * d.Nang;
    fills[i] = FillArgs{K[i], nbf, nbf, ldk[i], fill_blocks(nbf, nbf)};
    sF[i + 1] = sF[i] + fills[i].nblk;
    if (d.h_act.empty() || d.n_act == 0) continue;
    const bool f = fp32[i] != 0, tl = d.tiled;
    const int cls = (f ? 2 : 0) + (tl ? 1 : 0);
    if (tl) {
      const int64_t n_tiles = (int64_t)d.Nel * d.Nel * d.n_src, wg = (n_tiles + NT / 64 - 1) / (NT / 64);
      if (f) {
        pf.push_back(PtileArgs<float>{P[i], ldp[i], d.Nel, d.Nrad, d.Nang, d.n_src, d.first.p, d.ni.p,
                                      d.src_blk.p, d.ptile32.p});
        spf.push_back(spf.back() + (int)wg);
      } else {
        pd.push_back(PtileArgs<double>{P[i], ldp[i], d.Nel, d.Nrad, d.Nang, d.n_src, d.first.p, d.ni.p,
                                       d.src_blk.p, d.ptile.p});
        spd.push_back(spd.back() + (int)wg);
      }
    }
    for (int part = 0; part < 2; ++part) {
      const int n_ef = part == 0 ? d.Nel * (d.Nel - 1) / (d.sym ? 2 : 1) : d.Nel;
      const int n_ef_pad = ((n_ef + 7) / 8) * 8;
      const int64_t grid = n_ef == 0 ? 0 : (int64_t)n_ef_pad * d.n_act;
      if ((int64_t)sk[cls][part].back() + grid > 0x7fffffff) throw Error(-1, "batched exchange grid too large");
      const int *efl = part == 0 ? (d.sym ? d.ef_low.p : d.ef_off.p) : d.ef_in.p;
      if (f)
        kf[tl][part].push_back(KArgs<float>{tl ? nullptr : d.P32.p, nbf, K[i], ldk[i], d.Nel, d.Nrad, d.Nang, d.N_L,
                                            d.n_act, d.first.p, d.ni.p, d.rs32.p, d.rb32.p, d.chol32.p,
                                            d.chol_off.p, d.chol_rank.p, d.act_jk.p, d.term_start.p, d.term_off.p,
                                            d.term_c32.p, efl, n_ef, nullptr, tl ? d.ptile32.p : nullptr,
                                            d.term_tile.p, d.n_src, (int)d.sym});
      else
        kd[tl][part].push_back(KArgs<double>{P[i], ldp[i], K[i], ldk[i], d.Nel, d.Nrad, d.Nang, d.N_L, d.n_act,
                                             d.first.p, d.ni.p, d.rs.p, d.rb.p, d.chol.p, d.chol_off.p,
                                             d.chol_rank.p, d.act_jk.p, d.term_start.p, d.term_off.p, d.term_c.p,
                                             efl, n_ef, nullptr, tl ? d.ptile.p : nullptr, d.term_tile.p, d.n_src,
                                             (int)d.sym});
      sk[cls][part].push_back(sk[cls][part].back() + (int)grid);
      flops[cls][part] += part == 0 ? d.flops_off : d.flops_in;
    }
  }
  TabImage img;
  const size_t oF = img.add(fills), osF = img.add(sF);
  const size_t opd = img.add(pd), ospd = img.add(spd), opf = img.add(pf), ospf = img.add(spf);
  size_t ok[4][2], osk[4][2];
  for (int c = 0; c < 4; ++c)
    for (int p = 0; p < 2; ++p) {
      const bool f = c >= 2, tl = c & 1;
      ok[c][p] = f ? img.add(kf[tl][p]) : img.add(kd[tl][p]);
      osk[c][p] = img.add(sk[c][p]);
    }
  char *base = upload_table(img, b.dtab_k, b.htab_k, b.ev_k, s);
  auto nsys = [&](int c, int p) { return (int)sk[c][p].size() - 1; };
  auto fill_all = [&] {
    hipLaunchKernelGGL(k_fill_zero_batch, dim3(sF[n]), dim3(NT), 0, s, (const FillArgs *)(base + oF),
                       (const int *)(base + osF), n);
    hip_check(hipGetLastError(), "fill_zero_batch");
  };
  auto ptiles = [&](KernelTimer *t) {
    TimedLaunch tl(t, "batch_ptile", s);
    if (spd.back() > 0)
      hipLaunchKernelGGL(k_atomic_ptile_batch<double>, dim3(spd.back()), dim3(NT), 0, s,
                         (const PtileArgs<double> *)(base + opd), (const int *)(base + ospd), (int)pd.size());
    if (spf.back() > 0)
      hipLaunchKernelGGL(k_atomic_ptile_batch<float>, dim3(spf.back()), dim3(NT), 0, s,
                         (const PtileArgs<float> *)(base + opf), (const int *)(base + ospf), (int)pf.size());
    hip_check(hipGetLastError(), "atomic_ptile_batch");
  };
  auto kpart = [&](int c, int p, int nw) {
    launch_k_batch(c, p, nw, base + ok[c][p], (const int *)(base + osk[c][p]), nsys(c, p), sk[c][p].back(), s);
  };
  // 4. the workgroup shape of each class and part: tuned once per size bucket
  // (log2 of the launch's workgroup count) on this device and kept for the
  // process, so batches whose composition changes every iteration (systems
  // converging, switching precision) reuse the choice instead of re-timing
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  bool tiles_ready = false;
  for (int c = 0; c < 4; ++c)
    for (int p = 0; p < 2; ++p) {
      const int wgs = sk[c][p].back();
      if (wgs == 0) continue;
      if (!b.autotune) {
        b.nw[c][p] = 4;
        continue;
      }
      const int bucket = 31 - __builtin_clz((unsigned)wgs);
      const int known = tuned_shape(dev, c, p, bucket);
      if (known > 0) {
        b.nw[c][p] = known;
        continue;
      }
      if (!tiles_ready) {
        ptiles(nullptr);
        tiles_ready = true;
      }
      hipEvent_t e0, e1;
      hip_check(hipEventCreate(&e0), "hipEventCreate");
      hip_check(hipEventCreate(&e1), "hipEventCreate");
      const int shapes[3] = {1, 2, 4};
      float best = 1e30f;
      for (int q = 0; q < 3; ++q) {
        float ms = 0.f;
        for (int rep = 0; rep < 2; ++rep) {  // the second run is timed
          fill_all();
          hip_check(hipEventRecord(e0, s), "hipEventRecord");
          kpart(c, p, shapes[q]);
          hip_check(hipEventRecord(e1, s), "hipEventRecord");
          hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
          hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        }
        b.tune_ms[c][p][q] = ms;
        if (ms < best) {
          best = ms;
          b.nw[c][p] = shapes[q];
        }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      remember_shape(dev, c, p, bucket, b.nw[c][p]);
    }
  // 5. the launch sequence: zero K, density tiles, cross- and in-element kernels
  std::vector<long long> sig{sF[n], spd.back(), spf.back()};
  for (int c = 0; c < 4; ++c)
    for (int p = 0; p < 2; ++p) {
      sig.push_back(sk[c][p].back());
      sig.push_back(b.nw[c][p]);
    }
  static const char *names[4][2] = {{"batch_k_off", "batch_k_in"},
                                    {"batch_k_off", "batch_k_in"},
                                    {"batch_k_off_f32", "batch_k_in_f32"},
                                    {"batch_k_off_f32", "batch_k_in_f32"}};
  auto launches = [&](KernelTimer *t) {
    fill_all();
    ptiles(t);
    for (int p = 0; p < 2; ++p)
      for (int c = 0; c < 4; ++c) {
        if (sk[c][p].back() == 0) continue;
        TimedLaunch tl(t, names[c][p], s, flops[c][p]);
        kpart(c, p, b.nw[c][p] ? b.nw[c][p] : 4);
      }
  };
  if (timer && timer->on)
    launches(timer);
  else
    run_graphed(b.use_graphs, b.gexec_k, b.sig_k, sig, s, [&] { launches(nullptr); });
}

}  // namespace hfg
