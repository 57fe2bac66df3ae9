// Diatomic (prolate spheroidal) Coulomb and exchange kernels (FP64), gfx950.
//
// Data layout (device):
//   P, J, K   dense Nbf x Nbf column-major in the pure basis; shell a's
//             radial function r is row poff[a] + r - skip[a] (r >= skip[a]).
//   pq        per ((L,|M|), element) four 16x16 zero-padded tiles: the
//             disjoint integrals P0, P2, Q0, Q2 (basis.cpp:352-394).
//   chol      per ((L,|M|), element) and Cholesky vector p the two channel
//             blocks M0_p(a,b) = B(a + b Ni, p), M2_p(a,b) = B(Ni^2 + a + b Ni, p)
//             of W = B diag(sigma) B' (basis.cpp:1492-1537), and sigma.
//
// Coulomb (basis.cpp:1627-1816): two density channels per (L,M),
// Paux0 = sum mod_coeff * P_kl and Paux2 = sum coeff * P_kl, of which only
// the element-diagonal bands matter.  Kernel 1 gathers both bands, reduces
// the P/Q traces and applies the in-element kernel through the Cholesky
// factor; kernel 2 prefix-sums the cross-element traces; kernel 3 expands
// the bands into the shell blocks of J.
// Exchange (basis.cpp:1818-2089): one workgroup per (output block (j,k),
// element pair (e,f)); for each coupled (L,|M|) the four Gaunt-weighted
// density tiles R00, R02, R20, R22 are formed in registers and contracted on
// the matrix cores with the disjoint P/Q factors (e != f) or the four-channel
// Cholesky factor (e == f).  The density tiles are first copied out of P into
// a tiled array (k_dia_ptile: one contiguous, aligned 2 KB tile per (element
// pair, density block) in operand order), so a term's gather is one aligned
// 32-byte read per lane and an element pair's slice (3.3 MB for N2 at lmax
// 16) stays in the L2 of the XCD that works on it.
#include "hfgpu_internal.hpp"
#include "device_common.hpp"

// terms in flight per lane in the cross-element gathers, and that launch's
// waves per SIMD
#ifndef HFG_DIA_OFF_UNROLL
#define HFG_DIA_OFF_UNROLL 4
#endif
#ifndef HFG_DIA_OFF_WAVES
#define HFG_DIA_OFF_WAVES 5
#endif

namespace hfg {
namespace {

using namespace dev;

__device__ inline long long pure_row(const int *skip, const long long *poff, int shell, int r) {
  const int s = skip[shell];
  return r >= s ? poff[shell] + (r - s) : -1;
}

// --- Coulomb -----------------------------------------------------------------

struct DJArgs {
  const double *P;
  int64_t ldp;
  int Nel, Nang;
  const int *first, *ni, *skip;
  const long long *poff;
  const double *pq, *chol, *sigma;
  const int *chol_off, *chol_rank;
  const int *LM_ilm;
  const double *LM_fac;
  const int *pstart, *pair_kl;
  const double *pair_c;
  double *Jin, *ds, *db;
};

// One workgroup per ((L,M), element).
__global__ void __launch_bounds__(NT) k_dia_j_gather(DJArgs A) {
  const int q = blockIdx.x / A.Nel, e = blockIdx.x - q * A.Nel;
  const int t = threadIdx.x, a = t & 15, b = t >> 4;
  const int ni = A.ni[e], f0 = A.first[e], ilm = A.LM_ilm[q];
  const double fac = A.LM_fac[q];
  __shared__ double s0[TILE2], s2[TILE2];
  __shared__ double sc[2 * TILE2];
  __shared__ double red[8];
  double v0 = 0.0, v2 = 0.0;
  if (a < ni && b < ni) {
    for (int p = A.pstart[q]; p < A.pstart[q + 1]; ++p) {
      const int kl = A.pair_kl[p], k = kl / A.Nang, l = kl - k * A.Nang;
      const long long r = pure_row(A.skip, A.poff, k, f0 + a), c = pure_row(A.skip, A.poff, l, f0 + b);
      if (r < 0 || c < 0) continue;
      const double x = A.P[r + c * A.ldp];
      v0 += A.pair_c[2 * p] * x;
      v2 += A.pair_c[2 * p + 1] * x;
    }
  }
  s0[t] = v0;
  s2[t] = v2;
  __syncthreads();
  // traces tr(X s) = sum_ab X(a,b) s(b,a)
  const size_t tb = ((size_t)ilm * A.Nel + e) * 4 * TILE2;
  const double x0 = s0[b + TILE * a], x2 = s2[b + TILE * a];
  double dsv = A.pq[tb + t] * x0 - A.pq[tb + TILE2 + t] * x2;
  double dbv = A.pq[tb + 2 * TILE2 + t] * x0 - A.pq[tb + 3 * TILE2 + t] * x2;
  block_sum2(dsv, dbv, red);
  if (t == 0) {
    A.ds[(size_t)q * A.Nel + e] = fac * dsv;
    A.db[(size_t)q * A.Nel + e] = fac * dbv;
  }
  // c_p = sigma_p <[M0_p; M2_p], [s0; s2]>, wave w handles p = w, w+4, ...
  const int rank = A.chol_rank[ilm * A.Nel + e], off = A.chol_off[ilm * A.Nel + e];
  const double *M = A.chol + (size_t)off * 2 * TILE2;
  const int lane = t & 63, w = t >> 6;
  for (int p = w; p < rank; p += 4) {
    const double *M0 = M + (size_t)p * 2 * TILE2, *M2 = M0 + TILE2;
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) s += M0[lane + 64 * u] * s0[lane + 64 * u] + M2[lane + 64 * u] * s2[lane + 64 * u];
    s = wave_sum(s);
    if (lane == 0) sc[p] = s * A.sigma[off + p];
  }
  __syncthreads();
  double j0 = 0.0, j2 = 0.0;
  for (int p = 0; p < rank; ++p) {
    const double cp = sc[p];
    j0 += M[(size_t)p * 2 * TILE2 + t] * cp;
    j2 += M[(size_t)p * 2 * TILE2 + TILE2 + t] * cp;
  }
  double *out = A.Jin + ((size_t)q * A.Nel + e) * 2 * TILE2;
  out[t] = j0 * fac;
  out[TILE2 + t] = j2 * fac;
}

// preA(q,e) = sum_{j>e} db(q,j), preB(q,e) = sum_{j<e} ds(q,j)
__global__ void k_dia_j_prefix(int n_q, int Nel, const double *ds, const double *db, double *preA, double *preB) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_q) return;
  double acc = 0.0;
  for (int e = 0; e < Nel; ++e) {
    preB[(size_t)q * Nel + e] = acc;
    acc += ds[(size_t)q * Nel + e];
  }
  acc = 0.0;
  for (int e = Nel - 1; e >= 0; --e) {
    preA[(size_t)q * Nel + e] = acc;
    acc += db[(size_t)q * Nel + e];
  }
}

struct DJOutArgs {
  double *J;
  int64_t ldj;
  int Nel, Nang;
  const int *first, *ni, *skip;
  const long long *poff;
  const double *pq, *Jin, *preA, *preB;
  const int *LM_ilm, *out_start, *out_q;
  const double *out_c;
};

// Ja0 = Jin0 + preA P0 + preB Q0, Ja2 = Jin2 - preA P2 - preB Q2 at entry t
__device__ inline void dia_band(const DJOutArgs &A, int q, int e, int t, double &x0, double &x2) {
  const size_t tb = ((size_t)A.LM_ilm[q] * A.Nel + e) * 4 * TILE2;
  const size_t qe = (size_t)q * A.Nel + e;
  const double pa = A.preA[qe], pb = A.preB[qe];
  const double *jin = A.Jin + qe * 2 * TILE2;
  x0 = jin[t] + pa * A.pq[tb + t] + pb * A.pq[tb + 2 * TILE2 + t];
  x2 = jin[TILE2 + t] - pa * A.pq[tb + TILE2 + t] - pb * A.pq[tb + 3 * TILE2 + t];
}

// One workgroup per (output shell block (i,j), element).
__global__ void __launch_bounds__(NT) k_dia_j_expand(DJOutArgs A) {
  const int ij = blockIdx.x / A.Nel, e = blockIdx.x - ij * A.Nel;
  const int o0 = A.out_start[ij], o1 = A.out_start[ij + 1];
  if (o0 == o1) return;
  const int t = threadIdx.x, a = t & 15, b = t >> 4;
  const int ni = A.ni[e], f0 = A.first[e];
  if (a >= ni || b >= ni) return;
  if (e > 0 && a == 0 && b == 0) return;  // shared corner, written by element e-1
  const int i = ij / A.Nang, j = ij - i * A.Nang;
  const long long row = pure_row(A.skip, A.poff, i, f0 + a), col = pure_row(A.skip, A.poff, j, f0 + b);
  if (row < 0 || col < 0) return;
  const bool corner = (a == ni - 1 && b == ni - 1 && e + 1 < A.Nel);
  double v = 0.0;
  for (int o = o0; o < o1; ++o) {
    const int q = A.out_q[o];
    double x0, x2;
    dia_band(A, q, e, t, x0, x2);
    if (corner) {
      double y0, y2;
      dia_band(A, q, e + 1, 0, y0, y2);
      x0 += y0;
      x2 += y2;
    }
    v += A.out_c[2 * o] * x0 + A.out_c[2 * o + 1] * x2;
  }
  A.J[row + col * A.ldj] = v;
}

// --- Exchange ----------------------------------------------------------------

// One pass over P per shell block pair i <= l and 32-row chunk ch of P_il:
// partial sums for the Frobenius norms of P_il and P_li (the reference's
// screen) and partial maxima for the symmetry test, part[(il * nch + ch) *
// 4 + {0..3}] = (sum P_il^2, sum P_li^2, max |P_il - P_li'|, max |P_il|),
// reduced on the host.  32x32 sub-tiles of both blocks are read coalesced
// into LDS and compared transposed.  Shell a spans rows poff[a] .. poff[a] +
// Nrad - skip[a] (a Nrad, no skip when the tables are null: atomic basis).
__device__ inline void block_stats_body(const double *P, int64_t ldp, int Nrad, int Nang, const int *skip,
                                        const long long *poff, double *part, int il, int ch, int nch) {
  const int i = il / Nang, l = il - i * Nang;
  if (i > l) return;  // covered by the workgroups of (l, i)
  __shared__ double A[32][33], B[32][33];
  __shared__ double red[16];
  double amax = 0.0, pmax = 0.0, ssa = 0.0, ssb = 0.0;
  const int ri = Nrad - (skip ? skip[i] : 0), rl = Nrad - (skip ? skip[l] : 0);
  const long long oi = poff ? poff[i] : (long long)i * Nrad, ol = poff ? poff[l] : (long long)l * Nrad;
  const int t = threadIdx.x, x = t & 31, y0 = t >> 5;
  for (int tr = 32 * ch; tr < ri; tr += 32 * nch)
    for (int tc = 0; tc < rl; tc += 32) {
      __syncthreads();
      // A[x][y] = P_il(tr + x, tc + y), B[x][y] = P_li(tc + x, tr + y)
      for (int y = y0; y < 32; y += NT / 32) {
        const bool va = tr + x < ri && tc + y < rl, vb = tc + x < rl && tr + y < ri;
        const double pa = va ? P[(oi + tr + x) + (ol + tc + y) * ldp] : 0.0;
        const double pb = vb ? P[(ol + tc + x) + (oi + tr + y) * ldp] : 0.0;
        A[x][y] = pa;
        B[x][y] = pb;
        ssa += pa * pa;
        ssb += pb * pb;
      }
      __syncthreads();
      for (int y = y0; y < 32; y += NT / 32) {
        const double dlt = fabs(A[x][y] - B[y][x]);
        amax = isnan(dlt) ? HUGE_VAL : fmax(amax, dlt);  // a NaN counts as asymmetric
        pmax = fmax(pmax, fabs(A[x][y]));
      }
    }
  for (int o = 32; o > 0; o >>= 1) {
    amax = fmax(amax, __shfl_xor(amax, o));
    pmax = fmax(pmax, __shfl_xor(pmax, o));
    ssa += __shfl_xor(ssa, o);
    ssb += __shfl_xor(ssb, o);
  }
  const int w = t >> 6;
  if ((t & 63) == 0) {
    red[w] = amax;
    red[4 + w] = pmax;
    red[8 + w] = ssa;
    red[12 + w] = ssb;
  }
  __syncthreads();
  if (t == 0) {
    double *o = part + ((size_t)il * nch + ch) * 4;
    o[0] = (red[8] + red[9]) + (red[10] + red[11]);
    o[1] = (red[12] + red[13]) + (red[14] + red[15]);
    o[2] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    o[3] = fmax(fmax(red[4], red[5]), fmax(red[6], red[7]));
  }
}

__global__ void __launch_bounds__(NT) k_block_stats(const double *P, int64_t ldp, int Nrad, int Nang,
                                                    const int *skip, const long long *poff, double *part) {
  block_stats_body(P, ldp, Nrad, Nang, skip, poff, part, blockIdx.x, blockIdx.y, gridDim.y);
}

// Many matrices in one launch: workgroup b of system s (start[s] <= b <
// start[s+1]) takes shell pair il = local / nch and chunk local % nch.
__global__ void __launch_bounds__(NT) k_block_stats_batch(const StatsArgs *args, const int *start, int n) {
  const int s = batch_find(start, n, blockIdx.x);
  const StatsArgs &a = args[s];
  const int loc = blockIdx.x - start[s], nch = a.nch;
  block_stats_body(a.P, a.ldp, a.Nrad, a.Nang, a.skip, a.poff, a.part, loc / nch, loc - (loc / nch) * nch, nch);
}

// Two products sharing the Y operand, issued interleaved so the two
// accumulation chains overlap: da += Xa Y, db += Xb Y (16x16x16 each).
template <typename T>
__device__ inline void mma_pair(const T xa[4], const T xb[4], const T ys[4], typename Mma<T>::acc_t &da,
                                typename Mma<T>::acc_t &db) {
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    da = Mma<T>::step(xa[s], ys[s], da);
    db = Mma<T>::step(xb[s], ys[s], db);
  }
}

// da += Xa Ya, db += Xb Yb with Ya, Yb result tiles, interleaved.
template <typename T>
__device__ inline void mma_pair2(const T xa[4], typename Mma<T>::acc_t ya, typename Mma<T>::acc_t &da,
                                 const T xb[4], typename Mma<T>::acc_t yb, typename Mma<T>::acc_t &db) {
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    da = Mma<T>::step(xa[s], ya[s], da);
    db = Mma<T>::step(xb[s], yb[s], db);
  }
}

// Cholesky channel tiles of vector p in both operand layouts:
// y = M(c, k) (M' as Y operand), x = M(k, c) (M' as X operand), k = row(g, s).
template <typename T>
__device__ inline void load_chol(const T *M, int p, const int kk[4], int c, T y0[4], T y2[4], T x0[4], T x2[4]) {
  const T *M0 = M + (size_t)p * 2 * TILE2, *M2 = M0 + TILE2;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    y0[s] = M0[c + TILE * kk[s]];
    y2[s] = M2[c + TILE * kk[s]];
    x0[s] = M0[kk[s] + TILE * c];
    x2[s] = M2[kk[s] + TILE * c];
  }
}

// Gaunt-weighted density tiles of one run, four channels, in matrix-operand
// layout (lane (g,c) holds R(c, row(g,s)) in entry s).  Lanes outside the
// element pair read a clamped entry (rloc, cloc) and are masked by cmask; the
// dropped radial function 0 of an m != 0 shell (only at the first radial
// index) is masked per term.  No branches, two terms' loads in flight.
template <typename T, int U = 2>
__device__ inline void gather_terms(const T *P, int64_t ldp, const int4 *tpos, const T *term_f, int q0, int q1,
                                    int rloc, const int cloc[4], const T cmask[4], T r00[4], T r02[4], T r20[4],
                                    T r22[4]) {
#pragma unroll
  for (int s = 0; s < 4; ++s) r00[s] = r02[s] = r20[s] = r22[s] = T(0);
  int q = q0;
  for (; q + U <= q1; q += U) {
    T x[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      // (poff[i], poff[l] - skip[l], skip[i], skip[l]) of the term's density block
      const int4 tp = tpos[q + u];
      const long long row = tp.x + max(rloc - tp.z, 0), cb = tp.y;
      const bool rv = rloc >= tp.z;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const T v = P[row + (cb + max(cloc[s], tp.w)) * ldp];
        x[u][s] = (rv && cloc[s] >= tp.w) ? v : T(0);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const T f00 = term_f[4 * (q + u)], f02 = term_f[4 * (q + u) + 1], f20 = term_f[4 * (q + u) + 2],
              f22 = term_f[4 * (q + u) + 3];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        r00[s] += f00 * x[u][s];
        r02[s] += f02 * x[u][s];
        r20[s] += f20 * x[u][s];
        r22[s] += f22 * x[u][s];
      }
    }
  }
  for (; q < q1; ++q) {
    const int4 tp = tpos[q];
    const long long row = tp.x + max(rloc - tp.z, 0), cb = tp.y;
    const bool rv = rloc >= tp.z;
    const T f00 = term_f[4 * q], f02 = term_f[4 * q + 1], f20 = term_f[4 * q + 2], f22 = term_f[4 * q + 3];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const T v = P[row + (cb + max(cloc[s], tp.w)) * ldp];
      const T x = (rv && cloc[s] >= tp.w) ? v : T(0);
      r00[s] += f00 * x;
      r02[s] += f02 * x;
      r20[s] += f20 * x;
      r22[s] += f22 * x;
    }
  }
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    r00[s] *= cmask[s];
    r02[s] *= cmask[s];
    r20[s] *= cmask[s];
    r22[s] *= cmask[s];
  }
}

// Density tiles in gather order: tile (ef = e*Nel+f, src) holds the 16x16
// block of P between element e's radial functions of shell i and element f's
// of shell l (src = i*Nang+l from the source list), stored lane-major in the
// matrix-core operand order: entry lane*4 + s is R(c, row(g,s)) for lane
// (g,c), zero where a radial index is outside the element or is the dropped
// radial function 0 of an m != 0 shell.  One wave per tile.
template <typename T>
__global__ void __launch_bounds__(NT) k_dia_ptile(const double *P, int64_t ldp, int Nel, const int *first,
                                                  const int *ni, const int *skip, const long long *poff,
                                                  int Nang, const int *src_blk, int n_src, T *out) {
  const int64_t tile = (int64_t)blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
  const int64_t n_tiles = (int64_t)Nel * Nel * n_src;
  if (tile >= n_tiles) return;
  const int ef = (int)(tile / n_src), src = (int)(tile - (int64_t)ef * n_src);
  const int e = ef / Nel, f = ef - e * Nel;
  const int blk = src_blk[src], i = blk / Nang, l = blk - i * Nang;
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const int rr = first[e] + c, si = skip[i], sl = skip[l];
  const bool rv = c < ni[e] && rr >= si;
  typename Mma<T>::acc_t v;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int kk = Mma<T>::row(g, s), cc = first[f] + kk;
    const bool ok = rv && kk < ni[f] && cc >= sl;
    v[s] = ok ? (T)P[(poff[i] + rr - si) + (poff[l] + cc - sl) * ldp] : T(0);
  }
  *reinterpret_cast<typename Mma<T>::acc_t *>(out + tile * TILE2 + lane * 4) = v;
}

// gather_terms over the tiled density: term q reads tile tt[q] of the
// element pair's slice pt (one 4-vector per lane), masks already applied.
template <typename T, int U = 2>
__device__ inline void gather_tiles(const T *pt, const int *tt, const T *term_f, int q0, int q1, T r00[4],
                                    T r02[4], T r20[4], T r22[4]) {
  typedef typename Mma<T>::acc_t vec;
#pragma unroll
  for (int s = 0; s < 4; ++s) r00[s] = r02[s] = r20[s] = r22[s] = T(0);
  int q = q0;
  for (; q + U <= q1; q += U) {
    vec x[U];
#pragma unroll
    for (int u = 0; u < U; ++u) x[u] = *reinterpret_cast<const vec *>(pt + (size_t)tt[q + u] * TILE2);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const T f00 = term_f[4 * (q + u)], f02 = term_f[4 * (q + u) + 1], f20 = term_f[4 * (q + u) + 2],
              f22 = term_f[4 * (q + u) + 3];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        r00[s] += f00 * x[u][s];
        r02[s] += f02 * x[u][s];
        r20[s] += f20 * x[u][s];
        r22[s] += f22 * x[u][s];
      }
    }
  }
  for (; q < q1; ++q) {
    const vec x = *reinterpret_cast<const vec *>(pt + (size_t)tt[q] * TILE2);
    const T f00 = term_f[4 * q], f02 = term_f[4 * q + 1], f20 = term_f[4 * q + 2], f22 = term_f[4 * q + 3];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      r00[s] += f00 * x[s];
      r02[s] += f02 * x[s];
      r20[s] += f20 * x[s];
      r22[s] += f22 * x[s];
    }
  }
}

template <typename T>
struct DKArgs {
  const T *P;
  int64_t ldp;
  double *K;
  int64_t ldk;
  int Nel, Nang, n_act;
  const int *first, *ni, *skip;
  const long long *poff;
  const T *pq, *chol, *sigma;
  const int *chol_off, *chol_rank;
  const int *act_jk, *act_lstart, *lrun_ilm, *lrun_tstart;
  const int4 *tpos;
  const T *term_f;
  const int *efl;  // element pairs e*Nel+f of this launch
  int n_ef;
  const double *lsig;  // FP32 only: log scale of the (ilm, element) P/Q tiles
  const T *pt;         // tiled density (k_dia_ptile) or null: gather from P
  const int *ttile;    // per term: its tile in an element pair's slice
  int n_src;
  int sym;  // symmetric density: e > f pairs (e == f: j <= k) also written to the mirror
};

// One workgroup per (active output block (j,k), element pair (e,f)), the
// four waves splitting the block's coupled (L,|M|) runs.  Workgroups sharing
// (e,f) are mapped to one XCD so the (e,f) slice of the density tiles (or of
// P, when the tiles do not fit) stays in its L2.
// T = double is the reference's arithmetic; T = float gathers and multiplies
// in FP32 and sums the waves' tiles in FP64.
template <bool INSIDE, bool TILED, typename T>
__device__ inline void dia_k_body(const DKArgs<T> &A) {
  typedef Mma<T> MM;
  typedef typename MM::acc_t acc_t;
  const int w = blockIdx.x;
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
  __shared__ double red[4 * TILE2];
  acc_t acc = {0, 0, 0, 0}, acc2 = {0, 0, 0, 0};
  // Lanes outside the element pair read a clamped entry and are masked once
  // per tile; the dropped radial function 0 of an m != 0 shell (only at the
  // first radial index) is masked per term.  The term loop has no branches.
  const int rloc = fi + min(c, ni - 1);
  int cloc[4], kk[4];
  T cmask[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    kk[s] = MM::row(g, s);
    cloc[s] = fj + min(kk[s], nj - 1);
    cmask[s] = (c < ni && kk[s] < nj) ? T(1) : T(0);
  }
  const int r0 = A.act_lstart[act], r1 = A.act_lstart[act + 1];
  const T *pt = TILED ? A.pt + (size_t)ef * A.n_src * TILE2 + lane * 4 : nullptr;
  for (int run = r0 + wv; run < r1; run += 4) {
    const int ilm = A.lrun_ilm[run];
    T r00[4], r02[4], r20[4], r22[4];
    if constexpr (TILED)
      gather_tiles<T, INSIDE ? 2 : HFG_DIA_OFF_UNROLL>(pt, A.ttile, A.term_f, A.lrun_tstart[run],
                                                       A.lrun_tstart[run + 1], r00, r02, r20, r22);
    else
      gather_terms<T>(A.P, A.ldp, A.tpos, A.term_f, A.lrun_tstart[run], A.lrun_tstart[run + 1], rloc, cloc, cmask,
                      r00, r02, r20, r22);
    const T *te = A.pq + ((size_t)ilm * A.Nel + e) * 4 * TILE2;
    if (!INSIDE) {
      const T *tf = A.pq + ((size_t)ilm * A.Nel + f) * 4 * TILE2;
      const bool up = e > f;
      // row side: Q of e if e > f else P of e; column side the other kind
      const T *i0 = te + (up ? 2 : 0) * TILE2, *i2 = te + (up ? 3 : 1) * TILE2;
      const T *j0 = tf + (up ? 0 : 2) * TILE2, *j2 = tf + (up ? 1 : 3) * TILE2;
      T a0[4], a2[4], b0[4], b2[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int x = c + TILE * kk[s];
        a0[s] = i0[x];
        a2[s] = i2[x];
        b0[s] = j0[x];
        b2[s] = j2[x];
      }
      // two independent chains: ya = R00 j0' + R02 j2', yb = R20 j0' + R22 j2'
      acc_t ya = {0, 0, 0, 0}, yb = {0, 0, 0, 0};
      mma_pair<T>(r00, r20, b0, ya, yb);
      mma_pair<T>(r02, r22, b2, ya, yb);
      if constexpr (sizeof(T) == 4) {
        // FP32 tiles hold P / s and Q * s per (ilm, element): restore the
        // pair's Q(e) P(f) or P(e) Q(f) scale, s_f / s_e or s_e / s_f
        const double le = A.lsig[ilm * A.Nel + e], lf = A.lsig[ilm * A.Nel + f];
        const T rho = (T)exp(up ? lf - le : le - lf);
        ya *= rho;
        yb *= rho;
      }
      mma_pair2<T>(a0, ya, acc, a2, yb, acc2);  // acc += i0 ya, acc2 += i2 yb
    } else {
      const int rank = A.chol_rank[ilm * A.Nel + e], off = A.chol_off[ilm * A.Nel + e];
      const T *M = A.chol + (size_t)off * 2 * TILE2;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        r02[s] = -r02[s];
        r20[s] = -r20[s];
      }
      // M0', M2' as Y operands (y0, y2) and as X operands (x0, x2); the next
      // vector's tiles are loaded while the current one is on the matrix cores
      T y0[4], y2[4], x0[4], x2[4];
      load_chol<T>(M, 0, kk, c, y0, y2, x0, x2);
      T sg = rank > 0 ? A.sigma[off] : T(0);
      for (int p = 0; p < rank; ++p) {
        T ny0[4], ny2[4], nx0[4], nx2[4];
        T nsg = 0;
        if (p + 1 < rank) {
          load_chol<T>(M, p + 1, kk, c, ny0, ny2, nx0, nx2);
          nsg = A.sigma[off + p + 1];
        }
        // ya = R00 M0' - R02 M2', yb = -R20 M0' + R22 M2' (signs folded into
        // r02, r20); sigma_p = +-1 goes on the X operands, off the MFMA chain
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          x0[s] *= sg;
          x2[s] *= sg;
        }
        acc_t ya = {0, 0, 0, 0}, yb = {0, 0, 0, 0};
        mma_pair<T>(r00, r20, y0, ya, yb);
        mma_pair<T>(r02, r22, y2, ya, yb);
        mma_pair2<T>(x0, ya, acc, x2, yb, acc2);  // acc += sigma M0' ya, acc2 += sigma M2' yb
        if (p + 1 < rank) {
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            y0[s] = ny0[s];
            y2[s] = ny2[s];
            x0[s] = nx0[s];
            x2[s] = nx2[s];
          }
          sg = nsg;
        }
      }
    }
  }
  // combine the four waves' partial tiles (lane holds acc(row(g,r), c)) in FP64
  acc += acc2;
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) red[wv * TILE2 + kk[rr] + TILE * c] = (double)acc[rr];
  __syncthreads();
  const int a = t & 15, b = t >> 4;
  if (a >= ni || b >= nj) return;
  const long long row = pure_row(A.skip, A.poff, j, fi + a), col = pure_row(A.skip, A.poff, k, fj + b);
  if (row < 0 || col < 0) return;
  const double v = red[t] + red[TILE2 + t] + red[2 * TILE2 + t] + red[3 * TILE2 + t];
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

// Cross-element launch held to 5 waves per SIMD (96 registers), which the
// two-term gather fits; the in-element fallback keeps the compiler's choice.
template <typename T, bool TILED>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(HFG_DIA_OFF_WAVES))) k_dia_k_off(DKArgs<T> A) {
  dia_k_body<false, TILED, T>(A);
}
template <typename T, bool TILED>
__global__ void __launch_bounds__(NT) k_dia_k_in(DKArgs<T> A) {
  dia_k_body<true, TILED, T>(A);
}

// --- Cross-element exchange in two steps ----------------------------------------
//
// One workgroup per (column shell k, element pair (e,f)); each wave owns up to
// three rows j of the column and keeps their output tiles in registers over
// all coupled (L,|M|) groups of k.  Per group (and parity class of l_i): the
// slot tiles V_b(i) = sum_l v_b(k,l) P(i,l) of the pair are formed in LDS
// (waves split the slots), then each owned row gathers its four channel tiles
// R00 = sum c0 V0, R02 = -sum c0 V1, R20 = -sum c1 V0, R22 = sum c1 V1 from LDS
// and applies the disjoint P/Q factors as k_dia_k_off does.  The rows share
// each V tile, so the L2 gathers drop from |S_j| |S_k| per row and group to
// |S_k| per slot.

template <typename T>
struct DK2Args {
  const T *pt;  // tiled density
  int n_src;
  double *K;
  int64_t ldk;
  int Nel, Nang;
  const int *first, *ni, *skip;
  const long long *poff;
  const T *pq;
  const int *wg, *kshell, *kj, *jshell, *kg, *gilm, *gslot, *gjt, *sl, *lt, *jto, *jts;
  const T *ltv, *jtc;
  int sym;
};

template <typename T>
__global__ void __launch_bounds__(K2_THREADS) k_dia_k2_off(DK2Args<T> A) {
  typedef Mma<T> MM;
  typedef typename MM::acc_t acc_t;
  typedef acc_t vec;
  constexpr int NW = K2_THREADS / 64, JPW = K2_JMAX / NW;
  __shared__ T V[K2_SLOTS * 2 * TILE2];  // slot tiles V0, V1 in lane-major operand order
  const int w = blockIdx.x;
  const int kq = A.wg[2 * w], ef = A.wg[2 * w + 1];
  if (kq < 0) return;
  const int e = ef / A.Nel, f = ef - e * A.Nel;
  const int k = A.kshell[kq], r0 = A.kj[kq], nr = A.kj[kq + 1] - r0;
  const int t = threadIdx.x, lane = t & 63, g = lane >> 4, c = lane & 15;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  int kk[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) kk[s] = MM::row(g, s);
  acc_t acc[JPW];  // one accumulator per owned row (registers: three rows per wave)
#pragma unroll
  for (int r = 0; r < JPW; ++r) acc[r] = acc_t{0, 0, 0, 0};
  const T *pt = A.pt + (size_t)ef * A.n_src * TILE2 + lane * 4;
  const bool up = e > f;
  for (int gi = A.kg[kq]; gi < A.kg[kq + 1]; ++gi) {
    const int ilm = A.gilm[gi], s0 = A.gslot[gi], ns = A.gslot[gi + 1] - s0;
    // the group's factor operands are fetched first and used after the slot
    // phase: row side Q of e if e > f else P of e, column side the other kind
    T a0[4], a2[4], b0[4], b2[4];
    {
      const T *te = A.pq + ((size_t)ilm * A.Nel + e) * 4 * TILE2, *tf = A.pq + ((size_t)ilm * A.Nel + f) * 4 * TILE2;
      const T *i0 = te + (up ? 2 : 0) * TILE2, *i2 = te + (up ? 3 : 1) * TILE2;
      const T *j0 = tf + (up ? 0 : 2) * TILE2, *j2 = tf + (up ? 1 : 3) * TILE2;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int x = c + TILE * kk[s];
        a0[s] = i0[x];
        a2[s] = i2[x];
        b0[s] = j0[x];
        b2[s] = j2[x];
      }
    }
    __syncthreads();  // the previous group's LDS tiles are no longer read
    for (int sq = wv; sq < ns; sq += NW) {
      vec v0 = {0, 0, 0, 0}, v1 = {0, 0, 0, 0};
      const int x1 = A.sl[s0 + sq + 1];
      int x = A.sl[s0 + sq];
      for (; x + 4 <= x1; x += 4) {
        vec tv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) tv[u] = *reinterpret_cast<const vec *>(pt + (size_t)A.lt[x + u] * TILE2);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          v0 += A.ltv[2 * (x + u)] * tv[u];
          v1 += A.ltv[2 * (x + u) + 1] * tv[u];
        }
      }
      for (; x < x1; ++x) {
        const vec ta = *reinterpret_cast<const vec *>(pt + (size_t)A.lt[x] * TILE2);
        v0 += A.ltv[2 * x] * ta;
        v1 += A.ltv[2 * x + 1] * ta;
      }
      *reinterpret_cast<vec *>(V + (size_t)(2 * sq) * TILE2 + lane * 4) = v0;
      *reinterpret_cast<vec *>(V + (size_t)(2 * sq + 1) * TILE2 + lane * 4) = v1;
    }
    __syncthreads();
    const int *jo = A.jto + A.gjt[gi];
#pragma unroll
    for (int r = 0; r < JPW; ++r) {
      const int q = wv + NW * r;
      if (q >= nr) break;
      const int y0 = jo[q], y1 = jo[q + 1];
      if (y0 == y1) continue;
      T r00[4] = {0, 0, 0, 0}, r02[4] = {0, 0, 0, 0}, r20[4] = {0, 0, 0, 0}, r22[4] = {0, 0, 0, 0};
      for (int y = y0; y < y1; ++y) {
        const int slq = A.jts[y];
        const T c0 = A.jtc[2 * y], c1 = A.jtc[2 * y + 1];
        const vec V0 = *reinterpret_cast<const vec *>(V + (size_t)(2 * slq) * TILE2 + lane * 4);
        const vec V1 = *reinterpret_cast<const vec *>(V + (size_t)(2 * slq + 1) * TILE2 + lane * 4);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          r00[s] += c0 * V0[s];
          r02[s] -= c0 * V1[s];
          r20[s] -= c1 * V0[s];
          r22[s] += c1 * V1[s];
        }
      }
      acc_t ya = {0, 0, 0, 0}, yb = {0, 0, 0, 0};
      mma_pair<T>(r00, r20, b0, ya, yb);
      mma_pair<T>(r02, r22, b2, ya, yb);
#pragma unroll
      for (int s = 0; s < 4; ++s) acc[r] = MM::step(a0[s], ya[s], acc[r]);  // acc += i0 ya
#pragma unroll
      for (int s = 0; s < 4; ++s) acc[r] = MM::step(a2[s], yb[s], acc[r]);  // acc += i2 yb
    }
  }
  // the owned rows' tiles (lane holds entry (row(g,s), c)) and their mirrors
  const int ni = A.ni[e], nj = A.ni[f], fi = A.first[e], fj = A.first[f];
  const int b = c;
  const long long col = pure_row(A.skip, A.poff, k, fj + min(b, nj - 1));
  const bool scol = (b == 0 && f > 0) || (b == nj - 1 && f + 1 < A.Nel);
#pragma unroll
  for (int r = 0; r < JPW; ++r) {
    const int q = wv + NW * r;
    if (q >= nr) break;
    const acc_t v = acc[r];
    const int j = A.jshell[r0 + q];
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int a = kk[rr];
      if (a >= ni || b >= nj || col < 0) continue;
      const long long row = pure_row(A.skip, A.poff, j, fi + a);
      if (row < 0) continue;
      const bool srow = (a == 0 && e > 0) || (a == ni - 1 && e + 1 < A.Nel);
      double *dst = A.K + row + col * A.ldk;
      if (srow || scol)
        atomicAdd(dst, -(double)v[rr]);
      else
        *dst = -(double)v[rr];
      if (A.sym) {
        double *mdst = A.K + col + row * A.ldk;
        if (srow || scol)
          atomicAdd(mdst, -(double)v[rr]);
        else
          *mdst = -(double)v[rr];
      }
    }
  }
}

// --- In-element exchange with the Cholesky factors resident in LDS -------------
//
// One 1024-thread workgroup per ((L,|M|), element e, batch of runs of that
// (L,|M|)): the rank x {M0_p, M2_p} channel tiles of (ilm, e) are read once
// into LDS (rows padded to 17 so both the M and M' operand reads are free of
// bank conflicts), then each of the 16 waves takes runs of the batch, gathers
// their density tiles and applies sum_p sigma_p [M0'(R00 M0' - R02 M2') +
// M2'(-R20 M0' + R22 M2')] from LDS.  Runs of one output block sit in
// different workgroups, so the tiles are added to K atomically.  Compared
// with k_dia_k_in (one workgroup per (block, element), factors re-read
// from L2 for every run) this moves the factor traffic from ~100 GB to a few
// GB per N2 build.

constexpr int KL_THREADS = 1024, KL_WAVES = 16, KL_RMAX = KL_RANK_MAX, KL_LD = 17;

template <typename T>
struct DKLArgs {
  const T *P;
  int64_t ldp;
  double *K;
  int64_t ldk;
  int Nel, Nang;
  const int *first, *ni, *skip;
  const long long *poff;
  const T *chol, *sigma;
  const int *chol_off, *chol_rank;
  const int *act_jk, *lrun_tstart;
  const int4 *tpos;
  const T *term_f;
  const int *wg;       // per workgroup: ilm, e, first entry in runs, count (or ilm = -1)
  const int *runs;     // runs grouped by ilm
  const int *run_act;  // output block of each run
  const T *pt;         // tiled density (k_dia_ptile) or null: gather from P
  const int *ttile;
  int n_src;
  int sym;  // symmetric density: runs of blocks j < k also written to the mirror
};

template <typename T, bool TILED>
__global__ void __launch_bounds__(KL_THREADS) k_dia_kin_lds(DKLArgs<T> A) {
  typedef Mma<T> MM;
  typedef typename MM::acc_t acc_t;
  constexpr int TL = TILE * KL_LD;  // one padded 16x16 tile
  __shared__ T Mt[KL_RMAX * 2 * TL];
  __shared__ T sg[KL_RMAX];
  const int w = blockIdx.x;
  const int ilm = A.wg[4 * w], e = A.wg[4 * w + 1], b0 = A.wg[4 * w + 2], nb = A.wg[4 * w + 3];
  if (ilm < 0) return;
  const int rank = A.chol_rank[ilm * A.Nel + e], off = A.chol_off[ilm * A.Nel + e];
  const T *M = A.chol + (size_t)off * 2 * TILE2;
  for (int q = threadIdx.x; q < rank * 2 * TILE2; q += KL_THREADS) {
    const int pc = q / TILE2, u = q - pc * TILE2;  // pc = 2 p + channel
    Mt[pc * TL + (u & 15) + KL_LD * (u >> 4)] = M[q];
  }
  for (int q = threadIdx.x; q < rank; q += KL_THREADS) sg[q] = A.sigma[off + q];
  __syncthreads();
  const int t = threadIdx.x, lane = t & 63, g = lane >> 4, c = lane & 15;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int n = A.ni[e], fe = A.first[e];
  const int rloc = fe + min(c, n - 1);
  int cloc[4], kk[4];
  T cmask[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    kk[s] = MM::row(g, s);
    cloc[s] = fe + min(kk[s], n - 1);
    cmask[s] = (c < n && kk[s] < n) ? T(1) : T(0);
  }
  const T *pt = TILED ? A.pt + (size_t)(e * A.Nel + e) * A.n_src * TILE2 + lane * 4 : nullptr;
  for (int i = wv; i < nb; i += KL_WAVES) {
    const int run = A.runs[b0 + i], act = A.run_act[run];
    T r00[4], r02[4], r20[4], r22[4];
    if constexpr (TILED)
      gather_tiles<T>(pt, A.ttile, A.term_f, A.lrun_tstart[run], A.lrun_tstart[run + 1], r00, r02, r20, r22);
    else
      gather_terms<T>(A.P, A.ldp, A.tpos, A.term_f, A.lrun_tstart[run], A.lrun_tstart[run + 1], rloc, cloc, cmask,
                      r00, r02, r20, r22);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      r02[s] = -r02[s];
      r20[s] = -r20[s];
    }
    acc_t acc = {0, 0, 0, 0}, acc2 = {0, 0, 0, 0};
    // sigma_p = +-1 is applied to the X operands (exact).  The loop stays
    // rolled: the 16 waves of the workgroup keep the matrix pipe busy, and
    // two vectors per step measured no faster.
    for (int p = 0; p < rank; ++p) {
      const T *M0 = Mt + 2 * p * TL, *M2 = M0 + TL;
      const T sp = sg[p];
      T y0[4], y2[4], x0[4], x2[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        y0[s] = M0[c + KL_LD * kk[s]];
        y2[s] = M2[c + KL_LD * kk[s]];
        x0[s] = sp * M0[kk[s] + KL_LD * c];
        x2[s] = sp * M2[kk[s] + KL_LD * c];
      }
      acc_t ya = {0, 0, 0, 0}, yb = {0, 0, 0, 0};
      mma_pair<T>(r00, r20, y0, ya, yb);
      mma_pair<T>(r02, r22, y2, ya, yb);
      mma_pair2<T>(x0, ya, acc, x2, yb, acc2);
    }
    acc += acc2;
    const int jk = A.act_jk[act], j = jk / A.Nang, k = jk - j * A.Nang;
    const long long col = pure_row(A.skip, A.poff, k, fe + c);
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int a = kk[rr];
      if (a >= n || c >= n) continue;
      const long long row = pure_row(A.skip, A.poff, j, fe + a);
      if (row >= 0 && col >= 0) {
        atomicAdd(A.K + row + col * A.ldk, -(double)acc[rr]);
        if (A.sym && j != k) atomicAdd(A.K + col + row * A.ldk, -(double)acc[rr]);
      }
    }
  }
}

// --- In-element exchange as a dense operator on the matrix cores ---------------
//
// For one ((L,|M|), element e) every coupled output block's in-element tile
// is Y(a,b) = sum_ch sum_cd T_ch((a,b),(c,d)) R_ch(c,d), the four channel
// tiles R_ch the run's Gaunt-weighted density sums and T_ch = sign_ch W~_ch
// with W~ = B diag(sigma) B' (basis.cpp:1992-2033 contracts the same
// factor).  Applied as a GEMM Y (225 x runs) = T (225 x 912) X (912 x runs) it
// costs 2 n^4 per channel and run against 12 r n^3 / 4 of the Cholesky form
// (r ~ 30: 3.4x fewer matrix operations).
//
// k_dia_op_build (at upload) writes T in the matrix-core operand order: for
// k-step ks (inner indices kappa = 4 ks + g, kappa = ch * 228 + c + 15 d) and
// row tile rt (rows rho = 16 rt + c, rho = a + 15 b), lane (g, c) holds
// T(rho, kappa); one aligned 512-byte read per wave and (ks, rt).
// k_dia_kin_op: one 256-thread workgroup per ((L,|M|), e, up to 16 runs):
// the waves gather the runs' four channel tiles from the tiled density into
// LDS (X[kappa][run], rows padded to 17), then each wave multiplies its row
// tiles of T (streamed from L2: the workgroups of one (L,|M|), e) run on one
// XCD) by X on the FP64 matrix cores and adds the tiles into K atomically
// (boundary rows shared with other launches; other (L,|M|) of the block in
// other workgroups), with the mirror of a symmetric density as k_dia_kin_lds.

constexpr int OP_NC = 16, OP_KC = 228, OP_K = 4 * OP_KC, OP_KS = OP_K / 4, OP_RT = 15, OP_XLD = 17;
constexpr size_t OP_TSIZE = (size_t)OP_KS * OP_RT * 64;  // doubles of one (ilm, element) operator

struct DOpBuildArgs {
  const double *chol, *sigma;
  const int *chol_off, *chol_rank, *ni;
  int Nel;
  double *T;  // [ilm*Nel + e][OP_TSIZE]
};

__global__ void __launch_bounds__(NT) k_dia_op_build(DOpBuildArgs A) {
  const int x = blockIdx.x / OP_KS, ks = blockIdx.x - x * OP_KS;
  const int e = x % A.Nel, n = A.ni[e];
  const int rank = A.chol_rank[x], off = A.chol_off[x];
  const double *M = A.chol + (size_t)off * 2 * TILE2;
  double *out = A.T + (size_t)x * OP_TSIZE + (size_t)ks * OP_RT * 64;
  for (int q = threadIdx.x; q < OP_RT * 64; q += NT) {
    const int rt = q >> 6, lane = q & 63, g = lane >> 4, c = lane & 15;
    const int rho = 16 * rt + c, kap = 4 * ks + g;
    const int ch = kap / OP_KC, kc = kap - ch * OP_KC;
    const int cc = kc % 15, dd = kc / 15, a = rho % 15, b = rho / 15;
    double v = 0.0;
    if (rho < 225 && kc < 225 && a < n && b < n && cc < n && dd < n) {
      // B(xo + cc + a n, p) B(yo + b + dd n, p): channel block x of M_p at (cc, a), y at (b, dd)
      const int xo = ch >> 1, yo = ch & 1;
      for (int p = 0; p < rank; ++p) {
        const double *Mp = M + (size_t)p * 2 * TILE2;
        v += A.sigma[off + p] * Mp[xo * TILE2 + cc + TILE * a] * Mp[yo * TILE2 + b + TILE * dd];
      }
      if (ch == 1 || ch == 2) v = -v;
    }
    out[q] = v;
  }
}

struct DOpArgs {
  double *K;
  int64_t ldk;
  int Nel, Nang;
  const int *first, *ni, *skip;
  const long long *poff;
  const double *T;
  const int *act_jk, *lrun_tstart;
  const double *term_f;
  const int *wg;       // per workgroup: ilm, e, first entry in runs, count (or ilm = -1)
  const int *runs, *run_act;
  const double *pt;    // tiled density
  const int *ttile;
  int n_src;
  int sym;
};

__global__ void __launch_bounds__(NT) k_dia_kin_op(DOpArgs A) {
  __shared__ double X[OP_K * OP_XLD];
  const int w = blockIdx.x;
  const int ilm = A.wg[4 * w], e = A.wg[4 * w + 1], b0 = A.wg[4 * w + 2], nb = A.wg[4 * w + 3];
  if (ilm < 0) return;
  const int t = threadIdx.x, lane = t & 63, g = lane >> 4, c = lane & 15;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int n = A.ni[e], fe = A.first[e];
  for (int q = t; q < OP_K * OP_XLD; q += NT) X[q] = 0.0;
  __syncthreads();
  const double *pt = A.pt + (size_t)(e * A.Nel + e) * A.n_src * TILE2 + lane * 4;
  for (int i = wv; i < nb; i += NT / 64) {
    const int run = A.runs[b0 + i];
    double r[4][4];
    gather_tiles<double>(pt, A.ttile, A.term_f, A.lrun_tstart[run], A.lrun_tstart[run + 1], r[0], r[1], r[2],
                         r[3]);
    // lane (g, c) holds R_ch(c, kk) in entry s, kk = row(g, s): kappa = ch * 228 + c + 15 kk
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int kk = Mma<double>::row(g, s);
      if (c < n && kk < n)
#pragma unroll
        for (int ch = 0; ch < 4; ++ch) X[(ch * OP_KC + c + 15 * kk) * OP_XLD + i] = r[ch][s];
    }
  }
  __syncthreads();
  // row tiles wv, wv + 4, wv + 8, wv + 12 (< 15)
  const int nrt = wv < 3 ? 4 : 3;
  v4d acc[4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
  // One wave per SIMD (the X tile takes the LDS): the operator reads are
  // issued OP_PF k-steps ahead so their L2 latency hides behind the matrix work.
  constexpr int OP_PF = 4;
  static_assert(OP_KS % OP_PF == 0, "k-steps in whole prefetch rounds");
  const double *Tx = A.T + (size_t)(ilm * A.Nel + e) * OP_TSIZE + lane + (size_t)wv * 64;
  double abuf[OP_PF][4];
#pragma unroll
  for (int u = 0; u < OP_PF; ++u)
#pragma unroll
    for (int j = 0; j < 4; ++j) abuf[u][j] = j < nrt ? Tx[(size_t)u * OP_RT * 64 + 4 * j * 64] : 0.0;
  for (int ks0 = 0; ks0 < OP_KS; ks0 += OP_PF) {
#pragma unroll
    for (int u = 0; u < OP_PF; ++u) {
      const int ks = ks0 + u;
      const double bv = X[(4 * ks + g) * OP_XLD + c];
      double av[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) av[j] = abuf[u][j];
      if (ks + OP_PF < OP_KS) {
        const double *Tn = Tx + (size_t)(ks + OP_PF) * OP_RT * 64;
#pragma unroll
        for (int j = 0; j < 4; ++j) abuf[u][j] = j < nrt ? Tn[4 * j * 64] : 0.0;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (j < nrt) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[j], bv, acc[j], 0, 0, 0);
    }
  }
  // lane (g, c) holds Y(rho = 16 rt + g + 4 r, run c) in acc[j][r]
  if (c >= nb) return;
  const int run = A.runs[b0 + c], act = A.run_act[run];
  const int jk = A.act_jk[act], j = jk / A.Nang, k = jk - j * A.Nang;
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) {
    if (jj >= nrt) break;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rho = 16 * (wv + 4 * jj) + g + 4 * r;
      const int a = rho % 15, b = rho / 15;
      if (rho >= 225 || a >= n || b >= n) continue;
      const long long row = pure_row(A.skip, A.poff, j, fe + a), col = pure_row(A.skip, A.poff, k, fe + b);
      if (row < 0 || col < 0) continue;
      atomicAdd(A.K + row + col * A.ldk, -acc[jj][r]);
      if (A.sym && j != k) atomicAdd(A.K + col + row * A.ldk, -acc[jj][r]);
    }
  }
}

// --- In-element exchange as a dense operator ----------------------------------
//
// For one ((L,|M|), element e) the in-element exchange of every output block
// coupled to it is Y(a,b) = sum_ch sum_cd T_ch((a,b),(c,d)) R_ch(c,d), R_ch the
// Gaunt-weighted density tiles of that block: a GEMM Y (256 x runs) =
// T (256 x 4 KC) X (4 KC x runs), run on the f64 matrix cores.  X is formed in
// LDS one channel at a time from the screened terms.  Partial tiles go to
// Ypart[run][e]; k_dia_kin_reduce sums them per output block.

constexpr int NCOL = 32;  // runs (GEMM columns) per workgroup

struct DKInArgs {
  const double *P;
  int64_t ldp;
  int Nel, Nang;
  const int *first, *ni, *skip;
  const long long *poff;
  const double *T;
  const int *wg, *ilm_rstart, *ilm_runs, *lrun_tstart, *term_il;
  const double *term_f;
  double *Y;
};

__global__ void __launch_bounds__(NT) k_dia_kin_gemm(DKInArgs A) {
  const int w = blockIdx.x;
  const int ilm = A.wg[3 * w], e = A.wg[3 * w + 1], c0 = A.wg[3 * w + 2];
  if (ilm < 0) return;
  const int r0 = A.ilm_rstart[ilm] + c0, r1 = min(A.ilm_rstart[ilm + 1], r0 + NCOL);
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, g = lane >> 4, c = lane & 15;
  const int n = A.ni[e], fe = A.first[e];
  __shared__ double X[KC * NCOL];
  __shared__ int srun[NCOL];
  if (t < NCOL) srun[t] = r0 + t < r1 ? A.ilm_runs[r0 + t] : -1;
  v4d acc[4][2];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) acc[rt][0] = acc[rt][1] = v4d{0.0, 0.0, 0.0, 0.0};
  const double *Tb = A.T + (size_t)(ilm * A.Nel + e) * 4 * KC * TILE2;
  const int cc = t % 15, dd = t / 15;
  const bool cdok = t < 225 && cc < n && dd < n;
  for (int ch = 0; ch < 4; ++ch) {
    __syncthreads();
    if (t < KC)
      for (int j = 0; j < NCOL; ++j) {
        const int run = srun[j];
        double v = 0.0;
        if (run >= 0 && cdok)
          for (int q = A.lrun_tstart[run]; q < A.lrun_tstart[run + 1]; ++q) {
            const int il = A.term_il[q], i = il / A.Nang, l = il - i * A.Nang;
            const long long row = pure_row(A.skip, A.poff, i, fe + cc), col = pure_row(A.skip, A.poff, l, fe + dd);
            if (row >= 0 && col >= 0) v += A.term_f[4 * q + ch] * A.P[row + col * A.ldp];
          }
        X[t * NCOL + j] = v;
      }
    __syncthreads();
    const double *Tc = Tb + (size_t)ch * KC * TILE2 + 64 * wv + c;
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
  }
  // lane (g, c) holds Y(64 wv + 16 rt + g + 4 r, 16 ct + c) in component r
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

struct DKRedArgs {
  double *K;
  int64_t ldk;
  int Nel, Nang;
  const int *first, *ni, *skip;
  const long long *poff;
  const int *act_jk, *act_lstart;
  const double *Y;
};

// One workgroup per (active output block, element): K tile = -sum over the
// block's runs of the partial tiles, shared boundary rows/cols accumulated
// atomically (the cross-element launch adds into the same entries).
__global__ void __launch_bounds__(NT) k_dia_kin_reduce(DKRedArgs A) {
  const int act = blockIdx.x / A.Nel, e = blockIdx.x - act * A.Nel;
  const int t = threadIdx.x, a = t & 15, b = t >> 4;
  const int n = A.ni[e], fe = A.first[e];
  if (a >= n || b >= n) return;
  const int jk = A.act_jk[act], j = jk / A.Nang, k = jk - j * A.Nang;
  const long long row = pure_row(A.skip, A.poff, j, fe + a), col = pure_row(A.skip, A.poff, k, fe + b);
  if (row < 0 || col < 0) return;
  double v = 0.0;
  for (int run = A.act_lstart[act]; run < A.act_lstart[act + 1]; ++run)
    v += A.Y[((size_t)run * A.Nel + e) * TILE2 + t];
  const bool srow = (a == 0 && e > 0) || (a == n - 1 && e + 1 < A.Nel);
  const bool scol = (b == 0 && e > 0) || (b == n - 1 && e + 1 < A.Nel);
  double *dst = A.K + row + col * A.ldk;
  if (srow || scol)
    atomicAdd(dst, -v);
  else
    *dst = -v;
}

// K(-mj,-mk) = K(mj,mk) for a +-m symmetric density (basis.cpp:2067-2086):
// one workgroup per mirrored shell block; m != 0 on both sides, so both
// blocks hold (Nrad-1) x (Nrad-1) pure entries.
__global__ void __launch_bounds__(NT) k_dia_mirror(double *K, int64_t ldk, int Nrad, const long long *poff,
                                                    const int *mirror, int Nang) {
  const int dst = mirror[2 * blockIdx.x], src = mirror[2 * blockIdx.x + 1];
  const int jd = dst / Nang, kd = dst - jd * Nang, js = src / Nang, ks = src - js * Nang;
  const int n = Nrad - 1;
  for (int q = threadIdx.x; q < n * n; q += NT) {
    const int cc = q / n, rr = q - cc * n;
    K[(poff[jd] + rr) + (poff[kd] + cc) * ldk] = K[(poff[js] + rr) + (poff[ks] + cc) * ldk];
  }
}

}  // namespace

void diatomic_op_build(DiatomicDevice &d, hipStream_t s) {
  d.op_in = false;
  d.Top2.release();
  const size_t nx = (size_t)d.n_lm * d.Nel;
  if (nx == 0 || d.max_rank > 2 * TILE2) return;
  // opt-in: on N2 (lmax 16) the Cholesky-form kernel is faster (13.1 against
  // 16.2 ms): each workgroup streams its 1.75 MB operator for 16 runs
  const char *env = std::getenv("HFGPU_EXCHANGE_OP");
  if (!env || std::atoi(env) == 0) return;
  size_t fr = 0, tot = 0;
  hip_check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
  const size_t bytes = nx * OP_TSIZE * sizeof(double);
  if (bytes > fr / 4) return;  // keep three quarters of the free memory for densities and tiles
  d.Top2.alloc(nx * OP_TSIZE);
  DOpBuildArgs a{d.chol.p, d.sigma.p, d.chol_off.p, d.chol_rank.p, d.ni.p, d.Nel, d.Top2.p};
  hipLaunchKernelGGL(k_dia_op_build, dim3((unsigned)(nx * OP_KS)), dim3(NT), 0, s, a);
  hip_check(hipGetLastError(), "diatomic_op_build");
  d.op_in = true;
}

void diatomic_coulomb_launch(const DiatomicDevice &d, const double *P, int64_t ldp, double *J, int64_t ldj,
                             hipStream_t s, KernelTimer *timer) {
  fill_zero_launch(J, d.Nbf, d.Nbf, ldj, s);
  if (d.n_LM == 0) return;
  DJArgs a{P, ldp, d.Nel, d.Nang, d.first.p, d.ni.p, d.skip.p, d.poff.p, d.pq.p, d.chol.p, d.sigma.p,
           d.chol_off.p, d.chol_rank.p, d.LM_ilm.p, d.LM_fac.p, d.pstart.p, d.pair_kl.p, d.pair_c.p,
           d.Jin.p, d.ds.p, d.db.p};
  {
    TimedLaunch tl(timer, "diatomic_j_gather", s);
    hipLaunchKernelGGL(k_dia_j_gather, dim3(d.n_LM * d.Nel), dim3(NT), 0, s, a);
    hip_check(hipGetLastError(), "diatomic_j_gather");
  }
  hipLaunchKernelGGL(k_dia_j_prefix, dim3((d.n_LM + 63) / 64), dim3(64), 0, s, d.n_LM, d.Nel, d.ds.p, d.db.p,
                     d.preA.p, d.preB.p);
  hip_check(hipGetLastError(), "diatomic_j_prefix");
  DJOutArgs o{J, ldj, d.Nel, d.Nang, d.first.p, d.ni.p, d.skip.p, d.poff.p, d.pq.p, d.Jin.p, d.preA.p,
              d.preB.p, d.LM_ilm.p, d.out_start.p, d.out_q.p, d.out_c.p};
  {
    TimedLaunch tl(timer, "diatomic_j_expand", s);
    hipLaunchKernelGGL(k_dia_j_expand, dim3(d.Nang * d.Nang * d.Nel), dim3(NT), 0, s, o);
    hip_check(hipGetLastError(), "diatomic_j_expand");
  }
}

int block_stats_chunks(int Nrad) { return (Nrad + 31) / 32; }

void block_stats_batch_launch(const StatsArgs *d_args, const int *d_start, int n, int total_wg, hipStream_t s) {
  if (total_wg <= 0) return;
  hipLaunchKernelGGL(k_block_stats_batch, dim3(total_wg), dim3(NT), 0, s, d_args, d_start, n);
  hip_check(hipGetLastError(), "block_stats_batch");
}

void block_stats_launch(const double *P, int64_t ldp, int Nrad, int Nang, const int *skip, const long long *poff,
                        double *d_part, double *h_part, hipStream_t s) {
  const int nch = block_stats_chunks(Nrad);
  hipLaunchKernelGGL(k_block_stats, dim3(Nang * Nang, nch), dim3(NT), 0, s, P, ldp, Nrad, Nang, skip, poff, d_part);
  hip_check(hipGetLastError(), "block_stats");
  hip_check(hipMemcpyAsync(h_part, d_part, sizeof(double) * 4 * nch * Nang * Nang, hipMemcpyDeviceToHost, s),
            "copy block statistics");
}

bool block_stats_reduce(const double *h_part, int Nrad, int Nang, double *bnorm) {
  const int nch = block_stats_chunks(Nrad);
  double amax = 0.0, pmax = 0.0;
  for (int i = 0; i < Nang; ++i)
    for (int l = i; l < Nang; ++l) {  // only the pairs i <= l are written
      const double *o = h_part + (size_t)(i * Nang + l) * nch * 4;
      double sa = 0.0, sb = 0.0;
      for (int c = 0; c < nch; ++c) {
        sa += o[4 * c];
        sb += o[4 * c + 1];
        amax = std::max(amax, o[4 * c + 2]);
        pmax = std::max(pmax, o[4 * c + 3]);
      }
      bnorm[i * Nang + l] = std::sqrt(sa);
      if (i != l) bnorm[l * Nang + i] = std::sqrt(sb);
    }
  return amax <= 1e-14 * pmax;
}

bool symmetric_exchange_allowed() {
  const char *env = std::getenv("HFGPU_EXCHANGE_SYMMETRIC");
  return !(env && std::atoi(env) == 0);
}

namespace {
template <typename T>
void launch_dia_k(int part, bool tiled, int64_t grid, hipStream_t s, const DKArgs<T> &a) {
  const dim3 gd((unsigned)grid), bd(NT);
  if (part == 0 && tiled)
    hipLaunchKernelGGL((k_dia_k_off<T, true>), gd, bd, 0, s, a);
  else if (part == 0)
    hipLaunchKernelGGL((k_dia_k_off<T, false>), gd, bd, 0, s, a);
  else if (tiled)
    hipLaunchKernelGGL((k_dia_k_in<T, true>), gd, bd, 0, s, a);
  else
    hipLaunchKernelGGL((k_dia_k_in<T, false>), gd, bd, 0, s, a);
}

// FP32 copy of the P/Q tiles with P kinds divided and Q kinds multiplied by
// exp(lsig) of their (ilm, element).
__global__ void k_scale_pq_f32(const double *pq, const double *lsig, float *pq32, int n_x) {
  const int x = blockIdx.x;
  if (x >= n_x) return;
  const double up = exp(-lsig[x]), dn = exp(lsig[x]);
  for (int u = threadIdx.x; u < 4 * TILE2; u += NT) {
    const int kind = u / TILE2;
    pq32[(size_t)x * 4 * TILE2 + u] = (float)(pq[(size_t)x * 4 * TILE2 + u] * (kind < 2 ? up : dn));
  }
}
}  // namespace

void diatomic_exchange_kernels(DiatomicDevice &d, const double *P, int64_t ldp, double *K, int64_t ldk,
                               hipStream_t s, KernelTimer *timer, bool fp32) {
  fill_zero_launch(K, d.Nbf, d.Nbf, ldk, s);
  if (d.n_act == 0) return;
  if (fp32) {
    if (!d.f32_tables) {
      d.pq32.alloc(d.pq.n);
      d.chol32.alloc(d.chol.n);
      d.sigma32.alloc(d.sigma.n);
      const int n_x = (int)(d.pq.n / (4 * TILE2));
      hipLaunchKernelGGL(k_scale_pq_f32, dim3(n_x), dim3(NT), 0, s, d.pq.p, d.lsig.p, d.pq32.p, n_x);
      hip_check(hipGetLastError(), "scale_pq_f32");
      to_f32_launch(d.chol.p, d.chol32.p, d.chol.n, s);
      to_f32_launch(d.sigma.p, d.sigma32.p, d.sigma.n, s);
      d.f32_tables = true;
    }
    if (d.term_f32_gen != d.list_gen) {
      d.term_f32.alloc(d.term_f.n);
      to_f32_launch(d.term_f.p, d.term_f32.p, d.term_f.n, s);
      d.term_f32_gen = d.list_gen;
    }
    if (!d.tiled) {
      d.P32.alloc((size_t)d.Nbf * d.Nbf);
      mat_to_f32_launch(P, ldp, d.P32.p, d.Nbf, s);
    }
  }
  if (d.tiled) {
    const int64_t n_tiles = (int64_t)d.Nel * d.Nel * d.n_src, n_wg = (n_tiles + NT / 64 - 1) / (NT / 64);
    if (n_wg > 0x7fffffff) throw Error(-1, "density tile grid too large");
    TimedLaunch tl(timer, fp32 ? "diatomic_ptile_f32" : "diatomic_ptile", s, 0.0);
    if (fp32) {
      d.ptile32.alloc((size_t)n_tiles * TILE2);
      hipLaunchKernelGGL((k_dia_ptile<float>), dim3((unsigned)n_wg), dim3(NT), 0, s, P, ldp, d.Nel, d.first.p,
                         d.ni.p, d.skip.p, d.poff.p, d.Nang, d.src_blk.p, d.n_src, d.ptile32.p);
    } else {
      d.ptile.alloc((size_t)n_tiles * TILE2);
      hipLaunchKernelGGL((k_dia_ptile<double>), dim3((unsigned)n_wg), dim3(NT), 0, s, P, ldp, d.Nel, d.first.p,
                         d.ni.p, d.skip.p, d.poff.p, d.Nang, d.src_blk.p, d.n_src, d.ptile.p);
    }
    hip_check(hipGetLastError(), "diatomic_ptile");
  }
  // two launches: element pairs e != f (disjoint P/Q factors) and the
  // in-element tiles (Cholesky factor); boundary rows/cols shared between
  // neighbouring tiles are accumulated atomically in both
  for (int part = 0; part < 2; ++part) {
    const int n_ef = part == 0 ? d.Nel * (d.Nel - 1) / (d.sym ? 2 : 1) : d.Nel;
    if (n_ef == 0) continue;
    const char *name = part == 0 ? (fp32 ? "diatomic_k_off_f32" : "diatomic_k_off")
                                 : (fp32 ? "diatomic_k_in_f32" : "diatomic_k_in");
    const double flops = part == 0 ? d.flops_off : d.flops_in;
    if (part == 1 && d.dense_in && !fp32) {
      DKInArgs g{P, ldp, d.Nel, d.Nang, d.first.p, d.ni.p, d.skip.p, d.poff.p, d.Top.p, d.wg_in.p,
                 d.ilm_rstart.p, d.ilm_runs.p, d.lrun_tstart.p, d.term_il.p, d.term_f.p, d.Ypart.p};
      DKRedArgs r{K, ldk, d.Nel, d.Nang, d.first.p, d.ni.p, d.skip.p, d.poff.p, d.act_jk.p, d.act_lstart.p,
                  d.Ypart.p};
      TimedLaunch tl(timer, name, s, flops);
      hipLaunchKernelGGL(k_dia_kin_gemm, dim3(d.n_wg_in), dim3(NT), 0, s, g);
      hip_check(hipGetLastError(), "diatomic_kin_gemm");
      hipLaunchKernelGGL(k_dia_kin_reduce, dim3(d.n_act * d.Nel), dim3(NT), 0, s, r);
      hip_check(hipGetLastError(), "diatomic_kin_reduce");
      continue;
    }
    if (part == 1 && d.op_in && d.tiled && !fp32 && d.n_wg_op > 0) {
      // timed against the Cholesky-form flop count of the same results (the
      // reference's algorithm), so the rate compares with the other form
      TimedLaunch tl(timer, name, s, flops);
      DOpArgs a{K, ldk, d.Nel, d.Nang, d.first.p, d.ni.p, d.skip.p, d.poff.p, d.Top2.p, d.act_jk.p,
                d.lrun_tstart.p, d.term_f.p, d.wg_op.p, d.kl_runs.p, d.run_act.p, d.ptile.p, d.term_tile.p,
                d.n_src, (int)d.sym};
      hipLaunchKernelGGL(k_dia_kin_op, dim3(d.n_wg_op), dim3(NT), 0, s, a);
      hip_check(hipGetLastError(), "diatomic_kin_op");
      continue;
    }
    if (part == 1 && d.n_wg_kl > 0) {
      TimedLaunch tl(timer, name, s, flops);
      if (fp32) {
        DKLArgs<float> a{d.P32.p, d.Nbf, K, ldk, d.Nel, d.Nang, d.first.p, d.ni.p, d.skip.p, d.poff.p,
                         d.chol32.p, d.sigma32.p, d.chol_off.p, d.chol_rank.p, d.act_jk.p, d.lrun_tstart.p,
                         d.term_pos.p, d.term_f32.p, d.wg_kl.p, d.kl_runs.p, d.run_act.p,
                         d.tiled ? d.ptile32.p : nullptr, d.term_tile.p, d.n_src, (int)d.sym};
        if (d.tiled)
          hipLaunchKernelGGL((k_dia_kin_lds<float, true>), dim3(d.n_wg_kl), dim3(KL_THREADS), 0, s, a);
        else
          hipLaunchKernelGGL((k_dia_kin_lds<float, false>), dim3(d.n_wg_kl), dim3(KL_THREADS), 0, s, a);
      } else {
        DKLArgs<double> a{P, ldp, K, ldk, d.Nel, d.Nang, d.first.p, d.ni.p, d.skip.p, d.poff.p, d.chol.p,
                          d.sigma.p, d.chol_off.p, d.chol_rank.p, d.act_jk.p, d.lrun_tstart.p, d.term_pos.p,
                          d.term_f.p, d.wg_kl.p, d.kl_runs.p, d.run_act.p,
                          d.tiled ? d.ptile.p : nullptr, d.term_tile.p, d.n_src, (int)d.sym};
        if (d.tiled)
          hipLaunchKernelGGL((k_dia_kin_lds<double, true>), dim3(d.n_wg_kl), dim3(KL_THREADS), 0, s, a);
        else
          hipLaunchKernelGGL((k_dia_kin_lds<double, false>), dim3(d.n_wg_kl), dim3(KL_THREADS), 0, s, a);
      }
      hip_check(hipGetLastError(), "diatomic_kin_lds");
      continue;
    }
    if (part == 0 && d.k2 && !fp32) {
      TimedLaunch tl(timer, name, s, d.flops_off_k2);
      DK2Args<double> a{d.ptile.p, d.n_src, K, ldk, d.Nel, d.Nang, d.first.p, d.ni.p, d.skip.p, d.poff.p, d.pq.p,
                        d.k2_wg.p, d.k2_kshell.p, d.k2_kj.p, d.k2_jshell.p, d.k2_kg.p, d.k2_gilm.p, d.k2_gslot.p,
                        d.k2_gjt.p, d.k2_sl.p, d.k2_lt.p, d.k2_jto.p, d.k2_jts.p, d.k2_ltv.p, d.k2_jtc.p,
                        (int)d.sym};
      hipLaunchKernelGGL((k_dia_k2_off<double>), dim3(d.n_wg_k2), dim3(K2_THREADS), 0, s, a);
      hip_check(hipGetLastError(), "diatomic_k2_off");
      continue;
    }
    const int n_ef_pad = ((n_ef + 7) / 8) * 8;
    const int64_t grid = (int64_t)n_ef_pad * d.n_act;
    if (grid > 0x7fffffff) throw Error(-1, "exchange grid too large");
    const int *efl = part == 0 ? (d.sym ? d.ef_low.p : d.ef_off.p) : d.ef_in.p;
    TimedLaunch tl(timer, name, s, flops);
    if (fp32) {
      DKArgs<float> a{d.P32.p, d.Nbf, K, ldk, d.Nel, d.Nang, d.n_act, d.first.p, d.ni.p, d.skip.p, d.poff.p,
                      d.pq32.p, d.chol32.p, d.sigma32.p, d.chol_off.p, d.chol_rank.p, d.act_jk.p, d.act_lstart.p,
                      d.lrun_ilm.p, d.lrun_tstart.p, d.term_pos.p, d.term_f32.p, efl, n_ef, d.lsig.p,
                      d.tiled ? d.ptile32.p : nullptr, d.term_tile.p, d.n_src, (int)d.sym};
      launch_dia_k<float>(part, d.tiled, grid, s, a);
    } else {
      DKArgs<double> a{P, ldp, K, ldk, d.Nel, d.Nang, d.n_act, d.first.p, d.ni.p, d.skip.p, d.poff.p, d.pq.p,
                       d.chol.p, d.sigma.p, d.chol_off.p, d.chol_rank.p, d.act_jk.p, d.act_lstart.p, d.lrun_ilm.p,
                       d.lrun_tstart.p, d.term_pos.p, d.term_f.p, efl, n_ef, nullptr,
                       d.tiled ? d.ptile.p : nullptr, d.term_tile.p, d.n_src, (int)d.sym};
      launch_dia_k<double>(part, d.tiled, grid, s, a);
    }
    hip_check(hipGetLastError(), "diatomic_k");
  }
  if (d.n_mirror > 0) {
    hipLaunchKernelGGL(k_dia_mirror, dim3(d.n_mirror), dim3(NT), 0, s, K, ldk, d.Nrad, d.poff.p, d.mirror.p,
                       d.Nang);
    hip_check(hipGetLastError(), "diatomic_mirror");
  }
}

}  // namespace hfg
