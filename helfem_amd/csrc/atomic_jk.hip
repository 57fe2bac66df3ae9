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

namespace hfg {
namespace {

using namespace dev;

__global__ void __launch_bounds__(NT) k_fill_zero(double *A, int64_t rows, int64_t cols, int64_t ld) {
  const int64_t n = rows * cols;
  for (int64_t q = (int64_t)blockIdx.x * NT + threadIdx.x; q < n; q += (int64_t)gridDim.x * NT) {
    const int64_t j = q / rows, i = q - j * rows;
    A[i + j * ld] = 0.0;
  }
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

// One workgroup per ((L,M), element).
__global__ void __launch_bounds__(NT) k_atomic_j_gather(JArgs A) {
  const int lm = blockIdx.x / A.Nel, e = blockIdx.x - lm * A.Nel;
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

// Cross-element weights: pre_a(lm,e) = sum_{j>e} jbig(j), pre_b = sum_{j<e} jsmall(j)
__global__ void k_atomic_j_prefix(int n_lm, int Nel, const double *jsm, const double *jbg, double *preA,
                                  double *preB) {
  const int lm = blockIdx.x * blockDim.x + threadIdx.x;
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

// One workgroup per (angular output block (i,j), element).
__global__ void __launch_bounds__(NT) k_atomic_j_expand(JOutArgs A) {
  const int ij = blockIdx.x / A.Nel, e = blockIdx.x - ij * A.Nel;
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

// --- Exchange ----------------------------------------------------------------

__global__ void __launch_bounds__(NT) k_block_norms(const double *P, int64_t ldp, int Nrad, int Nang,
                                                     double *bnorm) {
  const int il = blockIdx.x, i = il / Nang, l = il - i * Nang;
  __shared__ double red[8];
  double s = 0.0, dummy = 0.0;
  const int64_t n = (int64_t)Nrad * Nrad;
  for (int64_t q = threadIdx.x; q < n; q += NT) {
    const int64_t c = q / Nrad, r = q - c * Nrad;
    const double v = P[(int64_t)(i * Nrad) + r + ((int64_t)l * Nrad + c) * ldp];
    s += v * v;
  }
  block_sum2(s, dummy, red);
  if (threadIdx.x == 0) bnorm[il] = sqrt(s);
}

struct KArgs {
  const double *P;
  int64_t ldp;
  double *K;
  int64_t ldk;
  int Nel, Nrad, Nang, N_L, n_act;
  const int *first, *ni;
  const double *rs, *rb, *chol;
  const int *chol_off, *chol_rank;
  const int *act_jk, *term_start, *term_il;
  const double *term_c;
  const int *efl;  // element pairs e*Nel+f of this launch
  int n_ef;
};

// One workgroup per (coupled output block (j,k), element pair (e,f)); the
// four waves split the multipoles L.  For each L a wave forms the
// Gaunt-weighted density tile R_L(e,f) = sum_terms c * P_il(e,f) directly in
// MFMA operand layout (lane (g,c) holds R(c, g+4s)), then contracts it with
// the radial factors on the matrix cores:
//   e != f:  K += A R B'       (A, B the r^L / r^-L-1 element matrices)
//   e == f:  K += sum_p M_p R M_p'   (M_p the Cholesky vectors as 16x16)
// Workgroups are mapped so that those sharing (e,f) run on one XCD.
template <bool INSIDE>
__global__ void __launch_bounds__(NT) k_atomic_k(KArgs A) {
  const int w = blockIdx.x;
  const int xcd = w & 7, loc = w >> 3;
  const int efi = (loc / A.n_act) * 8 + xcd;
  const int act = loc - (loc / A.n_act) * A.n_act;
  if (efi >= A.n_ef) return;
  const int ef = A.efl[efi], e = ef / A.Nel, f = ef - e * A.Nel;
  const int jk = A.act_jk[act], j = jk / A.Nang, k = jk - j * A.Nang;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, g = lane >> 4, c = lane & 15;
  const int ni = A.ni[e], nj = A.ni[f], fi = A.first[e], fj = A.first[f];
  __shared__ double red[4 * TILE2];
  v4d acc = {0.0, 0.0, 0.0, 0.0}, acc2 = {0.0, 0.0, 0.0, 0.0};
  const bool rowok = c < ni;
  bool colok[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) colok[s] = (g + 4 * s) < nj;
  for (int L = wv; L < A.N_L; L += 4) {
    const int q0 = A.term_start[act * A.N_L + L], q1 = A.term_start[act * A.N_L + L + 1];
    if (q0 == q1) continue;
    double r[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = q0; q < q1; ++q) {
      const int il = A.term_il[q];
      const double cq = A.term_c[q];
      const int i = il / A.Nang, l = il - i * A.Nang;
      const double *Pb = A.P + (int64_t)(i * A.Nrad + fi + c) + (int64_t)(l * A.Nrad + fj + g) * A.ldp;
      if (rowok) {
#pragma unroll
        for (int s = 0; s < 4; ++s)
          if (colok[s]) r[s] += cq * Pb[(int64_t)(4 * s) * A.ldp];
      }
    }
    const size_t te = ((size_t)L * A.Nel + e) * TILE2, tf = ((size_t)L * A.Nel + f) * TILE2;
    if (!INSIDE) {
      const double *At = (e > f) ? A.rb + te : A.rs + te;
      const double *Bt = (e > f) ? A.rs + tf : A.rb + tf;
      double as[4], bs[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        as[s] = At[c + TILE * (g + 4 * s)];
        bs[s] = Bt[c + TILE * (g + 4 * s)];
      }
      v4d y = {0.0, 0.0, 0.0, 0.0};
      y = mfma_16x16x16(r, bs, y);           // Y = R B'
      const double ys[4] = {y[0], y[1], y[2], y[3]};
      acc = mfma_16x16x16(as, ys, acc);      // K += A Y
    } else {
      const int rank = A.chol_rank[L * A.Nel + e];
      const double *M = A.chol + A.chol_off[L * A.Nel + e];
      // two Cholesky vectors per step, so two independent chains overlap;
      // M_p(c, k) serves as both the Y operand of R M_p' and the X operand of M_p Y
      for (int p = 0; p < rank; p += 2) {
        const double *Ma = M + (size_t)p * TILE2;
        const double *Mb = M + (size_t)(p + 1 < rank ? p + 1 : p) * TILE2;
        const double wb = p + 1 < rank ? 1.0 : 0.0;
        double ma[4], mb[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          ma[s] = Ma[c + TILE * (g + 4 * s)];
          mb[s] = wb * Mb[c + TILE * (g + 4 * s)];
        }
        v4d ya = {0.0, 0.0, 0.0, 0.0}, yb = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          ya = __builtin_amdgcn_mfma_f64_16x16x4f64(r[s], ma[s], ya, 0, 0, 0);  // R M_a'
          yb = __builtin_amdgcn_mfma_f64_16x16x4f64(r[s], mb[s], yb, 0, 0, 0);  // R M_b'
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ma[s], ya[s], acc, 0, 0, 0);   // M_a Y_a
          acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(mb[s], yb[s], acc2, 0, 0, 0); // M_b Y_b
        }
      }
    }
  }
  // combine the four waves' partial tiles (lane holds K(g+4r, c))
  acc += acc2;
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) red[wv * TILE2 + (g + 4 * rr) + TILE * c] = acc[rr];
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

}  // namespace

void fill_zero_launch(double *A, int64_t rows, int64_t cols, int64_t ld, hipStream_t s) {
  const int64_t n = rows * cols;
  if (n == 0) return;
  int64_t blocks = (n + NT - 1) / NT;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(NT), 0, s, A, rows, cols, ld);
  hip_check(hipGetLastError(), "fill_zero");
}

void atomic_coulomb_launch(const AtomicDevice &d, const double *P, int64_t ldp, double *J, int64_t ldj,
                           hipStream_t s, KernelTimer *timer) {
  const int64_t nbf = (int64_t)d.Nrad * d.Nang;
  fill_zero_launch(J, nbf, nbf, ldj, s);
  if (d.n_lm == 0) return;
  JArgs a{P, ldp, d.Nel, d.Nrad, d.Nang, d.first.p, d.ni.p, d.rs.p, d.rb.p, d.chol.p, d.chol_off.p,
          d.chol_rank.p, d.lm_L.p, d.lm_pstart.p, d.pair_kl.p, d.pair_c.p, d.Jin.p, d.jsm.p, d.jbg.p};
  {
    TimedLaunch tl(timer, "atomic_j_gather", s);
    hipLaunchKernelGGL(k_atomic_j_gather, dim3(d.n_lm * d.Nel), dim3(NT), 0, s, a);
    hip_check(hipGetLastError(), "atomic_j_gather");
  }
  hipLaunchKernelGGL(k_atomic_j_prefix, dim3((d.n_lm + 63) / 64), dim3(64), 0, s, d.n_lm, d.Nel, d.jsm.p,
                     d.jbg.p, d.preA.p, d.preB.p);
  hip_check(hipGetLastError(), "atomic_j_prefix");
  JOutArgs o{J, ldj, d.Nel, d.Nrad, d.Nang, d.first.p, d.ni.p, d.lm_L.p, d.rs.p, d.rb.p, d.Jin.p,
             d.preA.p, d.preB.p, d.out_start.p, d.out_lm.p, d.out_c.p};
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
  if (mask == d.screen_mask && !d.screen_mask.empty()) return;
  std::vector<int> act, tstart(1, 0), til;
  std::vector<double> tc;
  double frad = 0.0, frad_in = 0.0;
  for (size_t a = 0; a < d.h_act.size(); ++a) {
    const size_t t0 = tstart.size(), n0 = til.size();
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
        frad += d.h_radL[L];
        frad_in += d.h_radL_in[L];
      }
      starts.push_back((int)til.size());
    }
    if (til.size() == n0) continue;  // nothing reaches this output block
    act.push_back(d.h_act[a]);
    tstart.insert(tstart.end(), starts.begin(), starts.end());
    (void)t0;
  }
  d.n_act = (int)act.size();
  d.n_terms = (int64_t)til.size();
  d.flops_k_radial = frad;
  d.flops_k_angular = 2.0 * (double)til.size() * (double)d.Nrad * (double)d.Nrad;
  // split between the in-element and cross-element launches, the angular
  // sums in proportion to the tile areas they cover
  double area_in = 0.0, area_all = 0.0;
  for (int e = 0; e < d.Nel; ++e) {
    area_in += (double)d.h_ni[e] * d.h_ni[e];
    for (int f = 0; f < d.Nel; ++f) area_all += (double)d.h_ni[e] * d.h_ni[f];
  }
  d.flops_in = frad_in + d.flops_k_angular * area_in / area_all;
  d.flops_off = d.flops_k_radial + d.flops_k_angular - d.flops_in;
  if (act.empty()) act.push_back(0);
  if (til.empty()) {
    til.push_back(0);
    tc.push_back(0.0);
  }
  d.act_jk.upload(act, s);
  d.term_start.upload(tstart, s);
  d.term_il.upload(til, s);
  d.term_c.upload(tc, s);
  hip_check(hipStreamSynchronize(s), "exchange list upload");
  d.screen_mask = mask;
}

void atomic_exchange_launch(AtomicDevice &d, const double *P, int64_t ldp, double *K, int64_t ldk,
                            hipStream_t s, KernelTimer *timer) {
  const int64_t nbf = (int64_t)d.Nrad * d.Nang;
  if (d.h_act.empty()) {
    fill_zero_launch(K, nbf, nbf, ldk, s);
    return;
  }
  hipLaunchKernelGGL(k_block_norms, dim3(d.Nang * d.Nang), dim3(NT), 0, s, P, ldp, d.Nrad, d.Nang, d.bnorm.p);
  hip_check(hipGetLastError(), "block_norms");
  hip_check(hipMemcpyAsync(d.h_bnorm.p, d.bnorm.p, sizeof(double) * d.Nang * d.Nang, hipMemcpyDeviceToHost, s),
            "copy block norms");
  hip_check(hipStreamSynchronize(s), "block norms sync");
  compact_exchange_lists(d, s);
  fill_zero_launch(K, nbf, nbf, ldk, s);
  if (d.n_act == 0) return;
  // two launches: element pairs e != f and the in-element tiles; boundary
  // rows/cols shared by neighbouring tiles are accumulated atomically in both
  for (int part = 0; part < 2; ++part) {
    const int n_ef = part == 0 ? d.Nel * (d.Nel - 1) : d.Nel;
    if (n_ef == 0) continue;
    const int n_ef_pad = ((n_ef + 7) / 8) * 8;
    KArgs a{P, ldp, K, ldk, d.Nel, d.Nrad, d.Nang, d.N_L, d.n_act, d.first.p, d.ni.p, d.rs.p, d.rb.p,
            d.chol.p, d.chol_off.p, d.chol_rank.p, d.act_jk.p, d.term_start.p, d.term_il.p, d.term_c.p,
            part == 0 ? d.ef_off.p : d.ef_in.p, n_ef};
    const int64_t grid = (int64_t)n_ef_pad * d.n_act;
    if (grid > 0x7fffffff) throw Error(-1, "exchange grid too large");
    TimedLaunch tl(timer, part == 0 ? "atomic_k_off" : "atomic_k_in", s, part == 0 ? d.flops_off : d.flops_in);
    if (part == 0)
      hipLaunchKernelGGL(k_atomic_k<false>, dim3((unsigned)grid), dim3(NT), 0, s, a);
    else
      hipLaunchKernelGGL(k_atomic_k<true>, dim3((unsigned)grid), dim3(NT), 0, s, a);
    hip_check(hipGetLastError(), "atomic_k");
  }
}

}  // namespace hfg
