// LDA exchange-correlation on the atomic and pure-m diatomic DFT grids
// (FP64), gfx950.
//
// Atomic grid (src/atomic/dftgrid.cpp:470-576): per element, nrad Chebyshev
// radial points x lang cos(theta) Chebyshev nodes x mang even phi points.
// Pure-m diatomic grid (src/diatomic/dftgrid_purem.cpp:87-199): per element,
// the mu points x lang nu points, the phi integral analytic.  In both the
// basis values factorise, chi_(a,r) = R_r(i) Theta_a(t) e^(i m_a phi_k), so
// instead of the reference's dense (functions x points) products the kernels
// contract one index at a time:
//   k_xc_rad_in   D_i(p)   = sum_rs R_r(i) P_(a,r),(b,s) R_s(i),  p = (a,b)
//   k_xc_point    rho(i,t,k) = sum_d C_d(i,t) cos(d phi_k),
//                 C_d(i,t) = sum_{p: m_a - m_b = d} Theta_a(t) Theta_b(t) D_i(p);
//                 LDA exc and v pointwise; V_d(i,t) = sum_k w v cos(d phi_k);
//                 W_i(p) = sum_t Theta_a(t) Theta_b(t) V_{m_a-m_b}(i,t)
//   k_xc_rad_out  H_(a,r),(b,s) = sum_i R_r(i) R_s(i) W_i(p), element bands
// which is the reference's rho = sum_uv P_uv Re(chi_u chi_v*) and
// H += (a .* w v) a' + (b .* w v) b' (dftgrid.cpp:51-128, 304-359;
// dftgrid_purem.cpp:212-260, 468-500) with the sums reordered.
#include "hfgpu_internal.hpp"
#include "device_common.hpp"

namespace hfg {
namespace {

using namespace dev;

// lda_x (id 1) and lda_c_vwn (id 7, VWN5), spin-unpolarised, as the libxc
// formulas: Slater e_x = -(3/4)(3/pi)^(1/3) rho^(1/3), v_x = 4/3 e_x; VWN5
// paramagnetic interpolation in x = sqrt(rs), v_c = e_c - (x/6) de_c/dx.
// Densities below thr give zero (libxc's density threshold).
__device__ inline void lda_point(int xf, int cf, double rho, double thr, double &exc, double &v) {
  exc = 0.0;
  v = 0.0;
  if (!(rho >= thr)) return;
  const double pi = 3.14159265358979323846;
  if (xf == 1) {
    const double ex = -0.75 * cbrt(3.0 / pi) * cbrt(rho);
    exc += ex;
    v += (4.0 / 3.0) * ex;
  }
  if (cf == 7) {
    const double A = 0.0310907, b = 3.72744, c = 12.9352, x0 = -0.10498;
    const double rs = cbrt(3.0 / (4.0 * pi * rho));
    const double x = sqrt(rs);
    const double Q = sqrt(4.0 * c - b * b);
    const double X = x * x + b * x + c, Xp = 2.0 * x + b, X0 = x0 * x0 + b * x0 + c;
    const double f1 = 2.0 * b / Q, f2 = b * x0 / X0, f3 = 2.0 * (2.0 * x0 + b) / Q;
    const double at = atan(Q / (2.0 * x + b));
    const double F = A * (log(x * x / X) + (f1 - f2 * f3) * at - f2 * log((x - x0) * (x - x0) / X));
    const double dF = A * (2.0 / x - Xp / X + (f1 - f2 * f3) * (-Q / (2.0 * X)) - f2 * (2.0 / (x - x0) - Xp / X));
    exc += F;
    v += F - x / 6.0 * dF;
  }
}

__device__ inline long long grid_row(const int *skip, const long long *poff, int shell, int r) {
  const int s = skip[shell];
  return r >= s ? poff[shell] + (r - s) : -1;
}

struct XCArgs {
  const double *P;
  int64_t ldp;
  double *H;
  int64_t ldh;
  int Nel, Nang, nrad, lang, mang, nd, xf, cf, npair;
  double thr;
  const int *first, *ni, *skip;
  const long long *poff;
  const int *rstart, *rb, *rp, *dstart, *da, *db, *dp;
  const double *R, *wpt, *th, *cs;
  double *D, *W, *part;
};

// One workgroup per (element, shell a); the waves take the pairs (a, b) of
// row a in turn.
__global__ void __launch_bounds__(NT) k_xc_rad_in(XCArgs A) {
  const int e = blockIdx.x / A.Nang, a = blockIdx.x - e * A.Nang;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int ni = A.ni[e], f0 = A.first[e];
  extern __shared__ double sm[];
  double *sR = sm;                    // [nrad][TILE]
  double *sP = sm + A.nrad * TILE;    // [4][TILE2]
  for (int q = t; q < A.nrad * TILE; q += NT) sR[q] = A.R[(size_t)e * A.nrad * TILE + q];
  const int q0 = A.rstart[a], q1 = A.rstart[a + 1];
  for (int qb = q0; qb < q1; qb += 4) {
    const int q = qb + wv;
    __syncthreads();
    if (q < q1) {
      const int b = A.rb[q];
      for (int u = 0; u < 4; ++u) {
        const int x = lane + 64 * u, r = x & 15, s = x >> 4;
        double v = 0.0;
        if (r < ni && s < ni) {
          const long long row = grid_row(A.skip, A.poff, a, f0 + r), col = grid_row(A.skip, A.poff, b, f0 + s);
          if (row >= 0 && col >= 0) v = A.P[row + col * A.ldp];
        }
        sP[wv * TILE2 + x] = v;
      }
    }
    __syncthreads();
    if (q >= q1) continue;
    const int p = A.rp[q];
    const double *Pt = sP + wv * TILE2;
    for (int i = lane; i < A.nrad; i += 64) {
      const double *Ri = sR + i * TILE;
      double d = 0.0;
      for (int r = 0; r < ni; ++r) {
        double tr = 0.0;
        for (int s = 0; s < ni; ++s) tr += Pt[r + TILE * s] * Ri[s];
        d += Ri[r] * tr;
      }
      A.D[((size_t)e * A.npair + p) * A.nrad + i] = d;
    }
  }
}

// One workgroup per (element, radial point).
__global__ void __launch_bounds__(NT) k_xc_point(XCArgs A) {
  const int e = blockIdx.x / A.nrad, i = blockIdx.x - e * A.nrad;
  const int t = threadIdx.x;
  const int Na = A.Nang, nd = A.nd, nang = A.lang * A.mang;
  extern __shared__ double sm[];
  double *sD = sm;                      // [npair]
  double *sC = sD + A.npair;            // [lang][nd]
  double *sV = sC + A.lang * nd;        // [lang][mang]
  __shared__ double red[8];
  const size_t eb = (size_t)e * A.npair * A.nrad;
  for (int p = t; p < A.npair; p += NT) sD[p] = A.D[eb + (size_t)p * A.nrad + i];
  __syncthreads();
  for (int q = t; q < A.lang * nd; q += NT) {
    const int tt = q / nd, d = q - tt * nd;
    const double *th = A.th + (size_t)tt * Na;
    double c = 0.0;
    for (int x = A.dstart[d]; x < A.dstart[d + 1]; ++x) c += th[A.da[x]] * th[A.db[x]] * sD[A.dp[x]];
    sC[q] = c;
  }
  __syncthreads();
  const double *wp = A.wpt + ((size_t)e * A.nrad + i) * A.lang;
  double ex = 0.0, ne = 0.0;
  for (int q = t; q < nang; q += NT) {
    const int tt = q / A.mang, k = q - tt * A.mang;
    double rho = 0.0;
    for (int d = 0; d < nd; ++d) rho += sC[tt * nd + d] * A.cs[(size_t)d * A.mang + k];
    double exc, v;
    lda_point(A.xf, A.cf, rho, A.thr, exc, v);
    const double w = wp[tt];
    ex += w * exc * rho;
    ne += w * rho;
    sV[q] = w * v;
  }
  block_sum2(ex, ne, red);
  if (t == 0) {
    A.part[2 * ((size_t)e * A.nrad + i)] = ex;
    A.part[2 * ((size_t)e * A.nrad + i) + 1] = ne;
  }
  __syncthreads();
  for (int q = t; q < A.lang * nd; q += NT) {
    const int tt = q / nd, d = q - tt * nd;
    double s = 0.0;
    for (int k = 0; k < A.mang; ++k) s += sV[tt * A.mang + k] * A.cs[(size_t)d * A.mang + k];
    sC[q] = s;
  }
  __syncthreads();
  for (int d = 0; d < nd; ++d)
    for (int x = A.dstart[d] + t; x < A.dstart[d + 1]; x += NT) {
      const int a = A.da[x], b = A.db[x];
      double w = 0.0;
      for (int tt = 0; tt < A.lang; ++tt) w += A.th[(size_t)tt * Na + a] * A.th[(size_t)tt * Na + b] * sC[tt * nd + d];
      A.W[eb + (size_t)A.dp[x] * A.nrad + i] = w;
    }
}

// One workgroup per (element, shell a); thread (r,s) of the element band,
// looping over the pairs (a, b) of row a.  The corner shared with element
// e+1 is written by e with both contributions; e > 0 skips its own (0,0).
__global__ void __launch_bounds__(NT) k_xc_rad_out(XCArgs A) {
  const int e = blockIdx.x / A.Nang, a = blockIdx.x - e * A.Nang;
  const int t = threadIdx.x, r = t & 15, s = t >> 4;
  const int ni = A.ni[e], f0 = A.first[e];
  extern __shared__ double sm[];
  double *sR = sm;  // [nrad][TILE] of e, then column 0 of e+1
  double *sN = sm + A.nrad * TILE;
  for (int q = t; q < A.nrad * TILE; q += NT) sR[q] = A.R[(size_t)e * A.nrad * TILE + q];
  const bool has_next = e + 1 < A.Nel;
  if (has_next)
    for (int q = t; q < A.nrad; q += NT) sN[q] = A.R[((size_t)(e + 1) * A.nrad + q) * TILE];
  __syncthreads();
  if (r >= ni || s >= ni) return;
  if (e > 0 && r == 0 && s == 0) return;
  const long long row = grid_row(A.skip, A.poff, a, f0 + r);
  if (row < 0) return;
  const bool corner = has_next && r == ni - 1 && s == ni - 1;
  const size_t eb = (size_t)e * A.npair * A.nrad;
  const size_t nb = (size_t)(e + 1) * A.npair * A.nrad;
  for (int q = A.rstart[a]; q < A.rstart[a + 1]; ++q) {
    const int b = A.rb[q], p = A.rp[q];
    const long long col = grid_row(A.skip, A.poff, b, f0 + s);
    if (col < 0) continue;
    const double *Wp = A.W + eb + (size_t)p * A.nrad;
    double v = 0.0;
    for (int i = 0; i < A.nrad; ++i) v += sR[i * TILE + r] * sR[i * TILE + s] * Wp[i];
    if (corner) {
      const double *Wn = A.W + nb + (size_t)p * A.nrad;
      for (int i = 0; i < A.nrad; ++i) v += sN[i] * sN[i] * Wn[i];
    }
    A.H[row + col * A.ldh] = v;
  }
}

}  // namespace

void xc_grid_launch(XCGridDevice &g, int xf, int cf, double thr, const double *P, int64_t ldp, double *H,
                    int64_t ldh, hipStream_t s, KernelTimer *timer) {
  fill_zero_launch(H, g.Nbf, g.Nbf, ldh, s);
  const int nd = 2 * g.dmax + 1;
  XCArgs a{P, ldp, H, ldh, g.Nel, g.Nang, g.nrad, g.lang, g.mang, nd, xf, cf, g.npair, thr,
           g.first.p, g.ni.p, g.skip.p, g.poff.p, g.rstart.p, g.rb.p, g.rp.p, g.dstart.p, g.da.p, g.db.p, g.dp.p,
           g.R.p, g.wpt.p, g.th.p, g.cs.p, g.D.p, g.W.p, g.part.p};
  const size_t sh_in = sizeof(double) * ((size_t)g.nrad * TILE + 4 * TILE2);
  const size_t sh_pt = sizeof(double) * ((size_t)g.npair + (size_t)g.lang * nd + (size_t)g.lang * g.mang);
  const size_t sh_out = sizeof(double) * ((size_t)g.nrad * TILE + g.nrad);
  {
    TimedLaunch tl(timer, "xc_rad_in", s);
    hipLaunchKernelGGL(k_xc_rad_in, dim3(g.Nel * g.Nang), dim3(NT), sh_in, s, a);
    hip_check(hipGetLastError(), "xc_rad_in");
  }
  {
    TimedLaunch tl(timer, "xc_point", s);
    hipLaunchKernelGGL(k_xc_point, dim3(g.Nel * g.nrad), dim3(NT), sh_pt, s, a);
    hip_check(hipGetLastError(), "xc_point");
  }
  {
    TimedLaunch tl(timer, "xc_rad_out", s);
    hipLaunchKernelGGL(k_xc_rad_out, dim3(g.Nel * g.Nang), dim3(NT), sh_out, s, a);
    hip_check(hipGetLastError(), "xc_rad_out");
  }
}

}  // namespace hfg
