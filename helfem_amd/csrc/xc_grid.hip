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
//
// GGA (atomic: dftgrid.cpp:75-86, 166-180, 320-334, 382-405; pure-m
// diatomic: dftgrid_purem.cpp:248-260, 382-395, 468-500; 3D diatomic:
// src/diatomic/dftgrid.cpp:75-100, 292-305; dftgrid_common.h:188-216).  The
// gradient in the grid's orthogonal components, g = 2 Re sum P (d chi) chi* /
// (h_1, h_2, h_3), factorises the same way, with c_k = 1/h_k per point
// (atomic (1, 1/r, 1/(r sin theta)); diatomic 1/h_mu = 1/h_nu =
// 1/(Rh sqrt(sinh^2 mu + sin^2 nu)), 1/h_phi = 1/(Rh sinh mu sin nu)):
//   g_1 = 2 c_1 sum_d G_d cos(d phi),  G_d(t) = sum_{x in d} Theta_a Theta_b E(x),
//         E(x) = sum_rs R'_r P_(a,r),(b,s) R_s (k_xc_rad_in alongside D)
//   g_2 = c_2 sum_d T_d cos(d phi),  T_d(t) = sum_{x in d} (Theta_a Theta_b)' D(x)
//   g_3 = -c_3 sum_d d C_d sin(d phi)      (zero on the pure-m grid: d = 0)
// (the last two use D(a,b) = D(b,a), P symmetric, as the reference's factor
// 2 does).  With U = 2 w vsigma grad rho (restricted) the potential term
// H += X + X', X = sum U . Re(grad chi_u) chi_v*, becomes per pair x = (a,b)
//   H_(a,r),(b,s) += sum_i R_r R_s W2(x) + R'_r R_s Y(x) + R_r R'_s Y(x')
//   W2(x) = sum_t Theta_a Theta_b [V_d - d Vphi_d] + (Theta_a Theta_b)' Vtheta_d
//   Y(x)  = sum_t Theta_a Theta_b Vr_d,   x' = (b, a)
// with V_d, Vr_d, Vtheta_d = sum_k (w vrho, c_1 U_1, c_2 U_2) cos(d phi_k) and
// Vphi_d = sum_k c_3 U_3 sin(d phi_k).
//
// Meta-GGA (tau; dftgrid.cpp:89-107, 336-349; dftgrid_purem.cpp:287-318,
// 482-494): tau = 1/2 sum_c c_c^2 Re sum P d_c chi d_c chi*, per pair
//   radial  F(x) = sum_rs R'_r P R'_s  with Theta_a Theta_b,
//   theta   D(x) with Theta'_a Theta'_b,   phi  D(x) with m_a m_b Theta_a Theta_b,
// and H += sum_c (w vtau/2) c_c^2 Re d_c chi_u d_c chi_v*: with
// S_d = sum_k (w vtau / 2) cos(d phi_k), per pair
//   H += sum_i R'_r R'_s Z(x) + R_r R_s sum_t (c_2^2 Theta'_a Theta'_b + c_3^2 m_a m_b Theta_a Theta_b) S_d,
//   Z(x) = sum_t c_1^2 Theta_a Theta_b S_d.
#include "hfgpu_internal.hpp"
#include "device_common.hpp"
#include "xc_func.hpp"
#include <algorithm>

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

// Spin-polarized form: Slater exchange by spin scaling, VWN5 correlation
// e_c = e_P + alpha_c f(z)(1 - z^4)/f''(0) + (e_F - e_P) f(z) z^4 with
// v_cs = e_c - (rs/3) de/drs - (z - s) de/dz (libxc lda_c_vwn).
__device__ inline void vwn_aux(double A, double b, double c, double x0, double x, double &F, double &dF) {
  const double Q = sqrt(4.0 * c - b * b);
  const double X = x * x + b * x + c, Xp = 2.0 * x + b, X0 = x0 * x0 + b * x0 + c;
  const double f1 = 2.0 * b / Q, f2 = b * x0 / X0, f3 = 2.0 * (2.0 * x0 + b) / Q;
  const double at = atan(Q / (2.0 * x + b));
  F = A * (log(x * x / X) + (f1 - f2 * f3) * at - f2 * log((x - x0) * (x - x0) / X));
  dF = A * (2.0 / x - Xp / X + (f1 - f2 * f3) * (-Q / (2.0 * X)) - f2 * (2.0 / (x - x0) - Xp / X));
}

__device__ inline void lda_point_pol(int xf, int cf, double ra, double rb, double thr, double &exc, double &va,
                                     double &vb) {
  exc = 0.0;
  va = 0.0;
  vb = 0.0;
  const double rho = ra + rb;
  if (!(rho >= thr)) return;
  const double pi = 3.14159265358979323846;
  if (xf == 1) {
    const double c6 = cbrt(6.0 / pi);
    const double a3 = ra > 0.0 ? cbrt(ra) : 0.0, b3 = rb > 0.0 ? cbrt(rb) : 0.0;
    exc += -0.75 * c6 * (ra * a3 + rb * b3) / rho;
    va += -c6 * a3;
    vb += -c6 * b3;
  }
  if (cf == 7) {
    const double rs = cbrt(3.0 / (4.0 * pi * rho)), x = sqrt(rs);
    const double z = fmax(-1.0, fmin(1.0, (ra - rb) / rho));
    double eP, dP, eF, dF, eA, dA;
    vwn_aux(0.0310907, 3.72744, 12.9352, -0.10498, x, eP, dP);
    vwn_aux(0.01554535, 7.06042, 18.0578, -0.32500, x, eF, dF);
    vwn_aux(-1.0 / (6.0 * pi * pi), 1.13107, 13.0045, -0.0047584, x, eA, dA);
    const double c43 = cbrt(2.0) * 2.0 - 2.0;
    const double opz = 1.0 + z, omz = 1.0 - z, opz13 = cbrt(opz), omz13 = cbrt(omz);
    const double fz = (opz * opz13 + omz * omz13 - 2.0) / c43;
    const double dfz = 4.0 / 3.0 * (opz13 - omz13) / c43;
    const double fpp = 4.0 / (9.0 * (cbrt(2.0) - 1.0));
    const double z3 = z * z * z, z4 = z3 * z;
    const double e = eP + eA * fz * (1.0 - z4) / fpp + (eF - eP) * fz * z4;
    const double dedx = dP + dA * fz * (1.0 - z4) / fpp + (dF - dP) * fz * z4;
    const double dedz = eA * (dfz * (1.0 - z4) - 4.0 * z3 * fz) / fpp + (eF - eP) * (dfz * z4 + 4.0 * z3 * fz);
    const double base = e - x / 6.0 * dedx;
    exc += e;
    va += base - (z - 1.0) * dedz;
    vb += base - (z + 1.0) * dedz;
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
  const int *rstart, *rb, *rx, *dstart, *xd;
  const double *R, *wpt, *thx, *thxt, *cs;
  double *D, *W, *part;
  double *Db, *Wb;  // second spin (polarized evaluation), else null
  // gradient terms (k_xc_rad_in<true>, k_xc_point_gga, k_xc_rad_out<true>);
  // gsc [Nel][nrad][lang][3] = 1/(h_1, h_2, h_3) per point
  const double *dR, *gsc, *sn, *thd, *thdt;
  const int *rxT;
  double *E, *Y, *Eb, *Yb;
  // kinetic energy density terms (meta-GGA)
  const double *thdd, *thddt, *thmm, *thmmt;
  double *F, *Z, *Fb, *Zb;
};

// One workgroup per (element, shell a).  The pairs (a, b) of row a go in
// groups of J = 256 / nrad (1 to 8, xc_pairs_per_group): their P tiles are staged in LDS, and
// thread t takes (pair t / nrad, radial point t % nrad), so every thread of
// the workgroup has a point (J nrad of 256 busy).  A thread keeps R_i in
// registers; the R rows sit in LDS at an odd stride (17), so neighbouring
// points read different banks, and the tile reads of a wave touch at most two
// tiles (broadcast).  R and the tiles are zero outside the element.
constexpr int XC_RS = TILE + 1;

__host__ __device__ inline int xc_pairs_per_group(int nrad) { return min(8, max(1, NT / nrad)); }

// MODE 0: D only (LDA); 1: also E (GGA); 2: also F = sum R'_r P R'_s (meta-GGA)
template <int MODE>
__global__ void __launch_bounds__(NT) k_xc_rad_in(XCArgs A) {
  const int e = blockIdx.x / A.Nang, a = blockIdx.x - e * A.Nang;
  const int t = threadIdx.x;
  const int ni = A.ni[e], f0 = A.first[e];
  const int J = xc_pairs_per_group(A.nrad);
  extern __shared__ double sm[];
  double *sR = sm;                                        // [nrad][XC_RS]
  double *sdR = sm + A.nrad * XC_RS;                      // [nrad][XC_RS] with GGA
  double *sP = sm + (MODE ? 2 : 1) * A.nrad * XC_RS;      // [J][TILE2]
  for (int q = t; q < A.nrad * TILE; q += NT) {
    sR[(q >> 4) * XC_RS + (q & 15)] = A.R[(size_t)e * A.nrad * TILE + q];
    if (MODE) sdR[(q >> 4) * XC_RS + (q & 15)] = A.dR[(size_t)e * A.nrad * TILE + q];
  }
  const int q0 = A.rstart[a], q1 = A.rstart[a + 1];
  for (int qb = q0; qb < q1; qb += J) {
    __syncthreads();
    for (int x = t; x < J * TILE2; x += NT) {
      const int j = x >> 8, y = x & 255, r = y & 15, s = y >> 4, q = qb + j;
      double v = 0.0;
      if (q < q1 && r < ni && s < ni) {
        const long long row = grid_row(A.skip, A.poff, a, f0 + r), col = grid_row(A.skip, A.poff, A.rb[q], f0 + s);
        if (row >= 0 && col >= 0) v = A.P[row + col * A.ldp];
      }
      sP[x] = v;
    }
    __syncthreads();
    for (int item = t; item < J * A.nrad; item += NT) {
      const int j = item / A.nrad, i = item - j * A.nrad, q = qb + j;
      if (q >= q1) continue;
      const double *Pt = sP + j * TILE2;
      const double *Ri = sR + i * XC_RS, *dRi = sdR + i * XC_RS;
      double ri[TILE];
#pragma unroll
      for (int s = 0; s < TILE; ++s) ri[s] = Ri[s];
      double d = 0.0, eg = 0.0, fk = 0.0;
#pragma unroll
      for (int r = 0; r < TILE; ++r) {
        double tr = 0.0, trd = 0.0;
#pragma unroll
        for (int s = 0; s < TILE; ++s) {
          tr += Pt[r + TILE * s] * ri[s];
          if (MODE == 2) trd += Pt[r + TILE * s] * dRi[s];
        }
        d += ri[r] * tr;
        if (MODE) eg += dRi[r] * tr;
        if (MODE == 2) fk += dRi[r] * trd;
      }
      const size_t o = ((size_t)e * A.nrad + i) * A.npair + A.rx[q];
      A.D[o] = d;
      if (MODE) A.E[o] = eg;
      if (MODE == 2) A.F[o] = fk;
    }
  }
}

// One workgroup per (element, radial point).  With A.Db set the density of
// the second spin is contracted alongside and the polarized functional used.
// The pairs x are numbered by class d = m_a - m_b, so the class sums read
// contiguous runs of D and of the Theta_a Theta_b table: thread (d, t) forms
// C_d(t) from the [x][t] table (lanes of one d read consecutive t), and
// thread x forms W(x) from the [t][x] table (lanes read consecutive x).
// LDA point kernel workgroup: its 37 KB of LDS per radial point (Xe-sized
// bases) lets four 256-thread groups share a CU, which measured fastest
// (Xe: 256 / 512 / 1024 threads 0.82 / 1.10 / 0.98 ms).
#ifndef HFG_XC_LDA_POINT_NT
#define HFG_XC_LDA_POINT_NT 256
#endif
constexpr int LP_NT = HFG_XC_LDA_POINT_NT, LP_WAVES = LP_NT / 64;
static_assert(LP_NT % 64 == 0 && LP_NT <= 1024, "LDA point kernel workgroup");

__global__ void __launch_bounds__(LP_NT) k_xc_point(XCArgs A) {
  const int e = blockIdx.x / A.nrad, i = blockIdx.x - e * A.nrad;
  const int t = threadIdx.x;
  const int np = A.npair, nd = A.nd, lang = A.lang, mang = A.mang, nang = lang * mang;
  const bool pol = A.Db != nullptr;
  const int ns = pol ? 2 : 1;
  extern __shared__ double sm[];
  double *sD = sm;                    // [ns][np]
  double *sC = sD + ns * np;          // [ns][nd][lang]
  double *sV = sC + ns * nd * lang;   // [ns][lang][mang]
  __shared__ double red[2 * LP_WAVES];
  const size_t row = ((size_t)e * A.nrad + i) * np;
  for (int x = t; x < np; x += LP_NT) {
    sD[x] = A.D[row + x];
    if (pol) sD[np + x] = A.Db[row + x];
  }
  __syncthreads();
  for (int q = t; q < ns * nd * lang; q += LP_NT) {
    const int sp = q / (nd * lang), qq = q - sp * nd * lang;
    const int d = qq / lang, tt = qq - d * lang;
    const double *D = sD + sp * np;
    double c = 0.0;
    for (int x = A.dstart[d]; x < A.dstart[d + 1]; ++x) c += A.thxt[(size_t)x * lang + tt] * D[x];
    sC[q] = c;
  }
  __syncthreads();
  const double *wp = A.wpt + ((size_t)e * A.nrad + i) * lang;
  double ex = 0.0, ne = 0.0;
  for (int q = t; q < nang; q += LP_NT) {
    const int tt = q / mang, k = q - tt * mang;
    double ra = 0.0, rb = 0.0;
    for (int d = 0; d < nd; ++d) ra += sC[d * lang + tt] * A.cs[(size_t)d * mang + k];
    if (pol)
      for (int d = 0; d < nd; ++d) rb += sC[nd * lang + d * lang + tt] * A.cs[(size_t)d * mang + k];
    const double w = wp[tt];
    double exc, va, vb = 0.0;
    if (pol)
      lda_point_pol(A.xf, A.cf, ra, rb, A.thr, exc, va, vb);
    else
      lda_point(A.xf, A.cf, ra, A.thr, exc, va);
    ex += w * exc * (ra + rb);
    ne += w * (ra + rb);
    sV[q] = w * va;
    if (pol) sV[nang + q] = w * vb;
  }
  ex = wave_sum(ex);
  ne = wave_sum(ne);
  if ((t & 63) == 0) {
    red[t >> 6] = ex;
    red[LP_WAVES + (t >> 6)] = ne;
  }
  __syncthreads();
  if (t == 0) {
    double sx = 0.0, sn = 0.0;
    for (int w = 0; w < LP_WAVES; ++w) {
      sx += red[w];
      sn += red[LP_WAVES + w];
    }
    A.part[3 * ((size_t)e * A.nrad + i)] = sx;
    A.part[3 * ((size_t)e * A.nrad + i) + 1] = sn;
    A.part[3 * ((size_t)e * A.nrad + i) + 2] = 0.0;
  }
  __syncthreads();
  for (int q = t; q < ns * nd * lang; q += LP_NT) {
    const int sp = q / (nd * lang), qq = q - sp * nd * lang;
    const int d = qq / lang, tt = qq - d * lang;
    const double *V = sV + sp * nang + tt * mang;
    double s = 0.0;
    for (int k = 0; k < mang; ++k) s += V[k] * A.cs[(size_t)d * mang + k];
    sC[q] = s;
  }
  __syncthreads();
  for (int sp = 0; sp < ns; ++sp) {
    double *Wout = sp == 0 ? A.W : A.Wb;
    const double *C = sC + sp * nd * lang;
    for (int x = t; x < np; x += LP_NT) {
      const double *Cd = C + A.xd[x] * lang;
      double w = 0.0;
      for (int tt = 0; tt < lang; ++tt) w += A.thx[(size_t)tt * np + x] * Cd[tt];
      Wout[((size_t)e * np + x) * A.nrad + i] = w;  // [element][x][point]: rad_out reads runs of points
    }
  }
}

// General point kernel: one workgroup per (element, radial point), both spins
// when polarized.  LDS: region A holds each spin's radial contractions (D, E
// and with tau F), then its point values (w vrho, c_1 U_1, c_2 U_2, c_3 U_3
// and with tau w vtau/2); region B holds the class sums (C, G, T and with tau
// the three tau sums), then the class transforms (V_d - d Vphi_d, Vr_d,
// Vtheta_d and with tau S_d).  Any supported functional pair runs here
// (vsigma = 0 without a GGA part, vtau = 0 without a meta-GGA part).
template <int MODE> struct GgaShape {
  static constexpr int nin = MODE == 2 ? 3 : 2;  // radial contractions per spin
  static constexpr int ncs = MODE == 2 ? 6 : 3;  // class sums per spin
  static constexpr int npv = MODE == 2 ? 5 : 4;  // point arrays per spin
};

size_t xc_gga_point_lds(int np, int nd, int lang, int mang, int ns, bool tau) {
  const int nin = tau ? 3 : 2, ncs = tau ? 6 : 3, npv = tau ? 5 : 4;
  const size_t na = (size_t)std::max(nin * np, npv * lang * mang);
  return sizeof(double) * ((size_t)ns * na + (size_t)ns * ncs * nd * lang);
}

// The GGA / meta-GGA point kernel holds one radial point's density and
// potential tables in LDS (63 KB for PBE, 101 KB for TPSS on a 36 x 36
// angular rule with 49 shells), so one or two workgroups fit a CU: it runs
// XP_NT threads over the same tables to keep more waves resident (Xe:
// 1024 threads take the PBE point kernel from 4.1 to 2.6 ms and TPSS from
// 7.8 to 4.3 ms against 256).
#ifndef HFG_XC_POINT_NT
#define HFG_XC_POINT_NT 1024
#endif
constexpr int XP_NT = HFG_XC_POINT_NT, XP_WAVES = XP_NT / 64;
static_assert(XP_NT % 64 == 0 && XP_NT <= 1024, "point kernel workgroup");

template <int MODE>
__global__ void __launch_bounds__(XP_NT) k_xc_point_gga(XCArgs A) {
  typedef GgaShape<MODE> G;
  const int e = blockIdx.x / A.nrad, i = blockIdx.x - e * A.nrad;
  const int t = threadIdx.x;
  const int np = A.npair, nd = A.nd, lang = A.lang, mang = A.mang, nang = lang * mang, dmax = (nd - 1) / 2;
  const bool pol = A.Db != nullptr;
  const int ns = pol ? 2 : 1;
  const bool gga = xcf::is_gga(A.xf) || xcf::is_gga(A.cf);
  const bool tau = MODE == 2;
  const int na = max(G::nin * np, G::npv * nang), ncl = nd * lang;
  extern __shared__ double sm[];
  double *sA = sm;            // [ns][na]
  double *sB = sm + ns * na;  // [ns][ncs][nd][lang]
  __shared__ double red[3 * XP_WAVES];
  const size_t row = ((size_t)e * A.nrad + i) * np;
  for (int x = t; x < np; x += XP_NT) {
    sA[x] = A.D[row + x];
    sA[np + x] = A.E[row + x];
    if (tau) sA[2 * np + x] = A.F[row + x];
    if (pol) {
      sA[na + x] = A.Db[row + x];
      sA[na + np + x] = A.Eb[row + x];
      if (tau) sA[na + 2 * np + x] = A.Fb[row + x];
    }
  }
  __syncthreads();
  // class sums: 0 C (Theta Theta D), 1 G (Theta Theta E), 2 T ((Theta Theta)' D),
  // with tau 3 (Theta Theta F), 4 (Theta' Theta' D), 5 (m m Theta Theta D)
  for (int q = t; q < ns * G::ncs * ncl; q += XP_NT) {
    const int sp = q / (G::ncs * ncl), r1 = q - sp * G::ncs * ncl, c = r1 / ncl, r2 = r1 - c * ncl;
    const int d = r2 / lang, tt = r2 - d * lang;
    const double *src = sA + sp * na + (c == 1 ? np : c == 3 ? 2 * np : 0);
    const double *tab = c == 2 ? A.thdt : c == 4 ? A.thddt : c == 5 ? A.thmmt : A.thxt;
    double acc = 0.0;
    for (int x = A.dstart[d]; x < A.dstart[d + 1]; ++x) acc += tab[(size_t)x * lang + tt] * src[x];
    sB[q] = acc;
  }
  __syncthreads();
  const double *wp = A.wpt + ((size_t)e * A.nrad + i) * lang;
  const double *gs = A.gsc + ((size_t)e * A.nrad + i) * lang * 3;
  double ex = 0.0, ne = 0.0, ek = 0.0;
  for (int q = t; q < nang; q += XP_NT) {
    const int tt = q / mang, k = q - tt * mang;
    const double c1 = gs[3 * tt], c2 = gs[3 * tt + 1], c3 = gs[3 * tt + 2];
    double rho[2] = {0.0, 0.0}, g[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}}, tk[2] = {0.0, 0.0};
    for (int sp = 0; sp < ns; ++sp) {
      const double *C = sB + sp * G::ncs * ncl;
      double r0 = 0.0, gr = 0.0, gt = 0.0, gp = 0.0, kf = 0.0, kt = 0.0, km = 0.0;
      for (int d = 0; d < nd; ++d) {
        const double cc = A.cs[(size_t)d * mang + k], cd = C[d * lang + tt];
        r0 += cd * cc;
        gr += C[ncl + d * lang + tt] * cc;
        gt += C[2 * ncl + d * lang + tt] * cc;
        gp += (double)(d - dmax) * cd * A.sn[(size_t)d * mang + k];
        if (tau) {
          kf += C[3 * ncl + d * lang + tt] * cc;
          kt += C[4 * ncl + d * lang + tt] * cc;
          km += C[5 * ncl + d * lang + tt] * cc;
        }
      }
      rho[sp] = r0;
      g[sp][0] = 2.0 * c1 * gr;
      g[sp][1] = c2 * gt;
      g[sp][2] = -c3 * gp;
      if (tau) tk[sp] = 0.5 * (c1 * c1 * kf + c2 * c2 * kt + c3 * c3 * km);
    }
    const double w = wp[tt];
    double f, va, vb = 0.0, vta = 0.0, vtb = 0.0, u[2][3];
    if (pol) {
      const double x7[7] = {rho[0], rho[1], g[0][0] * g[0][0] + g[0][1] * g[0][1] + g[0][2] * g[0][2],
                            g[0][0] * g[1][0] + g[0][1] * g[1][1] + g[0][2] * g[1][2],
                            g[1][0] * g[1][0] + g[1][1] * g[1][1] + g[1][2] * g[1][2], tk[0], tk[1]};
      double v[7];
      xcf::eval_pol(A.xf, A.cf, x7, A.thr, gga, tau, f, v);
      va = v[0];
      vb = v[1];
      vta = v[5];
      vtb = v[6];
      for (int c = 0; c < 3; ++c) {
        u[0][c] = w * (2.0 * v[2] * g[0][c] + v[3] * g[1][c]);
        u[1][c] = w * (2.0 * v[4] * g[1][c] + v[3] * g[0][c]);
      }
    } else {
      const double x3[3] = {rho[0], g[0][0] * g[0][0] + g[0][1] * g[0][1] + g[0][2] * g[0][2], tk[0]};
      double v[3];
      xcf::eval_unpol(A.xf, A.cf, x3, A.thr, gga, tau, f, v);
      va = v[0];
      vta = v[2];
      for (int c = 0; c < 3; ++c) u[0][c] = 2.0 * w * v[1] * g[0][c];
    }
    ex += w * f;
    ne += w * (rho[0] + rho[1]);
    ek += w * (tk[0] + tk[1]);
    sA[q] = w * va;
    sA[nang + q] = c1 * u[0][0];
    sA[2 * nang + q] = c2 * u[0][1];
    sA[3 * nang + q] = c3 * u[0][2];
    if (tau) sA[4 * nang + q] = 0.5 * w * vta;
    if (pol) {
      sA[na + q] = w * vb;
      sA[na + nang + q] = c1 * u[1][0];
      sA[na + 2 * nang + q] = c2 * u[1][1];
      sA[na + 3 * nang + q] = c3 * u[1][2];
      if (tau) sA[na + 4 * nang + q] = 0.5 * w * vtb;
    }
  }
  ex = wave_sum(ex);
  ne = wave_sum(ne);
  ek = wave_sum(ek);
  if ((t & 63) == 0) {
    red[t >> 6] = ex;
    red[XP_WAVES + (t >> 6)] = ne;
    red[2 * XP_WAVES + (t >> 6)] = ek;
  }
  __syncthreads();
  if (t == 0) {
    double sx = 0.0, sn = 0.0, sk = 0.0;
    for (int w = 0; w < XP_WAVES; ++w) {
      sx += red[w];
      sn += red[XP_WAVES + w];
      sk += red[2 * XP_WAVES + w];
    }
    A.part[3 * ((size_t)e * A.nrad + i)] = sx;
    A.part[3 * ((size_t)e * A.nrad + i) + 1] = sn;
    A.part[3 * ((size_t)e * A.nrad + i) + 2] = sk;
  }
  __syncthreads();
  // class transforms: 0 Q = V_d - d Vphi_d, 1 Vr_d, 2 Vtheta_d, with tau 3 S_d
  const int ntr = tau ? 4 : 3;
  for (int q = t; q < ns * ntr * ncl; q += XP_NT) {
    const int sp = q / (ntr * ncl), r1 = q - sp * ntr * ncl, c = r1 / ncl, r2 = r1 - c * ncl;
    const int d = r2 / lang, tt = r2 - d * lang;
    const double *V = sA + sp * na + tt * mang;
    const double *cs = A.cs + (size_t)d * mang, *snd = A.sn + (size_t)d * mang;
    double acc = 0.0;
    if (c == 0) {
      const double fphi = (double)(d - dmax);
      const double *Vp = V + 3 * nang;
      for (int k = 0; k < mang; ++k) acc += V[k] * cs[k] - fphi * Vp[k] * snd[k];
    } else {
      const double *Vc = V + (c == 1 ? 1 : c == 2 ? 2 : 4) * nang;
      for (int k = 0; k < mang; ++k) acc += Vc[k] * cs[k];
    }
    sB[sp * G::ncs * ncl + c * ncl + d * lang + tt] = acc;
  }
  __syncthreads();
  for (int sp = 0; sp < ns; ++sp) {
    double *Wout = sp == 0 ? A.W : A.Wb, *Yout = sp == 0 ? A.Y : A.Yb, *Zout = sp == 0 ? A.Z : A.Zb;
    const double *Q = sB + sp * G::ncs * ncl, *Vr = Q + ncl, *Vt = Vr + ncl, *S = Vt + ncl;
    for (int x = t; x < np; x += XP_NT) {
      const int d = A.xd[x];
      const double *Qd = Q + d * lang, *Vrd = Vr + d * lang, *Vtd = Vt + d * lang, *Sd = S + d * lang;
      double w2 = 0.0, y = 0.0, z = 0.0;
      for (int tt = 0; tt < lang; ++tt) {
        const double th = A.thx[(size_t)tt * np + x];
        w2 += th * Qd[tt] + A.thd[(size_t)tt * np + x] * Vtd[tt];
        y += th * Vrd[tt];
        if (tau) {
          const double c1 = gs[3 * tt], c2 = gs[3 * tt + 1], c3 = gs[3 * tt + 2];
          w2 += (c2 * c2 * A.thdd[(size_t)tt * np + x] + c3 * c3 * A.thmm[(size_t)tt * np + x]) * Sd[tt];
          z += c1 * c1 * th * Sd[tt];
        }
      }
      Wout[((size_t)e * np + x) * A.nrad + i] = w2;
      Yout[((size_t)e * np + x) * A.nrad + i] = y;
      if (tau) Zout[((size_t)e * np + x) * A.nrad + i] = z;
    }
  }
}

// Workgroups per row of the output kernel (below): the pairs of one row are
// spread over this many, so the launch holds more waves for the latency of
// its dependent radial loads (Xe, split 1 / 2 / 4 / 8: LDA 0.437 / 0.405 /
// 0.392 / 0.401 ms, PBE 1.18 / 1.12 / 1.10 / 1.14 ms).
#ifndef HFG_XC_OUT_SPLIT
#define HFG_XC_OUT_SPLIT 4
#endif
constexpr int XO_SPLIT = HFG_XC_OUT_SPLIT;

// One workgroup per (element, shell a); thread (r,s) of the element band,
// looping over the pairs (a, b) of row a.  The corner shared with element
// e+1 is written by e with both contributions; e > 0 skips its own (0,0).
template <int MODE>
__global__ void __launch_bounds__(NT) k_xc_rad_out(XCArgs A) {
  // XO_SPLIT workgroups share row a, taking its pairs (a, b) in turn; each
  // pair's H block is written by one of them
  const int blk = blockIdx.x / XO_SPLIT, part = blockIdx.x - blk * XO_SPLIT;
  const int e = blk / A.Nang, a = blk - e * A.Nang;
  const int t = threadIdx.x, r = t & 15, s = t >> 4;
  const int ni = A.ni[e], f0 = A.first[e];
  extern __shared__ double sm[];
  double *sR = sm;  // [nrad][TILE] of e, then column 0 of e+1; with GGA the same of d/dr R
  double *sN = sm + A.nrad * TILE;
  double *sdR = sN + A.nrad, *sdN = sdR + A.nrad * TILE;
  for (int q = t; q < A.nrad * TILE; q += NT) {
    sR[q] = A.R[(size_t)e * A.nrad * TILE + q];
    if (MODE) sdR[q] = A.dR[(size_t)e * A.nrad * TILE + q];
  }
  const bool has_next = e + 1 < A.Nel;
  if (has_next)
    for (int q = t; q < A.nrad; q += NT) {
      sN[q] = A.R[((size_t)(e + 1) * A.nrad + q) * TILE];
      if (MODE) sdN[q] = A.dR[((size_t)(e + 1) * A.nrad + q) * TILE];
    }
  __syncthreads();
  if (r >= ni || s >= ni) return;
  if (e > 0 && r == 0 && s == 0) return;
  const long long row = grid_row(A.skip, A.poff, a, f0 + r);
  if (row < 0) return;
  const bool corner = has_next && r == ni - 1 && s == ni - 1;
  const size_t np = A.npair;
  for (int q = A.rstart[a] + part; q < A.rstart[a + 1]; q += XO_SPLIT) {
    const int b = A.rb[q], x = A.rx[q];
    const long long col = grid_row(A.skip, A.poff, b, f0 + s);
    if (col < 0) continue;
    const double *Wp = A.W + ((size_t)e * np + x) * A.nrad;
    double v = 0.0;
    if (MODE) {
      const int xT = A.rxT[q];
      const double *Yp = A.Y + ((size_t)e * np + x) * A.nrad, *YT = A.Y + ((size_t)e * np + xT) * A.nrad;
      const double *Zp = MODE == 2 ? A.Z + ((size_t)e * np + x) * A.nrad : nullptr;
      for (int i = 0; i < A.nrad; ++i) {
        const double Rr = sR[i * TILE + r], Rs = sR[i * TILE + s], Dr = sdR[i * TILE + r], Ds = sdR[i * TILE + s];
        v += Rr * Rs * Wp[i] + Dr * Rs * Yp[i] + Rr * Ds * YT[i];
        if (MODE == 2) v += Dr * Ds * Zp[i];
      }
      if (corner) {
        const double *Wn = A.W + ((size_t)(e + 1) * np + x) * A.nrad;
        const double *Yn = A.Y + ((size_t)(e + 1) * np + x) * A.nrad, *YTn = A.Y + ((size_t)(e + 1) * np + xT) * A.nrad;
        for (int i = 0; i < A.nrad; ++i) v += sN[i] * (sN[i] * Wn[i] + sdN[i] * (Yn[i] + YTn[i]));
        if (MODE == 2) {
          const double *Zn = A.Z + ((size_t)(e + 1) * np + x) * A.nrad;
          for (int i = 0; i < A.nrad; ++i) v += sdN[i] * sdN[i] * Zn[i];
        }
      }
    } else {
      for (int i = 0; i < A.nrad; ++i) v += sR[i * TILE + r] * sR[i * TILE + s] * Wp[i];
      if (corner) {
        const double *Wn = A.W + ((size_t)(e + 1) * np + x) * A.nrad;
        for (int i = 0; i < A.nrad; ++i) v += sN[i] * sN[i] * Wn[i];
      }
    }
    A.H[row + col * A.ldh] = v;
  }
}

}  // namespace

bool xc_generic(int xf, int cf) { return !((xf == 0 || xf == 1) && (cf == 0 || cf == 7)); }
bool xc_tau(int xf, int cf) { return xcf::is_mgga(xf) || xcf::is_mgga(cf); }

size_t xc_point_lds(const XCGridDevice &g, int xf, int cf, bool pol) {
  const int ns = pol ? 2 : 1, nd = 2 * g.dmax + 1;
  if (xc_generic(xf, cf)) return xc_gga_point_lds(g.npair, nd, g.lang, g.mang, ns, xc_tau(xf, cf));
  return sizeof(double) * ns * ((size_t)g.npair + (size_t)g.lang * nd + (size_t)g.lang * g.mang);
}

namespace {
template <int MODE>
void launch_general(const XCGridDevice &g, const XCArgs &a, const XCArgs &b, bool pol, size_t sh_in, size_t sh_pt,
                    size_t sh_out, hipStream_t s, KernelTimer *timer) {
  {
    TimedLaunch tl(timer, "xc_rad_in", s);
    hipLaunchKernelGGL(k_xc_rad_in<MODE>, dim3(g.Nel * g.Nang), dim3(NT), sh_in, s, a);
    if (pol) hipLaunchKernelGGL(k_xc_rad_in<MODE>, dim3(g.Nel * g.Nang), dim3(NT), sh_in, s, b);
    hip_check(hipGetLastError(), "xc_rad_in");
  }
  {
    TimedLaunch tl(timer, "xc_point", s);
    if (MODE == 0) {
      hipLaunchKernelGGL(k_xc_point, dim3(g.Nel * g.nrad), dim3(LP_NT), sh_pt, s, a);
    } else {
      if (sh_pt > 64 * 1024)
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_xc_point_gga<MODE>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh_pt),
                  "xc_point_gga LDS size");
      hipLaunchKernelGGL(k_xc_point_gga<MODE>, dim3(g.Nel * g.nrad), dim3(XP_NT), sh_pt, s, a);
    }
    hip_check(hipGetLastError(), "xc_point");
  }
  {
    TimedLaunch tl(timer, "xc_rad_out", s);
    hipLaunchKernelGGL(k_xc_rad_out<MODE>, dim3(g.Nel * g.Nang * XO_SPLIT), dim3(NT), sh_out, s, a);
    if (pol) hipLaunchKernelGGL(k_xc_rad_out<MODE>, dim3(g.Nel * g.Nang * XO_SPLIT), dim3(NT), sh_out, s, b);
    hip_check(hipGetLastError(), "xc_rad_out");
  }
}
}  // namespace

void xc_grid_launch(XCGridDevice &g, int xf, int cf, double thr, const double *Pa, int64_t ldpa, const double *Pb,
                    int64_t ldpb, double *Ha, int64_t ldha, double *Hb, int64_t ldhb, hipStream_t s,
                    KernelTimer *timer) {
  const bool pol = Pb != nullptr;
  const int mode = xc_tau(xf, cf) ? 2 : xc_generic(xf, cf) ? 1 : 0;
  fill_zero_launch(Ha, g.Nbf, g.Nbf, ldha, s);
  if (pol) fill_zero_launch(Hb, g.Nbf, g.Nbf, ldhb, s);
  const int nd = 2 * g.dmax + 1;
  XCArgs a{Pa, ldpa, Ha, ldha, g.Nel, g.Nang, g.nrad, g.lang, g.mang, nd, xf, cf, g.npair, thr,
           g.first.p, g.ni.p, g.skip.p, g.poff.p, g.rstart.p, g.rb.p, g.rx.p, g.dstart.p, g.xd.p,
           g.R.p, g.wpt.p, g.thx.p, g.thxt.p, g.cs.p, g.D.p, g.W.p, g.part.p, pol ? g.Db.p : nullptr,
           pol ? g.Wb.p : nullptr, g.dR.p, g.gsc.p, g.sn.p, g.thd.p, g.thdt.p, g.rxT.p,
           g.E.p, g.Y.p, pol ? g.Eb.p : nullptr, pol ? g.Yb.p : nullptr,
           g.thdd.p, g.thddt.p, g.thmm.p, g.thmmt.p, g.F.p, g.Z.p, pol ? g.Fb.p : nullptr, pol ? g.Zb.p : nullptr};
  XCArgs b = a;  // second spin: its density in, its matrix out
  if (pol) {
    b.P = Pb;
    b.ldp = ldpb;
    b.H = Hb;
    b.ldh = ldhb;
    b.D = g.Db.p;
    b.W = g.Wb.p;
    b.E = g.Eb.p;
    b.Y = g.Yb.p;
    b.F = g.Fb.p;
    b.Z = g.Zb.p;
  }
  const int rf = mode ? 2 : 1;
  const size_t sh_in =
      sizeof(double) * ((size_t)rf * g.nrad * XC_RS + (size_t)xc_pairs_per_group(g.nrad) * TILE2);
  const size_t sh_pt = xc_point_lds(g, xf, cf, pol);
  const size_t sh_out = sizeof(double) * rf * ((size_t)g.nrad * TILE + g.nrad);
  if (mode == 2)
    launch_general<2>(g, a, b, pol, sh_in, sh_pt, sh_out, s, timer);
  else if (mode == 1)
    launch_general<1>(g, a, b, pol, sh_in, sh_pt, sh_out, s, timer);
  else
    launch_general<0>(g, a, b, pol, sh_in, sh_pt, sh_out, s, timer);
}

}  // namespace hfg
