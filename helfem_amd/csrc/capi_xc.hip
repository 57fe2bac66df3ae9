// C-ABI for the DFT grids (include/hfgpu.h): grid tables and the LDA
// exchange-correlation matrix (kernels in xc_grid.hip).
#include "capi_common.hpp"
#include "xc_func.hpp"
#include <cmath>
#include <memory>

using namespace hfg;

struct hf_dftgrid {
  hf_ctx *ctx = nullptr;
  const hfg::AtomicSetup *atomic = nullptr;
  const hfg::DiatomicSetup *diatomic = nullptr;
  int lang = 0, mang = 0;
  bool full3d = false;  // diatomic: the general (mu, nu, phi) grid instead of the pure-m one
  int rule = 0;         // atomic: 0 the reference's rule, 1 Gauss-Legendre cos(theta) x Gauss-Lobatto r
  hfg::XCGridDevice dev;
  bool uploaded = false;
  int64_t Nbf() const { return atomic ? atomic->Nbf() : diatomic->Nbf(); }
};

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;

// Theta_a(x) = Y_a(arccos x, phi = 0) from std::sph_legendre (Condon-Shortley
// phase; Y_l^-m = (-1)^m conj Y_l^m, spherical_harmonics.cpp:19-35).
double theta_value(int l, int m, double x) {
  const int ma = std::abs(m);
  double y = std::sph_legendre((unsigned)l, (unsigned)ma, std::acos(x));
  if (m < 0 && (ma & 1)) y = -y;
  return y;
}

// dTheta_a/dtheta at x = cos(theta): m cot(theta) Theta_lm + sqrt((l - m)(l + m + 1))
// Theta_l,m+1, the phi = 0 value of the reference's d/dtheta Y_lm
// (TwoDBasis.cpp:1181-1199).
double theta_deriv(int l, int m, double x) {
  const double sinth = std::sqrt(std::max((1.0 - x) * (1.0 + x), 0.0));
  const double cotth = sinth > 0.0 ? x / sinth : 0.0;
  double v = m * cotth * theta_value(l, m, x);
  if (m < l) v += std::sqrt(double((l - m) * (l + m + 1))) * theta_value(l, m + 1, x);
  return v;
}

// Grid tables shared by both bases.  Atomic (dftgrid.cpp:470-576): R = B/r
// at the radial Chebyshev points, weight w_cth 2pi/mang * h w_i * r_i^2 with
// the compound angular rule of angular.cpp:20-72.  Pure-m diatomic
// (dftgrid_purem.cpp:87-120): R = B at the mu points, weight
// 2pi w_nu h w_i Rh^3 sinh(mu) (sinh^2 mu + sin^2 nu), one phi point, only
// shell pairs of equal m.  General 3D diatomic (src/diatomic/dftgrid.cpp:
// 414-460): R = B, weight w_nu 2pi/mang h w_i Rh^3 sinh(mu) (sinh^2 mu +
// sin^2 nu) on the same compound rule as the atomic grid, all shell pairs.
void xc_upload(hf_dftgrid *g) {
  XCGridDevice &d = g->dev;
  hipStream_t s = g->ctx->cur;
  const RadialFE &rad = g->atomic ? g->atomic->rad : g->diatomic->rad;
  const std::vector<int> &lv = g->atomic ? g->atomic->lval : g->diatomic->lval;
  const std::vector<int> &mv = g->atomic ? g->atomic->mval : g->diatomic->mval;
  const int Nel = rad.Nel(), Na = (int)lv.size(), nrad = rad.nquad;
  need(nrad >= 1 && nrad <= 1024, "radial quadrature size out of range");
  d.Nel = Nel;
  d.Nrad = rad.Nrad();
  d.Nang = Na;
  d.Nbf = g->Nbf();
  d.nrad = nrad;
  d.lang = g->lang;
  d.mang = g->mang;
  int mmax = 0;
  for (int m : mv) mmax = std::max(mmax, std::abs(m));
  d.dmax = (g->atomic || g->full3d) ? 2 * mmax : 0;
  const int nd = 2 * d.dmax + 1;
  // shell pairs within reach of the phi grid, numbered row by row
  std::vector<int> rstart(Na + 1, 0), rb, rp;
  std::vector<std::vector<int>> byd_a(nd), byd_b(nd), byd_p(nd);
  int np = 0;
  for (int a = 0; a < Na; ++a) {
    for (int b = 0; b < Na; ++b) {
      const int dd = mv[a] - mv[b];
      if (std::abs(dd) > d.dmax) continue;
      rb.push_back(b);
      rp.push_back(np);
      byd_a[dd + d.dmax].push_back(a);
      byd_b[dd + d.dmax].push_back(b);
      byd_p[dd + d.dmax].push_back(np);
      ++np;
    }
    rstart[a + 1] = (int)rb.size();
  }
  d.npair = np;
  // pairs renumbered by class d (x = position in the d-sorted order): the
  // work arrays D, W are indexed by x, and rx maps each row-list entry to it
  std::vector<int> dstart(nd + 1, 0), da, db, px(np), xd;
  for (int dd = 0; dd < nd; ++dd) {
    for (size_t j = 0; j < byd_p[dd].size(); ++j) px[byd_p[dd][j]] = (int)da.size() + (int)j;
    da.insert(da.end(), byd_a[dd].begin(), byd_a[dd].end());
    db.insert(db.end(), byd_b[dd].begin(), byd_b[dd].end());
    xd.insert(xd.end(), byd_a[dd].size(), dd);
    dstart[dd + 1] = (int)da.size();
  }
  std::vector<int> rx(rp.size());
  for (size_t q = 0; q < rp.size(); ++q) rx[q] = px[rp[q]];
  const size_t lds = sizeof(double) * ((size_t)np + (size_t)g->lang * nd + (size_t)g->lang * g->mang);
  if (lds > 64 * 1024)
    throw Error(HF_EINVAL, "angular basis and DFT grid too large for one workgroup's LDS (need shell pairs + "
                           "lang*(4 mmax + 1) + lang*mang <= 8192)");
  d.lds_pol_ok = 2 * lds <= 64 * 1024;
  if (sizeof(double) * ((size_t)nrad * TILE + 4 * TILE2) > 64 * 1024) throw Error(HF_EINVAL, "radial quadrature too large");
  std::vector<double> xq, wq, xt, wt;
  if (g->rule == 1) {
    gauss_lobatto(nrad, xq, wq);
    gauss_legendre(g->lang, xt, wt);
  } else {
    gauss_chebyshev2(nrad, xq, wq);
    gauss_chebyshev2(g->lang, xt, wt);
  }
  const double dphi = 2.0 * kPi / g->mang;
  std::vector<double> R((size_t)Nel * nrad * TILE, 0.0), wpt((size_t)Nel * nrad * g->lang);
  std::vector<double> th((size_t)g->lang * Na), cs((size_t)nd * g->mang);
  for (int t = 0; t < g->lang; ++t)
    for (int a = 0; a < Na; ++a) th[(size_t)t * Na + a] = theta_value(lv[a], mv[a], xt[t]);
  for (int dd = 0; dd < nd; ++dd)
    for (int k = 0; k < g->mang; ++k) cs[(size_t)dd * g->mang + k] = std::cos(double(dd - d.dmax) * (k * dphi));
  // Theta_a(t) Theta_b(t) per pair x, in both orders ([t][x] and [x][t])
  std::vector<double> thx((size_t)g->lang * np), thxt((size_t)np * g->lang);
  for (int t = 0; t < g->lang; ++t)
    for (int x = 0; x < np; ++x) {
      const double v = th[(size_t)t * Na + da[x]] * th[(size_t)t * Na + db[x]];
      thx[(size_t)t * np + x] = v;
      thxt[(size_t)x * g->lang + t] = v;
    }
  for (int e = 0; e < Nel; ++e) {
    const Blk v = g->atomic ? rad.rvalues(xq, e) : rad.values(0, xq, e);
    const double h = rad.half_len(e), mid = rad.mid(e);
    for (int i = 0; i < nrad; ++i) {
      for (int c = 0; c < v.c; ++c) R[((size_t)e * nrad + i) * TILE + c] = v(i, c);
      const double x = mid + h * xq[i], wr = h * wq[i];
      double *w = &wpt[((size_t)e * nrad + i) * g->lang];
      if (g->atomic) {
        for (int t = 0; t < g->lang; ++t) w[t] = (wt[t] * dphi) * wr * std::pow(x, 2);
      } else {
        const double shmu = std::sinh(x), Rh3 = std::pow(g->diatomic->Rh, 3);
        for (int t = 0; t < g->lang; ++t) {
          const double sth = std::sqrt(std::max((1.0 - xt[t]) * (1.0 + xt[t]), 0.0));
          const double wa = g->full3d ? wt[t] * dphi : 2.0 * kPi * wt[t];
          w[t] = wa * wr * Rh3 * shmu * (shmu * shmu + sth * sth);
        }
      }
    }
  }
  std::vector<int> skip(Na, 0);
  std::vector<long long> poff(Na + 1, 0);
  for (int a = 0; a < Na; ++a) {
    skip[a] = g->atomic ? 0 : g->diatomic->skip(a);
    poff[a] = g->atomic ? (long long)a * d.Nrad : (long long)g->diatomic->poff[a];
  }
  d.first.upload(rad.first, s);
  d.ni.upload(rad.ni, s);
  d.skip.upload(skip, s);
  d.poff.upload(poff, s);
  d.R.upload(R, s);
  d.wpt.upload(wpt, s);
  d.cs.upload(cs, s);
  d.rstart.upload(rstart, s);
  d.rb.upload(rb, s);
  d.rx.upload(rx, s);
  d.dstart.upload(dstart, s);
  d.xd.upload(xd, s);
  d.thx.upload(thx, s);
  d.thxt.upload(thxt, s);
  // gradient tables (GGA): d/dr R (atomic, R = B/r) or d/dmu R (diatomic,
  // R = B), the inverse scale factors of the three gradient components per
  // point, sin(d phi), d/dtheta (Theta_a Theta_b), and the transposed pair of
  // each row-list entry; the host vectors live until the stream sync below
  std::vector<double> dR((size_t)Nel * nrad * TILE, 0.0), gsc((size_t)Nel * nrad * g->lang * 3);
  std::vector<double> sn((size_t)nd * g->mang), thd((size_t)g->lang * np), thdt((size_t)np * g->lang);
  std::vector<double> thdd((size_t)g->lang * np), thddt((size_t)np * g->lang), thmm((size_t)g->lang * np),
      thmmt((size_t)np * g->lang);
  std::vector<int> rxT(rb.size());
  {
    for (int e = 0; e < Nel; ++e) {
      const Blk v = g->atomic ? rad.rderiv(xq, e) : rad.values(1, xq, e);
      const double h = rad.half_len(e), mid = rad.mid(e);
      for (int i = 0; i < nrad; ++i) {
        for (int c = 0; c < v.c; ++c) dR[((size_t)e * nrad + i) * TILE + c] = v(i, c);
        const double x = mid + h * xq[i];
        for (int t = 0; t < g->lang; ++t) {
          const double sth = std::sqrt(std::max((1.0 - xt[t]) * (1.0 + xt[t]), 0.0));
          double *c3 = &gsc[(((size_t)e * nrad + i) * g->lang + t) * 3];
          if (g->atomic) {
            c3[0] = 1.0;
            c3[1] = 1.0 / x;
            c3[2] = 1.0 / (x * sth);
          } else {
            const double Rh = g->diatomic->Rh, shmu = std::sinh(x);
            c3[0] = c3[1] = 1.0 / (Rh * std::sqrt(shmu * shmu + sth * sth));
            // 1/h_phi: the phi gradient of a pure-m density vanishes (d = 0), but
            // its kinetic energy density keeps the m^2 / h_phi^2 term
            c3[2] = 1.0 / (Rh * shmu * sth);
          }
        }
      }
    }
    for (int dd = 0; dd < nd; ++dd)
      for (int k = 0; k < g->mang; ++k) sn[(size_t)dd * g->mang + k] = std::sin(double(dd - d.dmax) * (k * dphi));
    std::vector<double> dth((size_t)g->lang * Na);
    for (int t = 0; t < g->lang; ++t)
      for (int a = 0; a < Na; ++a) dth[(size_t)t * Na + a] = theta_deriv(lv[a], mv[a], xt[t]);
    for (int t = 0; t < g->lang; ++t)
      for (int x = 0; x < np; ++x) {
        const size_t A = (size_t)t * Na + da[x], B = (size_t)t * Na + db[x];
        const double v = dth[A] * th[B] + th[A] * dth[B];
        thd[(size_t)t * np + x] = v;
        thdt[(size_t)x * g->lang + t] = v;
        const double vd = dth[A] * dth[B], vm = double(mv[da[x]] * mv[db[x]]) * th[A] * th[B];
        thdd[(size_t)t * np + x] = vd;
        thddt[(size_t)x * g->lang + t] = vd;
        thmm[(size_t)t * np + x] = vm;
        thmmt[(size_t)x * g->lang + t] = vm;
      }
    std::vector<int> xid((size_t)Na * Na, -1);
    for (int a = 0; a < Na; ++a)
      for (int q = rstart[a]; q < rstart[a + 1]; ++q) xid[(size_t)a * Na + rb[q]] = rx[q];
    for (int a = 0; a < Na; ++a)
      for (int q = rstart[a]; q < rstart[a + 1]; ++q) {
        rxT[q] = xid[(size_t)rb[q] * Na + a];
        need(rxT[q] >= 0, "pair list is not symmetric");
      }
    d.dR.upload(dR, s);
    d.gsc.upload(gsc, s);
    d.sn.upload(sn, s);
    d.thd.upload(thd, s);
    d.thdt.upload(thdt, s);
    d.rxT.upload(rxT, s);
    d.thdd.upload(thdd, s);
    d.thddt.upload(thddt, s);
    d.thmm.upload(thmm, s);
    d.thmmt.upload(thmmt, s);
  }
  d.D.alloc((size_t)Nel * np * nrad);
  d.W.alloc((size_t)Nel * np * nrad);
  d.part.alloc((size_t)3 * Nel * nrad);
  d.h_part.alloc((size_t)3 * Nel * nrad);
  hip_check(hipStreamSynchronize(s), "dft grid upload");
  g->uploaded = true;
}

// lda_x (1) with lda_c_vwn (7) through the LDA point kernel; PBE exchange
// (101), PW92 (12, 13) and PBE (130) correlation, alone or mixed with those,
// through the general point kernel (xc_func.hpp)
void check_functionals(const hf_dftgrid *g, int x_func, int c_func) {
  (void)g;
  if (!xcf::exchange_ok(x_func))
    throw Error(HF_EINVAL, "unsupported exchange functional (0, 1 = lda_x, 101 = gga_x_pbe, 106 = gga_x_b88, "
                           "202 = mgga_x_tpss, 402 = hyb_gga_xc_b3lyp, 406 = hyb_gga_xc_pbeh, "
                           "433 = hyb_gga_xc_cam_b3lyp, 470 = hyb_gga_xc_camy_b3lyp)");
  if (!xcf::correlation_ok(c_func))
    throw Error(HF_EINVAL, "unsupported correlation functional (0, 7 = lda_c_vwn, 8 = lda_c_vwn_rpa, 12 = lda_c_pw, "
                           "13 = lda_c_pw_mod, 130 = gga_c_pbe, 131 = gga_c_lyp, 231 = mgga_c_tpss)");
}

// Pb == nullptr: restricted; else spin-polarized (Hb gets the beta potential,
// zeroed when `beta` is false, as the reference leaves it).
void eval_fxc(hf_dftgrid *g, int x_func, int c_func, const double *P, int64_t ldp, const double *Pb, int64_t ldpb,
              double *H, int64_t ldh, double *Hb, int64_t ldhb, bool beta, double thr, double *out) {
  hipStream_t s = g->ctx->cur;
  XCGridDevice &d = g->dev;
  const bool gen = xc_generic(x_func, c_func);
  if (gen) {
    if (xc_point_lds(d, x_func, c_func, Pb != nullptr) > 160 * 1024)
      throw Error(HF_EINVAL, "angular basis and DFT grid too large for one workgroup's LDS in the GGA evaluation");
    d.E.alloc(d.D.n);
    d.Y.alloc(d.W.n);
    if (xc_tau(x_func, c_func)) {
      d.F.alloc(d.D.n);
      d.Z.alloc(d.W.n);
    }
  }
  if (Pb) {
    if (!gen && !d.lds_pol_ok) throw Error(HF_EINVAL, "DFT grid too large for the spin-polarized evaluation");
    d.Db.alloc(d.D.n);
    d.Wb.alloc(d.W.n);
    if (gen) {
      d.Eb.alloc(d.D.n);
      d.Yb.alloc(d.W.n);
    }
    if (xc_tau(x_func, c_func)) {
      d.Fb.alloc(d.D.n);
      d.Zb.alloc(d.W.n);
    }
  }
  xc_grid_launch(d, x_func, c_func, thr, P, ldp, Pb, ldpb, H, ldh, Hb, ldhb, s, &g->ctx->timer);
  if (Pb && !beta) fill_zero_launch(Hb, d.Nbf, d.Nbf, ldhb, s);
  hip_check(hipMemcpyAsync(d.h_part.p, d.part.p, sizeof(double) * 3 * d.Nel * d.nrad, hipMemcpyDeviceToHost, s),
            "copy XC partial sums");
  hip_check(hipStreamSynchronize(s), "XC sync");
  double exc = 0.0, nel = 0.0, ekin = 0.0;
  for (int q = 0; q < d.Nel * d.nrad; ++q) {
    exc += d.h_part.p[3 * q];
    nel += d.h_part.p[3 * q + 1];
    ekin += d.h_part.p[3 * q + 2];
  }
  out[0] = exc;
  out[1] = nel;
  out[2] = ekin;  // integral of the kinetic energy density: zero without a meta-GGA
}

hf_dftgrid *ready(hf_dftgrid *g) {
  need(g != nullptr, "NULL grid");
  need(g->ctx != nullptr, "basis was created without a device context");
  return g;
}

}  // namespace

extern "C" {

int hf_atomic_dftgrid_create_rule(hf_atomic *h, int lang, int mang, int rule, hf_dftgrid **out) {
  return guarded([&] {
    need(h && out, "NULL argument");
    need(lang >= 1 && mang >= 1, "grid orders must be positive");
    need(rule == 0 || rule == 1, "rule must be 0 (reference) or 1 (Gauss-Legendre x Gauss-Lobatto)");
    need(rule == 0 || h->host.rad.nquad >= 2, "a Gauss-Lobatto radial rule needs at least 2 points");
    std::unique_ptr<hf_dftgrid> g(new hf_dftgrid);
    g->ctx = h->ctx;
    g->atomic = &h->host;
    g->lang = lang;
    g->mang = mang;
    g->rule = rule;
    *out = g.release();
  });
}

int hf_atomic_dftgrid_create(hf_atomic *h, int lang, int mang, hf_dftgrid **out) {
  return hf_atomic_dftgrid_create_rule(h, lang, mang, 0, out);
}

int hf_quadrature(int kind, int n, double *x, double *w) {
  return guarded([&] {
    need(x && w, "NULL argument");
    need(n >= 1, "n must be positive");
    std::vector<double> xv, wv;
    if (kind == 0)
      gauss_chebyshev2(n, xv, wv);
    else if (kind == 1)
      gauss_lobatto(n, xv, wv);
    else if (kind == 2)
      gauss_legendre(n, xv, wv);
    else
      throw Error(HF_EINVAL, "kind must be 0 (Chebyshev), 1 (Gauss-Lobatto) or 2 (Gauss-Legendre)");
    std::copy(xv.begin(), xv.end(), x);
    std::copy(wv.begin(), wv.end(), w);
  });
}

int hf_diatomic_dftgrid_create(hf_diatomic *h, int lang, hf_dftgrid **out) {
  return guarded([&] {
    need(h && out, "NULL argument");
    need(lang >= 1, "grid order must be positive");
    std::unique_ptr<hf_dftgrid> g(new hf_dftgrid);
    g->ctx = h->ctx;
    g->diatomic = &h->host;
    g->lang = lang;
    g->mang = 1;
    *out = g.release();
  });
}

int hf_diatomic_dftgrid3d_create(hf_diatomic *h, int lang, int mang, hf_dftgrid **out) {
  return guarded([&] {
    need(h && out, "NULL argument");
    need(lang >= 1 && mang >= 1, "grid orders must be positive");
    std::unique_ptr<hf_dftgrid> g(new hf_dftgrid);
    g->ctx = h->ctx;
    g->diatomic = &h->host;
    g->lang = lang;
    g->mang = mang;
    g->full3d = true;
    *out = g.release();
  });
}

int hf_dftgrid_destroy(hf_dftgrid *g) {
  return guarded([&] {
    if (!g) return;
    if (g->ctx) {
      DeviceGuard dg(g->ctx->device);
      (void)hipStreamSynchronize(g->ctx->cur);
    }
    delete g;
  });
}

int hf_dftgrid_eval_fxc_dev(hf_dftgrid *g, int x_func, int c_func, const double *dP, int64_t ldp, double *dH,
                            int64_t ldh, double thr, double out[3]) {
  return guarded([&] {
    ready(g);
    need(out != nullptr, "out is NULL");
    check_functionals(g, x_func, c_func);
    DeviceGuard dg(g->ctx->device);
    const int64_t n = g->Nbf();
    check_matrix(dP, ldp, n, "P");
    check_matrix(dH, ldh, n, "H");
    if (!g->uploaded) xc_upload(g);
    eval_fxc(g, x_func, c_func, dP, ldp, nullptr, 0, dH, ldh, nullptr, 0, false, thr, out);
  });
}

int hf_dftgrid_eval_fxc(hf_dftgrid *g, int x_func, int c_func, const double *P, int64_t ldp, double *H,
                        int64_t ldh, double thr, double out[3]) {
  return guarded([&] {
    ready(g);
    need(out != nullptr, "out is NULL");
    check_functionals(g, x_func, c_func);
    DeviceGuard dg(g->ctx->device);
    if (!g->uploaded) xc_upload(g);
    host_roundtrip(g->ctx->cur, g->Nbf(), g->dev.hostP, g->dev.hostOut, P, ldp, H, ldh,
                   [&](const double *dp, int64_t lp, double *dout, int64_t lo) {
                     eval_fxc(g, x_func, c_func, dp, lp, nullptr, 0, dout, lo, nullptr, 0, false, thr, out);
                   });
  });
}

int hf_dftgrid_eval_fxc_pol_dev(hf_dftgrid *g, int x_func, int c_func, const double *dPa, int64_t ldpa,
                                const double *dPb, int64_t ldpb, double *dHa, int64_t ldha, double *dHb, int64_t ldhb,
                                int beta, double thr, double out[3]) {
  return guarded([&] {
    ready(g);
    need(out != nullptr, "out is NULL");
    check_functionals(g, x_func, c_func);
    DeviceGuard dg(g->ctx->device);
    const int64_t n = g->Nbf();
    check_matrix(dPa, ldpa, n, "Pa");
    check_matrix(dPb, ldpb, n, "Pb");
    check_matrix(dHa, ldha, n, "Ha");
    check_matrix(dHb, ldhb, n, "Hb");
    if (!g->uploaded) xc_upload(g);
    eval_fxc(g, x_func, c_func, dPa, ldpa, dPb, ldpb, dHa, ldha, dHb, ldhb, beta != 0, thr, out);
  });
}

int hf_dftgrid_eval_fxc_pol(hf_dftgrid *g, int x_func, int c_func, const double *Pa, int64_t ldpa, const double *Pb,
                            int64_t ldpb, double *Ha, int64_t ldha, double *Hb, int64_t ldhb, int beta, double thr,
                            double out[3]) {
  return guarded([&] {
    ready(g);
    need(out != nullptr, "out is NULL");
    check_functionals(g, x_func, c_func);
    DeviceGuard dg(g->ctx->device);
    const int64_t n = g->Nbf();
    check_matrix(Pa, ldpa, n, "Pa");
    check_matrix(Pb, ldpb, n, "Pb");
    check_matrix(Ha, ldha, n, "Ha");
    check_matrix(Hb, ldhb, n, "Hb");
    if (!g->uploaded) xc_upload(g);
    XCGridDevice &d = g->dev;
    hipStream_t s = g->ctx->cur;
    const size_t nn = (size_t)n * n;
    d.hostP.alloc(2 * nn);
    d.hostOut.alloc(2 * nn);
    hip_check(hipMemcpy2DAsync(d.hostP.p, n * sizeof(double), Pa, ldpa * sizeof(double), n * sizeof(double), n,
                               hipMemcpyHostToDevice, s), "copy Pa");
    hip_check(hipMemcpy2DAsync(d.hostP.p + nn, n * sizeof(double), Pb, ldpb * sizeof(double), n * sizeof(double),
                               n, hipMemcpyHostToDevice, s), "copy Pb");
    eval_fxc(g, x_func, c_func, d.hostP.p, n, d.hostP.p + nn, n, d.hostOut.p, n, d.hostOut.p + nn, n, beta != 0,
             thr, out);
    hip_check(hipMemcpy2DAsync(Ha, ldha * sizeof(double), d.hostOut.p, n * sizeof(double), n * sizeof(double), n,
                               hipMemcpyDeviceToHost, s), "copy Ha");
    hip_check(hipMemcpy2DAsync(Hb, ldhb * sizeof(double), d.hostOut.p + nn, n * sizeof(double), n * sizeof(double),
                               n, hipMemcpyDeviceToHost, s), "copy Hb");
    hip_check(hipStreamSynchronize(s), "stream sync");
  });
}

int hf_xc_exact_exchange(int x_func, double out[4]) {
  return guarded([&] {
    need(out != nullptr, "out is NULL");
    if (!xcf::exchange_ok(x_func)) throw Error(HF_EINVAL, "unsupported exchange functional");
    int kernel = 0;
    xcf::exact_exchange(x_func, out[0], out[1], out[2], kernel);
    out[3] = kernel;
  });
}

}  // extern "C"
