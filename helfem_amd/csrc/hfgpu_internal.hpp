// Internal types shared by the C-ABI layer and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include <stdexcept>
#include <cstdint>
#include <map>
#include "host/fem.hpp"
#include "host/diatomic.hpp"

namespace hfg {

// Element tiles are padded to 16 x 16 (LIP-15 elements hold 14 or 15
// functions), the MFMA f64 tile edge on CDNA.
constexpr int TILE = 16;
constexpr int TILE2 = TILE * TILE;
// operator extent per channel: c + 15 d for c, d < 15, padded to 4
constexpr int KC = 228;
// largest Cholesky rank the LDS-resident in-element diatomic exchange holds
constexpr int KL_RANK_MAX = 32;
// runs per workgroup of that kernel
constexpr int KL_BATCH = 64;
// two-step cross-element diatomic exchange: slot tiles per (L,|M|) group in
// LDS, rows per column shell (three per wave), threads per workgroup
#ifndef HFG_K2_THREADS
#define HFG_K2_THREADS 512
#endif
constexpr int K2_SLOTS = 17, K2_JMAX = 24, K2_THREADS = HFG_K2_THREADS;
static_assert(K2_THREADS % 64 == 0 && K2_JMAX % (K2_THREADS / 64) == 0, "rows split evenly over the waves");

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw Error(-3, std::string(what) + ": " + hipGetErrorString(e));
}

// Optional per-kernel timing with HIP events recorded on the launch stream.
struct KernelTimer {
  bool on = false;
  struct Pending {
    const char *name;
    hipEvent_t a, b;
    double work;
  };
  std::vector<Pending> pending;
  struct Acc {
    double ms = 0.0, work = 0.0;
    long count = 0;
  };
  std::map<std::string, Acc> acc;
  void begin(const char *name, hipStream_t s, double work, hipEvent_t *ev) {
    if (!on) return;
    hip_check(hipEventCreate(&ev[0]), "hipEventCreate");
    hip_check(hipEventCreate(&ev[1]), "hipEventCreate");
    hip_check(hipEventRecord(ev[0], s), "hipEventRecord");
    pending.push_back({name, ev[0], ev[1], work});
  }
  void end(hipStream_t s, hipEvent_t *ev) {
    if (!on) return;
    hip_check(hipEventRecord(ev[1], s), "hipEventRecord");
  }
  void collect() {
    for (auto &p : pending) {
      hip_check(hipEventSynchronize(p.b), "hipEventSynchronize");
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
      Acc &a = acc[p.name];
      a.ms += ms;
      a.work += p.work;
      a.count += 1;
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    pending.clear();
  }
};

// Scoped timing of one kernel launch.
struct TimedLaunch {
  KernelTimer *t;
  hipStream_t s;
  hipEvent_t ev[2];
  TimedLaunch(KernelTimer *t_, const char *name, hipStream_t s_, double work = 0.0) : t(t_), s(s_) {
    if (t) t->begin(name, s, work, ev);
  }
  ~TimedLaunch() {
    if (t && t->on) (void)hipEventRecord(ev[1], s);
  }
};

// Bytes the tiled density copy of an exchange may take: the environment
// variable HFGPU_EXCHANGE_TILE_GIB (GiB; 0 turns the tiles off), else half of
// the free device memory plus what the handle's tiles already hold.
double exchange_tile_budget(size_t held_bytes);

// n x n device matrix from its nonzero blocks (meta: row0, col0, rows, cols,
// offset per block; vals column-major), zero elsewhere; synchronises s.
// block-diagonal f M + l (l + 1) X per shell of n radial functions (atomic
// one-electron matrices; X and lval may be null)
void shell_diag_launch(double *out, int64_t ld, int64_t nbf, int n, const double *M, const double *X,
                       const int *lval, double f, hipStream_t s);
void scatter_blocks_launch(double *out, int64_t ld, int64_t n, const std::vector<long long> &meta,
                           const std::vector<double> &vals, hipStream_t s);

// Device buffer owned by a handle.
template <typename T>
struct DevArray {
  T *p = nullptr;
  size_t n = 0;
  DevArray() {}
  DevArray(const DevArray &) = delete;
  DevArray &operator=(const DevArray &) = delete;
  ~DevArray() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count <= n && p) return;
    release();
    if (count == 0) return;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
      p = nullptr;
      throw Error(-4, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
    }
    n = count;
  }
  void upload(const std::vector<T> &v, hipStream_t s) {
    alloc(v.size());
    if (!v.empty()) hip_check(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
  }
};

// Pinned host buffer (hipHostMalloc).
template <typename T>
struct PinnedArray {
  T *p = nullptr;
  size_t n = 0;
  PinnedArray() {}
  PinnedArray(const PinnedArray &) = delete;
  PinnedArray &operator=(const PinnedArray &) = delete;
  ~PinnedArray() {
    if (p) (void)hipHostFree(p);
  }
  void alloc(size_t count) {
    if (count <= n && p) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
    if (hipHostMalloc((void **)&p, count * sizeof(T), hipHostMallocDefault) != hipSuccess)
      throw Error(-4, "pinned host allocation failed");
    n = count;
  }
};

// Everything the atomic J/K kernels read, flattened for the device.
struct AtomicDevice {
  int Nel = 0, Nrad = 0, Nang = 0, N_L = 0, n_lm = 0, n_act = 0;
  DevArray<int> first, ni;                // per element
  DevArray<double> rs, rb;                // [N_L][Nel][TILE2]
  DevArray<double> chol;                  // packed M_p tiles
  DevArray<int> chol_off, chol_rank;      // [N_L*Nel]
  // Coulomb tables
  DevArray<int> lm_L, lm_pstart, pair_kl; // pair_kl = k*Nang + l
  DevArray<double> pair_c;
  DevArray<int> out_start, out_lm;        // per output (i,j): lm list
  DevArray<double> out_c;                 // gaunt * 4pi/(2L+1)
  // Exchange tables: the static (density-independent) coupling lists live on
  // the host; the device holds the lists compacted to the density blocks
  // that survive the reference's norm screen (TwoDBasis.cpp:949-952).
  std::vector<int> h_act, h_tstart, h_til;
  std::vector<double> h_tc, h_radL, h_radL_in;  // radial flops per coupled (block, L): all / in-element
  std::vector<double> h_radL_low;         // ... over the element pairs e > f
  std::vector<int> h_ni;
  DevArray<int> ef_in, ef_off;            // element pairs e*Nel+f of the two exchange launches
  DevArray<int> ef_low;                   // the pairs e > f (symmetric-density exchange)
  // Symmetric densities (max |P - P'| <= 1e-14 max |P|) take the symmetric
  // exchange: only tiles with e > f, or e == f and j <= k, are computed, each
  // also written transposed to its mirror (k,j,f,e).
  DevArray<double> stats;                 // block statistics of k_block_stats, and their host copy
  PinnedArray<double> h_stats;
  bool sym = false, list_sym = false;     // this call's mode; the mode the lists were built for
  std::vector<unsigned char> screen_mask;  // mask the compacted lists were built for
  DevArray<int> act_jk;                   // active output blocks j*Nang+k
  DevArray<int> term_start;               // [n_act*N_L + 1]
  DevArray<int> term_il;                  // i*Nang + l
  DevArray<int2> term_off;                // (i*Nrad, l*Nrad)
  DevArray<double> term_c;                // 4pi/(2L+1) * G_j * G_k
  PinnedArray<double> h_bnorm;
  // work buffers
  DevArray<double> Jin, jsm, jbg, preA, preB;
  DevArray<double> hostP, hostOut;        // staging for host-pointer entry points
  int max_rank = 0;
  int64_t n_terms = 0, n_pairs = 0;
  // algorithmic work per launch (flops), for the roofline report
  double flops_k_radial = 0.0, flops_k_angular = 0.0, flops_j = 0.0;
  double flops_in = 0.0, flops_off = 0.0;
  // In-element exchange as one dense operator per (L, element):
  // T[k = c + 15 d][row = a + 16 b] = (L L')(a + c n, b + d n), applied to
  // the density tiles of many (output block, L) runs at once.
  bool dense_in = false;
  DevArray<double> Top;                   // [N_L*Nel][KC][TILE2]
  DevArray<int> run_q, run_L, act_rstart; // runs (act, L) with terms, act-major
  DevArray<int> L_rstart, L_runs;         // runs grouped by L
  DevArray<int> wg_in;                    // per workgroup (L, e, first run in group) or -1
  DevArray<double> Ypart;                 // [n_runs][Nel][TILE2]
  // range-separated exchange (rs_kind 1 Yukawa: y_* factor tiles; 2 erfc:
  // element-pair operators e_T [N_L][Nel][Nel][TILE2][TILE2]); per-L scale
  // of the stored Gaunt weights
  int rs_kind = 0;
  DevArray<double> y_rs, y_rb, y_chol, e_T, rs_Lscale;
  DevArray<int> y_chol_off, y_chol_rank;
  // Density tiles of the screened source blocks in gather order
  // ([e*Nel+f][src][lane][4], see k_atomic_ptile), rebuilt per exchange call;
  // used when they fit the tile budget, else the gathers read P.
  DevArray<double> ptile;
  DevArray<float> ptile32;
  DevArray<int> src_blk, term_tile;
  int n_src = 0;
  bool tiled = false;
  // FP32 copies for the reduced-precision exchange (built on first use)
  DevArray<float> rs32, rb32, chol32, term_c32, P32;
  bool f32_tables = false;
  int64_t list_gen = 0, term_c32_gen = -1; // exchange-list version, and the one term_c32 holds
  int n_wg_in = 0, n_runs = 0;
};

// Everything the diatomic J/K kernels read.  Matrices are in the pure basis
// (Nbf x Nbf); shell a's radial function r sits at poff[a] + r - skip[a] when
// r >= skip[a] (m != 0 shells drop r = 0, basis.cpp:717-737).
struct DiatomicDevice {
  int Nel = 0, Nrad = 0, Nang = 0, n_lm = 0, n_LM = 0, n_act = 0, max_rank = 0;
  int64_t Nbf = 0;
  bool absm = false;
  DevArray<int> first, ni, skip;
  DevArray<long long> poff;
  DevArray<double> pq;                    // [n_lm][Nel][4][TILE2]: P0, P2, Q0, Q2
  DevArray<double> chol;                  // per (ilm, e): rank x {M0_p, M2_p} tiles
  DevArray<double> sigma;                 // per (ilm, e): rank signs
  DevArray<int> chol_off, chol_rank;      // [n_lm*Nel]; chol_off in units of 2*TILE2
  // Coulomb tables
  DevArray<int> LM_ilm;                   // per (L,M): index into the (L,|M|) caches
  DevArray<double> LM_fac;                // LMfac(L,M)
  DevArray<int> pstart, pair_kl;          // per (L,M) density pairs k*Nang + l
  DevArray<double> pair_c;                // (c0, c2) per pair
  DevArray<int> out_start, out_q;         // per output (i,j): (L,M) list
  DevArray<double> out_c;                 // (c0, c2) per entry
  // Exchange: per-shell Gaunt tables on the host; coupling lists built for
  // the density blocks that pass the norm screen (basis.cpp:1858-1875).
  std::vector<double> h_g0, h_g2;         // [j][i][L]: gaunt_mod / gaunt3 (M = mj - mi)
  int gL = 0;                             // L extent of h_g0/h_g2
  std::vector<unsigned char> screen_mask;
  std::vector<double> h_rad_in, h_rad_off; // radial flops per coupled (jk, ilm): in / across elements
  DevArray<int> ef_in, ef_off;            // element pairs e*Nel+f of the two exchange launches
  DevArray<int> ef_low;                   // the pairs e > f (symmetric-density exchange)
  // Symmetric densities (max |P - P'| <= 1e-14 max |P|, from k_block_stats)
  // take the symmetric exchange: only tiles with e > f, or e == f and j <= k,
  // are computed, each also written transposed to its mirror (k,j,f,e).
  DevArray<double> stats;                 // block statistics of k_block_stats, and their host copy
  PinnedArray<double> h_stats;
  bool sym = false, list_sym = false;     // this call's mode; the mode the lists were built for
  std::vector<double> h_rad_low;          // radial flops per coupled (jk, ilm) over the pairs e > f
  DevArray<int> act_jk, act_lstart;       // active output blocks; ilm runs per block
  DevArray<int> lrun_ilm, lrun_tstart;    // per run: ilm and first term
  DevArray<int> term_il;
  DevArray<int4> term_pos;                // per term (poff[i], poff[l] - skip[l], skip[i], skip[l])
  DevArray<double> term_f;                // 4 per term: f00, f02, f20, f22 (LMfac included)
  DevArray<int> mirror;                   // absm: pairs (dst jk, src jk)
  int n_mirror = 0;
  PinnedArray<double> h_bnorm;
  DevArray<double> Jin, ds, db, preA, preB;
  DevArray<double> hostP, hostOut;
  int64_t n_terms = 0, n_runs = 0, n_pairs = 0;
  double flops_k_radial = 0.0, flops_k_angular = 0.0;
  double flops_in = 0.0, flops_off = 0.0;  // split of the two above between the launches
  // In-element exchange as one dense operator per ((L,|M|), element):
  // T[k = ch*KC + c + 15 d][row = a + 16 b] = sign_ch W~_ch(c + a n, b + d n),
  // W~ = B diag(sigma) B' (the factorisation the reference contracts),
  // applied to the Gaunt-weighted density tiles of many output blocks at once.
  bool dense_in = false;
  DevArray<double> Top;                   // [n_lm*Nel][4*KC][TILE2]
  DevArray<int> ilm_rstart, ilm_runs;     // runs grouped by (L,|M|)
  DevArray<int> wg_in;                    // per workgroup (ilm, e, first run in group) or -1
  DevArray<double> Ypart;                 // [n_runs][Nel][TILE2] in-element partial tiles
  // in-element exchange with LDS-resident factors: per workgroup (ilm, e,
  // first entry in kl_runs, count); runs grouped by ilm; output block per run
  DevArray<int> wg_kl, kl_runs, run_act;
  int n_wg_kl = 0;
  // in-element exchange as a dense operator on the matrix cores
  // (k_dia_kin_op): the operators of every (ilm, element) in operand order,
  // built on the device at upload when they fit the budget; per workgroup
  // (ilm, e, first entry in kl_runs, count <= 16)
  bool op_in = false;
  DevArray<double> Top2;
  DevArray<int> wg_op;
  int n_wg_op = 0;
  // Density tiles of the screened source blocks in gather order
  // ([e*Nel+f][src][lane][4], see k_dia_ptile), rebuilt per exchange call;
  // src_blk lists the blocks i*Nang+l, term_tile maps each term to its tile.
  // Used when the tiles fit tile_budget bytes, else gathers read P.
  DevArray<double> ptile;
  DevArray<float> ptile32;
  DevArray<int> src_blk, term_tile;
  int n_src = 0;
  bool tiled = false;
  // FP32 copies for the reduced-precision exchange (built on first use)
  DevArray<float> pq32, chol32, sigma32, term_f32, P32;
  DevArray<double> lsig;                  // [n_lm*Nel] log of the P/Q scale of the FP32 tiles
  bool f32_tables = false;
  int64_t list_gen = 0, term_f32_gen = -1; // exchange-list version, and the one term_f32 holds
  int n_wg_in = 0;
  // Cross-element exchange in two steps (k_dia_k2_off), used when every
  // column block has at most K2_JMAX coupled rows and every group at most
  // K2_SLOTS slots: one workgroup per (column shell k, element pair); per
  // coupled ((L,|M|), parity of l_i) group of k, slot tiles V_b(i) = sum_l
  // v_b(k,l) P(i,l) (b: mod / plain Gaunt channel) in LDS, then each row j's
  // four channel tiles R = fac sum_i u(j,i) (x) V(i) and its radial products.
  bool k2 = false;
  int n_wg_k2 = 0;
  DevArray<int> k2_wg;                    // per workgroup: (k index, element pair) or (-1, -1)
  DevArray<int> k2_kshell, k2_kj, k2_jshell, k2_kg;  // per k: shell, rows (offsets + shells), groups
  DevArray<int> k2_gilm, k2_gslot, k2_gjt;           // per group: ilm, slot range, offset into k2_jto
  DevArray<int> k2_sl, k2_lt, k2_jto, k2_jts;        // slot lt ranges; lt tile ids; per (group,row) jt ranges; jt slots
  DevArray<double> k2_ltv, k2_jtc;                   // (v0, v1) per lt; (fac u0, fac u1) per jt
  double flops_off_k2 = 0.0;
};



// DFT grid of a two-dimensional basis, evaluated by sum factorisation:
// basis function (a, r) at point (i, t, k) is R_r(i) Theta_a(t) e^(i m_a phi_k).
// Atomic grid: R = B/r at the radial Chebyshev points, Theta_a = Y_a(theta, 0)
// on the cos(theta) rule, mang even phi points.  Pure-m diatomic grid: R = B
// at the mu points, Theta_a = Y_a(nu, 0), one phi "point" (the phi integral is
// analytic), only shell pairs with m_a = m_b.  Shell pairs (a, b) with
// |m_a - m_b| <= dmax are numbered p = 0..npair-1.
struct XCGridDevice {
  int Nel = 0, Nrad = 0, Nang = 0, nrad = 0, lang = 0, mang = 0, dmax = 0, npair = 0;
  int64_t Nbf = 0;
  bool lds_pol_ok = false;   // the polarized point kernel's LDS fits
  DevArray<int> first, ni, skip;
  DevArray<long long> poff;
  DevArray<double> R;        // [Nel][nrad][TILE]
  DevArray<double> wpt;      // [Nel][nrad][lang] quadrature weight (phi weight included)
  DevArray<double> cs;       // [2 dmax + 1][mang] cos(d phi_k)
  DevArray<int> rstart, rb, rx;       // pairs of row a: b and pair index x
  DevArray<int> dstart, xd;           // pairs x bucketed by d = m_a - m_b + dmax; d of each x
  DevArray<double> thx, thxt;         // Theta_a(t) Theta_b(t): [lang][npair] and [npair][lang]
  DevArray<double> D, W;     // D [Nel][nrad][npair], W [Nel][npair][nrad]
  DevArray<double> Db, Wb;   // second spin of a polarized evaluation
  // gradient (GGA) tables and work arrays
  DevArray<double> dR;       // [Nel][nrad][TILE] d/dr (atomic) or d/dmu (diatomic) of R
  DevArray<double> gsc;      // [Nel][nrad][lang][3] inverse scale factors 1/h of the gradient components
  DevArray<double> sn;       // [2 dmax + 1][mang] sin(d phi_k)
  DevArray<double> thd, thdt;  // d/dtheta (Theta_a Theta_b): [lang][npair] and [npair][lang]
  DevArray<int> rxT;         // per row-list entry: pair index of the transposed pair (b, a)
  DevArray<double> E, Y;     // E [Nel][nrad][npair] (d/dr on the row), Y [Nel][npair][nrad]
  DevArray<double> Eb, Yb;
  // kinetic energy density (meta-GGA): Theta'_a Theta'_b and m_a m_b Theta_a Theta_b
  // ([lang][npair] and [npair][lang]), F = R' P R' [Nel][nrad][npair], Z [Nel][npair][nrad]
  DevArray<double> thdd, thddt, thmm, thmmt;
  DevArray<double> F, Z, Fb, Zb;
  DevArray<double> part;     // [Nel*nrad][3]: Exc, Nel and Ekin (integral of tau) partial sums
  PinnedArray<double> h_part;
  DevArray<double> hostP, hostOut;
};

// launchers (atomic_jk.hip)
// work: null (the handle's buffers) or atomic_coulomb_work_doubles(d)
// doubles of caller-owned device memory.
void atomic_coulomb_launch(const AtomicDevice &d, const double *P, int64_t ldp, double *J, int64_t ldj,
                           hipStream_t s, KernelTimer *timer, double *work = nullptr);
size_t atomic_coulomb_work_doubles(const AtomicDevice &d);
// Device address of the allocation-free entry point's scratch array, looked
// up for `device` (the current device); once per device, at context creation.
void noalloc_scratch_init(int device);

// compute_tei with the in-element Slater tensors and their Cholesky factors
// built on the device (tei_atomic.hip); fills the same host caches.
// seconds (optional) gets the wall time.  Synchronises s.
void atomic_tei_device(AtomicSetup &S, hipStream_t s, double *seconds);
// the diatomic compute_tei with the two-channel kernels and their signed
// Cholesky factors built on the device (tei_diatomic.hip); P/Q integrals,
// rule sizes and outer-node Q values on the host.  Synchronises s.
void diatomic_tei_device(DiatomicSetup &S, hipStream_t s, double *seconds);
// fp32: gathers and matrix products in FP32 (the result stays FP64, with
// about 1e-6 relative error); otherwise the reference's FP64 arithmetic.
void atomic_exchange_launch(AtomicDevice &d, const double *P, int64_t ldp, double *K, int64_t ldk,
                            hipStream_t s, KernelTimer *timer, bool fp32 = false);
// Range-separated exchange -K_sr with the tables of atomic_upload_rs.
void atomic_rs_exchange_launch(AtomicDevice &d, const double *P, int64_t ldp, double *K, int64_t ldk,
                               hipStream_t s, KernelTimer *timer);
// Dense FP64 -> FP32 conversions (column-major with leading dimensions).
void to_f32_launch(const double *src, float *dst, size_t n, hipStream_t s);
void mat_to_f32_launch(const double *A, int64_t lda, float *B, int64_t n, hipStream_t s);
void fill_zero_launch(double *A, int64_t rows, int64_t cols, int64_t ld, hipStream_t s);
void atomic_build_operator(const AtomicSetup &S, AtomicDevice &d, hipStream_t s);

// Arguments of one matrix of a batched k_block_stats launch.
struct StatsArgs {
  const double *P;
  int64_t ldp;
  int Nrad, Nang, nch;
  const int *skip;
  const long long *poff;
  double *part;
};
void block_stats_batch_launch(const StatsArgs *d_args, const int *d_start, int n, int total_wg, hipStream_t s);

// Many atomic systems whose Coulomb and exchange builds run as one launch
// per kernel (atomic_jk.hip): per-system argument records and workgroup
// offsets live in device tables rewritten each call; the launch sequence
// after the host-side screening is captured once in a HIP graph and replayed
// while the grid totals and workgroup shapes stay the same.  The cross- and
// in-element exchange kernels come in three workgroup shapes (1, 2 or 4
// waves, splitting the multipoles L); an autotuner times each on the batch
// and keeps the fastest per (precision, tiled) class and composition.
struct AtomicBatch {
  std::vector<AtomicDevice *> sys;
  // device argument tables, written from a pinned staging copy
  DevArray<char> dtab_j, dtab_k, dtab_s;
  PinnedArray<char> htab_j, htab_k, htab_s;
  hipEvent_t ev_j = nullptr, ev_k = nullptr, ev_s = nullptr;  // the last copy out of each staging buffer
  DevArray<double> stats;
  PinnedArray<double> h_stats;
  std::vector<size_t> stats_off;
  // captured launch sequences and the grid signature each was captured for
  hipGraphExec_t gexec_j = nullptr, gexec_k = nullptr;
  std::vector<long long> sig_j, sig_k;
  bool use_graphs = true;
  // waves per workgroup of the exchange kernels used by the last call, per
  // class (fp32 * 2 + tiled) and launch part (0 cross-element, 1
  // in-element); autotuned once per size bucket and device for the process
  int nw[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  bool autotune = true;
  double tune_ms[4][2][3] = {};  // launch times of the 1, 2, 4-wave shapes, last tuning of this batch
  ~AtomicBatch();
};
void atomic_batch_coulomb(AtomicBatch &b, const double *const *P, const int64_t *ldp, double *const *J,
                          const int64_t *ldj, hipStream_t s, KernelTimer *timer);
// fp32[s] != 0: system s in reduced precision (as atomic_exchange_launch).
void atomic_batch_exchange(AtomicBatch &b, const double *const *P, const int64_t *ldp, double *const *K,
                           const int64_t *ldk, const int *fp32, hipStream_t s, KernelTimer *timer);

// launchers (xc_grid.hip); exc/nel summed on the host from h_part
// Pb == nullptr: restricted (P the total density, H the potential); else
// spin-polarized with Pa, Pb in and Ha, Hb out.
void xc_grid_launch(XCGridDevice &g, int xf, int cf, double thr, const double *Pa, int64_t ldpa, const double *Pb,
                    int64_t ldpb, double *Ha, int64_t ldha, double *Hb, int64_t ldhb, hipStream_t s,
                    KernelTimer *timer);
// true when (xf, cf) is not the plain lda_x + lda_c_vwn pair: the general
// point kernel (GGA terms, any supported functional) runs
bool xc_generic(int xf, int cf);
// dynamic LDS of the point kernel that xc_grid_launch will use
size_t xc_point_lds(const XCGridDevice &g, int xf, int cf, bool pol);
// true when the pair needs the kinetic energy density (meta-GGA)
bool xc_tau(int xf, int cf);

// launchers (diatomic_jk.hip)
void diatomic_coulomb_launch(const DiatomicDevice &d, const double *P, int64_t ldp, double *J, int64_t ldj,
                             hipStream_t s, KernelTimer *timer);
// One pass over P (k_block_stats): per shell block pair i <= l and 32-row
// chunk, partial squared norms of P_il and P_li and partial maxima of |P_il -
// P_li'| and |P_il|, copied to h_part (4 * chunks * Nang^2 doubles, pinned)
// on stream s (skip/poff null: shell a is rows a*Nrad .. a*Nrad + Nrad).
int block_stats_chunks(int Nrad);
void block_stats_launch(const double *P, int64_t ldp, int Nrad, int Nang, const int *skip, const long long *poff,
                        double *d_part, double *h_part, hipStream_t s);
// After the copy: the Frobenius norm of every shell block into bnorm[Nang^2];
// returns max |P - P'| <= 1e-14 max |P| (a NaN fails).
bool block_stats_reduce(const double *h_part, int Nrad, int Nang, double *bnorm);
// false when the environment variable HFGPU_EXCHANGE_SYMMETRIC is 0
bool symmetric_exchange_allowed();
// fp32 as for atomic_exchange_launch.
void diatomic_exchange_kernels(DiatomicDevice &d, const double *P, int64_t ldp, double *K, int64_t ldk,
                               hipStream_t s, KernelTimer *timer, bool fp32 = false);
void diatomic_build_operator(const DiatomicSetup &S, DiatomicDevice &d, hipStream_t s);
// the matrix-core operators of k_dia_kin_op from the device Cholesky tiles
// (opt-in with HFGPU_EXCHANGE_OP=1; sets d.op_in when they fit the budget)
void diatomic_op_build(DiatomicDevice &d, hipStream_t s);
constexpr int OP_RUNS = 16;  // runs per k_dia_kin_op workgroup

}  // namespace hfg
