/*
 * hfgpu: MI355X (gfx950) Fock-build engine for HelFEM -- C ABI.
 *
 * The entry points replace the Fock-build members of HelFEM's basis and
 * grid classes.  Each declaration names the reference interface it stands
 * in for (paths relative to the HelFEM source tree).
 *
 * Conventions
 *   - Matrices are column-major doubles with an explicit leading dimension,
 *     as helfem::Mat<double> (an Eigen column-major matrix,
 *     libhelfem/include/Matrix.h:61).
 *   - Outputs are caller-allocated.  Every function returns 0 on success and
 *     a negative HF_E* code on failure; hf_last_error() then returns a
 *     thread-local message.  No C++ exception crosses this boundary.
 *   - One handle is stream-ordered and must not be used from two host
 *     threads at once; different handles may be used concurrently.
 *   - Functions ending in _dev take device pointers and queue their work on
 *     the context stream.  The exchange forms first read the density's block
 *     norms back to the host (the reference's screen decides which blocks
 *     are contracted), so they synchronise the stream once before queueing
 *     the contraction; the others do not synchronise.  All other functions
 *     take host pointers and return when the result is in host memory.
 */
#ifndef HFGPU_H
#define HFGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HF_OK 0
#define HF_EINVAL (-1)   /* bad argument (shape, index, null pointer) */
#define HF_ESTATE (-2)   /* call out of order, e.g. J/K before compute_tei */
#define HF_EDEVICE (-3)  /* HIP runtime error */
#define HF_ENOMEM (-4)   /* device or host allocation failed */
#define HF_EINTERNAL (-5)

typedef struct hf_ctx hf_ctx;
typedef struct hf_atomic hf_atomic;
typedef struct hf_diatomic hf_diatomic;
typedef struct hf_dftgrid hf_dftgrid;

/* Thread-local text of the last error. */
const char *hf_last_error(void);
/* Library version string. */
const char *hf_version(void);

/* ---------------------------------------------------------------- context */

/* Bind device `device` and create a stream-ordered context.  flags: 0. */
int hf_ctx_create(int device, unsigned flags, hf_ctx **out);
int hf_ctx_destroy(hf_ctx *ctx);
/* Run subsequent work of this context on `stream`, a hipStream_t (NULL = the
 * default stream).  A new context starts on a stream of its own. */
int hf_ctx_set_stream(hf_ctx *ctx, void *stream);
/* The hipStream_t work is issued on. */
int hf_ctx_get_stream(hf_ctx *ctx, void **stream);
int hf_ctx_synchronize(hf_ctx *ctx);
/* Per-kernel timing: with timing on, every kernel launch of this context is
 * bracketed by HIP events on its stream.  hf_ctx_kernel_stats returns, for
 * kernel `name`, out[0] = launches, out[1] = summed milliseconds, out[2] =
 * summed algorithmic flops.  Turning timing off clears the totals.  (No
 * counterpart in the reference, which has only wall-clock FockTimer,
 * src/general/scf_driver_common.h:69-170.) */
int hf_ctx_set_timing(hf_ctx *ctx, int on);
int hf_ctx_kernel_stats(hf_ctx *ctx, const char *name, double out[3]);

/* Element boundaries of the radial (atomic: r, diatomic: mu) grid on
 * [0, rmax]: igrid 1 linear, 2 quadratic, 3 polynomial (exponent zexp),
 * 4 generalised exponential; out gets nel+1 values.  Replaces
 * helfem::utils::get_grid (libhelfem/include/grid.h:37-101), called by
 * atomic::basis::normal_grid (src/atomic/basis.cpp:45-53). */
int hf_element_grid(double rmax, int nel, int igrid, double zexp, double *out);

/* ----------------------------------------------------------- atomic basis */

/* Radial finite elements x complex Y_lm.  Replaces the constructor
 * helfem::atomic::basis::TwoDBasisT<double>::TwoDBasisT
 * (src/atomic/TwoDBasis.h:142-149) for a point nucleus, LIP shape functions
 * on `nnodes` Gauss-Lobatto nodes (primbas 4), zero value at r=0 and
 * r=bval[nbval-1].  bval: element boundaries (nbval = Nel+1); lval/mval:
 * angular shells; nquad: radial rule size (0 -> 5*nnodes, main.cpp:265).
 * ctx may be NULL: the handle then supports the host-side setup
 * (compute_tei, caches, one-electron matrices) but not J/K. */
int hf_atomic_create(hf_ctx *ctx, int Z, int nnodes, int nquad, const double *bval, int nbval,
                     const int *lval, const int *mval, int nang, hf_atomic **out);
int hf_atomic_destroy(hf_atomic *h);
/* A second handle on the same basis and context, with a copy of the host-side
 * setup (two-electron caches included) and device tables of its own, so two
 * densities of one basis (the alpha and beta spins of an unrestricted SCF)
 * can sit in one batch.  No counterpart in the reference (a TwoDBasis copy). */
int hf_atomic_clone(const hf_atomic *h, hf_atomic **out);

/* Sizes: out[0]=Nbf, out[1]=Nrad, out[2]=Nang, out[3]=Nel, out[4]=N_L. */
int hf_atomic_sizes(const hf_atomic *h, int64_t out[5]);
/* First radial index and size of every element (FEMRadialBasisT::get_idx,
 * libhelfem/include/RadialBasis.h:128). */
int hf_atomic_elements(const hf_atomic *h, int *first, int *nprim);
/* Algorithmic work of the most recent exchange build, which depends on the
 * density through the block-norm screen (TwoDBasis.cpp:949-952): out[0] =
 * flops of the angular (Gaunt-weighted density) sums over full Nrad x Nrad
 * blocks, out[1] = flops of the radial RI-K contraction, out[2] = number of
 * (output block, L, density block) terms, out[3] = number of coupled output
 * blocks.  Counted as the reference's TwoDBasis.cpp:918-995 performs them;
 * all zero before the first exchange call. */
int hf_atomic_work(hf_atomic *h, double out[4]);

/* Build the per-(L, element) factor caches on the host and upload them:
 * r^L and r^-L-1 element matrices and the pivoted Cholesky factor of the
 * in-element Slater tensor.  Replaces TwoDBasisT::compute_tei
 * (src/atomic/TwoDBasis.h:181, TwoDBasis.cpp:709-735). */
int hf_atomic_compute_tei(hf_atomic *h, int exchange);

/* The same caches with the in-element Slater tensors (Lobatto order
 * refinement of the split-at-r=r' quadrature) and their pivoted Cholesky
 * factors built on the GPU; the r^L / r^-L-1 element matrices stay on the
 * host.  Replaces TwoDBasisT::compute_tei like hf_atomic_compute_tei
 * (RadialBasis.cpp:644-716 for the tensor and factor); needs a context;
 * synchronises its stream.  seconds (may be NULL) gets the wall time. */
int hf_atomic_compute_tei_device(hf_atomic *h, double *seconds);

/* Supply one cached factor computed elsewhere (e.g. by HelFEM's own
 * compute_tei), instead of hf_atomic_compute_tei.  kind 0: disjoint_L
 * (r^L, Ni x Ni), 1: disjoint_m1L (r^-L-1, Ni x Ni, ignored for iel 0),
 * 2: prim_chol (Ni^2 x rank).  Caches are indexed like TwoDBasis.h:83-104.
 * After every (kind, L, iel) has been set, call hf_atomic_commit_tei. */
int hf_atomic_set_cache(hf_atomic *h, int kind, int L, int iel, int rows, int cols, const double *data);
int hf_atomic_commit_tei(hf_atomic *h);
/* Read a cached factor back.  dims gets (rows, cols); data may be NULL. */
int hf_atomic_get_cache(const hf_atomic *h, int kind, int L, int iel, int dims[2], double *data);

/* One-electron matrices (Nbf x Nbf): which 0 overlap (TwoDBasis.cpp:321),
 * 1 kinetic (:354), 2 nuclear attraction (:377). */
int hf_atomic_one_electron(const hf_atomic *h, int which, double *out, int64_t ld);
/* The same into a device matrix on the handle's stream (needs a context):
 * the radial factors (Nrad x Nrad) are built on the host and sent once per
 * handle and kind, then every call expands them into the block-diagonal
 * matrix on the device. */
int hf_atomic_one_electron_dev(const hf_atomic *h, int which, double *dOut, int64_t ld);
/* Build and send those radial factors for all three kinds now (host work
 * that callers preparing many bases can spread over threads, one handle per
 * thread); synchronises the handle's stream.  No counterpart in the
 * reference, which assembles the matrices on the host when asked. */
int hf_atomic_one_electron_prepare(hf_atomic *h);

/* Range-separated two-electron caches: the Yukawa kernel exp(-lambda r12)/r12
 * (TwoDBasisT::compute_yukawa, TwoDBasis.h:183, TwoDBasis.cpp:738-760) or
 * the erfc kernel erfc(mu r12)/r12 (compute_erfc, TwoDBasis.h:185,
 * TwoDBasis.cpp:762-771).  The latest call decides the kernel. */
int hf_atomic_compute_yukawa(hf_atomic *h, double lambda);
int hf_atomic_compute_erfc(hf_atomic *h, double mu);
/* Short-range exchange -K_sr(P) of the spin density P with that kernel.
 * Replaces TwoDBasisT::rs_exchange (TwoDBasis.h:227, TwoDBasis.cpp:1002-1131;
 * erfc element pairs as CoulombExchangeFE.h:412-440).  HF_ESTATE before
 * either compute call. */
int hf_atomic_rs_exchange(hf_atomic *h, const double *P, int64_t ldp, double *K, int64_t ldk);
int hf_atomic_rs_exchange_dev(hf_atomic *h, const double *dP, int64_t ldp, double *dK, int64_t ldk);

/* Coulomb matrix J(P) for the total density P.  Replaces
 * TwoDBasisT::coulomb (TwoDBasis.h:223, TwoDBasis.cpp:774-877). */
int hf_atomic_coulomb(hf_atomic *h, const double *P, int64_t ldp, double *J, int64_t ldj);
/* Exchange contribution for the spin density P, returned with the sign the
 * reference gives it (-K, to be added to the Fock matrix).  Replaces
 * TwoDBasisT::exchange (TwoDBasis.h:225, TwoDBasis.cpp:880-999). */
int hf_atomic_exchange(hf_atomic *h, const double *P, int64_t ldp, double *K, int64_t ldk);
/* Device-pointer forms of the two above, asynchronous on the context stream. */
int hf_atomic_coulomb_dev(hf_atomic *h, const double *dP, int64_t ldp, double *dJ, int64_t ldj);
int hf_atomic_exchange_dev(hf_atomic *h, const double *dP, int64_t ldp, double *dK, int64_t ldk);
/* Allocation-free form of hf_atomic_coulomb_dev.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops).  Its work space (the in-element J
 * bands and multipole traces, n_lm * Nel * 260 doubles) comes from a
 * 16 MiB array every GPU holds from module load to process end, so the call
 * makes no allocation; a basis needing more returns HF_EINVAL.  The handle's
 * tables must already be on the device (one ordinary J or K call uploads
 * them; HF_ESTATE before).  All calls on one GPU share that array: calls that
 * overlap on one GPU (two streams or host threads) race. */
int hf_atomic_coulomb_dev_noalloc(hf_atomic *h, const double *dP, int64_t ldp, double *dJ, int64_t ldj);
/* Reduced-precision form of hf_atomic_exchange_dev for the early iterations
 * of a mixed-precision SCF: the density, the Gaunt weights and the radial
 * factors are rounded to FP32, gathered and multiplied on the FP32 matrix
 * cores, and the partial tiles summed in FP64; the result is FP64 with a
 * relative error of about 1e-6.  Same screening and output as the FP64 form
 * of TwoDBasisT::exchange (TwoDBasis.cpp:880-999). */
int hf_atomic_exchange_fp32_dev(hf_atomic *h, const double *dP, int64_t ldp, double *dK, int64_t ldk);

/* ----------------------------------------------------------- atomic batch */

/* Many atomic bases whose Fock builds run together: each kernel of the
 * Coulomb and of the exchange build is one launch over every system (the
 * systems' workgroups concatenated, a device table of per-system arguments),
 * and the launch sequence after the density screen is captured in a HIP
 * graph and replayed while the grid sizes stay the same (graphs need a
 * context stream other than the legacy null stream).  Batched forms of
 * TwoDBasisT::coulomb / exchange (src/atomic/TwoDBasis.h:223-225,
 * TwoDBasis.cpp:774-999); the reference builds one system at a time and has
 * no counterpart of the batch itself.  The handles must share ctx, have
 * their two-electron caches computed, stay alive while the batch exists and
 * not be used by other calls concurrently. */
typedef struct hf_atomic_batch hf_atomic_batch;
int hf_atomic_batch_create(hf_ctx *ctx, hf_atomic *const *handles, int n, hf_atomic_batch **out);
int hf_atomic_batch_destroy(hf_atomic_batch *b);
/* use_graphs: capture and replay the launch sequences (default 1);
 * autotune: time the 1-, 2- and 4-wave workgroup shapes of the exchange
 * kernels and keep the fastest (default 1), once per size bucket (log2 of
 * the launch's workgroup count), class and device for the whole process, so
 * batches whose composition changes reuse the choice; 0 fixes the 4-wave
 * shape. */
int hf_atomic_batch_set_options(hf_atomic_batch *b, int use_graphs, int autotune);
/* The autotuner's state: nw[c*2 + p] waves per workgroup the last call used
 * for class c (fp32 * 2 + tiled density) and part p (0 across elements, 1
 * inside), and ms[(c*2 + p)*3 + q] the launch times this batch last measured
 * for 1, 2, 4 waves (zero when every shape it needed was already tuned). */
int hf_atomic_batch_tuning(const hf_atomic_batch *b, int nw[8], double ms[24]);
/* J_s = J(P_s) for every system s (device pointers, asynchronous on the
 * context stream).  Arrays of n pointers and leading dimensions. */
int hf_atomic_batch_coulomb_dev(hf_atomic_batch *b, const double *const *dP, const int64_t *ldp, double *const *dJ,
                                const int64_t *ldj);
/* K_s = -K(P_s) for every system s, fp32[s] != 0 selecting the
 * reduced-precision form for that system (fp32 may be NULL: all FP64).  The
 * density screen of every system is decided on the host from one batched
 * statistics launch: the call returns after one stream synchronisation,
 * with the exchange launches queued. */
int hf_atomic_batch_exchange_dev(hf_atomic_batch *b, const double *const *dP, const int64_t *ldp, double *const *dK,
                                 const int64_t *ldk, const int *fp32);

/* ------------------------------------------------------- atomic DFT grid */

/* DFT quadrature of an atomic basis: the basis' radial Chebyshev points in
 * every element x `lang` cos(theta) Chebyshev nodes x `mang` even phi points.
 * Replaces the constructor helfem::atomic::dftgrid::DFTGrid(basis, lang,
 * mang) (src/atomic/dftgrid.h:140-173; the driver's defaults are lang =
 * 4 lmax + 12, mang = 4 mmax + 12, atomic/main.cpp:329-330).  The grid keeps
 * a pointer to h: destroy it before the basis. */
int hf_atomic_dftgrid_create(hf_atomic *h, int lang, int mang, hf_dftgrid **out);
/* The same grid sizes on another rule: rule 0 is the reference's (above);
 * rule 1 takes Gauss-Legendre nodes in cos(theta) and Gauss-Lobatto radial
 * points per element (the same counts), with uniform phi.  No counterpart in
 * the reference, which has only rule 0 (src/general/angular.cpp:63-70,
 * src/atomic/dftgrid.cpp:470-576). */
int hf_atomic_dftgrid_create_rule(hf_atomic *h, int lang, int mang, int rule, hf_dftgrid **out);
/* Quadrature rules on [-1, 1], nodes increasing: kind 0 the Chebyshev rule of
 * the radial and cos(theta) grids (libhelfem/include/chebyshev.h:47-50),
 * 1 Gauss-Lobatto (lobatto.h), 2 Gauss-Legendre. */
int hf_quadrature(int kind, int n, double *x, double *w);
int hf_dftgrid_destroy(hf_dftgrid *g);
/* Pure-m DFT quadrature of a diatomic basis: the mu Chebyshev points of every
 * element x `lang` nu Chebyshev nodes, the phi integral done analytically
 * (2 pi), for densities without m mixing.  Replaces the constructor
 * helfem::diatomic::dftgrid_purem::PureMDFTGrid(basis, lang)
 * (src/diatomic/dftgrid_purem.h; the driver's default is lang = 4 lmax + 12,
 * diatomic/main.cpp:316).  Matrices are in the pure basis.  Destroy the grid
 * before the basis. */
int hf_diatomic_dftgrid_create(hf_diatomic *h, int lang, hf_dftgrid **out);
/* General 3D DFT quadrature of a diatomic basis: the mu points of every
 * element x the compound (cos nu, phi) rule of `lang` Chebyshev x `mang`
 * uniform points, all shell pairs (densities may mix m).  Replaces the
 * constructor helfem::diatomic::dftgrid::DFTGrid(basis, lang, mang)
 * (src/diatomic/dftgrid.h, dftgrid.cpp:523-528; the driver uses it when the
 * pure-m path is off, diatomic/main.cpp:326-330, defaults lang = 4 lmax + 12,
 * mang = 4 mmax + 12).  Evaluated with hf_dftgrid_eval_fxc[_pol][_dev]
 * (dftgrid.cpp:533-620). */
int hf_diatomic_dftgrid3d_create(hf_diatomic *h, int lang, int mang, hf_dftgrid **out);
/* Restricted exchange-correlation matrix H (Nbf x Nbf, overwritten) of the
 * total density P, with out[0] = Exc, out[1] = Nel = integral of the
 * density, out[2] = Ekin = integral of the kinetic energy density tau (0
 * without a meta-GGA).  x_func / c_func are libxc functional ids: exchange
 * 0 (none), 1 (lda_x), 101 (gga_x_pbe), 106 (gga_x_b88), 202 (mgga_x_tpss)
 * and the hybrids 402 (hyb_gga_xc_b3lyp), 406 (hyb_gga_xc_pbeh, PBE0), 433
 * (hyb_gga_xc_cam_b3lyp) and 470 (hyb_gga_xc_camy_b3lyp), whose
 * density-functional part this evaluates (their exact-exchange fractions
 * come from hf_xc_exact_exchange); correlation 0, 7 (lda_c_vwn,
 * VWN5), 8 (lda_c_vwn_rpa), 12 (lda_c_pw), 13 (lda_c_pw_mod), 130
 * (gga_c_pbe), 131 (gga_c_lyp), 231 (mgga_c_tpss), on every grid; others
 * return HF_EINVAL.  Densities below thr are skipped.  Replaces DFTGrid::eval_Fxc
 * (src/atomic/dftgrid.h:156, dftgrid.cpp:591-653, with the GGA gradient
 * terms of dftgrid.cpp:75-86, 320-334 and the meta-GGA tau terms of
 * :89-107, 336-349), PureMDFTGrid::eval_Fxc
 * (src/diatomic/dftgrid_purem.h:154, dftgrid_purem.cpp:675-703, GGA terms
 * :248-260, 468-500) and the 3D diatomic DFTGrid::eval_Fxc for functionals
 * without external parameters. */
int hf_dftgrid_eval_fxc(hf_dftgrid *g, int x_func, int c_func, const double *P, int64_t ldp, double *H,
                        int64_t ldh, double thr, double out[3]);
/* Spin-polarized form: densities Pa, Pb in, potentials Ha, Hb out (Hb zeroed
 * when beta == 0, i.e. no beta electrons), out as above.  Replaces the
 * unrestricted DFTGrid::eval_Fxc (src/atomic/dftgrid.h:159,
 * dftgrid.cpp:655-712) and PureMDFTGrid::eval_Fxc (dftgrid_purem.h:158,
 * dftgrid_purem.cpp:705-740). */
int hf_dftgrid_eval_fxc_pol(hf_dftgrid *g, int x_func, int c_func, const double *Pa, int64_t ldpa, const double *Pb,
                            int64_t ldpb, double *Ha, int64_t ldha, double *Hb, int64_t ldhb, int beta, double thr,
                            double out[3]);
int hf_dftgrid_eval_fxc_pol_dev(hf_dftgrid *g, int x_func, int c_func, const double *dPa, int64_t ldpa,
                                const double *dPb, int64_t ldpb, double *dHa, int64_t ldha, double *dHb, int64_t ldhb,
                                int beta, double thr, double out[3]);
/* Device-pointer form; synchronises the context stream to return out. */
int hf_dftgrid_eval_fxc_dev(hf_dftgrid *g, int x_func, int c_func, const double *dP, int64_t ldp, double *dH,
                            int64_t ldh, double thr, double out[3]);

/* Exact-exchange fractions of functional x_func: out[0] = kfrac (full
 * range), out[1] = kshort (short range), out[2] = omega (range separation),
 * out[3] = kernel of the short-range part (0 none, 1 erfc, 2 Yukawa):
 * 0.20/0/0/0 for hyb_gga_xc_b3lyp, 0.25/0/0/0 for hyb_gga_xc_pbeh,
 * 0.65/-0.46/0.33/1 for hyb_gga_xc_cam_b3lyp, 0.65/-0.46/0.34/2 for
 * hyb_gga_xc_camy_b3lyp, zeros for the pure functionals.  The Fock matrix
 * then takes kfrac * exchange(P) + kshort * rs_exchange(P), the latter after
 * hf_atomic_compute_erfc / _compute_yukawa(omega).  Replaces
 * range_separation and is_range_separated (src/general/dftfuncs.cpp:470,
 * 511, used at src/atomic/main.cpp:237-243, 336-337, 424-437). */
int hf_xc_exact_exchange(int x_func, double out[4]);

/* ------------------------------------------------- batched eigensolver */
/* Eigenvalues (ascending, dW[b*n + i]) and eigenvectors (overwriting each
 * column-major A_b = dA + b*strideA with leading dimension lda) of a batch of
 * symmetric n x n matrices (lower triangles read), on the context stream
 * with rocSOLVER's dsyevd_strided_batched; dE is n*batch doubles of
 * workspace, dinfo batch ints (0: converged).  Asynchronous; the SCF drivers
 * stack their Fock blocks into one call.  Replaces the per-block host
 * eigensolves of eig_gsym / eig_gsym_sub (src/general/scf_helpers.cpp:41-80),
 * applied to the orthogonalised blocks. */
int hf_eigh_batched(hf_ctx *ctx, int n, double *dA, int64_t lda, int64_t strideA, double *dW, double *dE,
                    int *dinfo, int batch);

/* --------------------------------------------------------- diatomic basis */

/* Prolate spheroidal basis: radial finite elements in mu x Y_lm(nu, phi)
 * for nuclei Z1 at z = -Rhalf and Z2 at z = +Rhalf.  Replaces the
 * constructor helfem::diatomic::basis::TwoDBasis::TwoDBasis
 * (src/diatomic/basis.h:222) with LIP shape functions on `nnodes`
 * Gauss-Lobatto nodes (primbas 4); the mu = 0 function is kept and the
 * mu = mu_max function dropped.  bval: element boundaries in mu (nbval =
 * Nel+1, bval[0] = 0); lval/mval: shells in lm_to_l_m order
 * (basis.cpp:500-516).  Matrices are in the pure basis, which drops radial
 * function 0 of every m != 0 shell (pure_indices, basis.h:250).  ctx may be
 * NULL: the handle then supports the host-side setup but not J/K. */
int hf_diatomic_create(hf_ctx *ctx, int Z1, int Z2, double Rhalf, int nnodes, int nquad, const double *bval,
                       int nbval, const int *lval, const int *mval, int nang, hf_diatomic **out);
int hf_diatomic_destroy(hf_diatomic *h);
/* out[0]=Nbf (pure), out[1]=Ndummy, out[2]=Nrad, out[3]=Nang, out[4]=Nel,
 * out[5]=size of the (L,|M|) map, out[6]=size of the (L,M) map. */
int hf_diatomic_sizes(const hf_diatomic *h, int64_t out[7]);
/* Dummy index (shell*Nrad + radial) of every pure function, Nbf entries
 * (TwoDBasis::pure_indices, basis.h:250). */
int hf_diatomic_pure_index(const hf_diatomic *h, int64_t *idx);
/* (L,|M|) map (2 ints per entry, |M|-major) and (L,M) map (L-major); either
 * pointer may be NULL (lm_map / LM_map, basis.cpp:612-647). */
int hf_diatomic_maps(const hf_diatomic *h, int *lm, int *LM);
/* With sym != 0 the density is taken to be symmetric under m -> -m: the
 * exchange builds only output blocks with mj, mk >= 0 and mirrors
 * K(-mj,-mk) = K(mj,mk).  Replaces TwoDBasis::set_absm_symmetric
 * (basis.h:344). */
int hf_diatomic_set_absm_symmetric(hf_diatomic *h, int sym);
/* Build the per-((L,|M|), element) caches on the host: disjoint integrals
 * P0, P2, Q0, Q2 and the sign-aware Cholesky factor of the in-element
 * two-channel kernel.  Replaces TwoDBasis::compute_tei (basis.h:265,
 * basis.cpp:1382-1547). */
int hf_diatomic_compute_tei(hf_diatomic *h, int exchange);
/* The same caches with the two-channel in-element kernels (the Chebyshev
 * quadrature of quadrature.cpp:162-258, symmetrised as basis.cpp:1470-1490)
 * and their signed pivoted Cholesky factors (basis.cpp:1492-1537) built on
 * the GPU; the P/Q integrals, the per-element rule sizes and the Q values at
 * the outer nodes stay on the host.  Replaces TwoDBasis::compute_tei like
 * hf_diatomic_compute_tei; needs a context; synchronises its stream.
 * seconds (may be NULL) gets the wall time. */
int hf_diatomic_compute_tei_device(hf_diatomic *h, double *seconds);
/* Read a cache back: kind 0 P0, 1 P2, 2 Q0, 3 Q2 (Ni x Ni), 4 the Cholesky
 * vectors B (2 Ni^2 x rank), 5 their signs (rank x 1).  dims gets
 * (rows, cols); data may be NULL. */
int hf_diatomic_get_cache(const hf_diatomic *h, int kind, int ilm, int iel, int dims[2], double *data);
/* Pure-basis one-electron matrices: which 0 overlap, 1 kinetic, 2 nuclear
 * attraction (basis.h:280-284, basis.cpp:1032-1156). */
int hf_diatomic_one_electron(const hf_diatomic *h, int which, double *out, int64_t ld);
/* The same into a device matrix on the handle's stream (needs a context):
 * only the nonzero shell-pair blocks are built on the host and sent. */
int hf_diatomic_one_electron_dev(const hf_diatomic *h, int which, double *dOut, int64_t ld);
/* Work of the most recent exchange build, as hf_atomic_work
 * (basis.cpp:1818-2089). */
int hf_diatomic_work(hf_diatomic *h, double out[4]);
/* Coulomb matrix of the total density, pure basis.  Replaces
 * TwoDBasis::coulomb (basis.h:300, basis.cpp:1627-1816). */
int hf_diatomic_coulomb(hf_diatomic *h, const double *P, int64_t ldp, double *J, int64_t ldj);
/* Exchange contribution (-K) of the spin density, pure basis.  Replaces
 * TwoDBasis::exchange (basis.h:302, basis.cpp:1818-2089). */
int hf_diatomic_exchange(hf_diatomic *h, const double *P, int64_t ldp, double *K, int64_t ldk);
/* Device-pointer forms, asynchronous on the context stream. */
int hf_diatomic_coulomb_dev(hf_diatomic *h, const double *dP, int64_t ldp, double *dJ, int64_t ldj);
int hf_diatomic_exchange_dev(hf_diatomic *h, const double *dP, int64_t ldp, double *dK, int64_t ldk);
/* Reduced-precision form, as hf_atomic_exchange_fp32_dev, of
 * TwoDBasis::exchange (basis.cpp:1818-2089). */
int hf_diatomic_exchange_fp32_dev(hf_diatomic *h, const double *dP, int64_t ldp, double *dK, int64_t ldk);

#ifdef __cplusplus
}
#endif

#endif /* HFGPU_H */
