"""Python face of the CPU oracle (test infrastructure only).

Loads oracle/_build/liboracle.so (built by oracle/Makefile) and exposes the
restated HelFEM atomic basis plus a plain DIIS SCF driver.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this.

The SCF here stands in for OpenOrbitalOptimizer (fetched at a3abae9 by the
reference, CMakeLists.txt:310-341, absent here).  It follows the Fock-builder
of src/atomic/main.cpp:369-470: per-symmetry-block Sinvh (scf_driver_common.h
:176-187, utils.cpp:120-159), P = sum_k Sinvh_k C diag(occ) C' Sinvh_k'
(scf_driver_common.h:407-425), Etot = tr(PT)+tr(PV)+1/2 tr(PJ)+Exc+Exx with
Exx = 1/2 tr(Pa Ka) doubled when restricted (:460-474).  Converged energies do
not depend on the iteration algorithm, only on reaching the same state.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "_build" / "liboracle.so"
        if not path.exists():
            raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
        L = ctypes.CDLL(str(path))
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_gaunt.restype = ctypes.c_double
        L.orc_mod_gaunt.restype = ctypes.c_double
        L.orc_bessel_il.restype = ctypes.c_double
        L.orc_bessel_il.argtypes = [ctypes.c_double, ctypes.c_int]
        L.orc_bessel_kl.restype = ctypes.c_double
        L.orc_bessel_kl.argtypes = [ctypes.c_double, ctypes.c_int]
        L.orc_erfc_phi.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
        L.orc_atomic_compute_yukawa.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_atomic_compute_erfc.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_atomic_rs_exchange.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_atomic_rs_cache.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_void_p]
        L.orc_atomic_create.restype = ctypes.c_void_p
        L.orc_atomic_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        for name in ("orc_atomic_destroy", "orc_atomic_sizes", "orc_atomic_elements", "orc_atomic_compute_tei",
                     "orc_atomic_matrix", "orc_atomic_coulomb", "orc_atomic_exchange", "orc_atomic_cache",
                     "orc_atomic_twoe", "orc_atomic_radial_integral", "orc_atomic_radial_bf",
                     "orc_atomic_exchange_blocks", "orc_legendre", "orc_diatomic_destroy",
                     "orc_diatomic_sizes", "orc_diatomic_pure_idx", "orc_diatomic_compute_tei",
                     "orc_diatomic_matrix", "orc_diatomic_coulomb", "orc_diatomic_exchange",
                     "orc_diatomic_maps", "orc_diatomic_cache", "orc_diatomic_exchange_blocks", "orc_atomic_xc"):
            getattr(L, name).argtypes = None
        L.orc_lda.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_double,
                              ctypes.c_void_p, ctypes.c_void_p]
        L.orc_xc_exact_exchange.argtypes = [ctypes.c_int, ctypes.c_void_p]
        L.orc_xc_eval.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 3 + [ctypes.c_double] + \
            [ctypes.c_void_p] * 4
        L.orc_atomic_xc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_lda_pol.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_atomic_xc_pol.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p]
        L.orc_diatomic_xc_pol.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p]
        L.orc_diatomic_xc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_diatomic_xc3d.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_diatomic_xc3d_pol.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p]
        L.orc_diatomic_create.restype = ctypes.c_void_p
        L.orc_diatomic_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_void_p]
        _LIB = L
    return _LIB


def set_threads(n):
    """OpenMP threads of the oracle's loops."""
    _check(lib().orc_set_threads(int(n)))


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


def gaunt(L, M, l, m, lp, mp):
    return lib().orc_gaunt(L, M, l, m, lp, mp)


def mod_gaunt(lj, mj, L, M, li, mi):
    return lib().orc_mod_gaunt(lj, mj, L, M, li, mi)


def lobatto(n):
    x = np.zeros(n)
    w = np.zeros(n)
    _check(lib().orc_lobatto(n, _ptr(x), _ptr(w)))
    return x, w


def chebyshev(n):
    x = np.zeros(n)
    w = np.zeros(n)
    _check(lib().orc_chebyshev(n, _ptr(x), _ptr(w)))
    return x, w


def lda(rho, x_func=1, c_func=7, thr=1e-12):
    """lda_x + lda_c_vwn (VWN5), spin-unpolarized: (exc per particle, vrho)."""
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    exc = np.zeros_like(rho)
    v = np.zeros_like(rho)
    _check(lib().orc_lda(x_func, c_func, rho.size, _ptr(rho), thr, _ptr(exc), _ptr(v)))
    return exc, v


def lda_pol(ra, rb, x_func=1, c_func=7, thr=1e-12):
    """Spin-polarized lda_x + lda_c_vwn: (exc per particle, v_alpha, v_beta)."""
    ra = np.ascontiguousarray(ra, dtype=np.float64)
    rb = np.ascontiguousarray(rb, dtype=np.float64)
    exc, va, vb = np.zeros_like(ra), np.zeros_like(ra), np.zeros_like(ra)
    _check(lib().orc_lda_pol(x_func, c_func, ra.size, _ptr(ra), _ptr(rb), thr, _ptr(exc), _ptr(va), _ptr(vb)))
    return exc, va, vb


FUNCTIONALS = {"lda_x": 1, "lda_c_vwn": 7, "lda_c_pw": 12, "lda_c_pw_mod": 13, "gga_x_pbe": 101,
               "gga_c_pbe": 130, "mgga_x_tpss": 202, "mgga_c_tpss": 231}


def exact_exchange(x_func):
    """(kfrac, kshort, omega, kernel) of a hybrid functional id: the full-range
    and short-range exact-exchange fractions, the range-separation parameter
    and the short-range kernel, "erfc", "yukawa" or None (range_separation,
    is_range_separated, src/general/dftfuncs.cpp:470-575); zeros for pure
    functionals."""
    out = np.zeros(4)
    _check(lib().orc_xc_exact_exchange(int(x_func), _ptr(out)))
    return float(out[0]), float(out[1]), float(out[2]), {0: None, 1: "erfc", 2: "yukawa"}[int(out[3])]


def xc_eval(rho, sigma, x_func, c_func, thr=1e-12, tau=None):
    """Functional values at points: unpolarized rho (n,), sigma (n,), tau (n,)
    -> (exc, vrho, vsigma); polarized rho (2, n) [a, b], sigma (3, n) [aa, ab,
    bb], tau (2, n) -> (exc (n,), vrho (2, n), vsigma (3, n)).  Energies per
    particle, potentials d(rho exc)/drho and d(rho exc)/dsigma (complex-step
    derivatives).  With tau given (meta-GGA), vtau = d(rho exc)/dtau is
    returned fourth."""
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    sigma = np.ascontiguousarray(sigma, dtype=np.float64)
    pol = rho.ndim == 2
    n = rho.shape[-1]
    exc = np.zeros(n)
    vr = np.zeros_like(rho)
    vs = np.zeros_like(sigma)
    t = None if tau is None else np.ascontiguousarray(tau, dtype=np.float64)
    vt = np.zeros_like(rho)
    _check(lib().orc_xc_eval(x_func, c_func, int(pol), n, _ptr(rho), _ptr(sigma), None if t is None else _ptr(t), thr,
                             _ptr(exc), _ptr(vr), _ptr(vs), _ptr(vt)))
    return (exc, vr, vs) if tau is None else (exc, vr, vs, vt)


def grid(rmax, nel, igrid=4, zexp=2.0):
    out = np.zeros(nel + 1)
    _check(lib().orc_grid(ctypes.c_double(rmax), nel, igrid, ctypes.c_double(zexp), _ptr(out)))
    return out


def lip_eval(nnodes, x, n=0):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros((nnodes, x.size)).T.copy(order="F")
    out = np.zeros(x.size * nnodes)
    _check(lib().orc_lip_eval(nnodes, x.size, _ptr(x), n, _ptr(out)))
    return out.reshape((nnodes, x.size)).T


def twoe_rule(rmin, rmax, nquad, nnodes, drop_first, drop_last, L):
    nb = nnodes - int(drop_first) - int(drop_last)
    out = np.zeros(nb ** 4)
    _check(lib().orc_twoe_rule(ctypes.c_double(rmin), ctypes.c_double(rmax), nquad, nnodes, int(drop_first),
                               int(drop_last), L, _ptr(out)))
    return out.reshape((nb * nb, nb * nb), order="F")


def angular_basis(lmax, mmax):
    """src/atomic/basis.cpp:179-203"""
    lval, mval = [], []
    for mabs in range(mmax + 1):
        for l in range(mabs, lmax + 1):
            lval.append(l)
            mval.append(mabs)
            if mabs > 0:
                lval.append(l)
                mval.append(-mabs)
    return np.array(lval, dtype=np.int32), np.array(mval, dtype=np.int32)


class AtomicBasis:
    """Oracle restatement of helfem::atomic::basis::TwoDBasis (TwoDBasis.h:52-283)."""

    def __init__(self, Z, lmax, mmax, nelem, rmax=40.0, igrid=4, zexp=2.0, nnodes=15, nquad=None,
                 lval=None, mval=None, bval=None):
        self.Z = Z
        if lval is None:
            lval, mval = angular_basis(lmax, mmax)
        self.lval = np.ascontiguousarray(lval, dtype=np.int32)
        self.mval = np.ascontiguousarray(mval, dtype=np.int32)
        self.bval = grid(rmax, nelem, igrid, zexp) if bval is None else np.ascontiguousarray(bval, np.float64)
        self.nnodes = nnodes
        self.nquad = 5 * nnodes if nquad is None else nquad
        h = lib().orc_atomic_create(Z, nnodes, self.nquad, self.bval.size, _ptr(self.bval), self.lval.size,
                                    _ptr(self.lval), _ptr(self.mval))
        if not h:
            raise RuntimeError(lib().orc_last_error().decode())
        self._h = ctypes.c_void_p(h)
        s = (ctypes.c_long * 5)()
        lib().orc_atomic_sizes(self._h, s)
        self.Nbf, self.Nrad, self.Nel, self.N_L, self.Nang = (int(v) for v in s)
        first = np.zeros(self.Nel, np.int32)
        nprim = np.zeros(self.Nel, np.int32)
        lib().orc_atomic_elements(self._h, _ptr(first), _ptr(nprim))
        self.first, self.nprim = first, nprim
        self._tei = False

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _LIB is not None:
            _LIB.orc_atomic_destroy(h)

    def compute_tei(self):
        _check(lib().orc_atomic_compute_tei(self._h))
        self._tei = True

    def _matrix(self, which):
        out = np.zeros((self.Nbf, self.Nbf), order="F")
        _check(lib().orc_atomic_matrix(self._h, which, _ptr(out)))
        return out

    def overlap(self):
        return self._matrix(0)

    def kinetic(self):
        return self._matrix(1)

    def nuclear(self):
        return self._matrix(2)

    def coulomb(self, P):
        P = np.asfortranarray(P, dtype=np.float64)
        J = np.zeros((self.Nbf, self.Nbf), order="F")
        _check(lib().orc_atomic_coulomb(self._h, _ptr(P), _ptr(J)))
        return J

    def exchange(self, P):
        P = np.asfortranarray(P, dtype=np.float64)
        K = np.zeros((self.Nbf, self.Nbf), order="F")
        _check(lib().orc_atomic_exchange(self._h, _ptr(P), _ptr(K)))
        return K

    def compute_yukawa(self, lam):
        """Yukawa range-separation caches (TwoDBasis.cpp:738-760)."""
        _check(lib().orc_atomic_compute_yukawa(self._h, float(lam)))

    def compute_erfc(self, mu):
        """erfc range-separation caches (TwoDBasis.cpp:762-771)."""
        _check(lib().orc_atomic_compute_erfc(self._h, float(mu)))

    def rs_exchange(self, P):
        """-K of the short-range kernel (TwoDBasis.cpp:1002-1131)."""
        P = np.asfortranarray(P, dtype=np.float64)
        K = np.zeros((self.Nbf, self.Nbf), order="F")
        _check(lib().orc_atomic_rs_exchange(self._h, _ptr(P), _ptr(K)))
        return K

    def rs_cache(self, kind, index):
        dims = np.zeros(2, dtype=np.int32)
        _check(lib().orc_atomic_rs_cache(self._h, kind, index, dims.ctypes.data, None))
        out = np.zeros((int(dims[0]), int(dims[1])), order="F")
        if out.size:
            _check(lib().orc_atomic_rs_cache(self._h, kind, index, dims.ctypes.data, _ptr(out)))
        return out

    def exchange_blocks(self, P, jk):
        """-K restricted to the output blocks j*Nang+k listed in jk."""
        P = np.asfortranarray(P, dtype=np.float64)
        jk = np.ascontiguousarray(jk, dtype=np.int32)
        K = np.zeros((self.Nbf, self.Nbf), order="F")
        _check(lib().orc_atomic_exchange_blocks(self._h, _ptr(P), _ptr(K), jk.size, _ptr(jk)))
        return K

    def xc(self, P, lang=None, mang=None, x_func=1, c_func=7, thr=1e-12):
        """DFTGrid::eval_Fxc (src/atomic/dftgrid.cpp:591-653), LDA, restricted.

        Grid defaults ldft = 4 lmax + 12, mdft = 4 mmax + 12 (atomic/main.cpp:329-330).
        Returns (Vxc, Exc, Nel)."""
        lmax, mmax = int(np.max(self.lval)), int(np.max(np.abs(self.mval)))
        lang = 4 * lmax + 12 if lang is None else lang
        mang = 4 * mmax + 12 if mang is None else mang
        P = np.asfortranarray(P, dtype=np.float64)
        H = np.zeros((self.Nbf, self.Nbf), order="F")
        out = np.zeros(3)
        _check(lib().orc_atomic_xc(self._h, lang, mang, x_func, c_func, _ptr(P), thr, _ptr(H), _ptr(out)))
        self.last_ekin = float(out[2])
        return H, float(out[0]), float(out[1])

    def xc_pol(self, Pa, Pb, lang=None, mang=None, x_func=1, c_func=7, thr=1e-12):
        """Spin-polarized DFTGrid::eval_Fxc (dftgrid.cpp:655-712): (Ha, Hb, Exc, Nel)."""
        lmax, mmax = int(np.max(self.lval)), int(np.max(np.abs(self.mval)))
        lang = 4 * lmax + 12 if lang is None else lang
        mang = 4 * mmax + 12 if mang is None else mang
        Pa = np.asfortranarray(Pa, dtype=np.float64)
        Pb = np.asfortranarray(Pb, dtype=np.float64)
        Ha = np.zeros((self.Nbf, self.Nbf), order="F")
        Hb = np.zeros((self.Nbf, self.Nbf), order="F")
        out = np.zeros(3)
        _check(lib().orc_atomic_xc_pol(self._h, lang, mang, x_func, c_func, _ptr(Pa), _ptr(Pb), thr, _ptr(Ha),
                                       _ptr(Hb), _ptr(out)))
        self.last_ekin = float(out[2])
        return Ha, Hb, float(out[0]), float(out[1])

    def cache(self, kind, L, iel):
        dims = (ctypes.c_int * 2)()
        _check(lib().orc_atomic_cache(self._h, kind, L, iel, dims, None))
        out = np.zeros(dims[0] * dims[1])
        if out.size:
            _check(lib().orc_atomic_cache(self._h, kind, L, iel, dims, _ptr(out)))
        return out.reshape((dims[0], dims[1]), order="F")

    def twoe(self, L, iel):
        n = int(self.nprim[iel]) ** 2
        out = np.zeros(n * n)
        _check(lib().orc_atomic_twoe(self._h, L, iel, _ptr(out)))
        return out.reshape((n, n), order="F")

    def radial_integral(self, n, iel):
        ni = int(self.nprim[iel])
        out = np.zeros(ni * ni)
        _check(lib().orc_atomic_radial_integral(self._h, n, iel, _ptr(out)))
        return out.reshape((ni, ni), order="F")

    def radial_bf(self, iel, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        ni = int(self.nprim[iel])
        out = np.zeros(x.size * ni)
        _check(lib().orc_atomic_radial_bf(self._h, iel, x.size, _ptr(x), _ptr(out)))
        return out.reshape((x.size, ni), order="F")

    def sym_idx(self, symm=1):
        """TwoDBasis.cpp:202-228: index list per unique m (symm=1) or one block."""
        Nrad = self.Nrad
        if symm == 0:
            return [np.arange(self.Nbf)]
        blocks = []
        for m in sorted(set(self.mval.tolist())):
            idx = [a * Nrad + i for a in range(self.Nang) if self.mval[a] == m for i in range(Nrad)]
            blocks.append(np.array(idx))
        return blocks


def legendre(lmax, mmax, x):
    """P_l^m(x), Q_l^m(x) as (lmax+1, mmax+1) arrays (src/legendre/Legendre.h)."""
    P = np.zeros((mmax + 1) * (lmax + 1))
    Q = np.zeros_like(P)
    _check(lib().orc_legendre(lmax, mmax, ctypes.c_double(x), _ptr(P), _ptr(Q)))
    return P.reshape((mmax + 1, lmax + 1)).T, Q.reshape((mmax + 1, lmax + 1)).T


def diatomic_angular(lmmax):
    """lm_to_l_m, src/diatomic/basis.cpp:500-516."""
    lval, mval = [], []
    for mabs, lm in enumerate(lmmax):
        for l in range(mabs, lm + 1):
            lval.append(l)
            mval.append(mabs)
            if mabs > 0:
                lval.append(l)
                mval.append(-mabs)
    return np.array(lval, dtype=np.int32), np.array(mval, dtype=np.int32)


class DiatomicBasis:
    """Oracle restatement of helfem::diatomic::basis::TwoDBasis (basis.h:127-432).

    Rbond in bohr; mu_max = arcosh(Rmax / Rhalf); elements on the igrid=4
    grid with zexp = 1 (the diatomic driver's defaults, main.cpp:70-72)."""

    def __init__(self, Z1, Z2, Rbond, lmmax, nelem, rmax=40.0, igrid=4, zexp=1.0, nnodes=15, nquad=None):
        self.Z1, self.Z2, self.Rbond = Z1, Z2, Rbond
        self.Rhalf = 0.5 * Rbond
        self.lval, self.mval = diatomic_angular(lmmax)
        self.mumax = float(np.arccosh(max(rmax / self.Rhalf, 1.0)))
        self.bval = grid(self.mumax, nelem, igrid, zexp)
        self.nquad = 5 * nnodes if nquad is None else nquad
        h = lib().orc_diatomic_create(Z1, Z2, self.Rhalf, nnodes, self.nquad, self.bval.size, _ptr(self.bval),
                                      self.lval.size, _ptr(self.lval), _ptr(self.mval))
        if not h:
            raise RuntimeError(lib().orc_last_error().decode())
        self._h = ctypes.c_void_p(h)
        sz = (ctypes.c_long * 7)()
        lib().orc_diatomic_sizes(self._h, sz)
        self.Nbf, self.Ndummy, self.Nrad, self.Nang, self.Nel, self.n_lm, self.n_LM = (int(v) for v in sz)
        self.pure_idx = np.zeros(self.Nbf, np.int32)
        lib().orc_diatomic_pure_idx(self._h, _ptr(self.pure_idx))
        self.Enucr = Z1 * Z2 / Rbond

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _LIB is not None:
            _LIB.orc_diatomic_destroy(h)

    def compute_tei(self):
        _check(lib().orc_diatomic_compute_tei(self._h))

    def _matrix(self, which):
        out = np.zeros((self.Nbf, self.Nbf), order="F")
        _check(lib().orc_diatomic_matrix(self._h, which, _ptr(out)))
        return out

    def overlap(self):
        return self._matrix(0)

    def kinetic(self):
        return self._matrix(1)

    def nuclear(self):
        return self._matrix(2)

    def coulomb(self, P):
        P = np.asfortranarray(P, dtype=np.float64)
        J = np.zeros((self.Nbf, self.Nbf), order="F")
        _check(lib().orc_diatomic_coulomb(self._h, _ptr(P), _ptr(J)))
        return J

    def exchange(self, P):
        P = np.asfortranarray(P, dtype=np.float64)
        K = np.zeros((self.Nbf, self.Nbf), order="F")
        _check(lib().orc_diatomic_exchange(self._h, _ptr(P), _ptr(K)))
        return K

    # The reference's diatomic basis has no range-separated exchange
    # (src/diatomic/basis.h declares none); the atomic oracle entry points
    # take an atomic handle, so these must not forward to them.
    def compute_yukawa(self, lam):
        raise NotImplementedError("range-separated exchange exists for atomic bases only")

    def compute_erfc(self, mu):
        raise NotImplementedError("range-separated exchange exists for atomic bases only")

    def rs_exchange(self, P):
        raise NotImplementedError("range-separated exchange exists for atomic bases only")

    def rs_cache(self, kind, index):
        raise NotImplementedError("range-separated exchange exists for atomic bases only")

    def exchange_blocks(self, P, jk):
        """-K restricted to the output blocks j*Nang+k listed in jk."""
        P = np.asfortranarray(P, dtype=np.float64)
        jk = np.ascontiguousarray(jk, dtype=np.int32)
        K = np.zeros((self.Nbf, self.Nbf), order="F")
        _check(lib().orc_diatomic_exchange_blocks(self._h, _ptr(P), _ptr(K), jk.size, _ptr(jk)))
        return K

    def xc(self, P, lang=None, x_func=1, c_func=7, thr=1e-12):
        """PureMDFTGrid::eval_Fxc (src/diatomic/dftgrid_purem.cpp:675-703), LDA,
        restricted; lang defaults to 4 lmax + 12 (diatomic/main.cpp:316).
        Returns (Vxc, Exc, Nel) in the pure basis."""
        lang = 4 * int(np.max(self.lval)) + 12 if lang is None else lang
        P = np.asfortranarray(P, dtype=np.float64)
        H = np.zeros((self.Nbf, self.Nbf), order="F")
        out = np.zeros(3)
        _check(lib().orc_diatomic_xc(self._h, lang, x_func, c_func, _ptr(P), thr, _ptr(H), _ptr(out)))
        self.last_ekin = float(out[2])
        return H, float(out[0]), float(out[1])

    def xc_pol(self, Pa, Pb, lang=None, x_func=1, c_func=7, thr=1e-12):
        """Spin-polarized PureMDFTGrid::eval_Fxc (dftgrid_purem.cpp:705-740): (Ha, Hb, Exc, Nel)."""
        lang = 4 * int(np.max(self.lval)) + 12 if lang is None else lang
        Pa = np.asfortranarray(Pa, dtype=np.float64)
        Pb = np.asfortranarray(Pb, dtype=np.float64)
        Ha = np.zeros((self.Nbf, self.Nbf), order="F")
        Hb = np.zeros((self.Nbf, self.Nbf), order="F")
        out = np.zeros(3)
        _check(lib().orc_diatomic_xc_pol(self._h, lang, x_func, c_func, _ptr(Pa), _ptr(Pb), thr, _ptr(Ha), _ptr(Hb),
                                         _ptr(out)))
        self.last_ekin = float(out[2])
        return Ha, Hb, float(out[0]), float(out[1])

    def xc3d(self, P, lang=None, mang=None, x_func=1, c_func=7, thr=1e-12):
        """General 3D diatomic DFTGrid::eval_Fxc (src/diatomic/dftgrid.cpp:533-572): (H, Exc, Nel)."""
        lang = 4 * int(np.max(self.lval)) + 12 if lang is None else lang
        mang = 4 * int(np.max(np.abs(self.mval))) + 12 if mang is None else mang
        P = np.asfortranarray(P, dtype=np.float64)
        H = np.zeros((self.Nbf, self.Nbf), order="F")
        out = np.zeros(3)
        _check(lib().orc_diatomic_xc3d(self._h, lang, mang, x_func, c_func, _ptr(P), thr, _ptr(H), _ptr(out)))
        self.last_ekin = float(out[2])
        return H, float(out[0]), float(out[1])

    def xc3d_pol(self, Pa, Pb, lang=None, mang=None, x_func=1, c_func=7, thr=1e-12):
        """Spin-polarized 3D DFTGrid::eval_Fxc (dftgrid.cpp:574-620): (Ha, Hb, Exc, Nel)."""
        lang = 4 * int(np.max(self.lval)) + 12 if lang is None else lang
        mang = 4 * int(np.max(np.abs(self.mval))) + 12 if mang is None else mang
        Pa = np.asfortranarray(Pa, dtype=np.float64)
        Pb = np.asfortranarray(Pb, dtype=np.float64)
        Ha = np.zeros((self.Nbf, self.Nbf), order="F")
        Hb = np.zeros((self.Nbf, self.Nbf), order="F")
        out = np.zeros(3)
        _check(lib().orc_diatomic_xc3d_pol(self._h, lang, mang, x_func, c_func, _ptr(Pa), _ptr(Pb), thr, _ptr(Ha),
                                           _ptr(Hb), _ptr(out)))
        self.last_ekin = float(out[2])
        return Ha, Hb, float(out[0]), float(out[1])

    def maps(self):
        lm = np.zeros(2 * self.n_lm, np.int32)
        LM = np.zeros(2 * self.n_LM, np.int32)
        lib().orc_diatomic_maps(self._h, _ptr(lm), _ptr(LM))
        return lm.reshape(-1, 2), LM.reshape(-1, 2)

    def cache(self, kind, ilm, iel):
        dims = (ctypes.c_int * 2)()
        _check(lib().orc_diatomic_cache(self._h, kind, ilm, iel, dims, None))
        out = np.zeros(dims[0] * dims[1])
        if out.size:
            _check(lib().orc_diatomic_cache(self._h, kind, ilm, iel, dims, _ptr(out)))
        return out.reshape((dims[0], dims[1]), order="F")

    def sym_idx(self, symm=1):
        """get_sym_idx(1), basis.cpp:913-935: pure indices grouped by m."""
        if symm == 0:
            return [np.arange(self.Nbf)]
        blocks, off = {}, 0
        for a in range(self.Nang):
            n = self.Nrad if self.mval[a] == 0 else self.Nrad - 1
            blocks.setdefault(int(self.mval[a]), []).extend(range(off, off + n))
            off += n
        return [np.array(blocks[m]) for m in sorted(blocks)]


def invh(S):
    """utils.cpp:120-159 with chol=false: normalised symmetric orthonormalisation."""
    d = 1.0 / np.sqrt(np.diag(S))
    Sn = d[:, None] * S * d[None, :]
    e, v = np.linalg.eigh(Sn)
    X = (v * (1.0 / np.sqrt(e))) @ v.T
    return d[:, None] * X


class SCFResult(dict):
    pass


def scf_rhf(basis, nelec, H1=None, symm=1, conv=1e-11, maxit=200, verbose=False, jk=None, xc=None, kfrac=0.0,
            kshort=0.0):
    """Restricted closed-shell HF (or DFT) with Pulay DIIS over per-m symmetry blocks.

    jk(P, Pa) -> (J, Ka) may be supplied to swap in another Fock engine; the
    default is the oracle's own coulomb/exchange.  xc(P) -> (Vxc, Exc) turns
    the run into Kohn-Sham DFT: Etot gains Exc, and a hybrid adds the
    fraction kfrac of exact exchange, K = kfrac exchange(P) with Exx scaled
    alike, and a range-separated one kshort rs_exchange(P) on the kernel set
    by compute_erfc / compute_yukawa (atomic/main.cpp:237-243,336-337,369-470)."""
    S = basis.overlap()
    T = basis.kinetic()
    V = basis.nuclear()
    H = T + V if H1 is None else H1
    blocks = basis.sym_idx(symm)
    X = [invh(S[np.ix_(b, b)]) for b in blocks]
    nocc = nelec // 2
    if jk is None:
        def jk(P, Pa):
            return basis.coulomb(P), basis.exchange(Pa)

    def density(F):
        eps, vecs = [], []
        for k, b in enumerate(blocks):
            Fo = X[k].T @ F[np.ix_(b, b)] @ X[k]
            e, c = np.linalg.eigh(Fo)
            eps.append(e)
            vecs.append(c)
        allev = sorted((e, k, i) for k, ev in enumerate(eps) for i, e in enumerate(ev))
        occ = [np.zeros(len(ev)) for ev in eps]
        for e, k, i in allev[:nocc]:
            occ[k][i] = 2.0
        P = np.zeros_like(F)
        for k, b in enumerate(blocks):
            C = X[k] @ vecs[k]
            P[np.ix_(b, b)] += (C * occ[k]) @ C.T
        return P, occ

    P, occ = density(H)
    diis_F, diis_e = [], []
    Eold = 0.0
    it = 0
    E = 0.0
    for it in range(1, maxit + 1):
        Pa = 0.5 * P
        Ekin = np.sum(P * T)
        Enuc = np.sum(P * V)
        if xc is None:
            J, Ka = jk(P, Pa)
            Exx = 2.0 * 0.5 * np.sum(Pa * Ka)
            Exc = 0.0
            F = H + J + Ka
        else:
            J = basis.coulomb(P)
            Vxc, Exc = xc(P)
            Exx = 0.0
            F = H + J + Vxc
            if kfrac or kshort:
                Ka = kfrac * basis.exchange(Pa) if kfrac else 0.0
                if kshort:
                    Ka = Ka + kshort * basis.rs_exchange(Pa)
                Exx = np.sum(Pa * Ka)
                F = F + Ka
        Ecoul = 0.5 * np.sum(P * J)
        E = Ekin + Enuc + Ecoul + Exx + Exc
        # orthonormal-basis commutator per block
        errs = []
        for k, b in enumerate(blocks):
            Fo = X[k].T @ F[np.ix_(b, b)] @ X[k]
            Sinv = np.linalg.pinv(X[k])
            Po = Sinv @ P[np.ix_(b, b)] @ Sinv.T
            errs.append((Fo @ Po - Po @ Fo).ravel())
        err = np.concatenate(errs)
        emax = np.max(np.abs(err))
        if verbose:
            print(f"it {it:3d} E {E:.12f} dE {E - Eold:.3e} err {emax:.3e}")
        if it > 1 and abs(E - Eold) < conv and emax < 1e-6:
            break
        Eold = E
        diis_F.append(F)
        diis_e.append(err)
        if len(diis_F) > 10:
            diis_F.pop(0)
            diis_e.pop(0)
        n = len(diis_F)
        B = -np.ones((n + 1, n + 1))
        B[n, n] = 0.0
        for i in range(n):
            for j in range(n):
                B[i, j] = diis_e[i] @ diis_e[j]
        rhs = np.zeros(n + 1)
        rhs[n] = -1.0
        try:
            c = np.linalg.solve(B, rhs)[:n]
            Fd = sum(ci * Fi for ci, Fi in zip(c, diis_F))
        except np.linalg.LinAlgError:
            Fd = F
        P, occ = density(Fd)
    E += getattr(basis, "Enucr", 0.0)
    return SCFResult(E=E, Ekin=Ekin, Enuc=Enuc, Ecoul=Ecoul, Exx=Exx, Exc=Exc, iterations=it, P=P)


def scf_uhf(basis, nela, nelb, symm=1, conv=1e-11, maxit=300, verbose=False, xc_pol=None, kfrac=0.0, kshort=0.0):
    """Unrestricted HF (or spin-polarized DFT) with aufbau per spin over the
    per-m symmetry blocks and Pulay DIIS on both spins' commutators.

    xc_pol(Pa, Pb) -> (Va, Vb, Exc) selects Kohn-Sham DFT, with the fraction
    kfrac of exact exchange for a hybrid and kshort of its short-range part.
    Energy as the reference's fock builders (atomic/main.cpp:369-470,
    diatomic/main.cpp:370-479): tr(P T) + tr(P V) + 1/2 tr(P J) + Exx + Exc
    + Enucr, Exx = 1/2 (tr(Pa Ka) + tr(Pb Kb))."""
    S = basis.overlap()
    T = basis.kinetic()
    V = basis.nuclear()
    H = T + V
    blocks = basis.sym_idx(symm)
    X = [invh(S[np.ix_(b, b)]) for b in blocks]

    def density(F, nocc):
        eps, vecs = [], []
        for k, b in enumerate(blocks):
            e, c = np.linalg.eigh(X[k].T @ F[np.ix_(b, b)] @ X[k])
            eps.append(e)
            vecs.append(c)
        allev = sorted((e, k, i) for k, ev in enumerate(eps) for i, e in enumerate(ev))
        occ = [np.zeros(len(ev)) for ev in eps]
        for e, k, i in allev[:nocc]:
            occ[k][i] = 1.0
        P = np.zeros_like(F)
        for k, b in enumerate(blocks):
            C = X[k] @ vecs[k]
            P[np.ix_(b, b)] += (C * occ[k]) @ C.T
        return P

    Pa, Pb = density(H, nela), density(H, nelb)
    diis_F, diis_e = [], []
    Eold, E = 0.0, 0.0
    for it in range(1, maxit + 1):
        P = Pa + Pb
        J = basis.coulomb(P)
        Ekin, Enuc, Ecoul = np.sum(P * T), np.sum(P * V), 0.5 * np.sum(P * J)
        if xc_pol is None:
            Ka, Kb = basis.exchange(Pa), basis.exchange(Pb)
            Exx, Exc = 0.5 * (np.sum(Pa * Ka) + np.sum(Pb * Kb)), 0.0
            Fa, Fb = H + J + Ka, H + J + Kb
        else:
            Va, Vb, Exc = xc_pol(Pa, Pb)
            Exx = 0.0
            Fa, Fb = H + J + Va, H + J + Vb
            if kfrac or kshort:
                def kx(Ps):
                    K = kfrac * basis.exchange(Ps) if kfrac else 0.0
                    return K + kshort * basis.rs_exchange(Ps) if kshort else K
                Ka, Kb = kx(Pa), kx(Pb)
                Exx = 0.5 * (np.sum(Pa * Ka) + np.sum(Pb * Kb))
                Fa, Fb = Fa + Ka, Fb + Kb
        E = Ekin + Enuc + Ecoul + Exx + Exc
        errs = []
        for F_, P_ in ((Fa, Pa), (Fb, Pb)):
            for k, b in enumerate(blocks):
                Fo = X[k].T @ F_[np.ix_(b, b)] @ X[k]
                Sinv = np.linalg.pinv(X[k])
                Po = Sinv @ P_[np.ix_(b, b)] @ Sinv.T
                errs.append((Fo @ Po - Po @ Fo).ravel())
        err = np.concatenate(errs)
        emax = np.max(np.abs(err))
        if verbose:
            print(f"it {it:3d} E {E:.12f} dE {E - Eold:.3e} err {emax:.3e}")
        if it > 1 and abs(E - Eold) < conv and emax < 1e-6:
            break
        Eold = E
        diis_F.append((Fa, Fb))
        diis_e.append(err)
        if len(diis_F) > 10:
            diis_F.pop(0)
            diis_e.pop(0)
        n = len(diis_F)
        B = -np.ones((n + 1, n + 1))
        B[n, n] = 0.0
        for i in range(n):
            for j in range(n):
                B[i, j] = diis_e[i] @ diis_e[j]
        rhs = np.zeros(n + 1)
        rhs[n] = -1.0
        try:
            c = np.linalg.solve(B, rhs)[:n]
            Fa_d = sum(ci * f[0] for ci, f in zip(c, diis_F))
            Fb_d = sum(ci * f[1] for ci, f in zip(c, diis_F))
        except np.linalg.LinAlgError:
            Fa_d, Fb_d = Fa, Fb
        Pa, Pb = density(Fa_d, nela), density(Fb_d, nelb)
    E += getattr(basis, "Enucr", 0.0)
    return SCFResult(E=E, Ekin=Ekin, Enuc=Enuc, Ecoul=Ecoul, Exx=Exx, Exc=Exc, iterations=it, Pa=Pa, Pb=Pb)


def bessel_il(x, L):
    """i_L(x) = sqrt(pi/(2x)) I_{L+1/2}(x) (libhelfem/include/math.h:52-89)."""
    return lib().orc_bessel_il(float(x), int(L))


def bessel_kl(x, L):
    """k_L(x) = sqrt(2/(pi x)) K_{L+1/2}(x) (libhelfem/include/math.h:91-115)."""
    return lib().orc_bessel_kl(float(x), int(L))


def erfc_phi(Lmax, x, y):
    """Phi_L(x, y), L = 0..Lmax: erfc(mu d)/d = mu sum_L Phi_L(mu r, mu r') P_L(cos g)."""
    out = np.zeros(Lmax + 1)
    _check(lib().orc_erfc_phi(int(Lmax), float(x), float(y), _ptr(out)))
    return out
