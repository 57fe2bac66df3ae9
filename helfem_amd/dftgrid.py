"""Python mirror of HelFEM's DFT grids: the atomic grid
(helfem::atomic::dftgrid::DFTGrid, src/atomic/dftgrid.h:140-173) and the
pure-m diatomic grid (helfem::diatomic::dftgrid_purem::PureMDFTGrid,
src/diatomic/dftgrid_purem.h), chosen by the basis type.

`DFTGrid(basis, lang, mang).eval_Fxc(x_func, c_func, P)` returns the
restricted exchange-correlation matrix with Exc, the integrated electron
count and the kinetic-energy-density integral, like DFTGrid::eval_Fxc
(dftgrid.cpp:591-653); `eval_Fxc_pol(x_func, c_func, Pa, Pb)` is the
unrestricted overload (dftgrid.cpp:655-712) returning (Ha, Hb, Exc, Nel, Ekin).  Functionals are libxc ids; parse_method turns the
driver's --method string (e.g. "lda_x-lda_c_vwn", "gga_x_pbe-gga_c_pbe", "mgga_x_tpss-mgga_c_tpss",
"hyb_gga_xc_b3lyp") into them; every grid evaluates all of them, and exact_exchange gives a hybrid's
exact-exchange fraction.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib
from .atomic import _f64_fortran

FUNCTIONALS = {"none": 0, "lda_x": 1, "lda_c_vwn": 7, "lda_c_vwn_rpa": 8, "lda_c_pw": 12, "lda_c_pw_mod": 13,
               "gga_x_pbe": 101, "gga_x_b88": 106, "gga_c_pbe": 130, "gga_c_lyp": 131, "mgga_x_tpss": 202,
               "mgga_c_tpss": 231, "hyb_gga_xc_b3lyp": 402, "hyb_gga_xc_pbeh": 406, "hyb_gga_xc_cam_b3lyp": 433,
               "hyb_gga_xc_camy_b3lyp": 470}


def exact_exchange(x_func: int):
    """(kfrac, kshort, omega, kernel): the full-range and short-range
    exact-exchange fractions of functional x_func, its range-separation
    parameter and the short-range kernel ("erfc", "yukawa" or None), as
    range_separation and is_range_separated (src/general/dftfuncs.cpp:470-575)
    give them.  The Fock matrix of a hybrid takes kfrac * exchange(P) +
    kshort * rs_exchange(P), the latter on the caches of compute_erfc /
    compute_yukawa(omega) (src/atomic/main.cpp:336-337, 424-437)."""
    out = np.zeros(4)
    check(lib().hf_xc_exact_exchange(int(x_func), out.ctypes.data))
    return float(out[0]), float(out[1]), float(out[2]), {0: None, 1: "erfc", 2: "yukawa"}[int(out[3])]


def parse_method(method: str):
    """parse_xc_func (src/general/dftfuncs.cpp:94-117): "x-c" or a single name."""
    x, _, c = method.partition("-")
    try:
        return FUNCTIONALS[x], FUNCTIONALS[c] if c else 0
    except KeyError as err:
        raise ValueError(f"unsupported functional in {method!r}; supported: {sorted(FUNCTIONALS)}") from err


class DFTGrid:
    """DFT quadrature on the GPU (hf_dftgrid): atomic grid for an AtomicBasis;
    for a DiatomicBasis the pure-m grid (purem=True, mang unused) or the
    general 3D grid (purem=False), as the diatomic driver's --purem switch."""

    def __init__(self, basis, lang: int | None = None, mang: int | None = None, purem: bool = True, rule: int = 0):
        from .diatomic import DiatomicBasis
        lmax = int(np.max(basis.lval))
        mmax = int(np.max(np.abs(basis.mval)))
        self.lang = 4 * lmax + 12 if lang is None else int(lang)
        self.basis = basis  # the grid points into the basis handle
        h = ctypes.c_void_p()
        if isinstance(basis, DiatomicBasis) and purem:
            self.mang = 1
            check(lib().hf_diatomic_dftgrid_create(basis._h, self.lang, ctypes.byref(h)))
        elif isinstance(basis, DiatomicBasis):
            self.mang = 4 * mmax + 12 if mang is None else int(mang)
            check(lib().hf_diatomic_dftgrid3d_create(basis._h, self.lang, self.mang, ctypes.byref(h)))
        else:
            self.mang = 4 * mmax + 12 if mang is None else int(mang)
            check(lib().hf_atomic_dftgrid_create_rule(basis._h, self.lang, self.mang, int(rule), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._LIB is not None:
            _lib._LIB.hf_dftgrid_destroy(h)
            self._h = ctypes.c_void_p()

    def eval_Fxc(self, x_func: int, c_func: int, P, thr: float = 1e-12, out=None):
        """(H, Exc, Nel, Ekin) for the total density P (numpy or float64 CUDA tensor)."""
        n = self.basis.Nbf
        res = np.zeros(3)
        if isinstance(P, np.ndarray):
            P = _f64_fortran(P)
            if P.shape != (n, n):
                raise ValueError(f"P must be {n}x{n}")
            H = np.zeros((n, n), order="F")
            check(lib().hf_dftgrid_eval_fxc(self._h, x_func, c_func, P.ctypes.data, n, H.ctypes.data, n, thr,
                                            res.ctypes.data))
            return H, float(res[0]), float(res[1]), float(res[2])
        import torch
        if not (isinstance(P, torch.Tensor) and P.is_cuda and P.dtype == torch.float64):
            raise TypeError("P must be a numpy array or a float64 CUDA tensor")
        if P.shape != (n, n) or not P.is_contiguous():
            raise ValueError(f"P must be a contiguous {n}x{n} tensor")
        self.basis.ctx.bind_torch_stream()
        H = torch.empty_like(P) if out is None else out
        check(lib().hf_dftgrid_eval_fxc_dev(self._h, x_func, c_func, P.data_ptr(), n, H.data_ptr(), n, thr,
                                            res.ctypes.data))
        return H, float(res[0]), float(res[1]), float(res[2])

    def eval_Fxc_pol(self, x_func: int, c_func: int, Pa, Pb, beta: bool = True, thr: float = 1e-12):
        """(Ha, Hb, Exc, Nel, Ekin) for spin densities Pa, Pb (numpy or float64
        CUDA tensors); Hb is zero when beta is False (no beta electrons)."""
        n = self.basis.Nbf
        res = np.zeros(3)
        if isinstance(Pa, np.ndarray):
            Pa, Pb = _f64_fortran(Pa), _f64_fortran(Pb)
            if Pa.shape != (n, n) or Pb.shape != (n, n):
                raise ValueError(f"Pa and Pb must be {n}x{n}")
            Ha = np.zeros((n, n), order="F")
            Hb = np.zeros((n, n), order="F")
            check(lib().hf_dftgrid_eval_fxc_pol(self._h, x_func, c_func, Pa.ctypes.data, n, Pb.ctypes.data, n,
                                                Ha.ctypes.data, n, Hb.ctypes.data, n, int(beta), thr,
                                                res.ctypes.data))
            return Ha, Hb, float(res[0]), float(res[1]), float(res[2])
        import torch
        for name, M in (("Pa", Pa), ("Pb", Pb)):
            if not (isinstance(M, torch.Tensor) and M.is_cuda and M.dtype == torch.float64):
                raise TypeError(f"{name} must be a numpy array or a float64 CUDA tensor")
            if M.shape != (n, n) or not M.is_contiguous():
                raise ValueError(f"{name} must be a contiguous {n}x{n} tensor")
        self.basis.ctx.bind_torch_stream()
        Ha = torch.empty_like(Pa)
        Hb = torch.empty_like(Pb)
        check(lib().hf_dftgrid_eval_fxc_pol_dev(self._h, x_func, c_func, Pa.data_ptr(), n, Pb.data_ptr(), n,
                                                Ha.data_ptr(), n, Hb.data_ptr(), n, int(beta), thr,
                                                res.ctypes.data))
        return Ha, Hb, float(res[0]), float(res[1]), float(res[2])


def quadrature(kind: str, n: int):
    """(x, w) of the Chebyshev ("chebyshev", the reference's radial and
    cos(theta) rule), Gauss-Lobatto ("lobatto") or Gauss-Legendre ("legendre")
    rule on [-1, 1] (hf_quadrature)."""
    k = {"chebyshev": 0, "lobatto": 1, "legendre": 2}[kind]
    x, w = np.zeros(n), np.zeros(n)
    check(lib().hf_quadrature(k, int(n), x.ctypes.data, w.ctypes.data))
    return x, w
