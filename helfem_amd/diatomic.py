"""Python mirror of HelFEM's diatomic TwoDBasis Fock-build interface.

`DiatomicBasis` follows helfem::diatomic::basis::TwoDBasis
(src/diatomic/basis.h:127-432): compute_tei(), set_absm_symmetric(),
coulomb(P), exchange(P), overlap(), kinetic(), nuclear(), get_sym_idx(),
all in the pure basis.  coulomb/exchange accept a numpy array (host path) or
a float64 CUDA tensor (device path on torch's current stream); the
row-major/column-major argument of atomic.py applies unchanged, since the
diatomic J and K also commute with transposition of P.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib
from .atomic import Context, _f64_fortran, element_grid


def diatomic_shells(lmmax):
    """lm_to_l_m, src/diatomic/basis.cpp:500-516: |m|-major, (+m, -m) pairs.

    lmmax[|m|] is the largest l of that |m|."""
    lval, mval = [], []
    for mabs, lmax in enumerate(lmmax):
        for l in range(mabs, lmax + 1):
            lval.append(l)
            mval.append(mabs)
            if mabs > 0:
                lval.append(l)
                mval.append(-mabs)
    return np.array(lval, dtype=np.int32), np.array(mval, dtype=np.int32)


def mu_max(rmax: float, Rhalf: float) -> float:
    """Largest mu of the grid, arcosh(Rmax / Rhalf) (src/diatomic/main.cpp:216)."""
    return float(np.arccosh(max(rmax / Rhalf, 1.0)))


class DiatomicBasis:
    """Prolate spheroidal finite elements x Y_lm (hf_diatomic).

    Z1 sits at z = -Rbond/2 and Z2 at z = +Rbond/2.  The mu grid has `nelem`
    elements up to mu_max(rmax, Rbond/2) on element grid `igrid` with
    exponent `zexp` (diatomic/main.cpp:70-72 defaults: igrid 4, zexp 1)."""

    def __init__(self, Z1, Z2, Rbond, lmmax=None, nelem=None, rmax=40.0, igrid=4, zexp=1.0, nnodes=15, nquad=0,
                 ctx: Context | None = None, lval=None, mval=None, bval=None):
        if lval is None:
            lval, mval = diatomic_shells(lmmax)
        self.lval = np.ascontiguousarray(lval, dtype=np.int32)
        self.mval = np.ascontiguousarray(mval, dtype=np.int32)
        self.Z1, self.Z2, self.Rbond = int(Z1), int(Z2), float(Rbond)
        self.Rhalf = 0.5 * self.Rbond
        self.Enucr = self.Z1 * self.Z2 / self.Rbond
        if bval is None:
            bval = element_grid(mu_max(rmax, self.Rhalf), nelem, igrid, zexp)
        self.bval = np.ascontiguousarray(bval, np.float64)
        self.ctx = ctx
        h = ctypes.c_void_p()
        check(lib().hf_diatomic_create(ctx.handle if ctx else None, self.Z1, self.Z2, self.Rhalf, nnodes, nquad,
                                       self.bval.ctypes.data, self.bval.size, self.lval.ctypes.data,
                                       self.mval.ctypes.data, self.lval.size, ctypes.byref(h)))
        self._h = h
        sz = np.zeros(7, dtype=np.int64)
        check(lib().hf_diatomic_sizes(self._h, sz.ctypes.data))
        self.Nbf, self.Ndummy, self.Nrad, self.Nang, self.Nel, self.n_lm, self.n_LM = (int(v) for v in sz)
        self.pure_idx = np.zeros(self.Nbf, dtype=np.int64)
        check(lib().hf_diatomic_pure_index(self._h, self.pure_idx.ctypes.data))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._LIB is not None:
            _lib._LIB.hf_diatomic_destroy(h)
            self._h = ctypes.c_void_p()

    # --- setup ---------------------------------------------------------
    def compute_tei(self, exchange: bool = True, device: bool = False):
        """The two-electron caches (basis.cpp:1382-1547).  device=True builds
        the two-channel in-element kernels and their signed Cholesky factors
        on the GPU (needs a context) and returns the wall time in seconds."""
        if not device:
            check(lib().hf_diatomic_compute_tei(self._h, int(exchange)))
            return None
        sec = ctypes.c_double(0.0)
        check(lib().hf_diatomic_compute_tei_device(self._h, ctypes.byref(sec)))
        return sec.value

    def set_absm_symmetric(self, sym: bool):
        check(lib().hf_diatomic_set_absm_symmetric(self._h, int(bool(sym))))

    def maps(self):
        lm = np.zeros(2 * self.n_lm, np.int32)
        LM = np.zeros(2 * self.n_LM, np.int32)
        check(lib().hf_diatomic_maps(self._h, lm.ctypes.data, LM.ctypes.data))
        return lm.reshape(-1, 2), LM.reshape(-1, 2)

    def get_cache(self, kind, ilm, iel):
        dims = np.zeros(2, dtype=np.int32)
        check(lib().hf_diatomic_get_cache(self._h, kind, ilm, iel, dims.ctypes.data, None))
        out = np.zeros((int(dims[0]), int(dims[1])), order="F")
        if out.size:
            check(lib().hf_diatomic_get_cache(self._h, kind, ilm, iel, dims.ctypes.data, out.ctypes.data))
        return out

    def _one_e(self, which):
        out = np.zeros((self.Nbf, self.Nbf), order="F")
        check(lib().hf_diatomic_one_electron(self._h, which, out.ctypes.data, self.Nbf))
        return out

    def one_electron_dev(self, which):
        """Overlap (0), kinetic (1) or nuclear attraction (2) as a float64 CUDA
        tensor, assembled on the device from its nonzero blocks (same values
        as the host form)."""
        import torch
        if self.ctx is None:
            raise RuntimeError("basis was created without a GPU context")
        self.ctx.bind_torch_stream()
        out = torch.empty((self.Nbf, self.Nbf), dtype=torch.float64, device="cuda")
        check(lib().hf_diatomic_one_electron_dev(self._h, which, out.data_ptr(), self.Nbf))
        return out.T  # column-major storage: the transpose view is the matrix

    def overlap(self):
        return self._one_e(0)

    def kinetic(self):
        return self._one_e(1)

    def nuclear(self):
        return self._one_e(2)

    def get_sym_idx(self, symm=1):
        """basis.cpp:913-935 (symm=1): pure indices grouped by m."""
        if symm == 0:
            return [np.arange(self.Nbf)]
        blocks, off = {}, 0
        for a in range(self.Nang):
            n = self.Nrad - (1 if self.mval[a] != 0 else 0)
            blocks.setdefault(int(self.mval[a]), []).extend(range(off, off + n))
            off += n
        return [np.array(blocks[m]) for m in sorted(blocks)]

    def work(self):
        out = np.zeros(4)
        check(lib().hf_diatomic_work(self._h, out.ctypes.data))
        return {"flops_angular": out[0], "flops_radial": out[1], "terms": int(out[2]), "blocks": int(out[3])}

    # --- Fock build ------------------------------------------------------
    def _apply(self, P, host_fn, dev_fn, out=None):
        if isinstance(P, np.ndarray):
            if host_fn is None:
                raise TypeError("this form takes a float64 CUDA tensor")
            P = _f64_fortran(P)
            if P.shape != (self.Nbf, self.Nbf):
                raise ValueError(f"P must be {self.Nbf}x{self.Nbf}")
            O = np.zeros((self.Nbf, self.Nbf), order="F")
            check(host_fn(self._h, P.ctypes.data, self.Nbf, O.ctypes.data, self.Nbf))
            return O
        import torch
        if not (isinstance(P, torch.Tensor) and P.is_cuda and P.dtype == torch.float64):
            raise TypeError("P must be a numpy array or a float64 CUDA tensor")
        if P.shape != (self.Nbf, self.Nbf) or not P.is_contiguous():
            raise ValueError(f"P must be a contiguous {self.Nbf}x{self.Nbf} tensor")
        if self.ctx is None:
            raise RuntimeError("basis was created without a GPU context")
        self.ctx.bind_torch_stream()
        O = torch.empty_like(P) if out is None else out
        check(dev_fn(self._h, P.data_ptr(), self.Nbf, O.data_ptr(), self.Nbf))
        return O

    def coulomb(self, P, out=None):
        """J(P), basis.cpp:1627-1816."""
        return self._apply(P, lib().hf_diatomic_coulomb, lib().hf_diatomic_coulomb_dev, out)

    def exchange(self, P, out=None, precision: str = "fp64"):
        """-K(P) for a spin density P, basis.cpp:1818-2089.

        precision="fp32" runs the gathers and matrix products in FP32 (device
        tensors only; about 1e-6 relative error), for early SCF iterations."""
        if precision == "fp32":
            return self._apply(P, None, lib().hf_diatomic_exchange_fp32_dev, out)
        if precision != "fp64":
            raise ValueError(f"precision must be 'fp64' or 'fp32', not {precision!r}")
        return self._apply(P, lib().hf_diatomic_exchange, lib().hf_diatomic_exchange_dev, out)
