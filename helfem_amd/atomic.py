"""Python mirror of HelFEM's atomic TwoDBasis Fock-build interface.

`AtomicBasis` follows helfem::atomic::basis::TwoDBasisT<double>
(src/atomic/TwoDBasis.h:52-283): compute_tei(), coulomb(P), exchange(P),
overlap(), kinetic(), nuclear(), get_sym_idx().  coulomb/exchange accept a
numpy array (host path: copied to the GPU and back) or a torch CUDA tensor
(device path: asynchronous on torch's current stream, result stays in HBM).

torch tensors are row-major; the kernels read column-major.  J and K commute
with transposition of P (the two-electron integrals are symmetric under
(ij|kl) = (ji|lk)), so reading a row-major P as its transpose and writing
the row-major result gives exactly J(P) and K(P).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib


def angular_basis(lmax: int, mmax: int):
    """src/atomic/basis.cpp:179-203: |m| outer, l inner, (+m, -m) pairs."""
    lval, mval = [], []
    for mabs in range(mmax + 1):
        for l in range(mabs, lmax + 1):
            lval.append(l)
            mval.append(mabs)
            if mabs > 0:
                lval.append(l)
                mval.append(-mabs)
    return np.array(lval, dtype=np.int32), np.array(mval, dtype=np.int32)


def element_grid(rmax: float, nelem: int, igrid: int = 4, zexp: float = 2.0) -> np.ndarray:
    """Element boundaries on [0, rmax] (hf_element_grid; libhelfem/include/grid.h:37-101)."""
    out = np.zeros(nelem + 1)
    check(lib().hf_element_grid(float(rmax), int(nelem), int(igrid), float(zexp), out.ctypes.data))
    return out


class Context:
    """One GPU and one stream (hf_ctx)."""

    def __init__(self, device: int = 0, use_torch_stream: bool = True):
        h = ctypes.c_void_p()
        check(lib().hf_ctx_create(device, 0, ctypes.byref(h)))
        self._h = h
        self.device = device
        self.use_torch_stream = use_torch_stream

    @property
    def handle(self):
        return self._h

    def bind_torch_stream(self):
        """Issue this context's work on torch's current stream."""
        import torch
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(lib().hf_ctx_set_stream(self._h, ctypes.c_void_p(s)))

    def synchronize(self):
        check(lib().hf_ctx_synchronize(self._h))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().hf_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _f64_fortran(a):
    return np.asfortranarray(a, dtype=np.float64)


class AtomicBasis:
    """Radial finite elements x complex spherical harmonics (hf_atomic)."""

    def __init__(self, Z, lmax=None, mmax=None, nelem=None, rmax=40.0, igrid=4, zexp=2.0, nnodes=15, nquad=0,
                 ctx: Context | None = None, lval=None, mval=None, bval=None):
        if lval is None:
            lval, mval = angular_basis(lmax, mmax)
        self.lval = np.ascontiguousarray(lval, dtype=np.int32)
        self.mval = np.ascontiguousarray(mval, dtype=np.int32)
        self.bval = element_grid(rmax, nelem, igrid, zexp) if bval is None else np.ascontiguousarray(bval, np.float64)
        self.Z = int(Z)
        self.ctx = ctx
        h = ctypes.c_void_p()
        check(lib().hf_atomic_create(ctx.handle if ctx else None, self.Z, nnodes, nquad,
                                     self.bval.ctypes.data, self.bval.size, self.lval.ctypes.data,
                                     self.mval.ctypes.data, self.lval.size, ctypes.byref(h)))
        self._h = h
        sz = np.zeros(5, dtype=np.int64)
        check(lib().hf_atomic_sizes(self._h, sz.ctypes.data))
        self.Nbf, self.Nrad, self.Nang, self.Nel, self.N_L = (int(v) for v in sz)
        self.first = np.zeros(self.Nel, np.int32)
        self.nprim = np.zeros(self.Nel, np.int32)
        check(lib().hf_atomic_elements(self._h, self.first.ctypes.data, self.nprim.ctypes.data))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._LIB is not None:
            _lib._LIB.hf_atomic_destroy(h)
            self._h = ctypes.c_void_p()

    def clone(self):
        """A second handle on this basis (hf_atomic_clone): same host setup
        and caches, device tables of its own."""
        h = ctypes.c_void_p()
        check(lib().hf_atomic_clone(self._h, ctypes.byref(h)))
        c = object.__new__(AtomicBasis)
        c.__dict__.update({k: v for k, v in self.__dict__.items() if k != "_h"})
        c._h = h
        return c

    # --- setup ---------------------------------------------------------
    def compute_tei(self, exchange: bool = True, device: bool = False):
        """The two-electron caches (TwoDBasis.cpp:709-735).  device=True builds
        the in-element Slater tensors and their Cholesky factors on the GPU
        (needs a context) and returns the wall time in seconds."""
        if not device:
            check(lib().hf_atomic_compute_tei(self._h, int(exchange)))
            return None
        sec = ctypes.c_double(0.0)
        check(lib().hf_atomic_compute_tei_device(self._h, ctypes.byref(sec)))
        return sec.value

    def set_cache(self, kind, L, iel, mat):
        mat = _f64_fortran(mat)
        check(lib().hf_atomic_set_cache(self._h, kind, L, iel, mat.shape[0], mat.shape[1], mat.ctypes.data))

    def commit_tei(self):
        check(lib().hf_atomic_commit_tei(self._h))

    def get_cache(self, kind, L, iel):
        dims = np.zeros(2, dtype=np.int32)
        check(lib().hf_atomic_get_cache(self._h, kind, L, iel, dims.ctypes.data, None))
        out = np.zeros((int(dims[0]), int(dims[1])), order="F")
        if out.size:
            check(lib().hf_atomic_get_cache(self._h, kind, L, iel, dims.ctypes.data, out.ctypes.data))
        return out

    def _one_e(self, which):
        out = np.zeros((self.Nbf, self.Nbf), order="F")
        check(lib().hf_atomic_one_electron(self._h, which, out.ctypes.data, self.Nbf))
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
        check(lib().hf_atomic_one_electron_dev(self._h, which, out.data_ptr(), self.Nbf))
        return out.T  # column-major storage: the transpose view is the matrix

    def prepare_one_electron(self):
        """Build and send the radial one-electron factors now (host work,
        safe to run for different bases on different threads), so later
        one_electron_dev calls only expand them on the device."""
        if self.ctx is None:
            raise RuntimeError("basis was created without a GPU context")
        check(lib().hf_atomic_one_electron_prepare(self._h))

    def overlap(self):
        return self._one_e(0)

    def kinetic(self):
        return self._one_e(1)

    def nuclear(self):
        return self._one_e(2)

    def get_sym_idx(self, symm=1):
        """TwoDBasis.cpp:200-227: one index list per unique m (symm=1) or per
        (l, m) shell (symm=2).  symm=3 splits each m by the parity of l (even
        before odd): a density of definite-parity orbitals never couples the
        two."""
        if symm == 0:
            return [np.arange(self.Nbf)]
        if symm == 2:  # one block per (l, m) shell, in shell order
            return [np.arange(a * self.Nrad, (a + 1) * self.Nrad) for a in range(self.Nang)]
        out = []
        if symm == 3:
            for m in sorted(set(self.mval.tolist())):
                for par in (0, 1):
                    sh = [a for a in range(self.Nang) if self.mval[a] == m and self.lval[a] % 2 == par]
                    if sh:
                        out.append(np.concatenate([np.arange(a * self.Nrad, (a + 1) * self.Nrad) for a in sh]))
            return out
        for m in sorted(set(self.mval.tolist())):
            out.append(np.concatenate([np.arange(a * self.Nrad, (a + 1) * self.Nrad)
                                       for a in range(self.Nang) if self.mval[a] == m]))
        return out

    def work(self):
        """Work of the most recent exchange build (hf_atomic_work)."""
        out = np.zeros(4)
        check(lib().hf_atomic_work(self._h, out.ctypes.data))
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

    def compute_yukawa(self, lam: float):
        """Yukawa range-separation caches (TwoDBasis.cpp:738-760)."""
        check(lib().hf_atomic_compute_yukawa(self._h, float(lam)))

    def compute_erfc(self, mu: float):
        """erfc range-separation caches (TwoDBasis.cpp:762-771)."""
        check(lib().hf_atomic_compute_erfc(self._h, float(mu)))

    def rs_exchange(self, P, out=None):
        """-K of the short-range kernel, TwoDBasis.cpp:1002-1131."""
        return self._apply(P, lib().hf_atomic_rs_exchange, lib().hf_atomic_rs_exchange_dev, out)

    def coulomb(self, P, out=None):
        """J(P), TwoDBasis.cpp:774-877."""
        return self._apply(P, lib().hf_atomic_coulomb, lib().hf_atomic_coulomb_dev, out)

    def exchange(self, P, out=None, precision: str = "fp64"):
        """-K(P) for a spin density P, TwoDBasis.cpp:880-999.

        precision="fp32" runs the gathers and matrix products in FP32 (device
        tensors only; about 1e-6 relative error), for early SCF iterations."""
        if precision == "fp32":
            return self._apply(P, None, lib().hf_atomic_exchange_fp32_dev, out)
        if precision != "fp64":
            raise ValueError(f"precision must be 'fp64' or 'fp32', not {precision!r}")
        return self._apply(P, lib().hf_atomic_exchange, lib().hf_atomic_exchange_dev, out)


def _ptr_array(tensors):
    return (ctypes.c_void_p * len(tensors))(*[t.data_ptr() for t in tensors])


def _ld_array(tensors):
    return (ctypes.c_int64 * len(tensors))(*[t.shape[0] for t in tensors])


class AtomicBatch:
    """Coulomb and exchange builds of many atomic bases as single launches
    (hf_atomic_batch: one launch per kernel for all systems, the launch
    sequence replayed from a HIP graph, workgroup shapes autotuned).

    The bases must share one Context and have compute_tei() done; a basis may
    appear once (use AtomicBasis.clone() for a second density of the same
    basis).  Matrices are float64 CUDA tensors on torch's current stream."""

    def __init__(self, bases, ctx: Context):
        self.bases = list(bases)
        self.ctx = ctx
        arr = (ctypes.c_void_p * len(self.bases))(*[b._h.value for b in self.bases])
        h = ctypes.c_void_p()
        ctx.bind_torch_stream()
        check(lib().hf_atomic_batch_create(ctx.handle, arr, len(self.bases), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._LIB is not None:
            _lib._LIB.hf_atomic_batch_destroy(h)
            self._h = ctypes.c_void_p()

    def set_options(self, use_graphs: bool = True, autotune: bool = True):
        check(lib().hf_atomic_batch_set_options(self._h, int(use_graphs), int(autotune)))

    def tuning(self):
        """Autotuner state: waves per workgroup per (class, part), and the
        measured launch times (ms) of the 1, 2, 4-wave shapes."""
        nw = np.zeros(8, np.int32)
        ms = np.zeros(24)
        check(lib().hf_atomic_batch_tuning(self._h, nw.ctypes.data, ms.ctypes.data))
        return nw.reshape(4, 2), ms.reshape(4, 2, 3)

    def _check(self, Ps, outs):
        if len(Ps) != len(self.bases) or len(outs) != len(self.bases):
            raise ValueError("one density and one output per basis")
        for b, P, O in zip(self.bases, Ps, outs):
            for X in (P, O):
                if X.shape != (b.Nbf, b.Nbf) or not X.is_contiguous() or not X.is_cuda:
                    raise ValueError(f"matrices must be contiguous {b.Nbf}x{b.Nbf} CUDA tensors")

    def coulomb(self, Ps, outs):
        """outs[s] = J(Ps[s]) for every basis s."""
        self._check(Ps, outs)
        self.ctx.bind_torch_stream()
        check(lib().hf_atomic_batch_coulomb_dev(self._h, _ptr_array(Ps), _ld_array(Ps), _ptr_array(outs),
                                                _ld_array(outs)))
        return outs

    def exchange(self, Ps, outs, fp32=None):
        """outs[s] = -K(Ps[s]) for every basis s; fp32[s] true selects the
        reduced-precision form for that system."""
        self._check(Ps, outs)
        self.ctx.bind_torch_stream()
        f = None
        if fp32 is not None:
            f = (ctypes.c_int * len(self.bases))(*[int(bool(x)) for x in fp32])
        check(lib().hf_atomic_batch_exchange_dev(self._h, _ptr_array(Ps), _ld_array(Ps), _ptr_array(outs),
                                                 _ld_array(outs), f))
        return outs
