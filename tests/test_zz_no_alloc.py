"""hf_atomic_coulomb_dev_noalloc: the allocation-free Coulomb entry point.

It runs the kernels of hf_atomic_coulomb_dev with its work space taken from
a per-GPU device array instead of the handle's buffers, so its J must equal
the ordinary entry point's bit for bit; two calls in a row (the second on a
smaller basis, which reuses the array with stale values beyond its extent)
both must.  A basis too large for the array is refused with HF_EINVAL, and a
handle whose tables are not yet on the device with HF_ESTATE.
"""
import ctypes

import numpy as np
import pytest

import helfem_amd as H
from helfem_amd._lib import lib
from helfem_amd.synthetic import random_symmetric

pytestmark = pytest.mark.gpu


def _noalloc(hb, Pd, Jd):
    n = hb.Nbf
    return lib().hf_atomic_coulomb_dev_noalloc(hb._h, ctypes.c_void_p(Pd.data_ptr()), n,
                                               ctypes.c_void_p(Jd.data_ptr()), n)


def test_noalloc_coulomb_equals_ordinary(gpu_ctx):
    import torch
    print("[gpu] allocation-free Coulomb vs hf_atomic_coulomb_dev", flush=True)
    big = H.AtomicBasis(36, 3, 2, 4, ctx=gpu_ctx)
    small = H.AtomicBasis(18, 1, 1, 3, ctx=gpu_ctx)
    cases = []
    for hb, seed in ((big, 3001), (small, 3002)):
        hb.compute_tei()
        Pd = torch.as_tensor(random_symmetric(hb.Nbf, seed=seed), device="cuda")
        Jref = hb.coulomb(Pd)  # ordinary entry point; uploads the tables
        torch.cuda.synchronize()
        cases.append((hb, Pd, Jref))
    gpu_ctx.bind_torch_stream()
    outs = []
    for hb, Pd, _ in cases:  # two calls in a row, the second on the smaller basis
        Jd = torch.full_like(Pd, np.nan)
        assert _noalloc(hb, Pd, Jd) == 0, lib().hf_last_error()
        outs.append(Jd)
    torch.cuda.synchronize()
    for (hb, _, Jref), Jd in zip(cases, outs):
        assert torch.equal(Jd, Jref), float(torch.max(torch.abs(Jd - Jref)))


def test_noalloc_coulomb_refusals(gpu_ctx):
    import torch
    print("[gpu] allocation-free Coulomb refusals", flush=True)
    huge = H.AtomicBasis(1, 8, 8, 30, ctx=gpu_ctx)  # n_lm 289 x 30 elements x 260 doubles > 16 MiB
    P = torch.zeros((huge.Nbf, huge.Nbf), dtype=torch.float64, device="cuda")
    assert _noalloc(huge, P, P.clone()) == -1
    assert b"work space" in lib().hf_last_error()
    fresh = H.AtomicBasis(2, 0, 0, 2, ctx=gpu_ctx)
    fresh.compute_tei()
    Q = torch.zeros((fresh.Nbf, fresh.Nbf), dtype=torch.float64, device="cuda")
    assert _noalloc(fresh, Q, Q.clone()) == -2
    assert b"not on the device" in lib().hf_last_error()
