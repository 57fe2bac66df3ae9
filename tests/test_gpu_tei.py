"""Atomic two-electron setup on the device (hf_atomic_compute_tei_device):
the in-element Slater tensors (split-at-r=r' quadrature with Lobatto order
refinement, libhelfem/src/quadrature.cpp:36-161, RadialBasis.cpp:644-663)
and their pivoted Cholesky factors (RadialBasis.cpp:671-709) built on the
GPU, against the host setup (bit-identical to the oracle) and, through J and
K, against the oracle.

The factors themselves may differ where two residual diagonals tie to
rounding (another pivot order); what is fixed is the reconstructed tensor
L L', which both setups hold to the Cholesky tolerance (1e-12 absolute on
the residual diagonal, so at most ~1e-12 on any entry) plus rounding of the
quadrature sums.  Bar: max |L_d L_d' - L_h L_h'| <= 4e-12 + 1e-13 max |T|
per (L, element); J and K from the device caches within 1e-12 of the oracle.
"""
import numpy as np
import pytest

import helfem_amd as H
import oracle as O
from helfem_amd.synthetic import random_symmetric

pytestmark = pytest.mark.gpu

# (Z, lmax, mmax, elements, LIP nodes)
CASES = [(2, 0, 0, 4, 15), (10, 2, 2, 5, 15), (36, 3, 2, 4, 10), (18, 1, 1, 12, 8)]


def _pair(ctx, Z, lmax, mmax, nel, nn):
    hb = H.AtomicBasis(Z, lmax, mmax, nel, nnodes=nn, ctx=ctx)
    hb.compute_tei()
    db = H.AtomicBasis(Z, lmax, mmax, nel, nnodes=nn, ctx=ctx)
    sec = db.compute_tei(device=True)
    return hb, db, sec


@pytest.mark.parametrize("case", CASES, ids=[f"Z{c[0]}-l{c[1]}m{c[2]}-{c[3]}el-n{c[4]}" for c in CASES])
def test_device_setup_tensors_match_host(case, gpu_ctx):
    print(f"[gpu] device two-electron setup {case}", flush=True)
    hb, db, sec = _pair(gpu_ctx, *case)
    worst, dr = 0.0, 0
    for L in range(hb.N_L):
        for e in range(hb.Nel):
            Fh, Fd = hb.get_cache(2, L, e), db.get_cache(2, L, e)
            Th, Td = Fh @ Fh.T, Fd @ Fd.T
            err = float(np.max(np.abs(Th - Td)))
            bound = 4e-12 + 1e-13 * float(np.max(np.abs(Th)))
            assert err <= bound, (L, e, err, bound)
            worst = max(worst, err / bound)
            dr = max(dr, abs(Fh.shape[1] - Fd.shape[1]))
            # the element matrices stay the host's own
            assert np.array_equal(hb.get_cache(0, L, e), db.get_cache(0, L, e))
    print(f"[gpu]   {sec * 1e3:.1f} ms; worst error / bound {worst:.2f}; rank difference <= {dr}", flush=True)
    assert dr <= 1


@pytest.mark.parametrize("case", CASES[1:3], ids=["Z10-l2m2", "Z36-l3m2"])
def test_device_setup_coulomb_exchange_vs_oracle(case, gpu_ctx):
    import torch
    Z, lmax, mmax, nel, nn = case
    print(f"[gpu] J and K from the device setup vs oracle {case}", flush=True)
    db = H.AtomicBasis(Z, lmax, mmax, nel, nnodes=nn, ctx=gpu_ctx)
    db.compute_tei(device=True)
    ob = O.AtomicBasis(Z, lmax, mmax, nel, nnodes=nn)
    ob.compute_tei()
    P = random_symmetric(db.Nbf, seed=4242)
    Pd = torch.as_tensor(P, device="cuda")
    J = db.coulomb(Pd).cpu().numpy()
    K = db.exchange(Pd).cpu().numpy()
    Jr, Kr = ob.coulomb(P), ob.exchange(P)
    ej = np.max(np.abs(J - Jr)) / np.max(np.abs(Jr))
    ek = np.max(np.abs(K - Kr)) / np.max(np.abs(Kr))
    print(f"[gpu]   rel err J {ej:.2e} K {ek:.2e}", flush=True)
    assert ej <= 1e-12 and ek <= 1e-12


# diatomic (Z1, Z2, R, lmax per m, elements, LIP nodes)
DCASES = [(1, 1, 1.4, [4], 3, 15), (7, 1, 2.0, [6, 4, 2], 3, 15), (7, 7, 2.074, [8, 8], 4, 10)]


def _dpair(ctx, Z1, Z2, R, lmm, nel, nn):
    hb = H.DiatomicBasis(Z1, Z2, R, lmm, nel, nnodes=nn, ctx=ctx)
    hb.compute_tei()
    db = H.DiatomicBasis(Z1, Z2, R, lmm, nel, nnodes=nn, ctx=ctx)
    sec = db.compute_tei(device=True)
    return hb, db, sec


@pytest.mark.parametrize("case", DCASES, ids=["H2-l4", "NH-l642", "N2-l88-n10"])
def test_diatomic_device_setup_kernels_match_host(case, gpu_ctx):
    """The two-channel kernels W = B diag(sigma) B' rebuilt from the device
    factors equal the host's to the signed Cholesky's stopping rule (1e-12
    of the largest initial |diagonal|, here bounded by 4e-12 max|W| plus
    rounding); the disjoint P/Q integrals are the host's own."""
    print(f"[gpu] diatomic device two-electron setup {case}", flush=True)
    hb, db, sec = _dpair(gpu_ctx, *case)
    worst, dr = 0.0, 0
    for q in range(hb.n_lm):
        for e in range(hb.Nel):
            Bh, Bd = hb.get_cache(4, q, e), db.get_cache(4, q, e)
            sh, sd = hb.get_cache(5, q, e)[:, 0], db.get_cache(5, q, e)[:, 0]
            Wh, Wd = (Bh * sh) @ Bh.T, (Bd * sd) @ Bd.T
            scale = float(np.max(np.abs(Wh)))
            err = float(np.max(np.abs(Wh - Wd)))
            bound = 4e-12 * scale + 1e-300
            assert err <= bound, (q, e, err, scale)
            worst = max(worst, err / bound)
            dr = max(dr, abs(Bh.shape[1] - Bd.shape[1]))
            assert np.array_equal(hb.get_cache(0, q, e), db.get_cache(0, q, e))
    print(f"[gpu]   {sec * 1e3:.1f} ms; worst error / bound {worst:.2f}; rank difference <= {dr}", flush=True)
    assert dr <= 1


@pytest.mark.parametrize("case", DCASES[:2], ids=["H2-l4", "NH-l642"])
def test_diatomic_device_setup_coulomb_exchange_vs_oracle(case, gpu_ctx):
    import torch
    Z1, Z2, R, lmm, nel, nn = case
    print(f"[gpu] diatomic J and K from the device setup vs oracle {case}", flush=True)
    db = H.DiatomicBasis(Z1, Z2, R, lmm, nel, nnodes=nn, ctx=gpu_ctx)
    db.compute_tei(device=True)
    ob = O.DiatomicBasis(Z1, Z2, R, lmm, nel, nnodes=nn)
    ob.compute_tei()
    P = random_symmetric(db.Nbf, seed=4343)
    Pd = torch.as_tensor(P, device="cuda")
    J = db.coulomb(Pd).cpu().numpy()
    K = db.exchange(Pd).cpu().numpy()
    Jr, Kr = ob.coulomb(P), ob.exchange(P)
    ej = np.max(np.abs(J - Jr)) / np.max(np.abs(Jr))
    ek = np.max(np.abs(K - Kr)) / np.max(np.abs(Kr))
    print(f"[gpu]   rel err J {ej:.2e} K {ek:.2e}", flush=True)
    assert ej <= 1e-12 and ek <= 1e-12
