"""GPU parity of the atomic Coulomb/exchange kernels against the oracle.

Both sides get the same factor caches (the oracle's, uploaded through
hf_atomic_set_cache), so the comparison isolates the kernels.  Bar: FP64
relative error max|X_gpu - X_oracle| / max|X_oracle| <= 1e-12.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import helfem_amd as H
import oracle as O
from helfem_amd.synthetic import random_symmetric, seeded_density

pytestmark = pytest.mark.gpu
TOL = 1e-12
REF = json.loads((Path(__file__).parent / "golden" / "reference_energies.json").read_text())

CASES = {
    # Ne, spherically averaged, 10 elements to 40 bohr (the single-L J/K case)
    "Ne-l0-10el": dict(Z=10, lmax=0, mmax=0, nelem=10),
    # Ne at lmax=mmax=1 (holds 2p6), 10 elements
    "Ne-l1-10el": dict(Z=10, lmax=1, mmax=1, nelem=10),
    # mmax < lmax, three elements (first/last element sizes 14, one interior)
    "Ar-l2m1-3el": dict(Z=18, lmax=2, mmax=1, nelem=3),
    # many couplings: lmax=3, mmax=2
    "Kr-l3m2-4el": dict(Z=36, lmax=3, mmax=2, nelem=4),
    # one element only (both end functions dropped from the same element)
    "He-l1-1el": dict(Z=2, lmax=1, mmax=1, nelem=1),
}
_cache = {}


def pair(name, ctx):
    if name not in _cache:
        cfg = CASES[name]
        ob = O.AtomicBasis(cfg["Z"], cfg["lmax"], cfg["mmax"], cfg["nelem"])
        ob.compute_tei()
        hb = H.AtomicBasis(cfg["Z"], cfg["lmax"], cfg["mmax"], cfg["nelem"], ctx=ctx)
        for kind in range(3):
            for L in range(ob.N_L):
                for e in range(ob.Nel):
                    m = ob.cache(kind, L, e)
                    if m.size:
                        hb.set_cache(kind, L, e, m)
        hb.commit_tei()
        _cache[name] = (ob, hb)
    return _cache[name]


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name", list(CASES))
def test_coulomb_exchange_random_symmetric(name, gpu_ctx):
    print(f"[gpu] atomic J/K random symmetric P: {name}", flush=True)
    ob, hb = pair(name, gpu_ctx)
    P = random_symmetric(ob.Nbf, seed=1001)
    Jr, Kr = ob.coulomb(P), ob.exchange(P)
    Jg, Kg = hb.coulomb(P), hb.exchange(P)
    assert relerr(Jg, Jr) <= TOL, relerr(Jg, Jr)
    assert relerr(Kg, Kr) <= TOL, relerr(Kg, Kr)


@pytest.mark.parametrize("name", ["Ne-l0-10el", "Ne-l1-10el", "Kr-l3m2-4el"])
def test_device_path_physical_density(name, gpu_ctx):
    import torch
    print(f"[gpu] atomic J/K device path, seeded density: {name}", flush=True)
    ob, hb = pair(name, gpu_ctx)
    S = ob.overlap()
    blocks = ob.sym_idx(1)
    nocc = [1 if i == 0 else 0 for i in range(len(blocks))]
    nocc[len(blocks) // 2] += 1
    P = seeded_density(S, blocks, nocc, seed=1001)
    Pd = torch.as_tensor(P, device="cuda")
    Jg = hb.coulomb(Pd)
    Kg = hb.exchange(0.5 * Pd)
    torch.cuda.synchronize()
    Jr, Kr = ob.coulomb(P), ob.exchange(0.5 * P)
    assert relerr(Jg.cpu().numpy(), Jr) <= TOL
    assert relerr(Kg.cpu().numpy(), Kr) <= TOL
    # m-block-diagonal density: the norm screening path of exchange
    assert np.max(np.abs(Kg.cpu().numpy() - Kg.cpu().numpy().T)) <= 1e-13 * np.max(np.abs(Kr))


def test_repeat_calls_are_stable(gpu_ctx):
    print("[gpu] atomic J/K repeated calls", flush=True)
    ob, hb = pair("Ar-l2m1-3el", gpu_ctx)
    P = random_symmetric(ob.Nbf, seed=1002)
    K1 = hb.exchange(P)
    K2 = hb.exchange(P)
    J1 = hb.coulomb(P)
    J2 = hb.coulomb(P)
    assert np.array_equal(J1, J2)
    assert relerr(K1, K2) <= 1e-15


def test_gpu_scf_reproduces_reference_energy(gpu_ctx):
    """Product setup + GPU Fock build + device SCF: Ne HF at lmax=1, 5
    elements, -128.5470981094 (tests/README.md:79-80); oracle SCF within 1e-10."""
    from helfem_amd.scf import rhf
    print("[gpu] Ne HF SCF through the GPU Fock build", flush=True)
    hb = H.AtomicBasis(10, 1, 1, 5, ctx=gpu_ctx)
    hb.compute_tei()
    r = rhf(hb, 10)
    ob = O.AtomicBasis(10, 1, 1, 5)
    ob.compute_tei()
    ro = O.scf_rhf(ob, 10)
    assert abs(r["E"] - ro["E"]) < 1e-10, (r["E"], ro["E"])
    assert abs(r["E"] + 128.5470981094) < 1e-9


def test_dense_operator_inside_elements(gpu_ctx, monkeypatch):
    """With a memory budget for it (HFGPU_EXCHANGE_OPERATOR_GIB) the in-element
    exchange runs as a dense-operator GEMM instead of the default Cholesky-form
    kernel; both agree with the oracle."""
    print("[gpu] atomic K, dense in-element operator", flush=True)
    monkeypatch.setenv("HFGPU_EXCHANGE_OPERATOR_GIB", "32")
    cfg = CASES["Kr-l3m2-4el"]
    ob = O.AtomicBasis(cfg["Z"], cfg["lmax"], cfg["mmax"], cfg["nelem"])
    ob.compute_tei()
    hb = H.AtomicBasis(cfg["Z"], cfg["lmax"], cfg["mmax"], cfg["nelem"], ctx=gpu_ctx)
    hb.compute_tei()
    P = random_symmetric(ob.Nbf, seed=1005)
    assert relerr(hb.exchange(P), ob.exchange(P)) <= TOL


@pytest.mark.parametrize("name", ["Kr-l3m2-4el", "Ar-l2m1-3el"])
def test_symmetric_and_general_density_paths(name, gpu_ctx, monkeypatch):
    """A symmetric density (|P - P'| <= 1e-14 |P|) takes the symmetric exchange
    (half the element pairs and blocks, mirrored); anything else the general
    one.  Both meet the oracle bar: a non-symmetric P, a symmetric one, the
    same with the symmetric path switched off (HFGPU_EXCHANGE_SYMMETRIC=0),
    and a symmetric P with rounding-level asymmetry."""
    import torch
    print(f"[gpu] atomic K, symmetric vs general density: {name}", flush=True)
    ob, hb = pair(name, gpu_ctx)
    rng = np.random.default_rng(1009)
    G = rng.standard_normal((ob.Nbf, ob.Nbf))
    S = random_symmetric(ob.Nbf, seed=1010)
    N = S * (1.0 + 1e-16 * rng.standard_normal(S.shape))
    for P in (G, S, N):
        Kg = hb.exchange(torch.as_tensor(P, device="cuda")).cpu().numpy()
        assert relerr(Kg, ob.exchange(P)) <= TOL
    monkeypatch.setenv("HFGPU_EXCHANGE_SYMMETRIC", "0")
    Kg = hb.exchange(torch.as_tensor(S, device="cuda")).cpu().numpy()
    assert relerr(Kg, ob.exchange(S)) <= TOL
    monkeypatch.delenv("HFGPU_EXCHANGE_SYMMETRIC")
    torch.cuda.synchronize()


@pytest.mark.parametrize("name", ["Kr-l3m2-4el", "He-l1-1el"])
def test_direct_gather_without_tile_budget(name, gpu_ctx, monkeypatch):
    """With no memory granted for the tiled density (HFGPU_EXCHANGE_TILE_GIB=0)
    the exchange gathers straight from P; FP64 still meets the oracle bar and
    FP32 agrees with the tiled FP32 path to rounding."""
    import torch
    print(f"[gpu] atomic K, direct density gathers: {name}", flush=True)
    cfg = CASES[name]
    ob = O.AtomicBasis(cfg["Z"], cfg["lmax"], cfg["mmax"], cfg["nelem"])
    ob.compute_tei()
    P = random_symmetric(ob.Nbf, seed=1006)
    Pd = torch.as_tensor(P, device="cuda")
    out = {}
    for budget in ("0", None):
        if budget is None:
            monkeypatch.delenv("HFGPU_EXCHANGE_TILE_GIB", raising=False)
        else:
            monkeypatch.setenv("HFGPU_EXCHANGE_TILE_GIB", budget)
        hb = H.AtomicBasis(cfg["Z"], cfg["lmax"], cfg["mmax"], cfg["nelem"], ctx=gpu_ctx)
        hb.compute_tei()
        out[budget] = (hb.exchange(Pd).cpu().numpy(), hb.exchange(Pd, precision="fp32").cpu().numpy())
    torch.cuda.synchronize()
    Kr = ob.exchange(P)
    assert relerr(out["0"][0], Kr) <= TOL
    assert relerr(out[None][0], Kr) <= TOL
    assert relerr(out["0"][1], out[None][1]) <= 1e-6


def test_gpu_uhf_hydrogen(gpu_ctx):
    """atomic-H-hf-u: one alpha electron; the exchange cancels the self-Coulomb
    exactly, E = -0.5 Eh to the basis limit (tests/refs/ci.json)."""
    from helfem_amd.scf import uhf
    print("[gpu] H UHF SCF", flush=True)
    hb = H.AtomicBasis(1, 0, 0, 5, ctx=gpu_ctx)
    hb.compute_tei()
    r = uhf(hb, 1, 0, conv=1e-12)
    assert abs(r.E - REF["atomic-H-hf-u"]["total"]) < 1e-9, r.E
    assert abs(r.Ecoul + r.Exx) < 1e-10


def test_one_electron_on_device(gpu_ctx):
    """hf_atomic_one_electron_dev assembles S, T, V on the device from their
    angular blocks: equal to the host matrices, which equal the oracle's."""
    import torch
    print("[gpu] atomic one-electron matrices on the device", flush=True)
    ob, hb = pair("Kr-l3m2-4el", gpu_ctx)
    for w, host in ((0, hb.overlap()), (1, hb.kinetic()), (2, hb.nuclear())):
        dev = hb.one_electron_dev(w).cpu().numpy()
        torch.cuda.synchronize()
        assert np.array_equal(dev, host), w
    assert np.array_equal(hb.overlap(), ob.overlap())
