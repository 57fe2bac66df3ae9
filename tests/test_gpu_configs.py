"""GPU parity at the full sizes of the benchmark workloads.

* N2 diatomic HF (R = 2.074 bohr, 8 mu-elements, LIP-15, lmax 16 for
  m = 0..3, Nbf 11894): the P/Q Legendre tiles span 1e-56 .. 1e75 at L = 34.
  Full J against the oracle's coulomb (src/diatomic/basis.cpp:1627-1816) and
  K on 128 output blocks against the oracle's exchange over those blocks
  (basis.cpp:1818-2089), for the seeded closed-shell density (symmetric: the
  mirrored exchange runs) and a non-symmetric density (the general one).
* Xe atomic (lmax = mmax = 6, 25 elements, Nbf 17101): the same J and
  sampled-K checks (src/atomic/TwoDBasis.cpp:774-999), the FP32 exchange's
  error against the oracle, and the LDA matrix against the oracle's values
  recorded by tests/golden/make_xe_lda_fixture.py (dftgrid.cpp:591-653).
* Ne HF (lmax = mmax = 1, 10 elements): the SCF energy against the oracle's
  SCF (1e-10 Eh) and the reference's -128.5470981094 (tests/README.md:80).

Bar: FP64 max|X_gpu - X_oracle| / max|X_oracle| <= 1e-12 over the compared
entries; the FP32 exchange within 2e-6.
"""
from pathlib import Path

import numpy as np
import pytest

import helfem_amd as H
import oracle as O
from helfem_amd.synthetic import seeded_density, uniform

pytestmark = pytest.mark.gpu
TOL = 1e-12
TOL32 = 2e-6
GOLDEN = Path(__file__).parent / "golden"

N2 = dict(Z1=7, Z2=7, R=2.074, lmm=[16, 16, 16, 16], nel=8, seed=1002, occ={0: 5, 1: 1, -1: 1})
XE = dict(Z=54, lmax=6, mmax=6, nel=25, seed=1003, occ_l={0: 5, 1: 4, 2: 2})
_cache = {}


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _n2(ctx):
    if "n2" not in _cache:
        ob = O.DiatomicBasis(N2["Z1"], N2["Z2"], N2["R"], N2["lmm"], N2["nel"])
        ob.compute_tei()
        hb = H.DiatomicBasis(N2["Z1"], N2["Z2"], N2["R"], N2["lmm"], N2["nel"], ctx=ctx)
        hb.compute_tei()
        S = ob.overlap()
        ms = sorted(set(int(m) for m in ob.mval))
        P = seeded_density(S, ob.sym_idx(1), [N2["occ"].get(m, 0) for m in ms], seed=N2["seed"])
        shell = np.asarray(hb.pure_idx) // hb.Nrad
        _cache["n2"] = (ob, hb, P, shell)
    return _cache["n2"]


def _xe(ctx):
    if "xe" not in _cache:
        ob = O.AtomicBasis(XE["Z"], XE["lmax"], XE["mmax"], XE["nel"])
        ob.compute_tei()
        hb = H.AtomicBasis(XE["Z"], XE["lmax"], XE["mmax"], XE["nel"], ctx=ctx)
        hb.compute_tei()
        S = ob.overlap()
        ms = sorted(set(int(m) for m in ob.mval))
        occ = [sum(c for l, c in XE["occ_l"].items() if abs(m) <= l) for m in ms]
        P = seeded_density(S, ob.sym_idx(1), occ, seed=XE["seed"])
        shell = np.arange(hb.Nbf) // hb.Nrad
        _cache["xe"] = (ob, hb, P, shell)
    return _cache["xe"]


def _nonsymmetric(P, blocks, seed):
    """P plus a non-symmetric rank-2 term inside every symmetry block of
    comparable size, so the general (unmirrored) exchange runs."""
    G = P.copy()
    for q, b in enumerate(blocks):
        b = np.asarray(b)
        X = uniform(seed + 31 * q, 4 * b.size).reshape(4, b.size)
        A = np.outer(X[0], X[1]) + np.outer(X[2], X[3])
        scale = np.max(np.abs(P[np.ix_(b, b)]))
        if scale > 0:
            G[np.ix_(b, b)] += 0.5 * scale / np.max(np.abs(A)) * A
    return G


def _sample_blocks(mval, count):
    """`count` output blocks j*Nang + k spread over the m-diagonal ones (the
    only blocks an m-block-diagonal density couples)."""
    na = len(mval)
    diag = [j * na + k for j in range(na) for k in range(na) if mval[j] == mval[k]]
    pick = np.linspace(0, len(diag) - 1, count).round().astype(int)
    return np.unique(np.asarray(diag)[pick]).astype(np.int32)


def _block_entries(M, shell, jk, na):
    out = []
    for q in jk:
        j, k = divmod(int(q), na)
        out.append(M[np.ix_(shell == j, shell == k)].ravel())
    return np.concatenate(out)


def _coulomb_and_sampled_exchange(ob, hb, shell, P):
    import torch
    Pd = torch.as_tensor(np.ascontiguousarray(P), device="cuda")
    Jg = hb.coulomb(Pd).cpu().numpy()
    Kg = hb.exchange(0.5 * Pd).cpu().numpy()
    torch.cuda.synchronize()
    del Pd
    assert relerr(Jg, ob.coulomb(P)) <= TOL
    mval = np.asarray(hb.mval)
    jk = _sample_blocks(mval, 128)
    assert jk.size >= 100
    Kr = ob.exchange_blocks(0.5 * P, jk)
    eg, er = _block_entries(Kg, shell, jk, hb.Nang), _block_entries(Kr, shell, jk, hb.Nang)
    assert np.max(np.abs(er)) > 0
    err = relerr(eg, er)
    # blocks the density cannot couple stay exactly zero
    off = mval[shell][:, None] != mval[shell][None, :]
    assert not np.any(Kg[off])
    return err


@pytest.mark.parametrize("density", ["seeded", "nonsymmetric"])
def test_n2_full_size_coulomb_exchange(density, gpu_ctx):
    print(f"[gpu] N2 lmax 16, 8 elements: J and sampled K vs oracle ({density} density)", flush=True)
    ob, hb, P, shell = _n2(gpu_ctx)
    if density == "nonsymmetric":
        P = _nonsymmetric(P, ob.sym_idx(1), seed=2002)
    err = _coulomb_and_sampled_exchange(ob, hb, shell, P)
    print(f"[gpu]   sampled K rel err {err:.2e}", flush=True)
    assert err <= TOL


@pytest.mark.parametrize("density", ["seeded", "nonsymmetric"])
def test_xe_full_size_coulomb_exchange(density, gpu_ctx):
    print(f"[gpu] Xe lmax=mmax=6, 25 elements: J and sampled K vs oracle ({density} density)", flush=True)
    ob, hb, P, shell = _xe(gpu_ctx)
    if density == "nonsymmetric":
        P = _nonsymmetric(P, ob.sym_idx(1), seed=2003)
    err = _coulomb_and_sampled_exchange(ob, hb, shell, P)
    print(f"[gpu]   sampled K rel err {err:.2e}", flush=True)
    assert err <= TOL


def test_n2_device_setup_coulomb_exchange(gpu_ctx):
    """The N2 two-electron setup with the two-channel kernels and their
    signed Cholesky factors of all 224 x 8 ((L,|M|), element) pairs built on
    the device: J and the sampled K from those caches meet the same bar
    against the oracle."""
    import time
    print("[gpu] N2 device two-electron setup: J and sampled K vs oracle", flush=True)
    ob, hb, P, shell = _n2(gpu_ctx)
    t = time.perf_counter()
    db = H.DiatomicBasis(N2["Z1"], N2["Z2"], N2["R"], N2["lmm"], N2["nel"], ctx=gpu_ctx)
    dev = db.compute_tei(device=True)
    total = time.perf_counter() - t
    err = _coulomb_and_sampled_exchange(ob, db, shell, P)
    print(f"[gpu]   device setup {dev:.3f} s ({total:.3f} s with the basis); sampled K rel err {err:.2e}",
          flush=True)
    assert err <= TOL


def test_xe_device_setup_coulomb_exchange(gpu_ctx):
    """The Xe two-electron setup built on the device (Slater tensors and
    Cholesky factors of all 13 x 25 (L, element) pairs): J and the sampled K
    from those caches meet the same bar against the oracle."""
    import time
    print("[gpu] Xe device two-electron setup: J and sampled K vs oracle", flush=True)
    ob, hb, P, shell = _xe(gpu_ctx)
    t = time.perf_counter()
    db = H.AtomicBasis(XE["Z"], XE["lmax"], XE["mmax"], XE["nel"], ctx=gpu_ctx)
    dev = db.compute_tei(device=True)
    total = time.perf_counter() - t
    err = _coulomb_and_sampled_exchange(ob, db, shell, P)
    print(f"[gpu]   device setup {dev:.3f} s ({total:.3f} s with the basis); sampled K rel err {err:.2e}",
          flush=True)
    assert err <= TOL


def test_xe_full_size_fp32_exchange_error(gpu_ctx):
    """The reduced-precision exchange at the Xe size, stated against the
    oracle on the sampled blocks."""
    import torch
    print("[gpu] Xe FP32 exchange vs oracle (sampled blocks)", flush=True)
    ob, hb, P, shell = _xe(gpu_ctx)
    Pd = torch.as_tensor(np.ascontiguousarray(0.5 * P), device="cuda")
    K32 = hb.exchange(Pd, precision="fp32").cpu().numpy()
    torch.cuda.synchronize()
    jk = _sample_blocks(np.asarray(hb.mval), 128)
    Kr = ob.exchange_blocks(0.5 * P, jk)
    err = relerr(_block_entries(K32, shell, jk, hb.Nang), _block_entries(Kr, shell, jk, hb.Nang))
    print(f"[gpu]   FP32 sampled K rel err {err:.2e}", flush=True)
    assert 1e-12 < err <= TOL32


def test_xe_full_size_lda(gpu_ctx):
    """LDA (Slater + VWN5) on the Xe grid at ldft = mdft = 36, 25 elements
    (97,200 points per element), against the oracle values recorded in
    tests/golden/xe_lda_l6m6_25el.npz."""
    import torch
    from helfem_amd.dftgrid import DFTGrid
    print("[gpu] Xe LDA Vxc, Exc, Nel at full size vs recorded oracle values", flush=True)
    ref = np.load(GOLDEN / "xe_lda_l6m6_25el.npz")
    ob, hb, P, shell = _xe(gpu_ctx)
    assert abs(float(np.sum(P * ob.overlap())) - float(ref["trPS"])) < 1e-12 * 54
    g = DFTGrid(hb, lang=36, mang=36)
    Pd = torch.as_tensor(np.ascontiguousarray(P), device="cuda")
    Vd, exc, nel, _ = g.eval_Fxc(1, 7, Pd)
    V = Vd.cpu().numpy()
    torch.cuda.synchronize()
    vmax = float(ref["vmax"])
    assert abs(exc - float(ref["exc"])) <= TOL * abs(float(ref["exc"])), (exc, float(ref["exc"]))
    assert abs(nel - float(ref["nel"])) <= TOL * 54, (nel, float(ref["nel"]))
    assert np.max(np.abs(np.diagonal(V) - ref["diag"])) <= TOL * vmax
    assert np.max(np.abs(V.sum(axis=1) - ref["rowsum"])) <= TOL * vmax * np.sqrt(hb.Nbf)
    W = uniform(XE["seed"] + 17, 4 * hb.Nbf).reshape(4, hb.Nbf)
    probes = (W @ V) @ W.T
    assert np.max(np.abs(probes - ref["probes"])) <= TOL * np.max(np.abs(ref["probes"])) * 10


def test_ne_scf_ten_elements(gpu_ctx):
    """Ne HF at lmax = mmax = 1 on the 10-element grid: the GPU SCF equals the
    oracle's SCF to 1e-10 Eh and the reference value -128.5470981094
    (tests/README.md:79-80) to 1e-9 Eh."""
    from helfem_amd.scf import rhf
    print("[gpu] Ne HF SCF, lmax=mmax=1, 10 elements", flush=True)
    hb = H.AtomicBasis(10, 1, 1, 10, ctx=gpu_ctx)
    hb.compute_tei()
    r = rhf(hb, 10)
    ob = O.AtomicBasis(10, 1, 1, 10)
    ob.compute_tei()
    ro = O.scf_rhf(ob, 10)
    assert abs(r.E - ro["E"]) < 1e-10, (r.E, ro["E"])
    assert abs(r.E + 128.5470981094) < 1e-9, r.E


def test_atoms_batch_distribution_cohort_equals_alone(gpu_ctx):
    """32 atoms drawn from the batched-atoms distribution (Z 1..86, 4..24
    elements of order 10 or 15, lmax = mmax 0..4): the mixed-precision cohort
    gives every atom the energy it gets alone to 1e-9 Eh (both converged to
    dE < 1e-10 Eh; the batched eigensolves pad differently, so the two runs
    take different rounding paths to the same state) and the FP64 cohort's
    energy to 1e-8 Eh."""
    from helfem_amd import batch
    from helfem_amd.cohort import Cohort
    print("[gpu] 32 random atoms: cohort vs alone", flush=True)
    systems = batch.random_atoms(32, seed=1004, ctx=gpu_ctx, zmax=86, nel_range=(4, 24), orders=(10, 15),
                                 lmax_range=(0, 4), setup_threads=16)
    occ = [s.occ for s in systems]
    mixed = Cohort(systems, gpu_ctx, precision="mixed", conv=1e-10, occupations=occ).solve()
    fp64 = Cohort(systems, gpu_ctx, precision="fp64", conv=1e-10, occupations=occ).solve()
    for s, a, b in zip(systems, mixed, fp64):
        assert a["converged"] and b["converged"], s.info
        assert abs(a["E"] - b["E"]) < 1e-8, (s.info, a["E"], b["E"])
    for s, a in zip(systems, mixed):
        alone = Cohort([s], gpu_ctx, precision="mixed", conv=1e-10, occupations=[s.occ]).solve()[0]
        assert abs(a["E"] - alone["E"]) < 1e-9, (s.info, a["E"], alone["E"])


def test_served_diatomic_distribution_equals_alone(gpu_ctx):
    """Six requests from the served diatomic stream's distribution (Z1, Z2
    1..18, R 1-6 bohr, 4..12 mu-elements, lmax 8..24 for m 0..mmax (0..3),
    HF or LDA), Poisson arrivals, continuous batching with the requests'
    Fock builds overlapping on the device: each served energy equals the
    request solved alone to 1e-8 Eh (both converged to dE < 1e-11 Eh and an
    orbital gradient below 1e-7: at 1e-6 the open-shell LDA requests, whose
    soft orbital rotations leave an energy error of order g^2 / 3e-5, still
    differed by 3e-8 Eh)."""
    from helfem_amd import batch
    print("[gpu] six served diatomic requests vs alone", flush=True)
    specs = batch.random_diatomic_specs(6, seed=1005)
    assert {s["lda"] for s in specs} == {True, False}
    res, wall = batch.serve(specs, rate_hz=2.0, seed=1006, ctx=gpu_ctx, conv=1e-11, setup_threads=16, gconv=1e-7)
    for spec, r in zip(specs, res):
        print(f"[gpu]   {spec} latency {r['latency_s']:.1f}s E {r['E']:.10f} in {r['iterations']} it", flush=True)
        alone, _ = batch.solve_sequential([batch.build_diatomic(spec, gpu_ctx)], conv=1e-11, gconv=1e-7)
        assert abs(r["E"] - alone[0].E) < 1e-8, (spec, r["E"], alone[0].E)


def test_xe_lda_configured_grid_variant(gpu_ctx):
    """The Xe LDA grid of the workload description (Gauss-Legendre cos(theta)
    x Gauss-Lobatto radial, same sizes) beside the reference's rule (Chebyshev
    x Chebyshev): both integrate the density to 54 electrons and their Exc
    differ by the quadrature error, stated here (2.5e-6 Eh on the seeded
    density; bound 1e-5 Eh)."""
    import torch
    from helfem_amd.dftgrid import DFTGrid
    print("[gpu] Xe LDA on the Gauss-Legendre x Gauss-Lobatto grid vs the reference rule", flush=True)
    ob, hb, P, shell = _xe(gpu_ctx)
    Pd = torch.as_tensor(np.ascontiguousarray(P), device="cuda")
    out = {}
    for rule in (0, 1):
        g = DFTGrid(hb, lang=36, mang=36, rule=rule)
        _, exc, nel, _ = g.eval_Fxc(1, 7, Pd)
        out[rule] = (exc, nel)
    torch.cuda.synchronize()
    ref = np.load(GOLDEN / "xe_lda_l6m6_25el.npz")
    assert abs(out[0][0] - float(ref["exc"])) <= TOL * abs(float(ref["exc"]))
    print(f"[gpu]   Exc reference rule {out[0][0]:.12f}, variant {out[1][0]:.12f}, "
          f"difference {out[1][0] - out[0][0]:.3e} Eh; Nel {out[0][1]:.10f} / {out[1][1]:.10f}", flush=True)
    for exc, nel in out.values():
        assert abs(nel - 54.0) < 1e-6
    assert abs(out[1][0] - out[0][0]) < 1e-5
