"""Batched atomic Fock builds (hf_atomic_batch): every kernel runs once for
all systems, the launch sequence is replayed from a HIP graph, and the
exchange workgroup shape is autotuned.  Each system's J and K must equal its
single-system build (the same kernels; the workgroup shape only reorders the
final sum over multipoles: 1e-14) and the oracle (1e-12,
src/atomic/TwoDBasis.cpp:774-999); FP32 systems in a mixed batch equal their
single-system FP32 build to rounding."""
import numpy as np
import pytest

import helfem_amd as H
import oracle as O
from helfem_amd.synthetic import random_symmetric, seeded_density

pytestmark = pytest.mark.gpu
TOL = 1e-12

# (Z, lmax, mmax, nelem, nnodes): spherical to non-spherical, LIP-10 and LIP-15
SYSTEMS = [(2, 0, 0, 4, 10), (10, 1, 1, 5, 15), (18, 2, 2, 3, 10), (36, 3, 2, 4, 15), (7, 4, 4, 6, 10),
           (54, 2, 1, 5, 15)]


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def bases(gpu_ctx):
    out = []
    for Z, lmax, mmax, nel, nn in SYSTEMS:
        b = H.AtomicBasis(Z, lmax, mmax, nel, nnodes=nn, ctx=gpu_ctx)
        b.compute_tei()
        out.append(b)
    return out


def _densities(bases, symmetric=True, seed=4000):
    import torch
    Ps = []
    for q, b in enumerate(bases):
        P = random_symmetric(b.Nbf, seed=seed + q)
        if not symmetric:
            P = P + 0.3 * np.triu(random_symmetric(b.Nbf, seed=seed + 100 + q))
        Ps.append(torch.as_tensor(np.ascontiguousarray(P), device="cuda"))
    return Ps


@pytest.mark.parametrize("symmetric", [True, False])
def test_batch_equals_single_system(bases, gpu_ctx, symmetric):
    import torch
    print(f"[gpu] batched atomic J/K vs single-system builds (symmetric={symmetric})", flush=True)
    Ps = _densities(bases, symmetric)
    bt = H.AtomicBatch(bases, gpu_ctx)
    J = bt.coulomb(Ps, [torch.empty_like(P) for P in Ps])
    K = bt.exchange(Ps, [torch.empty_like(P) for P in Ps])
    for b, P, Jb, Kb in zip(bases, Ps, J, K):
        Js, Ks = b.coulomb(P), b.exchange(P)
        assert torch.equal(Jb, Js)
        assert relerr(Kb.cpu().numpy(), Ks.cpu().numpy()) <= 1e-14
    nw, ms = bt.tuning()
    assert set(nw.ravel().tolist()) <= {0, 1, 2, 4}
    assert np.any(nw > 0)


def test_batch_against_oracle(bases, gpu_ctx):
    import torch
    print("[gpu] batched atomic J/K vs oracle", flush=True)
    Ps = _densities(bases, True, seed=4100)
    bt = H.AtomicBatch(bases, gpu_ctx)
    J = bt.coulomb(Ps, [torch.empty_like(P) for P in Ps])
    K = bt.exchange(Ps, [torch.empty_like(P) for P in Ps])
    for (Z, lmax, mmax, nel, nn), P, Jb, Kb in list(zip(SYSTEMS, Ps, J, K))[2:5]:
        ob = O.AtomicBasis(Z, lmax, mmax, nel, nnodes=nn)
        ob.compute_tei()
        Pn = P.cpu().numpy()
        assert relerr(Jb.cpu().numpy(), ob.coulomb(Pn)) <= TOL
        assert relerr(Kb.cpu().numpy(), ob.exchange(Pn)) <= TOL


def test_graph_replay_and_new_buffers(bases, gpu_ctx):
    """Repeated calls replay the captured launches; new tensors (other device
    pointers, same shapes) and a changed screening pattern are picked up."""
    import torch
    print("[gpu] batched atomic J/K graph replays", flush=True)
    bt = H.AtomicBatch(bases, gpu_ctx)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        Ps = _densities(bases, True, seed=4200)
        ref_K = [b.exchange(P) for b, P in zip(bases, Ps)]
        ref_J = [b.coulomb(P) for b, P in zip(bases, Ps)]
        for rep in range(3):
            Pr = [P.clone() for P in Ps]  # new buffers each replay
            J = bt.coulomb(Pr, [torch.empty_like(P) for P in Pr])
            K = bt.exchange(Pr, [torch.empty_like(P) for P in Pr])
            torch.cuda.synchronize()
            for a, b_ in zip(J, ref_J):
                assert torch.equal(a, b_)
            for a, b_ in zip(K, ref_K):
                assert relerr(a.cpu().numpy(), b_.cpu().numpy()) <= 1e-14
        # an m-block-diagonal density screens out most blocks: new lists, new graph
        Pm = []
        for b in bases:
            S = b.overlap()
            blocks = b.get_sym_idx(1)
            Pm.append(torch.as_tensor(seeded_density(S, blocks, [1] * len(blocks), seed=4300), device="cuda"))
        K = bt.exchange(Pm, [torch.empty_like(P) for P in Pm])
        torch.cuda.synchronize()
        for b, P, Kb in zip(bases, Pm, K):
            assert relerr(Kb.cpu().numpy(), b.exchange(P).cpu().numpy()) <= 1e-14


def test_mixed_precision_batch(bases, gpu_ctx):
    import torch
    print("[gpu] batched atomic exchange with FP32 and FP64 systems", flush=True)
    Ps = _densities(bases, True, seed=4400)
    bt = H.AtomicBatch(bases, gpu_ctx)
    fp32 = [q % 2 == 1 for q in range(len(bases))]
    K = bt.exchange(Ps, [torch.empty_like(P) for P in Ps], fp32=fp32)
    for b, P, Kb, f in zip(bases, Ps, K, fp32):
        Ks = b.exchange(P, precision="fp32" if f else "fp64")
        e = relerr(Kb.cpu().numpy(), Ks.cpu().numpy())
        assert e <= (1e-6 if f else 1e-14), (f, e)


def test_clone_holds_a_second_density(bases, gpu_ctx):
    """A clone of a basis carries its own device tables, so the alpha and beta
    densities of one basis can be built in the same batch."""
    import torch
    print("[gpu] batched exchange of a basis and its clone", flush=True)
    b = bases[3]
    c = b.clone()
    Pa = torch.as_tensor(random_symmetric(b.Nbf, seed=4500), device="cuda")
    Pb = torch.as_tensor(random_symmetric(b.Nbf, seed=4501), device="cuda")
    bt = H.AtomicBatch([b, c], gpu_ctx)
    Ka, Kb = bt.exchange([Pa, Pb], [torch.empty_like(Pa), torch.empty_like(Pb)])
    assert relerr(Ka.cpu().numpy(), b.exchange(Pa).cpu().numpy()) <= 1e-14
    assert relerr(Kb.cpu().numpy(), b.exchange(Pb).cpu().numpy()) <= 1e-14
