"""GPU parity of the range-separated exchange (hf_atomic_rs_exchange,
TwoDBasisT::rs_exchange, src/atomic/TwoDBasis.cpp:1002-1131) against the
oracle, whose kernels are pinned in test_oracle_rsx.py (closed-form hydrogen
self-exchange, mpmath erfc multipoles, the small-parameter Coulomb limit).
Bar: FP64 relative error <= 1e-12 on seeded densities; the GPU hydrogen
self-exchange equals the closed form as the oracle's does."""
import numpy as np
import pytest
from scipy.integrate import quad

import helfem_amd as H
import oracle as O
from helfem_amd.synthetic import seeded_density

pytestmark = pytest.mark.gpu
TOL = 1e-12


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = {
    "He-l0-5el": dict(Z=2, lmax=0, mmax=0, nelem=5, nquad=30),
    "Ne-l2-3el": dict(Z=10, lmax=2, mmax=2, nelem=3, nquad=20),
}


@pytest.mark.parametrize("kind", ["yukawa", "erfc"])
@pytest.mark.parametrize("name", list(CASES))
def test_rs_exchange_matches_oracle(name, kind, gpu_ctx):
    print(f"[gpu] range-separated exchange {kind}: {name}", flush=True)
    c = CASES[name]
    ob = O.AtomicBasis(c["Z"], c["lmax"], c["mmax"], c["nelem"], nquad=c["nquad"])
    hb = H.AtomicBasis(c["Z"], c["lmax"], c["mmax"], c["nelem"], nquad=c["nquad"], ctx=gpu_ctx)
    ob.compute_tei()
    hb.compute_tei()
    S = ob.overlap()
    blocks = ob.sym_idx(1)
    P = seeded_density(S, blocks, [1] * len(blocks), seed=1501)
    if kind == "yukawa":
        ob.compute_yukawa(0.4)
        hb.compute_yukawa(0.4)
    else:
        ob.compute_erfc(0.4)
        hb.compute_erfc(0.4)
    Kr = ob.rs_exchange(P)
    Kg = hb.rs_exchange(P)
    assert relerr(Kg, Kr) <= TOL, relerr(Kg, Kr)
    # the plain exchange of the same handle is unaffected
    assert relerr(hb.exchange(P), ob.exchange(P)) <= TOL


@pytest.mark.parametrize("kind", ["yukawa", "erfc"])
def test_rs_exchange_general_density(kind, gpu_ctx):
    """A non-symmetric density takes the general (all element pairs) path."""
    print(f"[gpu] range-separated exchange {kind}, non-symmetric P", flush=True)
    c = CASES["Ne-l2-3el"]
    ob = O.AtomicBasis(c["Z"], c["lmax"], c["mmax"], c["nelem"], nquad=c["nquad"])
    hb = H.AtomicBasis(c["Z"], c["lmax"], c["mmax"], c["nelem"], nquad=c["nquad"], ctx=gpu_ctx)
    ob.compute_tei()
    hb.compute_tei()
    getattr(ob, "compute_" + kind)(0.4)
    getattr(hb, "compute_" + kind)(0.4)
    P = np.random.default_rng(1502).standard_normal((ob.Nbf, ob.Nbf))
    assert relerr(hb.rs_exchange(P), ob.rs_exchange(P)) <= TOL


def test_rs_exchange_hydrogen_closed_form(gpu_ctx):
    """Yukawa self-exchange of the hydrogen 1s density through the GPU path."""
    from helfem_amd.scf import uhf
    print("[gpu] Yukawa hydrogen self-exchange", flush=True)
    hb = H.AtomicBasis(1, 0, 0, 5, ctx=gpu_ctx)
    hb.compute_tei()
    r = uhf(hb, 1, 0, conv=1e-12)
    P = r.Pa.cpu().numpy()
    lam = 0.3
    hb.compute_yukawa(lam)
    j = -float(np.sum(P * hb.rs_exchange(P)))
    f = lambda k: k * k * (16.0 / (4.0 + k * k) ** 2) ** 2 * 4 * np.pi / (k * k + lam * lam)
    ref = quad(f, 0, np.inf, limit=400, epsabs=1e-16, epsrel=1e-14)[0] / (2 * np.pi ** 2)
    assert abs(j - ref) < 1e-11, (j, ref)


def test_rs_exchange_requires_caches(gpu_ctx):
    hb = H.AtomicBasis(2, 0, 0, 3, ctx=gpu_ctx)
    hb.compute_tei()
    with pytest.raises(H.HfError):
        hb.rs_exchange(np.eye(hb.Nbf))
