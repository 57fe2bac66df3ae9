"""Host logic of the device SCF driver (helfem_amd/scf.py) on the CPU: the
Fock build is delegated to the oracle (J/K of TwoDBasis.cpp:774-999 in
plain C++), the matrices are CPU torch tensors, and the driver's own parts
are checked -- symmetry-block eigensolves with the shell-group reduction,
aufbau, DIIS, energy expression and the lockstep engine -- against the
oracle's independent SCF and the reference energies."""
import json
from pathlib import Path

import numpy as np
import pytest

import helfem_amd as H
import oracle as O

torch = pytest.importorskip("torch")
REF = json.loads((Path(__file__).parent / "golden" / "reference_energies.json").read_text())


class OracleFock:
    """A host-built product basis whose J/K come from the oracle."""

    def __init__(self, Z, lmax, nelem):
        self.h = H.AtomicBasis(Z, lmax, lmax, nelem, ctx=None)
        self.o = O.AtomicBasis(Z, lmax, lmax, nelem)
        self.o.compute_tei()
        self.lval, self.mval, self.Nrad, self.Nbf = self.h.lval, self.h.mval, self.h.Nrad, self.h.Nbf

    def overlap(self):
        return self.h.overlap()

    def kinetic(self):
        return self.h.kinetic()

    def nuclear(self):
        return self.h.nuclear()

    def get_sym_idx(self, symm=1):
        return self.h.get_sym_idx(symm)

    def coulomb(self, P):
        return torch.as_tensor(self.o.coulomb(P.numpy()))

    def exchange(self, P, precision="fp64"):
        return torch.as_tensor(self.o.exchange(P.numpy()))


@pytest.fixture(scope="module")
def systems():
    return {name: OracleFock(*args) for name, args in
            (("He", (2, 0, 5)), ("Be", (4, 0, 5)), ("Ne", (10, 1, 5)))}


@pytest.mark.parametrize("name,Z,key", [("He", 2, "atomic-He-hf-r"), ("Be", 4, "atomic-Be-hf-r")])
def test_rhf_energy_on_cpu(systems, name, Z, key):
    from helfem_amd.scf import rhf
    r = rhf(systems[name], Z, device="cpu", conv=1e-11)
    assert abs(r.E - REF[key]["total"]) < 1e-9, r.E


def test_ne_rhf_matches_oracle_scf(systems):
    """Ne at lmax 1 (the 2p shell splits the blocks by m): the driver's
    symmetry reductions reproduce the oracle's plain per-block SCF."""
    from helfem_amd.scf import rhf
    r = rhf(systems["Ne"], 10, device="cpu", conv=1e-11)
    ro = O.scf_rhf(systems["Ne"].o, 10)
    assert abs(r.E - ro["E"]) < 1e-10
    assert abs(r.E + 128.5470981094) < 1e-9
    assert r.reused_solves > 0  # identical shell groups were solved once


def test_engine_lockstep_equals_single(systems):
    """Three states advanced together by one engine (shared batched
    eigensolves) give the same energies as each alone."""
    from helfem_amd.scf import SCFEngine, SCFState, rhf
    alone = {n: rhf(systems[n], Z, device="cpu", conv=1e-11).E for n, Z in (("He", 2), ("Be", 4), ("Ne", 10))}
    eng = SCFEngine()
    states = {n: SCFState(systems[n], (Z // 2,), True, conv=1e-11, device="cpu", tag=n)
              for n, Z in (("He", 2), ("Be", 4), ("Ne", 10))}
    for st in states.values():
        eng.add(st)
    finished = []
    while eng.active:
        finished += eng.step()
    assert sorted(st.tag for st in finished) == ["Be", "He", "Ne"]
    for n, st in states.items():
        assert abs(st.result().E - alone[n]) < 1e-12, n


def test_uhf_hydrogen_on_cpu():
    from helfem_amd.scf import uhf
    r = uhf(OracleFock(1, 0, 5), 1, 0, device="cpu", conv=1e-12)
    assert abs(r.E - REF["atomic-H-hf-u"]["total"]) < 1e-9
    assert abs(r.Ecoul + r.Exx) < 1e-10


def test_weyl_skip_of_unoccupied_blocks(monkeypatch):
    """Ne at lmax 2: the d blocks hold no occupied orbital, so once their
    lowest level minus the change of their matrix stays above the highest
    occupied level they are not diagonalised; the energy equals the SCF that
    solves every block (GAP = inf switches the bound off), and the bound is
    not trusted blindly: UHF with an empty beta channel skips everything of
    it after the first iteration and still reproduces the full solve."""
    from helfem_amd import scf
    from helfem_amd.scf import rhf, uhf
    ne = OracleFock(10, 2, 4)
    r = rhf(ne, 10, device="cpu", conv=1e-11)
    assert r.skipped_solves > 0
    monkeypatch.setattr(scf._Blocks, "GAP", np.inf)
    full = rhf(ne, 10, device="cpu", conv=1e-11)
    assert full.skipped_solves == 0
    assert abs(r.E - full.E) < 1e-10
    assert r.iterations == full.iterations
    monkeypatch.setattr(scf._Blocks, "GAP", 0.05)
    h = OracleFock(1, 1, 4)
    u = uhf(h, 1, 0, device="cpu", conv=1e-12)
    monkeypatch.setattr(scf._Blocks, "GAP", np.inf)
    uf = uhf(h, 1, 0, device="cpu", conv=1e-12)
    assert u.skipped_solves > 0 and uf.skipped_solves == 0
    assert abs(u.E - uf.E) < 1e-10


def test_blockwise_invh_equals_full():
    """The half inverse taken over the shell groups S couples (with shared
    eigensolves for identical groups) equals invh of the whole symmetry block
    (libhelfem/src/utils.cpp:121-159) and orthonormalises it."""
    from helfem_amd.scf import _Blocks, _invh, _one_electron
    ne = OracleFock(10, 2, 4)
    S, _, _ = _one_electron(ne, torch.device("cpu"))
    bs = _Blocks(ne, S, 1, torch.device("cpu"))
    for Sk, Xk in zip(bs.S, bs.X):
        full = _invh(Sk)
        assert float(torch.max(torch.abs(Xk - full))) <= 1e-10 * float(torch.max(torch.abs(full)))
        eye = torch.eye(Sk.shape[0], dtype=Sk.dtype)
        assert float(torch.max(torch.abs(Xk.T @ Sk @ Xk - eye))) <= 1e-12


def test_padded_batched_eigh():
    """Mixed sizes share padded batched eigensolves: the eigenvalues equal
    each matrix's own and the eigenvectors span the same vectors (up to sign),
    also for a spectrum spanning 1e5."""
    from helfem_amd.scf import batched_eigh
    g = torch.Generator().manual_seed(7)
    mats = []
    for n in (40, 37, 31, 40, 12, 29):
        A = torch.randn(n, n, generator=g, dtype=torch.float64)
        A = A + A.T
        A[0, 0] += 1e5
        mats.append(A)
    for A, (e, v) in zip(mats, batched_eigh(mats)):
        e0, v0 = torch.linalg.eigh(A)
        assert v.shape == A.shape and len(e) == A.shape[0]
        assert np.max(np.abs(e - e0.numpy())) <= 1e-10
        s = torch.sign((v * v0).sum(0))
        assert float(torch.max(torch.abs(v * s - v0))) <= 1e-10


def test_frozen_occupations_in_aufbau():
    """finish() with per-block orbital counts puts exactly that many of each
    block's lowest orbitals in the density; with the aufbau's own counts it
    reproduces the aufbau density."""
    from helfem_amd.scf import _Blocks, _one_electron
    be = OracleFock(4, 1, 4)
    S, T, V = _one_electron(be, torch.device("cpu"))
    H = T + V
    bs = _Blocks(be, S, 1, torch.device("cpu"))
    Hb = bs.blocks_of(H)
    P0, Pb0, nk0 = bs.density(Hb, 2, 2.0, H)
    assert sum(nk0) == 2
    plan = bs.prepare(Hb, 1)
    from helfem_amd.scf import batched_eigh
    sols = batched_eigh(plan["mats"])
    P1, _, nk1 = bs.finish(plan, sols, 2, 2.0, H, frozen=list(nk0))
    assert nk1 == nk0 and float(torch.max(torch.abs(P1 - P0))) < 1e-12
    want = [1 if k == 0 else 0 for k in range(len(nk0))]
    want[len(nk0) // 2] += 1
    plan = bs.prepare(Hb, 2)
    P2, _, nk2 = bs.finish(plan, batched_eigh(plan["mats"]), 2, 2.0, H, frozen=want)
    assert nk2 == want
    Sd = S.numpy()
    assert abs(float(np.sum(P2.numpy() * Sd)) - 4.0) < 1e-10


def test_fractional_frozen_occupations():
    """a fractional count occupies the block's last orbital by its fraction:
    0.5 in each of m = -1 and +1 gives the average of the two integer
    densities (the cylindrically averaged open shell of the reference's
    --symmetry=3 runs), with the electron count kept"""
    from helfem_amd.scf import _Blocks, _one_electron, batched_eigh
    be = OracleFock(4, 1, 4)
    S, T, V = _one_electron(be, torch.device("cpu"))
    H = T + V
    bs = _Blocks(be, S, 1, torch.device("cpu"))
    Hb = bs.blocks_of(H)
    plan = bs.prepare(Hb, 0)
    sols = batched_eigh(plan["mats"])
    Pa, _, _ = bs.finish(plan, sols, 2, 1.0, H, frozen=[1, 1, 0])
    Pb, _, _ = bs.finish(plan, sols, 2, 1.0, H, frozen=[0, 1, 1])
    Ph, _, nk = bs.finish(plan, sols, 2, 1.0, H, frozen=[0.5, 1, 0.5])
    assert nk == [1, 1, 1]
    assert float(torch.max(torch.abs(Ph - 0.5 * (Pa + Pb)))) < 1e-13
    assert abs(float(np.sum(Ph.numpy() * S.numpy())) - 2.0) < 1e-10
