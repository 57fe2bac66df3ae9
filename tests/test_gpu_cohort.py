"""Vectorised cohort SCF (helfem_amd/cohort.py): many atoms iterated together
with batched Fock builds, stacked block algebra and batched eigensolves.

Each system's energy must be the same in the cohort and alone (1e-10 Eh),
match the per-system SCF driver (1e-9 Eh; the two take the orthogonaliser
over whole blocks or over shell groups, which differ by rounding) with the
same fixed ground-state occupations, and the reference's recorded values
(tests/refs/ci.json, tests/README.md:80); the mixed-precision cohort (FP32
exchange until the orbital gradient is below 1e-3) must reach the FP64
energies to 1e-9 Eh.  Aufbau occupations (no fixed pattern) are covered by
the cohort-vs-alone comparison of a second set."""
import json
from pathlib import Path

import pytest

import helfem_amd as H
from helfem_amd import batch
from helfem_amd.cohort import Cohort

pytestmark = pytest.mark.gpu
REF = json.loads((Path(__file__).parent / "golden" / "reference_energies.json").read_text())

# (Z, lmax, nelem, nnodes): closed and open shells, spherical and not
ATOMS = [(2, 0, 4, 15), (10, 1, 5, 15), (1, 0, 5, 15), (7, 1, 4, 10), (4, 1, 4, 15), (8, 2, 3, 10),
         (26, 2, 4, 10), (22, 3, 4, 10)]


@pytest.fixture(scope="module")
def systems(gpu_ctx):
    out = []
    for Z, lmax, nel, nn in ATOMS:
        b = H.AtomicBasis(Z, lmax, lmax, nel, nnodes=nn, ctx=gpu_ctx)
        b.compute_tei()
        out.append(batch.atom_system(b, Z, info=dict(Z=Z, lmax=lmax, nelem=nel, nnodes=nn)))
    return out


def _occ(systems):
    return [s.occ for s in systems]


@pytest.fixture(scope="module")
def together(systems, gpu_ctx):
    return Cohort(systems, gpu_ctx, conv=1e-10, occupations=_occ(systems)).solve()


def test_cohort_equals_alone(systems, together, gpu_ctx):
    print("[gpu] cohort SCF: together vs alone", flush=True)
    for s, r in zip(systems, together):
        alone = Cohort([s], gpu_ctx, conv=1e-10, occupations=[s.occ]).solve()[0]
        assert r["converged"] and alone["converged"], s.info
        assert abs(r["E"] - alone["E"]) < 1e-10, (s.info, r["E"], alone["E"])


def test_cohort_aufbau_equals_alone(systems, gpu_ctx):
    """Without fixed occupations (aufbau each iteration) as well."""
    print("[gpu] cohort SCF with aufbau occupations: together vs alone", flush=True)
    sub = systems[:6]
    res = Cohort(sub, gpu_ctx, conv=1e-10).solve()
    for s, r in zip(sub, res):
        alone = Cohort([s], gpu_ctx, conv=1e-10).solve()[0]
        assert abs(r["E"] - alone["E"]) < 1e-10, (s.info, r["E"], alone["E"])


def test_cohort_reference_energies(systems, together):
    print("[gpu] cohort SCF: reference energies", flush=True)
    E = {s.info["Z"]: r["E"] for s, r in zip(systems, together)}
    assert abs(E[10] + 128.5470981094) < 1e-9, E[10]
    assert abs(E[1] - REF["atomic-H-hf-u"]["total"]) < 1e-9, E[1]


def test_cohort_matches_per_system_driver(systems, together):
    print("[gpu] cohort SCF vs the per-system driver", flush=True)
    seq, _ = batch.solve_sequential(systems, conv=1e-10)
    for s, r, ref in zip(systems, together, seq):
        assert abs(r["E"] - ref.E) < 1e-9, (s.info, r["E"], ref.E)


def test_mixed_precision_cohort(systems, together, gpu_ctx):
    print("[gpu] mixed-precision cohort SCF", flush=True)
    mixed = Cohort(systems, gpu_ctx, precision="mixed", conv=1e-10, occupations=_occ(systems)).solve()
    for s, a, b in zip(systems, mixed, together):
        assert abs(a["E"] - b["E"]) < 1e-9, (s.info, a["E"], b["E"])
    assert any(r["iterations_fp32"] > 0 for r in mixed)


def test_memory_bounded_cohorts_equal_one_cohort(systems, together, gpu_ctx):
    """solve_cohorts under a budget that splits the systems into several
    cohorts gives each system the energy of the single cohort."""
    from helfem_amd.cohort import cohort_bytes, solve_cohorts
    print("[gpu] memory-bounded cohorts vs one cohort", flush=True)
    budget = 2 * max(cohort_bytes(s) for s in systems)
    res, _ = solve_cohorts(systems, gpu_ctx, budget_bytes=budget, conv=1e-10, occupations=_occ(systems))
    assert len({r["cohort"] for r in res}) > 1
    for s, a, b in zip(systems, res, together):
        assert a["converged"], s.info
        assert abs(a["E"] - b["E"]) < 1e-10, (s.info, a["E"], b["E"])
