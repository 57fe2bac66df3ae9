"""Host-side pieces of the vectorised cohort that need no GPU: the aufbau
over block levels must pick the same orbitals as the per-system driver's
(helfem_amd/scf.py _Blocks.finish): the nocc lowest, chains of levels closer
than 1e-9 (relative) filled in block order, frozen occupations per block."""
import numpy as np
import pytest

from helfem_amd.cohort import DEGEN, aufbau_levels


def reference_aufbau(lev, nocc):
    """The per-system driver's rule, written out level by level."""
    lv = sorted((float(e), k, j) for k in range(lev.shape[0]) for j, e in enumerate(lev[k]) if np.isfinite(e))
    order, start = [], 0
    for q in range(1, len(lv) + 1):
        if q == len(lv) or lv[q][0] - lv[q - 1][0] > DEGEN * max(1.0, abs(lv[q][0])):
            order += sorted(lv[start:q], key=lambda x: (x[1], x[2]))
            start = q
    counts = np.zeros(lev.shape[0], dtype=np.int64)
    for _, k, _j in order[:nocc]:
        counts[k] += 1
    return counts, (order[nocc - 1][0] if nocc else -np.inf)


@pytest.mark.parametrize("seed", range(40))
def test_aufbau_matches_per_system_rule(seed):
    rng = np.random.default_rng(seed)
    nb, km = int(rng.integers(1, 9)), int(rng.integers(1, 12))
    base = np.sort(rng.uniform(-50, 5, size=(nb, km)), axis=1)
    # plant degenerate levels across blocks (exact and within the tolerance)
    for _ in range(int(rng.integers(0, 6))):
        a, b = rng.integers(0, nb, size=2)
        j = int(rng.integers(0, km))
        base[b, j] = base[a, j] * (1.0 + float(rng.choice([0.0, 3e-10, -4e-10])))
    base = np.sort(base, axis=1)
    # some blocks hold fewer levels
    for k in range(nb):
        if rng.random() < 0.3:
            base[k, int(rng.integers(1, km + 1)):] = np.inf
    nlev = int(np.isfinite(base).sum())
    nocc = int(rng.integers(0, nlev + 1))
    got, homo = aufbau_levels(base, nocc)
    ref, rhomo = reference_aufbau(base, nocc)
    assert np.array_equal(got, ref)
    assert homo == rhomo


def test_frozen_occupations_take_each_block_lowest():
    lev = np.array([[-3.0, -1.0, np.inf], [-2.0, 0.5, 1.0]])
    got, homo = aufbau_levels(lev, 3, frozen=[1, 3])
    assert got.tolist() == [1, 3] and homo == np.inf
    got, _ = aufbau_levels(lev, 3, frozen=[5, 0])
    assert got.tolist() == [2, 0]


def test_too_few_levels_is_an_error():
    with pytest.raises(RuntimeError):
        aufbau_levels(np.array([[-1.0, np.inf]]), 2)


class _FakeBasis:
    def __init__(self, nbf, blocks):
        self.Nbf = nbf
        self._blocks = blocks

    def get_sym_idx(self, symm):
        out, o = [], 0
        for b in self._blocks:
            out.append(np.arange(o, o + b))
            o += b
        return out


class _FakeSystem:
    def __init__(self, nbf, blocks, nela, nelb):
        self.basis, self.nela, self.nelb = _FakeBasis(nbf, blocks), nela, nelb


def test_cohort_groups_respect_budget():
    """Memory-bounded cohorts: consecutive runs under the budget, every
    system once, in order; an oversized system alone."""
    from helfem_amd.cohort import cohort_bytes, cohort_groups
    rng = np.random.default_rng(7)
    systems = []
    for _ in range(50):
        blocks = [int(x) for x in rng.integers(5, 60, size=int(rng.integers(1, 6)))]
        na = int(rng.integers(1, 10))
        systems.append(_FakeSystem(sum(blocks), blocks, na, na - int(rng.integers(0, 2))))
    sizes = [cohort_bytes(s) for s in systems]
    # restricted systems hold one channel, open-shell ones two
    r = _FakeSystem(100, [100], 3, 3)
    u = _FakeSystem(100, [100], 3, 2)
    assert cohort_bytes(u) - cohort_bytes(r) == 8 * 100 * 100 * (2 + 30)
    budget = 3 * max(sizes)
    groups = cohort_groups(systems, budget)
    assert [i for g in groups for i in g] == list(range(len(systems)))
    for g in groups:
        assert sum(sizes[i] for i in g) <= budget
    for a, b in zip(groups, groups[1:]):
        assert sum(sizes[i] for i in a) + sizes[b[0]] > budget  # greedy: the next one did not fit
    assert cohort_groups(systems, 1) == [[i] for i in range(len(systems))]
    assert cohort_groups(systems, 1 << 62) == [list(range(len(systems)))]
