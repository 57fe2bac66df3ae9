"""Ground-state configurations used as fixed per-block occupations
(helfem_amd/occupations.py; the reference's --readocc input,
src/atomic/main.cpp:526-583).  Checked against known ground-state
configurations and their spin multiplicities, and for the restricted
filling of bases that lack some occupied l."""
import numpy as np
import pytest

from helfem_amd import occupations as oc
from helfem_amd.atomic import angular_basis

# (Z, unpaired electrons of the Madelung/Hund configuration; Gd: 4f8 6s2 there, 6 unpaired)
UNPAIRED = [(1, 1), (2, 0), (5, 1), (6, 2), (7, 3), (8, 2), (10, 0), (21, 1), (25, 5), (26, 4), (30, 0),
            (36, 0), (54, 0), (57, 1), (64, 6), (86, 0)]


@pytest.mark.parametrize("Z,unpaired", UNPAIRED)
def test_spins_of_known_ground_states(Z, unpaired):
    na, nb = oc.spins(Z)
    assert na + nb == Z and na - nb == unpaired


@pytest.mark.parametrize("Z", range(1, 87))
@pytest.mark.parametrize("lmax", [0, 1, 2, 3])
def test_restricted_configuration(Z, lmax):
    """Over the subshells with l <= lmax: every electron placed, no subshell
    above its capacity, Madelung order (a subshell holds electrons only when
    every earlier allowed one is full), Hund's rule in the open one; the
    per-(l, m) counts of a basis with all m of those l add up to the spins."""
    cfg = oc.configuration(Z, lmax)
    assert sum(len(a) + len(b) for _, _, a, b in cfg) == Z
    assert all(l <= lmax for _, l, _, _ in cfg)
    for q, (n, l, a, b) in enumerate(cfg):
        cap = 2 * l + 1
        assert len(a) <= cap and len(b) <= cap and len(b) <= len(a)
        if q + 1 < len(cfg):
            assert len(a) == cap and len(b) == cap  # only the last subshell is open
    lval, mval = angular_basis(lmax, lmax)
    counts = oc.block_counts(Z, lval, mval, 2, lmax=lmax)
    na, nb = oc.spins(Z, lmax)
    assert counts is not None and sum(counts[0]) == na and sum(counts[1]) == nb
    # the unrestricted configuration fits exactly when no occupied l exceeds lmax
    fits = max(l for _, l, _, _ in oc.configuration(Z)) <= lmax
    assert (oc.block_counts(Z, lval, mval, 2) is not None) == fits
    if fits:
        assert oc.block_counts(Z, lval, mval, 2) == counts
