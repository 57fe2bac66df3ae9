"""Ground-state aufbau occupations of neutral atoms, per symmetry block.

Subshells fill in Madelung (n + l, n) order; an open subshell holds its
electrons by Hund's rule (alpha first, up to 2l + 1), each spin occupying
m = l, l - 1, ... in turn.  The counts of occupied orbitals per spin and
symmetry block -- per (l, m) shell (the reference's symmetry=2), per m
(symmetry=1) or per (m, parity of l) -- are what the reference's --readocc
feeds its optimiser as fixed per-block occupations
(src/atomic/main.cpp:526-583, four columns nocca, noccb, l, m under
symmetry=2): the SCF then converges to the state of that configuration
instead of whichever aufbau pattern its path happens to reach, which for
open-shell atoms depends on rounding (and so on the precision of the early
iterations).  Only per-(l, m) blocks pin the configuration itself: a larger
block fills its lowest levels, which may trade an s for a d orbital.
"""
from __future__ import annotations

import numpy as np

MADELUNG = sorted(((n, l) for n in range(1, 8) for l in range(0, n)), key=lambda s: (s[0] + s[1], s[0]))
# the same order continued to high n, for fillings restricted to low l (an
# s-only basis holds Rn's 86 electrons in 43 s orbitals)
MADELUNG_LONG = sorted(((n, l) for n in range(1, 61) for l in range(0, n)), key=lambda s: (s[0] + s[1], s[0]))


def configuration(Z, lmax=None):
    """[(n, l, alpha m list, beta m list)] of the Madelung/Hund ground state;
    with lmax, of the Madelung/Hund filling of the subshells with l <= lmax
    alone (the configuration aufbau reaches in a basis without the higher l)."""
    left, out = Z, []
    for n, l in (MADELUNG if lmax is None else MADELUNG_LONG):
        if left <= 0:
            break
        if lmax is not None and l > lmax:
            continue
        cap = 2 * (2 * l + 1)
        k = min(left, cap)
        na = min(k, 2 * l + 1)
        nb = k - na
        ms = list(range(l, -l - 1, -1))
        out.append((n, l, ms[:na], ms[:nb]))
        left -= k
    return out


def spins(Z, lmax=None):
    """(nela, nelb) of the configuration."""
    cfg = configuration(Z, lmax)
    return sum(len(a) for _, _, a, _ in cfg), sum(len(b) for _, _, _, b in cfg)


def block_counts(Z, lval, mval, symm, lmax=None):
    """Occupied orbitals per spin and symmetry block of an atomic basis with
    shells (lval, mval), blocks ordered as AtomicBasis.get_sym_idx(symm):
    symm 2 one per shell (shell order), 1 one per m (ascending), 3 one per m
    and parity of l (even before odd).  Returns (alpha counts, beta counts),
    or None when a subshell of the configuration has no (l, m) shell in the
    basis (lmax or mmax too small).  lmax restricts the configuration to the
    subshells with l <= lmax (configuration(Z, lmax))."""
    lval, mval = np.asarray(lval).tolist(), np.asarray(mval).tolist()
    have = set(zip(lval, mval))
    if symm == 2:
        keys = list(zip(lval, mval))
        key_of = lambda l, m: (l, m)  # noqa: E731
    elif symm == 1:
        keys = sorted(set(mval))
        key_of = lambda l, m: m  # noqa: E731
    elif symm == 3:
        keys = [(m, p) for m in sorted(set(mval)) for p in (0, 1)
                if any(mm == m and ll % 2 == p for ll, mm in have)]
        key_of = lambda l, m: (m, l % 2)  # noqa: E731
    else:
        raise ValueError("symm must be 1, 2 or 3")
    index = {k: i for i, k in enumerate(keys)}
    ca, cb = [0] * len(keys), [0] * len(keys)
    for _, l, ma, mb in configuration(Z, lmax):
        for counts, ms in ((ca, ma), (cb, mb)):
            for m in ms:
                if (l, m) not in have:
                    return None
                counts[index[key_of(l, m)]] += 1
    return ca, cb
