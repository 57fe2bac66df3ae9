"""Oracle Gaunt coefficients against the reference's golden values.

tests/golden/gaunt_reference.json holds the 550 (call, value) pairs hard-coded
in HelFEM's src/general/gaunt_test.cpp (extracted by make_gaunt_fixture.py).
The reference test tolerance is DBL_EPSILON*(1+|ref|) (gaunt_test.cpp:11).
"""
import json
import math
from pathlib import Path

import oracle as O

GOLD = json.loads((Path(__file__).parent / "golden" / "gaunt_reference.json").read_text())["entries"]


def test_gaunt_golden_values():
    assert len(GOLD) == 550
    eps = 2.220446049250313e-16
    bad = []
    for e in GOLD:
        f = O.gaunt if e["kind"] == "coeff" else O.mod_gaunt
        v = f(*e["args"])
        if abs(v - e["value"]) >= eps * (1 + abs(e["value"])):
            bad.append((e, v))
    assert not bad, bad[:5]


def test_gaunt_closed_forms():
    # int Y00* Y00 Y00 = 1/sqrt(4 pi)   (atomic/precision.cpp:291-306)
    assert abs(O.gaunt(0, 0, 0, 0, 0, 0) - 1 / math.sqrt(4 * math.pi)) < 1e-17
    # int Y_10* Y_00 Y_10 = 1/sqrt(4 pi)
    assert abs(O.gaunt(1, 0, 0, 0, 1, 0) - 1 / math.sqrt(4 * math.pi)) < 1e-17
    # int Y_00* Y_1m Y_1,-m = (-1)^m / sqrt(4 pi)  (Y_1^-m = (-1)^m Y_1^m*)
    for m in (-1, 0, 1):
        assert abs(O.gaunt(0, 0, 1, m, 1, -m) - (-1) ** m / math.sqrt(4 * math.pi)) < 1e-17
    # selection rules
    assert O.gaunt(1, 0, 1, 0, 1, 0) == 0.0          # parity
    assert O.gaunt(4, 0, 1, 0, 1, 0) == 0.0          # triangle
    assert O.gaunt(2, 1, 1, 0, 1, 0) == 0.0          # m-sum


def test_gaunt_sympy_exact():
    sympy = __import__("pytest").importorskip("sympy")
    from sympy.physics.wigner import gaunt as sgaunt
    for (L, M, l, m, lp) in [(2, 1, 3, -1, 3), (4, -2, 2, 1, 4), (6, 3, 5, 2, 7), (12, 5, 6, 3, 6), (20, 4, 11, -3, 13)]:
        mp = M - m
        # int Y_L^M* Y_l^m Y_lp^mp = (-1)^M gaunt(L,-M,l,m,lp,mp)
        ref = float((-1) ** M * sgaunt(L, l, lp, -M, m, mp))
        assert abs(O.gaunt(L, M, l, m, lp, mp) - ref) <= 2e-16 * max(1.0, abs(ref)), (L, M, l, m, lp)
