"""Pin the oracle's B3LYP pieces and the hybrid SCF: Becke 88 exchange
(libxc gga_x_b88, id 106), Lee-Yang-Parr correlation (gga_c_lyp, 131), VWN's
RPA fit (lda_c_vwn_rpa, 8) and their B3LYP mix (hyb_gga_xc_b3lyp, 402, with
0.20 exact exchange; src/atomic/main.cpp:237-243, 424-437).

libxc is absent, so each piece is pinned by what its construction fixes:
B88 by the spin-scaling identity of exchange and its gradient coefficient
beta at small reduced gradients (Becke, PRA 38, 3098 (1988)); LYP by the
original closed-shell formula at zero gradient (Lee, Yang, Parr, PRB 37, 785
(1988)) and by vanishing for a fully polarized density; the RPA fit by the
exact high-density RPA limit A ln rs, A = (1 - ln 2)/pi^2 unpolarized and
half that fully polarized; the B3LYP mix and its exact-exchange fraction by
the reference's He energies, restricted and unrestricted
(tests/refs/ci.json atomic-He-hybrid-r/-u, -2.9152186629).

The range-separated hybrids CAM-B3LYP (433, erfc, omega 0.33) and
CAMY-B3LYP (470, Yukawa, omega 0.34): the short-range attenuation of the
uniform-gas exchange against direct quadrature of its defining integral,
(8/3) int_0^1 w(x) (1 - 3x/2 + x^3/2) dx, over densities that take both the
closed forms and the large-a series; the whole functional with 0.65 exact
exchange and -0.46 of its short-range part by the reference's H energies
(atomic-H-rsh-yukawa-u -0.4980726135 to 1e-9 Eh; atomic-H-rsh-erfc-u
-0.4990948394 agrees to 9.2e-9 Eh and is bounded at 2e-8: the Yukawa twin,
which shares every other term, agrees to 2e-10, which places the residual
in the erfc short-range exchange, whose multipole kernel the reference
evaluates by a series it describes as noisy for small arguments
(erfc_expn.cpp:171-235) where the oracle integrates it to 1e-13).  The fitted
b, c, x0 of the ferromagnetic RPA curve enter no pinned quantity above
(He is unpolarized): parity unpinned for them beyond the high-density limit.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle as O

REF = json.loads((Path(__file__).parent / "golden" / "reference_energies.json").read_text())
CX = 0.75 * (6 / np.pi) ** (1 / 3)  # spin-scaled Slater constant


def _f(rho, sigma, xf, cf):
    """energy per volume"""
    r = np.asarray(rho)
    tot = r if r.ndim == 1 else r[0] + r[1]
    return tot * O.xc_eval(rho, sigma, xf, cf)[0]


def test_b88_spin_scaling():
    """E_x[rho_a, rho_b] = (E_x[2 rho_a] + E_x[2 rho_b]) / 2 for exchange"""
    rng = np.random.default_rng(3)
    ra = np.geomspace(1e-3, 30, 12)
    rb = ra * rng.uniform(0.05, 0.9, ra.size)
    saa = rng.uniform(0.01, 2.0, ra.size) * ra ** (8 / 3)
    sbb = rng.uniform(0.01, 2.0, ra.size) * rb ** (8 / 3)
    pol = _f(np.stack([ra, rb]), np.stack([saa, 0.3 * np.sqrt(saa * sbb), sbb]), 106, 0)
    unp = 0.5 * (_f(2 * ra, 4 * saa, 106, 0) + _f(2 * rb, 4 * sbb, 106, 0))
    assert np.max(np.abs(pol / unp - 1)) < 1e-14


def test_b88_small_gradient_coefficient():
    """e_x = -C_x rho_s^(4/3) (1 + (beta / C_x) x^2 + O(x^3)), x = |grad rho_s| / rho_s^(4/3), beta 0.0042"""
    rho = np.array([0.05, 1.0, 20.0])
    for x in (1e-2, 3e-3):
        rs = rho / 2  # each spin
        sigma = 4 * (x * rs ** (4 / 3)) ** 2  # total sigma = 4 sigma_ss
        ratio = _f(rho, sigma, 106, 0) / _f(rho, 0 * sigma, 106, 0)
        # the next term is relatively -6 beta x^2 (2.5e-6 at x = 1e-2)
        assert np.max(np.abs((ratio - 1) / x ** 2 / (0.0042 / CX) - 1)) < 1e-5
    # zero gradient: Slater exchange
    assert np.max(np.abs(_f(rho, 0 * rho, 106, 0) / _f(rho, 0 * rho, 1, 0) - 1)) < 1e-15


def test_lyp_zero_gradient_is_the_closed_shell_formula():
    """at sigma = 0 the Miehlich form reduces to LYP's closed-shell
    -a rho (1 + b C_F exp(-c rho^-1/3)) / (1 + d rho^-1/3)"""
    a, b, c, d = 0.04918, 0.132, 0.2533, 0.349
    cf = 0.3 * (3 * np.pi ** 2) ** (2 / 3)
    rho = np.geomspace(1e-5, 1e3, 17)
    rm = rho ** (-1 / 3)
    ref = -a * rho * (1 + b * cf * np.exp(-c * rm)) / (1 + d * rm)
    assert np.max(np.abs(_f(rho, 0 * rho, 0, 131) / ref - 1)) < 1e-13


def test_lyp_vanishes_for_one_spin():
    rng = np.random.default_rng(5)
    ra = np.geomspace(1e-3, 30, 9)
    saa = rng.uniform(0.01, 3.0, ra.size) * ra ** (8 / 3)
    z = np.zeros_like(ra)
    e, vr, vs = O.xc_eval(np.stack([ra, z]), np.stack([saa, z, z]), 0, 131)
    assert np.max(np.abs(e)) < 1e-14 * np.max(np.abs(O.xc_eval(ra, saa, 0, 131)[0]))


def test_vwn_rpa_high_density_limit():
    """eps_c -> A ln rs + const as rs -> 0: A = (1 - ln 2)/pi^2 unpolarized, A/2 fully polarized"""
    A = (1 - np.log(2)) / np.pi ** 2
    rs = np.array([1e-10, 1e-12])
    rho = 3 / (4 * np.pi * rs ** 3)
    e0 = O.xc_eval(rho, 0 * rho, 0, 8)[0]
    assert abs((e0[0] - e0[1]) / np.log(rs[0] / rs[1]) - A) < 1e-6 * A
    z = np.zeros_like(rho)
    e1 = O.xc_eval(np.stack([rho, z]), np.stack([z, z, z]), 0, 8)[0]
    assert abs((e1[0] - e1[1]) / np.log(rs[0] / rs[1]) - A / 2) < 1e-6 * A


def test_exact_exchange_fractions():
    assert O.exact_exchange(402) == (0.20, 0.0, 0.0, None)
    assert O.exact_exchange(406) == (0.25, 0.0, 0.0, None)
    assert O.exact_exchange(101) == (0.0, 0.0, 0.0, None)
    assert O.exact_exchange(433) == (0.65, -0.46, 0.33, "erfc")
    assert O.exact_exchange(470) == (0.65, -0.46, 0.34, "yukawa")


@pytest.mark.parametrize("xf,omega,kernel", [(433, 0.33, "erfc"), (470, 0.34, "yukawa")])
def test_short_range_attenuation_against_quadrature(xf, omega, kernel):
    """at zero gradient the short-range B88 is Slater exchange times the
    attenuation F(a), a = omega / (2 k_F) per spin channel"""
    from scipy.integrate import quad
    rho = np.geomspace(1e-9, 1e3, 25)  # a from ~50 down to ~0.004
    z = np.zeros_like(rho)
    sr = (_f(rho, z, xf, 0) - 0.35 * _f(rho, z, 1, 0) - 0.19 * _f(rho, z, 0, 7) - 0.81 * _f(rho, z, 0, 131)) / 0.46
    kf = (6 * np.pi ** 2 * rho / 2) ** (1 / 3)
    scale = np.abs(_f(rho, z, 1, 0)) / 0.46  # the subtraction's own magnitude
    for r, k, e, sc in zip(rho, kf, sr, scale):
        a = omega / (2 * k)
        if kernel == "erfc":
            w = lambda x: -np.expm1(-x * x / (4 * a * a))
        else:
            w = lambda x: x * x / (x * x + a * a)
        F = 8 / 3 * quad(lambda x: w(x) * (1 - 1.5 * x + 0.5 * x ** 3), 0, 1, epsabs=0, epsrel=1e-13, limit=200)[0]
        ref = -2 * CX * (r / 2) ** (4 / 3) * F
        assert abs(e - ref) < 1e-12 * abs(ref) + 1e-14 * sc, (r, a, e, ref)


@pytest.mark.parametrize("xf,name,tol", [(470, "atomic-H-rsh-yukawa-u", 1e-9), (433, "atomic-H-rsh-erfc-u", 2e-8)])
def test_h_range_separated_hybrid_energies(xf, name, tol):
    b = O.AtomicBasis(1, 0, 0, 5)
    b.compute_tei()
    kf, ks, om, ker = O.exact_exchange(xf)
    (b.compute_erfc if ker == "erfc" else b.compute_yukawa)(om)
    u = O.scf_uhf(b, 1, 0, xc_pol=lambda Pa, Pb: b.xc_pol(Pa, Pb, x_func=xf, c_func=0)[:3], kfrac=kf, kshort=ks)
    assert abs(u["E"] - REF[name]["total"]) < tol, (u["E"], REF[name]["total"])


def test_he_b3lyp_energies():
    b = O.AtomicBasis(2, 0, 0, 5)
    b.compute_tei()
    kf = O.exact_exchange(402)[0]
    r = O.scf_rhf(b, 2, xc=lambda P: b.xc(P, x_func=402, c_func=0)[:2], kfrac=kf)
    assert abs(r["E"] - REF["atomic-He-hybrid-r"]["total"]) < 1e-9, r["E"]
    u = O.scf_uhf(b, 1, 1, xc_pol=lambda Pa, Pb: b.xc_pol(Pa, Pb, x_func=402, c_func=0)[:3], kfrac=kf)
    assert abs(u["E"] - REF["atomic-He-hybrid-u"]["total"]) < 1e-9, u["E"]
