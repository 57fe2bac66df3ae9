"""Pin the oracle's GGA exchange-correlation path: PBE exchange (libxc
gga_x_pbe, id 101) and correlation (gga_c_pbe, id 130) on the atomic DFT
grid (src/atomic/dftgrid.cpp:51-128, 304-359, 470-576;
src/general/dftgrid_common.h:188-216).

libxc is absent, so the functionals are pinned by what their construction
fixes (Perdew, Burke, Ernzerhof, PRL 77, 3865 (1996)): the uniform-gas limit
(PBE exchange = Slater, PBE correlation = PW92 at zero gradient), the
gradient-expansion coefficients mu and beta at small gradients, the
Lieb-Oxford saturation 1 + kappa of the exchange enhancement and the
vanishing correlation of rapidly varying densities; the potentials by finite
differences of the energy; the grid's gradient terms by the exchange-
correlation matrix being the derivative of Exc with respect to P; and the
whole path by the reference's PBE energies of He, restricted and
unrestricted (tests/refs/ci.json atomic-He-gga-r/-u, -2.8929348668).
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle as O

REF = json.loads((Path(__file__).parent / "golden" / "reference_energies.json").read_text())
KAPPA, MU, BETA = 0.804, 0.06672455060314922 * np.pi ** 2 / 3, 0.06672455060314922


def test_zero_gradient_is_the_uniform_gas():
    rho = np.geomspace(1e-6, 1e3, 9)
    z = np.zeros_like(rho)
    ex, vx, _ = O.xc_eval(rho, z, 101, 0)
    e0, v0 = O.lda(rho, 1, 0)
    assert np.max(np.abs(ex / e0 - 1)) < 1e-14 and np.max(np.abs(vx / v0 - 1)) < 1e-14
    ec, vc, _ = O.xc_eval(rho, z, 0, 130)
    ep, vp, _ = O.xc_eval(rho, z, 0, 13)
    assert np.max(np.abs(ec / ep - 1)) < 1e-13 and np.max(np.abs(vc / vp - 1)) < 1e-13
    # polarized, zeta = 1/3 and fully polarized
    for f in (2 / 3, 1.0):
        r2 = np.stack([f * rho, (1 - f) * rho])
        s3 = np.zeros((3, rho.size))
        ec, _, _ = O.xc_eval(r2, s3, 0, 130)
        ep, _, _ = O.xc_eval(r2, s3, 0, 13)
        assert np.max(np.abs(ec / ep - 1)) < 1e-13, f


def test_exchange_enhancement_limits():
    rho = np.full(3, 0.3)
    e0, _ = O.lda(rho, 1, 0)
    # small s: F = 1 + mu s^2 + O(s^4)
    s = np.array([1e-3, 2e-3, 4e-3])
    sigma = (s * 2 * (3 * np.pi ** 2) ** (1 / 3) * rho ** (4 / 3)) ** 2
    ex, _, _ = O.xc_eval(rho, sigma, 101, 0)
    assert np.allclose((ex / e0 - 1) / s ** 2, MU, rtol=1e-4)
    # large s: F -> 1 + kappa (Lieb-Oxford)
    ex, _, _ = O.xc_eval(rho, np.array([1e6, 1e8, 1e10]), 101, 0)
    F = ex / e0
    assert np.all(np.diff(F) > 0) and abs(F[-1] - (1 + KAPPA)) < 1e-6


def test_correlation_gradient_limits():
    rho = np.array([1e-3, 0.05, 2.0])
    ep, _, _ = O.xc_eval(rho, np.zeros(3), 0, 13)
    # small t: H = beta t^2 + O(t^4) (phi = 1); t^2 = sigma / (4 ks^2 rho^2), ks^2 = 4 kF / pi
    ks2 = 4 * (3 * np.pi ** 2 * rho) ** (1 / 3) / np.pi
    t2 = 1e-6
    ec, _, _ = O.xc_eval(rho, t2 * 4 * ks2 * rho ** 2, 0, 130)
    assert np.allclose((ec - ep) / t2, BETA, rtol=1e-4)
    # t -> infinity: the correlation energy vanishes
    ec, _, _ = O.xc_eval(rho, 1e12 * 4 * ks2 * rho ** 2, 0, 130)
    assert np.max(np.abs(ec / ep)) < 1e-5


def test_pw92_near_vwn5():
    """PW92 and VWN5 fit the same Ceperley-Alder energies: within 1 mEh per
    electron over rs 0.5..20, unpolarized and fully polarized; the modified
    parameters differ from the printed ones by less than 1e-6 Eh."""
    rs = np.geomspace(0.5, 20, 12)
    rho = 3 / (4 * np.pi * rs ** 3)
    e7, _, _ = O.xc_eval(rho, np.zeros_like(rho), 0, 7)
    e12, _, _ = O.xc_eval(rho, np.zeros_like(rho), 0, 12)
    e13, _, _ = O.xc_eval(rho, np.zeros_like(rho), 0, 13)
    assert np.max(np.abs(e12 - e7)) < 1e-3 and np.max(np.abs(e13 - e12)) < 1e-6
    r2, s3 = np.stack([rho, np.zeros_like(rho)]), np.zeros((3, rho.size))
    f7, _, _ = O.xc_eval(r2, s3, 0, 7)
    f13, _, _ = O.xc_eval(r2, s3, 0, 13)
    assert np.max(np.abs(f13 - f7)) < 1e-3


def _fd(f, x, h):
    """fourth-order central difference of f at x (step h, elementwise)"""
    return (8 * (f(x + h) - f(x - h)) - (f(x + 2 * h) - f(x - 2 * h))) / (12 * h)


@pytest.mark.parametrize("xf,cf", [(101, 0), (0, 130), (1, 0), (0, 7), (0, 12), (0, 13), (106, 0), (0, 131),
                                   (0, 8), (402, 0), (433, 0), (470, 0)])
def test_potentials_are_energy_derivatives(xf, cf):
    """vrho and vsigma against fourth-order finite differences of rho*exc, one
    functional at a time (PBE exchange and correlation cancel in vsigma to
    three digits, which a difference quotient of their sum cannot resolve)."""
    rng = np.random.default_rng(11)
    rho = np.geomspace(1e-4, 50, 15)
    sigma = rng.uniform(0.01, 3.0, rho.size) * rho ** (8 / 3)
    e, vr, vs = O.xc_eval(rho, sigma, xf, cf)
    fr = _fd(lambda x: x * O.xc_eval(x, sigma, xf, cf)[0], rho, 1e-3 * rho)
    assert np.max(np.abs(fr - vr) / np.abs(vr)) < 1e-8
    if xc_gga := xf in (101, 106, 402, 406, 433, 470) or cf in (130, 131):
        fs = _fd(lambda x: rho * O.xc_eval(rho, x, xf, cf)[0], sigma, 1e-3 * sigma)
        assert np.max(np.abs(fs - vs) / np.abs(vs)) < 1e-8
    else:
        assert np.all(vs == 0.0)
    # polarized: each of the five inputs
    r2 = np.stack([0.7 * rho, 0.3 * rho])
    s3 = np.stack([0.49 * sigma, 0.15 * sigma, 0.09 * sigma])
    e, vr, vs = O.xc_eval(r2, s3, xf, cf)
    for k in range(5 if xc_gga else 2):
        def f(x, k=k):
            a, b = r2.copy(), s3.copy()
            (a if k < 2 else b)[k if k < 2 else k - 2] = x
            return (a[0] + a[1]) * O.xc_eval(a, b, xf, cf)[0]
        x = r2[k] if k < 2 else s3[k - 2]
        v = vr[k] if k < 2 else vs[k - 2]
        fd = _fd(f, x, 1e-3 * x)
        assert np.max(np.abs(fd - v) / (np.abs(v) + 1e-300)) < 1e-8, k


@pytest.fixture(scope="module")
def ne_basis():
    b = O.AtomicBasis(10, 2, 1, 3, nnodes=8)
    b.compute_tei()
    return b


def _density(b, seed, k=6):
    rng = np.random.default_rng(seed)
    C = 0.3 * rng.standard_normal((b.Nbf, k))
    D = rng.standard_normal((b.Nbf, b.Nbf))
    return C @ C.T, D + D.T


def test_gga_matrix_is_energy_derivative(ne_basis):
    """<H, D> = dExc(P + h D)/dh: the gradient terms (d/dr, d/dtheta, d/dphi
    of the basis, the 1/r and 1/(r sin theta) scales and the X + X' form) carry
    their factors correctly on a non-spherical density (lmax 2, mmax 1)."""
    b = ne_basis
    P, D = _density(b, 5)
    H, E, _ = b.xc(P, x_func=101, c_func=130)
    h = 1e-5
    fd = (b.xc(P + h * D, x_func=101, c_func=130)[1] - b.xc(P - h * D, x_func=101, c_func=130)[1]) / (2 * h)
    assert abs(fd / np.sum(H * D) - 1) < 1e-8
    Pa, Pb = 0.6 * P, 0.4 * P
    Ha, Hb, _, _ = b.xc_pol(Pa, Pb, x_func=101, c_func=130)
    for which, Hs in ((0, Ha), (1, Hb)):
        e = []
        for sgn in (1, -1):
            A, B = (Pa + sgn * h * D, Pb) if which == 0 else (Pa, Pb + sgn * h * D)
            e.append(b.xc_pol(A, B, x_func=101, c_func=130)[2])
        assert abs((e[0] - e[1]) / (2 * h) / np.sum(Hs * D) - 1) < 1e-8, which


def test_polarized_equal_spins_is_restricted(ne_basis):
    b = ne_basis
    P, _ = _density(b, 6)
    H, E, N = b.xc(P, x_func=101, c_func=130)
    Ha, Hb, E2, N2 = b.xc_pol(0.5 * P, 0.5 * P, x_func=101, c_func=130)
    s = np.max(np.abs(H))
    assert abs(E2 - E) < 1e-13 * abs(E) and np.max(np.abs(Ha - H)) < 1e-13 * s and np.max(np.abs(Hb - H)) < 1e-13 * s


def test_he_pbe_energies():
    b = O.AtomicBasis(2, 0, 0, 5)
    b.compute_tei()
    r = O.scf_rhf(b, 2, xc=lambda P: b.xc(P, x_func=101, c_func=130)[:2])
    assert abs(r["E"] - REF["atomic-He-gga-r"]["total"]) < 1e-9, r["E"]
    u = O.scf_uhf(b, 1, 1, xc_pol=lambda Pa, Pb: b.xc_pol(Pa, Pb, x_func=101, c_func=130)[:3])
    assert abs(u["E"] - REF["atomic-He-gga-u"]["total"]) < 1e-9, u["E"]


# ---- diatomic grids: pure-m (dftgrid_purem.cpp:87-199, 248-260, 382-395,
# 468-500) and the general 3D grid (src/diatomic/dftgrid.cpp:75-100, 292-305,
# 414-450; basis.cpp:2204-2243)

@pytest.fixture(scope="module")
def nh_basis():
    b = O.DiatomicBasis(7, 1, 1.958, [5, 4, 3], 3, nnodes=8)
    b.compute_tei()
    return b


@pytest.fixture(scope="module")
def nh_small():
    """a smaller basis for the 3D grid, whose lang x mang rule is the costly one"""
    b = O.DiatomicBasis(7, 1, 1.958, [3, 3, 2], 2, nnodes=7)
    b.compute_tei()
    return b


def _m_diagonal(b, P):
    Pd = np.zeros_like(P)
    for blk in b.sym_idx(1):
        Pd[np.ix_(blk, blk)] = P[np.ix_(blk, blk)]
    return Pd


@pytest.mark.parametrize("grid", ["purem", "3d"])
def test_diatomic_gga_matrix_is_energy_derivative(nh_basis, nh_small, grid):
    """<H, D> = dExc(P + h D)/dh on both diatomic grids, restricted and both
    spins: the mu and nu derivatives, the scale factor h_mu = h_nu and, on the
    3D grid, the phi component carry their factors correctly."""
    b = nh_basis if grid == "purem" else nh_small
    P, D = _density(b, 21)
    if grid == "purem":
        P, D = _m_diagonal(b, P), _m_diagonal(b, D)
        xc = lambda A: b.xc(A, x_func=101, c_func=130)
        xcp = lambda A, B: b.xc_pol(A, B, x_func=101, c_func=130)
    else:
        xc = lambda A: b.xc3d(A, x_func=101, c_func=130)
        xcp = lambda A, B: b.xc3d_pol(A, B, x_func=101, c_func=130)
    H, E, _ = xc(P)
    h = 1e-5
    fd = (xc(P + h * D)[1] - xc(P - h * D)[1]) / (2 * h)
    assert abs(fd / np.sum(H * D) - 1) < 1e-8
    Pa, Pb = 0.6 * P, 0.4 * P
    Ha, Hb, _, _ = xcp(Pa, Pb)
    for which, Hs in ((0, Ha), (1, Hb)):
        e = [xcp(Pa + s * h * D, Pb)[2] if which == 0 else xcp(Pa, Pb + s * h * D)[2] for s in (1, -1)]
        assert abs((e[0] - e[1]) / (2 * h) / np.sum(Hs * D) - 1) < 1e-8, which
    Ha, Hb, E2, _ = xcp(0.5 * P, 0.5 * P)
    assert abs(E2 - E) < 1e-13 * abs(E) and np.max(np.abs(Ha - H)) < 1e-13 * np.max(np.abs(H))


def test_3d_grid_gga_equals_purem_on_m_diagonal_density(nh_basis):
    b = nh_basis
    P = _m_diagonal(b, _density(b, 22)[0])
    H3, E3, N3 = b.xc3d(P, x_func=101, c_func=130)
    Hp, Ep, Np = b.xc(P, x_func=101, c_func=130)
    assert np.max(np.abs(H3 - Hp)) < 1e-11 * np.max(np.abs(Hp))
    assert abs(E3 - Ep) < 1e-12 * abs(Ep) and abs(N3 - Np) < 1e-12 * abs(Np)


def test_ne_dummy_centre_pbe_energy():
    """diatomic-Ne-dummy-gga (tests/refs/ci.json, -128.866426259): Ne through
    the diatomic code with a zero-charge centre at 1 bohr, lmax 13 (m=0) and 9
    (|m|=1), 3 elements, PBE on the pure-m grid."""
    b = O.DiatomicBasis(10, 0, 1.0, [13, 9], 3)
    b.compute_tei()
    r = O.scf_rhf(b, 10, xc=lambda P: b.xc(P, x_func=101, c_func=130)[:2])
    assert abs(r["E"] - REF["diatomic-Ne-dummy-gga"]["total"]) < 1e-9, r["E"]


# ---- meta-GGA: TPSS exchange (mgga_x_tpss, 202) and correlation (mgga_c_tpss,
# 231), Tao, Perdew, Staroverov, Scuseria, PRL 91, 146401 (2003); the grids'
# kinetic energy density (dftgrid.cpp:89-107, 336-349; dftgrid_purem.cpp:
# 287-318, 482-494)

def _tau_unif(rho):
    return 0.3 * (3 * np.pi ** 2) ** (2 / 3) * rho ** (5 / 3)


def test_tpss_uniform_gas_limits():
    """zero gradient and tau = tau_unif (alpha = 1, z = 0, p = 0): TPSS exchange
    is Slater's and TPSS correlation is PW92 (modified constants)."""
    rho = np.geomspace(1e-5, 1e3, 9)
    z, tu = np.zeros_like(rho), _tau_unif(rho)
    ex, vx, _, _ = O.xc_eval(rho, z, 202, 0, tau=tu)
    e0, v0 = O.lda(rho, 1, 0)
    assert np.max(np.abs(ex / e0 - 1)) < 1e-13
    ec, _, _, _ = O.xc_eval(rho, z, 0, 231, tau=tu)
    ep, _, _ = O.xc_eval(rho, z, 0, 13)
    assert np.max(np.abs(ec / ep - 1)) < 1e-13


def test_tpss_correlation_vanishes_for_one_electron_densities():
    """A fully polarized density with tau = tau_W (one orbital) has no TPSS
    correlation energy: revPKZB's self-interaction correction cancels PBE's."""
    rho = np.geomspace(1e-4, 10, 8)
    sigma = np.random.default_rng(3).uniform(0.1, 2.0, rho.size) * rho ** (8 / 3)
    r2 = np.stack([rho, np.zeros_like(rho)])
    s3 = np.stack([sigma, np.zeros_like(rho), np.zeros_like(rho)])
    t2 = np.stack([sigma / (8 * rho), np.zeros_like(rho)])
    ec, _, _, _ = O.xc_eval(r2, s3, 0, 231, tau=t2)
    ep, _, _ = O.xc_eval(r2, s3, 0, 130)
    assert np.max(np.abs(ec)) < 1e-10 * np.max(np.abs(ep)), ec  # the 1e-20 tau floor moves z off 1 by ~1e-13


@pytest.mark.parametrize("xf,cf", [(202, 0), (0, 231)])
def test_tpss_potentials_are_energy_derivatives(xf, cf):
    rng = np.random.default_rng(12)
    rho = np.geomspace(1e-3, 20, 12)
    sigma = rng.uniform(0.05, 1.0, rho.size) * rho ** (8 / 3)
    tau = sigma / (8 * rho) * rng.uniform(1.2, 3.0, rho.size)   # z = tau_W / tau in (1/3, 5/6)
    e, vr, vs, vt = O.xc_eval(rho, sigma, xf, cf, tau=tau)
    for k, (x, v) in enumerate(((rho, vr), (sigma, vs), (tau, vt))):
        def f(y, k=k):
            a = [rho, sigma, tau]
            a[k] = y
            return a[0] * O.xc_eval(a[0], a[1], xf, cf, tau=a[2])[0]
        fd = _fd(f, x, 1e-4 * x)
        assert np.max(np.abs(fd - v) / np.abs(v)) < 1e-7, k
    r2 = np.stack([0.6 * rho, 0.4 * rho])
    s3 = np.stack([0.36 * sigma, 0.2 * sigma, 0.16 * sigma])
    t2 = np.stack([0.55 * tau, 0.45 * tau])
    e, vr, vs, vt = O.xc_eval(r2, s3, xf, cf, tau=t2)
    for k in range(7):
        def f(y, k=k):
            a, b, c = r2.copy(), s3.copy(), t2.copy()
            (a if k < 2 else b if k < 5 else c)[k if k < 2 else k - 2 if k < 5 else k - 5] = y
            return (a[0] + a[1]) * O.xc_eval(a, b, xf, cf, tau=c)[0]
        x = r2[k] if k < 2 else s3[k - 2] if k < 5 else t2[k - 5]
        v = vr[k] if k < 2 else vs[k - 2] if k < 5 else vt[k - 5]
        fd = _fd(f, x, 1e-4 * x)
        assert np.max(np.abs(fd - v) / (np.abs(v) + 1e-300)) < 1e-7, k


@pytest.mark.parametrize("grid", ["atomic", "purem", "3d"])
def test_tpss_matrix_is_energy_derivative(ne_basis, nh_basis, nh_small, grid):
    """<H, D> = dExc(P + h D)/dh with TPSS on every grid: the kinetic energy
    density (radial, angular and, on the diatomic grids, the m^2/h_phi^2 term)
    and the vtau/2 potential terms carry their factors correctly."""
    if grid == "atomic":
        b, P, D = ne_basis, *_density(ne_basis, 31)
        xc = lambda A: b.xc(A, x_func=202, c_func=231)
        xcp = lambda A, B: b.xc_pol(A, B, x_func=202, c_func=231)
    elif grid == "purem":
        b = nh_basis
        P, D = (_m_diagonal(b, M) for M in _density(b, 32))
        xc = lambda A: b.xc(A, x_func=202, c_func=231)
        xcp = lambda A, B: b.xc_pol(A, B, x_func=202, c_func=231)
    else:
        b, P, D = nh_small, *_density(nh_small, 33)
        xc = lambda A: b.xc3d(A, x_func=202, c_func=231)
        xcp = lambda A, B: b.xc3d_pol(A, B, x_func=202, c_func=231)
    # TPSS is only piecewise smooth (the z <= 1 cap of libxc), so the bound is
    # looser than for PBE; a missing factor would show at order one
    H, E, _ = xc(P)
    h = 1e-5
    fd = (xc(P + h * D)[1] - xc(P - h * D)[1]) / (2 * h)
    assert abs(fd / np.sum(H * D) - 1) < 1e-6
    Pa, Pb = 0.6 * P, 0.4 * P
    Ha, Hb, _, _ = xcp(Pa, Pb)
    for which, Hs in ((0, Ha), (1, Hb)):
        e = [xcp(Pa + s * h * D, Pb)[2] if which == 0 else xcp(Pa, Pb + s * h * D)[2] for s in (1, -1)]
        assert abs((e[0] - e[1]) / (2 * h) / np.sum(Hs * D) - 1) < 1e-6, which


def test_he_tpss_energies():
    b = O.AtomicBasis(2, 0, 0, 5)
    b.compute_tei()
    r = O.scf_rhf(b, 2, xc=lambda P: b.xc(P, x_func=202, c_func=231)[:2])
    assert abs(r["E"] - REF["atomic-He-mgga-r"]["total"]) < 1e-9, r["E"]
    u = O.scf_uhf(b, 1, 1, xc_pol=lambda Pa, Pb: b.xc_pol(Pa, Pb, x_func=202, c_func=231)[:3])
    assert abs(u["E"] - REF["atomic-He-mgga-u"]["total"]) < 1e-9, u["E"]
    # the kinetic energy density integrates to the kinetic energy of the orbitals
    _, _, _ = b.xc(r["P"], x_func=202, c_func=231)
    T = np.sum(r["P"] * b.kinetic())
    assert abs(b.last_ekin - T) < 1e-6 * T, (b.last_ekin, T)


def test_ne_dummy_centre_tpss_energy():
    """diatomic-Ne-dummy-mgga (tests/refs/ci.json, -128.981103006): TPSS on the
    pure-m grid, Z1 = 10, Z2 = 0, R = 1 bohr, lmax 13 (m=0) and 9 (|m|=1)."""
    b = O.DiatomicBasis(10, 0, 1.0, [13, 9], 3)
    b.compute_tei()
    r = O.scf_rhf(b, 10, xc=lambda P: b.xc(P, x_func=202, c_func=231)[:2])
    assert abs(r["E"] - REF["diatomic-Ne-dummy-mgga"]["total"]) < 1e-9, r["E"]
