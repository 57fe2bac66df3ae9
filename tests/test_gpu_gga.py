"""GPU parity of the GGA exchange-correlation path (PBE, and the other
functional pairs the general point kernel evaluates) against the oracle, on
the atomic, pure-m diatomic and 3D diatomic grids.

The oracle restates the reference's dense grids with the gradient terms
(src/atomic/dftgrid.cpp:51-128, 304-359, 470-576; src/diatomic/
dftgrid_purem.cpp:87-500; src/diatomic/dftgrid.cpp:75-100, 292-305, 414-450;
dftgrid_common.h:188-216)
and differentiates the functionals by complex step; the kernels contract the
factorised basis one index at a time and differentiate by forward-mode dual
numbers, so the two agree to rounding.  Bar: max|H_gpu - H_oracle| /
max|H_oracle| <= 1e-12, Exc and Nel within 1e-12 relative; the PBE SCF
energies of He (restricted and unrestricted) within 1e-9 Eh of the
reference's -2.8929348668 (tests/refs/ci.json atomic-He-gga-r/-u) and of Ne
through the diatomic code, -128.866426259 (diatomic-Ne-dummy-gga); TPSS
likewise; the B3LYP SCF of He within 1e-9 Eh of -2.9152186629
(atomic-He-hybrid-r/-u) with 0.20 exact exchange.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import helfem_amd as H
import oracle as O
from helfem_amd.synthetic import seeded_density

pytestmark = pytest.mark.gpu
TOL = 1e-12
REF = json.loads((Path(__file__).parent / "golden" / "reference_energies.json").read_text())

CASES = {
    "He-l0-5el": dict(Z=2, lmax=0, mmax=0, nelem=5),
    "Ne-l2m1-3el-n8": dict(Z=10, lmax=2, mmax=1, nelem=3, nnodes=8),
    "Ar-l2m2-3el": dict(Z=18, lmax=2, mmax=2, nelem=3),
    "Kr-l3m3-2el-n10": dict(Z=36, lmax=3, mmax=3, nelem=2, nnodes=10),
}
PAIRS = [(101, 130), (101, 0), (0, 130), (1, 13), (1, 12), (101, 7)]


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _bases(name, ctx):
    c = dict(CASES[name])
    args = (c.pop("Z"), c.pop("lmax"), c.pop("mmax"), c.pop("nelem"))
    return O.AtomicBasis(*args, **c), H.AtomicBasis(*args, ctx=ctx, **c)


def _density(ob, seed, scale=1.0, nocc=1):
    S = ob.overlap()
    blocks = ob.sym_idx(1)
    return scale * seeded_density(S, blocks, [nocc] * len(blocks), seed=seed)


@pytest.mark.parametrize("name", list(CASES))
def test_gga_matrix_matches_oracle(name, gpu_ctx):
    print(f"[gpu] atomic GGA Vxc: {name}", flush=True)
    ob, hb = _bases(name, gpu_ctx)
    P = _density(ob, 2003)
    grid = H.DFTGrid(hb)
    for xf, cf in PAIRS if name.startswith("Ne") else PAIRS[:1]:
        Hr, Er, Nr = ob.xc(P, x_func=xf, c_func=cf)
        Hg, Eg, Ng, Kg = grid.eval_Fxc(xf, cf, P)
        print(f"[gpu]   ({xf},{cf}) H {relerr(Hg, Hr):.2e} Exc {abs(Eg - Er) / abs(Er):.2e}", flush=True)
        assert relerr(Hg, Hr) <= TOL, (xf, cf, relerr(Hg, Hr))
        assert abs(Eg - Er) <= TOL * abs(Er) and abs(Ng - Nr) <= TOL * abs(Nr)
        assert Kg == 0.0


@pytest.mark.parametrize("name", ["He-l0-5el", "Ne-l2m1-3el-n8", "Kr-l3m3-2el-n10"])
def test_polarized_gga_matrix_matches_oracle(name, gpu_ctx):
    print(f"[gpu] atomic polarized GGA Vxc: {name}", flush=True)
    ob, hb = _bases(name, gpu_ctx)
    Pa, Pb = _density(ob, 2105), _density(ob, 2106, 0.6)
    grid = H.DFTGrid(hb)
    Har, Hbr, Er, Nr = ob.xc_pol(Pa, Pb, x_func=101, c_func=130)
    Hag, Hbg, Eg, Ng, _ = grid.eval_Fxc_pol(101, 130, Pa, Pb)
    print(f"[gpu]   Ha {relerr(Hag, Har):.2e} Hb {relerr(Hbg, Hbr):.2e}", flush=True)
    assert relerr(Hag, Har) <= TOL and relerr(Hbg, Hbr) <= TOL
    assert abs(Eg - Er) <= TOL * abs(Er) and abs(Ng - Nr) <= TOL * abs(Nr)
    # fully polarized (no beta density), as for the H atom
    Har, Hbr, Er, _ = ob.xc_pol(Pa, 0 * Pb, x_func=101, c_func=130)
    Hag, Hbg, Eg, _, _ = grid.eval_Fxc_pol(101, 130, Pa, 0 * Pb)
    assert relerr(Hag, Har) <= TOL and abs(Eg - Er) <= TOL * abs(Er)
    assert np.max(np.abs(Hbg - Hbr)) <= TOL * np.max(np.abs(Har))


DIA = {
    "N2-l432-3el": (7, 7, 2.074, [4, 3, 2], 3),
    "LiH-l32-4el": (3, 1, 3.0, [3, 2], 4),
    "HeH-l21-1el": (2, 1, 1.5, [2, 1], 1),
}


@pytest.mark.parametrize("name", list(DIA))
def test_purem_gga_matrix_matches_oracle(name, gpu_ctx):
    """Pure-m diatomic grid (dftgrid_purem.cpp:248-260, 382-395, 468-500):
    mu and nu gradient components with h_mu = h_nu, restricted and polarized."""
    print(f"[gpu] diatomic pure-m GGA Vxc: {name}", flush=True)
    Z1, Z2, R, lmm, nel = DIA[name]
    ob = O.DiatomicBasis(Z1, Z2, R, lmm, nel)
    hb = H.DiatomicBasis(Z1, Z2, R, lmm, nel, ctx=gpu_ctx)
    P, Pb = _density(ob, 2205), _density(ob, 2206, 0.5)
    grid = H.DFTGrid(hb)
    Hr, Er, Nr = ob.xc(P, x_func=101, c_func=130)
    Hg, Eg, Ng, _ = grid.eval_Fxc(101, 130, P)
    assert relerr(Hg, Hr) <= TOL, relerr(Hg, Hr)
    assert abs(Eg - Er) <= TOL * abs(Er) and abs(Ng - Nr) <= TOL * abs(Nr)
    Har, Hbr, Er, Nr = ob.xc_pol(P, Pb, x_func=101, c_func=130)
    Hag, Hbg, Eg, Ng, _ = grid.eval_Fxc_pol(101, 130, P, Pb)
    assert relerr(Hag, Har) <= TOL and relerr(Hbg, Hbr) <= TOL, (relerr(Hag, Har), relerr(Hbg, Hbr))
    assert abs(Eg - Er) <= TOL * abs(Er) and abs(Ng - Nr) <= TOL * abs(Nr)


@pytest.mark.parametrize("name", ["LiH-l32-4el", "HeH-l21-1el"])
def test_3d_gga_matrix_matches_oracle(name, gpu_ctx):
    """General 3D diatomic grid (src/diatomic/dftgrid.cpp:75-100, 292-305):
    mu, nu and phi components on a density that mixes m."""
    from helfem_amd.synthetic import random_symmetric
    print(f"[gpu] diatomic 3D-grid GGA Vxc: {name}", flush=True)
    Z1, Z2, R, lmm, nel = DIA[name]
    ob = O.DiatomicBasis(Z1, Z2, R, lmm, nel)
    hb = H.DiatomicBasis(Z1, Z2, R, lmm, nel, ctx=gpu_ctx)
    S = ob.overlap()
    X = random_symmetric(ob.Nbf, seed=1411)
    P = X @ S @ X.T
    P *= 4.0 / np.sum(P * S)
    grid = H.DFTGrid(hb, purem=False)
    Hr, Er, Nr = ob.xc3d(P, x_func=101, c_func=130)
    Hg, Eg, Ng, _ = grid.eval_Fxc(101, 130, P)
    assert relerr(Hg, Hr) <= TOL, relerr(Hg, Hr)
    assert abs(Eg - Er) <= TOL * abs(Er) and abs(Ng - Nr) <= TOL * abs(Nr)
    Pb = 0.3 * P
    Har, Hbr, Er, Nr = ob.xc3d_pol(P, Pb, x_func=101, c_func=130)
    Hag, Hbg, Eg, Ng, _ = grid.eval_Fxc_pol(101, 130, P, Pb)
    assert relerr(Hag, Har) <= TOL and relerr(Hbg, Hbr) <= TOL, (relerr(Hag, Har), relerr(Hbg, Hbr))
    assert abs(Eg - Er) <= TOL * abs(Er) and abs(Ng - Nr) <= TOL * abs(Nr)


def test_ne_dummy_centre_pbe_scf_on_gpu(gpu_ctx):
    """diatomic-Ne-dummy-gga: PBE on the pure-m grid, Z1=10, Z2=0, R=1 bohr, 3
    elements, lmax 13 (m=0), 9 (|m|=1); reference -128.866426259."""
    import torch
    from helfem_amd.scf import rhf
    print("[gpu] diatomic Ne dummy-centre PBE SCF", flush=True)
    hb = H.DiatomicBasis(10, 0, 1.0, [13, 9], 3, ctx=gpu_ctx)
    hb.compute_tei()
    grid = H.DFTGrid(hb)
    r = rhf(hb, 10, xc=lambda P: grid.eval_Fxc(101, 130, P)[:2], conv=1e-12)
    torch.cuda.synchronize()
    ref = REF["diatomic-Ne-dummy-gga"]["total"]
    print(f"[gpu]   E = {r.E:.10f} (reference {ref})", flush=True)
    assert abs(r.E - ref) < 1e-9, (r.E, ref)


@pytest.mark.parametrize("restricted", [True, False], ids=["r", "u"])
def test_he_pbe_scf_on_gpu(restricted, gpu_ctx):
    import torch
    from helfem_amd.scf import rhf, uhf
    name = "atomic-He-gga-" + ("r" if restricted else "u")
    print(f"[gpu] atomic PBE SCF: {name}", flush=True)
    hb = H.AtomicBasis(2, 0, 0, 5, ctx=gpu_ctx)
    hb.compute_tei()
    grid = H.DFTGrid(hb)
    if restricted:
        r = rhf(hb, 2, xc=lambda P: grid.eval_Fxc(101, 130, P)[:2], conv=1e-12)
    else:
        def xc_pol(Pa, Pb):
            Ha, Hb, Exc, _, _ = grid.eval_Fxc_pol(101, 130, Pa, Pb)
            return Ha, Hb, Exc
        r = uhf(hb, 1, 1, xc_pol=xc_pol, conv=1e-12)
    torch.cuda.synchronize()
    print(f"[gpu]   E = {r.E:.10f} (reference {REF[name]['total']})", flush=True)
    assert abs(r.E - REF[name]["total"]) < 1e-9, (r.E, REF[name]["total"])


# ---- meta-GGA (TPSS, mgga_x_tpss 202 + mgga_c_tpss 231): kinetic energy
# density and the vtau terms on every grid; Ekin (the integral of tau) is the
# third output, as DFTGrid::eval_Fxc returns it.  The densities hold two
# orbitals per block: a one-orbital density has tau = tau_W exactly, z = 1, on
# libxc's sigma <= 8 rho tau cap, where the split of the potential between
# vsigma and vtau follows the rounding (equal energies, different matrices)

@pytest.mark.parametrize("name", ["He-l0-5el", "Ne-l2m1-3el-n8", "Kr-l3m3-2el-n10"])
def test_tpss_matrix_matches_oracle(name, gpu_ctx):
    print(f"[gpu] atomic TPSS Vxc: {name}", flush=True)
    ob, hb = _bases(name, gpu_ctx)
    P = _density(ob, 2403, nocc=2)
    grid = H.DFTGrid(hb)
    Hr, Er, Nr = ob.xc(P, x_func=202, c_func=231)
    Kr = ob.last_ekin
    Hg, Eg, Ng, Kg = grid.eval_Fxc(202, 231, P)
    print(f"[gpu]   H {relerr(Hg, Hr):.2e} Exc {abs(Eg - Er) / abs(Er):.2e} Ekin {abs(Kg - Kr) / abs(Kr):.2e}", flush=True)
    assert relerr(Hg, Hr) <= TOL, relerr(Hg, Hr)
    assert abs(Eg - Er) <= TOL * abs(Er) and abs(Ng - Nr) <= TOL * abs(Nr) and abs(Kg - Kr) <= TOL * abs(Kr)
    Pb = _density(ob, 2404, 0.6, nocc=2)
    Har, Hbr, Er, Nr = ob.xc_pol(P, Pb, x_func=202, c_func=231)
    Kr = ob.last_ekin
    Hag, Hbg, Eg, Ng, Kg = grid.eval_Fxc_pol(202, 231, P, Pb)
    print(f"[gpu]   Ha {relerr(Hag, Har):.2e} Hb {relerr(Hbg, Hbr):.2e}", flush=True)
    assert relerr(Hag, Har) <= TOL and relerr(Hbg, Hbr) <= TOL
    assert abs(Eg - Er) <= TOL * abs(Er) and abs(Kg - Kr) <= TOL * abs(Kr)


@pytest.mark.parametrize("name", ["N2-l432-3el", "LiH-l32-4el"])
def test_diatomic_tpss_matrix_matches_oracle(name, gpu_ctx):
    """Pure-m grid (tau's m^2 / h_phi^2 term, dftgrid_purem.cpp:300-318) and the
    3D grid, restricted and polarized."""
    from helfem_amd.synthetic import random_symmetric
    print(f"[gpu] diatomic TPSS Vxc: {name}", flush=True)
    Z1, Z2, R, lmm, nel = DIA[name]
    ob = O.DiatomicBasis(Z1, Z2, R, lmm, nel)
    hb = H.DiatomicBasis(Z1, Z2, R, lmm, nel, ctx=gpu_ctx)
    P, Pb = _density(ob, 2405, nocc=2), _density(ob, 2406, 0.5, nocc=2)
    grid = H.DFTGrid(hb)
    Hr, Er, _ = ob.xc(P, x_func=202, c_func=231)
    Kr = ob.last_ekin
    Hg, Eg, _, Kg = grid.eval_Fxc(202, 231, P)
    assert relerr(Hg, Hr) <= TOL and abs(Eg - Er) <= TOL * abs(Er) and abs(Kg - Kr) <= TOL * abs(Kr)
    Har, Hbr, Er, _ = ob.xc_pol(P, Pb, x_func=202, c_func=231)
    Hag, Hbg, Eg, _, _ = grid.eval_Fxc_pol(202, 231, P, Pb)
    assert relerr(Hag, Har) <= TOL and relerr(Hbg, Hbr) <= TOL and abs(Eg - Er) <= TOL * abs(Er)
    S = ob.overlap()
    X = random_symmetric(ob.Nbf, seed=1511)
    P3 = X @ S @ X.T
    P3 *= 4.0 / np.sum(P3 * S)
    g3 = H.DFTGrid(hb, purem=False)
    Hr, Er, _ = ob.xc3d(P3, x_func=202, c_func=231)
    Hg, Eg, _, _ = g3.eval_Fxc(202, 231, P3)
    print(f"[gpu]   3D grid H {relerr(Hg, Hr):.2e}", flush=True)
    assert relerr(Hg, Hr) <= TOL and abs(Eg - Er) <= TOL * abs(Er)
    Har, Hbr, Er, _ = ob.xc3d_pol(P3, 0.3 * P3, x_func=202, c_func=231)
    Hag, Hbg, Eg, _, _ = g3.eval_Fxc_pol(202, 231, P3, 0.3 * P3)
    assert relerr(Hag, Har) <= TOL and relerr(Hbg, Hbr) <= TOL and abs(Eg - Er) <= TOL * abs(Er)


@pytest.mark.parametrize("restricted", [True, False], ids=["r", "u"])
def test_he_tpss_scf_on_gpu(restricted, gpu_ctx):
    import torch
    from helfem_amd.scf import rhf, uhf
    name = "atomic-He-mgga-" + ("r" if restricted else "u")
    print(f"[gpu] atomic TPSS SCF: {name}", flush=True)
    hb = H.AtomicBasis(2, 0, 0, 5, ctx=gpu_ctx)
    hb.compute_tei()
    grid = H.DFTGrid(hb)
    if restricted:
        r = rhf(hb, 2, xc=lambda P: grid.eval_Fxc(202, 231, P)[:2], conv=1e-12)
    else:
        r = uhf(hb, 1, 1, xc_pol=lambda Pa, Pb: grid.eval_Fxc_pol(202, 231, Pa, Pb)[:3], conv=1e-12)
    torch.cuda.synchronize()
    print(f"[gpu]   E = {r.E:.10f} (reference {REF[name]['total']})", flush=True)
    assert abs(r.E - REF[name]["total"]) < 1e-9, (r.E, REF[name]["total"])


def test_ne_dummy_centre_tpss_scf_on_gpu(gpu_ctx):
    """diatomic-Ne-dummy-mgga: TPSS on the pure-m grid; reference -128.981103006."""
    import torch
    from helfem_amd.scf import rhf
    print("[gpu] diatomic Ne dummy-centre TPSS SCF", flush=True)
    hb = H.DiatomicBasis(10, 0, 1.0, [13, 9], 3, ctx=gpu_ctx)
    hb.compute_tei()
    grid = H.DFTGrid(hb)
    r = rhf(hb, 10, xc=lambda P: grid.eval_Fxc(202, 231, P)[:2], conv=1e-12)
    torch.cuda.synchronize()
    ref = REF["diatomic-Ne-dummy-mgga"]["total"]
    print(f"[gpu]   E = {r.E:.10f} (reference {ref})", flush=True)
    assert abs(r.E - ref) < 1e-9, (r.E, ref)


# ---- B3LYP pieces and the global hybrids: B88 exchange (106), LYP (131),
# VWN RPA (8), hyb_gga_xc_b3lyp (402) and hyb_gga_xc_pbeh (406, PBE0), whose
# density-functional part the grid evaluates; the SCF adds the exact-exchange
# fraction (hf_xc_exact_exchange) of exchange(P), as src/atomic/main.cpp:424-437
HYB_PAIRS = [(402, 0), (406, 0), (106, 131), (0, 8)]


@pytest.mark.parametrize("name", ["He-l0-5el", "Ne-l2m1-3el-n8", "Kr-l3m3-2el-n10"])
def test_hybrid_xc_matrix_matches_oracle(name, gpu_ctx):
    print(f"[gpu] atomic B3LYP/PBE0 Vxc: {name}", flush=True)
    ob, hb = _bases(name, gpu_ctx)
    P, Pb = _density(ob, 2303), _density(ob, 2304, 0.4)
    grid = H.DFTGrid(hb)
    for xf, cf in HYB_PAIRS:
        Hr, Er, Nr = ob.xc(P, x_func=xf, c_func=cf)
        Hg, Eg, Ng, _ = grid.eval_Fxc(xf, cf, P)
        print(f"[gpu]   ({xf},{cf}) H {relerr(Hg, Hr):.2e} Exc {abs(Eg - Er) / abs(Er):.2e}", flush=True)
        assert relerr(Hg, Hr) <= TOL, (xf, cf, relerr(Hg, Hr))
        assert abs(Eg - Er) <= TOL * abs(Er) and abs(Ng - Nr) <= TOL * abs(Nr)
        Har, Hbr, Er, _ = ob.xc_pol(P, Pb, x_func=xf, c_func=cf)
        Hag, Hbg, Eg, _, _ = grid.eval_Fxc_pol(xf, cf, P, Pb)
        assert relerr(Hag, Har) <= TOL and relerr(Hbg, Hbr) <= TOL, (xf, cf, relerr(Hag, Har), relerr(Hbg, Hbr))
        assert abs(Eg - Er) <= TOL * abs(Er)
    # fully polarized B3LYP (LYP vanishes, B88 of one spin)
    Har, Hbr, Er, _ = ob.xc_pol(P, 0 * Pb, x_func=402, c_func=0)
    Hag, Hbg, Eg, _, _ = grid.eval_Fxc_pol(402, 0, P, 0 * Pb)
    assert relerr(Hag, Har) <= TOL and abs(Eg - Er) <= TOL * abs(Er)


@pytest.mark.parametrize("name", ["N2-l432-3el", "LiH-l32-4el"])
def test_diatomic_hybrid_xc_matrix_matches_oracle(name, gpu_ctx):
    """B3LYP's density-functional part on the pure-m and 3D diatomic grids"""
    from helfem_amd.synthetic import random_symmetric
    print(f"[gpu] diatomic B3LYP Vxc: {name}", flush=True)
    Z1, Z2, R, lmm, nel = DIA[name]
    ob = O.DiatomicBasis(Z1, Z2, R, lmm, nel)
    hb = H.DiatomicBasis(Z1, Z2, R, lmm, nel, ctx=gpu_ctx)
    P, Pb = _density(ob, 2405), _density(ob, 2406, 0.5)
    grid = H.DFTGrid(hb)
    Hr, Er, Nr = ob.xc(P, x_func=402, c_func=0)
    Hg, Eg, Ng, _ = grid.eval_Fxc(402, 0, P)
    assert relerr(Hg, Hr) <= TOL, relerr(Hg, Hr)
    assert abs(Eg - Er) <= TOL * abs(Er) and abs(Ng - Nr) <= TOL * abs(Nr)
    Har, Hbr, Er, _ = ob.xc_pol(P, Pb, x_func=402, c_func=0)
    Hag, Hbg, Eg, _, _ = grid.eval_Fxc_pol(402, 0, P, Pb)
    assert relerr(Hag, Har) <= TOL and relerr(Hbg, Hbr) <= TOL, (relerr(Hag, Har), relerr(Hbg, Hbr))
    if name != "LiH-l32-4el":
        return
    S = ob.overlap()
    X = random_symmetric(ob.Nbf, seed=1412)
    P3 = X @ S @ X.T
    P3 *= 4.0 / np.sum(P3 * S)
    g3 = H.DFTGrid(hb, purem=False)
    Hr, Er, Nr = ob.xc3d(P3, x_func=402, c_func=0)
    Hg, Eg, Ng, _ = g3.eval_Fxc(402, 0, P3)
    assert relerr(Hg, Hr) <= TOL, relerr(Hg, Hr)
    assert abs(Eg - Er) <= TOL * abs(Er) and abs(Ng - Nr) <= TOL * abs(Nr)


@pytest.mark.parametrize("restricted", [True, False], ids=["r", "u"])
def test_he_b3lyp_scf_on_gpu(restricted, gpu_ctx):
    """atomic-He-hybrid-r/-u: B3LYP with 0.20 exact exchange; reference -2.9152186629"""
    import torch
    from helfem_amd.dftgrid import exact_exchange
    from helfem_amd.scf import rhf, uhf
    name = "atomic-He-hybrid-" + ("r" if restricted else "u")
    print(f"[gpu] atomic B3LYP SCF: {name}", flush=True)
    hb = H.AtomicBasis(2, 0, 0, 5, ctx=gpu_ctx)
    hb.compute_tei()
    grid = H.DFTGrid(hb)
    kfrac = exact_exchange(402)[0]
    if restricted:
        r = rhf(hb, 2, xc=lambda P: grid.eval_Fxc(402, 0, P)[:2], conv=1e-12, kfrac=kfrac)
    else:
        r = uhf(hb, 1, 1, xc_pol=lambda Pa, Pb: grid.eval_Fxc_pol(402, 0, Pa, Pb)[:3], conv=1e-12, kfrac=kfrac)
    torch.cuda.synchronize()
    print(f"[gpu]   E = {r.E:.10f} (reference {REF[name]['total']}), Exx {r.Exx:.10f}", flush=True)
    assert abs(r.E - REF[name]["total"]) < 1e-9, (r.E, REF[name]["total"])
    assert abs(r.Exx - REF[name]["Exx"]) < 1e-6  # components carry first-order density error


def test_ne_b3lyp_scf_matches_oracle(gpu_ctx):
    """Ne B3LYP (lmax 1, 5 elements): GPU SCF equal to the oracle's to 1e-9 Eh"""
    import torch
    from helfem_amd.dftgrid import exact_exchange
    from helfem_amd.scf import rhf
    print("[gpu] atomic Ne B3LYP SCF vs oracle", flush=True)
    kfrac = exact_exchange(402)[0]
    ob = O.AtomicBasis(10, 1, 1, 5)
    ob.compute_tei()
    ro = O.scf_rhf(ob, 10, xc=lambda P: ob.xc(P, x_func=402, c_func=0)[:2], kfrac=kfrac)
    hb = H.AtomicBasis(10, 1, 1, 5, ctx=gpu_ctx)
    hb.compute_tei()
    grid = H.DFTGrid(hb)
    r = rhf(hb, 10, xc=lambda P: grid.eval_Fxc(402, 0, P)[:2], conv=1e-12, kfrac=kfrac)
    torch.cuda.synchronize()
    print(f"[gpu]   E = {r.E:.10f} oracle {ro['E']:.10f}", flush=True)
    assert abs(r.E - ro["E"]) < 1e-9


# ---- range-separated hybrids: CAM-B3LYP (433, erfc, omega 0.33) and
# CAMY-B3LYP (470, Yukawa, omega 0.34); the SCF adds 0.65 exchange(P) - 0.46
# rs_exchange(P) on the basis' erfc or Yukawa caches (main.cpp:336-337, 424-437)

@pytest.mark.parametrize("name", ["He-l0-5el", "Ne-l2m1-3el-n8"])
def test_range_separated_xc_matrix_matches_oracle(name, gpu_ctx):
    print(f"[gpu] atomic CAM/CAMY-B3LYP Vxc: {name}", flush=True)
    ob, hb = _bases(name, gpu_ctx)
    P, Pb = _density(ob, 2503), _density(ob, 2504, 0.4)
    grid = H.DFTGrid(hb)
    for xf in (433, 470):
        Hr, Er, Nr = ob.xc(P, x_func=xf, c_func=0)
        Hg, Eg, Ng, _ = grid.eval_Fxc(xf, 0, P)
        print(f"[gpu]   ({xf}) H {relerr(Hg, Hr):.2e} Exc {abs(Eg - Er) / abs(Er):.2e}", flush=True)
        assert relerr(Hg, Hr) <= TOL and abs(Eg - Er) <= TOL * abs(Er)
        Har, Hbr, Er, _ = ob.xc_pol(P, Pb, x_func=xf, c_func=0)
        Hag, Hbg, Eg, _, _ = grid.eval_Fxc_pol(xf, 0, P, Pb)
        assert relerr(Hag, Har) <= TOL and relerr(Hbg, Hbr) <= TOL and abs(Eg - Er) <= TOL * abs(Er)


@pytest.mark.parametrize("xf,name,tol", [(470, "atomic-H-rsh-yukawa-u", 1e-9), (433, "atomic-H-rsh-erfc-u", 2e-8)])
def test_h_range_separated_scf_on_gpu(xf, name, tol, gpu_ctx):
    """atomic-H-rsh-yukawa-u (-0.4980726135) to 1e-9 Eh; atomic-H-rsh-erfc-u
    (-0.4990948394) to 2e-8 Eh, where the oracle lands (9.2e-9; the residual
    sits in the erfc short-range exchange, the Yukawa twin agreeing to 2e-10);
    both equal to the oracle's SCF to 1e-10 Eh"""
    import torch
    from helfem_amd.dftgrid import exact_exchange
    from helfem_amd.scf import uhf
    print(f"[gpu] atomic range-separated hybrid SCF: {name}", flush=True)
    kf, ks, om, ker = exact_exchange(xf)
    ob = O.AtomicBasis(1, 0, 0, 5)
    ob.compute_tei()
    (ob.compute_erfc if ker == "erfc" else ob.compute_yukawa)(om)
    ro = O.scf_uhf(ob, 1, 0, xc_pol=lambda Pa, Pb: ob.xc_pol(Pa, Pb, x_func=xf, c_func=0)[:3], kfrac=kf, kshort=ks)
    hb = H.AtomicBasis(1, 0, 0, 5, ctx=gpu_ctx)
    hb.compute_tei()
    (hb.compute_erfc if ker == "erfc" else hb.compute_yukawa)(om)
    grid = H.DFTGrid(hb)
    r = uhf(hb, 1, 0, xc_pol=lambda Pa, Pb: grid.eval_Fxc_pol(xf, 0, Pa, Pb)[:3], conv=1e-12, kfrac=kf, kshort=ks)
    torch.cuda.synchronize()
    print(f"[gpu]   E = {r.E:.10f} oracle {ro['E']:.10f} (reference {REF[name]['total']})", flush=True)
    assert abs(r.E - ro["E"]) < 1e-10
    assert abs(r.E - REF[name]["total"]) < tol, (r.E, REF[name]["total"])


def test_he_dummy_centre_b3lyp_scf_on_gpu(gpu_ctx):
    """diatomic-He-dummy-hybrid: He through the diatomic code (Z2 = 0 at 1
    bohr, 3 elements, lmax 6) with B3LYP must reproduce the atomic solver:
    equal to the oracle's diatomic SCF to 1e-10 Eh and within 2e-8 Eh of the
    atomic reference -2.9152186629 (the oracle lands 5.8e-9 away; the two
    codes' grids differ, as the Hartree-Fock pair differs by 8e-10)"""
    import torch
    from helfem_amd.dftgrid import exact_exchange
    from helfem_amd.scf import rhf
    print("[gpu] diatomic He dummy-centre B3LYP SCF", flush=True)
    kfrac = exact_exchange(402)[0]
    ob = O.DiatomicBasis(2, 0, 1.0, [6], 3)
    ob.compute_tei()
    ro = O.scf_rhf(ob, 2, xc=lambda P: ob.xc(P, x_func=402, c_func=0)[:2], kfrac=kfrac)
    hb = H.DiatomicBasis(2, 0, 1.0, [6], 3, ctx=gpu_ctx)
    hb.compute_tei()
    grid = H.DFTGrid(hb)
    r = rhf(hb, 2, xc=lambda P: grid.eval_Fxc(402, 0, P)[:2], conv=1e-12, kfrac=kfrac)
    torch.cuda.synchronize()
    ref = REF["atomic-He-hybrid-r"]["total"]
    print(f"[gpu]   E = {r.E:.10f} oracle {ro['E']:.10f} (atomic reference {ref})", flush=True)
    assert abs(r.E - ro["E"]) < 1e-10
    assert abs(r.E - ref) < 2e-8
