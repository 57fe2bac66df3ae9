"""Device-resident restricted SCF around the GPU Fock build.

Counterpart of the Fock-builder of src/atomic/main.cpp:369-470 together with
the SCF glue of src/general/scf_driver_common.h:
  * per-symmetry-block half-inverse overlap (build_per_block_Sinvh, :176-187;
    utils::invh, libhelfem/src/utils.cpp:120-159, chol=false),
  * P = sum_k Sinvh_k C diag(occ) C' Sinvh_k' (accumulate_density_block, :407-425),
  * Etot = tr(PT) + tr(PV) + 1/2 tr(PJ) + Exc + Exx, Exx = 1/2 tr(Pa Ka)
    doubled when restricted (assemble_hf_exchange, :460-474),
  * F_k = Sinvh_k' (H1 + J + XC + K) Sinvh_k (assemble_fock_blocks, :490-520).
The reference iterates with OpenOrbitalOptimizer (DIIS/ODA/CG); this driver
uses aufbau over all blocks and Pulay DIIS on the orthonormal commutators.
Converged energies depend only on the state reached, not on the iteration.

All matrices stay in HBM (torch float64); the per-block eigensolves run on
the GPU (torch.linalg.eigh), J and K through the HIP kernels of libhfgpu.
"""
from __future__ import annotations

import time

import numpy as np


def _invh(S):
    import torch
    d = 1.0 / torch.sqrt(torch.diagonal(S))
    Sn = d[:, None] * S * d[None, :]
    e, v = torch.linalg.eigh(Sn)
    return d[:, None] * ((v * (1.0 / torch.sqrt(e))) @ v.T)


class SCFResult(dict):
    __getattr__ = dict.get


def rhf(basis, nelec: int, symm: int = 1, conv: float = 1e-10, maxit: int = 100, diis_len: int = 10,
        verbose: bool = False, xc=None, device: str = "cuda", precision: str = "fp64", switch: float = 1e-3):
    """Closed-shell SCF; returns energies, iterations and the final density.

    xc(P) -> (Vxc, Exc) adds a density functional (None: Hartree-Fock).
    precision="mixed" builds the exchange in FP32 while the orbital gradient
    (the largest DIIS commutator element) is above `switch`, then restarts
    DIIS and finishes in FP64; overlap, orthogonalisation, eigensolves and
    energy sums are FP64 throughout."""
    import torch
    dev = torch.device(device)
    f64 = torch.float64
    t0 = time.perf_counter()
    S = torch.as_tensor(np.ascontiguousarray(basis.overlap()), dtype=f64, device=dev)
    T = torch.as_tensor(np.ascontiguousarray(basis.kinetic()), dtype=f64, device=dev)
    V = torch.as_tensor(np.ascontiguousarray(basis.nuclear()), dtype=f64, device=dev)
    H = T + V
    blocks = [torch.as_tensor(b, dtype=torch.long, device=dev) for b in basis.get_sym_idx(symm)]
    X = [_invh(S[b][:, b]) for b in blocks]
    nocc = nelec // 2
    have_exx = xc is None
    Enucr = float(getattr(basis, "Enucr", 0.0))  # nuclear repulsion of a diatomic basis

    def density(F):
        evs, vecs = [], []
        for k, b in enumerate(blocks):
            Fo = X[k].T @ F[b][:, b] @ X[k]
            e, c = torch.linalg.eigh(Fo)
            evs.append(e)
            vecs.append(c)
        ev_host = [e.cpu().numpy() for e in evs]
        order = sorted((float(e), k, i) for k, ev in enumerate(ev_host) for i, e in enumerate(ev))
        nk = [0] * len(blocks)
        for _, k, _i in order[:nocc]:
            nk[k] += 1
        P = torch.zeros_like(F)
        for k, b in enumerate(blocks):
            if nk[k]:
                C = X[k] @ vecs[k][:, :nk[k]]
                P[b[:, None], b[None, :]] += 2.0 * (C @ C.T)
        return P, nk, ev_host

    if precision not in ("fp64", "mixed"):
        raise ValueError(f"precision must be 'fp64' or 'mixed', not {precision!r}")
    fp32 = precision == "mixed"
    it_fp32 = 0
    P, nk, _ = density(H)
    diis_F, diis_e = [], []
    Eold = None
    hist = []
    E = 0.0
    it = 0
    for it in range(1, maxit + 1):
        J = basis.coulomb(P)
        Ecoul = 0.5 * torch.sum(P * J)
        F = H + J
        Exx = torch.zeros((), dtype=f64, device=dev)
        Exc = torch.zeros((), dtype=f64, device=dev)
        if have_exx:
            Pa = 0.5 * P
            Ka = basis.exchange(Pa, precision="fp32" if fp32 else "fp64")
            Exx = torch.sum(Pa * Ka)
            F = F + Ka
        else:
            Vxc, Exc = xc(P)
            F = F + Vxc
        Ekin = torch.sum(P * T)
        Enuc = torch.sum(P * V)
        E = float(Ekin + Enuc + Ecoul + Exx + Exc) + Enucr
        errs = []
        for k, b in enumerate(blocks):
            Fk, Pk, Sk = F[b][:, b], P[b][:, b], S[b][:, b]
            FPS = Fk @ Pk @ Sk
            errs.append((X[k].T @ (FPS - FPS.T) @ X[k]).reshape(-1))
        err = torch.cat(errs)
        emax = float(torch.max(torch.abs(err)))
        hist.append(E)
        if verbose:
            print(f"it {it:3d} E {E:.12f} err {emax:.3e}{' (fp32 K)' if fp32 else ''}", flush=True)
        if fp32:
            it_fp32 = it
            if emax < switch:
                # FP64 from here on: the FP32 Fock matrices leave the DIIS
                # history and this step is a plain diagonalisation
                fp32 = False
                diis_F, diis_e, Eold = [], [], None
                P, nk, _ = density(F)
                continue
        elif Eold is not None and abs(E - Eold) < conv and emax < 1e-6:
            break
        else:
            Eold = E
        diis_F.append(F)
        diis_e.append(err)
        if len(diis_F) > diis_len:
            diis_F.pop(0)
            diis_e.pop(0)
        c = _diis(diis_F, diis_e, dev)
        Fd = F if c is None else sum(c[i] * Fi for i, Fi in enumerate(diis_F))
        P, nk, _ = density(Fd)
    torch.cuda.synchronize(dev)
    return SCFResult(E=E, Ekin=float(Ekin), Enuc=float(Enuc), Ecoul=float(Ecoul), Exx=float(Exx), Exc=float(Exc), Enucr=Enucr,
                     iterations=it, iterations_fp32=it_fp32, P=P, occupations=nk,
                     seconds=time.perf_counter() - t0, history=hist)


def _diis(diis_F, diis_e, dev):
    """Pulay extrapolation coefficients over the stored commutators."""
    import torch
    n = len(diis_e)
    E_mat = torch.stack(diis_e)
    B = torch.empty((n + 1, n + 1), dtype=torch.float64, device=dev)
    B[:n, :n] = E_mat @ E_mat.T
    B[n, :] = -1.0
    B[:, n] = -1.0
    B[n, n] = 0.0
    rhs = torch.zeros(n + 1, dtype=torch.float64, device=dev)
    rhs[n] = -1.0
    try:
        return torch.linalg.solve(B, rhs)[:n]
    except RuntimeError:
        return None


def uhf(basis, nela: int, nelb: int, symm: int = 1, conv: float = 1e-10, maxit: int = 200, diis_len: int = 10,
        verbose: bool = False, xc_pol=None, device: str = "cuda", precision: str = "fp64", switch: float = 1e-3):
    """Unrestricted SCF (the reference's fock builders with Pa != Pb,
    atomic/main.cpp:369-470, diatomic/main.cpp:370-479).

    xc_pol(Pa, Pb) -> (Va, Vb, Exc) selects spin-polarized Kohn-Sham DFT;
    None is UHF with Exx = 1/2 (tr(Pa Ka) + tr(Pb Kb)).  precision="mixed"
    as for rhf."""
    import torch
    dev = torch.device(device)
    f64 = torch.float64
    t0 = time.perf_counter()
    S = torch.as_tensor(np.ascontiguousarray(basis.overlap()), dtype=f64, device=dev)
    T = torch.as_tensor(np.ascontiguousarray(basis.kinetic()), dtype=f64, device=dev)
    V = torch.as_tensor(np.ascontiguousarray(basis.nuclear()), dtype=f64, device=dev)
    H = T + V
    blocks = [torch.as_tensor(b, dtype=torch.long, device=dev) for b in basis.get_sym_idx(symm)]
    X = [_invh(S[b][:, b]) for b in blocks]
    Enucr = float(getattr(basis, "Enucr", 0.0))

    def density(F, nocc):
        evs, vecs = [], []
        for k, b in enumerate(blocks):
            e, c = torch.linalg.eigh(X[k].T @ F[b][:, b] @ X[k])
            evs.append(e)
            vecs.append(c)
        ev_host = [e.cpu().numpy() for e in evs]
        order = sorted((float(e), k, i) for k, ev in enumerate(ev_host) for i, e in enumerate(ev))
        nk = [0] * len(blocks)
        for _, k, _i in order[:nocc]:
            nk[k] += 1
        P = torch.zeros_like(F)
        for k, b in enumerate(blocks):
            if nk[k]:
                C = X[k] @ vecs[k][:, :nk[k]]
                P[b[:, None], b[None, :]] += C @ C.T
        return P, nk

    if precision not in ("fp64", "mixed"):
        raise ValueError(f"precision must be 'fp64' or 'mixed', not {precision!r}")
    fp32 = precision == "mixed"
    it_fp32 = 0
    (Pa, nka), (Pb, nkb) = density(H, nela), density(H, nelb)
    diis_F, diis_e = [], []
    Eold, E, it = None, 0.0, 0
    hist = []
    for it in range(1, maxit + 1):
        P = Pa + Pb
        J = basis.coulomb(P)
        Ecoul = 0.5 * torch.sum(P * J)
        Exx = torch.zeros((), dtype=f64, device=dev)
        Exc = torch.zeros((), dtype=f64, device=dev)
        if xc_pol is None:
            kp = "fp32" if fp32 else "fp64"
            Ka = basis.exchange(Pa, precision=kp)
            Kb = basis.exchange(Pb, precision=kp) if nelb else torch.zeros_like(Pb)
            Exx = 0.5 * (torch.sum(Pa * Ka) + torch.sum(Pb * Kb))
            Fa, Fb = H + J + Ka, H + J + Kb
        else:
            Va, Vb, Exc = xc_pol(Pa, Pb)
            Fa, Fb = H + J + Va, H + J + Vb
        Ekin = torch.sum(P * T)
        Enuc = torch.sum(P * V)
        E = float(Ekin + Enuc + Ecoul + Exx + Exc) + Enucr
        errs = []
        for F_, P_ in ((Fa, Pa), (Fb, Pb)):
            for k, b in enumerate(blocks):
                FPS = F_[b][:, b] @ P_[b][:, b] @ S[b][:, b]
                errs.append((X[k].T @ (FPS - FPS.T) @ X[k]).reshape(-1))
        err = torch.cat(errs)
        emax = float(torch.max(torch.abs(err)))
        hist.append(E)
        if verbose:
            print(f"it {it:3d} E {E:.12f} err {emax:.3e}{' (fp32 K)' if fp32 else ''}", flush=True)
        if fp32:
            it_fp32 = it
            if emax < switch:
                fp32 = False
                diis_F, diis_e, Eold = [], [], None
                (Pa, nka), (Pb, nkb) = density(Fa, nela), density(Fb, nelb)
                continue
        elif Eold is not None and abs(E - Eold) < conv and emax < 1e-6:
            break
        else:
            Eold = E
        diis_F.append((Fa, Fb))
        diis_e.append(err)
        if len(diis_F) > diis_len:
            diis_F.pop(0)
            diis_e.pop(0)
        c = _diis(diis_F, diis_e, dev)
        if c is None:
            Fad, Fbd = Fa, Fb
        else:
            Fad = sum(c[i] * f[0] for i, f in enumerate(diis_F))
            Fbd = sum(c[i] * f[1] for i, f in enumerate(diis_F))
        (Pa, nka), (Pb, nkb) = density(Fad, nela), density(Fbd, nelb)
    torch.cuda.synchronize(dev)
    return SCFResult(E=E, Ekin=float(Ekin), Enuc=float(Enuc), Ecoul=float(Ecoul), Exx=float(Exx), Exc=float(Exc),
                     Enucr=Enucr, iterations=it, iterations_fp32=it_fp32, Pa=Pa, Pb=Pb, occupations=(nka, nkb),
                     seconds=time.perf_counter() - t0, history=hist)
