"""Device-resident SCF (restricted and unrestricted) around the GPU Fock build.

Counterpart of the Fock builders of src/atomic/main.cpp:369-470 and
src/diatomic/main.cpp:370-479 together with the SCF glue of
src/general/scf_driver_common.h:
  * per-symmetry-block half-inverse overlap (build_per_block_Sinvh, :176-187;
    utils::invh, libhelfem/src/utils.cpp:120-159, chol=false),
  * P = sum_k Sinvh_k C diag(occ) C' Sinvh_k' (accumulate_density_block, :407-425),
  * Etot = tr(PT) + tr(PV) + 1/2 tr(PJ) + Exc + Exx + Enucr, Exx = 1/2 sum_s
    tr(P_s K_s) (assemble_hf_exchange, :460-474),
  * F_k = Sinvh_k' (H1 + J + XC + K) Sinvh_k (assemble_fock_blocks, :490-520).
The reference iterates with OpenOrbitalOptimizer (DIIS/ODA/CG); this driver
uses aufbau over all blocks and Pulay DIIS on the orthonormal commutators.
Converged energies depend only on the state reached, not on the iteration.

All matrices stay in HBM (torch float64); the per-block eigensolves run on
the GPU (torch.linalg.eigh), J and K through the HIP kernels of libhfgpu.
Only the symmetry blocks of the Fock matrix enter the eigensolves, so the
DIIS history holds blocks, not Nbf x Nbf matrices.
"""
from __future__ import annotations

import time

import numpy as np


def _one_electron(basis, dev):
    """S, T, V on the device: assembled there from their nonzero blocks when
    the basis has a GPU context, else built on the host and sent (the host
    matrices are Fortran-ordered, so their transposes are C-contiguous views:
    those are sent and transposed back, avoiding a host-side copy)."""
    import torch
    if dev.type == "cuda" and getattr(basis, "ctx", None) is not None and hasattr(basis, "one_electron_dev"):
        return [basis.one_electron_dev(w).contiguous() for w in (0, 1, 2)]
    out = []
    for M in (basis.overlap(), basis.kinetic(), basis.nuclear()):
        out.append(torch.from_numpy(np.asarray(M).T).to(dev, dtype=torch.float64).T.contiguous())
    return out


def _sync(dev):
    import torch
    if dev.type == "cuda":
        torch.cuda.synchronize(dev)


class _Clock:
    """Per-phase wall time, synchronised with the device at phase ends."""

    def __init__(self, dev):
        self.dev = dev
        self.t = {}
        self.last = time.perf_counter()

    def tick(self, name):
        _sync(self.dev)
        now = time.perf_counter()
        self.t[name] = self.t.get(name, 0.0) + now - self.last
        self.last = now


def _invh(S):
    """D (D S D)^-1/2 with D = diag(S)^-1/2 (invh with chol = false,
    libhelfem/src/utils.cpp:121-159)."""
    import torch
    d = 1.0 / torch.sqrt(torch.diagonal(S))
    Sn = d[:, None] * S * d[None, :]
    e, v = torch.linalg.eigh(Sn)
    return d[:, None] * ((v * (1.0 / torch.sqrt(e))) @ v.T)


class SCFResult(dict):
    __getattr__ = dict.get


class _Blocks:
    """Per-symmetry-block orthogonalisation and eigensolves
    (build_per_block_Sinvh, scf_driver_common.h:176-187).

    Two exact reductions of the eigenproblems, both re-checked every call:
      * a block (one m) splits into independent sub-problems wherever the
        Fock and overlap matrices have no coupling between groups of shells
        (the l shells of an atom with a spherical density);
      * sub-problems holding the same matrix (the same l in every m of such
        an atom, the m and -m blocks of a +-m symmetric density) are solved
        once.
    "No coupling" and "the same" are judged at 1e-12 of the block's largest
    element, so an open shell that breaks a symmetry falls back to the full
    eigensolve of its block.

    A third, also exact: a sub-problem solved before whose lowest eigenvalue,
    lowered by the Frobenius norm of the change of its matrix since (Weyl's
    bound), stays above the previous highest occupied level by GAP is not
    solved: it holds no occupied orbital.  After the aufbau over the solved
    ones the bound is checked against the new highest occupied level, and any
    sub-problem that fails it is solved after all."""

    TOL = 1e-12
    GAP = 0.05
    DEGEN = 1e-9

    def __init__(self, basis, S, symm, dev):
        import torch
        self.torch = torch
        idx = basis.get_sym_idx(symm)
        self.b = [torch.as_tensor(b, dtype=torch.long, device=dev) for b in idx]
        self.S = [S[b][:, b] for b in self.b]
        self.X = [None] * len(self.S)
        self.dev = dev
        self.ctx = getattr(basis, "ctx", None)
        # shells of each block, in block order: (l, local start, size)
        lval = np.asarray(basis.lval).tolist()
        mval = np.asarray(basis.mval).tolist()
        self.shells = []
        if symm == 1:
            for m in sorted(set(mval)):
                sh, off = [], 0
                for a in range(len(lval)):
                    if mval[a] != m:
                        continue
                    n = basis.Nrad - (1 if (getattr(basis, "Ndummy", None) is not None and m != 0) else 0)
                    sh.append((lval[a], off, n))
                    off += n
                self.shells.append(sh)
        elif symm == 3:  # (m, parity of l) blocks of an atomic basis
            for m in sorted(set(mval)):
                for par in (0, 1):
                    sh, off = [], 0
                    for a in range(len(lval)):
                        if mval[a] == m and lval[a] % 2 == par:
                            sh.append((lval[a], off, basis.Nrad))
                            off += basis.Nrad
                    if sh:
                        self.shells.append(sh)
        else:
            self.shells = [[(-1, 0, len(b))] for b in idx]
        for k, sh in enumerate(self.shells):
            if sum(x[2] for x in sh) != len(idx[k]):
                self.shells[k] = [(-1, 0, len(idx[k]))]
        # overlap couplings between shells (fixed)
        self.Scpl = [self._coupling(Sk, k) for k, Sk in enumerate(self.S)]
        self.X = self._invh_blocks()
        self.solves = 0
        self.paired_solves = 0
        self.skipped_solves = 0
        self.memo = {}  # (slot, block, shells) -> (matrix last solved, its lowest eigenvalue)
        self.homo = {}  # slot -> highest occupied level of the last aufbau

    def _invh_blocks(self):
        """invh of every overlap block, taken over the groups of shells that S
        couples (S is block diagonal over them, and so is its half inverse):
        one eigensolve per distinct group, all in padded batched calls."""
        torch = self.torch
        groups = []  # (block, local indices, index of the distinct group)
        uniq = []    # distinct overlap sub-blocks
        for k, Sk in enumerate(self.S):
            if self.Scpl[k] is None:
                comps = [list(range(len(self.shells[k])))]
            else:
                host = self.Scpl[k].cpu().numpy()
                ns = len(self.shells[k])
                parent = list(range(ns))

                def find(x):
                    while parent[x] != x:
                        parent[x] = parent[parent[x]]
                        x = parent[x]
                    return x

                for s_ in range(ns):
                    for t in range(s_ + 1, ns):
                        if host[s_, t] != 0.0:
                            parent[find(s_)] = find(t)
                cc = {}
                for s_ in range(ns):
                    cc.setdefault(find(s_), []).append(s_)
                comps = sorted(cc.values())
            for comp in comps:
                li = self._local(k, comp)
                Ssub = Sk[li][:, li]
                u = next((q for q, Sp in enumerate(uniq) if Sp.shape == Ssub.shape and torch.equal(Sp, Ssub)), None)
                if u is None:
                    u = len(uniq)
                    uniq.append(Ssub)
                groups.append((k, li, u))
        # D (D S D)^-1/2 per distinct group (invh, utils.cpp:121-159)
        ds = [1.0 / torch.sqrt(torch.diagonal(Sp)) for Sp in uniq]
        sols = batched_eigh([d[:, None] * Sp * d[None, :] for d, Sp in zip(ds, uniq)], self.ctx)
        Xu = []
        for d, (e, v) in zip(ds, sols):
            ev = torch.as_tensor(e, dtype=v.dtype, device=v.device)
            Xu.append(d[:, None] * ((v * (1.0 / torch.sqrt(ev))) @ v.T))
        X = [torch.zeros_like(Sk) for Sk in self.S]
        for k, li, u in groups:
            X[k][li[:, None], li[None, :]] = Xu[u]
        return X

    def _coupling(self, M, k):
        """(ns x ns) host matrix of max |M_st| over shell pairs (None when the
        block is one shell group)."""
        sh = self.shells[k]
        if len(sh) == 1:
            return None
        n = sh[0][2]
        if any(x[2] != n for x in sh):
            return None
        ns = len(sh)
        return M.reshape(ns, n, ns, n).abs().amax(dim=(1, 3))

    def blocks_of(self, F):
        return [F[b][:, b] for b in self.b]

    def _subproblems(self, Fb):
        """Independent (block, shells) groups and, for each, the index of the
        group whose solution it reuses (itself if solved)."""
        torch = self.torch
        cpl = [self._coupling(F, k) for k, F in enumerate(Fb)]
        scale = torch.stack([torch.max(torch.abs(F)) for F in Fb])
        host = [c.cpu().numpy() if c is not None else None for c in cpl]
        scale = scale.cpu().numpy()
        shost = [c.cpu().numpy() if c is not None else None for c in self.Scpl]
        subs = []
        for k, sh in enumerate(self.shells):
            ns = len(sh)
            if host[k] is None:
                subs.append((k, list(range(ns))))
                continue
            thr = self.TOL * scale[k]
            sthr = self.TOL * float(np.max(shost[k]))
            parent = list(range(ns))

            def find(x):
                while parent[x] != x:
                    parent[x] = parent[parent[x]]
                    x = parent[x]
                return x

            for s_ in range(ns):
                for t in range(s_ + 1, ns):
                    if host[k][s_, t] > thr or host[k][t, s_] > thr or shost[k][s_, t] > sthr:
                        parent[find(s_)] = find(t)
            comps = {}
            for s_ in range(ns):
                comps.setdefault(find(s_), []).append(s_)
            for c in sorted(comps.values()):
                subs.append((k, c))
        # candidate duplicates: same shell l's and sizes
        keys = {}
        for i, (k, c) in enumerate(subs):
            key = tuple((self.shells[k][s_][0], self.shells[k][s_][2]) for s_ in c)
            keys.setdefault(key, []).append(i)
        loc = [self._local(k, c) for k, c in subs]
        reuse = list(range(len(subs)))
        cand = [(i, j) for members in keys.values() for a, i in enumerate(members) for j in members[:a]]
        if cand:
            diffs = []
            for i, j in cand:
                (ki, _), (kj, _) = subs[i], subs[j]
                li, lj = loc[i], loc[j]
                dF = torch.max(torch.abs(Fb[ki][li][:, li] - Fb[kj][lj][:, lj]))
                dX = torch.max(torch.abs(self.X[ki][li][:, li] - self.X[kj][lj][:, lj]))
                diffs.append(torch.stack([dF, dX]))
            d = torch.stack(diffs).cpu().numpy()
            xscale = [float(torch.max(torch.abs(X))) for X in self.X]
            same = {(i, j): (dF <= self.TOL * scale[subs[i][0]] and dX <= self.TOL * xscale[subs[i][0]])
                    for (i, j), (dF, dX) in zip(cand, d)}
            for members in keys.values():
                for a, i in enumerate(members):
                    for j in members[:a]:
                        if reuse[j] == j and same[(i, j)]:
                            reuse[i] = j
                            break
        return subs, loc, reuse

    def _local(self, k, comp):
        sh = self.shells[k]
        return self.torch.cat([self.torch.arange(sh[s_][1], sh[s_][1] + sh[s_][2], device=self.dev)
                               for s_ in comp])

    def prepare(self, Fb, slot=0, frozen=None):
        """Independent eigenproblems X' F X of this Fock matrix (one per
        distinct shell group and not excluded by the Weyl bound), with the
        bookkeeping `finish` needs.  `slot` names the spin channel; with
        `frozen` (orbitals per block) blocks holding none are not solved."""
        torch = self.torch
        subs, loc, reuse = self._subproblems(Fb)
        uniq = [i for i in range(len(subs)) if reuse[i] == i]
        X, mats = {}, {}
        for i in uniq:
            k, li = subs[i][0], loc[i]
            X[i] = self.X[k][li][:, li]
            mats[i] = X[i].T @ Fb[k][li][:, li] @ X[i]
        keys = {i: (slot, subs[i][0], tuple(subs[i][1])) for i in uniq}
        skipped = {}
        if frozen is not None:
            # solutions another sub-problem reuses stay needed
            needed = {reuse[j] for j in range(len(subs)) if frozen[subs[j][0]] > 0}
            for i in uniq:
                if frozen[subs[i][0]] == 0 and i not in needed:
                    skipped[i] = np.inf
        homo = self.homo.get(slot)
        cand = [i for i in uniq if homo is not None and i not in skipped and keys[i] in self.memo
                and self.memo[keys[i]][0].shape == mats[i].shape]
        if cand:
            dn = torch.stack([torch.linalg.matrix_norm(mats[i] - self.memo[keys[i]][0]) for i in cand]).cpu().numpy()
            for i, d in zip(cand, dn):
                lb = self.memo[keys[i]][1] - float(d)
                if lb > homo + self.GAP:
                    skipped[i] = lb
        solve = [i for i in uniq if i not in skipped]
        self.solves += len(solve)
        self.skipped_solves += len(skipped)
        self.paired_solves += len(subs) - len(uniq)
        return {"subs": subs, "loc": loc, "reuse": reuse, "uniq": solve, "X": X, "mats": [mats[i] for i in solve],
                "all": mats, "keys": keys, "skipped": skipped, "slot": slot}

    def finish(self, plan, sols, nocc, weight, like, frozen=None):
        """Aufbau over all sub-problems from their eigenpairs (`sols`, in the
        order of plan["mats"]); returns the full P (for J/K), its symmetry
        blocks and the occupied orbitals per block.  `frozen` (orbitals per
        block) replaces the aufbau's choice of blocks: each block then takes
        its own lowest levels; a fractional count occupies the block's last
        orbital by its fraction (0.5 of a pi electron in each of m = +-1 is
        the cylindrically averaged state)."""
        subs, loc, reuse = plan["subs"], plan["loc"], plan["reuse"]
        ev, vecs = {}, {}

        def take(idx, solutions):
            for i, (e, v) in zip(idx, solutions):
                ev[i] = e
                vecs[i] = plan["X"][i] @ v
                self.memo[plan["keys"][i]] = (plan["all"][i], float(e[0]))

        def aufbau():
            # levels within DEGEN of each other (a chain of such gaps) count as
            # degenerate and are filled in sub-problem order, so the occupied
            # set does not hinge on rounding (open shells: which m fills first)
            lv = sorted((float(e), i, j) for i in range(len(subs)) if reuse[i] in ev
                        for j, e in enumerate(ev[reuse[i]]))
            out, start = [], 0
            for q in range(1, len(lv) + 1):
                if q == len(lv) or lv[q][0] - lv[q - 1][0] > self.DEGEN * max(1.0, abs(lv[q][0])):
                    out += sorted(lv[start:q], key=lambda x: (x[1], x[2]))
                    start = q
            return out

        take(plan["uniq"], sols)
        order = aufbau()
        homo = order[nocc - 1][0] if nocc > 0 else -np.inf
        late = [i for i, lb in plan["skipped"].items()
                if (frozen[subs[i][0]] > 0 if frozen is not None else not lb > homo)]
        if late:  # the occupied levels rose past a bound: solve those after all
            self.solves += len(late)
            self.skipped_solves -= len(late)
            take(late, batched_eigh([plan["all"][i] for i in late], self.ctx))
            order = aufbau()
            homo = order[nocc - 1][0] if nocc > 0 else -np.inf
        nsub = [0] * len(subs)
        frac = [None] * len(subs)  # occupation of each taken orbital when fixed counts are fractional
        if frozen is None:
            self.homo[plan["slot"]] = homo
            for _, i, _j in order[:nocc]:
                nsub[i] += 1
        else:
            self.homo[plan["slot"]] = np.inf  # no Weyl skipping under fixed occupations
            left = [float(x) for x in frozen]
            for _, i, _j in order:
                k = subs[i][0]
                if left[k] > 1e-12:
                    f = min(1.0, left[k])
                    left[k] -= f
                    nsub[i] += 1
                    if f < 1.0 or frac[i] is not None:
                        frac[i] = (frac[i] or [1.0] * (nsub[i] - 1)) + [f]
        P = like.new_zeros(like.shape)
        Pb = [like.new_zeros((len(b), len(b))) for b in self.b]
        nk = [0] * len(self.b)
        for i, (k, c) in enumerate(subs):
            if not nsub[i]:
                continue
            Cs = vecs[reuse[i]][:, :nsub[i]]
            li = loc[i]
            if frac[i] is None:
                Pb[k][li[:, None], li[None, :]] = weight * (Cs @ Cs.T)
            else:
                w = Cs.new_tensor(frac[i])
                Pb[k][li[:, None], li[None, :]] = weight * ((Cs * w) @ Cs.T)
            nk[k] += nsub[i]
        for k, b in enumerate(self.b):
            if nk[k]:
                P[b[:, None], b[None, :]] = Pb[k]
        return P, Pb, nk

    def density(self, Fb, nocc, weight, like, slot=0):
        plan = self.prepare(Fb, slot)
        return self.finish(plan, batched_eigh(plan["mats"], self.ctx), nocc, weight, like)

    def commutator(self, Fb, Pb):
        """Orthonormal-basis commutators X_k' (F P S - S P F) X_k per block."""
        out = []
        for k in range(len(self.b)):
            FPS = Fb[k] @ Pb[k] @ self.S[k]
            out.append(self.X[k].T @ (FPS - FPS.T) @ self.X[k])
        return out


PAD_RATIO = 0.75
# Smallest stack size handed to the batched eigensolver.  The rocSOLVER build
# that ships with this PyTorch (and that the library binds to once torch is
# loaded, the two sharing a soname) solves batches of n <= 64 matrices an
# order of magnitude slower than n = 72 (178 blocks: 50 ms at n = 53, 2 ms at
# n = 71), so smaller blocks are padded up.
EIGH_MIN_N = 72


def batched_eigh(mats, ctx=None):
    """Eigenpairs of symmetric matrices, ascending.  Matrices whose sizes lie
    within PAD_RATIO of the largest of a group share one batched eigensolve:
    each is padded to the group size with a diagonal c = 2 |A|_F + 1 above its
    spectrum and no coupling, so its eigenpairs are the first n of the padded
    matrix (the Householder reduction keeps the zero coupling exact).  The
    eigensolves are launch-bound at these sizes, so fewer, larger batches
    win.  With a Context the stacks go to hf_eigh_batched (rocSOLVER's
    batched dsyevd), else to torch.linalg.eigh.  All eigenvalues (and the
    solver statuses) come to the host in one transfer."""
    import torch
    from . import linalg
    if not mats:
        return []
    order = sorted(range(len(mats)), key=lambda j: -mats[j].shape[0])
    groups = []
    for j in order:
        if groups and mats[j].shape[0] >= PAD_RATIO * mats[groups[-1][0]].shape[0]:
            groups[-1].append(j)
        else:
            groups.append([j])
    vals, vecs = [None] * len(mats), [None] * len(mats)
    evs, infos = [], []
    for members in groups:
        # stacks below EIGH_MIN_N are padded up to it: the batched eigensolver
        # is an order of magnitude slower below that size
        N = max(mats[members[0]].shape[0], EIGH_MIN_N)
        A0 = mats[members[0]]
        if all(mats[j].shape[0] == N for j in members):
            stack = torch.stack([mats[j] for j in members])
        else:
            stack = A0.new_zeros((len(members), N, N))
            for q, j in enumerate(members):
                A = mats[j]
                n = A.shape[0]
                stack[q, :n, :n] = A
                if n < N:
                    idx = torch.arange(n, N, device=A.device)
                    stack[q, idx, idx] = 2.0 * torch.linalg.matrix_norm(A) + 1.0
        if ctx is not None and stack.is_cuda:
            e, v, info = linalg.eigh(stack, ctx)
            infos.append(info)
        else:
            e, v = torch.linalg.eigh(stack)
        for q, j in enumerate(members):
            n = mats[j].shape[0]
            vecs[j] = v[q, :n, :n]
        evs.append((members, e))
    parts = [e.reshape(-1) for _, e in evs]
    nval = sum(p.numel() for p in parts)
    host = torch.cat(parts + [i.to(torch.float64) for i in infos]).cpu().numpy()
    if np.any(host[nval:] != 0):
        raise RuntimeError("batched eigensolver did not converge")
    off = 0
    for members, e in evs:
        N = e.shape[1]
        for q, j in enumerate(members):
            n = mats[j].shape[0]
            vals[j] = host[off + q * N:off + q * N + n]
        off += len(members) * N
    return list(zip(vals, vecs))


class _DIIS:
    """Pulay DIIS over lists of block tensors; the B matrix is kept on the
    host and extended by one row of dot products per iteration."""

    def __init__(self, size):
        self.size = size
        self.F, self.e = [], []
        self.B = np.zeros((0, 0))

    def reset(self):
        self.F, self.e = [], []
        self.B = np.zeros((0, 0))

    def push(self, F, e):
        import torch
        if len(self.F) == self.size:
            self.F.pop(0)
            self.e.pop(0)
            self.B = self.B[1:, 1:]
        dots = [sum(torch.sum(a * b) for a, b in zip(e, ei)) for ei in self.e] + [sum(torch.sum(a * a) for a in e)]
        row = torch.stack([torch.as_tensor(d) for d in dots]).cpu().numpy()
        n = len(self.F) + 1
        B = np.zeros((n, n))
        B[:n - 1, :n - 1] = self.B
        B[n - 1, :] = row
        B[:, n - 1] = row
        self.B = B
        self.F.append(F)
        self.e.append(e)

    def extrapolate(self):
        n = len(self.F)
        A = -np.ones((n + 1, n + 1))
        A[:n, :n] = self.B
        A[n, n] = 0.0
        rhs = np.zeros(n + 1)
        rhs[n] = -1.0
        try:
            c = np.linalg.solve(A, rhs)[:n]
        except np.linalg.LinAlgError:
            return self.F[-1]
        return [sum(float(c[i]) * self.F[i][k] for i in range(n)) for k in range(len(self.F[-1]))]


class SCFState:
    """One SCF in flight: the setup, the Fock step of every iteration and the
    convergence test; the eigensolves it asks for are done by the caller
    (`SCFEngine`), so many states can share batched eigensolves.

    Restricted: nel_spin = (nocc,), xc(P) -> (Vxc, Exc).  Unrestricted:
    nel_spin = (nela, nelb), xc(Pa, Pb) -> (Va, Vb, Exc)."""

    def __init__(self, basis, nel_spin, restricted, symm=1, conv=1e-10, maxit=100, diis_len=10, verbose=False,
                 xc=None, device="cuda", precision="fp64", switch=1e-3, tag=None, occupations=None, gconv=1e-6,
                 kfrac=None, kshort=0.0):
        import torch
        if precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', not {precision!r}")
        self.torch = torch
        self.basis, self.nel_spin, self.restricted = basis, tuple(nel_spin), restricted
        self.conv, self.maxit, self.verbose, self.xc, self.switch = conv, maxit, verbose, xc, switch
        # fraction of exact exchange in the Fock matrix: 1 for Hartree-Fock, 0
        # for a pure functional, the hybrid's (exact_exchange) otherwise
        self.kfrac = (1.0 if xc is None else 0.0) if kfrac is None else float(kfrac)
        # and of its short-range part (range-separated hybrids; the basis holds
        # the erfc or Yukawa caches of the functional's omega)
        self.kshort = float(kshort)
        self.gconv = gconv  # largest orbital-gradient (commutator) entry at convergence
        self.tag = tag
        self.dev = torch.device(device)
        self.t0 = time.perf_counter()
        self.clk = _Clock(self.dev)
        self.S, self.T, self.V = _one_electron(basis, self.dev)
        self.H = self.T + self.V
        self.bs = _Blocks(basis, self.S, symm, self.dev)
        self.Enucr = float(getattr(basis, "Enucr", 0.0))  # nuclear repulsion of a diatomic basis
        self.clk.tick("setup")
        self.nspin = 1 if restricted else 2
        self.weight = 2.0 if restricted else 1.0
        # reduced precision applies to the exchange; a pure DFT build (xc set,
        # no exchange) has none, so its mixed mode is the FP64 one (no switch,
        # no DIIS restart)
        self.mixed = precision == "mixed" and xc is None
        self.fp32 = self.mixed
        self.it_fp32 = 0
        # incremental exchange of the mixed mode after the FP32 phase:
        # K(P) = K64(P_ref) + K32(P - P_ref) against the last FP64 build, while
        # max|P - P_ref| <= INC_MAX max|P|; convergence is confirmed by a full
        # FP64 build
        self.kref = [None] * (1 if restricted else 2)
        self.final_check = False
        self.it_inc = 0
        self.inc_now = False
        self.diis = _DIIS(diis_len)
        self.Eold, self.E, self.it, self.hist = None, 0.0, 0, []
        z = torch.zeros((), dtype=torch.float64, device=self.dev)
        self.terms = dict(Ekin=z, Enuc=z, Ecoul=z, Exx=z, Exc=z)
        self.dens = [None] * self.nspin
        # occupations per block of every delivered density; frozen once the
        # aufbau flip-flops between patterns (near-degenerate frontier levels)
        self.occ_hist = [[] for _ in range(self.nspin)]
        self.frozen = [None] * self.nspin
        if occupations is not None:
            # fixed orbitals per block and spin (the reference's --readocc)
            self.frozen = [list(occupations[sp]) for sp in range(self.nspin)]
        Hb = self.bs.blocks_of(self.H)
        self.pending = [(s, Hb) for s in range(self.nspin)]
        self.done = False
        self.stream = None  # the stream of this state's Fock builds in an overlapping engine
        self._built = None

    def requests(self):
        """(spin, Fock blocks, occupied orbitals) still to be diagonalised."""
        return [(s, Fb, self.nel_spin[s]) for s, Fb in self.pending]

    def deliver(self, s, dens):
        self.dens[s] = dens
        self.occ_hist[s].append(tuple(dens[2]))

    INC_MAX = 1e-2  # largest max|P - P_ref| / max|P| of an incremental exchange build
    FLIPS = 3       # occupation changes within the last WINDOW iterations that freeze them
    WINDOW = 10

    def _check_flips(self):
        """Aufbau oscillating between occupation patterns (a frontier level
        crossing back and forth) never converges: after FLIPS changes within
        WINDOW iterations keep the pattern of the lowest energy seen so far."""
        if self.frozen[0] is not None or self.it < self.WINDOW:
            return
        n = min(len(self.hist), *(len(h) for h in self.occ_hist))
        pats = [tuple(self.occ_hist[s][t] for s in range(self.nspin)) for t in range(n)]
        recent = pats[-self.WINDOW:]
        if sum(1 for a, b in zip(recent, recent[1:]) if a != b) < self.FLIPS:
            return
        best = min(range(n), key=lambda t: self.hist[t])
        self.frozen = [list(pats[best][s]) for s in range(self.nspin)]
        self.diis.reset()
        self.Eold = None

    def advance(self):
        """One Fock build, energy and convergence test on the delivered
        densities; leaves the next eigenproblems in `pending` or sets done."""
        if self.launch():
            self.complete()

    def launch(self, stream=None):
        """Queue this iteration's Fock build (J, and K or Vxc) -- on `stream`
        when given, so the builds of several states overlap on the device
        (the exchange's screen synchronises only that stream).  False when
        the state is done."""
        import contextlib
        torch = self.torch
        self.pending = []
        if self.it >= self.maxit:
            self.done = True
            return False
        self.it += 1
        clk, nspin = self.clk, self.nspin
        tick = clk.tick if stream is None else (lambda name: None)
        Ps = [d[0] for d in self.dens]
        scope = torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext()
        made = []
        with scope:
            # the total density is formed on the build's own stream: `stream`
            # waited for the densities (made on the caller's stream) before
            # this call, not for work queued on the caller's stream after it
            P = Ps[0] if self.restricted else Ps[0] + Ps[1]
            if not self.restricted:
                made.append(P)  # made on this stream, used by complete() on the caller's
            if stream is not None:
                for X in Ps + [P]:
                    X.record_stream(stream)
            J = self.basis.coulomb(P)
            made.append(J)
            tick("coulomb")
            Ks, Vs, Exc = [None] * nspin, [None] * nspin, None
            if self.kfrac or self.kshort:
                spins = [s for s in range(nspin) if self.nel_spin[s] > 0]
                Psps = {s: (0.5 * Ps[s] if self.restricted else Ps[s]) for s in spins}
                # incremental only when every spin's change from its reference is small
                inc = (self.mixed and not self.fp32 and not self.final_check
                       and all(self.kref[s] is not None for s in spins))
                if inc:
                    rel = torch.stack([torch.max(torch.abs(Psps[s] - self.kref[s][0])) / torch.max(torch.abs(Psps[s]))
                                       for s in spins])
                    inc = bool(torch.max(rel) <= self.INC_MAX)
                self.inc_now = inc
                for s in spins:
                    Psp = Psps[s]
                    if not self.kfrac:
                        K = None
                    elif self.fp32:
                        K = self.basis.exchange(Psp, precision="fp32")
                    elif inc:
                        K = self.kref[s][1] + self.basis.exchange(Psp - self.kref[s][0], precision="fp32")
                    else:
                        K = self.basis.exchange(Psp, precision="fp64")
                        if self.mixed:
                            self.kref[s] = (Psp, K)
                    if K is not None and self.kfrac != 1.0:
                        K = self.kfrac * K
                    if self.kshort:
                        Ksr = self.kshort * self.basis.rs_exchange(Psp)
                        K = Ksr if K is None else K + Ksr
                    Ks[s] = (Psp, K)
                    made += list(Ks[s])
                if inc:
                    self.it_inc += 1
                elif not self.fp32:
                    self.final_check = False  # a full FP64 build: the next test is final
                tick("exchange")
            if self.xc is not None:
                if self.restricted:
                    Vs[0], Exc = self.xc(P)
                else:
                    Vs[0], Vs[1], Exc = self.xc(Ps[0], Ps[1])
                made += [V for V in Vs if V is not None]
                tick("xc")
        self._built = (Ps, P, J, Ks, Vs, Exc, stream, made)
        return True

    def complete(self):
        """Energy, orbital gradient, DIIS and convergence test of the Fock
        build queued by launch()."""
        torch = self.torch
        Ps, P, J, Ks, Vs, Exc, stream, made = self._built
        self._built = None
        if stream is not None:
            cur = torch.cuda.current_stream(self.dev)
            cur.wait_stream(stream)
            for X in made:
                X.record_stream(cur)
        clk, bs, nspin = self.clk, self.bs, self.nspin
        f64 = torch.float64
        Ecoul = 0.5 * torch.sum(P * J)
        Fs = [self.H + J for _ in range(nspin)]
        Exx = torch.zeros((), dtype=f64, device=self.dev)
        if Exc is None:
            Exc = torch.zeros((), dtype=f64, device=self.dev)
        for s in range(nspin):
            if Ks[s] is not None:
                Psp, K = Ks[s]
                Exx = Exx + (2.0 if self.restricted else 1.0) * 0.5 * torch.sum(Psp * K)
                Fs[s] = Fs[s] + K
            if Vs[s] is not None:
                Fs[s] = Fs[s] + Vs[s]
        Ekin = torch.sum(P * self.T)
        Enuc = torch.sum(P * self.V)
        self.terms = dict(Ekin=Ekin, Enuc=Enuc, Ecoul=Ecoul, Exx=Exx, Exc=Exc)
        E = float(Ekin + Enuc + Ecoul + Exx + Exc) + self.Enucr
        self.E = E
        Fb = [bs.blocks_of(F) for F in Fs]
        err = []
        for s in range(nspin):
            err += bs.commutator(Fb[s], self.dens[s][1])
        emax = float(torch.max(torch.stack([torch.max(torch.abs(x)) for x in err])))
        self.hist.append(E)
        self._check_flips()
        clk.tick("energy+diis")
        if self.verbose:
            print(f"it {self.it:3d} E {E:.12f} err {emax:.3e}{' (fp32 K)' if self.fp32 else ''}", flush=True)
        if self.fp32:
            self.it_fp32 = self.it
            if emax < self.switch:
                # FP64 from here on: the FP32 Fock matrices leave the DIIS
                # history and this step is a plain diagonalisation
                self.fp32 = False
                self.diis.reset()
                self.Eold = None
                self.pending = [(s, Fb[s]) for s in range(nspin)]
                return
        elif self.Eold is not None and abs(E - self.Eold) < self.conv and emax < self.gconv and not self.inc_now:
            self.done = True
            return
        else:
            if self.inc_now and self.Eold is not None and abs(E - self.Eold) < self.conv and emax < self.gconv:
                self.final_check = True  # converged on an incremental build: confirm with a full FP64 one
            self.Eold = E
        nb = len(bs.b)
        self.diis.push([x for F in Fb for x in F], err)
        Fd = self.diis.extrapolate()
        clk.tick("energy+diis")
        self.pending = [(s, Fd[s * nb:(s + 1) * nb]) for s in range(nspin)]

    def result(self):
        _sync(self.dev)
        t = self.terms
        out = SCFResult(E=self.E, Ekin=float(t["Ekin"]), Enuc=float(t["Enuc"]), Ecoul=float(t["Ecoul"]),
                        Exx=float(t["Exx"]), Exc=float(t["Exc"]), Enucr=self.Enucr, iterations=self.it,
                        iterations_fp32=self.it_fp32, iterations_incremental=self.it_inc,
                        seconds=time.perf_counter() - self.t0, history=self.hist,
                        phases=self.clk.t, eigensolves=self.bs.solves, reused_solves=self.bs.paired_solves,
                        skipped_solves=self.bs.skipped_solves, frozen_occupations=self.frozen[0] is not None,
                        tag=self.tag)
        if self.restricted:
            out.update(P=self.dens[0][0], occupations=self.dens[0][2])
        else:
            out.update(Pa=self.dens[0][0], Pb=self.dens[1][0], occupations=(self.dens[0][2], self.dens[1][2]))
        return out


class SCFEngine:
    """Advances many SCFStates in lockstep: every step diagonalises all the
    pending Fock blocks of all states together (equal-size eigenproblems in
    one batched call), then runs each state's Fock build.  States can join
    between steps and leave when they converge (continuous batching).  With
    `overlap` (default) the Fock builds of several states are queued on one
    stream each before any is completed, so they run concurrently."""

    def __init__(self, overlap=True):
        self.active = []
        self.steps = 0
        self.eig_seconds = 0.0
        self.overlap = overlap

    def add(self, state):
        self.active.append(state)

    def step(self):
        """One iteration of every active state; returns the states that
        finished in it."""
        import torch
        devs = {st.dev for st in self.active}
        t = time.perf_counter()
        reqs = []
        for st in self.active:
            for s, Fb, nocc in st.requests():
                reqs.append((st, s, st.bs.prepare(Fb, s, st.frozen[s]), nocc))
        for dev in devs:
            _sync(dev)
        t1 = time.perf_counter()
        sols = batched_eigh([m for r in reqs for m in r[2]["mats"]], self.active[0].bs.ctx if self.active else None)
        for dev in devs:
            _sync(dev)
        t2 = time.perf_counter()
        off = 0
        for st, s, plan, nocc in reqs:
            n = len(plan["mats"])
            st.deliver(s, st.bs.finish(plan, sols[off:off + n], nocc, st.weight, st.H, st.frozen[s]))
            off += n
        for dev in devs:
            _sync(dev)
        t3 = time.perf_counter()
        self.eig_seconds += t2 - t1
        for st in self.active:
            ph = st.clk.t
            ph["eigensolve"] = ph.get("eigensolve", 0.0) + (t2 - t1)
            ph["density"] = ph.get("density", 0.0) + (t1 - t) + (t3 - t2)
            st.clk.last = time.perf_counter()
        if self.overlap and len(self.active) > 1 and all(st.dev.type == "cuda" for st in self.active):
            # every state's Fock build queued on a stream of its own before
            # any is completed, so the builds of the states overlap on the
            # device; the densities were made on the current stream
            cur = torch.cuda.current_stream()
            launched = []
            for st in self.active:
                if st.stream is None:
                    st.stream = torch.cuda.Stream(st.dev)
                st.stream.wait_stream(cur)
                if st.launch(st.stream):
                    launched.append(st)
            for st in launched:
                st.complete()
        else:
            for st in self.active:
                st.advance()
        self.steps += 1
        done = [st for st in self.active if st.done]
        self.active = [st for st in self.active if not st.done]
        return done


def _scf(basis, nel_spin, restricted, symm, conv, maxit, diis_len, verbose, xc, device, precision, switch,
         occupations=None, kfrac=None, kshort=0.0):
    st = SCFState(basis, nel_spin, restricted, symm, conv, maxit, diis_len, verbose, xc, device, precision, switch,
                  occupations=occupations, kfrac=kfrac, kshort=kshort)
    eng = SCFEngine()
    eng.add(st)
    while eng.active:
        eng.step()
    return st.result()


def rhf(basis, nelec: int, symm: int = 1, conv: float = 1e-10, maxit: int = 100, diis_len: int = 10,
        verbose: bool = False, xc=None, device: str = "cuda", precision: str = "fp64", switch: float = 1e-3,
        occupations=None, kfrac=None, kshort=0.0):
    """Closed-shell SCF; returns energies, iterations and the final density.

    xc(P) -> (Vxc, Exc) adds a density functional (None: Hartree-Fock); a
    hybrid also passes its exact-exchange fraction kfrac
    (dftgrid.exact_exchange), so the Fock matrix gains kfrac * K and the
    energy kfrac * Exx (src/atomic/main.cpp:424-437); a range-separated one
    also kshort * rs_exchange(P) on the basis' erfc or Yukawa caches.
    precision="mixed" builds the exchange in FP32 while the orbital gradient
    (the largest DIIS commutator element) is above `switch`, then restarts
    DIIS on FP64 and split-precision builds: a full FP64 exchange K64(P_ref)
    is kept and later iterations add the FP32 exchange of the density's
    change, K64(P_ref) + K32(P - P_ref) (a fresh FP64 build whenever
    max|P - P_ref| exceeds 1e-2 max|P|); convergence reached on such a build
    is confirmed by a full FP64 one, so the converged energy carries no FP32
    error.  Overlap, orthogonalisation, eigensolves and
    energy sums are FP64 throughout.  occupations: fixed doubly occupied
    orbitals per symmetry block ([counts]) instead of the aufbau."""
    return _scf(basis, (nelec // 2,), True, symm, conv, maxit, diis_len, verbose, xc, device, precision, switch,
                occupations, kfrac, kshort)


def uhf(basis, nela: int, nelb: int, symm: int = 1, conv: float = 1e-10, maxit: int = 200, diis_len: int = 10,
        verbose: bool = False, xc_pol=None, device: str = "cuda", precision: str = "fp64", switch: float = 1e-3,
        occupations=None, kfrac=None, kshort=0.0):
    """Unrestricted SCF (the reference's fock builders with Pa != Pb,
    atomic/main.cpp:369-470, diatomic/main.cpp:370-479).

    xc_pol(Pa, Pb) -> (Va, Vb, Exc) selects spin-polarized Kohn-Sham DFT;
    None is UHF with Exx = 1/2 (tr(Pa Ka) + tr(Pb Kb)).  precision="mixed",
    occupations ([alpha counts, beta counts]) and a hybrid's kfrac, kshort as for rhf."""
    return _scf(basis, (nela, nelb), False, symm, conv, maxit, diis_len, verbose, xc_pol, device, precision, switch,
                occupations, kfrac, kshort)
