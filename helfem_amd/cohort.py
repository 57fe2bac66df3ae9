"""Many small atomic SCF solves advanced as one vectorised cohort.

The per-system engine (scf.SCFState) issues a few hundred small device
operations and several host round trips per system and iteration, which is
what a stream of small atoms is bound by.  A cohort keeps every matrix of
every system in a handful of flat device buffers and every symmetry block in
padded stacks, so one iteration of all systems is a fixed number of
operations whatever the system count:

  * J of every system and -K of every spin density in one batched build each
    (hf_atomic_batch: one launch per kernel, HIP-graph replay, autotuned
    workgroup shapes);
  * Fock blocks F_b = H_b + J_b + K_b gathered into stacks, orthonormalised
    (X_b' F_b X_b with X_b the reference's invh of the overlap block,
    libhelfem/src/utils.cpp:120-159), diagonalised in batched eigensolves;
  * aufbau on the host from one transfer of the lowest levels;
  * densities, energies (Etot = tr(PT) + tr(PV) + 1/2 tr(PJ) + Exx + Enucr,
    src/atomic/main.cpp:369-470), orbital-gradient commutators and Pulay
    DIIS as batched products over the stacks.

The iteration is the one of scf.SCFState (aufbau over the symmetry blocks,
DIIS on the orthonormal commutators, the FP32 exchange of the mixed mode
until the orbital gradient drops below `switch`, then FP64 with a fresh DIIS,
Weyl-bound skipping of blocks that cannot hold occupied orbitals, frozen
occupations when the aufbau oscillates), so a system's converged energy is
the same alone or in a cohort.  Restricted systems (nela == nelb) carry one
spin channel of weight 2, the others an alpha channel and, with beta
electrons, a beta channel on a clone of the basis (so both densities go into
one exchange batch).  Systems leave the stacks and the batched builds as
soon as they converge.
"""
from __future__ import annotations

import time

import numpy as np

from . import linalg
from .scf import EIGH_MIN_N

PAD_RATIO = 0.75
DEGEN = 1e-9
DEGEN_FP32 = 1e-5  # levels from an FP32 exchange carry ~1e-6 relative error
GAP = 0.05
SHIFT_AFTER, SHIFT = 30, 0.3  # iterations before a level shift, and its size (Eh)

# per-block tensors of a class (first dimension: the block), compacted together
_BLOCK_KEYS = ("gs", "gc", "pad", "sys_t", "w", "S", "X", "Hs", "ev", "V", "Alast")
_HIST_KEYS = ("Fh", "Eh")


class _Channel:
    __slots__ = ("sys", "spin", "weight", "nocc", "basis", "blocks", "off")

    def __init__(self, sys, spin, weight, nocc, basis, off):
        self.sys, self.spin, self.weight, self.nocc, self.basis, self.off = sys, spin, weight, nocc, basis, off
        self.blocks = []


def aufbau_levels(lev, nocc, frozen=None, degen=DEGEN):
    """Occupied orbitals per block from the lowest levels of each block
    (lev: [blocks, k], inf where a block has no more levels): the nocc lowest
    overall, levels within `degen` of each other (a chain of such gaps) taken in
    block order, so which of several degenerate orbitals is occupied does not
    hinge on rounding (scf._Blocks.finish); with `frozen` (orbitals per block)
    each block takes its own lowest.  Returns (counts, highest occupied level)."""
    nb, km = lev.shape
    counts = np.zeros(nb, dtype=np.int64)
    if frozen is not None:
        avail = np.isfinite(lev).sum(axis=1)
        return np.minimum(np.asarray(frozen, dtype=np.int64), avail), np.inf
    if nocc <= 0:
        return counts, -np.inf
    v = lev.ravel()
    kk = np.repeat(np.arange(nb), km)
    jj = np.tile(np.arange(km), nb)
    fin = np.isfinite(v)
    v, kk, jj = v[fin], kk[fin], jj[fin]
    if v.size < nocc:
        raise RuntimeError(f"{nocc} occupied orbitals but only {v.size} levels in the blocks")
    o = np.lexsort((jj, kk, v))
    v, kk, jj = v[o], kk[o], jj[o]
    brk = np.nonzero(np.diff(v) > degen * np.maximum(1.0, np.abs(v[1:])))[0] + 1
    starts = np.concatenate([[0], brk])
    ends = np.concatenate([brk, [v.size]])
    g = np.searchsorted(starts, nocc - 1, side="right") - 1
    s0, e0 = int(starts[g]), int(ends[g])
    inner = np.lexsort((jj[s0:e0], kk[s0:e0]))
    take = np.concatenate([kk[:s0], kk[s0:e0][inner][:nocc - s0]])
    counts += np.bincount(take, minlength=nb)
    return counts, float(v[s0:e0][inner][nocc - 1 - s0])


class Cohort:
    """SCF of many atomic systems in lockstep (see the module docstring).

    systems: objects with .basis (AtomicBasis with compute_tei done, on
    `ctx`), .nela, .nelb.  precision "fp64" or "mixed".  symm: the symmetry
    blocks, 2 (per (l, m) shell, default) or 3 (per m and parity of l).
    occupations: per system None (aufbau) or (alpha, beta) occupied orbitals
    per block in get_sym_idx(symm) order (occupations.block_counts), held
    fixed as the reference's --readocc does; blocks holding none are never
    diagonalised."""

    def __init__(self, systems, ctx, precision="fp64", conv=1e-10, maxit=200, diis_len=10, switch=1e-3,
                 use_graphs=True, autotune=True, verbose=False, timing=True, occupations=None, symm=2):
        if symm not in (2, 3):
            raise ValueError("symm must be 2 (per (l, m)) or 3 (per m and parity)")
        self.symm = symm
        import torch
        if precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', not {precision!r}")
        self.torch = torch
        self.dev = torch.device("cuda", ctx.device)
        self.stream = torch.cuda.Stream(self.dev)
        self.systems = list(systems)
        self.ctx = ctx
        self.conv, self.maxit, self.H, self.switch = conv, maxit, diis_len, switch
        self.verbose, self.timing, self.use_graphs, self.autotune = verbose, timing, use_graphs, autotune
        self.t0 = time.perf_counter()
        self._last = self.t0
        self.phases = {}
        self._infos = []
        with torch.cuda.stream(self.stream):
            self._setup(precision)
            self._batches()
        if occupations is not None:
            for q, oc in enumerate(occupations):
                if oc is not None:
                    ca, cb = (tuple(int(x) for x in o) for o in oc)
                    if len(ca) != len(self.chan[q].blocks) or sum(ca) != self.chan[q].nocc:
                        raise ValueError(f"occupations of system {q} do not match its blocks or electrons")
                    self.frozen[q] = (ca, cb)
        self._tick("setup", force=True)

    def _tick(self, name, force=False):
        if not (self.timing or force):
            return
        self.torch.cuda.synchronize(self.dev)
        now = time.perf_counter()
        self.phases[name] = self.phases.get(name, 0.0) + now - self._last
        self._last = now

    # ------------------------------------------------------------------ setup
    def _setup(self, precision):
        torch = self.torch
        f64, dev = torch.float64, self.dev
        nsys = len(self.systems)
        self.nbf = np.array([s.basis.Nbf for s in self.systems], dtype=np.int64)
        sz = self.nbf ** 2
        self.soff = np.concatenate([[0], np.cumsum(sz)])
        ntot = int(self.soff[-1])
        # channels: the alpha channel of system q shares the system's offset
        # (the first ntot entries of the channel buffer mirror the system
        # buffers); beta channels follow
        self.chan = []
        self.restricted = np.array([s.nela == s.nelb for s in self.systems])
        for q, s in enumerate(self.systems):
            self.chan.append(_Channel(q, 0, 2.0 if self.restricted[q] else 1.0, s.nela, s.basis, int(self.soff[q])))
        self.beta_of = np.full(nsys, -1)
        off = ntot
        for q, s in enumerate(self.systems):
            if not self.restricted[q] and s.nelb > 0:
                self.beta_of[q] = len(self.chan)
                beta = s.beta_basis() if hasattr(s, "beta_basis") else s.basis.clone()
                self.chan.append(_Channel(q, 1, 1.0, s.nelb, beta, off))
                off += int(sz[q])
        ctot = off
        self.ntot, self.ctot = ntot, ctot
        # flat buffers; the last element of each is a zero that padded stack
        # entries gather from (and padded density entries, all zero, go to)
        self.Sf = torch.zeros(ntot + 1, dtype=f64, device=dev)
        self.Tf = torch.zeros(ntot + 1, dtype=f64, device=dev)
        self.Vf = torch.zeros(ntot + 1, dtype=f64, device=dev)
        self.Jf = torch.zeros(ntot + 1, dtype=f64, device=dev)
        self.Ptf = torch.zeros(ntot + 1, dtype=f64, device=dev)
        self.Pcf = torch.zeros(ctot + 1, dtype=f64, device=dev)
        self.Kf = torch.zeros(ctot + 1, dtype=f64, device=dev)
        for q, s in enumerate(self.systems):
            for buf, w in ((self.Sf, 0), (self.Tf, 1), (self.Vf, 2)):
                self._view(buf, int(self.soff[q]), q).copy_(s.basis.one_electron_dev(w))
        Hf = self.Tf + self.Vf
        # symmetry blocks of every channel (AtomicBasis.get_sym_idx): per (l, m)
        # shell (symm 2, the reference's symmetry=2: orbitals of pure l and m)
        # or per (m, parity of l) (symm 3: a density of definite-parity
        # orbitals has only even multipoles, whose Gaunt couplings join shells
        # of equal parity, so F never couples the two parities)
        blocks = []  # (channel, rows)
        for ci, c in enumerate(self.chan):
            for idx in c.basis.get_sym_idx(self.symm):
                c.blocks.append(len(blocks))
                blocks.append((ci, np.asarray(idx, dtype=np.int64)))
        self.nblk = len(blocks)
        self.blk_sys = np.array([self.chan[c].sys for c, _ in blocks])
        self.blk_spin = np.array([self.chan[c].spin for c, _ in blocks])
        order = sorted(range(self.nblk), key=lambda b: -blocks[b][1].size)
        groups = []
        for b in order:
            if groups and blocks[b][1].size >= PAD_RATIO * blocks[groups[-1][0]][1].size:
                groups[-1].append(b)
            else:
                groups.append([b])
        self.kmax = max(1, max(c.nocc for c in self.chan))
        self.cls = []
        for members in groups:
            # stacks below EIGH_MIN_N are padded up to it: the batched
            # eigensolver is an order of magnitude slower below that size
            N = max(blocks[members[0]][1].size, EIGH_MIN_N)
            B = len(members)
            rows = np.full((B, N), -1, dtype=np.int64)
            for p, b in enumerate(members):
                rows[p, :blocks[b][1].size] = blocks[b][1]
            ch = np.array([blocks[b][0] for b in members])
            sysq = self.blk_sys[members]
            rows_t = torch.as_tensor(rows, device=dev)
            valid = rows_t >= 0
            r = rows_t.clamp(min=0)
            lin = r[:, :, None] * torch.as_tensor(self.nbf[sysq], device=dev)[:, None, None] + r[:, None, :]
            vmask = valid[:, :, None] & valid[:, None, :]
            C = dict(members=np.array(members), N=N, nb=np.array([blocks[b][1].size for b in members]))
            C["gs"] = torch.where(vmask, torch.as_tensor(self.soff[sysq], device=dev)[:, None, None] + lin,
                                  torch.full_like(lin, ntot))
            C["gc"] = torch.where(vmask, torch.as_tensor([self.chan[c].off for c in ch], device=dev)[:, None, None]
                                  + lin, torch.full_like(lin, ctot))
            C["pad"] = (~valid).to(f64)
            C["sys_t"] = torch.as_tensor(sysq, device=dev)
            C["w"] = torch.as_tensor([self.chan[c].weight for c in ch], dtype=f64, device=dev)
            C["S"] = self.Sf[C["gs"]]
            C["Hs"] = Hf[C["gs"]]
            d = torch.diagonal(C["S"], dim1=1, dim2=2)
            d = torch.where(d > 0, 1.0 / torch.sqrt(d.clamp(min=1e-300)), torch.zeros_like(d))
            Sn = d[:, :, None] * C["S"] * d[:, None, :] + torch.diag_embed(C["pad"])
            e, v = self._eigh(Sn)
            X = d[:, :, None] * ((v * (1.0 / torch.sqrt(e))[:, None, :]) @ v.transpose(1, 2))
            C["X"] = X * (1.0 - C["pad"])[:, :, None] * (1.0 - C["pad"])[:, None, :]
            C["kc"] = kc = min(self.kmax, N)
            C["ev"] = torch.full((B, kc), float("inf"), dtype=f64, device=dev)
            C["V"] = torch.zeros((B, N, kc), dtype=f64, device=dev)
            C["Alast"] = torch.zeros((B, N, N), dtype=f64, device=dev)
            C["Fh"] = torch.zeros((self.H, B, N, N), dtype=f64, device=dev)
            C["Eh"] = torch.zeros_like(C["Fh"])
            self.cls.append(C)
        del Hf
        self._index_classes()
        # per-system state
        self.fp32 = np.full(nsys, precision == "mixed")
        self.it = np.zeros(nsys, dtype=np.int64)
        self.it_fp32 = np.zeros(nsys, dtype=np.int64)
        self.E = np.zeros(nsys)
        self.Eold = np.full(nsys, np.nan)
        self.done = np.zeros(nsys, dtype=bool)
        self.terms = np.zeros((nsys, 3))
        self.hist = [[] for _ in range(nsys)]
        self.occ_hist = [[] for _ in range(nsys)]
        self.frozen = [None] * nsys
        self.last_emax = np.full(nsys, np.inf)
        # DIIS: a ring of H slots shared by all systems; each system uses the
        # slots written since its last reset
        self.slot = 0
        self.valid = [[] for _ in range(nsys)]
        self.Bm = [np.zeros((0, 0)) for _ in range(nsys)]
        self.solved = np.zeros(self.nblk, dtype=bool)
        self.lam_min = np.full(self.nblk, np.inf)
        self.homo = None
        self.nsolves = 0
        self.nskipped = 0
        self.active = np.ones(nsys, dtype=bool)
        self.shift = np.zeros(nsys)
        self.occ_last = None

    def _index_classes(self):
        """Position of every block in its class, and, per class, the alpha
        blocks with the position of their beta partner (for the total
        density)."""
        torch = self.torch
        self.blk_pos = np.full((self.nblk, 2), -1, dtype=np.int64)
        for cl, C in enumerate(self.cls):
            for p, b in enumerate(C["members"]):
                self.blk_pos[b] = (cl, p)
        for cl, C in enumerate(self.cls):
            alpha = [p for p, b in enumerate(C["members"]) if self.blk_spin[b] == 0]
            partner = []
            for p in alpha:
                b = C["members"][p]
                q = self.blk_sys[b]
                pb = -1
                if self.beta_of[q] >= 0:
                    # the beta block of the same sub-problem (equal size, so the same class)
                    k = self.chan[q].blocks.index(b)
                    bb = self.chan[self.beta_of[q]].blocks[k]
                    cl2, p2 = self.blk_pos[bb]
                    if cl2 != cl:
                        raise RuntimeError("alpha and beta blocks of one sub-problem in different classes")
                    pb = int(p2)
                partner.append(pb)
            C["alpha"] = torch.as_tensor(np.array(alpha, dtype=np.int64), device=self.dev)
            pt = np.array(partner, dtype=np.int64)
            C["has_beta"] = torch.as_tensor(pt >= 0, device=self.dev)
            C["beta"] = torch.as_tensor(np.maximum(pt, 0), device=self.dev)

    def _view(self, buf, off, q):
        n = int(self.nbf[q])
        return buf[off:off + n * n].view(n, n)

    def _batches(self):
        """Batched J over the active systems and K over their channels."""
        from .atomic import AtomicBatch
        act = np.nonzero(self.active)[0]
        self.jsys = act.tolist()
        self.kch = [ci for ci, c in enumerate(self.chan) if self.active[c.sys]]
        self.jbatch = AtomicBatch([self.systems[q].basis for q in self.jsys], self.ctx)
        self.kbatch = AtomicBatch([self.chan[c].basis for c in self.kch], self.ctx)
        for b in (self.jbatch, self.kbatch):
            b.set_options(self.use_graphs, self.autotune)
        self.Pt_views = [self._view(self.Ptf, int(self.soff[q]), q) for q in self.jsys]
        self.J_views = [self._view(self.Jf, int(self.soff[q]), q) for q in self.jsys]
        self.Pc_views = [self._view(self.Pcf, self.chan[c].off, self.chan[c].sys) for c in self.kch]
        self.K_views = [self._view(self.Kf, self.chan[c].off, self.chan[c].sys) for c in self.kch]

    def _compact(self):
        """Drop the blocks of systems that left from every class (and the
        classes left empty) and rebuild the batched builds."""
        torch = self.torch
        keep_cls = []
        for C in self.cls:
            keep = np.nonzero(self.active[self.blk_sys[C["members"]]])[0]
            if keep.size == 0:
                continue
            if keep.size < len(C["members"]):
                kt = torch.as_tensor(keep, device=self.dev)
                for key in _BLOCK_KEYS:
                    C[key] = C[key].index_select(0, kt)
                for key in _HIST_KEYS:
                    C[key] = C[key].index_select(1, kt)
                C["members"] = C["members"][keep]
                C["nb"] = C["nb"][keep]
            keep_cls.append(C)
        self.cls = keep_cls
        self._index_classes()
        self._batches()

    # -------------------------------------------------------------- iteration
    def _fock_stacks(self, core_only=False):
        if core_only:
            return [C["Hs"].clone() for C in self.cls]
        return [C["Hs"] + self.Jf[C["gs"]] + self.Kf[C["gc"]] for C in self.cls]

    def _orthonormal(self, Fs):
        """A_b = X_b' F_b X_b per class (symmetrised; zero in padded entries),
        plus the level shift sigma (1 - C_occ C_occ') of a system that has not
        converged within SHIFT_AFTER iterations (the virtual levels raised by
        sigma against the previous occupied orbitals, which damps the
        occupied-virtual rotations DIIS over-steps; the converged state is
        that of the unshifted matrix, whose commutator the test uses)."""
        torch = self.torch
        out = []
        for C, F in zip(self.cls, Fs):
            X = C["X"]
            A = X.transpose(1, 2) @ F @ X
            A = 0.5 * (A + A.transpose(1, 2))
            if self.occ_last is not None and self.shift.any():
                sig = self.shift[self.blk_sys[C["members"]]]
                o = self.occ_last[C["members"]]
                k = int(o.max()) if o.size else 0
                if sig.any():
                    st = torch.as_tensor(sig, dtype=A.dtype, device=self.dev)
                    proj = torch.diag_embed(1.0 - C["pad"])
                    if k > 0:
                        ot = torch.as_tensor(o, device=self.dev)
                        W = C["V"][:, :, :k] * (torch.arange(k, device=self.dev)[None, :] < ot[:, None]).to(A.dtype)[:, None, :]
                        proj = proj - W @ W.transpose(1, 2)
                    A = A + st[:, None, None] * proj
            out.append(A)
        return out

    def _solve(self, As, todo):
        """Eigensolve the blocks flagged in `todo` (global block ids, bool),
        padded entries held above the spectrum (c = 2|A|_F + 1); keep each
        solved block's lowest kc eigenpairs and the matrix it was solved for."""
        torch = self.torch
        for C, A in zip(self.cls, As):
            pos = np.nonzero(todo[C["members"]])[0]
            if pos.size == 0:
                continue
            pt = torch.as_tensor(pos, device=self.dev)
            As_ = A.index_select(0, pt) if pos.size < len(C["members"]) else A
            c = 2.0 * torch.linalg.matrix_norm(As_) + 1.0
            e, v = self._eigh(As_ + torch.diag_embed(C["pad"].index_select(0, pt) * c[:, None]))
            kc = C["kc"]
            C["ev"].index_copy_(0, pt, e[:, :kc])
            C["V"].index_copy_(0, pt, v[:, :, :kc])
            C["Alast"].index_copy_(0, pt, As_)
            self.solved[C["members"][pos]] = True
            self.nsolves += int(pos.size)

    def _eigh(self, A):
        """Batched eigensolve (hf_eigh_batched, rocSOLVER dsyevd); its status
        is checked with the next transfer of the levels."""
        e, v, info = linalg.eigh(A, self.ctx)
        self._infos.append(info)
        return e, v

    def _levels(self):
        """Stored lowest levels of every block, on the host ([nblk, kmax])."""
        torch = self.torch
        parts = [C["ev"].reshape(-1) for C in self.cls]
        nval = sum(p.numel() for p in parts)
        host = torch.cat(parts + [i.to(torch.float64) for i in self._infos]).cpu().numpy()
        if np.any(host[nval:] != 0):
            raise RuntimeError("batched eigensolver did not converge")
        self._infos = []
        out = np.full((self.nblk, self.kmax), np.inf)
        o = 0
        for C in self.cls:
            kc, B = C["kc"], len(C["members"])
            blk = host[o:o + B * kc].reshape(B, kc)
            o += B * kc
            lim = np.minimum(C["nb"], kc)
            rows = np.where(np.arange(kc)[None, :] < lim[:, None], blk, np.inf)
            out[C["members"], :kc] = rows
        return out

    def _aufbau(self, lev, skipped):
        occ = np.zeros(self.nblk, dtype=np.int64)
        homo = np.full(len(self.chan), -np.inf)
        for ci, c in enumerate(self.chan):
            if not self.active[c.sys]:
                continue
            bl = np.asarray(c.blocks)
            rows = lev[bl].copy()
            rows[skipped[bl]] = np.inf
            fz = self.frozen[c.sys]
            cnt, homo[ci] = aufbau_levels(rows, c.nocc, None if fz is None else fz[c.spin],
                                          DEGEN_FP32 if self.fp32[c.sys] else DEGEN)
            occ[bl] = cnt
        return occ, homo

    def _diagonalise(self, Fs):
        """Occupations of the Fock matrices Fs.  A block solved before whose
        lowest level, lowered by the Frobenius norm of the change of its matrix
        (Weyl's bound), stays GAP above its channel's previous highest occupied
        level holds no occupied orbital and is not solved; any whose bound
        fails against the new highest occupied level is solved after all, so
        the occupations are exactly those of solving every block."""
        torch = self.torch
        As = self._orthonormal(Fs)
        skip = np.zeros(self.nblk, dtype=bool)
        bound = np.full(self.nblk, -np.inf)
        live = np.zeros(self.nblk, dtype=bool)
        for C in self.cls:
            live[C["members"]] = True
        # fixed occupations: blocks holding no orbital are never needed
        for c in self.chan:
            fz = self.frozen[c.sys]
            if self.active[c.sys] and fz is not None:
                bl = np.asarray(c.blocks)
                skip[bl] = np.asarray(fz[c.spin]) == 0
                bound[bl] = np.inf
        if self.homo is not None and self.solved.any():
            dn = torch.cat([torch.linalg.matrix_norm(A - C["Alast"]) for C, A in zip(self.cls, As)]).cpu().numpy()
            o = 0
            for C in self.cls:
                B = len(C["members"])
                fixed = np.array([self.frozen[self.blk_sys[b]] is not None for b in C["members"]])
                bound[C["members"]] = np.where(fixed, bound[C["members"]], self.lam_min[C["members"]] - dn[o:o + B])
                o += B
            for ci, c in enumerate(self.chan):
                if not self.active[c.sys] or self.frozen[c.sys] is not None:
                    continue
                bl = np.asarray(c.blocks)
                skip[bl] = self.solved[bl] & (bound[bl] > self.homo[ci] + GAP)
        self._solve(As, live & ~skip)
        lev = self._levels()
        self.lam_min = np.where(live, lev[:, 0], self.lam_min)
        occ, homo = self._aufbau(lev, skip)
        late = np.zeros(self.nblk, dtype=bool)
        for ci, c in enumerate(self.chan):
            if self.frozen[c.sys] is not None:
                continue  # fixed occupations: the skipped blocks hold none
            bl = np.asarray(c.blocks)
            late[bl] = skip[bl] & ~(bound[bl] > homo[ci])
        if late.any():
            self._solve(As, late)
            skip &= ~late
            lev = self._levels()
            self.lam_min = np.where(live, lev[:, 0], self.lam_min)
            occ, homo = self._aufbau(lev, skip)
        self.nskipped += int(skip.sum())
        self.homo = homo
        self.occ_last = occ
        return occ

    def _densities(self, occ):
        """Spin densities of every block into the channel buffer (K input) and
        the weighted sum of a system's channels into the total-density buffer
        (J input).  Entries outside the blocks are never written (zero)."""
        torch = self.torch
        self.Pb = []
        for C in self.cls:
            o = occ[C["members"]]
            k = int(o.max()) if o.size else 0
            if k == 0:
                P = torch.zeros((len(C["members"]), C["N"], C["N"]), dtype=torch.float64, device=self.dev)
            else:
                ot = torch.as_tensor(o, device=self.dev)
                mask = (torch.arange(k, device=self.dev)[None, :] < ot[:, None]).to(torch.float64)
                W = C["X"] @ (C["V"][:, :, :k] * mask[:, None, :])
                P = W @ W.transpose(1, 2)
            self.Pb.append(P)
            self.Pcf.index_put_((C["gc"].reshape(-1),), P.reshape(-1))
            a = C["alpha"]
            if a.numel():
                Pa = P.index_select(0, a)
                Pbeta = P.index_select(0, C["beta"]) * C["has_beta"].to(P.dtype)[:, None, None]
                Pt = C["w"].index_select(0, a)[:, None, None] * Pa + Pbeta
                self.Ptf.index_put_((C["gs"].index_select(0, a).reshape(-1),), Pt.reshape(-1))
        self.Pcf[-1] = 0.0
        self.Ptf[-1] = 0.0

    def _fock_build(self):
        self.jbatch.coulomb(self.Pt_views, self.J_views)
        self.kbatch.exchange(self.Pc_views, self.K_views,
                             fp32=[bool(self.fp32[self.chan[c].sys]) for c in self.kch])

    def _step_terms(self, Fs, Js, Ks):
        """Per system on the device: (tr(P H), 1/2 tr(P J), Exx), the largest
        orbital-gradient entry, and the DIIS dot products of the new error with
        every ring slot.  Errors are X'(F P S - S P F)X per block with P the
        spin density times the channel weight."""
        torch = self.torch
        nsys = len(self.systems)
        s = self.slot
        terms = torch.zeros((3, nsys), dtype=torch.float64, device=self.dev)
        emax = torch.zeros(nsys, dtype=torch.float64, device=self.dev)
        dots = torch.zeros((self.H + 1, nsys), dtype=torch.float64, device=self.dev)
        for C, F, J, K, P in zip(self.cls, Fs, Js, Ks, self.Pb):
            wP = C["w"][:, None, None] * P
            e3 = torch.stack([(wP * C["Hs"]).sum(dim=(1, 2)), 0.5 * (wP * J).sum(dim=(1, 2)),
                              0.5 * (wP * K).sum(dim=(1, 2))])
            terms.index_add_(1, C["sys_t"], e3)
            FPS = F @ wP @ C["S"]
            E = C["X"].transpose(1, 2) @ (FPS - FPS.transpose(1, 2)) @ C["X"]
            emax.scatter_reduce_(0, C["sys_t"], E.abs().amax(dim=(1, 2)), reduce="amax")
            C["Fh"][s] = F
            C["Eh"][s] = E
            dots[:self.H].index_add_(1, C["sys_t"], (C["Eh"] * E[None]).sum(dim=(2, 3)))
        dots[self.H] = dots[s]
        return terms, emax, dots

    def _extrapolate(self, coef):
        """F_b = sum_t coef[sys(b), t] F_hist[t]_b per class."""
        torch = self.torch
        cf = torch.as_tensor(coef, dtype=torch.float64, device=self.dev)  # [nsys, H]
        out = []
        for C in self.cls:
            w = cf.index_select(0, C["sys_t"]).transpose(0, 1)  # [H, B]
            out.append(torch.einsum("hb,hbij->bij", w, C["Fh"]))
        return out

    def _check_flips(self, q):
        """Aufbau flipping between occupation patterns never converges: after
        3 changes within 10 iterations keep the pattern of the lowest energy
        (as scf.SCFState._check_flips)."""
        if self.frozen[q] is not None or self.it[q] < 10:
            return False
        pats = self.occ_hist[q]
        n = min(len(self.hist[q]), len(pats))
        recent = pats[n - 10:n]
        if sum(1 for a, b in zip(recent, recent[1:]) if a != b) < 3:
            return False
        best = min(range(n), key=lambda t: self.hist[q][t])
        self.frozen[q] = pats[best]
        return True

    def _record_patterns(self, occ):
        for q in np.nonzero(self.active)[0]:
            pat = [tuple(int(x) for x in occ[self.chan[q].blocks])]
            if self.beta_of[q] >= 0:
                pat.append(tuple(int(x) for x in occ[self.chan[self.beta_of[q]].blocks]))
            elif not self.restricted[q]:
                pat.append(tuple(0 for _ in pat[0]))
            self.occ_hist[q].append(tuple(pat))

    def solve(self):
        """Iterate until every system converged (or hit maxit); returns one
        result dict per system, in input order."""
        torch = self.torch
        nsys = len(self.systems)
        with torch.cuda.stream(self.stream):
            Fs = self._fock_stacks(core_only=True)
            while True:
                occ = self._diagonalise(Fs)
                self._record_patterns(occ)
                self._tick("eigensolve+aufbau")
                self._densities(occ)
                self._fock_build()
                self._tick("fock build")
                Js = [self.Jf[C["gs"]] for C in self.cls]
                Ks = [self.Kf[C["gc"]] for C in self.cls]
                Fs = [C["Hs"] + J + K for C, J, K in zip(self.cls, Js, Ks)]
                terms, emax, dots = self._step_terms(Fs, Js, Ks)
                host = torch.cat([terms.reshape(-1), emax, dots.reshape(-1)]).cpu().numpy()
                del Js, Ks
                self._tick("energy+diis")
                terms_h = host[:3 * nsys].reshape(3, nsys).T
                emax_h = host[3 * nsys:4 * nsys]
                dots_h = host[4 * nsys:].reshape(self.H + 1, nsys)
                coef = np.zeros((nsys, self.H))
                coef[:, self.slot] = 1.0
                for q in np.nonzero(self.active)[0]:
                    self._advance(q, terms_h[q], emax_h[q], dots_h[:, q], coef)
                self.slot = (self.slot + 1) % self.H
                left = self.active & ~self.done
                if not left.any():
                    break
                if (left != self.active).any():
                    self.active = left
                    self._compact()
                Fs = self._extrapolate(coef)
        self._tick("energy+diis")
        return [self._result(q) for q in range(nsys)]

    def _advance(self, q, terms, emax, dq, coef):
        """System q's energy, convergence test, precision switch and DIIS
        (the decisions of scf.SCFState.advance)."""
        self.it[q] += 1
        self.terms[q] = terms
        E = float(terms.sum()) + float(getattr(self.systems[q].basis, "Enucr", 0.0))
        self.E[q] = E
        self.last_emax[q] = emax
        self.hist[q].append(E)
        if self.verbose:
            print(f"sys {q} it {self.it[q]} E {E:.12f} err {emax:.3e}{' (fp32 K)' if self.fp32[q] else ''}",
                  flush=True)
        if self._check_flips(q):
            self._reset_diis(q)
            self.Eold[q] = np.nan
        if self.fp32[q]:
            self.it_fp32[q] = self.it[q]
            if emax < self.switch:
                # FP64 from here on: the FP32 Fock matrices leave the DIIS
                # history and this step is a plain diagonalisation
                self.fp32[q] = False
                self.Eold[q] = np.nan
                self._reset_diis(q)
                return
        elif not np.isnan(self.Eold[q]) and abs(E - self.Eold[q]) < self.conv and emax < 1e-6:
            self.done[q] = True
            return
        else:
            self.Eold[q] = E
        if self.it[q] >= self.maxit:
            self.done[q] = True
            return
        if self.it[q] >= SHIFT_AFTER and self.shift[q] == 0.0:
            # slow to converge: level-shift the virtual orbitals from here on,
            # with a fresh DIIS over the shifted Fock matrices
            self.shift[q] = SHIFT
            self._reset_diis(q)
        coef[q] = self._diis_coef(q, dq)

    def _reset_diis(self, q):
        self.valid[q] = []
        self.Bm[q] = np.zeros((0, 0))

    def _diis_coef(self, q, dq):
        """Extend system q's DIIS matrix by the new slot and solve for the
        extrapolation coefficients (one per ring slot)."""
        s = self.slot
        keep = [i for i, t in enumerate(self.valid[q]) if t != s]  # the slot being reused leaves
        B = self.Bm[q][np.ix_(keep, keep)] if keep else np.zeros((0, 0))
        v = [self.valid[q][i] for i in keep] + [s]
        n = len(v)
        Bn = np.zeros((n, n))
        Bn[:n - 1, :n - 1] = B
        row = np.array([dq[t] for t in v[:-1]] + [dq[self.H]])
        Bn[n - 1, :] = row
        Bn[:, n - 1] = row
        self.valid[q], self.Bm[q] = v, Bn
        A = -np.ones((n + 1, n + 1))
        A[:n, :n] = Bn
        A[n, n] = 0.0
        rhs = np.zeros(n + 1)
        rhs[n] = -1.0
        out = np.zeros(self.H)
        try:
            c = np.linalg.solve(A, rhs)[:n]
        except np.linalg.LinAlgError:
            out[s] = 1.0
            return out
        for t, ct in zip(v, c):
            out[t] = ct
        return out

    def _result(self, q):
        t = self.terms[q]
        return dict(E=float(self.E[q]), E1=float(t[0]), Ecoul=float(t[1]), Exx=float(t[2]),
                    iterations=int(self.it[q]), iterations_fp32=int(self.it_fp32[q]), history=list(self.hist[q]),
                    converged=bool(self.done[q] and self.it[q] < self.maxit), orbital_gradient=float(self.last_emax[q]),
                    frozen_occupations=self.frozen[q] is not None, phases=dict(self.phases),
                    eigensolves=self.nsolves, skipped_eigensolves=self.nskipped)


def cohort_bytes(system, diis_len=10, symm=2):
    """Device bytes a system holds in a cohort: its flat matrices (S, T, V, J,
    total density; density and exchange per spin channel) and its padded
    block stacks (overlap, orthogonaliser, core Hamiltonian, last solved
    matrix, DIIS history of Fock matrices and errors, and the transient Fock
    and gathered J, K stacks of an iteration)."""
    b = system.basis
    nch = 1 if system.nela == system.nelb else (2 if system.nelb > 0 else 1)
    flat = 8 * b.Nbf ** 2 * (5 + 2 * nch)
    blk = sum(len(i) ** 2 for i in b.get_sym_idx(symm))
    return flat + 8 * nch * blk * (2 * diis_len + 10)


def cohort_groups(systems, budget_bytes, diis_len=10, symm=2):
    """Consecutive runs of system indices whose cohort_bytes sum stays within
    `budget_bytes` (a system larger than the budget gets a cohort of its own)."""
    groups, chunk, used = [], [], 0
    for i, s in enumerate(systems):
        need = cohort_bytes(s, diis_len, symm)
        if chunk and used + need > budget_bytes:
            groups.append(chunk)
            chunk, used = [], 0
        chunk.append(i)
        used += need
    if chunk:
        groups.append(chunk)
    return groups


def solve_cohorts(systems, ctx, budget_bytes=96 << 30, **kw):
    """All systems through cohorts of at most `budget_bytes` of device memory
    each, one after another (input order); returns the results in input order
    (each with the index of its cohort under "cohort") and the wall time."""
    t0 = time.perf_counter()
    occ = kw.pop("occupations", None)
    groups = cohort_groups(systems, budget_bytes, kw.get("diis_len", 10), kw.get("symm", 2))
    res = [None] * len(systems)
    import gc
    import torch
    for c, g in enumerate(groups):
        co = Cohort([systems[i] for i in g], ctx, occupations=None if occ is None else [occ[i] for i in g], **kw)
        for i, r in zip(g, co.solve()):
            r["cohort"] = c
            res[i] = r
        # hand the cohort's memory back: a cohort holds up to the budget in
        # padded stacks, and blocks cached from the last one would leave the
        # next one's allocations to fail, free the whole cache and retry
        # (seconds per cohort)
        del co
        gc.collect()
        torch.cuda.empty_cache()
    return res, time.perf_counter() - t0
