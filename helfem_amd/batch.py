"""Batched and served SCF solves on one GPU.

Many independent systems (atoms of different Z, element counts and lmax;
diatomic molecules) are solved together by one `SCFEngine`: each step
diagonalises the pending Fock blocks of every system in shared batched
eigensolves and then runs each system's Fock build on the GPU.  Systems are
admitted while their estimated device footprint fits a memory budget and
leave as soon as they converge, so new ones can join mid-flight (continuous
batching).  `serve` feeds the engine from a stream of timed arrivals and
prepares each request's basis (host two-electron setup) on a thread pool
while the GPU keeps stepping the admitted ones.
"""
from __future__ import annotations

import heapq
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from .scf import SCFEngine, SCFState


@dataclass
class System:
    """One SCF problem: a basis with its two-electron setup done, the
    electron counts per spin and an optional DFT grid (LDA)."""
    basis: object
    nela: int
    nelb: int
    lda: bool = False
    tag: object = None
    grid: object = None
    arrival: float = 0.0
    info: dict = field(default_factory=dict)
    # fixed occupations (alpha, beta) per (l, m) shell block of the ground-state
    # configuration (occupations.block_counts, symmetry 2), or None: aufbau
    occ: tuple = None

    @property
    def restricted(self):
        return self.nela == self.nelb

    def beta_basis(self):
        """A second handle on the basis for the beta density of an
        unrestricted solve, made once and kept (its device tables with it),
        so later solves of this system reuse it."""
        b = self.__dict__.get("_beta")
        if b is None:
            b = self.basis.clone()
            self.__dict__["_beta"] = b
        return b

    def footprint(self):
        """Device bytes held while in flight: S, T, V, H, P (per spin), J, K
        or Vxc, F and its DIIS blocks, about 14 Nbf^2 doubles."""
        return 8 * 14 * self.basis.Nbf ** 2


def _xc_for(sys_):
    if not sys_.lda:
        return None
    g = sys_.grid
    if sys_.restricted:
        def xc(P):
            H, Exc, _, _ = g.eval_Fxc(1, 7, P)
            return H, Exc
        return xc

    def xc_pol(Pa, Pb):
        Ha, Hb, Exc, _, _ = g.eval_Fxc_pol(1, 7, Pa, Pb, beta=sys_.nelb > 0)
        return Ha, Hb, Exc
    return xc_pol


def state_for(sys_, precision="fp64", conv=1e-10, maxit=200, gconv=1e-6):
    """SCFState of a system: RHF/RKS with nela doubly occupied orbitals when
    the spins pair up, UHF/UKS otherwise; converged when the energy changes by
    less than conv and the largest orbital-gradient entry is below gconv."""
    nel = (sys_.nela,) if sys_.restricted else (sys_.nela, sys_.nelb)
    from .atomic import AtomicBasis
    # atoms: orbitals of pure (l, m) (the reference's symmetry=2), with the
    # configuration's fixed occupations when the basis holds them (as the
    # vectorised cohort)
    atomic = isinstance(sys_.basis, AtomicBasis)
    occ = None
    if atomic and sys_.occ is not None:
        occ = sys_.occ[:1] if sys_.restricted else sys_.occ
    return SCFState(sys_.basis, nel, sys_.restricted, symm=2 if atomic else 1, conv=conv, maxit=maxit,
                    xc=_xc_for(sys_), precision=precision, tag=sys_.tag, occupations=occ, gconv=gconv)


def solve_batched(systems, precision="fp64", conv=1e-10, budget_bytes=64 << 30):
    """All systems through one engine, admitted by footprint.  Returns the
    results (in input order) and the wall time."""
    t0 = time.perf_counter()
    eng = SCFEngine()
    queue = list(enumerate(systems))
    results = [None] * len(systems)
    in_use = {}
    while queue or eng.active:
        used = sum(in_use.values())
        while queue and (not eng.active or used + queue[0][1].footprint() <= budget_bytes):
            i, s = queue.pop(0)
            st = state_for(s, precision, conv)
            st.tag = i
            eng.add(st)
            in_use[i] = s.footprint()
            used += in_use[i]
        for st in eng.step():
            results[st.tag] = st.result()
            del in_use[st.tag]
    return results, time.perf_counter() - t0


def solve_sequential(systems, precision="fp64", conv=1e-10, gconv=1e-6):
    """One engine per system, one after another (the plain drop-in way)."""
    t0 = time.perf_counter()
    out = []
    for i, s in enumerate(systems):
        eng = SCFEngine()
        st = state_for(s, precision, conv, gconv=gconv)
        eng.add(st)
        while eng.active:
            eng.step()
        out.append(st.result())
    return out, time.perf_counter() - t0


# --- workload generators ------------------------------------------------------

_MADELUNG = sorted(((n, l) for n in range(1, 8) for l in range(0, n)), key=lambda s: (s[0] + s[1], s[0]))


def _spins(Z):
    """Ground-state-like spin counts: Hund's rule on the last open subshell of
    the Madelung filling (nela - nelb unpaired electrons); equal to
    occupations.spins."""
    left = Z
    for n, l in _MADELUNG:
        cap = 2 * (2 * l + 1)
        if left <= cap:
            unpaired = left if left <= cap // 2 else cap - left
            return (Z + unpaired) // 2, (Z - unpaired) // 2
        left -= cap
    return (Z + 1) // 2, Z // 2


def random_atoms(n, seed, ctx, zmax=86, nel_range=(4, 24), orders=(10, 15), lmax_range=(0, 4), setup_threads=8,
                 device_setup=True):
    """Seeded atomic systems: Z ~ U{1..zmax}, elements ~ U{nel_range}, LIP
    order from `orders`, lmax = mmax ~ U{lmax_range}; neutral, spins from
    Hund's rule.  The bases are built on a thread pool, their two-electron
    setups on the GPU (device_setup, the default when ctx is given) or on
    the host."""
    from .atomic import AtomicBasis
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n):
        Z = int(rng.integers(1, zmax + 1))
        nel = int(rng.integers(nel_range[0], nel_range[1] + 1))
        nnodes = int(rng.choice(orders))
        lmax = int(rng.integers(lmax_range[0], lmax_range[1] + 1))
        specs.append((Z, nel, nnodes, lmax))

    def build(spec):
        Z, nel, nnodes, lmax = spec
        b = AtomicBasis(Z, lmax, lmax, nel, nnodes=nnodes, ctx=ctx)
        b.compute_tei(device=device_setup and ctx is not None)
        if ctx is not None:
            b.prepare_one_electron()
        return b

    with ThreadPoolExecutor(setup_threads) as ex:
        bases = list(ex.map(build, specs))
    return [atom_system(b, spec[0], tag=spec, info=dict(Z=spec[0], nelem=spec[1], nnodes=spec[2], lmax=spec[3],
                                                        Nbf=b.Nbf)) for spec, b in zip(specs, bases)]


def atom_system(basis, Z, tag=None, info=None):
    """A neutral atom with the fixed occupations (and Hund's-rule spins) of
    its Madelung ground-state configuration where the basis holds every
    occupied (l, m); in a basis without some occupied l, of the Madelung/Hund
    filling of the subshells it holds (what aufbau reaches there, pinned so
    the state does not hinge on rounding); None (aufbau) if neither fits."""
    from . import occupations as oc
    occ = oc.block_counts(Z, basis.lval, basis.mval, symm=2)
    na, nb = oc.spins(Z)
    info = dict(info or {})
    if occ is None:
        lmax = int(max(basis.lval))
        occ = oc.block_counts(Z, basis.lval, basis.mval, symm=2, lmax=lmax)
        if occ is not None:
            na, nb = oc.spins(Z, lmax)
            info["configuration"] = f"Madelung/Hund over l <= {lmax}"
    return System(basis, na, nb, tag=tag, info=info, occ=occ)


def random_diatomic_specs(n, seed, zmax=18, r_range=(1.0, 6.0), nel_range=(4, 12), lmax_range=(8, 24),
                          mmax_range=(0, 3)):
    """Seeded diatomic requests: Z1, Z2 ~ U{1..zmax}, R ~ U(r_range) bohr,
    mu-elements ~ U{nel_range}, lmax ~ U{lmax_range} for every |m| <= mmax ~
    U{mmax_range}; HF or LDA with equal odds; neutral, restricted when the
    electron count is even (else one unpaired electron)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        Z1, Z2 = int(rng.integers(1, zmax + 1)), int(rng.integers(1, zmax + 1))
        R = float(rng.uniform(*r_range))
        nel = int(rng.integers(nel_range[0], nel_range[1] + 1))
        lmax = int(rng.integers(lmax_range[0], lmax_range[1] + 1))
        mmax = int(rng.integers(mmax_range[0], mmax_range[1] + 1))
        lda = bool(rng.integers(0, 2))
        out.append(dict(Z1=Z1, Z2=Z2, R=R, nelem=nel, lmax=lmax, mmax=mmax, lda=lda))
    return out


def build_diatomic(spec, ctx, device_setup=True):
    """A served request's system: the basis with its two-electron setup (on
    the GPU by default) and, for LDA, its pure-m DFT grid."""
    from .diatomic import DiatomicBasis
    from .dftgrid import DFTGrid
    b = DiatomicBasis(spec["Z1"], spec["Z2"], spec["R"], [spec["lmax"]] * (spec["mmax"] + 1), spec["nelem"], ctx=ctx)
    b.compute_tei(exchange=not spec["lda"], device=device_setup and ctx is not None)
    ne = spec["Z1"] + spec["Z2"]
    s = System(b, (ne + 1) // 2, ne // 2, lda=spec["lda"], tag=spec, info=dict(spec, Nbf=b.Nbf))
    if spec["lda"]:
        s.grid = DFTGrid(b)
    return s


def serve(specs, rate_hz, seed, ctx, precision="fp64", conv=1e-10, budget_bytes=64 << 30, setup_threads=8,
          build=build_diatomic, gconv=1e-6, on_done=None, overlap=True):
    """Poisson arrivals at `rate_hz` of the requests in `specs`; each request's
    basis is prepared on a host thread pool as it arrives and admitted into
    the running engine when ready and when its footprint fits the budget.
    Returns per-request results with arrival, admission and completion times;
    on_done(index, result) is called as each request completes."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_hz, size=len(specs))
    arrivals = np.cumsum(gaps) - gaps[0]
    t0 = time.perf_counter()
    ex = ThreadPoolExecutor(setup_threads)
    futures = {}
    ready = []  # heap of (arrival, index)
    eng = SCFEngine()
    in_use, admitted_at, results = {}, {}, {}
    nxt = 0
    eng.overlap = overlap
    while len(results) < len(specs):
        now = time.perf_counter() - t0
        while nxt < len(specs) and arrivals[nxt] <= now:
            futures[nxt] = ex.submit(build, specs[nxt], ctx)
            nxt += 1
        for i in [i for i, f in futures.items() if f.done()]:
            s = futures.pop(i).result()
            s.arrival = float(arrivals[i])
            heapq.heappush(ready, (s.arrival, i, s))
        used = sum(in_use.values())
        while ready and (not eng.active or used + ready[0][2].footprint() <= budget_bytes):
            _, i, s = heapq.heappop(ready)
            st = state_for(s, precision, conv, gconv=gconv)
            st.tag = (i, s)
            eng.add(st)
            in_use[i] = s.footprint()
            used += in_use[i]
            admitted_at[i] = time.perf_counter() - t0
        if not eng.active:
            # nothing to step: wait for the next arrival or setup
            wait = [arrivals[nxt] - now] if nxt < len(specs) else []
            time.sleep(max(0.0, min(wait + [0.005])))
            continue
        for st in eng.step():
            i, s = st.tag
            r = st.result()
            r["arrival_s"] = s.arrival
            r["admitted_s"] = admitted_at[i]
            r["done_s"] = time.perf_counter() - t0
            r["latency_s"] = r["done_s"] - s.arrival
            r["spec"] = s.info
            results[i] = r
            if on_done is not None:
                on_done(i, r)
            del in_use[i]
    ex.shutdown()
    return [results[i] for i in range(len(specs))], time.perf_counter() - t0


def latency_summary(results):
    lat = np.array([r["latency_s"] for r in results])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}
