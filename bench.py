#!/usr/bin/env python3
"""Fock-build benchmark on one MI355X.

A step is one Fock build of the two-electron part: J(P) and K(P/2) on a
seeded density resident in HBM (FP64 plain mode, the drop-in path).  The
default workload is diatomic N2 at R = 2.074 bohr in prolate spheroidal
coordinates: 8 mu-elements to mu_max = arcosh(40 / Rhalf), LIP-15, lmax 16
for m = 0..3 (Nbf 11894), 14 electrons in sigma/pi blocks.  Atomic
workloads (Xe lmax=mmax=6, Ne) are selectable with --workload.

Prints one JSON line (rank 0):
  value        Fock builds/s over the timed steps (inputs already in HBM)
  roofline     dominant kernel (exchange): algorithmic FP64 flops per launch
               / average launch time from HIP events on its stream, against
               the 78.6 TFLOP/s FP64 matrix peak
  cpu_baseline the CPU oracle (same loop structure as the reference) on the
               host cores, on a bounded sample of the same workload
  latency_ms   p50/p99 of per-step Fock-build latency
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = json.loads((ROOT / "BASELINE.json").read_text())["metric"]
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0

# shells fill in Madelung (n+l, n) order
_MADELUNG = sorted(((n, l) for n in range(1, 8) for l in range(0, n)), key=lambda s: (s[0] + s[1], s[0]))

WORKLOADS = {
    "n2-hf-diatomic": dict(kind="diatomic", Z1=7, Z2=7, R=2.074, lmmax=[16, 16, 16, 16], nelem=8, seed=1002,
                           occ={0: 5, 1: 1, -1: 1},
                           desc="N2 HF diatomic, R=2.074 bohr, 8 mu-elements to r~40 bohr, LIP-15, lmax=16 for "
                                "m=0..3, Nbf 11894"),
    "xe-hf-atomic": dict(kind="atomic", Z=54, lmax=6, mmax=6, nelem=25, seed=1003,
                         desc="Xe HF atomic, lmax=mmax=6, 25 LIP-15 elements to 40 bohr (exp grid), Nbf 17101"),
    "ne-hf-atomic": dict(kind="atomic", Z=10, lmax=0, mmax=0, nelem=10, seed=1001,
                         desc="Ne HF atomic, lmax=0, 10 LIP-15 elements to 40 bohr (exp grid), Nbf 139"),
    "ne-hf-atomic-l1": dict(kind="atomic", Z=10, lmax=1, mmax=1, nelem=10, seed=1001,
                            desc="Ne HF atomic, lmax=mmax=1, 10 LIP-15 elements to 40 bohr, Nbf 556"),
}
# exchange launches (timed together for the roofline) and the Coulomb kernels
KERNELS = {"atomic": (("atomic_k_off", "atomic_k_in"), ("atomic_ptile", "atomic_j_gather", "atomic_j_expand")),
           "diatomic": (("diatomic_k_off", "diatomic_k_in"),
                        ("diatomic_ptile", "diatomic_j_gather", "diatomic_j_expand"))}


def occupied_shells(Z: int):
    """Number of occupied shells per l for the neutral atom (Madelung filling)."""
    left, per_l = Z, {}
    for n, l in _MADELUNG:
        if left <= 0:
            break
        per_l[l] = per_l.get(l, 0) + 1
        left -= 2 * (2 * l + 1)
    return per_l


def block_occupations(Z, mvals, lmax):
    """Doubly-occupied orbitals per m-block (blocks in ascending m)."""
    shells = occupied_shells(Z)
    out = []
    for m in sorted(set(mvals)):
        out.append(sum(c for l, c in shells.items() if l <= lmax and abs(m) <= l))
    return out


def measured_traffic(workload):
    """HBM bytes per exchange launch from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE x 2 and WRITE_SIZE x 1, the factors
    a plain device copy of known size gives), or None when none is committed."""
    f = ROOT / "profiles" / f"traffic_{workload}.json"
    if not f.exists():
        return None
    d = json.loads(f.read_text())
    return d.get("exchange_bytes_per_build")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="n2-hf-diatomic", choices=sorted(WORKLOADS) + ["batch-atoms", "serve-diatomic"],
                    help="a Fock-build workload, or batch-atoms / serve-diatomic (many SCF solves on one GPU)")
    ap.add_argument("--systems", type=int, default=0,
                    help="systems in batch-atoms (default 32) or requests in serve-diatomic (default 16)")
    ap.add_argument("--rate", type=float, default=2.0, help="serve-diatomic: Poisson arrival rate (1/s)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-scf", action="store_true", help="skip the timed full SCF solve")
    ap.add_argument("--cpu-sample-stride", type=int, default=0,
                    help="time every n-th exchange output block on the CPU and extrapolate "
                         "(0: 24 for atomic, 64 for diatomic workloads)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain device copy of known size (for PMC byte calibration)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_basis(H, cfg, ctx):
    if cfg["kind"] == "atomic":
        return H.AtomicBasis(cfg["Z"], cfg["lmax"], cfg["mmax"], cfg["nelem"], ctx=ctx)
    return H.DiatomicBasis(cfg["Z1"], cfg["Z2"], cfg["R"], cfg["lmmax"], cfg["nelem"], ctx=ctx)


def seeded_occupations(cfg, hb):
    """Doubly occupied orbitals per m-block (blocks in ascending m)."""
    ms = sorted(set(hb.mval.tolist()))
    if cfg["kind"] == "atomic":
        return block_occupations(cfg["Z"], hb.mval.tolist(), cfg["lmax"])
    return [cfg["occ"].get(m, 0) for m in ms]


def cpu_baseline(cfg, P, nblocks_total, stride):
    """Oracle J (full) and K (strided sample of output blocks, extrapolated)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if cfg["kind"] == "atomic":
        ob = O.AtomicBasis(cfg["Z"], cfg["lmax"], cfg["mmax"], cfg["nelem"])
    else:
        ob = O.DiatomicBasis(cfg["Z1"], cfg["Z2"], cfg["R"], cfg["lmmax"], cfg["nelem"])
    t = time.perf_counter()
    ob.compute_tei()
    log(f"oracle setup {time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    ob.coulomb(P)
    tJ = time.perf_counter() - t
    sample = np.arange(0, nblocks_total, stride, dtype=np.int32)
    t = time.perf_counter()
    ob.exchange_blocks(0.5 * P, sample)
    tKs = time.perf_counter() - t
    tK = tKs * nblocks_total / sample.size
    log(f"oracle J {tJ:.2f}s, K sample {sample.size}/{nblocks_total} blocks {tKs:.2f}s -> K {tK:.1f}s")
    return {"value": 1.0 / (tJ + tK), "unit": "Fock builds/s", "cores": threads, "kind": "port",
            "sample": f"oracle/ (C++ -O3, OpenMP {threads} threads): one full coulomb + exchange over "
                      f"{sample.size} of {nblocks_total} output blocks (every {stride}th), "
                      f"extrapolated; measured {tJ + tKs:.1f}s of CPU work",
            "seconds": {"coulomb": tJ, "exchange_estimated": tK}}


def main_batch(args):
    """Many independent atomic SCF solves through one lockstep engine (shared
    padded batched eigensolves, admission by HBM footprint), against the same
    systems solved one after another."""
    import helfem_amd as H
    from helfem_amd import batch
    n = args.systems or 32
    ctx = H.Context(0)
    t = time.perf_counter()
    systems = batch.random_atoms(n, seed=1004, ctx=ctx, zmax=86, nel_range=(4, 24), orders=(10, 15),
                                 lmax_range=(0, 4), setup_threads=16)
    log(f"{n} atoms, Nbf {min(s.info['Nbf'] for s in systems)}..{max(s.info['Nbf'] for s in systems)}; "
        f"host setup {time.perf_counter() - t:.1f}s")
    out = {}
    for prec in ("mixed", "fp64"):
        rb, tb = batch.solve_batched(systems, precision=prec, conv=1e-8)
        log(f"{prec}: batched {tb:.1f}s")
        rs, ts = batch.solve_sequential(systems, precision=prec, conv=1e-8)
        out[prec] = {"batched_solves_per_s": n / tb, "sequential_solves_per_s": n / ts,
                     "iterations": sum(r.iterations for r in rb),
                     "max_abs_dE_batched_vs_sequential": max(abs(a.E - b.E) for a, b in zip(rb, rs))}
        log(f"{prec}: batched {n / tb:.2f} solves/s, sequential {n / ts:.2f}")
    print(json.dumps({
        "metric": METRIC, "value": out["mixed"]["batched_solves_per_s"], "unit": "SCF solves/s",
        "n_gpus": 1, "steps": n, "warmup": 0, "ms_per_step": 1e3 / out["mixed"]["batched_solves_per_s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64 (mixed: f32 exchange early)",
        "data": "synthetic",
        "config": {"workload": f"{n} random neutral atoms (seed 1004): Z 1..86, 4..24 LIP elements of order 10 or "
                               f"15, lmax = mmax 0..4, Hund spins; SCF to dE < 1e-8 Eh",
                   "parallelism": "single GPU, one lockstep engine"},
        "solves": out, "cpu_baseline": None}), flush=True)


def main_serve(args):
    """A Poisson stream of diatomic SCF requests served by one engine: host
    setup on a thread pool as requests arrive, admission by HBM footprint."""
    import helfem_amd as H
    from helfem_amd import batch
    n = args.systems or 16
    ctx = H.Context(0)
    specs = batch.random_diatomic_specs(n, seed=1005)
    res, wall = batch.serve(specs, args.rate, seed=1006, ctx=ctx, precision="mixed", conv=1e-8, setup_threads=16)
    lat = batch.latency_summary(res)
    log(f"{n} requests at {args.rate}/s: {n / wall:.2f} solves/s, latency p50 {lat['p50_s']:.2f}s p99 "
        f"{lat['p99_s']:.2f}s")
    print(json.dumps({
        "metric": METRIC, "value": n / wall, "unit": "SCF solves/s", "n_gpus": 1, "steps": n, "warmup": 0,
        "ms_per_step": 1e3 * wall / n, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (mixed: f32 exchange early)", "data": "synthetic",
        "config": {"workload": f"{n} diatomic requests (seed 1005), Poisson arrivals at {args.rate}/s: Z1, Z2 "
                               f"1..18, R 1-6 bohr, 4..12 mu-elements, lmax 8..24 for m 0..mmax (0..3), HF or "
                               f"LDA; SCF to dE < 1e-8 Eh",
                   "parallelism": "single GPU, continuous batching"},
        "latency_s": lat, "wall_s": wall,
        "requests": [{"spec": r["spec"], "latency_s": r["latency_s"], "iterations": r["iterations"],
                      "E": r["E"]} for r in res],
        "cpu_baseline": None}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is a "
                                   "design note, not a measured path)"}))
        sys.exit(2)
    if args.workload == "batch-atoms":
        return main_batch(args)
    if args.workload == "serve-diatomic":
        return main_serve(args)
    import torch
    import helfem_amd as H
    from helfem_amd import _lib
    from helfem_amd.synthetic import seeded_density

    cfg = WORKLOADS[args.workload]
    kind = cfg["kind"]
    knames, jnames = KERNELS[kind]
    ctx = H.Context(0)
    t = time.perf_counter()
    hb = make_basis(H, cfg, ctx)
    hb.compute_tei()
    log(f"basis Nbf={hb.Nbf} Nrad={hb.Nrad} Nang={hb.Nang}; host setup {time.perf_counter() - t:.1f}s")
    S = hb.overlap()
    blocks = hb.get_sym_idx(1)
    nocc = seeded_occupations(cfg, hb)
    P = seeded_density(S, blocks, nocc, seed=cfg["seed"])
    nelec = float(np.sum(P * S))
    log(f"seeded density: tr(PS) = {nelec:.6f}, occupied orbitals per m-block {nocc}")

    dev = torch.device("cuda:0")
    Pd = torch.as_tensor(P, device=dev)
    Pa = 0.5 * Pd
    J = torch.empty_like(Pd)
    K = torch.empty_like(Pd)
    ctx.bind_torch_stream()
    t = time.perf_counter()
    for _ in range(args.warmup):
        hb.coulomb(Pd, out=J)
        hb.exchange(Pa, out=K)
    torch.cuda.synchronize()
    log(f"warmup {args.warmup} steps {time.perf_counter() - t:.2f}s (includes device upload of the caches)")
    _lib.check(_lib.lib().hf_ctx_set_timing(ctx.handle, 1))
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record()
        hb.coulomb(Pd, out=J)
        hb.exchange(Pa, out=K)
        ends[i].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    lat = np.array([s.elapsed_time(e) for s, e in zip(starts, ends)])
    stats = {}
    for name in knames + jnames:
        out = np.zeros(3)
        _lib.check(_lib.lib().hf_ctx_kernel_stats(ctx.handle, name.encode(), out.ctypes.data))
        stats[name] = out
    _lib.check(_lib.lib().hf_ctx_set_timing(ctx.handle, 0))
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    k_ms = sum(stats[n][1] / max(stats[n][0], 1) for n in knames)
    k_flops = sum(stats[n][2] / max(stats[n][0], 1) for n in knames)
    achieved = k_flops / (k_ms * 1e-3) / 1e12 if k_ms > 0 else 0.0
    log(f"{args.steps} steps {ms_per_step:.3f} ms/step; exchange kernel {k_ms:.3f} ms, "
        f"{k_flops / 1e12:.4f} TFLOP/launch -> {achieved:.2f} TFLOP/s")

    if args.copy_calibration:
        nbytes = 1 << 30
        a = torch.empty(nbytes // 8, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copy calibration: 3 x {nbytes} bytes copied")

    Ecoul = 0.5 * float(torch.sum(Pd * J))
    Exx = float(torch.sum(Pa * K))
    log(f"Ecoul {Ecoul:.10f} Exx {Exx:.10f}")

    result = {
        "metric": METRIC,
        "value": args.steps / (t1 - t0),
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": cfg["desc"], "mode": "plain FP64 J+K (HF Fock build)", "Nbf": hb.Nbf,
                   "density": f"seeded SplitMix64 (seed {cfg['seed']}), P = 2CC' per m-block, "
                              f"tr(PS) = {nelec:.3f}",
                   "parallelism": "single GPU"},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "roofline": {"bound": "mfma", "kernel": {"atomic": "k_atomic_k_off_f64 + k_atomic_k_in_f64",
                                                 "diatomic": "k_dia_k2_off<double> + k_dia_kin_lds<double>"}[kind] +
                                                " (exchange: cross- and in-element launches of one build)",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": measured_traffic(args.workload),
                     "avg_launch_ms": k_ms, "flops_per_launch": k_flops},
        "kernels_ms": {k: (v[1] / max(v[0], 1)) for k, v in stats.items()},
        "kernels_tflops": {k: (v[2] / max(v[1], 1e-30) / 1e9) for k, v in stats.items() if v[2] > 0},
        "energies": {"coulomb": Ecoul, "exchange": Exx},
        "exchange_work": hb.work(),
    }
    if not args.no_scf:
        from helfem_amd.scf import rhf
        nel = int(round(nelec))
        scf = {}
        for prec in ("fp64", "mixed"):
            t = time.perf_counter()
            r = rhf(hb, nel, conv=1e-10, precision=prec)
            tsolve = time.perf_counter() - t
            log(f"SCF {prec}: E = {r.E:.10f} in {r.iterations} iterations ({r.iterations_fp32} with FP32 "
                f"exchange), {tsolve:.2f}s")
            scf[prec] = {"solves_per_s": 1.0 / tsolve, "seconds": tsolve, "iterations": r.iterations,
                         "iterations_fp32_exchange": r.iterations_fp32, "energy": r.E,
                         "phases_s": {k: round(v, 4) for k, v in r.phases.items()},
                         "eigensolves": r.eigensolves, "reused_eigensolves": r.reused_solves,
                         "skipped_eigensolves": r.skipped_solves}
        scf["mixed_vs_fp64"] = {"speedup": scf["mixed"]["solves_per_s"] / scf["fp64"]["solves_per_s"],
                                "energy_difference_Eh": scf["mixed"]["energy"] - scf["fp64"]["energy"]}
        scf["convergence"] = "dE < 1e-10 Eh and orbital gradient < 1e-6, core-Hamiltonian guess"
        scf["precision"] = ("fp64: FP64 throughout; mixed: FP32 exchange until the orbital gradient is below "
                            "1e-3, FP64 after (overlap, eigensolves and energies always FP64)")
        result["scf"] = scf
    if not args.no_cpu_baseline:
        stride = args.cpu_sample_stride or (24 if kind == "atomic" else 64)
        result["cpu_baseline"] = cpu_baseline(cfg, P, hb.Nang * hb.Nang, stride)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
