"""Seeded synthetic inputs for parity tests and the benchmark.

RNG: SplitMix64 (Steele, Lea, Flood 2014), uniforms from the top 53 bits.
Seeded density: X (Nbf x nocc) ~ U[-1,1) per symmetry block, orthonormalised
in the S metric (C = S^-1/2 Q with Q from a QR of S^1/2-free Gram-Schmidt in
the orthonormal basis), P = 2 C C' summed over blocks.  tr(P S) = 2*nocc.
"""
from __future__ import annotations

import numpy as np

_M64 = (1 << 64) - 1


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n raw 64-bit outputs of SplitMix64 started from `seed`."""
    x = np.uint64(seed & _M64)
    out = np.empty(n, dtype=np.uint64)
    gamma = np.uint64(0x9E3779B97F4A7C15)
    m1 = np.uint64(0xBF58476D1CE4E5B9)
    m2 = np.uint64(0x94D049BB133111EB)
    with np.errstate(over="ignore"):
        for i in range(n):
            x = x + gamma
            z = x
            z = (z ^ (z >> np.uint64(30))) * m1
            z = (z ^ (z >> np.uint64(27))) * m2
            out[i] = z ^ (z >> np.uint64(31))
    return out


def splitmix64_fast(seed: int, n: int) -> np.ndarray:
    """Vectorised SplitMix64 (same stream as splitmix64)."""
    gamma = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        x = np.uint64(seed & _M64) + gamma * np.arange(1, n + 1, dtype=np.uint64)
        z = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int, lo=-1.0, hi=1.0) -> np.ndarray:
    """Uniforms in [lo, hi) from the top 53 bits of SplitMix64."""
    u = (splitmix64_fast(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def seeded_density(S: np.ndarray, blocks, nocc_per_block, seed: int) -> np.ndarray:
    """P = sum_k 2 C_k C_k' with C_k (|block| x nocc_k) orthonormal in S."""
    n = S.shape[0]
    P = np.zeros((n, n))
    off = 0
    for b, nocc in zip(blocks, nocc_per_block):
        if nocc == 0:
            continue
        b = np.asarray(b)
        Sb = S[np.ix_(b, b)]
        X = uniform(seed + off, b.size * nocc).reshape((b.size, nocc), order="F")
        off += 7919
        # Cholesky-based S-orthonormalisation: C = X (X' S X)^-1/2
        G = X.T @ Sb @ X
        e, v = np.linalg.eigh(G)
        C = X @ (v / np.sqrt(e)) @ v.T
        P[np.ix_(b, b)] += 2.0 * C @ C.T
    return P


def random_symmetric(n: int, seed: int) -> np.ndarray:
    """Dense symmetric matrix with U[-1,1) entries (exercises every coupling)."""
    A = uniform(seed, n * n).reshape((n, n))
    return 0.5 * (A + A.T)
