"""Batched symmetric eigensolver on the GPU (hf_eigh_batched: rocSOLVER's
divide-and-conquer dsyevd over a strided batch, on the context's stream).

The SCF drivers diagonalise many small Fock blocks per iteration
(eig_gsym / eig_gsym_sub, src/general/scf_helpers.cpp:41-80, one host
eigensolve per block in the reference) as padded stacks.  The call is
asynchronous: the solver status rides along with the next host transfer of
the eigenvalues instead of a synchronisation per call (as torch.linalg.eigh
makes to check its errors).
"""
from __future__ import annotations

from ._lib import check, lib


def eigh(A, ctx):
    """Eigenvalues (ascending) and eigenvectors of a (B, N, N) float64 CUDA
    stack of symmetric matrices, like torch.linalg.eigh (upper triangle read),
    plus the solver's per-matrix status (int32 CUDA tensor, 0 = converged;
    check it with the next host transfer: eigh_failed)."""
    import torch
    if A.dim() != 3 or A.shape[-1] != A.shape[-2] or A.dtype != torch.float64 or not A.is_cuda:
        raise ValueError("eigh expects a (B, N, N) float64 CUDA tensor")
    B, N, _ = A.shape
    V = A.contiguous().clone()  # overwritten with the eigenvectors (columns of the column-major copy)
    W = torch.empty((B, N), dtype=torch.float64, device=A.device)
    E = torch.empty_like(W)
    info = torch.empty(B, dtype=torch.int32, device=A.device)
    ctx.bind_torch_stream()
    check(lib().hf_eigh_batched(ctx.handle, N, V.data_ptr(), N, N * N, W.data_ptr(), E.data_ptr(), info.data_ptr(),
                                B))
    return W, V.transpose(1, 2), info


def eigh_failed(infos) -> bool:
    """True if any solver status in the list of info tensors is nonzero (one
    device-to-host transfer)."""
    import torch
    if not infos:
        return False
    return bool(torch.cat([i.reshape(-1) for i in infos]).ne(0).any().item())
