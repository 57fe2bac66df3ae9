"""hf_eigh_batched (rocSOLVER's batched dsyevd) against torch.linalg.eigh
and the eigen-equation itself, on the stack shapes the SCF drivers use:
eigenvalues within 1e-12 of the largest |eigenvalue|, residual
|A V - V diag(w)| and orthonormality |V'V - I| within 1e-12 (times |A|)."""
import numpy as np
import pytest

import helfem_amd as H
from helfem_amd import linalg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("B,N", [(1, 1), (3, 2), (7, 53), (5, 35), (3, 200), (2, 335)])
def test_batched_eigh_matches_torch(B, N, gpu_ctx):
    import torch
    print(f"[gpu] batched eigensolver B={B} N={N}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(17 + N)
    A = torch.randn(B, N, N, dtype=torch.float64, device="cuda", generator=g)
    A = A + A.transpose(1, 2)
    w, V, info = linalg.eigh(A, gpu_ctx)
    we, _ = torch.linalg.eigh(A)
    torch.cuda.synchronize()
    assert not linalg.eigh_failed([info])
    scale = float(we.abs().max())
    assert float((w - we).abs().max()) <= 1e-12 * scale
    nA = float(torch.linalg.matrix_norm(A, ord=2).max())
    res = A @ V - V * w[:, None, :]
    assert float(res.abs().max()) <= 1e-12 * nA
    eye = torch.eye(N, dtype=torch.float64, device="cuda")
    assert float((V.transpose(1, 2) @ V - eye).abs().max()) <= 1e-12
    # the input is left as it was
    assert torch.equal(A, A.transpose(1, 2))
