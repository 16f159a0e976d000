"""Test-matrix generators for the solver workloads (own seeded generators).

* ``svd_matrix``: A = U diag(sigma) V^T with U, V Haar-distributed orthogonal
  (QR of a Gaussian matrix with the signs of R's diagonal moved into Q) and
  sigma geometric from 1 to 1/kappa, so ||A||_2 = 1 and kappa_2(A) = kappa.
* ``rhs_from_solution``: x_true ~ U[-1, 1), b = A x_true, so forward errors
  can be measured against x_true.

Host versions use numpy; ``svd_matrix_device`` builds the same kind of
matrix on the GPU (torch's Gaussian and QR for U and V, the library's own
DGEMM for the products) for sizes where the host would be slow.
"""
from __future__ import annotations

import numpy as np


def haar(n: int, rng: np.random.Generator) -> np.ndarray:
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    return q * np.sign(np.diag(r))


def sigma(n: int, kappa: float) -> np.ndarray:
    if n == 1:
        return np.ones(1)
    return kappa ** (-np.arange(n) / (n - 1))


def svd_matrix(n: int, kappa: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    u, v = haar(n, rng), haar(n, rng)
    return np.asfortranarray((u * sigma(n, kappa)) @ v.T)


def rhs_from_solution(a: np.ndarray, nrhs: int, seed: int):
    rng = np.random.default_rng(seed + 7919)
    x = np.asfortranarray(rng.uniform(-1.0, 1.0, (a.shape[1], nrhs)))
    return np.asfortranarray(a @ x), x


def svd_matrix_device(n: int, kappas, seed: int):
    """Device matrices A_kappa = U diag(sigma_kappa) V^T for each kappa (same
    U, V), as flens_amd.DeviceMatrix objects, plus x_true and b = A x_true per
    kappa (nrhs = 1)."""
    import torch
    import flens_amd as F
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)

    def haar_t():
        z = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
        q, r = torch.linalg.qr(z)
        return q * torch.sign(torch.diagonal(r))

    u, v = haar_t(), haar_t()             # row-major n x n tensors
    # column-major device views: storage of a column-major M is M^T row-major
    U = F.DeviceMatrix(0, n, n, n, u.T.contiguous())
    U.ptr = U.storage.data_ptr()
    Vt = F.DeviceMatrix(0, n, n, n, v.contiguous())       # V^T column-major = V row-major
    Vt.ptr = Vt.storage.data_ptr()
    del u, v
    xt = torch.empty(1, n, dtype=torch.float64, device="cuda").uniform_(-1.0, 1.0, generator=g)
    out = []
    for kappa in kappas:
        s = torch.from_numpy(sigma(n, kappa)).to("cuda")
        US = F.DeviceMatrix(0, n, n, n, U.storage * s[:, None])    # column j scaled by s_j
        US.ptr = US.storage.data_ptr()
        A = F.DeviceMatrix.empty(n, n)
        A.storage.zero_()                  # beta = 0 still reads C (0 * C)
        F.blas.mm(0, 0, 1.0, US, Vt, 0.0, A)
        del US
        X = F.DeviceMatrix(0, n, 1, n, xt.clone())
        X.ptr = X.storage.data_ptr()
        B = F.DeviceMatrix.empty(n, 1)
        B.storage.zero_()
        F.blas.mm(0, 0, 1.0, A, X, 0.0, B)
        out.append((kappa, A, X, B))
    return out
