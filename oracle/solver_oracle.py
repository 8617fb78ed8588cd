"""CPU FP64 restatement of the batched solver behind Solver::solve.

TEST INFRASTRUCTURE ONLY (see capon_oracle.py): imported by ``tests/`` and bench.py's CPU
leg, never by what ships.  Paths are relative to the reference root.

  solve_batch  Gaussian elimination with partial pivoting, framework/beamformer/capon/
               RealTimeCapon/nvidia_solve.cu:2198-2312 (zfsolve_batch's kernel), pivot
               measure |re| + |im| (RealTimeCapon/operations.h:35,198-205), the systems
               column-major as Solver::solve takes them (RealTimeCapon/Solver.cpp:23-116)

Pinning (tests/test_solver_oracle.py): against LAPACK's solve (whose pivot search,
izamax, uses the same |re| + |im| measure), on systems that force a row exchange at every
step (a zero diagonal), on permutation matrices (exact), on a triangular system (no
elimination), and on the pivot choice itself where two rows tie or where the modulus and
|re| + |im| order two candidates differently.
"""
from __future__ import annotations

import numpy as np


def pivot_measure(a: np.ndarray) -> np.ndarray:
    """|re| + |im| (operations.h:198-205 with FAST_COMPLEX_ARITH, defined at :35)."""
    return np.abs(a.real) + np.abs(a.imag)


def solve_batch(A: np.ndarray, b: np.ndarray, return_pivots: bool = False):
    """x[s] solving A_s x = b_s for column-major A [batch][N*N] (A_s[i][j] = A[s][i + N j])
    and b [batch][N].  At step k the pivot is the first row i >= k of largest |re|+|im| in
    column k (nvidia_solve.cu:2256-2296), the row is exchanged, scaled by the pivot's
    reciprocal (:2298-2302) and eliminated below (:2304-2310); x from the unit upper
    triangle by back substitution (:2316-2325).  Vectorised over the batch in complex128."""
    A = np.asarray(A, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    nb, N = b.shape
    M = np.empty((nb, N, N + 1), dtype=np.complex128)
    M[:, :, :N] = A.reshape(nb, N, N).transpose(0, 2, 1)   # column-major storage
    M[:, :, N] = b
    rows = np.arange(nb)
    piv = np.empty((nb, N), dtype=np.int64)
    for k in range(N):
        p = k + np.argmax(pivot_measure(M[:, k:, k]), axis=1)   # argmax keeps the first maximum
        piv[:, k] = p
        rk = M[rows, k, :].copy()
        M[rows, k, :] = M[rows, p, :]
        M[rows, p, :] = rk
        M[:, k, k + 1:] /= M[:, k, k:k + 1]
        M[:, k, k] = 1.0
        M[:, k + 1:, k + 1:] -= M[:, k + 1:, k:k + 1] * M[:, k:k + 1, k + 1:]
        M[:, k + 1:, k] = 0.0
    x = M[:, :, N].copy()
    for j in range(N - 1, 0, -1):
        x[:, :j] -= M[:, :j, j] * x[:, j:j + 1]
    return (x, piv) if return_pivots else x
