"""Pins of the batched-solver oracle (oracle/solver_oracle.py) on the CPU."""
import numpy as np

from oracle import solver_oracle as S


def colmajor(A):
    """[batch][N][N] row-indexed matrices as Solver::solve's column-major [batch][N*N]."""
    return np.ascontiguousarray(np.swapaxes(A, 1, 2)).reshape(A.shape[0], -1)


def test_matches_lapack():
    rng = np.random.default_rng(1)
    for N in (1, 2, 3, 10, 33, 72):
        A = rng.standard_normal((4, N, N)) + 1j * rng.standard_normal((4, N, N))
        b = rng.standard_normal((4, N)) + 1j * rng.standard_normal((4, N))
        x = S.solve_batch(colmajor(A), b)
        ref = np.linalg.solve(A, b[..., None])[..., 0]
        assert np.abs(x - ref).max() <= 1e-9 * np.abs(ref).max(), N


def test_zero_diagonal_forces_exchanges():
    # a cyclic shift with a zero diagonal: every step must exchange rows
    N = 9
    rng = np.random.default_rng(2)
    P = np.roll(np.eye(N), 1, axis=0) * (2 + rng.uniform(size=N))[:, None]
    A = P[None].astype(complex)
    b = rng.standard_normal((1, N)) + 0j
    x, piv = S.solve_batch(colmajor(A), b, return_pivots=True)
    # column k's only candidate is the row that held element k + 1 of the shift
    np.testing.assert_array_equal(piv[0, :N - 1], np.arange(1, N))
    assert np.abs(A[0] @ x[0] - b[0]).max() <= 1e-12


def test_permutation_is_exact():
    N = 12
    perm = np.random.default_rng(3).permutation(N)
    A = np.eye(N)[perm][None].astype(complex)
    b = (np.arange(N) + 1j * np.arange(N)[::-1])[None].astype(complex)
    x = S.solve_batch(colmajor(A), b)
    np.testing.assert_array_equal(A[0] @ x[0], b[0])
    np.testing.assert_array_equal(x[0][perm], b[0])


def test_upper_triangular_needs_no_elimination():
    # a diagonally dominant upper triangle keeps every pivot on the diagonal
    N = 6
    U = np.triu(np.ones((N, N)) * (1 + 1j)) + 5 * np.eye(N)
    b = np.ones((1, N), complex)
    x, piv = S.solve_batch(colmajor(U[None]), b, return_pivots=True)
    np.testing.assert_array_equal(piv[0], np.arange(N))
    assert np.abs(U @ x[0] - b[0]).max() <= 1e-13


def test_pivot_measure_and_first_of_equals():
    # column 0 holds 3 (|.| = 3, |re|+|im| = 3) and 2+2j (|.| = 2.83, |re|+|im| = 4):
    # |re|+|im| picks row 1 where the modulus would pick row 0
    A = np.array([[3, 1], [2 + 2j, 5]], complex)
    _, piv = S.solve_batch(colmajor(A[None]), np.ones((1, 2), complex), return_pivots=True)
    assert piv[0, 0] == 1
    # two rows of equal measure: the first one wins (strict comparison in a forward scan)
    A = np.array([[1, 0, 0], [1j, 1, 0], [-1, 0, 1]], complex)
    _, piv = S.solve_batch(colmajor(A[None]), np.ones((1, 3), complex), return_pivots=True)
    assert piv[0, 0] == 0
