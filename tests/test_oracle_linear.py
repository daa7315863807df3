"""Pins the oracle's Jacobi-PCG (oracle/krylov.cpp) with the reference's manufactured-solution
solves (tests/linear/test_petsc_linear_solver.cpp:25-101, test_assembly_backend_independent.cpp:21-80),
their 7-point analogue, and dense brute force."""
import numpy as np
import pytest

import oracle_lib as ol
from frehg_amd import abi


def laplacian(ndim, nx, ny, nz=1, diag_val=None):
    """Dirichlet Laplacian in the reference's owned ordering (diag 2*ndim, neighbours -1)."""
    n = nx * ny * (nz if ndim == 3 else 1)
    diag = np.full(n, float(diag_val if diag_val is not None else 2 * ndim))
    offs = [np.zeros(n) for _ in range(6 if ndim == 3 else 4)]
    cols = []
    for row in range(n):
        if ndim == 3:
            k = row % nz
            i = (row // nz) % nx
            j = row // (nz * nx)
        else:
            i, j, k = row % nx, row // nx, 0
        nb = []
        for d, (di, dj, dk) in enumerate([(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)][: len(offs)]):
            ii, jj, kk = i + di, j + dj, k + dk
            if 0 <= ii < nx and 0 <= jj < ny and (ndim == 2 or 0 <= kk < nz):
                offs[d][row] = -1.0
                col = ((jj * nx + ii) * nz + kk) if ndim == 3 else jj * nx + ii
                nb.append(col)
        cols.append(nb)
    return diag, offs, cols


def matvec(diag, offs, cols, x):
    y = diag * x
    for row, nb in enumerate(cols):
        y[row] -= sum(x[c] for c in nb)
    return y


@pytest.mark.parametrize("xexact", [lambda r: 1.5 + np.sin(0.1 * r), lambda r: 0.5 + 0.001 * r])
def test_5pt_manufactured(xexact):
    nx, ny = 6, 5
    diag, offs, cols = laplacian(2, nx, ny)
    xs = xexact(np.arange(nx * ny, dtype=float))
    b = matvec(diag, offs, cols, xs)
    x, st = ol.ls_solve(2, nx, ny, 1, diag, offs, b, abi.solver_config(1e-13))
    assert st.converged and np.linalg.norm(x - xs) < 1e-9


def test_7pt_manufactured_and_dense():
    nx, ny, nz = 5, 4, 6
    diag, offs, cols = laplacian(3, nx, ny, nz)
    n = nx * ny * nz
    xs = 1.5 + np.sin(0.1 * np.arange(n))
    b = matvec(diag, offs, cols, xs)
    x, st = ol.ls_solve(3, nx, ny, nz, diag, offs, b, abi.solver_config(1e-13))
    assert st.converged and np.linalg.norm(x - xs) < 1e-9
    A = np.diag(diag)
    for row, nb in enumerate(cols):
        for c in nb:
            A[row, c] = -1.0
    np.testing.assert_allclose(x, np.linalg.solve(A, b), atol=1e-10)


def test_pcg_iteration_count_jacobi_is_cg_on_scaled_system():
    # Jacobi-PCG on a constant-diagonal operator is plain CG: CG on an n-dim SPD system with k
    # distinct eigenvalues terminates in <= k iterations; a 1D Laplacian chain of 8 has 8.
    nx, ny = 8, 1
    diag, offs, cols = laplacian(2, nx, ny, diag_val=2.0)
    b = np.ones(nx)
    x, st = ol.ls_solve(2, nx, ny, 1, diag, offs, b, abi.solver_config(1e-14))
    assert st.iterations <= 8
    A = np.diag(diag) - np.eye(nx, k=1) - np.eye(nx, k=-1)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), atol=1e-12)


def test_max_it_exhaustion_raises():
    # KSP divergence / max_it exhaustion -> error (PetscLinearSolver.cpp:206-211)
    nx, ny, nz = 6, 6, 6
    diag, offs, cols = laplacian(3, nx, ny, nz)
    b = np.random.default_rng(0).standard_normal(nx * ny * nz)
    with pytest.raises(RuntimeError):
        ol.ls_solve(3, nx, ny, nz, diag, offs, b, abi.solver_config(1e-14, max_it=2))
