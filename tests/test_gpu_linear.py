"""LinearSolver seam (frehg_ls_solve): manufactured-solution KATs of the reference
(tests/linear/test_petsc_linear_solver.cpp:25-101, test_assembly_backend_independent.cpp:21-80),
their 7-point analogue, an oracle comparison and the error paths."""
import numpy as np
import pytest

import oracle_lib as ol
from frehg_amd import abi, gpu
from test_oracle_linear import dense, indefinite_surface_operator, laplacian, matvec, random_stencil

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ndim,dims", [(2, (6, 5, 1)), (3, (5, 4, 6)), (3, (70, 33, 9))])
def test_manufactured(ndim, dims):
    nx, ny, nz = dims
    diag, offs, cols = laplacian(ndim, nx, ny, nz)
    n = len(diag)
    ctx = gpu.Context(0)
    for xs in (1.5 + np.sin(0.1 * np.arange(n)), 0.5 + 0.001 * np.arange(n)):
        b = matvec(diag, offs, cols, xs)
        x, st = ctx.ls_solve(ndim, nx, ny, nz, diag, offs, b, abi.solver_config(1e-13))
        assert st.converged and np.linalg.norm(x - xs) < 1e-9 * max(1.0, np.sqrt(n / 30))
        xo, _ = ol.ls_solve(ndim, nx, ny, nz, diag, offs, b, abi.solver_config(1e-13))
        assert np.abs(x - xo).max() < 1e-10


@pytest.mark.parametrize("ndim,dims", [(2, (6, 5, 1)), (3, (70, 33, 9))])
def test_zero_rhs_ends_before_the_first_update(ndim, dims):
    """b = 0: the solve ends at its start (0 iterations) and x is exactly zero. In 3D without fixed
    rows the start kernel does not clear x (the first update writes it), so this is the path where
    the solver clears x after a solve that had no update."""
    nx, ny, nz = dims
    diag, offs, cols = laplacian(ndim, nx, ny, nz)
    n = len(diag)
    ctx = gpu.Context(0)
    xs = 1.5 + np.sin(0.1 * np.arange(n))
    x, st = ctx.ls_solve(ndim, nx, ny, nz, diag, offs, matvec(diag, offs, cols, xs), abi.solver_config(1e-13))
    assert st.converged
    x, st = ctx.ls_solve(ndim, nx, ny, nz, diag, offs, np.zeros(n), abi.solver_config(1e-13))
    assert st.converged and st.iterations == 0
    assert np.array_equal(x, np.zeros(n))


def test_decoupled_rows_2d():
    """Rows with no off-diagonals (dry cells at rest) are solved exactly and folded out."""
    nx, ny = 7, 6
    diag, offs, cols = laplacian(2, nx, ny)
    for d in range(4):
        offs[d] = offs[d].copy()
    row = 2 * nx + 3
    for d in range(4):
        offs[d][row] = 0.0
    diag[row] = 6400.0
    b = np.ones(nx * ny)
    b[row] = 6400.0 * 0.37
    ctx = gpu.Context(0)
    x, st = ctx.ls_solve(2, nx, ny, 1, diag, offs, b, abi.solver_config(1e-13))
    xo, _ = ol.ls_solve(2, nx, ny, 1, diag, offs, b, abi.solver_config(1e-13))
    assert x[row] == pytest.approx(0.37, rel=1e-15)
    assert np.abs(x - xo).max() < 1e-10


def test_errors_raise():
    ctx = gpu.Context(0)
    nx, ny, nz = 6, 6, 6
    diag, offs, cols = laplacian(3, nx, ny, nz)
    b = np.random.default_rng(0).standard_normal(nx * ny * nz)
    with pytest.raises(gpu.FrehgError) as e:
        ctx.ls_solve(3, nx, ny, nz, diag, offs, b, abi.solver_config(1e-14, max_it=2))
    assert e.value.code == abi.ERR_NOT_CONVERGED
    offs[1] = offs[1] * 2.0  # x+ couplings no longer mirror the x- ones
    with pytest.raises(gpu.FrehgError) as e:
        ctx.ls_solve(3, nx, ny, nz, diag, offs, b, abi.solver_config(1e-10))
    assert e.value.code == abi.ERR_INVALID_ARG


BCGS = dict(ksp_type=abi.KSP_BICGSTAB)


@pytest.mark.parametrize("ndim,dims", [(2, (7, 5, 1)), (3, (4, 3, 5))])
def test_bicgstab_nonsymmetric_dense(ndim, dims):
    nx, ny, nz = dims
    diag, offs = random_stencil(ndim, nx, ny, nz, seed=3)
    A = dense(ndim, nx, ny, nz, diag, offs)
    b = np.random.default_rng(4).standard_normal(len(diag))
    x, st = gpu.Context(0).ls_solve(ndim, nx, ny, nz, diag, offs, b, abi.solver_config(1e-13, **BCGS))
    assert st.converged
    np.testing.assert_allclose(x, np.linalg.solve(A, b), atol=1e-10)


def test_bicgstab_nonsymmetric_large_vs_oracle():
    nx, ny, nz = 70, 33, 9
    diag, offs = random_stencil(3, nx, ny, nz, seed=11, dominance=1.02)
    xs = 1.0 + np.cos(0.01 * np.arange(len(diag)))
    # too large for a dense check: manufactured solution b = A xs in the owned ordering
    b = diag * xs
    idx = np.arange(len(diag)).reshape(ny, nx, nz)
    steps = {0: (0, -1, 0), 1: (0, 1, 0), 2: (-1, 0, 0), 3: (1, 0, 0), 4: (0, 0, -1), 5: (0, 0, 1)}
    for d, (dj, di, dk) in steps.items():
        src = np.roll(idx, shift=(-dj, -di, -dk), axis=(0, 1, 2))
        b += offs[d] * xs[src.ravel()]  # off-domain couplings are zero, so the wrap-around is inert
    cfg = abi.solver_config(1e-13, **BCGS)
    x, st = gpu.Context(0).ls_solve(3, nx, ny, nz, diag, offs, b, cfg)
    xo, sto = ol.ls_solve(3, nx, ny, nz, diag, offs, b, cfg)
    assert st.converged and sto.converged
    assert np.abs(x - xs).max() < 1e-9 and np.abs(x - xo).max() < 1e-10


def test_bicgstab_indefinite_surface_row():
    """The indefinite surface system of a wetting front: BiCGSTAB solves it on the GPU."""
    nx, ny = 9, 7
    diag, offs = indefinite_surface_operator(nx, ny)
    A = dense(2, nx, ny, 1, diag, offs)
    b = np.random.default_rng(6).uniform(0.5, 1.5, nx * ny)
    x, st = gpu.Context(0).ls_solve(2, nx, ny, 1, diag, offs, b, abi.solver_config(1e-13, **BCGS))
    np.testing.assert_allclose(x, np.linalg.solve(A, b), atol=1e-9)


def test_bad_solver_config_rejected():
    diag, offs, _ = laplacian(2, 4, 4)
    b = np.ones(16)
    for kw in (dict(ksp_type=7), dict(pc_type=3), dict(max_it=0)):
        with pytest.raises(gpu.FrehgError) as e:
            gpu.Context(0).ls_solve(2, 4, 4, 1, diag, offs, b, abi.solver_config(1e-10, **kw))
        assert e.value.code == abi.ERR_INVALID_ARG
