"""frehg_ls_solve_noalloc (the allocation-free variant of the LinearSolver seam) against
frehg_ls_solve on the same operators: device arrays in, solution out, two calls in a row on one
GPU with the second operator smaller than the first (the scratch array keeps the first call's
contents), in 3D and 2D; and the size limit and the FP64-CG-only scope rejected with an error."""
import numpy as np
import pytest
import torch

from frehg_amd import abi, gpu

pytestmark = pytest.mark.gpu


def spd_stencil(ndim, nx, ny, nz, seed):
    """A random symmetric positive definite stencil operator in the reference's owned ordering
    (3D rows z-fastest, 2D rows x-fastest): face couplings -a with a > 0 random, diagonal = the sum
    of its couplings plus a random storage term."""
    rng = np.random.default_rng(seed)
    nzz = nz if ndim == 3 else 1
    shape = (ny, nx, nzz) if ndim == 3 else (ny, nx)  # owned order: index = (j * nx + i) * nz + k
    axes = [1, 0, 2] if ndim == 3 else [1, 0]  # x, y, (z) axis in `shape`
    diag = 0.1 + rng.random(shape)
    offs = []
    for ax in axes:
        a = 0.2 + rng.random(shape)
        sl_last = [slice(None)] * len(shape)
        sl_last[ax] = slice(-1, None)
        a[tuple(sl_last)] = 0.0  # no +face on the domain's far edge
        minus = np.zeros(shape)
        src = [slice(None)] * len(shape)
        dst = [slice(None)] * len(shape)
        src[ax] = slice(0, -1)
        dst[ax] = slice(1, None)
        minus[tuple(dst)] = a[tuple(src)]  # the -face of a row is its neighbour's +face
        diag += a + minus
        offs += [-minus.ravel(), -a.ravel()]
    b = rng.standard_normal(diag.size)
    return diag.ravel(), offs, b


def _dev(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")


def _solve_noalloc(ctx, ndim, nx, ny, nz, diag, offs, b, cfg):
    d, o, bb = _dev(diag), [_dev(v) for v in offs], _dev(b)
    x = torch.full_like(bb, float("nan"))  # the call must write every entry
    torch.cuda.synchronize()
    st = ctx.ls_solve_noalloc(ndim, nx, ny, nz, d.data_ptr(), [v.data_ptr() for v in o], bb.data_ptr(),
                              x.data_ptr(), cfg)
    return x.cpu().numpy(), st


def test_noalloc_matches_ls_solve_on_consecutive_calls():
    ctx = gpu.Context(0)
    cfg = abi.solver_config(1e-12)
    cases = [(3, 96, 70, 40, 1), (3, 21, 13, 7, 2), (2, 150, 97, 1, 3), (2, 9, 8, 1, 4)]
    refs = []
    for ndim, nx, ny, nz, seed in cases:
        diag, offs, b = spd_stencil(ndim, nx, ny, nz, seed)
        x, st = ctx.ls_solve(ndim, nx, ny, nz, diag, offs, b, cfg)
        assert st.converged
        refs.append((diag, offs, b, x, st))
    # consecutive allocation-free calls, each smaller than the one before within a dimension
    for (ndim, nx, ny, nz, _), (diag, offs, b, xr, sr) in zip(cases, refs):
        print(f"no-alloc solve {ndim}D {nx}x{ny}x{nz}", flush=True)
        x, st = _solve_noalloc(ctx, ndim, nx, ny, nz, diag, offs, b, cfg)
        assert st.converged and st.residual_norm <= 1e-12 * st.rhs_norm * (1 + 1e-9)
        assert abs(st.iterations - sr.iterations) <= 1
        assert st.rhs_norm == pytest.approx(sr.rhs_norm, rel=1e-12)
        np.testing.assert_allclose(x, xr, rtol=0, atol=1e-9 * np.abs(xr).max())


def test_noalloc_rejects_what_it_does_not_serve():
    ctx = gpu.Context(0)
    diag, offs, b = spd_stencil(3, 8, 7, 6, 5)
    # mixed precision and BiCGSTAB need buffers of their own
    for cfg in (abi.solver_config(1e-10, precision=abi.PREC_MIXED),
                abi.solver_config(1e-10, ksp_type=abi.KSP_BICGSTAB)):
        with pytest.raises(gpu.FrehgError) as e:
            _solve_noalloc(ctx, 3, 8, 7, 6, diag, offs, b, cfg)
        assert e.value.code == abi.ERR_INVALID_ARG
    # an operator larger than the scratch array (checked before any device work)
    one = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    p = one.data_ptr()
    with pytest.raises(gpu.FrehgError) as e:
        ctx.ls_solve_noalloc(3, 512, 512, 30, p, [p] * 6, p, p, abi.solver_config(1e-10))
    assert e.value.code == abi.ERR_INVALID_ARG and "scratch" in str(e.value)
    # a nonsymmetric operator
    offs[1] = offs[1] * 1.5
    with pytest.raises(gpu.FrehgError) as e:
        _solve_noalloc(ctx, 3, 8, 7, 6, diag, offs, b, abi.solver_config(1e-10))
    assert "symmetric" in str(e.value)
