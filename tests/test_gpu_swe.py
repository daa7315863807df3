"""Shallow-water step: HIP path (through the C ABI) vs the oracle and the reference's golden
vector benchmarks/b1-sw/frehg2_np1_depth_reference.txt."""
import numpy as np
import pytest

import oracle_lib as ol
from cases import flow_2d_inputs, b1_inputs, gaussian_bump_bed, interior2, interp_series
from frehg_amd import abi, gpu
from gpu_common import compare_swe, say, swe_copy_state

pytestmark = pytest.mark.gpu


def test_b1_sw_golden_vector_on_gpu():
    say("SWE b1-sw golden on GPU")
    bath, rain, ref, offset, p = b1_inputs()
    ctx = gpu.Context(0)
    ctx.swe_init(1, 10, 80.0, 80.0, p)
    ctx.swe_set_bathymetry(bath + offset)
    ctx.swe_initialize_state(-2.0)
    t, worst = 0.0, 0.0
    for step in range(1, 3601):
        t += 5.0
        ctx.swe_step(interp_series(rain[:, 0], rain[:, 1], t), 0.0)
        if step % 360 == 0:
            row = ref[step // 360 - 1]
            d = interior2(ctx.swe_get("dept"), 1, 10)[:, 0]
            worst = max(worst, np.linalg.norm(d - row[1:]) / np.linalg.norm(row[1:]))
    say(f"  worst rel-L2 {worst:.3e}")
    assert worst < 1e-10


def test_init_matches_oracle():
    nx, ny = 13, 9
    rng = np.random.default_rng(2)
    bed = 0.3 * rng.random((ny, nx)) + 0.05 * np.arange(nx)[None, :]
    p = abi.swe_params(manning=0.03, dt=0.5, offset=0.1)
    orc = ol.OracleSwe(nx, ny, 1.0, 1.0, p)
    orc.set_bathymetry(bed.ravel())
    orc.initialize_state(0.25)
    ctx = gpu.Context(0)
    ctx.swe_init(nx, ny, 1.0, 1.0, p)
    ctx.swe_set_bathymetry(bed.ravel())
    ctx.swe_initialize_state(0.25)
    compare_swe(ctx, orc, rtol=0, atol=0, label="init")


@pytest.mark.parametrize("visc,ksp", [(0.0, abi.KSP_CG), (1e-3, abi.KSP_CG), (1e-3, abi.KSP_BICGSTAB)])
def test_wet_dry_rain_2d(visc, ksp):
    """2D wetting/drying with rain, viscosity and every limiter branch active."""
    nx, ny = 24, 17
    say(f"SWE 2D wet/dry visc={visc} ksp={ksp}")
    rng = np.random.default_rng(4)
    x = np.arange(nx)[None, :]
    bed = 0.02 * x + 0.05 * rng.standard_normal((ny, nx))
    p = abi.swe_params(manning=0.03, dt=0.5, visc_x=visc, visc_y=visc, hE=0.001)
    p.solver.ksp_type = ksp
    orc = ol.OracleSwe(nx, ny, 2.0, 1.5, p)
    orc.set_bathymetry(bed.ravel())
    orc.initialize_state(0.2)
    ctx = gpu.Context(0)
    ctx.swe_init(nx, ny, 2.0, 1.5, p)
    ctx.swe_set_bathymetry(bed.ravel())
    ctx.swe_initialize_state(0.2)
    for s in range(150):
        rain = 2e-4 if s < 80 else 0.0
        orc.step(rain, 1e-6)
        ctx.swe_step(rain, 1e-6)
        if s % 50 == 0 or s == 149:
            compare_swe(ctx, orc, rtol=1e-9, atol=1e-11, label=f"step {s}")
    d = interior2(ctx.swe_get("dept"), nx, ny)
    assert (d == 0).any() and (d > 0).any(), "case must contain wet and dry cells"


@pytest.mark.parametrize("transposed", [0, 1])
def test_flow_2d_x_terms_match_oracle(transposed):
    """The rain-free 2D wet/dry flow whose transpose symmetry pins the oracle's x-direction terms
    (a mound of water running over dry ground, dx != dy, velocities near 1 m/s): every SWE field
    of the HIP path against the oracle, on the problem and on its transpose."""
    bed, eta0 = flow_2d_inputs()
    dx, dy, vx, vy = 1.0, 1.5, 1e-2, 3e-2
    if transposed:
        bed, eta0, dx, dy, vx, vy = bed.T.copy(), eta0.T.copy(), dy, dx, vy, vx
    ny, nx = bed.shape
    say(f"SWE 2D flow {nx}x{ny} transposed={transposed}")
    p = abi.swe_params(manning=0.03, dt=0.2, visc_x=vx, visc_y=vy, solver=abi.solver_config(1e-13))
    orc = ol.OracleSwe(nx, ny, dx, dy, p)
    orc.set_bathymetry(bed.ravel())
    orc.set_initial_eta(eta0.ravel())
    ctx = gpu.Context(0)
    ctx.swe_init(nx, ny, dx, dy, p)
    ctx.swe_set_bathymetry(bed.ravel())
    ctx.swe_set_initial_eta(eta0.ravel())
    compare_swe(ctx, orc, rtol=1e-14, atol=1e-16, label="initial eta")
    umax = 0.0
    for s in range(60):
        orc.step(0.0, 0.0)
        ctx.swe_step(0.0, 0.0)
        umax = max(umax, np.abs(ctx.swe_get("uu")).max(), np.abs(ctx.swe_get("vv")).max())
        if s % 20 == 0 or s == 59:
            compare_swe(ctx, orc, rtol=1e-8, atol=1e-10, label=f"step {s}")
    assert umax > 0.3
    d = interior2(ctx.swe_get("dept"), nx, ny)
    assert (d == 0).any() and (d > 0).any()


def test_b0_lake_at_rest():
    p = abi.swe_params(manning=0.02, dt=1.0, solver=abi.solver_config(1e-13))
    ctx = gpu.Context(0)
    ctx.swe_init(20, 20, 1.0, 1.0, p)
    ctx.swe_set_bathymetry(gaussian_bump_bed().ravel())
    ctx.swe_initialize_state(1.0)
    for _ in range(200):
        ctx.swe_step(0.0, 0.0)
    assert np.abs(interior2(ctx.swe_get("eta"), 20, 20) - 1.0).max() < 1e-8
    assert np.abs(interior2(ctx.swe_get("uu"), 20, 20)).max() < 1e-8


def test_max_cfl():
    nx, ny = 10, 8
    p = abi.swe_params(dt=0.5)
    orc = ol.OracleSwe(nx, ny, 1.0, 1.0, p)
    orc.set_bathymetry(np.linspace(0, 0.3, nx * ny))
    orc.initialize_state(0.2)
    ctx = gpu.Context(0)
    ctx.swe_init(nx, ny, 1.0, 1.0, p)
    swe_copy_state(orc, ctx)
    for _ in range(5):
        orc.step(1e-4)
        ctx.swe_step(1e-4)
    assert ctx.swe_max_cfl() == pytest.approx(orc.max_cfl(), rel=1e-10)


def test_step_graphs_match_eager_launches():
    """The surface step replayed from captured graphs (iteration budget, re-capture when the
    forcing changes, eager continuation past the budget) gives the same fields, bit for bit, as
    eager launches; the replays alternate with the host's own synchronisations and reads."""
    import os
    nx, ny = 40, 23
    rng = np.random.default_rng(6)
    bed = 0.02 * np.arange(nx)[None, :] + 0.04 * rng.standard_normal((ny, nx))
    p = abi.swe_params(manning=0.03, dt=0.5, visc_x=1e-3, visc_y=1e-3, hE=0.001)
    ctxs = {}
    for mode in ("0", "1"):
        os.environ["FREHG_SW_GRAPH"] = mode
        try:
            c = gpu.Context(0)
            c.swe_init(nx, ny, 1.0, 1.0, p)
        finally:
            del os.environ["FREHG_SW_GRAPH"]
        c.swe_set_bathymetry(bed.ravel())
        c.swe_initialize_state(0.18)
        ctxs[mode] = c
    rains = [0.0] * 6 + [2e-5] * 8 + [0.0] * 6  # the forcing change forces a new capture
    for s, rain in enumerate(rains):
        say(f"SWE graphs vs eager step {s} rain {rain}")
        st = {m: c.swe_step(rain, 0.0) for m, c in ctxs.items()}
        assert st["0"].iterations == st["1"].iterations
        for f in ("eta", "dept", "uu", "vv", "Vs"):
            assert np.array_equal(ctxs["0"].swe_get(f), ctxs["1"].swe_get(f)), f"step {s} field {f}"
