"""SW<->GW exchange and the coupled step (Orchestrator::step): HIP path vs the oracle, plus
the exchange's exact conservation (tests/coupling/test_coupled_timestep.cpp:64-121)."""
import numpy as np
import pytest

import oracle_lib as ol
from cases import interior2, interior3
from frehg_amd import abi, gpu
from gpu_common import compare_re, compare_swe, say

pytestmark = pytest.mark.gpu


def make_pair(nx, ny, nz, dx, bed, eta0, h0, rtol_gw=1e-12, gw_dt=1.0, adaptive=0):
    sp = abi.swe_params(manning=0.03, dt=1.0)
    soil = abi.soil_params(alpha=2.0, n=1.8, theta_s=0.40, theta_r=0.05, Ks_x=1e-5, Ks_y=1e-5, Ks_z=1e-5, Ss=1e-5)
    rp = abi.re_params(soil=soil, solver=abi.solver_config(rtol_gw), bc_type=(0, 0, 0, 0, 0, 0), dx=dx, dy=dx, dz=0.05,
                       dt=gw_dt, adaptive_dt=adaptive, dt_min=0.05, dt_max=0.6, co_max=2.0, use_full3d=1)
    cp = abi.coupling_params(dx=dx, dy=dx)
    os_ = ol.OracleSwe(nx, ny, dx, dx, sp)
    os_.set_bathymetry(bed.ravel())
    os_.initialize_state(eta0)
    orr = ol.OracleRe(nx, ny, nz, rp)
    orr.set_initial(h=h0)
    ctx = gpu.Context(0)
    ctx.swe_init(nx, ny, dx, dx, sp)
    ctx.swe_set_bathymetry(bed.ravel())
    ctx.swe_initialize_state(eta0)
    ctx.re_init(nx, ny, nz, rp)
    ctx.re_set_initial(h=h0)
    return os_, orr, ctx, cp


def test_exchange_matches_oracle_and_conserves():
    nx, ny, nz = 12, 10, 6
    rng = np.random.default_rng(1)
    bed = 0.05 * rng.random((ny, nx))
    h0 = -0.5 + 0.4 * rng.random(nx * ny * nz)
    os_, orr, ctx, cp = make_pair(nx, ny, nz, 2.0, bed, 0.04, h0)
    sw0, gw0 = ctx.water_volume()
    vg, qg = ctx.couple_exchange(cp, 1.0, want_q=True)
    vo, qo = ol.couple_exchange(os_, orr, cp, 1.0, want_q=True)
    np.testing.assert_allclose(qg, qo, rtol=1e-12, atol=1e-20)
    assert vg == pytest.approx(vo, rel=1e-12)
    sw1, gw1 = ctx.water_volume()
    assert abs((sw1 - sw0) - vg) < 1e-9 and abs((gw1 - gw0) + vg) < 1e-9
    compare_swe(ctx, os_, fields=("eta", "dept"), rtol=1e-13, atol=1e-15)
    compare_re(ctx, orr, fields=("h", "wc"), rtol=1e-12, atol_h=1e-12)


@pytest.mark.parametrize("gw", ["same_dt", "substeps", "adaptive"])
def test_coupled_run_matches_oracle(gw):
    """A small version of the 512x512 coupled configuration: tilted bed with a random field,
    uniform rain, dry start, 1 s surface steps. The groundwater runs at the surface dt, at a
    quarter of it (four substeps per coupled step: gwCatchUpTo, src/driver/Orchestrator.cpp:1243-1252),
    or with its own adaptive dt (substep counts vary, the last substep may overshoot t + dt)."""
    nx, ny, nz = 24, 20, 8
    say(f"coupled 24x20x8 vs oracle, groundwater {gw}")
    rng = np.random.default_rng(5)
    bed = 0.01 * np.arange(nx)[None, :] + 0.02 * rng.standard_normal((ny, nx))
    h0 = np.full(nx * ny * nz, -0.3)
    gw_dt, adaptive = {"same_dt": (1.0, 0), "substeps": (0.25, 0), "adaptive": (0.3, 1)}[gw]
    os_, orr, ctx, cp = make_pair(nx, ny, nz, 1.0, bed, -10.0, h0, gw_dt=gw_dt, adaptive=adaptive)
    oc = ol.OracleCoupled(os_, orr, cp, 100.0)
    rain = 5e-4
    subs = []
    for s in range(60):
        before = int(oc.state[3])  # the oracle counts substeps cumulatively
        oc.step(1.0, rain)
        st = ctx.coupled_step(cp, 1.0, rain, 0.0, 100.0)
        subs.append(st.gw_substeps)
        assert st.gw_substeps == int(oc.state[3]) - before, s
        assert abs(st.gw_time - oc.state[1]) <= 1e-9, (s, st.gw_time, oc.state[1])
    compare_swe(ctx, os_, fields=("eta", "dept", "uu", "vv"), rtol=1e-7, atol=1e-9, label="coupled")
    compare_re(ctx, orr, fields=("h", "wc"), rtol=1e-7, atol_h=1e-7, label="coupled")
    d = interior2(ctx.swe_get("dept"), nx, ny)
    assert (d > 0).any(), "rain must pond somewhere"
    if gw == "same_dt":
        assert set(subs) == {1}
    elif gw == "substeps":
        assert set(subs) == {4}
    else:
        assert len(set(subs)) > 1 and max(subs) > 1, subs
