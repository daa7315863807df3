"""The large coupled configuration at its full size: a 1024x1024 surface at dx = 2 m over
1024x1024x64 soil at dz = 0.04 m (67M subsurface cells on one GPU), with the critical-depth
outflow on the low-x edge and the solute tracer (concentration 1 in a 100 x 100 m surface patch,
molecular diffusion 1e-9 m^2/s, minmod MUSCL advection). One coupled step against the FP64
oracle (heads, depths and both concentrations), then the FP64 water ledger of every step closes
to round-off."""
import numpy as np
import pytest

import oracle_lib as ol
from cases import interior2, interior3
from frehg_amd import abi, gpu, workloads
from gpu_common import say

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


def test_large_configuration_one_step_against_the_oracle_and_ledger():
    say("large configuration 1024x1024x64: building inputs")
    case = workloads.CoupledCase.large()
    nx, ny, nz = case.nx, case.ny, case.nz
    ctx = gpu.Context(0)
    case.setup_gpu(ctx, abi.PREC_FP64, abi.GUESS_ZERO, picard_max=50)
    sp = abi.solute_params(muscl=1, D=1e-9)
    patch = case.solute_patch(side=100.0)
    ctx.solute_init(sp)
    ctx.solute_set_initial(patch, None)
    rain = case.rain * 200  # the first steps of the storm at a rate that ponds in one step
    st = ctx.coupled_step(case.cp, case.dt, rain, 0.0, case.t_end)
    assert st.gw.picard_converged == 1
    h1, d1 = ctx.re_get("h"), ctx.swe_get("dept")
    say("  GPU step done; FP64 oracle step")
    s = ol.OracleSwe(nx, ny, case.dx, case.dx, case.swe)
    s.set_bathymetry(case.bed_with_offset())
    s.initialize_state(-1.0e3)
    s.set_boundary_polygons(case.boundary_regions())
    r = ol.OracleRe(nx, ny, nz, case.re_params(picard_max=50))
    r.set_soil([case.soil], None, True, case.ks.ravel())
    r.set_initial(h=case.h0)
    sol = ol.OracleSolute(s, r, sp)
    sol.set("surface", np.pad(patch.reshape(ny, nx), 1).ravel())
    oc = ol.OracleCoupled(s, r, case.cp, case.t_end, solute=sol)
    oc.step(case.dt, rain)
    eh = float(np.abs(interior3(h1, nx, ny, nz) - interior3(r.get("h"), nx, ny, nz)).max())
    dd = interior2(s.get("dept"), nx, ny)
    ed = float(np.abs(interior2(d1, nx, ny) - dd).max())
    say(f"  max |dh| {eh:.2e} m, max |d depth| {ed:.2e} m, ponded cells {(dd > 0).sum()}")
    assert eh <= 1e-9 and ed <= 1e-12
    assert (dd > 0).sum() > 1000
    es = float(np.abs(interior2(ctx.solute_get("surface"), nx, ny) - interior2(sol.get("surface"), nx, ny)).max())
    cg_o = interior3(sol.get("subsurface"), nx, ny, nz)[0]
    eg = float(np.abs(interior3(ctx.solute_get("subsurface"), nx, ny, nz)[0] - cg_o).max())
    say(f"  solute: max |dC| surface {es:.2e}, top soil {eg:.2e}; top soil cells reached {(cg_o > 0).sum()}")
    assert es <= 1e-8 and eg <= 1e-8 and (cg_o > 0).sum() > 100
    del s, r, sol, oc
    worst = abs(st.gw.mass_residual) / st.gw.mass_before
    for _ in range(3):
        st = ctx.coupled_step(case.cp, case.dt, rain, 0.0, case.t_end)
        worst = max(worst, abs(st.gw.mass_residual) / st.gw.mass_before)
    say(f"  worst relative ledger residual {worst:.2e}")
    assert worst < 1e-13
