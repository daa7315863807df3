"""The HIP solute against the FP64 oracle: rain mixing, decay, upwind and MUSCL advection with the
substep doubling and the refusal, implicit diffusion through the device CG, on the surface, in
the soil and on both, then inside coupled steps (the SW<->GW exchange carries solute), with the
solute mass and its ledger."""
import numpy as np
import pytest

import oracle_lib as ol
from cases import interior3
from frehg_amd import abi, gpu, workloads
from gpu_common import say
from solute_cases import Flow, inner2, inner3, pad2, pad3, params

pytestmark = pytest.mark.gpu


def random_flow(nx, ny, nz, surface, courant, seed):
    """Random face velocities / fluxes scaled so the largest outgoing Courant number at dt = 1 is
    about `courant`, depths with dry and barely wet cells, random water contents."""
    rng = np.random.default_rng(seed)
    dx, dy, dz = 1.5, 1.0, 0.05
    depth = rng.uniform(0.0, 0.2, (ny, nx))
    depth[rng.uniform(size=(ny, nx)) < 0.15] = 0.0
    depth[rng.uniform(size=(ny, nx)) < 0.05] = 1e-3 + 5e-9
    kw = {}
    if nz is not None:
        kw = dict(wc=rng.uniform(0.1, 0.4, (nz, ny, nx)), qx=rng.uniform(-1, 1, (nz, ny, nx)) * courant * dx / 6,
                  qy=rng.uniform(-1, 1, (nz, ny, nx)) * courant * dy / 6,
                  qz=rng.uniform(-1, 1, (nz, ny, nx)) * courant * dz / 6)
    return Flow(nx, ny, nz, dx=dx, dy=dy, dz=dz, surface=surface, depth=depth,
                u=rng.uniform(-1, 1, (ny, nx)) * courant * dx / 4, v=rng.uniform(-1, 1, (ny, nx)) * courant * dy / 4,
                **kw)


def initial(f, seed):
    rng = np.random.default_rng(seed)
    s = pad2(rng.uniform(0, 2, (f.ny, f.nx)), f.nx, f.ny) if f.surface else None
    g = pad3(rng.uniform(0, 2, (f.nz, f.ny, f.nx)), f.nx, f.ny, f.nz) if f.subsurface else None
    return s, g


def compare(ctx, orc, f, atol, label):
    worst = 0.0
    if f.surface:
        e = float(np.abs(inner2(ctx.solute_get("surface"), f.nx, f.ny) - inner2(orc.get("surface"), f.nx, f.ny)).max())
        worst = max(worst, e)
    if f.subsurface:
        a, b = ctx.solute_get("subsurface"), orc.get("subsurface")
        e = float(np.abs(inner3(a, f.nx, f.ny, f.nz) - inner3(b, f.nx, f.ny, f.nz)).max())
        worst = max(worst, e)
    m_g, st_g = ctx.solute_mass()
    m_o, st_o = orc.mass()
    say(f"  {label}: max |dC| {worst:.2e}, mass {m_g:.12e} vs {m_o:.12e}, substeps {st_g.substeps}/{st_o.substeps}")
    assert worst <= atol, label
    assert m_g == pytest.approx(m_o, rel=1e-12, abs=1e-300)
    assert st_g.substeps == st_o.substeps and st_g.refused_steps == st_o.refused_steps
    assert st_g.max_courant == pytest.approx(st_o.max_courant, rel=1e-13)
    assert st_g.rain_added == pytest.approx(st_o.rain_added, rel=1e-12, abs=1e-300)
    assert st_g.decay_loss == pytest.approx(st_o.decay_loss, rel=1e-9, abs=1e-14)


DOMAINS = {  # (nx, ny, nz, surface): tile edges in x (64-wide tiles) and y (4 rows) are crossed
    "surface": (70, 37, None, True),
    "soil": (67, 9, 11, False),
    "both": (33, 21, 7, True),
}


@pytest.mark.parametrize("muscl", [0, 1])
@pytest.mark.parametrize("domain", list(DOMAINS))
def test_apply_matches_the_oracle(domain, muscl):
    nx, ny, nz, surface = DOMAINS[domain]
    say(f"solute apply {domain} {nx}x{ny}x{nz} muscl={muscl}")
    f = random_flow(nx, ny, nz, surface, courant=2.6, seed=11 + muscl)
    p = params(muscl=muscl, c_rain=3.0, k_decay=1e-3, D=1e-3, implicit_diffusion=1)
    cs, cg = initial(f, 5)
    ctx = gpu.Context(0)
    f.gpu(ctx, p)
    _, _, orc = f.oracle(p)
    for name, a in (("surface", cs), ("subsurface", cg)):
        if a is not None:
            ctx.solute_set(name, a)
            orc.set(name, a)
    for s in range(4):
        ctx.solute_apply(float(s), 1.0, 2e-3)
        orc.apply(float(s), 1.0, 2e-3)
    assert orc.mass()[1].substeps > 4  # the Courant number above 1 doubled the substeps
    compare(ctx, orc, f, 1e-10, domain)
    ctx.close()


def test_unit_courant_shift_refusal_and_no_diffusion_match_exactly():
    say("solute: unit-Courant shift, refusal at 4096 substeps")
    f = Flow(80, 3, u=1.0, v=0.0)
    p = params()
    ctx = gpu.Context(0)
    f.gpu(ctx, p)
    _, _, orc = f.oracle(p)
    c0 = pad2(np.exp(-0.5 * ((np.arange(80) - 10.0) / 2.0) ** 2)[None, :], 80, 3)
    ctx.solute_set("surface", c0)
    orc.set("surface", c0)
    for s in range(6):
        ctx.solute_apply(float(s), 1.0)
        orc.apply(float(s), 1.0)
    assert np.array_equal(ctx.solute_get("surface"), orc.get("surface"))
    ctx.swe_set("uu", pad2(5000.0, 80, 3))
    before = ctx.solute_get("surface")
    ctx.solute_apply(6.0, 1.0)
    assert np.array_equal(ctx.solute_get("surface"), before)
    assert ctx.solute_mass()[1].refused_steps == 1
    ctx.close()


def test_initial_concentrations_and_field_round_trip():
    f = Flow(9, 7, 5)
    ctx = gpu.Context(0)
    f.gpu(ctx, params())
    rng = np.random.default_rng(3)
    s, g = rng.uniform(size=63), rng.uniform(size=315)
    ctx.solute_set_initial(s, g)
    np.testing.assert_array_equal(inner2(ctx.solute_get("surface"), 9, 7).ravel(), s)
    np.testing.assert_array_equal(inner3(ctx.solute_get("subsurface"), 9, 7, 5).ravel(), g)
    with pytest.raises(gpu.FrehgError):
        ctx.solute_get("nonsense")
    ctx.close()


def coupled_pair(case, p, patch):
    """The ponded coupled case with a solute on the GPU and in the oracle."""
    nx, ny, nz = case.nx, case.ny, case.nz
    eta = case.bed.ravel() + 0.02
    ctx = gpu.Context(0)
    case.setup_gpu(ctx, abi.PREC_FP64, abi.GUESS_ZERO, picard_max=30)
    ctx.swe_set_initial_eta(eta)
    ctx.solute_init(p)
    ctx.solute_set_initial(patch, None)
    s = ol.OracleSwe(nx, ny, case.dx, case.dx, case.swe)
    s.set_bathymetry(case.bed_with_offset())
    s.initialize_state(-1.0e3)
    s.set_boundary_polygons(case.boundary_regions())
    s.set_initial_eta(eta)
    r = ol.OracleRe(nx, ny, nz, case.re_params(picard_max=30))
    r.set_soil([case.soil], None, True, case.ks.ravel())
    r.set_initial(h=case.h0)
    sol = ol.OracleSolute(s, r, p)
    sol.set("surface", pad2(patch.reshape(ny, nx), nx, ny))
    return ctx, (s, r, sol, ol.OracleCoupled(s, r, case.cp, case.t_end, solute=sol))


def test_coupled_steps_with_solute_match_the_oracle():
    """A ponded 40x32 surface over 40x32x10 soil under rain of concentration 0.5, a solute patch,
    MUSCL, implicit diffusion: concentrations on both sides after 12 coupled steps."""
    case = workloads.CoupledCase(nx=40, ny=32, nz=10, seed_bed=21, seed_ks=22, rain=20 * workloads.RAIN_50MM_PER_H)
    say("coupled steps with solute: 40x32x10")
    p = params(muscl=1, c_rain=0.5, D=1e-4, implicit_diffusion=1)
    ctx, (s, r, sol, oc) = coupled_pair(case, p, case.solute_patch(side=12.0))
    for _ in range(12):
        st = ctx.coupled_step(case.cp, case.dt, case.rain, 0.0, case.t_end)
        assert st.gw.picard_converged == 1
        oc.step(case.dt, case.rain)
    nx, ny, nz = case.nx, case.ny, case.nz
    es = float(np.abs(inner2(ctx.solute_get("surface"), nx, ny) - inner2(sol.get("surface"), nx, ny)).max())
    cg_o = interior3(sol.get("subsurface"), nx, ny, nz)
    eg = float(np.abs(interior3(ctx.solute_get("subsurface"), nx, ny, nz) - cg_o).max())
    m_g, st_g = ctx.solute_mass()
    m_o, st_o = sol.mass()
    say(f"  max |dC| surface {es:.2e} soil {eg:.2e}; mass {m_g:.10e} vs {m_o:.10e}; soil max C {cg_o.max():.3f}")
    assert cg_o.max() > 0.05  # solute reached the soil
    assert es <= 1e-8 and eg <= 1e-8
    assert m_g == pytest.approx(m_o, rel=1e-9)
    assert st_g.rain_added == pytest.approx(st_o.rain_added, rel=1e-9)
    assert st_g.substeps == st_o.substeps
    ctx.close()
