"""The configuration front end on the GPU: the reference's b1-sw file run unchanged but for the
step cap reproduces the golden depth vector (benchmarks/b1-sw/frehg2_np1_depth_reference.txt, the
gate of tests/swe/test_swe_b1_mpi.cpp: rel-L2 < 1e-10); the b2-gw file runs the Warrick column as
the oracle does with the reference test's hand-coded parameters (tests/re/test_re_b2_gw.cpp);
and a coupled configuration with every data file kind (bathymetry list, head file, boundary
polygon, soil map) gives the same bits as the same run driven call by call through the ABI, and
the oracle's fields."""
import glob
import os

import numpy as np
import pytest

import oracle_lib as ol
from cases import b1_inputs, interior2, interior3, interp_series
from frehg_amd import abi, gpu
from gpu_common import say

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def snapshot(d, var, t):
    return np.loadtxt(os.path.join(d, f"{var}_{t}.txt"))


def summary(d):
    out = {}
    for line in open(os.path.join(d, "simulation_summary.txt")):
        k, _, v = line.strip().partition(" ")
        out[k] = v
    return out


def test_b1_sw_yaml_reproduces_the_golden_vector(tmp_path):
    say("config front end: b1-sw.yaml")
    ctx = gpu.Context(0)
    st = ctx.run_config(os.path.join(GOLDEN, "b1_sw", "b1-sw.yaml"), ["time.max_steps=0"], out_dir=tmp_path)
    assert (st.steps, st.outputs) == (3600, 11) and st.t == 18000.0
    ref = np.loadtxt(os.path.join(GOLDEN, "b1_sw", "frehg2_np1_depth_reference.txt"))
    worst = 0.0
    for row in ref:
        d = snapshot(tmp_path, "water_depth", int(row[0]))
        worst = max(worst, np.linalg.norm(d - row[1:]) / np.linalg.norm(row[1:]))
    say(f"  worst rel-L2 vs golden {worst:.3e}")
    assert worst < 1e-10
    # The rain added (every row but the last, as the surface step adds it) and the volume gained.
    # The legacy scheme does not conserve exactly (its wet/dry limiters); the oracle run of the
    # same case gains the same volume.
    s = summary(tmp_path)
    assert int(s["time_steps_completed"]) == 3600 and int(s["output_intervals_completed"]) == 11
    bath, rain, _, offset, p = b1_inputs()
    orc = ol.OracleSwe(1, 10, 80.0, 80.0, p)
    orc.set_bathymetry(bath + offset)
    orc.initialize_state(-2.0)
    t, rain_in = 0.0, 0.0
    for _ in range(3600):
        t += 5.0
        r = interp_series(rain[:, 0], rain[:, 1], t)
        orc.step(r, 0.0)
        rain_in += r * 5.0 * 80.0 * 80.0 * 9
    gained = float(np.maximum(interior2(orc.get("eta"), 1, 10) - interior2(orc.get("bottom"), 1, 10), 0).sum()) * 6400
    assert st.rain_volume == pytest.approx(rain_in, rel=1e-14) and float(s["water_rain_volume_in"]) == st.rain_volume
    assert st.water_initial == 0.0
    assert st.water_final - st.water_initial == pytest.approx(gained, rel=1e-10)


def test_b1_sw_yaml_step_cap_and_restarted_context(tmp_path):
    """max_steps as written in the file (100 steps) ends the run at t = 500 s with the initial and
    the final snapshot; a context that already holds a solver is refused."""
    ctx = gpu.Context(0)
    st = ctx.run_config(os.path.join(GOLDEN, "b1_sw", "b1-sw.yaml"), out_dir=tmp_path)
    assert (st.steps, st.outputs, st.t) == (100, 2, 500.0)
    assert sorted(os.path.basename(f) for f in glob.glob(str(tmp_path / "eta_*.txt"))) == ["eta_0.txt", "eta_500.txt"]
    with pytest.raises(gpu.FrehgError) as e:
        ctx.run_config(os.path.join(GOLDEN, "b1_sw", "b1-sw.yaml"), out_dir=tmp_path)
    assert e.value.code == abi.ERR_STATE


def test_b2_gw_yaml_matches_the_oracle(tmp_path):
    say("config front end: b2-gw.yaml")
    ctx = gpu.Context(0)
    nsteps = 300
    st = ctx.run_config(os.path.join(GOLDEN, "b2_gw", "b2-gw.yaml"), [f"time.max_steps={nsteps}"], out_dir=tmp_path)
    assert st.steps == nsteps and st.gw_substeps == nsteps
    # tests/re/test_re_b2_gw.cpp:245-275 with the Orchestrator's rtol default (1e-10)
    soil = abi.soil_params(alpha=1.43, n=1.56, theta_s=0.33, theta_r=0.0, Ks_z=2.89e-6, Ss=1e-5)
    p = abi.re_params(soil=soil, solver=abi.solver_config(1e-10), bc_type=(0, 0, 0, 0, 0, 1), dx=1.0, dy=1.0, dz=0.01,
                      botz=-1.0, dt=1e-4, dt_min=1e-4, dt_max=2.0, co_max=2.0, htop=0.0)
    orc = ol.OracleRe(1, 1, 100, p)
    orc.set_initial(wc=np.full(100, 0.033))
    t = 0.0
    for _ in range(nsteps):
        t += orc.step().dt_used
    assert st.t == pytest.approx(t, rel=1e-9)
    final = [f for f in glob.glob(str(tmp_path / "moisture_*.txt")) if not f.endswith("moisture_0.txt")]
    assert len(final) == 1
    wc = np.loadtxt(final[0])
    want = interior3(orc.get("wc"), 1, 1, 100).ravel()
    say(f"  max |d theta| vs oracle {np.abs(wc - want).max():.3e} after {nsteps} steps (t = {st.t:.4g} s)")
    np.testing.assert_allclose(wc, want, rtol=0, atol=1e-8)
    np.testing.assert_array_equal(snapshot(tmp_path, "moisture", 0), np.full(100, 0.033))


COUPLED = """\
schema_version: '2.0'
simulation: {{id: coupled-yaml, mode: coupled}}
domain:
  nx: {nx}
  ny: {ny}
  nz: {nz}
  dx: 1.0
  dy: 1.0
  dz: 0.05
  botz: 0.0
  bathymetry: {{from_file: true, file: bed.txt}}
time: {{dt: 1.0, t_end: 30.0, max_steps: 0, output_interval: 10}}
modules: {{surface_water: true, groundwater: true, solute: false}}
coupling: {{mode: sync}}
surface_water:
  manning: 0.03
  min_depth: 1.0e-8
  solver: {{ksp_type: cg, rtol: 1.0e-12}}
groundwater:
  solver: {{ksp_type: cg, rtol: 1.0e-12}}
  full_3d: true
  adaptive_dt: false
  dt_max: 2.0
  specific_storage: 1.0e-5
  bc_type_gw: [0, 0, 0, 0, 0, 0]
soil:
  map: {{from_file: true, file: soil.txt}}
  types:
  - {{id: 0, theta_s: 0.40, theta_r: 0.05, vg: {{alpha: 2.0, n: 1.8}}, k_sat: {{x: 1.0e-5, y: 1.0e-5, z: 1.0e-5}}}}
  - {{id: 1, theta_s: 0.35, theta_r: 0.08, vg: {{alpha: 1.5, n: 1.6}}, k_sat: {{x: 3.0e-6, y: 3.0e-6, z: 3.0e-6}}}}
initial_conditions:
  surface_water: {{eta: -10.0}}
  groundwater:
    h: {{type: file, file: h0.txt}}
sources:
  surface:
    rainfall: {{from_file: false, rate: 5.0e-4}}
boundaries:
- name: outlet
  type: bc_critical
  vertices: [[-1.0, -1.0], [1.0, -1.0], [1.0, {yt}], [-1.0, {yt}]]
output: {{format: hdf5, filename: out/run.h5, variables: [water_depth, eta, head, moisture]}}
"""


@pytest.fixture
def pinned_shape():
    """One SpMV launch shape for every context: the autotuner times the candidates, and two
    shapes sum a row in different orders (the last bits of a solve may differ)."""
    os.environ["FREHG_SPMV_SHAPE"] = "2:8"
    yield
    del os.environ["FREHG_SPMV_SHAPE"]


def test_coupled_yaml_equals_the_call_by_call_run_and_the_oracle(tmp_path, pinned_shape):
    nx, ny, nz = 24, 20, 8
    say("config front end: coupled 24x20x8 with a boundary polygon and a soil map")
    rng = np.random.default_rng(5)
    bed = 0.01 * np.arange(nx)[None, :] + 0.02 * rng.standard_normal((ny, nx))
    cls = (np.arange(nx)[None, :] + np.arange(ny)[:, None]) % 2
    np.savetxt(tmp_path / "bed.txt", bed.ravel())
    np.savetxt(tmp_path / "soil.txt", cls.ravel(), fmt="%d")
    np.savetxt(tmp_path / "h0.txt", np.full(nx * ny, -0.3))
    cfg = tmp_path / "coupled.yaml"
    cfg.write_text(COUPLED.format(nx=nx, ny=ny, nz=nz, yt=ny + 1.0))
    setup = gpu.config_setup(cfg)
    assert (setup.coupled, setup.soil_map, setup.boundary_regions, setup.gw_initial_head) == (1, 2, 1, 1)

    ctx = gpu.Context(0)
    st = ctx.run_config(cfg)
    out = tmp_path / "out"
    assert st.steps == 30 and st.outputs == 4 and st.boundary_outflow > 0

    # the same run, call by call, in the front end's order
    ref = gpu.Context(0)
    offset = -bed.min()
    ref.swe_init(nx, ny, 1.0, 1.0, setup.swe)
    ref.swe_set_bathymetry(bed.ravel() + offset)
    ref.re_init(nx, ny, nz, setup.re)
    ref.re_set_soil([setup.re.soil, gpu_soil(setup)], cls.ravel(), False)
    ref.swe_set_boundary_polygons([(abi.BC_CRITICAL, 0.0, [(-1, -1), (1, -1), (1, ny + 1), (-1, ny + 1)])])
    ref.swe_set_initial_eta(np.full(nx * ny, -10.0))
    ref.re_set_initial(h=np.full(nx * ny * nz, -0.3))
    outflow = 0.0
    for _ in range(30):
        cs = ref.coupled_step(setup.coupling, 1.0, 5.0e-4, 0.0, 30.0)
        outflow += cs.boundary_outflow
    assert outflow == st.boundary_outflow
    np.testing.assert_array_equal(snapshot(out, "eta", 30), interior2(ref.swe_get("eta"), nx, ny).ravel())
    np.testing.assert_array_equal(snapshot(out, "head", 30), interior3(ref.re_get("h"), nx, ny, nz).ravel())
    np.testing.assert_array_equal(snapshot(out, "moisture", 30), interior3(ref.re_get("wc"), nx, ny, nz).ravel())

    # and the oracle, driven with the reference's own sequence
    s = ol.OracleSwe(nx, ny, 1.0, 1.0, setup.swe)
    s.set_bathymetry(bed.ravel() + offset)
    s.set_initial_eta(np.full(nx * ny, -10.0))
    s.set_boundary_polygons([(abi.BC_CRITICAL, 0.0, [(-1, -1), (1, -1), (1, ny + 1), (-1, ny + 1)])])
    r = ol.OracleRe(nx, ny, nz, setup.re)
    r.set_soil([setup.re.soil, gpu_soil(setup)], cls.ravel(), False)
    r.set_initial(h=np.full(nx * ny * nz, -0.3))
    oc = ol.OracleCoupled(s, r, setup.coupling, 30.0)
    for _ in range(30):
        oc.step(1.0, 5.0e-4)
    d = snapshot(out, "water_depth", 30)
    want = np.maximum(interior2(s.get("eta"), nx, ny) - interior2(s.get("bottom"), nx, ny), 0).ravel()
    say(f"  max |d depth| vs oracle {np.abs(d - want).max():.3e}, outflow {st.boundary_outflow:.4e} m^3")
    np.testing.assert_allclose(d, want, rtol=0, atol=1e-9)
    np.testing.assert_allclose(snapshot(out, "head", 30), interior3(r.get("h"), nx, ny, nz).ravel(), rtol=0, atol=1e-8)
    assert st.boundary_outflow == pytest.approx(oc.state[4], rel=1e-7)


def gpu_soil(setup):
    """The second soil class of the coupled configuration (the first is setup.re.soil)."""
    return abi.soil_params(theta_s=0.35, theta_r=0.08, alpha=1.5, n=1.6, Ks_x=3.0e-6, Ks_y=3.0e-6, Ks_z=3.0e-6,
                           Ss=1e-5, use_vg=1, use_mvg=0, air_entry=-0.02)


def test_coupled_yaml_with_solute_equals_the_call_by_call_run_and_the_oracle(tmp_path, pinned_shape):
    """The coupled configuration with the solute block (MUSCL, decay, diffusion, rain of
    concentration 0.8, an initial soil concentration from a file): the concentration snapshots
    equal the call-by-call run bit for bit and the oracle's to 1e-8, and the summary carries the
    solute ledger of writeSummary (:1666-1675)."""
    nx, ny, nz = 24, 20, 8
    say("config front end: coupled 24x20x8 with solute")
    rng = np.random.default_rng(6)
    bed = 0.01 * np.arange(nx)[None, :] + 0.02 * rng.standard_normal((ny, nx))
    cls = (np.arange(nx)[None, :] + np.arange(ny)[:, None]) % 2
    c0 = rng.uniform(0.0, 1.0, nx * ny * nz)
    np.savetxt(tmp_path / "bed.txt", bed.ravel())
    np.savetxt(tmp_path / "soil.txt", cls.ravel(), fmt="%d")
    np.savetxt(tmp_path / "h0.txt", np.full(nx * ny, -0.3))
    np.savetxt(tmp_path / "c0.txt", c0)
    text = COUPLED.format(nx=nx, ny=ny, nz=nz, yt=ny + 1.0)
    text = text.replace("solute: false}", "solute: true}")
    text = text.replace("    h: {type: file, file: h0.txt}\n",
                        "    h: {type: file, file: h0.txt}\n  solute: {surface: 0.0, subsurface: {type: file, file: c0.txt}}\n")
    text = text.replace("variables: [water_depth, eta, head, moisture]", "variables: [water_depth, head, concentration]")
    text += ("solute: {enabled: true, c_rain: 0.8, k_decay: 1.0e-4, D: 1.0e-4, advection_scheme: muscl, "
             "solver: {rtol: 1.0e-12}}\n")
    cfg = tmp_path / "coupled.yaml"
    cfg.write_text(text)
    setup = gpu.config_setup(cfg)
    assert setup.solute == 1 and setup.solute_params.muscl == 1
    ctx = gpu.Context(0)
    st = ctx.run_config(cfg)
    out = tmp_path / "out"
    summ = summary(out)
    assert summ["modules"] == "{surface_water=true, groundwater=true, solute=true}"
    assert float(summ["solute_initial_mass"]) == st.solute_initial_mass > 0
    assert float(summ["solute_added_by_rain"]) == st.solute_added_by_rain > 0
    np.testing.assert_array_equal(snapshot(out, "subsurface_concentration", 0), c0)

    ref = gpu.Context(0)
    offset = -bed.min()
    ref.swe_init(nx, ny, 1.0, 1.0, setup.swe)
    ref.swe_set_bathymetry(bed.ravel() + offset)
    ref.re_init(nx, ny, nz, setup.re)
    ref.re_set_soil([setup.re.soil, gpu_soil(setup)], cls.ravel(), False)
    ref.swe_set_boundary_polygons([(abi.BC_CRITICAL, 0.0, [(-1, -1), (1, -1), (1, ny + 1), (-1, ny + 1)])])
    ref.swe_set_initial_eta(np.full(nx * ny, -10.0))
    ref.re_set_initial(h=np.full(nx * ny * nz, -0.3))
    ref.solute_init(setup.solute_params)
    ref.solute_set_initial(np.zeros(nx * ny), c0)
    for _ in range(30):
        ref.coupled_step(setup.coupling, 1.0, 5.0e-4, 0.0, 30.0)
    cs_ref = interior2(ref.solute_get("surface"), nx, ny).ravel()
    cg_ref = interior3(ref.solute_get("subsurface"), nx, ny, nz).ravel()
    np.testing.assert_array_equal(snapshot(out, "surface_concentration", 30), cs_ref)
    np.testing.assert_array_equal(snapshot(out, "subsurface_concentration", 30), cg_ref)
    m, tot = ref.solute_mass()
    assert m == st.solute_final_mass and tot.rain_added == st.solute_added_by_rain

    s = ol.OracleSwe(nx, ny, 1.0, 1.0, setup.swe)
    s.set_bathymetry(bed.ravel() + offset)
    s.set_initial_eta(np.full(nx * ny, -10.0))
    s.set_boundary_polygons([(abi.BC_CRITICAL, 0.0, [(-1, -1), (1, -1), (1, ny + 1), (-1, ny + 1)])])
    r = ol.OracleRe(nx, ny, nz, setup.re)
    r.set_soil([setup.re.soil, gpu_soil(setup)], cls.ravel(), False)
    r.set_initial(h=np.full(nx * ny * nz, -0.3))
    sol = ol.OracleSolute(s, r, setup.solute_params)
    sol.set("subsurface", np.pad(c0.reshape(nz, ny, nx), 1).ravel())
    oc = ol.OracleCoupled(s, r, setup.coupling, 30.0, solute=sol)
    for _ in range(30):
        oc.step(1.0, 5.0e-4)
    es = np.abs(cs_ref - interior2(sol.get("surface"), nx, ny).ravel()).max()
    eg = np.abs(cg_ref - interior3(sol.get("subsurface"), nx, ny, nz).ravel()).max()
    say(f"  max |dC| vs oracle: surface {es:.2e}, soil {eg:.2e}; ledger error {st.solute_relative_mass_error:.3e}")
    assert es <= 1e-8 and eg <= 1e-8
    assert st.solute_added_by_rain == pytest.approx(sol.mass()[1].rain_added, rel=1e-9)


def test_surface_only_and_groundwater_only_yaml_runs_with_solute(tmp_path):
    """The solute in the surface-only branch (applySolute after the surface step, rain of
    concentration c_rain on the file's initial concentration 15; b1-sw.yaml) and in the groundwater-only branch (after each adaptive
    groundwater step, with an initial soil concentration; b2-gw.yaml), both switched on by
    overrides, against the oracle driven with the reference's sequence (Orchestrator::step,
    :1337-1352)."""
    say("config front end: b1-sw and b2-gw with solute")
    sol_keys = ["solute.enabled=true", "solute.c_rain=2.0", "solute.k_decay=1e-5", "solute.D=1e-3",
                "solute.advection_scheme=muscl", "output.variables=[water_depth, concentration]"]
    st = gpu.Context(0).run_config(os.path.join(GOLDEN, "b1_sw", "b1-sw.yaml"), ["time.max_steps=200"] + sol_keys,
                                   out_dir=tmp_path / "b1")
    assert st.steps == 200 and st.solute_added_by_rain > 0
    bath, rain, _, offset, p = b1_inputs()
    sp = gpu.config_setup(os.path.join(GOLDEN, "b1_sw", "b1-sw.yaml"), sol_keys).solute_params
    orc = ol.OracleSwe(1, 10, 80.0, 80.0, p)
    orc.set_bathymetry(bath + offset)
    orc.initialize_state(-2.0)
    sol = ol.OracleSolute(orc, None, sp)
    # the file's own initial_conditions.solute.surface (15)
    c0 = gpu.config_field(os.path.join(GOLDEN, "b1_sw", "b1-sw.yaml"), "solute_surface", sol_keys)
    assert np.all(c0 == 15.0)
    sol.set("surface", np.pad(c0.reshape(10, 1), 1).ravel())
    t = 0.0
    for _ in range(200):
        r = interp_series(rain[:, 0], rain[:, 1], t + 5.0)
        orc.step(r, 0.0)
        sol.apply(t, 5.0, r)
        t += 5.0
    cs = snapshot(tmp_path / "b1", "surface_concentration", 1000)
    want = interior2(sol.get("surface"), 1, 10).ravel()
    say(f"  b1: max |dC| {np.abs(cs - want).max():.2e}, C max {want.max():.3f}")
    assert want.min() < 14.0 and np.abs(cs - want).max() <= 1e-9
    assert st.solute_added_by_rain == pytest.approx(sol.mass()[1].rain_added, rel=1e-12)

    gw_keys = ["time.max_steps=200", "solute.enabled=true", "solute.D=1e-6", "initial_conditions.solute.subsurface=1.5",
               "output.variables=[moisture, conc]"]
    st = gpu.Context(0).run_config(os.path.join(GOLDEN, "b2_gw", "b2-gw.yaml"), gw_keys, out_dir=tmp_path / "b2")
    setup = gpu.config_setup(os.path.join(GOLDEN, "b2_gw", "b2-gw.yaml"), gw_keys)
    np.testing.assert_array_equal(gpu.config_field(os.path.join(GOLDEN, "b2_gw", "b2-gw.yaml"), "solute_subsurface",
                                                   gw_keys), np.full(100, 1.5))
    soil = abi.soil_params(alpha=1.43, n=1.56, theta_s=0.33, theta_r=0.0, Ks_z=2.89e-6, Ss=1e-5)
    pr = abi.re_params(soil=soil, solver=abi.solver_config(1e-10), bc_type=(0, 0, 0, 0, 0, 1), dx=1.0, dy=1.0,
                       dz=0.01, botz=-1.0, dt=1e-4, dt_min=1e-4, dt_max=2.0, co_max=2.0, htop=0.0)
    r = ol.OracleRe(1, 1, 100, pr)
    r.set_initial(wc=np.full(100, 0.033))
    sol = ol.OracleSolute(None, r, setup.solute_params)
    sol.set("subsurface", np.pad(np.full((100, 1, 1), 1.5), 1).ravel())
    t = 0.0
    for _ in range(200):
        used = r.step().dt_used
        sol.apply(t, used, 0.0)
        t += used
    final = [f for f in glob.glob(str(tmp_path / "b2" / "subsurface_concentration_*.txt"))
             if not f.endswith("_0.txt")]
    assert len(final) == 1
    cg = np.loadtxt(final[0])
    want = interior3(sol.get("subsurface"), 1, 1, 100).ravel()
    say(f"  b2: max |dC| {np.abs(cg - want).max():.2e}, mass {st.solute_final_mass:.6e} vs {sol.mass()[0]:.6e}")
    assert np.abs(cg - want).max() <= 1e-8
    assert st.solute_final_mass == pytest.approx(sol.mass()[0], rel=1e-9)
