"""The configuration front end on the CPU (no GPU): the YAML-subset reader against what YAML
fixes for each construct, the reference's own benchmark files (benchmarks/b1-sw/b1-sw.yaml,
benchmarks/b2-gw/b2-gw.yaml) mapped to the parameters the reference's tests hand-code for the
same cases (tests/swe/b1_sw_runner.hpp:99-112, tests/re/test_re_b2_gw.cpp:245-275), the schema
and capability gates (tests/integration/test_schema_gate.cpp, test_capability_gate.cpp), and the
data files a configuration reads (ESRI grids north row first, plain lists, soil class maps,
recharge polygons, rainfall series)."""
import os

import numpy as np
import pytest

from frehg_amd import abi, gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
B1 = os.path.join(GOLDEN, "b1_sw", "b1-sw.yaml")
B2 = os.path.join(GOLDEN, "b2_gw", "b2-gw.yaml")


def write(tmp_path, text, name="c.yaml"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def q(path, key, overrides=()):
    return gpu.config_query(path, key, overrides)


# ---- the reader ---------------------------------------------------------------------------------
READER = """\
# leading comment
---
a:
  b: 1            # trailing comment
  c: 'it''s: #1'
  d: "tab\\there"
  url: http://x.org/a#frag
  e: ~
  f:
  g: null
seq_at_key_indent: &anchor1
- 1
- 2
alias: *anchor1
items:
  - face: x+
    legacy_code: 0
  - face: y-
    legacy_code: 2
  - - 7
    - 8
flow: [[0, 0], [1.5, 0], [1.5, 2]]
flow_map: {k: v, n: [2, 3], e: }
multi: [1,
        2, 3]
yes_no: [yes, No, on, OFF, y, N, true, False]
num: [1.0e-08, -.inf, .nan, +3, 0x10]
scalar_anchor: &s 42
scalar_alias: *s
"""


def test_reader_constructs(tmp_path):
    p = write(tmp_path, READER)
    assert q(p, "a.b") == (1, "1")
    assert q(p, "a.c") == (1, "it's: #1")
    assert q(p, "a.d") == (1, "tab\there")
    assert q(p, "a.url") == (1, "http://x.org/a#frag")
    for k in ("a.e", "a.f", "a.g", "a.missing", "nothere.x"):
        assert q(p, k) == (0, ""), k
    assert q(p, "seq_at_key_indent") == (2, "2")
    assert q(p, "alias") == (2, "2")
    assert q(p, "items") == (2, "3")
    assert q(p, "flow") == (2, "3")
    assert q(p, "flow_map") == (3, "3")
    assert q(p, "flow_map.k") == (1, "v")
    assert q(p, "flow_map.n") == (2, "2")
    assert q(p, "flow_map.e") == (0, "")
    assert q(p, "multi") == (2, "3")
    assert q(p, "scalar_alias") == (1, "42")
    assert q(p, "a") == (3, "7")


def test_reader_sequence_items_and_overrides(tmp_path):
    """Overrides reach into the parsed tree (dotted keys, YAML values) and create missing maps."""
    p = write(tmp_path, READER)
    assert q(p, "a.b", ["a.b=[3, 4]"]) == (2, "2")
    assert q(p, "new.deep.key", ["new.deep.key='x y'"]) == (1, "x y")
    assert q(p, "a.b", ["a.b="]) == (0, "")
    with pytest.raises(gpu.FrehgError, match="key=value"):
        q(p, "a.b", ["novalue"])
    with pytest.raises(gpu.FrehgError, match="not a mapping"):
        q(p, "a.b.c", ["a.b.c=1"])


@pytest.mark.parametrize("text,line,msg", [
    ("a:\n\tb: 1\n", 2, "tab"),
    ("a: 1\n   b: 2\n", 2, "indentation"),
    ("a: *nope\n", 1, "unknown alias"),
    ("a: [1, 2\nb: 3\n", 1, "unterminated"),
    ("a: |\n  text\n", 1, "block scalars"),
    ("a: !!int 3\n", 1, "tags"),
    ("a: 1\n---\nb: 2\n", 2, "several documents"),
    ("a:\n  - 1\n  b: 2\n", 3, "indentation"),
    ("a: 1\n- 2\n", 2, "mapping key"),
    ("a: 'open\n", 1, "unterminated quoted"),
])
def test_reader_errors_name_the_line(tmp_path, text, line, msg):
    p = write(tmp_path, text)
    with pytest.raises(gpu.FrehgError) as e:
        q(p, "a")
    assert f":{line}: " in str(e.value) and msg in str(e.value), str(e.value)


# ---- the reference's benchmark files ---------------------------------------------------------------
def test_b1_sw_maps_to_the_reference_runner_parameters():
    """b1-sw.yaml -> the SweParams tests/swe/b1_sw_runner.hpp:99-112 hand-codes for the same case,
    the datum offset -min(bath) (Orchestrator.cpp:424-428) and the 4-sample rain series."""
    s = gpu.config_setup(B1)
    assert (s.surface_water, s.groundwater, s.solute, s.coupled) == (1, 0, 0, 0)
    assert (s.nx, s.ny, s.dt, s.t_end, s.max_steps, s.output_interval) == (1, 10, 5.0, 18000.0, 100, 1800.0)
    p = s.swe
    assert (p.gravity, p.manning, p.min_depth, p.visc_x, p.visc_y, p.hD, p.wtfh, p.dt) == \
        (9.81, 0.019, 1.0e-8, 1.0e-6, 1.0e-6, 0.1, 1.0e-8, 5.0)
    bath = np.loadtxt(os.path.join(GOLDEN, "b1_sw", "bath"))
    assert p.offset == -bath.min() == 0.4
    np.testing.assert_array_equal(gpu.config_field(B1, "bathymetry"), bath + 0.4)
    np.testing.assert_array_equal(gpu.config_field(B1, "eta0"), np.full(10, -2.0))  # initializeState(-2.0)
    rain = np.loadtxt(os.path.join(GOLDEN, "b1_sw", "rain")).reshape(-1, 2)
    np.testing.assert_array_equal(gpu.config_field(B1, "rain_t"), rain[:, 0])
    np.testing.assert_array_equal(gpu.config_field(B1, "rain_v"), rain[:, 1])
    assert p.solver.rtol == 1e-12 and p.solver.ksp_type == abi.KSP_CG  # Orchestrator.cpp:433-437


def test_b2_gw_maps_to_the_reference_test_parameters():
    """b2-gw.yaml -> the ReParams tests/re/test_re_b2_gw.cpp:245-275 hand-codes (the YAML path keeps
    the Orchestrator's GW rtol default 1e-10, Orchestrator.cpp:487-490), initial water content 0.033."""
    s = gpu.config_setup(B2)
    assert (s.surface_water, s.groundwater, s.nx, s.ny, s.nz) == (0, 1, 1, 1, 100)
    r = s.re
    soil = r.soil
    assert (soil.alpha, soil.n, soil.theta_s, soil.theta_r, soil.Ks_x, soil.Ks_y, soil.Ks_z, soil.Ss) == \
        (1.43, 1.56, 0.33, 0.0, 0.0, 0.0, 2.89e-6, 1.0e-5)
    assert (soil.use_vg, soil.use_mvg, soil.air_entry) == (1, 0, -0.02)
    assert (r.dx, r.dy, r.dz, r.botz, r.dt, r.dt_min, r.dt_max, r.co_max) == (1.0, 1.0, 0.01, -1.0, 1e-4, 1e-4, 2.0, 2.0)
    assert (r.adaptive_dt, r.use_corrector, r.use_full3d, r.sim_shallowwater) == (1, 1, 0, 0)
    assert list(r.bc_type) == [0, 0, 0, 0, 0, 1] and r.htop == 0.0
    assert r.solver.rtol == 1e-10 and r.picard_max == 1 and r.initial_guess == abi.GUESS_ZERO
    assert s.gw_initial_head == 0
    np.testing.assert_array_equal(gpu.config_field(B2, "gw_initial"), np.full(100, 0.033))


# ---- gates ----------------------------------------------------------------------------------------
def sw_yaml(sw_extra="", evap_model="0", top_extra="", out_vars=""):
    """tests/integration/test_capability_gate.cpp:21-42."""
    return ("schema_version: '2.0'\n"
            "simulation: {id: gate, mode: surface_water}\n"
            "domain: {nx: 3, ny: 3, nz: 1, dx: 10.0, dy: 10.0, dz: 0.1, botz: 0.0}\n"
            "time: {dt: 1.0, t_end: 1.0, max_steps: 1, output_interval: 0}\n"
            "modules: {surface_water: true, groundwater: false, solute: false}\n"
            "surface_water: {gravity: 9.81, manning: 0.02, min_depth: 1.0e-8, "
            "viscosity: {x: 1.0e-6, y: 1.0e-6}, h_diffusion_ref: 0.1, waterfall_depth: 1.0e-8" + sw_extra + "}\n"
            "initial_conditions: {surface_water: {eta: 0.5}}\n"
            "sources: {surface: {rainfall: {from_file: false, rate: 0.0}, evaporation: {rate: 0.0, model: "
            + evap_model + "}}}\n" + top_extra +
            "output: {format: hdf5, filename: out/gate_sw.h5, io_mode: serial_gather" +
            (", variables: " + out_vars if out_vars else "") + "}\n")


def gw_yaml(bc, gw_extra="", solver="pca"):
    """tests/integration/test_capability_gate.cpp:45-62."""
    return ("schema_version: '2.0'\n"
            "simulation: {id: gate, mode: groundwater}\n"
            "domain: {nx: 1, ny: 1, nz: 6, dx: 1.0, dy: 1.0, dz: 0.01, dz_incre: 1.0, botz: -1.0}\n"
            "time: {dt: 0.1, t_end: 0.1, max_steps: 1, output_interval: 0}\n"
            "modules: {surface_water: false, groundwater: true, solute: false}\n"
            "groundwater: {solver: " + solver + ", full_3d: false, adaptive_dt: false, use_corrector: true, "
            "dt_max: 2.0, dt_min: 0.1, co_max: 2.0, specific_storage: 1.0e-5, bc_type_gw: " + bc + gw_extra + "}\n"
            "soil: {types: [{id: 0, theta_s: 0.33, theta_r: 0.0, vg: {alpha: 1.43, n: 1.56}, "
            "k_sat: {x: 0.0, y: 0.0, z: 2.89e-6}}]}\n"
            "initial_conditions: {groundwater: {wc: 0.2}}\n"
            "output: {format: hdf5, filename: out/gate_gw.h5, io_mode: serial_gather}\n")


def refused(tmp_path, text, expect):
    p = write(tmp_path, text)
    with pytest.raises(gpu.FrehgError) as e:
        gpu.config_setup(p)
    assert expect in str(e.value), str(e.value)


def test_gate_schema_v2(tmp_path):
    """tests/integration/test_schema_gate.cpp:55-83."""
    good = sw_yaml()
    gpu.config_setup(write(tmp_path, good))
    refused(tmp_path, "\n".join(l for l in good.splitlines() if "schema_version" not in l), "schema_version")
    refused(tmp_path, good.replace("'2.0'", "'1.0'"), "schema_version")
    refused(tmp_path, "\n".join(l for l in good.splitlines() if not l.startswith("output:")), "output")
    refused(tmp_path, "\n".join(l for l in good.splitlines() if not l.startswith("modules:")), "modules")
    refused(tmp_path, good.replace("t_end: 1.0", "t_end: soon"), "time.t_end")
    refused(tmp_path, good.replace("nz: 1,", "nz: 0,"), "domain.nz")


def test_gate_surface_water_features(tmp_path):
    """tests/integration/test_capability_gate.cpp:84-95."""
    refused(tmp_path, sw_yaml(", diffusive_wave: true"), "diffusive_wave")
    refused(tmp_path, sw_yaml(", wind: {enabled: true}"), "wind")
    refused(tmp_path, sw_yaml(", subgrid: {enabled: true}"), "subgrid")
    refused(tmp_path, sw_yaml("", "1"), "evaporation.model")
    refused(tmp_path, sw_yaml("", "0", "solute: {baroclinic: true}\n"), "baroclinic")


def test_gate_groundwater_modes_and_schemes(tmp_path):
    """tests/integration/test_capability_gate.cpp:97-112,127-136. A bottom code 3 is free drainage,
    which this backend adds (the reference refuses it)."""
    for i, bc in enumerate(["[1,0,0,0,0,1]", "[0,1,0,0,0,1]", "[0,0,2,0,0,1]", "[0,0,0,2,0,1]"]):
        refused(tmp_path, gw_yaml(bc), f"bc_type_gw[{i}]")
    refused(tmp_path, gw_yaml("[0,0,0,0,4,1]"), "bc_type_gw[4]")
    refused(tmp_path, gw_yaml("[0,0,0,0,0,3]"), "bc_type_gw[5]")
    refused(tmp_path, gw_yaml("[0,0,0,0,0,1]", ", iter_solve: 1"), "iter_solve")
    refused(tmp_path, gw_yaml("[0,0,0,0,0,1]", "", "newton"), "solver")
    for bc, extra in [("[0,0,0,0,0,1]", ""), ("[0,0,0,0,1,1]", ""), ("[0,0,0,0,2,0]", ", qbot: 1.0e-7"),
                      ("[0,0,0,0,0,2]", ", qtop: -1.0e-7"), ("[0,0,0,0,3,1]", "")]:
        s = gpu.config_setup(write(tmp_path, gw_yaml(bc, extra)))
        assert list(s.re.bc_type) == [int(v) for v in bc.strip("[]").split(",")]
    assert gpu.config_setup(write(tmp_path, gw_yaml("[0,0,0,0,2,0]", ", qbot: 1.0e-7"))).re.qbot == 1.0e-7


def test_gate_output_variables(tmp_path):
    """tests/integration/test_capability_gate.cpp:114-125."""
    refused(tmp_path, sw_yaml("", "0", "", "[water_depth, bogus_field]"), "unknown field")
    refused(tmp_path, sw_yaml("", "0", "", "[water_depth, qx]"), "groundwater")
    gpu.config_setup(write(tmp_path, sw_yaml("", "0", "", "[depth, WSE, u, vv]")))


@pytest.mark.parametrize("edit,key", [
    (lambda t: t + "monitors: {probes: [{i: 0, j: 0}]}\n", "monitors"),
    (lambda t: t.replace("{eta: 0.5}", "{eta: {type: formula, formula: 'x'}}"), "formula"),
    (lambda t: t.replace("{eta: 0.5}", "{eta: {type: polygon, default: 0}}"), "polygon"),
    (lambda t: t.replace("initial_conditions: {", "initial_conditions: {regions: [], "), "regions"),
    (lambda t: t.replace("{eta: 0.5}", "{eta: {type: raster, file: e.h5}}"), "HDF5"),
    (lambda t: t.replace("io_mode: serial_gather", "io_mode: parallel"), "io_mode"),
])
def test_gate_refuses_what_this_backend_does_not_implement(tmp_path, edit, key):
    refused(tmp_path, edit(sw_yaml()), key)


def test_polygon_sources_refused(tmp_path):
    text = sw_yaml().replace("sources: {surface: {rainfall: {from_file: false, rate: 0.0}, evaporation: "
                             "{rate: 0.0, model: 0}}}\n",
                             "sources:\n- {type: inflow_rate, rate: 1.0, vertices: [[0, 0], [1, 0], [1, 1]]}\n")
    refused(tmp_path, text, "polygon sources")


# ---- data files -----------------------------------------------------------------------------------
def test_bathymetry_raster_north_row_first_and_nodata(tmp_path):
    """ESRI grids list the north row first (src/io/AsciiRaster.cpp); the model's j grows
    northward (Orchestrator.cpp:404-411); NODATA takes the lowest valid elevation (:386-402)."""
    (tmp_path / "dem.asc").write_text("NCOLS 3\nnrows 2\nxllcenter 0\nYLLCORNER 0\ncellsize 10\nNODATA_value -9999\n"
                                      "1 2 3\n4 -9999 6\n")
    p = write(tmp_path, sw_yaml().replace("botz: 0.0}", "botz: 0.0, bathymetry: {from_file: true, file: dem.asc}}")
              .replace("nx: 3, ny: 3", "nx: 3, ny: 2"))
    s = gpu.config_setup(p)
    # south row (j = 0) is the raster's last line; NODATA -> min valid = 1; offset = -min = -1
    assert s.swe.offset == -1.0
    np.testing.assert_array_equal(gpu.config_field(p, "bathymetry"), np.array([4, 1, 6, 1, 2, 3]) - 1.0)
    (tmp_path / "dem.asc").write_text("ncols 3\nnrows 3\nxllcorner 0\nyllcorner 0\ncellsize 10\n" + "0 " * 9)
    refused(tmp_path, open(p).read(), "raster DEM")


def test_bathymetry_list_and_column(tmp_path):
    (tmp_path / "bath").write_text("0.5 0.25\n-0.5")
    p = write(tmp_path, sw_yaml().replace("botz: 0.0}", "botz: 0.0, bathymetry: {from_file: true}}")
              .replace("nx: 3, ny: 3", "nx: 1, ny: 3"))
    np.testing.assert_array_equal(gpu.config_field(p, "bathymetry"), np.array([1.0, 0.75, 0.0]))
    (tmp_path / "bath").write_text("1 2")
    refused(tmp_path, open(p).read(), "bathymetry file")


def test_soil_map_layers_recharge_polygon_and_head_ic(tmp_path):
    """A per-cell class map from one file per layer (Orchestrator.cpp:547-568), class ids checked
    against soil.types (:514-521), recharge inside a polygon by cell centre (:593-628) and a head IC
    from a file (ICApply.cpp:208-214; a 2D field repeats over the layers)."""
    for k, ids in enumerate(["0 1\n1 0", "1 1\n0 0"]):
        (tmp_path / f"layer{k}").write_text(ids)
    (tmp_path / "h0").write_text("-1 -2 -3 -4")
    text = gw_yaml("[0,0,0,0,0,2]", ", recharge: {rate: -2.0e-6, polygon: [[0, 0], [1.2, 0], [1.2, 2.5], [0, 2.5]]}")
    text = text.replace("nx: 1, ny: 1, nz: 6", "nx: 2, ny: 2, nz: 2")
    text = text.replace("soil: {types: [", "soil: {map: {from_file: true, layers: [layer0, layer1]}, types: [")
    text = text.replace("}}]}\n", "}}, {id: 1, theta_s: 0.4, theta_r: 0.05, vg: {alpha: 2.0, n: 1.8}, "
                                  "k_sat: {x: 1.0e-5, y: 1.0e-5, z: 1.0e-5}}]}\n")
    text = text.replace("{groundwater: {wc: 0.2}}", "{groundwater: {wc: 0.2, h: {type: file, file: h0}}}")
    p = write(tmp_path, text)
    s = gpu.config_setup(p)
    assert (s.soil_classes, s.soil_map, s.recharge, s.gw_initial_head) == (2, 3, 1, 1)
    np.testing.assert_array_equal(gpu.config_field(p, "soil_class"), [0, 1, 1, 0, 1, 1, 0, 0])
    # centres x = 0.5, 1.5 (dx 1); only i = 0 is inside x < 1.2
    np.testing.assert_array_equal(gpu.config_field(p, "top_flux"), [-2e-6, 0, -2e-6, 0])
    np.testing.assert_array_equal(gpu.config_field(p, "gw_initial"), [-1, -2, -3, -4] * 2)
    (tmp_path / "layer1").write_text("1 2\n0 0")
    refused(tmp_path, text, "soil class id 2")


def test_rainfall_series(tmp_path):
    """The rainfall file holds (time, rate) pairs; an odd count is refused (Orchestrator.cpp:317-330)."""
    (tmp_path / "rain").write_text("0 1e-5\n100 2e-5\n")
    text = sw_yaml().replace("rainfall: {from_file: false, rate: 0.0}", "rainfall: {from_file: true}")
    p = write(tmp_path, text)
    s = gpu.config_setup(p)
    assert (s.rain_from_file, s.rain_samples) == (1, 2)
    (tmp_path / "rain").write_text("0 1e-5 100")
    refused(tmp_path, text, "even number")


def test_cli_check_mode(tmp_path):
    """frehg_run --check loads, gates and maps without a GPU; a refused configuration exits 1."""
    import subprocess
    exe = os.path.join(os.path.dirname(gpu.LIB_PATH), "frehg_run")
    r = subprocess.run([exe, "--check", B1], capture_output=True, text=True)
    assert r.returncode == 0 and "1x10x30" in r.stdout, r.stderr
    r = subprocess.run([exe, "--check", "--set", "output.io_mode=parallel", B1], capture_output=True, text=True)
    assert r.returncode == 1 and "io_mode" in r.stderr


def test_scalar_conversions(tmp_path):
    """YAML 1.1 booleans as yaml-cpp reads them, numbers with exponents and signs, and integers
    that must be integers."""
    p = write(tmp_path, gw_yaml("[0,0,0,0,0,1]"))
    for word, v in [("yes", 1), ("No", 0), ("on", 1), ("OFF", 0), ("y", 1), ("N", 0), ("True", 1), ("false", 0)]:
        assert gpu.config_setup(p, [f"groundwater.adaptive_dt={word}"]).re.adaptive_dt == v, word
    assert gpu.config_setup(p, ["time.dt=+2.5e-03"]).dt == 2.5e-3
    assert gpu.config_setup(p, ["time.max_steps=-7"]).max_steps == -7
    with pytest.raises(gpu.FrehgError, match="time.max_steps' must be an integer"):
        gpu.config_setup(p, ["time.max_steps=2.5"])
    with pytest.raises(gpu.FrehgError, match="adaptive_dt' must be a boolean"):
        gpu.config_setup(p, ["groundwater.adaptive_dt=1"])
    with pytest.raises(gpu.FrehgError, match="domain.nx' must be an integer"):
        gpu.config_setup(p, ["domain.nx=one"])


def test_checkpoint_and_restart_keys(tmp_path):
    """time.dt_checkpoint / time.max_checkpoints (Orchestrator.cpp:253-254) and a restart initial
    condition, which needs a file (src/ic/ICConfig.cpp:66-74)."""
    p = write(tmp_path, sw_yaml().replace("output_interval: 0}", "output_interval: 0, dt_checkpoint: 5.0, max_checkpoints: 0}"))
    s = gpu.config_setup(p)
    assert (s.dt_checkpoint, s.max_checkpoints, s.restart) == (5.0, 1, 0)
    p = write(tmp_path, sw_yaml().replace("initial_conditions: {", "initial_conditions: {type: restart, file: run.ckpt.5.bin, "))
    assert gpu.config_setup(p).restart == 1
    refused(tmp_path, sw_yaml().replace("initial_conditions: {", "initial_conditions: {type: restart, "), "requires file")


# ---- the solute block -------------------------------------------------------------------------------
def test_yaml_solute_defaults_overrides_and_enable_key(tmp_path):
    """tests/solute/test_yaml_solute.cpp: defaults when keys are omitted, explicit values override
    them, and no solute without the block; solute.enabled is the key that builds one (modules.solute
    alone does not, Orchestrator.cpp:749-760)."""
    base = sw_yaml()
    s = gpu.config_setup(write(tmp_path, base + "solute: {enabled: true}\n"))
    p = s.solute_params
    assert s.solute == 1
    assert (p.c_rain, p.k_decay, p.D, p.muscl, p.implicit_diffusion, p.cfl_max) == (0.0, 0.0, 1.0e-9, 0, 1, 1.0)
    assert (p.min_depth, p.start_time, p.substeps) == (1.0e-8, 0.0, 1)
    assert (p.solver.ksp_type, p.solver.rtol) == (abi.KSP_CG, 1.0e-12)
    s = gpu.config_setup(write(tmp_path, base + "solute: {enabled: true, c_rain: 12.5, k_decay: 0.02, D: 1.0e-6, "
                                             "advection_scheme: muscl, diffusion_scheme: implicit, cfl_max: 0.8, "
                                             "substeps: 4, start_time: 60, solver: {rtol: 1.0e-10}}\n"))
    p = s.solute_params
    assert (p.c_rain, p.k_decay, p.D, p.muscl, p.cfl_max, p.substeps, p.start_time) == (12.5, 0.02, 1.0e-6, 1, 0.8, 4, 60.0)
    assert p.solver.rtol == 1.0e-10
    s = gpu.config_setup(write(tmp_path, base + "solute: {enabled: true, advection_scheme: tvd, diffusion_scheme: none}\n"))
    assert (s.solute_params.muscl, s.solute_params.implicit_diffusion) == (0, 0)
    assert gpu.config_setup(write(tmp_path, base)).solute == 0
    assert gpu.config_setup(write(tmp_path, base.replace("solute: false", "solute: true"))).solute == 0
    refused(tmp_path, base + "solute: {enabled: true, solver: {precision: mixed}}\n", "solute.solver.precision")


def test_yaml_solute_initial_conditions_and_outputs(tmp_path):
    """initial_conditions.solute.surface / .subsurface (src/ic/ICConfig.cpp:126-129) as a constant or
    a file; `concentration` needs the solute (Orchestrator.cpp:1150-1155)."""
    (tmp_path / "c0").write_text("\n".join(str(v) for v in range(9)))
    text = sw_yaml().replace("initial_conditions: {surface_water: {eta: 0.5}}",
                             "initial_conditions: {surface_water: {eta: 0.5}, solute: {surface: {type: file, file: c0}}}")
    p = write(tmp_path, text + "solute: {enabled: true}\n")
    np.testing.assert_array_equal(gpu.config_field(p, "solute_surface"), np.arange(9.0))
    assert gpu.config_field(p, "solute_subsurface").size == 0
    p = write(tmp_path, sw_yaml().replace("{eta: 0.5}}", "{eta: 0.5}, solute: {surface: 2.5}}") + "solute: {enabled: true}\n")
    np.testing.assert_array_equal(gpu.config_field(p, "solute_surface"), np.full(9, 2.5))
    refused(tmp_path, sw_yaml("", "0", "", "[water_depth, concentration]"), "solute.enabled")
    gpu.config_setup(write(tmp_path, sw_yaml("", "0", "", "[water_depth, conc]") + "solute: {enabled: true}\n"))
