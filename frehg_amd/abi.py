"""ctypes mirror of include/frehg_gpu.h (the C ABI of the MI355X backend).

Struct layouts, enum values and defaults follow the header field for field. Defaults follow the
reference's parameter structs: SoilParams (include/frehg2/re/SoilParams.hpp:13-25), ReParams
(src/re/ReSolver.hpp:24-46), SweParams (src/swe/SweSolver.hpp:28-39), CouplingParams
(src/coupling/Coupling.hpp:26-31) and the per-operator solver defaults the Orchestrator installs
(src/driver/Orchestrator.cpp:434-437 for SW: cg/jacobi/1e-12; :487-490 for GW: cg/sor/1e-10).
"""
import ctypes as C

ABI_VERSION = 5

OK = 0
INFO_GUESS_FALLBACK = 1
INFO_PICARD_NOT_CONVERGED = 2
ERR_INVALID_ARG = -1
ERR_OOM = -2
ERR_HIP = -3
ERR_NOT_CONVERGED = -4
ERR_NONFINITE = -5
ERR_STATE = -6
ERR_NO_DEVICE = -7

BC_DISCHARGE = 0
BC_DEPTH = 1
BC_CRITICAL = 2

KSP_CG = 0
KSP_BICGSTAB = 1
PC_JACOBI = 0
PREC_FP64 = 0
PREC_MIXED = 1
GUESS_ZERO = 0
GUESS_PREVIOUS = 1
GUESS_LEARNED = 2
GUESS_EXTRAPOLATED = 3

KT_GW_SPMV = 0
KT_GW_SPMV32 = 1
KT_GW_UPDATE = 2
KT_GW_ASSEMBLE = 3
KT_GW_CORRECT = 4
MLP_INPUTS = 11
MLP_PARAMS = 4993
MLP_NORM = 24
KT_SW_SPMV = 5
KT_SW_UPDATE = 6
KT_GW_REFINE = 7
KT_GW_GUESS = 8
KT_COUNT = 9

BC_FREE_DRAINAGE = 3


class SoilParams(C.Structure):
    _fields_ = [("alpha", C.c_double), ("n", C.c_double), ("theta_s", C.c_double),
                ("theta_r", C.c_double), ("Ks_x", C.c_double), ("Ks_y", C.c_double),
                ("Ks_z", C.c_double), ("Ss", C.c_double), ("use_vg", C.c_int32),
                ("use_mvg", C.c_int32), ("air_entry", C.c_double)]


class SolverConfig(C.Structure):
    _fields_ = [("ksp_type", C.c_int32), ("pc_type", C.c_int32), ("rtol", C.c_double),
                ("atol", C.c_double), ("max_it", C.c_int32), ("precision", C.c_int32),
                ("inner_rtol", C.c_double), ("max_refine", C.c_int32)]


class ReParams(C.Structure):
    _fields_ = [("soil", SoilParams), ("dx", C.c_double), ("dy", C.c_double),
                ("dz", C.c_double), ("botz", C.c_double), ("dt", C.c_double),
                ("dt_min", C.c_double), ("dt_max", C.c_double), ("co_max", C.c_double),
                ("adaptive_dt", C.c_int32), ("use_corrector", C.c_int32),
                ("use_full3d", C.c_int32), ("sim_shallowwater", C.c_int32),
                ("bc_type", C.c_int32 * 6), ("htop", C.c_double), ("hbot", C.c_double),
                ("qtop", C.c_double), ("qbot", C.c_double), ("solver", SolverConfig),
                ("picard_max", C.c_int32), ("picard_tol", C.c_double),
                ("initial_guess", C.c_int32)]


class SweParams(C.Structure):
    _fields_ = [("gravity", C.c_double), ("manning", C.c_double), ("min_depth", C.c_double),
                ("visc_x", C.c_double), ("visc_y", C.c_double), ("hD", C.c_double),
                ("hE", C.c_double), ("wtfh", C.c_double), ("dt", C.c_double),
                ("offset", C.c_double), ("solver", SolverConfig)]


class CouplingParams(C.Structure):
    _fields_ = [("dx", C.c_double), ("dy", C.c_double), ("visc", C.c_double),
                ("min_depth", C.c_double)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("refinements", C.c_int32),
                ("residual_norm", C.c_double), ("rhs_norm", C.c_double),
                ("converged", C.c_int32), ("fallback", C.c_int32)]


class ReStats(C.Structure):
    _fields_ = [("solve", SolveStats), ("dt_used", C.c_double), ("dt_next", C.c_double),
                ("picard_iterations", C.c_int32), ("guess_fallback", C.c_int32),
                ("max_dh", C.c_double), ("mass_before", C.c_double), ("mass_after", C.c_double),
                ("mass_boundary_in", C.c_double), ("mass_storage_ss", C.c_double), ("mass_adjust", C.c_double),
                ("mass_residual", C.c_double), ("picard_converged", C.c_int32)]


class EnsStats(C.Structure):
    _fields_ = [("max_iterations", C.c_int32), ("max_refinements", C.c_int32), ("converged", C.c_int32),
                ("graph", C.c_int32), ("extra_launches", C.c_int32), ("realloc_sweeps", C.c_int32),
                ("max_rel_mass_residual", C.c_double), ("mass_before", C.c_double), ("mass_after", C.c_double)]


class CoupledStats(C.Structure):
    _fields_ = [("sw", SolveStats), ("gw", ReStats), ("gw_substeps", C.c_int32),
                ("exchange_volume", C.c_double), ("t", C.c_double), ("gw_time", C.c_double),
                ("boundary_outflow", C.c_double)]


class SoluteParams(C.Structure):
    """frehg_solute_params."""
    _fields_ = [("c_rain", C.c_double), ("k_decay", C.c_double), ("D", C.c_double), ("cfl_max", C.c_double),
                ("min_depth", C.c_double), ("start_time", C.c_double), ("muscl", C.c_int32),
                ("implicit_diffusion", C.c_int32), ("substeps", C.c_int32), ("solver", SolverConfig)]


class SoluteStats(C.Structure):
    """frehg_solute_stats: cumulative solute totals."""
    _fields_ = [("rain_added", C.c_double), ("decay_loss", C.c_double), ("max_courant", C.c_double),
                ("substeps", C.c_int64), ("refused_steps", C.c_int32), ("diffusion_iterations", C.c_int32)]


class RunSetup(C.Structure):
    """frehg_run_setup: a configuration as the front end maps it."""
    _fields_ = [("surface_water", C.c_int32), ("groundwater", C.c_int32), ("solute", C.c_int32),
                ("coupled", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("botz", C.c_double),
                ("dt", C.c_double), ("t_end", C.c_double), ("output_interval", C.c_double),
                ("max_steps", C.c_int64), ("evaporation_rate", C.c_double), ("rain_from_file", C.c_int32),
                ("rain_samples", C.c_int32), ("rain_rate", C.c_double), ("swe", SweParams), ("re", ReParams),
                ("coupling", CouplingParams), ("soil_classes", C.c_int32), ("soil_map", C.c_int32),
                ("recharge", C.c_int32), ("boundary_regions", C.c_int32), ("gw_initial_head", C.c_int32),
                ("restart", C.c_int32), ("dt_checkpoint", C.c_double), ("max_checkpoints", C.c_int32),
                ("solute_params", SoluteParams)]


class CheckpointInfo(C.Structure):
    """frehg_checkpoint_info: what frehg_checkpoint_load restored."""
    _fields_ = [("t", C.c_double), ("gw_time", C.c_double), ("gw_dt", C.c_double), ("has_sw", C.c_int32),
                ("has_gw", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("params_match", C.c_int32), ("has_solute", C.c_int32)]


class RunStats(C.Structure):
    """frehg_run_stats: the outcome of frehg_run_config."""
    _fields_ = [("t", C.c_double), ("steps", C.c_int64), ("gw_substeps", C.c_int64), ("outputs", C.c_int32),
                ("infos", C.c_int32), ("water_initial", C.c_double), ("water_final", C.c_double),
                ("rain_volume", C.c_double), ("boundary_outflow", C.c_double), ("exchange_volume", C.c_double),
                ("wall_seconds", C.c_double), ("solute_initial_mass", C.c_double), ("solute_final_mass", C.c_double),
                ("solute_added_by_rain", C.c_double), ("solute_decay_loss", C.c_double),
                ("solute_relative_mass_error", C.c_double)]


def soil_params(**kw):
    """SoilParams with the reference defaults (SoilParams.hpp:13-25)."""
    s = SoilParams(alpha=1.43, n=1.56, theta_s=0.33, theta_r=0.0, Ks_x=0.0, Ks_y=0.0,
                   Ks_z=2.89e-6, Ss=1.0e-5, use_vg=1, use_mvg=0, air_entry=-0.02)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def solver_config(rtol, **kw):
    c = SolverConfig(ksp_type=KSP_CG, pc_type=PC_JACOBI, rtol=rtol, atol=1e-50, max_it=10000,
                     precision=PREC_FP64, inner_rtol=1e-5, max_refine=8)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def re_params(soil=None, solver=None, bc_type=(0, 0, 0, 0, 0, 1), **kw):
    """ReParams with the reference defaults (ReSolver.hpp:24-46) and GW solver rtol 1e-10."""
    p = ReParams()
    p.soil = soil if soil is not None else soil_params()
    p.dx, p.dy, p.dz, p.botz = 1.0, 1.0, 0.01, -1.0
    p.dt, p.dt_min, p.dt_max, p.co_max = 1.0e-4, 1.0e-4, 2.0, 2.0
    p.adaptive_dt, p.use_corrector, p.use_full3d, p.sim_shallowwater = 1, 1, 0, 0
    for i, b in enumerate(bc_type):
        p.bc_type[i] = b
    p.htop = p.hbot = p.qtop = p.qbot = 0.0
    p.solver = solver if solver is not None else solver_config(1e-10)
    p.picard_max = 1
    p.picard_tol = 1e-8
    p.initial_guess = GUESS_ZERO
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def swe_params(solver=None, **kw):
    """SweParams with the reference defaults (SweSolver.hpp:28-39) and SW solver rtol 1e-12."""
    p = SweParams(gravity=9.81, manning=0.0, min_depth=1.0e-8, visc_x=0.0, visc_y=0.0, hD=0.1,
                  hE=0.0, wtfh=1.0e-8, dt=1.0, offset=0.0)
    p.solver = solver if solver is not None else solver_config(1e-12)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def solute_params(solver=None, **kw):
    """SoluteParams with the reference defaults (src/solute/SoluteParams.hpp:18-25) and the diffusion
    solver of SoluteStepper (rtol 1e-12)."""
    p = SoluteParams(c_rain=0.0, k_decay=0.0, D=1.0e-9, cfl_max=1.0, min_depth=1.0e-8, start_time=0.0,
                     muscl=0, implicit_diffusion=1, substeps=1)
    p.solver = solver if solver is not None else solver_config(1e-12)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def coupling_params(**kw):
    c = CouplingParams(dx=1.0, dy=1.0, visc=1.0, min_depth=1.0e-8)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


# Field names that cross the ABI (GwFields / SweFields names).
RE_FIELDS = ("h", "hn", "wc", "wcn", "wch", "hwc", "ch", "Kx", "Ky", "Kz", "qx", "qy", "qz",
             "Gct", "Grhs")
SWE_FIELDS = ("eta", "etan", "bottom", "bottomXP", "bottomYP", "dept", "deptx", "depty", "uu",
              "vv", "un", "vn", "uy", "vx", "Vs", "Vsn", "Asx", "Asy", "Asz", "Aszx", "Aszy",
              "Vsx", "Vsy", "CDx", "CDy", "Dx", "Dy", "Ex", "Ey", "cflx", "cfly", "Fu", "Fv",
              "Sxp", "Sxm", "Syp", "Sym", "Sct", "Srhs", "cfl_active", "wtfx", "wtfy")
# SWE fields that carry state from one step to the next (everything read before it is written).
SWE_STATE_FIELDS = ("eta", "bottom", "bottomXP", "bottomYP", "dept", "uu", "vv", "uy", "vx",
                    "Vs", "Asx", "Asy", "Asz", "Aszx", "Aszy", "Vsx", "Vsy", "CDx", "CDy",
                    "cflx", "cfly", "cfl_active")
