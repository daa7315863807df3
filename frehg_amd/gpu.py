"""Python binding of the gfx950 backend's C ABI (include/frehg_gpu.h) via ctypes.

Classes mirror the reference's method sets: ReSolver (src/re/ReSolver.hpp:48-105),
SweSolver (src/swe/SweSolver.hpp:41-94), Coupling (src/coupling/Coupling.hpp:84) and
LinearSolver (include/frehg2/linear/LinearSolver.hpp:18-33). Errors raise RuntimeError with the
library's message, as the reference raises std::runtime_error. There is no CPU fallback: if the
HIP library is missing or no GPU is present, construction fails.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfrehg_gpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "frehg_gpu.h")

P = C.POINTER
dp = P(C.c_double)
_lib = None


class FrehgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"frehg error {code}: {msg}")
        self.code = code


def load():
    """Load libfrehg_gpu.so (built by frehg_amd.build); raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP backend not built: {LIB_PATH} missing (run python -m frehg_amd.build)")
    L = C.CDLL(LIB_PATH)
    ctx = C.c_void_p
    sig = {
        "frehg_abi_version": (C.c_int, []),
        "frehg_create": (C.c_int, [P(ctx), C.c_int]),
        "frehg_destroy": (None, [ctx]),
        "frehg_last_error": (C.c_char_p, [ctx]),
        "frehg_synchronize": (C.c_int, [ctx]),
        "frehg_re_init": (C.c_int, [ctx, C.c_int, C.c_int, C.c_int, P(abi.ReParams)]),
        "frehg_re_set_params": (C.c_int, [ctx, P(abi.ReParams)]),
        "frehg_re_set_soil": (C.c_int, [ctx, P(abi.SoilParams), C.c_int, P(C.c_int32), C.c_int, dp]),
        "frehg_re_set_top_flux": (C.c_int, [ctx, dp]),
        "frehg_re_set_initial": (C.c_int, [ctx, dp, dp]),
        "frehg_re_set_field": (C.c_int, [ctx, C.c_char_p, dp]),
        "frehg_re_get_field": (C.c_int, [ctx, C.c_char_p, dp]),
        "frehg_re_step": (C.c_int, [ctx, P(abi.ReStats)]),
        "frehg_re_dt": (C.c_double, [ctx]),
        "frehg_re_prime_dt": (C.c_int, [ctx, C.c_double]),
        "frehg_swe_init": (C.c_int, [ctx, C.c_int, C.c_int, C.c_double, C.c_double, P(abi.SweParams)]),
        "frehg_swe_set_params": (C.c_int, [ctx, P(abi.SweParams)]),
        "frehg_swe_set_bathymetry": (C.c_int, [ctx, dp]),
        "frehg_swe_initialize_state": (C.c_int, [ctx, C.c_double]),
        "frehg_swe_set_initial_eta": (C.c_int, [ctx, dp]),
        "frehg_swe_set_boundary_polygons": (C.c_int, [ctx, C.c_int, P(C.c_int32), dp, P(C.c_int32), dp, C.c_double,
                                                      C.c_double]),
        "frehg_swe_apply_boundaries": (C.c_int, [ctx, C.c_double, dp]),
        "frehg_swe_set_field": (C.c_int, [ctx, C.c_char_p, dp]),
        "frehg_swe_get_field": (C.c_int, [ctx, C.c_char_p, dp]),
        "frehg_swe_step": (C.c_int, [ctx, C.c_double, C.c_double, P(abi.SolveStats)]),
        "frehg_swe_max_cfl": (C.c_int, [ctx, dp]),
        "frehg_couple_exchange": (C.c_int, [ctx, P(abi.CouplingParams), C.c_double, dp, dp]),
        "frehg_coupled_step": (C.c_int, [ctx, P(abi.CouplingParams), C.c_double, C.c_double, C.c_double,
                                         C.c_double, P(abi.CoupledStats)]),
        "frehg_water_volume": (C.c_int, [ctx, dp, dp]),
        "frehg_ls_solve": (C.c_int, [ctx, C.c_int, C.c_int, C.c_int, C.c_int, dp, P(dp), dp, dp,
                                     P(abi.SolverConfig), P(abi.SolveStats)]),
        "frehg_profile_enable": (C.c_int, [ctx, C.c_int]),
        "frehg_profile_select": (C.c_int, [ctx, C.c_uint32]),
        "frehg_profile_read": (C.c_int, [ctx, C.c_int, dp, P(C.c_int64)]),
        "frehg_profile_reset": (C.c_int, [ctx]),
        "frehg_spmv_shape": (C.c_int, [ctx, C.c_int, P(C.c_int), P(C.c_int)]),
        "frehg_ens_init": (C.c_int, [ctx, C.c_int, P(C.c_int32), P(abi.ReParams)]),
        "frehg_ens_set_soil": (C.c_int, [ctx, C.c_int, P(abi.SoilParams), C.c_int, P(C.c_int32), C.c_int, dp]),
        "frehg_ens_set_initial": (C.c_int, [ctx, C.c_int, dp, dp]),
        "frehg_ens_get_field": (C.c_int, [ctx, C.c_int, C.c_char_p, dp]),
        "frehg_ens_step": (C.c_int, [ctx, C.c_int, P(abi.EnsStats), P(abi.ReStats)]),
        "frehg_mlp_init": (C.c_int, [ctx, C.c_uint64]),
        "frehg_mlp_log": (C.c_int, [ctx, C.c_int, C.c_int64]),
        "frehg_mlp_logged": (C.c_int, [ctx, P(C.c_int64)]),
        "frehg_mlp_train": (C.c_int, [ctx, C.c_int, C.c_int, C.c_double, C.c_uint64, dp, dp]),
        "frehg_mlp_get_params": (C.c_int, [ctx, P(C.c_float), P(C.c_float)]),
        "frehg_mlp_set_params": (C.c_int, [ctx, P(C.c_float), P(C.c_float)]),
        "frehg_mlp_loss_grad": (C.c_int, [ctx, P(C.c_float), P(C.c_float), C.c_int, P(C.c_float), dp]),
        "frehg_mlp_predict": (C.c_int, [ctx, P(C.c_float), C.c_int, P(C.c_float)]),
        "frehg_mlp_increment": (C.c_int, [ctx, dp]),
        "frehg_mlp_gate": (C.c_int, [ctx, C.c_double]),
        "frehg_mlp_start": (C.c_int, [ctx, dp]),
        "frehg_mlp_inferred_cells": (C.c_int, [ctx, P(C.c_int64)]),
        "frehg_ls_solve_noalloc": (C.c_int, [ctx, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                             P(C.c_void_p), C.c_void_p, C.c_void_p, P(abi.SolverConfig),
                                             P(abi.SolveStats)]),
        "frehg_config_setup": (C.c_int, [C.c_char_p, P(C.c_char_p), C.c_int, P(abi.RunSetup), C.c_char_p, C.c_int]),
        "frehg_config_query": (C.c_int, [C.c_char_p, P(C.c_char_p), C.c_int, C.c_char_p, C.c_char_p, C.c_int]),
        "frehg_run_config": (C.c_int, [ctx, C.c_char_p, P(C.c_char_p), C.c_int, C.c_char_p, P(abi.RunStats)]),
        "frehg_solute_init": (C.c_int, [ctx, P(abi.SoluteParams)]),
        "frehg_solute_set_initial": (C.c_int, [ctx, dp, dp]),
        "frehg_solute_set_field": (C.c_int, [ctx, C.c_char_p, dp]),
        "frehg_solute_get_field": (C.c_int, [ctx, C.c_char_p, dp]),
        "frehg_solute_apply": (C.c_int, [ctx, C.c_double, C.c_double, C.c_double]),
        "frehg_solute_mass": (C.c_int, [ctx, dp, P(abi.SoluteStats)]),
        "frehg_stream_probe": (C.c_int, [C.c_int, C.c_int64, C.c_int, dp, dp]),
        "frehg_checkpoint_save": (C.c_int, [ctx, C.c_char_p]),
        "frehg_checkpoint_load": (C.c_int, [ctx, C.c_char_p, P(abi.CheckpointInfo)]),
        "frehg_config_field": (C.c_int64, [C.c_char_p, P(C.c_char_p), C.c_int, C.c_char_p, dp, C.c_int64,
                                           C.c_char_p, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if L.frehg_abi_version() != abi.ABI_VERSION:
        raise RuntimeError("libfrehg_gpu.so ABI version mismatch")
    _lib = L
    return L


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(dp)


def _overrides(overrides):
    items = [str(o).encode() for o in (overrides or ())]
    return (C.c_char_p * max(1, len(items)))(*items), len(items)


def config_setup(path, overrides=()):
    """frehg_config_setup: load, gate and map a v2 YAML configuration (no GPU). Raises
    FrehgError with the front end's message when the configuration is refused."""
    arr, n = _overrides(overrides)
    out = abi.RunSetup()
    err = C.create_string_buffer(4096)
    rc = load().frehg_config_setup(str(path).encode(), arr, n, C.byref(out), err, len(err))
    if rc != 0:
        raise FrehgError(rc, err.value.decode())
    return out


def config_field(path, name, overrides=()):
    """frehg_config_field: a resolved per-cell input of a configuration as a float64 array."""
    arr, n = _overrides(overrides)
    err = C.create_string_buffer(4096)
    cnt = load().frehg_config_field(str(path).encode(), arr, n, name.encode(), None, 0, err, len(err))
    if cnt < 0:
        raise FrehgError(int(cnt), err.value.decode())
    out = np.zeros(int(cnt))
    if cnt:
        load().frehg_config_field(str(path).encode(), arr, n, name.encode(), out.ctypes.data_as(dp), cnt, err, len(err))
    return out


def stream_probe(cells, reps=50, device=0):
    """frehg_stream_probe: (GB/s, ms per launch) of the SpMV's six-read / two-write mix."""
    gbs, ms = C.c_double(), C.c_double()
    rc = load().frehg_stream_probe(device, int(cells), int(reps), C.byref(gbs), C.byref(ms))
    if rc != 0:
        raise FrehgError(rc, "frehg_stream_probe failed")
    return gbs.value, ms.value


def config_query(path, key, overrides=()):
    """frehg_config_query: (kind, text) of a dotted key; kind 0 missing/null, 1 scalar,
    2 sequence, 3 mapping (text = size)."""
    arr, n = _overrides(overrides)
    buf = C.create_string_buffer(4096)
    rc = load().frehg_config_query(str(path).encode(), arr, n, key.encode(), buf, len(buf))
    if rc < 0:
        raise FrehgError(rc, buf.value.decode())
    return rc, buf.value.decode()


class Context:
    """One HIP device + stream holding a GW solver, a SW solver and their coupling."""

    def __init__(self, device=0):
        self.lib = load()
        self.h = C.c_void_p()
        rc = self.lib.frehg_create(C.byref(self.h), device)
        if rc != 0:
            raise FrehgError(rc, "frehg_create failed (no GPU?)")
        self.re_dims = None
        self.swe_dims = None
        self.last_rc = 0  # return code of the last checked call (> 0: FREHG_INFO_*)

    def close(self):
        if self.h:
            self.lib.frehg_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc, allow_info=True):
        self.last_rc = rc
        if rc < 0 or (rc > 0 and not allow_info):
            raise FrehgError(rc, self.lib.frehg_last_error(self.h).decode())
        return rc

    # ---- ReSolver ------------------------------------------------------------------------------
    def re_init(self, nx, ny, nz, params):
        self.check(self.lib.frehg_re_init(self.h, nx, ny, nz, C.byref(params)))
        self.re_dims = (nx, ny, nz)

    def re_set_params(self, params):
        self.check(self.lib.frehg_re_set_params(self.h, C.byref(params)))

    def re_set_soil(self, classes, class_idx=None, idx3d=False, ks_cell=None):
        arr = (abi.SoilParams * len(classes))(*classes)
        ci = None
        if class_idx is not None:
            class_idx = np.ascontiguousarray(class_idx, dtype=np.int32).ravel()
            ci = class_idx.ctypes.data_as(P(C.c_int32))
        ks = None if ks_cell is None else np.ascontiguousarray(ks_cell, dtype=np.float64).ravel()
        self.check(self.lib.frehg_re_set_soil(self.h, arr, len(classes), ci, int(idx3d), _ptr(ks)))

    def re_set_top_flux(self, q):
        self.check(self.lib.frehg_re_set_top_flux(self.h, _ptr(np.ascontiguousarray(q, dtype=np.float64).ravel())))

    def re_set_initial(self, h=None, wc=None):
        hh = None if h is None else np.ascontiguousarray(h, dtype=np.float64).ravel()
        ww = None if wc is None else np.ascontiguousarray(wc, dtype=np.float64).ravel()
        self.check(self.lib.frehg_re_set_initial(self.h, _ptr(hh), _ptr(ww)))

    def re_ncell(self):
        nx, ny, nz = self.re_dims
        return (nx + 2) * (ny + 2) * (nz + 2)

    def re_get(self, name):
        out = np.empty(self.re_ncell(), dtype=np.float64)
        self.check(self.lib.frehg_re_get_field(self.h, name.encode(), _ptr(out)))
        return out

    def re_set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        assert a.size == self.re_ncell()
        self.check(self.lib.frehg_re_set_field(self.h, name.encode(), _ptr(a)))

    def re_step(self):
        st = abi.ReStats()
        self.check(self.lib.frehg_re_step(self.h, C.byref(st)))
        return st

    def re_dt(self):
        return self.lib.frehg_re_dt(self.h)

    # ---- SweSolver -----------------------------------------------------------------------------
    def swe_init(self, nx, ny, dx, dy, params):
        self.check(self.lib.frehg_swe_init(self.h, nx, ny, dx, dy, C.byref(params)))
        self.swe_dims = (nx, ny)

    def swe_set_bathymetry(self, bed):
        self.check(self.lib.frehg_swe_set_bathymetry(self.h, _ptr(np.ascontiguousarray(bed, dtype=np.float64).ravel())))

    def swe_initialize_state(self, eta0):
        self.check(self.lib.frehg_swe_initialize_state(self.h, eta0))

    def swe_set_initial_eta(self, eta_phys):
        a = np.ascontiguousarray(eta_phys, dtype=np.float64).ravel()
        assert a.size == self.swe_dims[0] * self.swe_dims[1]
        self.check(self.lib.frehg_swe_set_initial_eta(self.h, _ptr(a)))

    def swe_set_boundary_polygons(self, regions, x0=0.0, y0=0.0):
        """PolygonBC regions: a list of (kind, rate, [(x, y), ...]) with kind abi.BC_*."""
        kind = np.array([r[0] for r in regions], dtype=np.int32)
        rate = np.array([float(r[1]) for r in regions], dtype=np.float64)
        nvert = np.array([len(r[2]) for r in regions], dtype=np.int32)
        xy = np.ascontiguousarray(np.concatenate([np.asarray(r[2], dtype=np.float64).ravel() for r in regions])
                                  if regions else np.zeros(1), dtype=np.float64)
        self.check(self.lib.frehg_swe_set_boundary_polygons(self.h, len(regions), kind.ctypes.data_as(P(C.c_int32)),
                                                             _ptr(rate), nvert.ctypes.data_as(P(C.c_int32)), _ptr(xy),
                                                             x0, y0))

    def swe_apply_boundaries(self, dt):
        v = C.c_double(0.0)
        self.check(self.lib.frehg_swe_apply_boundaries(self.h, dt, C.byref(v)))
        return v.value

    def swe_nsurf(self):
        nx, ny = self.swe_dims
        return (nx + 2) * (ny + 2)

    def swe_get(self, name):
        out = np.empty(self.swe_nsurf(), dtype=np.float64)
        self.check(self.lib.frehg_swe_get_field(self.h, name.encode(), _ptr(out)))
        return out

    def swe_set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        assert a.size == self.swe_nsurf()
        self.check(self.lib.frehg_swe_set_field(self.h, name.encode(), _ptr(a)))

    def swe_step(self, rain, evap=0.0):
        st = abi.SolveStats()
        self.check(self.lib.frehg_swe_step(self.h, rain, evap, C.byref(st)))
        return st

    def swe_max_cfl(self):
        v = C.c_double()
        self.check(self.lib.frehg_swe_max_cfl(self.h, C.byref(v)))
        return v.value

    # ---- Coupling / Orchestrator::step -------------------------------------------------------------
    def couple_exchange(self, cp, dt, want_q=False):
        vol = C.c_double()
        q = np.empty(self.swe_dims[0] * self.swe_dims[1]) if want_q else None
        self.check(self.lib.frehg_couple_exchange(self.h, C.byref(cp), dt, C.byref(vol), _ptr(q)))
        return (vol.value, q) if want_q else vol.value

    def coupled_step(self, cp, dt, rain, evap=0.0, t_end=0.0):
        st = abi.CoupledStats()
        self.check(self.lib.frehg_coupled_step(self.h, C.byref(cp), dt, rain, evap, t_end, C.byref(st)))
        return st

    def water_volume(self):
        a, b = C.c_double(), C.c_double()
        self.check(self.lib.frehg_water_volume(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def synchronize(self):
        self.check(self.lib.frehg_synchronize(self.h))

    # ---- LinearSolver seam ---------------------------------------------------------------------------
    def ls_solve(self, ndim, nx, ny, nz, diag, offdiag, b, cfg):
        x = np.zeros_like(np.ascontiguousarray(b, dtype=np.float64))
        offs = [np.ascontiguousarray(o, dtype=np.float64) for o in offdiag]
        arr = (dp * len(offs))(*[_ptr(o) for o in offs])
        st = abi.SolveStats()
        self.check(self.lib.frehg_ls_solve(self.h, ndim, nx, ny, nz, _ptr(np.ascontiguousarray(diag, dtype=np.float64)),
                                           arr, _ptr(np.ascontiguousarray(b, dtype=np.float64)), _ptr(x),
                                           C.byref(cfg), C.byref(st)))
        return x, st

    def ls_solve_noalloc(self, ndim, nx, ny, nz, diag_dev, offdiag_dev, b_dev, x_dev, cfg):
        """frehg_ls_solve_noalloc on device arrays given by address (owned ordering); x_dev
        receives the solution."""
        arr = (C.c_void_p * len(offdiag_dev))(*[C.c_void_p(int(p)) for p in offdiag_dev])
        st = abi.SolveStats()
        self.check(self.lib.frehg_ls_solve_noalloc(self.h, ndim, nx, ny, nz, C.c_void_p(int(diag_dev)), arr,
                                                   C.c_void_p(int(b_dev)), C.c_void_p(int(x_dev)), C.byref(cfg),
                                                   C.byref(st)))
        return st

    # ---- solute --------------------------------------------------------------------------------------
    def solute_init(self, params):
        self.check(self.lib.frehg_solute_init(self.h, C.byref(params)))

    def solute_set_initial(self, surface=None, subsurface=None):
        s = None if surface is None else np.ascontiguousarray(surface, dtype=np.float64).ravel()
        g = None if subsurface is None else np.ascontiguousarray(subsurface, dtype=np.float64).ravel()
        self.check(self.lib.frehg_solute_set_initial(self.h, _ptr(s), _ptr(g)))

    def _solute_size(self, name):
        if name == "surface":
            nx, ny = self.swe_dims
            return (nx + 2) * (ny + 2)
        return self.re_ncell()

    def solute_get(self, name):
        out = np.empty(self._solute_size(name), dtype=np.float64)
        self.check(self.lib.frehg_solute_get_field(self.h, name.encode(), _ptr(out)))
        return out

    def solute_set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        assert a.size == self._solute_size(name)
        self.check(self.lib.frehg_solute_set_field(self.h, name.encode(), _ptr(a)))

    def solute_apply(self, t, dt, rain=0.0):
        self.check(self.lib.frehg_solute_apply(self.h, t, dt, rain))

    def solute_mass(self):
        m = C.c_double()
        st = abi.SoluteStats()
        self.check(self.lib.frehg_solute_mass(self.h, C.byref(m), C.byref(st)))
        return m.value, st

    # ---- checkpoint / restart --------------------------------------------------------------------------
    def checkpoint_save(self, path):
        self.check(self.lib.frehg_checkpoint_save(self.h, str(path).encode()))

    def checkpoint_load(self, path):
        info = abi.CheckpointInfo()
        self.check(self.lib.frehg_checkpoint_load(self.h, str(path).encode(), C.byref(info)))
        return info

    # ---- configuration front end -------------------------------------------------------------------
    def run_config(self, path, overrides=(), out_dir=None):
        """frehg_run_config: run a v2 YAML configuration on this (empty) context."""
        arr, n = _overrides(overrides)
        st = abi.RunStats()
        self.check(self.lib.frehg_run_config(self.h, str(path).encode(), arr, n,
                                             None if out_dir is None else str(out_dir).encode(), C.byref(st)))
        return st

    # ---- measurement -----------------------------------------------------------------------------------
    def profile_enable(self, on=True):
        self.check(self.lib.frehg_profile_enable(self.h, int(on)))

    def profile_select(self, kinds):
        """Time only the kernel types in `kinds` (an iterable of abi.KT_* ids) while enabled."""
        mask = 0
        for k in kinds:
            mask |= 1 << int(k)
        self.check(self.lib.frehg_profile_select(self.h, mask))

    def profile_read(self, kt):
        ms, n = C.c_double(), C.c_int64()
        self.check(self.lib.frehg_profile_read(self.h, kt, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def profile_reset(self):
        self.check(self.lib.frehg_profile_reset(self.h))

    # ---- batched ensemble ----
    def ens_init(self, dims, params):
        dims = np.ascontiguousarray(np.asarray(dims, dtype=np.int32).reshape(-1, 3))
        arr = (abi.ReParams * len(params))(*params)
        self.check(self.lib.frehg_ens_init(self.h, len(dims), dims.ctypes.data_as(P(C.c_int32)), arr))
        self.ens_dims = [tuple(int(v) for v in d) for d in dims]

    def ens_set_soil(self, box, classes, class_idx=None, idx3d=False, ks_cell=None):
        arr = (abi.SoilParams * len(classes))(*classes)
        ci = None
        if class_idx is not None:
            class_idx = np.ascontiguousarray(class_idx, dtype=np.int32).ravel()
            ci = class_idx.ctypes.data_as(P(C.c_int32))
        ks = None if ks_cell is None else np.ascontiguousarray(ks_cell, dtype=np.float64).ravel()
        self.check(self.lib.frehg_ens_set_soil(self.h, box, arr, len(classes), ci, int(idx3d),
                                               None if ks is None else ks.ctypes.data_as(dp)))

    def ens_set_initial(self, box, h=None, wc=None):
        hh = None if h is None else np.ascontiguousarray(h, dtype=np.float64).ravel()
        ww = None if wc is None else np.ascontiguousarray(wc, dtype=np.float64).ravel()
        self.check(self.lib.frehg_ens_set_initial(self.h, box, None if hh is None else hh.ctypes.data_as(dp),
                                                  None if ww is None else ww.ctypes.data_as(dp)))

    def ens_get(self, box, name):
        nx, ny, nz = self.ens_dims[box]
        out = np.zeros((nx + 2) * (ny + 2) * (nz + 2))
        self.check(self.lib.frehg_ens_get_field(self.h, box, name.encode(), out.ctypes.data_as(dp)))
        return out

    def ens_step(self, use_graph=True, per_box=False):
        st = abi.EnsStats()
        pb = (abi.ReStats * len(self.ens_dims))() if per_box else None
        self.check(self.lib.frehg_ens_step(self.h, int(use_graph), C.byref(st), pb))
        return (st, list(pb)) if per_box else st

    # ---- learned initial guess ----
    def mlp_init(self, seed=1):
        self.check(self.lib.frehg_mlp_init(self.h, seed))

    def mlp_log(self, stride, capacity):
        self.check(self.lib.frehg_mlp_log(self.h, stride, capacity))

    def mlp_logged(self):
        n = C.c_int64()
        self.check(self.lib.frehg_mlp_logged(self.h, C.byref(n)))
        return n.value

    def mlp_train(self, steps, batch=65536, lr=1e-3, seed=7):
        a, b = C.c_double(), C.c_double()
        self.check(self.lib.frehg_mlp_train(self.h, steps, batch, lr, seed, C.byref(a), C.byref(b)))
        return a.value, b.value

    def mlp_get_params(self):
        p = np.zeros(abi.MLP_PARAMS, np.float32)
        n = np.zeros(abi.MLP_NORM, np.float32)
        self.check(self.lib.frehg_mlp_get_params(self.h, _fp(p), _fp(n)))
        return p, n

    def mlp_set_params(self, params, norm):
        p = np.ascontiguousarray(params, dtype=np.float32)
        n = np.ascontiguousarray(norm, dtype=np.float32)
        self.check(self.lib.frehg_mlp_set_params(self.h, _fp(p), _fp(n)))

    def mlp_loss_grad(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(abi.MLP_PARAMS, np.float32)
        loss = C.c_double()
        self.check(self.lib.frehg_mlp_loss_grad(self.h, _fp(x), _fp(y), len(y), _fp(g), C.byref(loss)))
        return loss.value, g

    def mlp_predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros(len(x), np.float32)
        self.check(self.lib.frehg_mlp_predict(self.h, _fp(x), len(x), _fp(out)))
        return out

    def mlp_inferred_cells(self):
        n = C.c_int64(0)
        self.check(self.lib.frehg_mlp_inferred_cells(self.h, C.byref(n)))
        return n.value

    def mlp_gate(self, gate):
        self.check(self.lib.frehg_mlp_gate(self.h, float(gate)))

    def mlp_increment(self):
        out = np.zeros(int(np.prod(self.re_dims)))
        self.check(self.lib.frehg_mlp_increment(self.h, out.ctypes.data_as(dp)))
        return out

    def mlp_start(self):
        """frehg_mlp_start: the learned start the next step would take (interior, x fastest)."""
        out = np.zeros(int(np.prod(self.re_dims)))
        self.check(self.lib.frehg_mlp_start(self.h, out.ctypes.data_as(dp)))
        return out

    def spmv_shape(self, precision):
        ty, kc = C.c_int(), C.c_int()
        self.check(self.lib.frehg_spmv_shape(self.h, precision, C.byref(ty), C.byref(kc)))
        return ty.value, kc.value


def _fp(a):
    return a.ctypes.data_as(P(C.c_float))


def declared_symbols(header=HEADER):
    """Function names declared in include/frehg_gpu.h."""
    import re
    text = open(header).read()
    return sorted(set(re.findall(r"\b(frehg_[a-z0-9_]+)\s*\(", text)))
