"""ctypes loader of the FP64 CPU oracle (oracle/liboracle_frehg.so) -- tests only.

The oracle is the checker: tests compare the HIP backend against it, never the other way round.
"""
import ctypes as C
import os

import numpy as np

from frehg_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle_frehg.so")

_lib = None
P = C.POINTER
dp = P(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.orc_last_error.restype = C.c_char_p
        fp = P(C.c_float)
        L.orc_mlp_predict.argtypes = [fp, fp, fp, C.c_int, C.c_int, dp]
        L.orc_mlp_loss_grad.argtypes = [fp, fp, fp, fp, C.c_int, C.c_int, dp, dp]
        L.orc_re_create.restype = C.c_void_p
        L.orc_re_create.argtypes = [C.c_int, C.c_int, C.c_int, P(abi.ReParams)]
        L.orc_re_destroy.argtypes = [C.c_void_p]
        L.orc_re_set_params.argtypes = [C.c_void_p, P(abi.ReParams)]
        L.orc_re_set_soil.argtypes = [C.c_void_p, P(abi.SoilParams), C.c_int, P(C.c_int32), C.c_int, dp]
        L.orc_re_set_top_flux.argtypes = [C.c_void_p, dp]
        L.orc_re_set_initial.argtypes = [C.c_void_p, dp, dp]
        L.orc_re_get.argtypes = [C.c_void_p, C.c_char_p, dp]
        L.orc_re_set.argtypes = [C.c_void_p, C.c_char_p, dp]
        L.orc_re_step.argtypes = [C.c_void_p, P(abi.ReStats)]
        L.orc_re_reallocate.argtypes = [C.c_void_p]
        L.orc_re_dt.restype = C.c_double
        L.orc_re_dt.argtypes = [C.c_void_p]
        L.orc_swe_create.restype = C.c_void_p
        L.orc_swe_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, P(abi.SweParams)]
        L.orc_swe_destroy.argtypes = [C.c_void_p]
        L.orc_swe_set_bathymetry.argtypes = [C.c_void_p, dp]
        L.orc_swe_initialize_state.argtypes = [C.c_void_p, C.c_double]
        L.orc_swe_set_initial_eta.argtypes = [C.c_void_p, dp]
        L.orc_swe_set_boundary_polygons.argtypes = [C.c_void_p, C.c_int, P(C.c_int32), dp, P(C.c_int32), dp,
                                                    C.c_double, C.c_double]
        L.orc_swe_apply_boundaries.argtypes = [C.c_void_p, C.c_double, dp]
        L.orc_swe_get.argtypes = [C.c_void_p, C.c_char_p, dp]
        L.orc_swe_set.argtypes = [C.c_void_p, C.c_char_p, dp]
        L.orc_swe_step.argtypes = [C.c_void_p, C.c_double, C.c_double, P(abi.SolveStats)]
        L.orc_swe_max_cfl.restype = C.c_double
        L.orc_swe_max_cfl.argtypes = [C.c_void_p]
        L.orc_couple_exchange.argtypes = [C.c_void_p, C.c_void_p, P(abi.CouplingParams), C.c_double, dp, dp]
        L.orc_coupled_step.argtypes = [C.c_void_p, C.c_void_p, P(abi.CouplingParams), dp, C.c_double,
                                       C.c_double, C.c_double, C.c_double]
        L.orc_coupled_step_solute.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, P(abi.CouplingParams), dp,
                                              C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_solute_create.restype = C.c_void_p
        L.orc_solute_create.argtypes = [C.c_void_p, C.c_void_p, P(abi.SoluteParams)]
        L.orc_solute_destroy.argtypes = [C.c_void_p]
        L.orc_solute_get.argtypes = [C.c_void_p, C.c_char_p, dp]
        L.orc_solute_set.argtypes = [C.c_void_p, C.c_char_p, dp]
        L.orc_solute_apply.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.orc_solute_step.argtypes = [C.c_void_p, C.c_double, dp]
        L.orc_solute_exchange.argtypes = [C.c_void_p, dp, C.c_int, C.c_double, C.c_double]
        L.orc_solute_mass.argtypes = [C.c_void_p, dp, P(abi.SoluteStats)]
        L.orc_ls_solve.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, dp, P(dp), dp, dp,
                                   P(abi.SolverConfig), P(abi.SolveStats)]
        L.orc_vg_eval.argtypes = [P(abi.SoilParams), C.c_int, dp, dp, dp, dp, dp, dp, dp]
        _lib = L
    return _lib


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(dp)


def check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


class OracleRe:
    """Oracle ReSolver (src/re/ReSolver.hpp method set)."""

    def __init__(self, nx, ny, nz, params):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.params = params
        self.h = lib().orc_re_create(nx, ny, nz, C.byref(params))
        if not self.h:
            raise RuntimeError(lib().orc_last_error().decode())
        self.ncell = (nx + 2) * (ny + 2) * (nz + 2)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_re_destroy(self.h)
            self.h = None

    def set_params(self, p):
        self.params = p
        check(lib().orc_re_set_params(self.h, C.byref(p)))

    def set_soil(self, classes, class_idx=None, idx3d=False, ks_cell=None):
        arr = (abi.SoilParams * len(classes))(*classes)
        ci = None
        if class_idx is not None:
            class_idx = np.ascontiguousarray(class_idx, dtype=np.int32)
            ci = class_idx.ctypes.data_as(P(C.c_int32))
        ks = None if ks_cell is None else np.ascontiguousarray(ks_cell, dtype=np.float64)
        check(lib().orc_re_set_soil(self.h, arr, len(classes), ci, int(idx3d), ptr(ks)))

    def set_top_flux(self, q):
        check(lib().orc_re_set_top_flux(self.h, ptr(np.ascontiguousarray(q, dtype=np.float64))))

    def set_initial(self, h=None, wc=None):
        hh = None if h is None else np.ascontiguousarray(h, dtype=np.float64).ravel()
        ww = None if wc is None else np.ascontiguousarray(wc, dtype=np.float64).ravel()
        check(lib().orc_re_set_initial(self.h, ptr(hh), ptr(ww)))

    def get(self, name):
        out = np.empty(self.ncell, dtype=np.float64)
        check(lib().orc_re_get(self.h, name.encode(), ptr(out)))
        return out

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        assert a.size == self.ncell
        check(lib().orc_re_set(self.h, name.encode(), ptr(a)))

    def step(self):
        st = abi.ReStats()
        check(lib().orc_re_step(self.h, C.byref(st)))
        return st

    def dt(self):
        return lib().orc_re_dt(self.h)

    def reallocate(self):
        """reallocateWaterContent + clampWaterContent + enforceMoistureBc on the current fields."""
        check(lib().orc_re_reallocate(self.h))


class OracleSwe:
    """Oracle SweSolver (src/swe/SweSolver.hpp method set)."""

    def __init__(self, nx, ny, dx, dy, params):
        self.nx, self.ny = nx, ny
        self.params = params
        self.h = lib().orc_swe_create(nx, ny, dx, dy, C.byref(params))
        if not self.h:
            raise RuntimeError(lib().orc_last_error().decode())
        self.nsurf = (nx + 2) * (ny + 2)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_swe_destroy(self.h)
            self.h = None

    def set_bathymetry(self, bed):
        check(lib().orc_swe_set_bathymetry(self.h, ptr(np.ascontiguousarray(bed, dtype=np.float64).ravel())))

    def initialize_state(self, eta0):
        check(lib().orc_swe_initialize_state(self.h, eta0))

    def set_initial_eta(self, eta_phys):
        """setInitialEtaPhysical + setInitialVelocityConstant(0, 0) + finalizeInitialState."""
        check(lib().orc_swe_set_initial_eta(self.h, ptr(np.ascontiguousarray(eta_phys, dtype=np.float64).ravel())))

    def set_boundary_polygons(self, regions, x0=0.0, y0=0.0):
        """PolygonBC regions: a list of (kind, rate, [(x, y), ...]) (kind 0 discharge, 1 depth,
        2 critical)."""
        check(set_polygons(lib().orc_swe_set_boundary_polygons, self.h, regions, x0, y0))

    def apply_boundaries(self, dt):
        v = C.c_double(0.0)
        check(lib().orc_swe_apply_boundaries(self.h, dt, C.byref(v)))
        return v.value

    def get(self, name):
        out = np.empty(self.nsurf, dtype=np.float64)
        check(lib().orc_swe_get(self.h, name.encode(), ptr(out)))
        return out

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        assert a.size == self.nsurf
        check(lib().orc_swe_set(self.h, name.encode(), ptr(a)))

    def step(self, rain, evap=0.0):
        st = abi.SolveStats()
        check(lib().orc_swe_step(self.h, rain, evap, C.byref(st)))
        return st

    def max_cfl(self):
        return lib().orc_swe_max_cfl(self.h)


def set_polygons(fn, handle, regions, x0, y0):
    """Calls a (handle, nregions, kind, rate, nvert, xy, x0, y0) setter with `regions` =
    [(kind, rate, [(x, y), ...]), ...] (shared by the oracle and the GPU bindings)."""
    kind = np.array([r[0] for r in regions], dtype=np.int32)
    rate = np.array([float(r[1]) for r in regions], dtype=np.float64)
    nvert = np.array([len(r[2]) for r in regions], dtype=np.int32)
    xy = np.ascontiguousarray(np.concatenate([np.asarray(r[2], dtype=np.float64).ravel() for r in regions])
                              if regions else np.zeros(0), dtype=np.float64)
    rc = fn(handle, len(regions), kind.ctypes.data_as(P(C.c_int32)), rate.ctypes.data_as(dp),
            nvert.ctypes.data_as(P(C.c_int32)), xy.ctypes.data_as(dp), x0, y0)
    return rc


def couple_exchange(swe, re, cp, dt, want_q=False):
    vol = C.c_double(0.0)
    q = np.empty(swe.nx * swe.ny, dtype=np.float64) if want_q else None
    check(lib().orc_couple_exchange(swe.h, re.h, C.byref(cp), dt, C.byref(vol), ptr(q)))
    return (vol.value, q) if want_q else vol.value


class OracleSolute:
    """Oracle passive solute over the given solvers (either may be None): SoluteStepper and
    Orchestrator::applySolute (src/solute/SoluteStepper.cpp, src/driver/Orchestrator.cpp:938-1056)."""

    def __init__(self, swe, re, params):
        self.swe, self.re, self.params = swe, re, params  # the solvers must outlive this object
        self.h = lib().orc_solute_create(None if swe is None else swe.h, None if re is None else re.h,
                                         C.byref(params))
        if not self.h:
            raise RuntimeError(lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_solute_destroy(self.h)
            self.h = None

    def _size(self, name):
        if name == "surface":
            return (self.swe.nx + 2) * (self.swe.ny + 2)
        return self.re.ncell

    def get(self, name):
        out = np.empty(self._size(name), dtype=np.float64)
        check(lib().orc_solute_get(self.h, name.encode(), ptr(out)))
        return out

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        assert a.size == self._size(name)
        check(lib().orc_solute_set(self.h, name.encode(), ptr(a)))

    def apply(self, t, dt, rain=0.0):
        check(lib().orc_solute_apply(self.h, t, dt, rain))

    def step(self, dt):
        """One stepper step (decay, advection, diffusion); returns (Courant number, accepted)."""
        out = np.zeros(2, dtype=np.float64)
        check(lib().orc_solute_step(self.h, dt, ptr(out)))
        return float(out[0]), bool(out[1])

    def exchange(self, q, dt, area):
        qq = np.ascontiguousarray(q, dtype=np.float64).ravel()
        check(lib().orc_solute_exchange(self.h, ptr(qq), qq.size, dt, area))

    def mass(self):
        m = C.c_double()
        st = abi.SoluteStats()
        check(lib().orc_solute_mass(self.h, C.byref(m), C.byref(st)))
        return m.value, st


class OracleCoupled:
    """Orchestrator::step, coupled branch (src/driver/Orchestrator.cpp:1322-1336), with the solute
    after the flow step when one is given."""

    def __init__(self, swe, re, cp, t_end, solute=None):
        self.swe, self.re, self.cp, self.t_end, self.solute = swe, re, cp, t_end, solute
        self.state = np.zeros(5, dtype=np.float64)  # t, gw_time, exchange volume, substeps, outflow

    def step(self, dt, rain, evap=0.0):
        if self.solute is None:
            check(lib().orc_coupled_step(self.swe.h, self.re.h, C.byref(self.cp), ptr(self.state), dt,
                                         rain, evap, self.t_end))
        else:
            check(lib().orc_coupled_step_solute(self.swe.h, self.re.h, self.solute.h, C.byref(self.cp),
                                                ptr(self.state), dt, rain, evap, self.t_end))


def ls_solve(ndim, nx, ny, nz, diag, offdiag, b, cfg):
    x = np.zeros_like(b)
    offs = [np.ascontiguousarray(o, dtype=np.float64) for o in offdiag]
    arr = (dp * len(offs))(*[ptr(o) for o in offs])
    st = abi.SolveStats()
    check(lib().orc_ls_solve(ndim, nx, ny, nz, ptr(np.ascontiguousarray(diag)), arr,
                             ptr(np.ascontiguousarray(b)), ptr(x), C.byref(cfg), C.byref(st)))
    return x, st


def vg_eval(soil, h=None, wc=None):
    n = len(h) if h is not None else len(wc)
    hh = np.ascontiguousarray(h if h is not None else np.zeros(n), dtype=np.float64)
    ww = np.ascontiguousarray(wc if wc is not None else np.full(n, 0.5 * (soil.theta_s + soil.theta_r)),
                              dtype=np.float64)
    th, cap, K, hd, dk = (np.empty(n) for _ in range(5))
    check(lib().orc_vg_eval(C.byref(soil), n, ptr(hh), ptr(ww), ptr(th), ptr(cap), ptr(K), ptr(hd), ptr(dk)))
    return {"theta": th, "C": cap, "K": K, "head": hd, "dKdwc": dk}


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def mlp_predict(params, norm, x, bf16):
    """CPU reference forward pass of the learned-guess MLP (raw increments). bf16: 0 FP64, 1 the
    training kernels' BF16 arithmetic, 2 the inference kernel's FP16 arithmetic."""
    params, norm, x = _f32(params), _f32(norm), _f32(x)
    out = np.zeros(len(x))
    fp = C.POINTER(C.c_float)
    lib().orc_mlp_predict(params.ctypes.data_as(fp), norm.ctypes.data_as(fp), x.ctypes.data_as(fp), len(x),
                          int(bf16), ptr(out))
    return out


def mlp_loss_grad(params, norm, x, y, bf16):
    """CPU reference loss and gradient of the learned-guess MLP."""
    params, norm, x, y = _f32(params), _f32(norm), _f32(x), _f32(y)
    grad = np.zeros(len(params))
    loss = C.c_double()
    fp = C.POINTER(C.c_float)
    lib().orc_mlp_loss_grad(params.ctypes.data_as(fp), norm.ctypes.data_as(fp), x.ctypes.data_as(fp),
                            y.ctypes.data_as(fp), len(x), int(bf16), ptr(grad), C.byref(loss))
    return loss.value, grad
