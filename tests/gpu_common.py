"""Helpers shared by the GPU parity tests (oracle = checker, HIP path = the thing tested)."""
import sys

import numpy as np

import oracle_lib as ol
from frehg_amd import abi, gpu

RE_COMPARE = ("h", "hn", "wc", "wcn", "wch", "hwc", "ch", "Kx", "Ky", "Kz", "qx", "qy", "qz", "Gct", "Grhs")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def max_rel_err(a, b, floor):
    return float(np.max(np.abs(a - b) / (np.abs(b) + floor))) if a.size else 0.0


def compare_re(ctx, orc, fields=RE_COMPARE, rtol=1e-9, atol_h=1e-9, label=""):
    """Every GwFields array of the HIP path against the oracle's."""
    worst = {}
    # Relative errors are floored at 1e-3 of the field group's scale: lateral fluxes or
    # conductivities of round-off size (1e-20) next to vertical ones (1e-7) are compared on the
    # group's scale, not their own.
    groups = {"q": ("qx", "qy", "qz"), "K": ("Kx", "Ky", "Kz")}
    scale = {}
    for gname, members in groups.items():
        scale[gname] = max(np.max(np.abs(orc.get(m))) for m in members)
    for f in fields:
        a, b = ctx.re_get(f), orc.get(f)
        if f in ("h", "hn", "hwc"):
            err = float(np.max(np.abs(a - b)))
            ok = err <= atol_h
        else:
            grouped = next((scale[g] for g, m in groups.items() if f in m), None)
            floor = max(grouped, 1e-300) * 1e-3 if grouped is not None else max(np.max(np.abs(b)), 1e-300) * 1e-6
            err = max_rel_err(a, b, floor)
            ok = err <= rtol
        worst[f] = err
        assert ok, f"{label} field {f}: error {err:.3e} (max|gpu|={np.max(np.abs(a)):.3e}, max|oracle|={np.max(np.abs(b)):.3e})"
    return worst


def re_pair(nx, ny, nz, params, soil_map=None, top_flux=None, h0=None, wc0=None):
    """Oracle and GPU Richards solvers with identical parameters and initial state."""
    orc = ol.OracleRe(nx, ny, nz, params)
    ctx = gpu.Context(0)
    ctx.re_init(nx, ny, nz, params)
    if soil_map is not None:
        orc.set_soil(*soil_map)
        ctx.re_set_soil(*soil_map)
    if top_flux is not None:
        orc.set_top_flux(top_flux)
        ctx.re_set_top_flux(top_flux)
    orc.set_initial(h=h0, wc=wc0)
    ctx.re_set_initial(h=h0, wc=wc0)
    return orc, ctx


def swe_copy_state(orc, ctx):
    for f in abi.SWE_STATE_FIELDS:
        ctx.swe_set(f, orc.get(f))


def compare_swe(ctx, orc, fields=abi.SWE_FIELDS, rtol=1e-10, atol=1e-12, label=""):
    for f in fields:
        a, b = ctx.swe_get(f), orc.get(f)
        err = np.abs(a - b)
        tol = atol + rtol * np.abs(b)
        bad = err > tol
        assert not bad.any(), f"{label} SW field {f}: max err {err.max():.3e} at {np.argmax(err - tol)}"
