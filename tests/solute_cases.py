"""Solute set-ups shared by the oracle and GPU solute tests: a frozen flow field (surface depth and
face velocities, soil water content and face Darcy fluxes) put into a pair of solvers, and the
padded-array helpers. Faces follow the solvers' storage: the value at cell (i, j[, k]) is the face
between that cell and its +x / +y / +z neighbour (src/solute/Advection.cpp:41-46, 112-118)."""
import numpy as np

import oracle_lib as ol
from frehg_amd import abi


def pad2(a, nx, ny):
    out = np.zeros((ny + 2, nx + 2))
    out[1:-1, 1:-1] = np.broadcast_to(a, (ny, nx))
    return out.ravel()


def pad3(a, nx, ny, nz):
    out = np.zeros((nz + 2, ny + 2, nx + 2))
    out[1:-1, 1:-1, 1:-1] = np.broadcast_to(a, (nz, ny, nx))
    return out.ravel()


def inner2(a, nx, ny):
    return a.reshape(ny + 2, nx + 2)[1:-1, 1:-1]


def inner3(a, nx, ny, nz):
    return a.reshape(nz + 2, ny + 2, nx + 2)[1:-1, 1:-1, 1:-1]


class Flow:
    """A frozen flow: surface (nx x ny at dx, dy; depth over a flat bed at 0, u, v) and/or
    subsurface (nx x ny x nz at dz; water content, qx, qy, qz). Arrays are interior-shaped,
    (ny, nx) and (nz, ny, nx); scalars broadcast."""

    def __init__(self, nx, ny, nz=None, dx=1.0, dy=1.0, dz=1.0, surface=True, depth=1.0, u=0.0, v=0.0,
                 wc=0.3, qx=0.0, qy=0.0, qz=0.0):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx, self.dy, self.dz = dx, dy, dz
        self.surface = surface
        self.depth, self.u, self.v = depth, u, v
        self.wc, self.qx, self.qy, self.qz = wc, qx, qy, qz

    @property
    def subsurface(self):
        return self.nz is not None

    def re_params(self):
        soil = abi.soil_params(alpha=2.0, n=1.6, theta_s=0.45, theta_r=0.05, Ks_x=1e-5, Ks_y=1e-5, Ks_z=1e-5)
        return abi.re_params(soil=soil, dx=self.dx, dy=self.dy, dz=self.dz, botz=-self.nz * self.dz,
                             dt=1.0, adaptive_dt=0)

    def _surface_fields(self):
        nx, ny = self.nx, self.ny
        return {"eta": pad2(self.depth, nx, ny), "uu": pad2(self.u, nx, ny), "vv": pad2(self.v, nx, ny)}

    def _soil_fields(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        return {"wc": pad3(self.wc, nx, ny, nz), "qx": pad3(self.qx, nx, ny, nz), "qy": pad3(self.qy, nx, ny, nz),
                "qz": pad3(self.qz, nx, ny, nz)}

    def oracle(self, params):
        """(OracleSwe or None, OracleRe or None, OracleSolute)."""
        s = r = None
        if self.surface:
            s = ol.OracleSwe(self.nx, self.ny, self.dx, self.dy, abi.swe_params())
            s.set_bathymetry(np.zeros(self.nx * self.ny))
            s.initialize_state(0.0)
            for k, a in self._surface_fields().items():
                s.set(k, a)
        if self.subsurface:
            p = self.re_params()
            r = ol.OracleRe(self.nx, self.ny, self.nz, p)
            r.set_soil([p.soil])
            r.set_initial(h=np.full(self.nx * self.ny * self.nz, -1.0))
            for k, a in self._soil_fields().items():
                r.set(k, a)
        return s, r, ol.OracleSolute(s, r, params)

    def gpu(self, ctx, params):
        """The same flow on a GPU context, with its solute initialised."""
        if self.surface:
            ctx.swe_init(self.nx, self.ny, self.dx, self.dy, abi.swe_params())
            ctx.swe_set_bathymetry(np.zeros(self.nx * self.ny))
            ctx.swe_initialize_state(0.0)
            for k, a in self._surface_fields().items():
                ctx.swe_set(k, a)
        if self.subsurface:
            p = self.re_params()
            ctx.re_init(self.nx, self.ny, self.nz, p)
            ctx.re_set_soil([p.soil])
            ctx.re_set_initial(h=np.full(self.nx * self.ny * self.nz, -1.0))
            for k, a in self._soil_fields().items():
                ctx.re_set(k, a)
        ctx.solute_init(params)


def params(**kw):
    """Solute parameters for the operator tests: no diffusion unless asked for."""
    kw.setdefault("implicit_diffusion", 0)
    return abi.solute_params(**kw)


def gaussian(x, x0, sd):
    return np.exp(-0.5 * ((x - x0) / sd) ** 2)
