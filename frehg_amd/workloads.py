"""Synthetic inputs of the benchmark configurations (no data files: everything is generated from
seeds so runs are reproducible).

Gaussian random fields use FFT spectral synthesis with a Gaussian covariance
C(r) = sigma^2 exp(-r^2 / l^2) on the grid's own spacing, driven by a counter-based RNG
(numpy's Philox bit generator), normalised to exactly zero mean and unit sample variance before
scaling.
"""
import numpy as np

from . import abi

RAIN_50MM_PER_H = 50e-3 / 3600.0


def grf(shape, spacing, corr_len, sigma, seed):
    """Zero-mean Gaussian random field with Gaussian covariance of correlation length
    `corr_len` (same units as `spacing`) and standard deviation `sigma`."""
    rng = np.random.Generator(np.random.Philox(seed))
    white = rng.standard_normal(shape)
    spec = np.fft.fftn(white)
    k2 = np.zeros(shape)
    for ax, (n, d) in enumerate(zip(shape, spacing)):
        k = 2.0 * np.pi * np.fft.fftfreq(n, d=d)
        sh = [1] * len(shape)
        sh[ax] = n
        k2 = k2 + (k.reshape(sh)) ** 2
    # spectral density of exp(-r^2/l^2) is proportional to exp(-k^2 l^2 / 4); amplitude = sqrt
    field = np.real(np.fft.ifftn(spec * np.exp(-k2 * corr_len ** 2 / 8.0)))
    field -= field.mean()
    sd = field.std()
    return sigma * field / sd if sd > 0 else field


class CoupledCase:
    """Coupled surface/subsurface benchmark case: a 512x512 surface over 512x512x30 soil (the
    defaults), or any size of the same recipe (large() gives the 1024x1024 surface at dx = 2 m over
    1024x1024x64 soil at dz = 0.04 m)."""

    @classmethod
    def large(cls, seed_bed=11, seed_ks=12, **kw):
        return cls(nx=1024, ny=1024, nz=64, dx=2.0, dz=0.04, seed_bed=seed_bed, seed_ks=seed_ks, **kw)

    def __init__(self, nx=512, ny=512, nz=30, dx=1.0, dz=0.05, seed_bed=1, seed_ks=2, slope=0.01,
                 bed_sigma=0.05, bed_corr=20.0, ks_median=1e-5, ks_sigma_ln=1.0, ks_corr=10.0,
                 alpha=2.0, n=1.8, theta_s=0.40, theta_r=0.05, ss=1e-5, h0=-2.0, manning=0.03,
                 rain=RAIN_50MM_PER_H, dt=1.0, t_end=1800.0, outflow=True):
        self.nx, self.ny, self.nz, self.dx, self.dz = nx, ny, nz, dx, dz
        self.dt, self.rain, self.t_end = dt, rain, t_end
        x = (np.arange(nx) + 0.5) * dx
        # bed: tilted plane (low edge at x = 0) plus a random field
        self.bed = slope * x[None, :] + grf((ny, nx), (dx, dx), bed_corr, bed_sigma, seed_bed)
        lnk = grf((nz, ny, nx), (dz, dx, dx), ks_corr, ks_sigma_ln, seed_ks)
        self.ks = ks_median * np.exp(lnk)
        self.soil = abi.soil_params(alpha=alpha, n=n, theta_s=theta_s, theta_r=theta_r, Ks_x=ks_median,
                                    Ks_y=ks_median, Ks_z=ks_median, Ss=ss)
        self.h0 = np.full(nx * ny * nz, h0)
        self.swe = abi.swe_params(manning=manning, dt=dt, offset=-float(self.bed.min()))
        self.cp = abi.coupling_params(dx=dx, dy=dx)
        # closed sides and bottom; the exchange is the only top flux (bc z- = 0)
        self.bc = (0, 0, 0, 0, 0, 0)
        self.outflow = outflow

    def solute_patch(self, side=100.0):
        """Initial surface concentration: 1 in a side x side m square at the domain centre, 0
        elsewhere (interior cells, x fastest)."""
        x = (np.arange(self.nx) + 0.5) * self.dx
        y = (np.arange(self.ny) + 0.5) * self.dx
        cx, cy = 0.5 * self.nx * self.dx, 0.5 * self.ny * self.dx
        inx = np.abs(x - cx) < 0.5 * side
        iny = np.abs(y - cy) < 0.5 * side
        return (iny[:, None] & inx[None, :]).astype(np.float64).ravel()

    def boundary_regions(self):
        """Outflow at the low-x edge: a critical-depth region (bc_critical) over the first column
        of surface cells (centres at x = dx/2), as a polygon strip in domain coordinates."""
        if not self.outflow:
            return []
        dx, ly = self.dx, self.ny * self.dx
        return [(abi.BC_CRITICAL, 0.0, [(-dx, -dx), (dx, -dx), (dx, ly + dx), (-dx, ly + dx)])]

    def re_params(self, precision=abi.PREC_FP64, guess=abi.GUESS_ZERO, rtol=1e-10, inner_rtol=1e-5, picard_max=1,
                  picard_tol=1e-8):
        return abi.re_params(soil=self.soil, solver=abi.solver_config(rtol, precision=precision, inner_rtol=inner_rtol,
                                                                      max_refine=8),
                             bc_type=self.bc, dx=self.dx, dy=self.dx, dz=self.dz, botz=-self.dz * self.nz, dt=self.dt,
                             adaptive_dt=0, use_full3d=1, initial_guess=guess, picard_max=picard_max,
                             picard_tol=picard_tol)

    @property
    def cells(self):
        return self.nx * self.ny * self.nz + self.nx * self.ny

    def bed_with_offset(self):
        return (self.bed + self.swe.offset).ravel()

    def setup_gpu(self, ctx, precision=abi.PREC_FP64, guess=abi.GUESS_ZERO, picard_max=1):
        ctx.swe_init(self.nx, self.ny, self.dx, self.dx, self.swe)
        ctx.swe_set_bathymetry(self.bed_with_offset())
        ctx.swe_initialize_state(-1.0e3)  # dry start: eta clamps to the bed
        ctx.swe_set_boundary_polygons(self.boundary_regions())
        ctx.re_init(self.nx, self.ny, self.nz, self.re_params(precision, guess, picard_max=picard_max))
        ctx.re_set_soil([self.soil], None, True, self.ks.ravel())
        ctx.re_set_initial(h=self.h0)


class InfiltrationCase:
    """The 40x40x50 infiltration benchmark: homogeneous van Genuchten soil (alpha 3.35 1/m, n 2.0,
    theta_s 0.368, theta_r 0.102, Ks 9.22e-5 m/s, Ss 1e-6 1/m), dx = dy = 0.1 m, dz = 0.02 m,
    initial head -10 m, top Dirichlet h = -0.75 m, free-drainage bottom (unit gradient, bc 3), no-flux
    sides, dt = 10 s fixed for 360 steps, Picard to 1e-8 m, CG on the 7-point system."""

    def __init__(self, nx=40, ny=40, nz=50, dt=10.0, steps=360, picard_max=50, picard_tol=1e-8):
        self.nx, self.ny, self.nz, self.dt, self.steps = nx, ny, nz, dt, steps
        self.soil = abi.soil_params(alpha=3.35, n=2.0, theta_s=0.368, theta_r=0.102, Ks_x=9.22e-5, Ks_y=9.22e-5,
                                    Ks_z=9.22e-5, Ss=1e-6)
        self.picard_max, self.picard_tol = picard_max, picard_tol
        self.h0 = np.full(nx * ny * nz, -10.0)

    def re_params(self, precision=abi.PREC_FP64, guess=abi.GUESS_ZERO, rtol=1e-10):
        return abi.re_params(soil=self.soil, solver=abi.solver_config(rtol, precision=precision, inner_rtol=1e-5,
                                                                      max_refine=8),
                             bc_type=(0, 0, 0, 0, 3, 1), dx=0.1, dy=0.1, dz=0.02, botz=-0.02 * self.nz, dt=self.dt,
                             adaptive_dt=0, use_full3d=1, htop=-0.75, initial_guess=guess,
                             picard_max=self.picard_max, picard_tol=self.picard_tol)

    @property
    def cells(self):
        return self.nx * self.ny * self.nz


class EnsembleCase:
    """Batched ensemble of independent Richards boxes of uneven size: nx, ny uniform in [32, 128],
    nz uniform in [20, 120], dx = dy = 0.5 m, dz = 0.05 m; per box (seeded) a lognormal Ks field
    (median 1e-5 m/s, sigma_ln = 1, Gaussian correlation 2 m), alpha uniform in [0.5, 5] 1/m,
    n uniform in [1.3, 2.8], theta_s = 0.40, theta_r = 0.05, Ss = 1e-5; a downward top flux of
    1e-6 m/s, a closed bottom, a uniform initial head per box drawn from [-5, -0.5] m; fixed dt."""

    def __init__(self, nbox=96, seed=3, dt=30.0, nxy=(32, 128), nz=(20, 120), precision=abi.PREC_MIXED,
                 guess=abi.GUESS_PREVIOUS, rtol=1e-10):
        rng = np.random.Generator(np.random.Philox(seed))
        self.dims, self.params, self.soils, self.ks, self.h0 = [], [], [], [], []
        for b in range(nbox):
            nx, ny = (int(v) for v in rng.integers(nxy[0], nxy[1] + 1, size=2))
            nzb = int(rng.integers(nz[0], nz[1] + 1))
            alpha = float(rng.uniform(0.5, 5.0))
            n = float(rng.uniform(1.3, 2.8))
            soil = abi.soil_params(alpha=alpha, n=n, theta_s=0.40, theta_r=0.05, Ks_x=1e-5, Ks_y=1e-5, Ks_z=1e-5,
                                   Ss=1e-5)
            lnk = grf((nzb, ny, nx), (0.05, 0.5, 0.5), 2.0, 1.0, seed * 1000 + b)
            p = abi.re_params(soil=soil, solver=abi.solver_config(rtol, precision=precision, inner_rtol=1e-5,
                                                                    max_refine=8),
                              bc_type=(0, 0, 0, 0, 0, 2), dx=0.5, dy=0.5, dz=0.05, botz=-0.05 * nzb, dt=dt,
                              adaptive_dt=0, use_full3d=1, qtop=-1e-6, initial_guess=guess)
            self.dims.append((nx, ny, nzb))
            self.params.append(p)
            self.soils.append(soil)
            self.ks.append(1e-5 * np.exp(lnk).ravel())
            self.h0.append(float(rng.uniform(-5.0, -0.5)))

    @property
    def cells(self):
        return int(sum(nx * ny * nz for nx, ny, nz in self.dims))

    def setup_gpu(self, ctx):
        ctx.ens_init(self.dims, self.params)
        for b, (nx, ny, nz) in enumerate(self.dims):
            ctx.ens_set_soil(b, [self.soils[b]], None, True, self.ks[b])
            ctx.ens_set_initial(b, h=np.full(nx * ny * nz, self.h0[b]))
