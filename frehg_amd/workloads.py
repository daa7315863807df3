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
    """Coupled surface/subsurface case: the BASELINE 512x512 surface over 512x512x30 soil."""

    def __init__(self, nx=512, ny=512, nz=30, dx=1.0, dz=0.05, seed_bed=1, seed_ks=2, slope=0.01,
                 bed_sigma=0.05, bed_corr=20.0, ks_median=1e-5, ks_sigma_ln=1.0, ks_corr=10.0,
                 alpha=2.0, n=1.8, theta_s=0.40, theta_r=0.05, ss=1e-5, h0=-2.0, manning=0.03,
                 rain=RAIN_50MM_PER_H, dt=1.0, t_end=1800.0):
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

    def re_params(self, precision=abi.PREC_FP64, guess=abi.GUESS_ZERO, rtol=1e-10, inner_rtol=1e-5):
        return abi.re_params(soil=self.soil, solver=abi.solver_config(rtol, precision=precision, inner_rtol=inner_rtol,
                                                                      max_refine=8),
                             bc_type=self.bc, dx=self.dx, dy=self.dx, dz=self.dz, botz=-self.dz * self.nz, dt=self.dt,
                             adaptive_dt=0, use_full3d=1, initial_guess=guess)

    @property
    def cells(self):
        return self.nx * self.ny * self.nz + self.nx * self.ny

    def bed_with_offset(self):
        return (self.bed + self.swe.offset).ravel()

    def setup_gpu(self, ctx, precision=abi.PREC_FP64, guess=abi.GUESS_ZERO):
        ctx.swe_init(self.nx, self.ny, self.dx, self.dx, self.swe)
        ctx.swe_set_bathymetry(self.bed_with_offset())
        ctx.swe_initialize_state(-1.0e3)  # dry start: eta clamps to the bed
        ctx.re_init(self.nx, self.ny, self.nz, self.re_params(precision, guess))
        ctx.re_set_soil([self.soil], None, True, self.ks.ravel())
        ctx.re_set_initial(h=self.h0)
