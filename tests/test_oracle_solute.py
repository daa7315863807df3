"""The solute oracle pinned by the reference's own solute tests (tests/solute/test_advection.cpp,
test_diffusion.cpp, test_source_sink.cpp, test_solute_stepper.cpp) and by closed forms: the
unit-Courant upwind step is an exact one-cell shift in each direction and sign, the discrete
Neumann Laplacian has cosine eigenvectors with eigenvalue -(4/h^2) sin^2(pi m / 2n), implicit
diffusion of a step approaches 0.5 erfc, decay is exp(-k t), rain mixing has c_rain as its fixed
point, the interface exchange conserves sum C * water, and MUSCL with the minmod limiter is
total-variation diminishing and more accurate than upwind on a smooth pulse at small Courant
number."""
import math

import numpy as np
import pytest

import oracle_lib as ol
from cases import interior3
from frehg_amd import abi, workloads
from solute_cases import Flow, gaussian, inner2, inner3, pad2, pad3, params


def surface_1d(n, u, depth=1.0, along="x", dx=1.0):
    if along == "x":
        return Flow(n, 1, dx=dx, u=u, depth=depth)
    return Flow(1, n, dy=dx, v=u, depth=depth)


def line(sol, f, along="x"):
    a = inner2(sol.get("surface"), f.nx, f.ny)
    return a[0, :] if along == "x" else a[:, 0]


def set_line(sol, f, vals, along="x"):
    a = vals[None, :] if along == "x" else vals[:, None]
    sol.set("surface", pad2(a, f.nx, f.ny))


@pytest.mark.parametrize("along", ["x", "y"])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_unit_courant_upwind_is_an_exact_shift(along, sign):
    """test_advection.cpp:40-102: 6 steps at Courant 1 move the pulse 6 cells (inflow is 0)."""
    n, steps = 24, 6
    f = surface_1d(n, sign * 1.0, along=along)
    _, _, sol = f.oracle(params())
    x0 = 6.0 if sign > 0 else n - 7.0
    c0 = gaussian(np.arange(n), x0, 1.6)
    set_line(sol, f, c0, along)
    for _ in range(steps):
        cfl, ok = sol.step(1.0)
        assert ok and cfl == pytest.approx(1.0, abs=1e-14)
    idx = np.arange(n) - sign * steps
    want = np.where((idx >= 0) & (idx < n), gaussian(idx, x0, 1.6), 0.0)
    np.testing.assert_allclose(line(sol, f, along), want, rtol=0, atol=1e-12)


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_unit_courant_shift_in_z_through_the_darcy_flux(sign):
    """The subsurface scheme with qz at Courant 1 (z increases downward, qz > 0 moves down)."""
    nz, dz, steps = 40, 0.05, 5
    f = Flow(1, 1, nz, dz=dz, surface=False, qz=sign * dz)
    _, _, sol = f.oracle(params())
    z0 = 10.0 if sign > 0 else nz - 11.0
    c0 = gaussian(np.arange(nz), z0, 1.5)
    sol.set("subsurface", pad3(c0[:, None, None], 1, 1, nz))
    for _ in range(steps):
        cfl, ok = sol.step(1.0)
        assert ok and cfl == pytest.approx(1.0, abs=1e-14)
    idx = np.arange(nz) - sign * steps
    want = np.where((idx >= 0) & (idx < nz), gaussian(idx, z0, 1.5), 0.0)
    np.testing.assert_allclose(inner3(sol.get("subsurface"), 1, 1, nz)[:, 0, 0], want, rtol=0, atol=1e-12)


def test_cfl_refusal_leaves_the_field_unchanged():
    """test_advection.cpp:104-125 and test_solute_stepper.cpp:138-154."""
    f = surface_1d(16, 1.0)
    _, _, sol = f.oracle(params())
    c0 = gaussian(np.arange(16), 8.0, 2.0)
    set_line(sol, f, c0)
    before = sol.get("surface")
    cfl, ok = sol.step(1.5)
    assert not ok and cfl == pytest.approx(1.5)
    assert np.array_equal(sol.get("surface"), before)


@pytest.mark.parametrize("muscl", [0, 1])
def test_closed_domain_advection_conserves_the_cell_sum(muscl):
    """test_advection.cpp:127-142 (Courant 0.7, 10 steps), and a 2D and a 3D field of random
    face velocities."""
    f = surface_1d(30, 0.7)
    _, _, sol = f.oracle(params(muscl=muscl))
    set_line(sol, f, gaussian(np.arange(30), 8.0, 2.0))
    m0 = line(sol, f).sum()
    for _ in range(10):
        assert sol.step(1.0)[1]
    assert abs(line(sol, f).sum() - m0) < 1e-12 * m0
    rng = np.random.default_rng(5)
    nx, ny, nz = 13, 11, 7
    f2 = Flow(nx, ny, nz, dx=2.0, dy=1.5, dz=0.1, u=rng.uniform(-0.4, 0.4, (ny, nx)),
              v=rng.uniform(-0.3, 0.3, (ny, nx)), qx=rng.uniform(-0.3, 0.3, (nz, ny, nx)),
              qy=rng.uniform(-0.2, 0.2, (nz, ny, nx)), qz=rng.uniform(-0.02, 0.02, (nz, ny, nx)))
    _, _, sol2 = f2.oracle(params(muscl=muscl))
    sol2.set("surface", pad2(rng.uniform(0, 1, (ny, nx)), nx, ny))
    sol2.set("subsurface", pad3(rng.uniform(0, 1, (nz, ny, nx)), nx, ny, nz))
    s0, g0 = inner2(sol2.get("surface"), nx, ny).sum(), inner3(sol2.get("subsurface"), nx, ny, nz).sum()
    for _ in range(8):
        cfl, ok = sol2.step(1.0)
        assert ok and 0.3 < cfl < 1.0
    assert abs(inner2(sol2.get("surface"), nx, ny).sum() - s0) < 1e-12 * s0
    assert abs(inner3(sol2.get("subsurface"), nx, ny, nz).sum() - g0) < 1e-12 * g0


def test_muscl_is_tvd_and_more_accurate_than_upwind():
    """Minmod MUSCL creates no new extrema and no new variation on a step at Courant 0.5, and on a
    smooth pulse carried 100 cells at Courant 0.1 its error is a third of first-order upwind's,
    whose error halves when the grid is refined. (The update is forward Euler on the
    reconstructed face values, as the reference's, so at larger Courant numbers its time error
    steepens smooth pulses; the test does not ask more of it.)"""
    n = 60
    f = surface_1d(n, 0.5)
    _, _, sol = f.oracle(params(muscl=1))
    c = np.where((np.arange(n) >= 10) & (np.arange(n) < 30), 1.0, 0.0)  # clear of both walls
    set_line(sol, f, c)
    tv0 = np.abs(np.diff(c)).sum()
    for _ in range(40):
        assert sol.step(1.0)[1]
        cc = line(sol, f)
        assert cc.max() <= 1.0 + 1e-15 and cc.min() >= -1e-15
        assert np.abs(np.diff(cc)).sum() <= tv0 + 1e-12

    def err(muscl, refine, nu=0.1):
        n = 300 * refine
        f = surface_1d(n, nu, dx=1.0 / refine)
        _, _, sol = f.oracle(params(muscl=muscl))
        x = (np.arange(n) + 0.5) / refine
        set_line(sol, f, gaussian(x, 60.0, 8.0))
        for _ in range(int(round(100 / nu)) * refine):
            assert sol.step(1.0 / refine)[1]
        return np.abs(line(sol, f) - gaussian(x, 160.0, 8.0)).max()

    e_up, e_up2, e_mu = err(0, 1), err(0, 2), err(1, 1)
    assert 1.4 < e_up / e_up2 < 2.2    # upwind converges at first order
    assert e_mu < e_up / 3.0           # the limited reconstruction removes most of its smearing


def diffuse_once(f, c_interior, D, dt, three):
    p = params(D=D, implicit_diffusion=1)
    _, _, sol = f.oracle(p)
    if three:
        sol.set("subsurface", pad3(c_interior, f.nx, f.ny, f.nz))
    else:
        sol.set("surface", pad2(c_interior, f.nx, f.ny))
    assert sol.step(dt)[1]
    return sol


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_implicit_diffusion_of_a_neumann_eigenvector(axis):
    """(I - dt D L) C' = C with C a cosine mode of the Neumann Laplacian along one axis:
    C' = C / (1 + dt D (4/h^2) sin^2(pi m / 2n))."""
    nx, ny, nz, dx, dy, dz, D, dt, m = 12, 10, 8, 2.0, 0.5, 0.05, 1e-3, 7.0, 3
    f = Flow(nx, ny, nz, dx=dx, dy=dy, dz=dz, surface=False)
    n, h = (nx, ny, nz)[axis], (dx, dy, dz)[axis]
    mode = np.cos(math.pi * m * (np.arange(n) + 0.5) / n)
    shape = [1, 1, 1]
    shape[2 - axis] = n
    c = np.broadcast_to(mode.reshape(shape), (nz, ny, nx))
    sol = diffuse_once(f, c, D, dt, True)
    lam = 4.0 / h ** 2 * math.sin(math.pi * m / (2 * n)) ** 2
    np.testing.assert_allclose(inner3(sol.get("subsurface"), nx, ny, nz), c / (1 + dt * D * lam), rtol=0, atol=1e-10)
    if axis < 2:
        f2 = Flow(nx, ny, dx=dx, dy=dy)
        c2 = np.broadcast_to(mode.reshape(shape[1:]), (ny, nx))
        sol2 = diffuse_once(f2, c2, D, dt, False)
        np.testing.assert_allclose(inner2(sol2.get("surface"), nx, ny), c2 / (1 + dt * D * lam), rtol=0, atol=1e-10)


def test_implicit_diffusion_of_a_step_approaches_erfc():
    """test_diffusion.cpp:24-65: a 200-cell column, D = 1e-5, 250 steps of 0.25 s."""
    nz, dz, D, dt, steps = 200, 0.01, 1e-5, 0.25, 250
    f = Flow(1, 1, nz, dz=dz, surface=False)
    _, _, sol = f.oracle(params(D=D, implicit_diffusion=1))
    c = np.where(np.arange(nz) < nz // 2, 1.0, 0.0)
    sol.set("subsurface", pad3(c[:, None, None], 1, 1, nz))
    for _ in range(steps):
        assert sol.step(dt)[1]
    z = (np.arange(nz) + 0.5) * dz
    from scipy.special import erfc
    want = 0.5 * erfc((z - nz // 2 * dz) / (2 * math.sqrt(D * dt * steps)))
    assert np.abs(inner3(sol.get("subsurface"), 1, 1, nz)[:, 0, 0] - want).max() < 2e-2


def test_neumann_diffusion_conserves_the_column_and_zero_d_is_a_no_op():
    """test_diffusion.cpp:67-123."""
    nz = 120
    f = Flow(1, 1, nz, dz=0.01, surface=False)
    p = params(D=5e-5, implicit_diffusion=1, solver=abi.solver_config(1e-13))
    _, _, sol = f.oracle(p)
    c = np.where(np.arange(nz) < nz // 2, 1.0, 0.0)
    sol.set("subsurface", pad3(c[:, None, None], 1, 1, nz))
    for _ in range(50):
        assert sol.step(0.5)[1]
    assert abs(inner3(sol.get("subsurface"), 1, 1, nz).sum() - c.sum()) < 1e-9 * c.sum()
    f0 = Flow(1, 1, 10, dz=0.01, surface=False)
    _, _, sol0 = f0.oracle(params(D=0.0, implicit_diffusion=1))
    c0 = pad3((np.arange(10) + 1.0)[:, None, None], 1, 1, 10)
    sol0.set("subsurface", c0)
    assert sol0.step(1.0)[1]
    assert np.array_equal(sol0.get("subsurface"), c0)


def test_advect_and_diffuse_pipeline_conserves_mass():
    """test_solute_stepper.cpp:156-187 (D = 1e-3, rtol 1e-13, Courant 0.5, 10 steps)."""
    f = surface_1d(24, 0.5)
    _, _, sol = f.oracle(params(D=1e-3, implicit_diffusion=1, solver=abi.solver_config(1e-13)))
    set_line(sol, f, gaussian(np.arange(24), 8.0, 2.0))
    m0 = line(sol, f).sum()
    for _ in range(10):
        assert sol.step(1.0)[1]
    assert abs(line(sol, f).sum() - m0) < 1e-8 * m0


def test_decay_follows_exp_minus_kt_and_is_booked():
    """test_source_sink.cpp:85-98 and test_solute_stepper.cpp:114-136 through applySolute: the
    mass falls as exp(-k t) under a conservative flow, and the decay loss is the mass lost."""
    f = surface_1d(30, 0.5)
    k, dt, n = 0.05, 1.0, 15
    _, _, sol = f.oracle(params(k_decay=k))
    set_line(sol, f, gaussian(np.arange(30), 10.0, 2.5))
    m0 = sol.mass()[0]
    for s in range(n):
        sol.apply(s * dt, dt)
    m, st = sol.mass()
    assert m == pytest.approx(m0 * math.exp(-k * dt * n), rel=1e-10)
    assert st.decay_loss == pytest.approx(m0 - m, rel=1e-10)
    g = surface_1d(1, 0.0)
    _, _, one = g.oracle(params(k_decay=0.1))
    one.set("surface", pad2(7.0, 1, 1))
    for s in range(1, 21):
        one.apply((s - 1) * 2.0, 2.0)
        assert one.get("surface")[4] == pytest.approx(7.0 * math.exp(-0.1 * 2.0 * s), abs=1e-10)


@pytest.mark.parametrize("c_rain", [5.0, 35.0])
def test_rain_drives_the_surface_to_c_rain(c_rain):
    """test_source_sink.cpp:20-38 through applySolute: 4000 steps of 1 mm rain on 0.1 m of water.
    The last row gets no rain (the surface step gives it none) and neither do dry cells."""
    depth = np.array([[0.1, 0.0, 1e-3 + 5e-9], [0.1, 0.1, 0.1]])
    f = Flow(3, 2, depth=depth)
    _, _, sol = f.oracle(params(c_rain=c_rain))
    for s in range(4000):
        sol.apply(float(s), 1.0, 1e-3)
    c = inner2(sol.get("surface"), 3, 2)
    assert c[0, 0] == pytest.approx(c_rain, abs=1e-3)
    assert c[0, 1] == 0.0 and c[0, 2] == 0.0 and np.all(c[1] == 0.0)
    st = sol.mass()[1]
    assert st.rain_added == pytest.approx(4000 * 1e-3 * c_rain, rel=1e-12)
    sol.set("surface", pad2(np.array([[c_rain, 0, 0], [0, 0, 0]]), 3, 2))
    sol.apply(4000.0, 1.0, 1e-3)
    assert inner2(sol.get("surface"), 3, 2)[0, 0] == pytest.approx(c_rain, abs=1e-12)


def test_start_time_defers_the_solute():
    f = Flow(3, 2, u=0.5, depth=0.1)
    _, _, sol = f.oracle(params(c_rain=3.0, k_decay=0.1, start_time=5.0))
    sol.set("surface", pad2(1.0, 3, 2))
    before = sol.get("surface")
    sol.apply(4.0, 1.0, 1e-3)
    assert np.array_equal(sol.get("surface"), before)
    sol.apply(5.0, 1.0, 1e-3)
    assert not np.array_equal(sol.get("surface"), before)


def test_substeps_double_until_the_courant_number_passes():
    """applySolute's retry: Courant 2.6 at the flow step is accepted at 4 substeps (0.65), and the
    result equals four stepper steps of dt/4; Courant above cfl_max at 4096 substeps is refused
    and leaves the solute unchanged."""
    f = surface_1d(40, 2.6)
    c0 = gaussian(np.arange(40), 10.0, 3.0)
    _, _, sol = f.oracle(params(muscl=1))
    set_line(sol, f, c0)
    sol.apply(0.0, 1.0)
    st = sol.mass()[1]
    assert st.substeps == 4 and st.max_courant == pytest.approx(0.65) and st.refused_steps == 0
    _, _, ref = f.oracle(params(muscl=1))
    set_line(ref, f, c0)
    for _ in range(4):
        assert ref.step(0.25)[1]
    assert np.array_equal(sol.get("surface"), ref.get("surface"))
    g = surface_1d(8, 5000.0)
    _, _, fast = g.oracle(params())
    set_line(fast, g, np.ones(8))
    before = fast.get("surface")
    fast.apply(0.0, 1.0)
    assert fast.mass()[1].refused_steps == 1
    assert np.array_equal(fast.get("surface"), before)


CASES = [  # cs, cg, surface depth before, soil water before, exchanged height (test_source_sink.cpp:48-53)
    (10.0, 0.0, 0.60, 0.040, 0.010),
    (2.0, 7.0, 0.50, 0.100, 0.030),
    (3.0, 9.0, 0.40, 0.200, -0.050),
    (5.0, 1.0, 0.30, 0.150, -0.020),
]


@pytest.mark.parametrize("cs,cg,ds0,hg0,vex", CASES)
def test_interface_exchange_conserves_solute(cs, cg, ds0, hg0, vex):
    """test_source_sink.cpp:40-83: after a water move of vex (positive into the soil) the
    exchange conserves sum C * water, and each concentration stays between the two."""
    dz, dt, area = 0.5, 2.0, 1.0
    f = Flow(1, 1, 1, dz=dz, depth=ds0 - vex, wc=(hg0 + vex) / dz)
    _, _, sol = f.oracle(params())
    sol.set("surface", pad2(cs, 1, 1))
    sol.set("subsurface", pad3(cg, 1, 1, 1))
    sol.exchange(np.array([-vex * area / dt]), dt, area)
    s, g = sol.get("surface")[4], sol.get("subsurface")[13]
    assert s * (ds0 - vex) + g * (hg0 + vex) == pytest.approx(cs * ds0 + cg * hg0, abs=1e-12)
    lo, hi = min(cs, cg), max(cs, cg)
    assert lo - 1e-12 <= s <= hi + 1e-12 and lo - 1e-12 <= g <= hi + 1e-12


def test_infiltration_carries_surface_solute_into_the_top_soil():
    """A coupled step on ponded water over dry soil: the infiltrating water takes the surface
    concentration into the top soil cells (the exchange's sign), and the soil away from the
    solute-carrying water stays clean."""
    case = workloads.CoupledCase(nx=12, ny=10, nz=6, seed_bed=3, seed_ks=4, rain=0.0)
    nx, ny, nz = case.nx, case.ny, case.nz
    s = ol.OracleSwe(nx, ny, case.dx, case.dx, case.swe)
    s.set_bathymetry(case.bed_with_offset())
    wet = np.zeros((ny, nx), dtype=bool)
    wet[:, : nx // 2] = True
    s.set_initial_eta(case.bed.reshape(ny, nx) + np.where(wet, 0.05, 0.0))
    r = ol.OracleRe(nx, ny, nz, case.re_params(picard_max=20))
    r.set_soil([case.soil], None, True, case.ks.ravel())
    r.set_initial(h=case.h0)
    sol = ol.OracleSolute(s, r, params(muscl=1))
    sol.set("surface", pad2(np.where(wet, 1.0, 0.0), nx, ny))
    oc = ol.OracleCoupled(s, r, case.cp, case.t_end, solute=sol)
    oc.step(case.dt, 0.0)
    assert sol.mass()[1].substeps == 1
    top = interior3(sol.get("subsurface"), nx, ny, nz)[0]
    assert np.all(top[wet] > 0.0) and np.all(top <= 1.0 + 1e-6)
    # one substep moves soil solute at most one column: beyond that nothing but water arrived
    assert np.all(top[:, nx // 2 + 1:] == 0.0)
    assert np.all(interior3(sol.get("subsurface"), nx, ny, nz)[2:] == 0.0)
