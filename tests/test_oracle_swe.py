"""Pins the oracle's shallow-water step (oracle/swe_oracle.cpp, restating src/swe/SweSolver.cpp)
against the reference's golden vector benchmarks/b1-sw/frehg2_np1_depth_reference.txt
(tests/swe/test_swe_b1_mpi.cpp:54-90, gate rel-L2 < 1e-10), the lake-at-rest test
(tests/swe/test_swe_b0.cpp:22-77) and the x<->y transpose symmetry of the scheme."""
import numpy as np

import oracle_lib as ol
from cases import b1_inputs, flow_2d_inputs, gaussian_bump_bed, interior2, interp_series
from frehg_amd import abi


def test_b1_sw_golden_vector():
    bath, rain, ref, offset, p = b1_inputs()
    s = ol.OracleSwe(1, 10, 80.0, 80.0, p)
    s.set_bathymetry(bath + offset)
    s.initialize_state(-2.0)
    t, worst = 0.0, 0.0
    for step in range(1, 3601):
        t += 5.0
        s.step(interp_series(rain[:, 0], rain[:, 1], t), 0.0)
        if step % 360 == 0:
            row = ref[step // 360 - 1]
            assert int(row[0]) == int(round(t))
            d = interior2(s.get("dept"), 1, 10)[:, 0]
            worst = max(worst, np.linalg.norm(d - row[1:]) / np.linalg.norm(row[1:]))
    assert worst < 1e-10


def test_b0_lake_at_rest():
    p = abi.swe_params(manning=0.02, dt=1.0, solver=abi.solver_config(1e-13))
    s = ol.OracleSwe(20, 20, 1.0, 1.0, p)
    s.set_bathymetry(gaussian_bump_bed().ravel())
    s.initialize_state(1.0)
    for _ in range(200):
        s.step(0.0, 0.0)
    assert np.abs(interior2(s.get("eta"), 20, 20) - 1.0).max() < 1e-8
    assert np.abs(interior2(s.get("uu"), 20, 20)).max() < 1e-8
    assert np.abs(interior2(s.get("vv"), 20, 20)).max() < 1e-8


def _flow_case(bed, eta0, dx, dy, visc_x, visc_y, steps=60):
    ny, nx = bed.shape
    p = abi.swe_params(manning=0.03, dt=0.2, visc_x=visc_x, visc_y=visc_y, solver=abi.solver_config(1e-13))
    s = ol.OracleSwe(nx, ny, dx, dy, p)
    s.set_bathymetry(bed.ravel())
    s.set_initial_eta(eta0.ravel())
    umax = 0.0
    for _ in range(steps):
        s.step(0.0, 0.0)
        umax = max(umax, np.abs(interior2(s.get("uu"), nx, ny)).max(), np.abs(interior2(s.get("vv"), nx, ny)).max())
    return s, umax


def test_transpose_symmetry_with_flow():
    """The scheme is the same in x and y (src/swe/SweSolver.cpp:159-792): a rain-free problem
    solved on the transposed grid, with dx <-> dy and the x/y viscosities swapped, gives the
    transposed depth and swaps u and v, to round-off. The y-direction terms are pinned by the
    b1-sw golden vector (nx = 1); this carries the pin over to every x-direction term in geometry,
    momentum, matrix, velocity and limiter, in a wet/dry run with velocities near 1 m/s. A wrong
    index, area or spacing in an x-direction term breaks the symmetry."""
    bed, eta0 = flow_2d_inputs()
    m, n = bed.shape
    a, umax = _flow_case(bed, eta0, 1.0, 1.5, 1e-2, 3e-2)
    b, _ = _flow_case(bed.T.copy(), eta0.T.copy(), 1.5, 1.0, 3e-2, 1e-2)
    assert umax > 0.3
    da, db = interior2(a.get("dept"), n, m), interior2(b.get("dept"), m, n)
    assert (da <= 0).sum() > 10 and (da > 0).sum() > 10  # wet and dry cells
    np.testing.assert_allclose(da, db.T, rtol=0, atol=1e-12)
    for fa, fb in (("uu", "vv"), ("vv", "uu"), ("CDx", "CDy"), ("Fu", "Fv"), ("cflx", "cfly")):
        # relative: the drag coefficient of a film a few mm deep magnifies depth round-off
        np.testing.assert_allclose(interior2(a.get(fa), n, m), interior2(b.get(fb), m, n).T, rtol=1e-8,
                                   atol=1e-11, err_msg=fa)
    np.testing.assert_allclose(interior2(a.get("eta"), n, m), interior2(b.get("eta"), m, n).T, atol=1e-12)


def test_transpose_symmetry_still_water():
    """Still water over an i <-> j symmetric bed stays symmetric (and at rest)."""
    n = 8
    bs = 0.1 * np.add.outer(np.arange(n), np.arange(n)) / n
    p = abi.swe_params(manning=0.03, dt=0.5)
    s = ol.OracleSwe(n, n, 1.0, 1.0, p)
    s.set_bathymetry(bs.ravel())
    s.initialize_state(0.5)
    for _ in range(30):
        s.step(0.0, 0.0)
    d = interior2(s.get("dept"), n, n)
    np.testing.assert_allclose(d, d.T, atol=1e-12)
    u = interior2(s.get("uu"), n, n)
    v = interior2(s.get("vv"), n, n)
    np.testing.assert_allclose(u, v.T, atol=1e-12)
