"""Pins the oracle's polygon boundary regions (oracle/swe_oracle.cpp PolygonBc, restating
src/bc/Polygon.hpp, src/bc/PolygonIndex.cpp:9-56 and src/bc/PolygonBC.cpp:16-78) with hand-worked
cases: cell-centre membership of a concave polygon (first match wins), each rule's closed form
(critical-depth weir, prescribed depth, discharge spread over the region) with its volume limit,
and the water budget of a coupled run with an outflow edge."""
import numpy as np
import pytest

import oracle_lib as ol
from cases import interior2, interior3
from frehg_amd import abi

G = 9.81


def basin(nx=6, ny=5, dx=2.0, dy=1.0, eta0=0.3, dt=0.5):
    s = ol.OracleSwe(nx, ny, dx, dy, abi.swe_params(dt=dt, manning=0.03))
    s.set_bathymetry(np.zeros(nx * ny))
    s.initialize_state(eta0)
    return s


def test_membership_concave_polygon_first_match():
    """Cell centres (x0 + (i + 1/2) dx, y0 + (j + 1/2) dy) inside an L-shaped ring take its rule;
    a second polygon overlapping it only gets the cells the first did not take."""
    s = basin()
    x0, y0 = 10.0, 20.0
    L = [(x0 + 0, y0 + 0), (x0 + 6.5, y0 + 0), (x0 + 6.5, y0 + 1.8), (x0 + 2.2, y0 + 1.8), (x0 + 2.2, y0 + 5),
         (x0 + 0, y0 + 5)]
    box = [(x0 + 0, y0 + 0), (x0 + 12, y0 + 0), (x0 + 12, y0 + 1.2), (x0 + 0, y0 + 1.2)]
    # L covers: row j = 0 (yc = 0.5) and j = 1 (yc = 1.5): i with xc = 1, 3, 5 (< 6.5); rows
    # j = 2..4: only i = 0 (xc = 1 < 2.2). The box covers row j = 0 only, all of which L already
    # took, except i = 3..5 (xc = 7, 9, 11 > 6.5).
    s.set_boundary_polygons([(1, 0.05, L), (1, 0.10, box)], x0, y0)
    s.apply_boundaries(0.5)
    eta = interior2(s.get("eta"), 6, 5)
    want = np.full((5, 6), 0.3)
    for j, i in [(0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (1, 2), (2, 0), (3, 0), (4, 0)]:
        want[j, i] = 0.05
    for i in (3, 4, 5):
        want[0, i] = 0.10
    np.testing.assert_array_equal(eta, want)


def test_rules_closed_form():
    """One application over dt per rule on a basin at eta = 0.3 (bed 0, area 2 m^2 per cell):
    critical q = sqrt(g) h^1.5 min(dx, dy), prescribed depth, discharge Q dt / ncells; the
    returned net outflow is the volume removed."""
    dt = 0.5
    strip = [(0, 0), (2.0, 0), (2.0, 5.0), (0, 5.0)]  # column i = 0 (5 cells)
    # critical
    s = basin(dt=dt)
    s.set_boundary_polygons([(2, 0.0, strip)])
    out = s.apply_boundaries(dt)
    vol = np.sqrt(G) * 0.3 ** 1.5 * min(2.0, 1.0) * dt
    eta = interior2(s.get("eta"), 6, 5)
    np.testing.assert_allclose(eta[:, 0], 0.3 - vol / 2.0, rtol=1e-15)
    np.testing.assert_array_equal(eta[:, 1:], 0.3)
    assert out == pytest.approx(5 * vol, rel=1e-14)
    np.testing.assert_allclose(interior2(s.get("dept"), 6, 5)[:, 0], 0.3 - vol / 2.0, rtol=1e-15)
    # critical, limited by the ponded volume: a long interval drains the cells to the bed
    s = basin(dt=dt)
    s.set_boundary_polygons([(2, 0.0, strip)])
    out = s.apply_boundaries(100.0)
    np.testing.assert_array_equal(interior2(s.get("eta"), 6, 5)[:, 0], 0.0)
    assert out == pytest.approx(5 * 0.3 * 2.0, rel=1e-14)
    # prescribed depth (adds water where it is deeper than the cell's)
    s = basin(dt=dt)
    s.set_boundary_polygons([(1, 0.45, strip)])
    out = s.apply_boundaries(dt)
    np.testing.assert_array_equal(interior2(s.get("eta"), 6, 5)[:, 0], 0.45)
    assert out == pytest.approx(-5 * 0.15 * 2.0, rel=1e-13)
    # discharge Q = 0.4 m^3/s over the 5 cells
    s = basin(dt=dt)
    s.set_boundary_polygons([(0, 0.4, strip)])
    out = s.apply_boundaries(dt)
    np.testing.assert_allclose(interior2(s.get("eta"), 6, 5)[:, 0], 0.3 - 0.4 * dt / 5 / 2.0, rtol=1e-15)
    assert out == pytest.approx(0.4 * dt, rel=1e-14)


def test_outflow_volume_accounting_on_a_flowing_surface():
    """Surface steps on a tilted bed with rain, each followed by a critical-depth outflow strip
    along the low-x edge (as the coupled step applies it): every application removes exactly the
    volume it reports from the surface, the edge drains a good part of the rain, and the depth and
    geometry are refreshed (the next step starts from eta - bottom)."""
    nx, ny, dx = 12, 8, 1.0
    bed = 0.02 * np.arange(nx)[None, :] + np.zeros((ny, 1))
    s = ol.OracleSwe(nx, ny, dx, dx, abi.swe_params(manning=0.03, dt=1.0))
    s.set_bathymetry(bed.ravel())
    s.initialize_state(0.05)
    s.set_boundary_polygons([(2, 0.0, [(-1.0, -1.0), (1.0, -1.0), (1.0, ny + 1.0), (-1.0, ny + 1.0)])])

    def surface():
        return np.maximum(interior2(s.get("eta") - s.get("bottom"), nx, ny), 0).sum() * dx * dx

    total = 0.0
    for _ in range(40):
        s.step(2e-4, 0.0)
        before = surface()
        out = s.apply_boundaries(1.0)
        assert before - surface() == pytest.approx(out, rel=1e-12, abs=1e-15)
        total += out
        d = interior2(s.get("dept"), nx, ny)
        np.testing.assert_allclose(d, np.maximum(interior2(s.get("eta") - s.get("bottom"), nx, ny), 0), atol=1e-15)
    assert total > 0.05 * 2e-4 * 40 * nx * (ny - 1)


def test_regions_end_with_their_solver():
    """A solver's regions are dropped when it is destroyed: fresh solvers (which the allocator
    often places at a freed solver's address) start without regions, so the coupled step and
    apply_boundaries leave them unchanged."""
    box = [(-1.0, -1.0), (3.0, -1.0), (3.0, 9.0), (-1.0, 9.0)]
    for _ in range(8):
        s = basin()
        s.set_boundary_polygons([(abi.BC_DEPTH, 0.05, box)])
        assert s.apply_boundaries(0.5) != 0.0
        del s
        fresh = basin()
        eta = fresh.get("eta").copy()
        assert fresh.apply_boundaries(0.5) == 0.0
        np.testing.assert_array_equal(fresh.get("eta"), eta)
        del fresh
