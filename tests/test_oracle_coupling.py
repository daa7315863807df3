"""Pins the oracle's SW<->GW exchange (oracle/swe_oracle.cpp couple_exchange, restating
src/coupling/Coupling.cpp:25-164) with the reference's conservation gate
(tests/coupling/test_coupled_timestep.cpp:64-121) and the sign / donor-limit special cases."""
import numpy as np

import oracle_lib as ol
from cases import interior2, interior3
from frehg_amd import abi


def setup(nx=10, ny=10, nz=5, eta0=0.5, wc0=0.2):
    dx = dy = 10.0
    sp = abi.swe_params(dt=1.0)
    s = ol.OracleSwe(nx, ny, dx, dy, sp)
    s.set_bathymetry(np.zeros(nx * ny))
    s.initialize_state(eta0)
    rp = abi.re_params(soil=abi.soil_params(theta_s=0.33, theta_r=0.0), dx=dx, dy=dy, dz=0.1)
    r = ol.OracleRe(nx, ny, nz, rp)
    r.set_initial(wc=np.full(nx * ny * nz, wc0))
    return s, r, abi.coupling_params(dx=dx, dy=dy)


def volumes(s, r, nx=10, ny=10, nz=5):
    sw = np.maximum(interior2(s.get("eta") - s.get("bottom"), nx, ny), 0).sum() * 100.0
    gw = interior3(r.get("wc"), nx, ny, nz).sum() * 100.0 * 0.1
    return sw, gw


def test_exchange_conserves_exactly():
    s, r, cp = setup()
    sw0, gw0 = volumes(s, r)
    vol, q = ol.couple_exchange(s, r, cp, 1.0, want_q=True)
    sw1, gw1 = volumes(s, r)
    assert abs(vol) > 1e-3
    assert abs((sw1 - sw0) - vol) < 1e-9
    assert abs((gw1 - gw0) + vol) < 1e-9
    assert abs((sw1 - sw0) + (gw1 - gw0)) < 1e-9
    assert (q < 0).all()  # ponded water over unsaturated soil infiltrates


def test_donor_limits():
    # dry surface: no infiltration (q >= 0); saturated top cell cannot take more water
    s, r, cp = setup(eta0=0.0)
    vol, q = ol.couple_exchange(s, r, cp, 1.0, want_q=True)
    assert (q >= 0).all()
    s, r, cp = setup(eta0=0.5, wc0=0.33)
    vol, q = ol.couple_exchange(s, r, cp, 1.0, want_q=True)
    assert np.all(q >= -1e-300)  # headroom (theta_s - theta) V / dt = 0 caps infiltration
    # surface-limited infiltration: a thin film cannot lose more than its volume
    s, r, cp = setup(eta0=1e-4, wc0=0.05)
    vol, q = ol.couple_exchange(s, r, cp, 1.0, want_q=True)
    np.testing.assert_allclose(q, -1e-4 * 100.0 / 1.0, rtol=1e-12)
    assert interior2(s.get("dept"), 10, 10).max() < 1e-15
