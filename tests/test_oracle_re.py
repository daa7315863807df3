"""Pins the oracle's Richards PCA step (oracle/re_oracle.cpp, restating src/re/ReSolver.cpp)
with the reference's RE tests (tests/re/test_re_3d.cpp:133-274, test_re_b2_gw.cpp:109-153)
and a brute-force 1x1x3 column solved by dense linear algebra."""
import csv
import os

import numpy as np
import pytest

import oracle_lib as ol
from cases import box3d_params, box3d_soil, box3d_wc, interior3
from frehg_amd import abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

NX, NY, NZ = 4, 4, 8
V = 2.0 * 1.0 * 0.05


def run_box(params, steps=40, soil_map=None):
    r = ol.OracleRe(NX, NY, NZ, params)
    if soil_map is not None:
        r.set_soil(*soil_map)
    r.set_initial(wc=box3d_wc())
    t0 = interior3(r.get("wc"), NX, NY, NZ).sum() * V
    for _ in range(steps):
        r.step()
    w = interior3(r.get("wc"), NX, NY, NZ).copy()
    return t0, w.sum() * V, w


def test_closed_box_conserves():  # test_re_3d.cpp:133-150
    soil = box3d_soil()
    t0, t1, w = run_box(box3d_params(soil))
    assert abs(t1 - t0) / t0 < 1e-10
    assert w.min() >= soil.theta_r - 1e-12 and w.max() <= soil.theta_s + 1e-12


def test_lateral_activity():  # test_re_3d.cpp:152-169
    a0, a1, wl = run_box(box3d_params(box3d_soil()))
    b0, b1, wv = run_box(box3d_params(box3d_soil(ks=(0.0, 0.0, 1e-5))))
    assert np.abs(wl - wv).max() > 1e-7
    assert abs(a1 - a0) / a0 < 1e-10 and abs(b1 - b0) / b0 < 1e-10


def test_single_class_map_bit_identical():  # test_re_3d.cpp:171-190
    soil = box3d_soil()
    _, _, wu = run_box(box3d_params(soil))
    idx = np.zeros(NX * NY * NZ, dtype=np.int32)
    _, _, wm = run_box(box3d_params(soil), soil_map=([soil], idx, True, None))
    assert np.array_equal(wu, wm)


def test_layered_map_differs():  # test_re_3d.cpp:192-218
    fast, slow = box3d_soil(ks=(5e-5,) * 3), box3d_soil(ks=(2e-6,) * 3)
    idx = np.zeros((NZ, NY, NX), dtype=np.int32)
    idx[NZ // 2:] = 1
    _, _, wf = run_box(box3d_params(fast))
    _, _, wm = run_box(box3d_params(fast), soil_map=([fast, slow], idx.ravel(), True, None))
    assert np.abs(wf - wm).max() > 1e-6


def test_per_cell_ks_equals_one_class_per_cell():
    # ks_cell is the reference's one-class-per-cell SoilMap (src/soil/SoilMap.hpp:33-46)
    rng = np.random.default_rng(3)
    ks = 1e-5 * np.exp(rng.standard_normal(NX * NY * NZ))
    base = box3d_soil()
    classes = [abi.soil_params(alpha=2.0, n=1.6, theta_s=0.45, theta_r=0.05, Ks_x=k, Ks_y=k, Ks_z=k, Ss=0.0)
               for k in ks]
    _, _, wa = run_box(box3d_params(base), soil_map=([base], None, True, ks))
    _, _, wb = run_box(box3d_params(base), soil_map=(classes, np.arange(NX * NY * NZ, dtype=np.int32), True, None))
    assert np.array_equal(wa, wb)


def test_qbot_adds_prescribed_mass():  # test_re_3d.cpp:220-274
    soil = box3d_soil()
    t0, t1, w = run_box(box3d_params(soil, bc=(0, 0, 0, 0, 2, 0), qbot=1e-7))
    expected = 1e-7 * 2.0 * 1.0 * 5.0 * 40 * NX * NY
    assert (t1 - t0) > 0 and abs((t1 - t0) - expected) / expected < 1e-6
    assert np.isfinite(w).all()


def test_b2gw_one_step_wets_top():  # test_re_b2_gw.cpp:109-153
    soil = abi.soil_params(alpha=1.43, n=1.56, theta_s=0.33, Ks_z=2.89e-6, Ss=1e-5)
    p = abi.re_params(soil=soil, solver=abi.solver_config(1e-8), bc_type=(0, 0, 0, 0, 0, 1), dz=0.01,
                      dt=1e-4, htop=0.0)
    r = ol.OracleRe(1, 1, 100, p)
    r.set_initial(wc=np.full(100, 0.033))
    c = (1 + 1 * 3) + 3 * 3  # idx(0,0,0)
    wc0, h0 = r.get("wc")[c], r.get("h")[c]
    r.step()
    assert r.get("h")[c] > h0 + 1e-6 and r.get("wc")[c] > wc0 + 1e-8
    assert r.dt() == pytest.approx(1e-4 * 1.25, rel=1e-14)  # adaptive ramp (ReSolver.cpp:921-927)


def test_b2gw_adaptive_dt_ramp_and_plateau():
    """tests/re/test_re_b2_gw.cpp:338-397: the b2-gw column with adaptive dt run to t = 46800 s.
    The legacy dt file the reference compares the first 12 steps with is absent; its content there
    is the deterministic x1.25 ramp, so dt_k = 1e-4 * 1.25^k is asserted as a closed form. Every dt
    stays in [dt_min, dt_max] and the run ends on the dt_max plateau."""
    soil = abi.soil_params(alpha=1.43, n=1.56, theta_s=0.33, Ks_z=2.89e-6, Ss=1e-5)
    p = abi.re_params(soil=soil, solver=abi.solver_config(1e-10), bc_type=(0, 0, 0, 0, 0, 1), dz=0.01,
                      dt=1e-4, dt_min=1e-4, dt_max=2.0, co_max=2.0, adaptive_dt=1, htop=0.0)
    r = ol.OracleRe(1, 1, 100, p)
    r.set_initial(wc=np.full(100, 0.033))
    t, dts, snap = 0.0, [], {}
    while t < 46800.0 + 1e-9:
        used = r.dt()
        st = r.step()
        assert st.dt_used == used
        dts.append(used)
        t += st.dt_used
        assert 1e-4 - 1e-300 <= r.dt() <= 2.0 + 1e-12
        for ts in (11700, 23400, 46800):
            if ts not in snap and t >= ts - 1e-6:
                snap[ts] = r.get("wc").reshape(102, 3, 3)[1:-1, 1, 1].copy()
    ramp = 1e-4 * 1.25 ** np.arange(12)
    np.testing.assert_allclose(dts[:12], ramp, rtol=1e-13)
    assert r.dt() == 2.0
    # below dt_max every step is x1.25 or a hold (0.01 <= dq_max <= 0.02 near the wetting front);
    # the one other ratio is the clamp to dt_max
    ratio = np.array(dts[1:]) / np.array(dts[:-1])
    below = np.array(dts[1:]) < 2.0
    assert np.all(np.isclose(ratio[below], 1.25, rtol=1e-12) | (ratio[below] == 1.0))
    assert np.sum(ratio[below] == 1.0) > 0
    k = int(np.argmax(np.array(dts) >= 2.0))
    assert all(d == 2.0 for d in dts[k:])
    # the Warrick profile (the reference's R-2 comparison, test_re_b2_gw.cpp:275-320, which it
    # prints without a gate): nine points on the wetting front at three times, the cell nearest
    # each elevation (kFromElevation). The oracle's front runs deeper than Warrick's (rel-L2 0.30);
    # the loose bound catches a grossly wrong infiltration rate (a front missing or at the bottom).
    with open(os.path.join(GOLDEN, "b2_gw", "warrick_water_content_profile.csv")) as fh:
        rows = list(csv.DictReader(fh))
    model, warrick = [], []
    for row in rows:
        z = float(row["z_m"])
        kz = int(np.argmin([abs(-0.01 - kk * 0.01 + 0.005 - z) for kk in range(100)]))
        model.append(snap[int(row["time_s"])][kz])
        warrick.append(float(row["water_content"]))
    model, warrick = np.array(model), np.array(warrick)
    rel = float(np.linalg.norm(model - warrick) / np.linalg.norm(warrick))
    assert len(rows) == 9 and rel < 0.4, rel


def _dt_statistics(r, p, nx, ny, nz, dt_prev):
    """The adaptive-dt statistics of the state after a step, computed here from the oracle's face
    fluxes and water contents: max over non-top cells of |net inflow per unit area| dt/dz, and
    min of co_max dz / (dK/dtheta) over the non-top unsaturated cells (ReSolver.cpp:890-914)."""
    sh = (nz + 2, ny + 2, nx + 2)
    qx, qy, qz, wc = (r.get(f).reshape(sh) for f in ("qx", "qy", "qz", "wc"))
    Ax, Ay, Az = p.dy * p.dz, p.dx * p.dz, p.dx * p.dy
    k, j, i = slice(2, nz + 1), slice(1, ny + 1), slice(1, nx + 1)  # k = 1..nz-1 (not the top layer)
    net = (qx[k, j, i] - qx[k, j, 0:nx]) / Ax + (qy[k, j, i] - qy[k, 0:ny, i]) / Ay + \
        (qz[k, j, i] - qz[1:nz, j, i]) / Az
    dq = float(np.abs(net).max() * dt_prev / p.dz)
    w = wc[k, j, i]
    w = w[w < p.soil.theta_s]
    dtco = float((p.co_max * p.dz / ol.vg_eval(p.soil, wc=w.ravel())["dKdwc"]).min()) if w.size else 1e8
    return dq, dtco


@pytest.mark.parametrize("case", ["ramp_reduce_hold", "courant_dtmin"])
def test_adaptive_dt_every_branch(case):
    """adaptiveTimeStep (ReSolver.cpp:880-929) on an infiltration box: the next dt is 0.75 dt when
    dq_max > 0.02, 1.25 dt when dq_max < 0.01, dt otherwise, then clamped to the Courant limit,
    dt_max and dt_min. Each rule is checked against statistics computed here from the fields, and
    the two cases make every branch fire (x0.75, hold, x1.25, Courant, dt_min)."""
    soil = abi.soil_params(alpha=3.35, n=2.0, theta_s=0.368, theta_r=0.102, Ks_x=9.22e-5, Ks_y=9.22e-5,
                           Ks_z=9.22e-5, Ss=1e-6)
    nx, ny, nz = 3, 2, 10
    kw = dict(dt=5.0, dt_min=1e-3, co_max=1e9) if case == "ramp_reduce_hold" else dict(dt=0.5, dt_min=0.4,
                                                                                        co_max=1e-4)
    p = abi.re_params(soil=soil, solver=abi.solver_config(1e-12), bc_type=(0, 0, 0, 0, 0, 1), htop=-0.1, dx=0.1,
                      dy=0.1, dz=0.02, dt_max=100.0, adaptive_dt=1, use_full3d=1, **kw)
    r = ol.OracleRe(nx, ny, nz, p)
    r.set_initial(h=np.full(nx * ny * nz, -3.0))
    fired = set()
    for _ in range(8):
        dt0 = r.dt()
        r.step()
        dq, dtco = _dt_statistics(r, p, nx, ny, nz, dt0)
        rule = dt0 * (0.75 if dq > 0.02 else (1.25 if dq < 0.01 else 1.0))
        fired.add("x0.75" if dq > 0.02 else ("x1.25" if dq < 0.01 else "hold"))
        expect = rule
        if expect > dtco:
            expect = dtco
            fired.add("courant")
        expect = min(expect, p.dt_max)
        if expect < p.dt_min:
            expect = p.dt_min
            fired.add("dt_min")
        assert r.dt() == pytest.approx(expect, rel=1e-12), (dt0, dq, dtco)
    want = {"x0.75", "hold", "x1.25"} if case == "ramp_reduce_hold" else {"x0.75", "courant", "dt_min"}
    assert want <= fired, fired


def _realloc_case(nx, nz, wc, h, kx_faces=(), bc5=0, htop=0.0, full3d=1):
    """Oracle with a hand-set state (i + k*nx order, ny = 1) and nonzero Kx on the listed +x faces."""
    soil = abi.soil_params(alpha=2.0, n=1.8, theta_s=0.40, theta_r=0.05, Ks_x=1e-5, Ks_y=1e-5, Ks_z=1e-5)
    p = abi.re_params(soil=soil, bc_type=(0, 0, 0, 0, 0, bc5), htop=htop, use_full3d=full3d, dx=1.0, dy=1.0, dz=0.1)
    r = ol.OracleRe(nx, 1, nz, p)
    sh = (nz + 2, 3, nx + 2)
    W = np.full(sh, 0.1)
    H = np.full(sh, -5.0)
    K = np.zeros(sh)
    for (i, k) in [(i, k) for k in range(nz) for i in range(nx)]:
        W[k + 1, 1, i + 1] = wc[k * nx + i]
        H[k + 1, 1, i + 1] = h[k * nx + i]
    for (i, k) in kx_faces:
        K[k + 1, 1, i + 1] = 1e-6
    r.set("wc", W.ravel())
    r.set("h", H.ravel())
    r.set("Kx", K.ravel())
    r.set("Ky", np.zeros(W.size))
    r.reallocate()
    g = lambda f: r.get(f).reshape(sh)[1:nz + 1, 1, 1:nx + 1].ravel()  # noqa: E731
    return soil, g("wc"), g("h")


def test_reallocation_column_branches_and_order():
    """reallocateWaterContent (ReSolver.cpp:788-837) on a hand-set 1x1x4 column, expected values
    worked out by hand. Branches: theta >= theta_s -> theta_s; no saturated neighbour -> h = h(theta);
    saturated neighbour -> theta = theta(h). The sweep runs top (k = 0) to bottom and reads the
    cells above it after their update: with a ponded top Dirichlet face the top cell takes
    theta(+0.01 m) = theta_s, and that saturates the neighbour test of the cell below it, which a
    sweep over the old values would not see."""
    ts = 0.40
    wc = [0.36, 0.38, 0.20, 0.42]
    h = [0.01, -1.0, -3.0, -0.3]
    soil, w, hh = _realloc_case(1, 4, wc, h, bc5=1, htop=0.0)
    th = lambda x: float(ol.vg_eval(soil, h=np.array([x]))["theta"][0])  # noqa: E731
    hd = lambda x: float(ol.vg_eval(soil, wc=np.array([x]))["head"][0])  # noqa: E731
    assert th(0.01) == ts
    np.testing.assert_array_equal(w, [ts, th(-1.0), th(-3.0), ts])
    np.testing.assert_array_equal(hh, h)
    # no ponded top: the top cell has no saturated neighbour, so its head follows its water
    # content, and so does the second cell's (its neighbours stay unsaturated)
    _, w, hh = _realloc_case(1, 4, wc, h, bc5=0)
    np.testing.assert_array_equal(w, [0.36, 0.38, th(-3.0), ts])
    np.testing.assert_array_equal(hh, [hd(0.36), hd(0.38), -3.0, -0.3])


def test_reallocation_lateral_neighbours():
    """2x1x2 box (i + k*nx): the lateral neighbour counts only across a face with Kx != 0, and the
    +x cell sees the -x cell after its update. Cell (0,1) sits below the saturated (0,0) and takes
    theta(+0.01 m) = theta_s; cell (1,1) is then next to a saturated cell across the open face."""
    wc = [0.41, 0.30, 0.30, 0.30]
    h = [-0.1, -0.5, 0.01, -0.9]
    soil, w, hh = _realloc_case(2, 2, wc, h, kx_faces=[(0, 1)])
    th = lambda x: float(ol.vg_eval(soil, h=np.array([x]))["theta"][0])  # noqa: E731
    hd = lambda x: float(ol.vg_eval(soil, wc=np.array([x]))["head"][0])  # noqa: E731
    np.testing.assert_array_equal(w, [0.40, 0.30, 0.40, th(-0.9)])
    np.testing.assert_array_equal(hh, [-0.1, hd(0.30), 0.01, -0.9])
    # the same face closed: (1,1) has no saturated neighbour
    _, w, hh = _realloc_case(2, 2, wc, h, kx_faces=[])
    np.testing.assert_array_equal(w, [0.40, 0.30, 0.40, 0.30])
    np.testing.assert_array_equal(hh, [-0.1, hd(0.30), 0.01, hd(0.30)])


def test_column_predictor_matches_dense_solve():
    """1x1x3 column, top Dirichlet: the predictor h equals the dense solution of the 3x3 system
    written from the PCA discretisation (ReSolver.cpp:397-549) by hand."""
    soil = abi.soil_params(alpha=2.0, n=1.8, theta_s=0.4, theta_r=0.05, Ks_x=1e-5, Ks_y=1e-5, Ks_z=1e-5, Ss=1e-5)
    dz, dt = 0.05, 10.0
    p = abi.re_params(soil=soil, solver=abi.solver_config(1e-14), bc_type=(0, 0, 0, 0, 0, 1), dx=1.0, dy=1.0,
                      dz=dz, dt=dt, adaptive_dt=0, htop=-0.5, use_full3d=1)
    r = ol.OracleRe(1, 1, 3, p)
    h0 = np.array([-2.0, -2.5, -3.0])
    r.set_initial(h=h0)
    r.step()
    # independent assembly
    ev = ol.vg_eval(soil, h=h0)
    K, Cap, th = ev["K"], ev["C"], ev["theta"]
    Vc = dz
    kf = [0.5 * (K[0] + K[1]), 0.5 * (K[1] + K[2]), K[2]]  # +z faces; bottom = own K, bc4 = 0 -> 0
    kf[2] = 0.0
    ktop = soil.Ks_z  # top Dirichlet face uses Ks_z (ReSolver.cpp:355-356)
    A = np.zeros((3, 3))
    b = np.zeros(3)
    for k in range(3):
        st = (Cap[k] + soil.Ss * th[k] / soil.theta_s) * Vc
        A[k, k] = st
        b[k] = st * h0[k]
        kup = ktop if k == 0 else kf[k - 1]
        gup = kup * dt * Vc / (dz * (0.5 * dz if k == 0 else dz))
        gdn = kf[k] * dt * Vc / (dz * dz)
        A[k, k] += gup + gdn
        if k > 0:
            A[k, k - 1] -= gup
        else:
            b[k] += gup * p.htop
        if k < 2:
            A[k, k + 1] -= gdn
        b[k] += -dt * Vc * kf[k] / dz + dt * Vc * kup / dz
    h_dense = np.linalg.solve(A, b)
    # The reallocation pass may replace h by h(theta) afterwards, so the predictor is checked
    # through the Darcy fluxes, which are evaluated from the solved head (ReSolver.cpp:947-948).
    cells = [3 * 3 * (k + 1) + 1 + 3 for k in range(3)]  # idx(0,0,k) on the 3x3 halo plane
    qz = r.get("qz")
    q_dense_top = 1.0 * ktop * ((h_dense[0] - p.htop) / (0.5 * dz) - 1.0)
    assert qz[1 + 3] == pytest.approx(q_dense_top, rel=1e-9)  # idx(0,0,-1): top ghost face
    q_dense_mid = 1.0 * 0.5 * (ol.vg_eval(soil, h=h_dense[:1])["K"][0] + ol.vg_eval(soil, h=h_dense[1:2])["K"][0]) * \
        ((h_dense[1] - h_dense[0]) / dz - 1.0)
    assert qz[cells[0]] == pytest.approx(q_dense_mid, rel=1e-9)


def _column_system(soil, p, h_kc, h_time, dz, dt):
    """3x3 system of the 1x1x3 top-Dirichlet column written from the discretisation by hand
    (ReSolver.cpp:397-549): K and C evaluated at h_kc, the storage and time terms at h_time."""
    ev = ol.vg_eval(soil, h=h_kc)
    K, Cap = ev["K"], ev["C"]
    th = ol.vg_eval(soil, h=h_time)["theta"]
    Vc = dz
    kf = [0.5 * (K[0] + K[1]), 0.5 * (K[1] + K[2]), 0.0]  # +z faces; closed bottom
    ktop = soil.Ks_z  # top Dirichlet face uses Ks_z (ReSolver.cpp:355-356)
    A = np.zeros((3, 3))
    b = np.zeros(3)
    for k in range(3):
        st = (Cap[k] + soil.Ss * th[k] / soil.theta_s) * Vc
        A[k, k] = st
        b[k] = st * h_time[k]
        kup = ktop if k == 0 else kf[k - 1]
        gup = kup * dt * Vc / (dz * (0.5 * dz if k == 0 else dz))
        gdn = kf[k] * dt * Vc / (dz * dz)
        A[k, k] += gup + gdn
        if k > 0:
            A[k, k - 1] -= gup
        else:
            b[k] += gup * p.htop
        if k < 2:
            A[k, k + 1] -= gdn
        b[k] += -dt * Vc * kf[k] / dz + dt * Vc * kup / dz
    return A, b, ktop


def test_column_picard_matches_dense_fixed_point():
    """Picard outer loop (picard_max > 1): the converged head of a 1x1x3 column is the fixed point
    h = A(h)^-1 b(h; h^n) of the hand-written system with K and C at the iterate and the storage
    terms at h^n, found here by dense solves; it differs from the predictor (one iterate)."""
    soil = abi.soil_params(alpha=2.0, n=1.8, theta_s=0.4, theta_r=0.05, Ks_x=1e-5, Ks_y=1e-5, Ks_z=1e-5, Ss=1e-5)
    dz, dt = 0.05, 50.0
    h0 = np.array([-2.0, -2.5, -3.0])

    def run(picard_max):
        p = abi.re_params(soil=soil, solver=abi.solver_config(1e-14), bc_type=(0, 0, 0, 0, 0, 1), dx=1.0, dy=1.0,
                          dz=dz, dt=dt, adaptive_dt=0, htop=-0.5, use_full3d=1, picard_max=picard_max,
                          picard_tol=1e-12)
        r = ol.OracleRe(1, 1, 3, p)
        r.set_initial(h=h0)
        st = r.step()
        return p, st, r.get("qz")

    p, st, qz = run(60)
    assert 2 <= st.picard_iterations < 60
    hm = h0.copy()
    for it in range(60):
        A, b, ktop = _column_system(soil, p, hm, h0, dz, dt)
        hn1 = np.linalg.solve(A, b)
        d = np.abs(hn1 - hm).max()
        hm = hn1
        if d <= 1e-12:
            break
    q_top = ktop * ((hm[0] - p.htop) / (0.5 * dz) - 1.0)
    k01 = ol.vg_eval(soil, h=hm[:2])["K"]
    q_mid = 0.5 * (k01[0] + k01[1]) * ((hm[1] - hm[0]) / dz - 1.0)
    assert qz[1 + 3] == pytest.approx(q_top, rel=1e-9)   # idx(0,0,-1): top ghost face
    assert qz[3 * 3 * 1 + 1 + 3] == pytest.approx(q_mid, rel=1e-9)  # idx(0,0,0)
    # one iterate is the reference's predictor, which is a different head here
    _, st1, qz1 = run(1)
    assert st1.picard_iterations == 1
    assert abs(qz1[1 + 3] - qz[1 + 3]) > 1e-3 * abs(qz[1 + 3])


@pytest.mark.parametrize("case", ["closed_ss", "infiltration", "saturating", "drainage"])
def test_water_ledger_closes(case):
    """The step's FP64 ledger: mass_after - mass_before = boundary inflow + specific-storage term +
    reallocation/clamp adjustments, to round-off (the corrector is an exact flux balance)."""
    if case == "closed_ss":
        p = box3d_params(box3d_soil(ss=1e-4))
        r = ol.OracleRe(NX, NY, NZ, p)
        r.set_initial(wc=box3d_wc())
    else:
        soil = abi.soil_params(alpha=3.35, n=2.0, theta_s=0.368, theta_r=0.102, Ks_x=9.22e-5, Ks_y=9.22e-5,
                               Ks_z=9.22e-5, Ss=1e-5)
        bc = {"infiltration": (0, 0, 0, 0, 0, 1), "saturating": (0, 0, 0, 0, 0, 1),
              "drainage": (0, 0, 0, 0, 3, 0)}[case]
        htop = 0.05 if case == "saturating" else -0.75
        p = abi.re_params(soil=soil, solver=abi.solver_config(1e-13), bc_type=bc, dx=0.1, dy=0.1, dz=0.02, dt=2.0,
                          adaptive_dt=0, use_full3d=1, htop=htop)
        r = ol.OracleRe(NX, NY, NZ, p)
        r.set_initial(h=np.full(NX * NY * NZ, -0.5 if case == "drainage" else -3.0))
    moved = adjusted = 0.0
    for _ in range(15):
        st = r.step()
        assert abs(st.mass_residual) <= 1e-12 * st.mass_before, (case, st.mass_residual, st.mass_before)
        moved += abs(st.mass_boundary_in) + abs(st.mass_storage_ss)
        adjusted += abs(st.mass_adjust)
        assert st.mass_before > 0 and st.mass_after > 0
    # each case exercises the term it is meant to
    assert moved > 0
    if case == "saturating":
        assert adjusted > 0
