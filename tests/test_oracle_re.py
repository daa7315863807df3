"""Pins the oracle's Richards PCA step (oracle/re_oracle.cpp, restating src/re/ReSolver.cpp)
with the reference's RE tests (tests/re/test_re_3d.cpp:133-274, test_re_b2_gw.cpp:109-153)
and a brute-force 1x1x3 column solved by dense linear algebra."""
import numpy as np
import pytest

import oracle_lib as ol
from cases import box3d_params, box3d_soil, box3d_wc, interior3
from frehg_amd import abi

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
    assert r.dt() == pytest.approx(1e-4 * 1.25) or r.dt() >= 1e-4  # adaptive ramp (ReSolver.cpp:921-927)


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
