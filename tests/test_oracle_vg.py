"""Pins the oracle's Van Genuchten relations (oracle/oracle.hpp VG, restating
src/re/VanGenuchten.hpp:18-137).

Checks: the reference's own unit tests (tests/re/test_re_constitutive.cpp:12-74), independent
closed forms of van Genuchten (1980) / Mualem (1976) written from the mathematics (using
Se^(1/m) = 1/(1+|ah|^n)), finite-difference derivatives, and the branch cut-offs.
"""
import numpy as np
import pytest

import oracle_lib as ol
from frehg_amd import abi


def ref_soil(**kw):
    return abi.soil_params(alpha=1.43, n=1.56, theta_s=0.33, theta_r=0.0, Ks_z=2.89e-6, **kw)


def test_closed_form_bounds():  # test_re_constitutive.cpp:12-30
    s = ref_soil()
    r = ol.vg_eval(s, h=np.array([-0.5]))
    assert s.theta_r < r["theta"][0] < s.theta_s
    assert 0.0 < r["K"][0] <= s.Ks_z


@pytest.mark.parametrize("wc0", [0.05, 0.1, 0.15, 0.2, 0.25, 0.3])
def test_roundtrip_theta_h_theta(wc0):  # test_re_constitutive.cpp:32-46
    s = ref_soil()
    h = ol.vg_eval(s, wc=np.array([wc0]))["head"]
    wc1 = ol.vg_eval(s, h=h)["theta"][0]
    assert abs(wc1 - wc0) < 1e-10


def test_mvg_roundtrip():  # test_re_constitutive.cpp:48-62
    s = ref_soil(use_mvg=1, air_entry=-0.02)
    h = ol.vg_eval(s, wc=np.array([0.15]))["head"]
    assert abs(ol.vg_eval(s, h=h)["theta"][0] - 0.15) < 1e-10


def independent_vg(s, h):
    """van Genuchten (1980) with Mualem (1976) K, from the formulas, not from the oracle."""
    m = 1.0 - 1.0 / s.n
    p = np.abs(s.alpha * h) ** s.n
    se = (1.0 + p) ** (-m)
    theta = s.theta_r + (s.theta_s - s.theta_r) * se
    cap = s.alpha * s.n * m * (s.theta_s - s.theta_r) * np.abs(s.alpha * h) ** (s.n - 1) * (1 + p) ** (-m - 1)
    kr = np.sqrt(se) * (1.0 - (p / (1.0 + p)) ** m) ** 2
    return theta, cap, s.Ks_z * kr


@pytest.mark.parametrize("alpha,n", [(1.43, 1.56), (3.35, 2.0), (2.0, 1.8), (0.5, 1.3), (5.0, 2.8)])
def test_independent_closed_forms(alpha, n):
    s = abi.soil_params(alpha=alpha, n=n, theta_s=0.4, theta_r=0.05, Ks_z=1e-5)
    h = -np.logspace(-1.5, 2, 60)  # below air entry (-0.02) so the VG branch is taken
    r = ol.vg_eval(s, h=h)
    th, cap, K = independent_vg(s, h)
    np.testing.assert_allclose(r["theta"], th, rtol=1e-12)
    np.testing.assert_allclose(r["C"], cap, rtol=1e-11)
    np.testing.assert_allclose(r["K"], K, rtol=1e-8, atol=1e-300)


def test_capacity_is_dtheta_dh():
    s = abi.soil_params(alpha=2.0, n=1.8, theta_s=0.4, theta_r=0.05)
    h = -np.linspace(0.1, 8.0, 40)
    eps = 1e-6
    dth = (ol.vg_eval(s, h=h + eps)["theta"] - ol.vg_eval(s, h=h - eps)["theta"]) / (2 * eps)
    np.testing.assert_allclose(ol.vg_eval(s, h=h)["C"], dth, rtol=1e-6)


def test_dKdwc_is_dK_dtheta():
    s = abi.soil_params(alpha=2.0, n=1.8, theta_s=0.4, theta_r=0.05, Ks_z=1e-5)
    wc = np.linspace(0.1, 0.35, 20)
    eps = 1e-7
    hp = ol.vg_eval(s, wc=wc + eps)["head"]
    hm = ol.vg_eval(s, wc=wc - eps)["head"]
    dK = (ol.vg_eval(s, h=hp)["K"] - ol.vg_eval(s, h=hm)["K"]) / (2 * eps)
    np.testing.assert_allclose(ol.vg_eval(s, wc=wc)["dKdwc"], dK, rtol=1e-5)


def test_branch_cutoffs():
    s = abi.soil_params(alpha=2.0, n=1.8, theta_s=0.4, theta_r=0.05, Ks_z=1e-5, air_entry=-0.02)
    r = ol.vg_eval(s, h=np.array([-0.019, -0.0, 0.5, -0.021]), wc=np.array([0.4, 0.45, 0.04, 0.2]))
    # theta = theta_s above the air-entry value even for plain VG (VanGenuchten.hpp:37-38)
    assert r["theta"][0] == s.theta_s and r["theta"][2] == s.theta_s
    assert r["theta"][3] < s.theta_s
    # C > 0 for air_entry < h <= 0 in plain VG, 0 for h > 0 (VanGenuchten.hpp:75-79)
    assert r["C"][0] > 0.0 and r["C"][2] == 0.0
    # K = Ks for h > 0 (VanGenuchten.hpp:105-109)
    assert r["K"][2] == s.Ks_z and r["K"][0] < s.Ks_z
    # h(theta) = 0 at/above saturation; theta - theta_r floored at 1e-7 (VanGenuchten.hpp:55-62)
    assert r["head"][0] == 0.0 and r["head"][1] == 0.0
    assert np.isfinite(r["head"][2]) and r["head"][2] < -10.0
