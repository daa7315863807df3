"""Pins the CPU reference of the learned-guess MLP (oracle/mlp_oracle.cpp): its back-propagated
gradient against central finite differences of its own loss in FP64, and its BF16-emulating mode
and the inference kernel's mode (FP16 operands, tanh folded into the next layer) against the FP64
mode."""
import numpy as np
import pytest

import oracle_lib as ol

NP, NI = 4993, 11
GROUPS = {"W1": (0, 704), "b1": (704, 768), "W2": (768, 4864), "b2": (4864, 4928), "w3": (4928, 4992), "b3": (4992, 4993)}


def random_case(seed, n=40, scale=0.5):
    rng = np.random.default_rng(seed)
    params = (scale * rng.standard_normal(NP)).astype(np.float32)
    params[GROUPS["W1"][0]:GROUPS["W1"][1]] *= 0.4
    params[GROUPS["W2"][0]:GROUPS["W2"][1]] *= 0.2
    x = rng.standard_normal((n, NI)).astype(np.float32) * np.float32(3.0) + np.float32(1.0)
    y = rng.standard_normal(n).astype(np.float32) * np.float32(0.01)
    norm = np.concatenate([np.r_[np.full(NI, 1.0), 0.0], np.r_[np.full(NI, 3.0), 0.01]]).astype(np.float32)
    return params, norm, x, y


def test_gradient_matches_finite_differences():
    params, norm, x, y = random_case(1)
    loss, g = ol.mlp_loss_grad(params, norm, x, y, bf16=False)
    assert loss > 0
    rng = np.random.default_rng(2)
    picks = [int(rng.integers(lo, hi)) for lo, hi in GROUPS.values() for _ in range(8)]
    for p in picks:
        h = np.float32(1e-3) * max(np.float32(1.0), abs(params[p]))
        pp, pm = params.copy(), params.copy()
        pp[p] = params[p] + h
        pm[p] = params[p] - h
        step = float(pp[p]) - float(pm[p])  # the representable perturbation
        fd = (ol.mlp_loss_grad(pp, norm, x, y, False)[0] - ol.mlp_loss_grad(pm, norm, x, y, False)[0]) / step
        assert abs(fd - g[p]) <= 1e-5 * max(1.0, np.abs(g).max()) + 1e-4 * abs(g[p]), (p, fd, g[p])


def test_prediction_is_consistent_with_loss():
    params, norm, x, y = random_case(3)
    pred = ol.mlp_predict(params, norm, x, bf16=False)
    t = (y.astype(np.float64) - norm[NI]) / norm[2 * NI + 1]
    yn = (pred - norm[NI]) / norm[2 * NI + 1]
    loss, _ = ol.mlp_loss_grad(params, norm, x, y, bf16=False)
    assert loss == pytest.approx(np.mean((yn - t) ** 2), rel=1e-10)


def test_bf16_mode_close_to_fp64():
    params, norm, x, y = random_case(4, n=200)
    l64, g64 = ol.mlp_loss_grad(params, norm, x, y, bf16=False)
    l16, g16 = ol.mlp_loss_grad(params, norm, x, y, bf16=True)
    assert l16 == pytest.approx(l64, rel=5e-2)
    for name, (lo, hi) in GROUPS.items():
        scale = np.abs(g64[lo:hi]).max()
        assert np.abs(g16[lo:hi] - g64[lo:hi]).max() <= 0.05 * scale, name
    p64 = ol.mlp_predict(params, norm, x, bf16=False)
    p16 = ol.mlp_predict(params, norm, x, bf16=True)
    assert np.abs(p16 - p64).max() <= 0.05 * np.abs(p64).max()


def test_inference_mode_close_to_fp64():
    """Mode 2 (r = (1 - tanh z) / 2 outputs, -2 W and b + W 1 in the next layer, FP16 operands) is
    the FP64 network up to FP16 rounding: about four times closer than the BF16 mode. A wrong fold
    (sign, factor or the bias shift) is off by O(1), BF16-sized rounding by ~4e-3."""
    for seed in (4, 6):
        params, norm, x, _ = random_case(seed, n=1000)
        p64 = ol.mlp_predict(params, norm, x, bf16=0)
        pin = ol.mlp_predict(params, norm, x, bf16=2)
        assert np.abs(pin - p64).max() <= 2e-3 * np.abs(p64).max()
