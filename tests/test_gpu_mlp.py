"""Learned initial guess on the GPU: the BF16 MFMA forward pass and back-propagated gradient of
frehg_amd/csrc/mlp.hip against the CPU reference (oracle/mlp_oracle.cpp, BF16-emulating mode),
training on pairs logged from the plain solver, and what the trained network gains over the
zero initial guess on held-out soil fields while converging to the same heads."""
import numpy as np
import pytest

import oracle_lib as ol
from frehg_amd import abi, gpu
from gpu_common import say
from test_oracle_mlp import GROUPS, random_case

pytestmark = pytest.mark.gpu


def test_forward_matches_oracle():
    say("MLP forward vs oracle")
    params, norm, x, _ = random_case(11, n=1000)
    ctx = gpu.Context(0)
    ctx.mlp_set_params(params, norm)
    p = ctx.mlp_predict(x).astype(np.float64)
    ref = ol.mlp_predict(params, norm, x, bf16=True)
    ref64 = ol.mlp_predict(params, norm, x, bf16=False)
    scale = np.abs(ref).max()
    assert np.abs(p - ref).max() <= 2e-3 * scale, np.abs(p - ref).max() / scale
    assert np.abs(p - ref64).max() <= 5e-2 * scale


@pytest.mark.parametrize("n", [64, 1000, 20000])
def test_gradient_matches_oracle(n):
    say(f"MLP gradient vs oracle n={n}")
    params, norm, x, y = random_case(12 + n, n=n)
    ctx = gpu.Context(0)
    ctx.mlp_set_params(params, norm)
    loss, g = ctx.mlp_loss_grad(x, y)
    lref, gref = ol.mlp_loss_grad(params, norm, x, y, bf16=True)
    assert loss == pytest.approx(lref, rel=1e-3)
    for name, (lo, hi) in GROUPS.items():
        scale = np.abs(gref[lo:hi]).max()
        err = np.abs(g[lo:hi] - gref[lo:hi]).max()
        assert err <= 5e-3 * scale, f"{name}: {err:.3e} vs scale {scale:.3e}"


def infiltration(nx, ny, nz, seed, guess, picard_max=1, use_mvg=0):
    rng = np.random.default_rng(seed)
    soil = abi.soil_params(alpha=2.0, n=1.8, theta_s=0.40, theta_r=0.05, Ks_x=1e-5, Ks_y=1e-5, Ks_z=1e-5, Ss=1e-5,
                           use_mvg=use_mvg)
    p = abi.re_params(soil=soil, solver=abi.solver_config(1e-10), bc_type=(0, 0, 0, 0, 3, 1), dx=1.0, dy=1.0,
                      dz=0.05, dt=2.0, adaptive_dt=0, use_full3d=1, htop=-0.2, initial_guess=guess,
                      picard_max=picard_max, picard_tol=1e-8)
    ks = np.exp(rng.normal(np.log(1e-5), 1.0, nx * ny * nz))
    h0 = np.full(nx * ny * nz, -2.0) + 0.2 * rng.standard_normal(nx * ny * nz)
    return p, soil, ks, h0


def run(ctx, nx, ny, nz, seed, guess, steps, picard_max=1, picards=None):
    p, soil, ks, h0 = infiltration(nx, ny, nz, seed, guess, picard_max)
    ctx.re_init(nx, ny, nz, p)
    ctx.re_set_soil([soil], None, True, ks)
    ctx.re_set_initial(h=h0)
    its, fallbacks = [], 0
    for _ in range(steps):
        st = ctx.re_step()
        assert st.solve.converged
        its.append(st.solve.iterations)
        fallbacks += st.guess_fallback
        if picards is not None:
            picards.append(st.picard_iterations)
    return np.array(its), fallbacks, ctx.re_get("h")


def test_training_and_learned_guess_gain():
    nx, ny, nz, steps = 40, 32, 24, 12
    ctx = gpu.Context(0)
    ctx.mlp_init(5)
    ctx.mlp_log(stride=1, capacity=2_000_000)
    say("MLP: logging plain-solver pairs on two training fields")
    for seed in (101, 102):
        run(ctx, nx, ny, nz, seed, abi.GUESS_ZERO, steps)
    # the first step of a run has no previous solve increment and logs nothing
    assert ctx.mlp_logged() == 2 * (steps - 1) * nx * ny * nz
    ctx.mlp_log(stride=0, capacity=0)
    say("MLP: training")
    first, last = ctx.mlp_train(steps=1500, batch=8192, lr=1e-3, seed=9)
    say(f"MLP loss {first:.4f} -> {last:.4f}")
    assert last < 0.5 * first
    say("MLP: held-out field, zero guess vs learned guess")
    its0, _, h0 = run(ctx, nx, ny, nz, 201, abi.GUESS_ZERO, steps)
    its1, fb, h1 = run(ctx, nx, ny, nz, 201, abi.GUESS_LEARNED, steps)
    itsp, _, hp = run(ctx, nx, ny, nz, 201, abi.GUESS_PREVIOUS, steps)
    say(f"Krylov iterations per step: zero {its0.mean():.2f}, previous-step {itsp.mean():.2f}, "
        f"learned {its1.mean():.2f}, fallbacks {fb}")
    # Jacobi-CG contracts the residual ~50x per iteration on this operator, so a guess that is
    # k orders of magnitude better saves about k/1.7 iterations; the network must save some.
    assert its1.sum() < its0.sum() and its1.sum() <= itsp.sum()
    assert fb == 0
    assert np.abs(h1 - h0).max() <= 1e-7 and np.abs(hp - h0).max() <= 1e-7
    say("MLP: held-out field with the Picard loop (tol 1e-8 m): from h^n vs from the learned head")
    # (Picard contracts by about 0.3-0.5 per iteration here, so a start that is closer by one order
    # of magnitude saves only one or two iterations of a step)
    pz, pl = [], []
    _, _, hz = run(ctx, nx, ny, nz, 201, abi.GUESS_ZERO, steps, picard_max=80, picards=pz)
    _, fb, hl = run(ctx, nx, ny, nz, 201, abi.GUESS_LEARNED, steps, picard_max=80, picards=pl)
    say(f"Picard iterations per step: from h^n {np.mean(pz):.2f}, from the learned head {np.mean(pl):.2f}")
    assert max(pz) > 1 and max(pz) < 80 and max(pl) < 80
    assert sum(pl) <= sum(pz)
    assert np.abs(hl - hz).max() <= 1e-6


def test_untrained_model_rejected_or_harmless():
    nx, ny, nz = 16, 12, 10
    ctx = gpu.Context(0)
    p, soil, ks, h0 = infiltration(nx, ny, nz, 7, abi.GUESS_LEARNED)
    ctx.re_init(nx, ny, nz, p)
    ctx.re_set_soil([soil], None, True, ks)
    ctx.re_set_initial(h=h0)
    with pytest.raises(gpu.FrehgError) as e:
        ctx.re_step()
    assert e.value.code == abi.ERR_INVALID_ARG
    # a random (untrained) network: the solve still converges to the zero-guess answer
    params, norm, _, _ = random_case(3)
    ctx.mlp_set_params(params, norm)
    st = ctx.re_step()
    assert st.solve.converged


@pytest.mark.parametrize("use_mvg", [0, 1])
def test_guess_kernel_matches_oracle_on_solver_state(use_mvg):
    """The inference kernel's features (h, six neighbour differences, ln K, C, theta, depth) and its
    transposed-layer FP16 network against host-side features fed to the CPU reference, in the
    reference's emulation of the kernel's arithmetic and in FP64. The kernel evaluates ln K and C
    itself: in single precision for van Genuchten soils, through the FP64 law for modified VG."""
    nx, ny, nz = 70, 9, 6
    say(f"MLP guess kernel vs oracle use_mvg={use_mvg}")
    ctx = gpu.Context(0)
    p, soil, ks, h0 = infiltration(nx, ny, nz, 31, abi.GUESS_ZERO, use_mvg=use_mvg)
    ctx.re_init(nx, ny, nz, p)
    ctx.re_set_soil([soil], None, True, ks)
    ctx.re_set_initial(h=h0)
    ctx.re_step()
    params, norm, _, _ = random_case(5)
    norm[:11] = [-1.5, 0, 0, 0, 0, 0, 0, -12.0, 0.3, 0.25, 0.0]
    norm[12:23] = [0.5, 0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 1.5, 0.3, 0.05, 1.0]
    ctx.mlp_set_params(params, norm)
    inc = ctx.mlp_increment()
    h = ctx.re_get("h").reshape(nz + 2, ny + 2, nx + 2)
    wc = ctx.re_get("wc").reshape(nz + 2, ny + 2, nx + 2)
    hc = h[1:-1, 1:-1, 1:-1]
    vg = ol.vg_eval(soil, h=hc.ravel())
    X = np.zeros((nx * ny * nz, 11))
    X[:, 0] = hc.ravel()
    nbrs = [h[1:-1, 1:-1, :-2], h[1:-1, 1:-1, 2:], h[1:-1, :-2, 1:-1], h[1:-1, 2:, 1:-1], h[:-2, 1:-1, 1:-1],
            h[2:, 1:-1, 1:-1]]
    for q, nb in enumerate(nbrs):
        X[:, 1 + q] = (nb - hc).ravel()
    X[:, 7] = np.log(np.maximum(ks * vg["K"] / soil.Ks_z, 1e-30))
    X[:, 8] = vg["C"]
    X[:, 9] = wc[1:-1, 1:-1, 1:-1].ravel()
    # the kernel's arithmetic (FP16 operands, folded tanh), and the FP64 network
    ref = ol.mlp_predict(params, norm, X.astype(np.float32), bf16=2)
    ref64 = ol.mlp_predict(params, norm, X.astype(np.float32), bf16=0)
    scale = np.abs(ref64).max()
    assert np.abs(inc - ref).max() <= 1e-3 * scale, np.abs(inc - ref).max() / scale
    assert np.abs(inc - ref64).max() <= 3e-3 * scale, np.abs(inc - ref64).max() / scale
