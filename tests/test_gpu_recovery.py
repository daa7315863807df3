"""Recovery paths of the HIP backend, forced from the host: a learned guess from a broken network
(huge or non-finite outputs), a learned Picard start in the regime where the corrector drains
cells to theta_r (heads near -8e5 m), a mixed-precision solve whose refinement stalls (FP64
fallback), and a Picard loop stopped by its iteration cap (reported, not silent). Every case still
converges to the answer of the plain method, which the oracle or a zero-guess run gives."""
import numpy as np
import pytest

import oracle_lib as ol
from frehg_amd import abi, gpu
from gpu_common import compare_re, re_pair, say
from test_gpu_mlp import infiltration
from test_oracle_mlp import GROUPS, random_case

pytestmark = pytest.mark.gpu


def _ctx(nx, ny, nz, p, soil, ks, h0, params=None, norm=None):
    ctx = gpu.Context(0)
    ctx.re_init(nx, ny, nz, p)
    ctx.re_set_soil([soil], None, True, ks)
    ctx.re_set_initial(h=h0)
    if params is not None:
        ctx.mlp_set_params(params, norm)
    return ctx


@pytest.mark.parametrize("picard_max", [1, 100])
@pytest.mark.parametrize("broken", ["scaled_1e6", "nan_weights"])
def test_broken_network_falls_back(broken, picard_max):
    """A network whose outputs are 1e6 times too large, or whose weights are NaN: the step reports
    guess_fallback = 1 (the guess residual is not below the zero guess's, or an output is not
    finite) and reaches the zero-guess heads."""
    nx, ny, nz, steps = 20, 16, 12, 4
    say(f"recovery: broken network {broken}, picard_max={picard_max}")
    params, norm, _, _ = random_case(3)
    norm = norm.copy()
    norm[:11] = [-2.0, 0, 0, 0, 0, 0, 0, -14.0, 0.03, 0.2, 0.0]
    norm[12:23] = [0.3, 0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 2.0, 0.03, 0.05, 1.0]
    if broken == "scaled_1e6":
        norm[23] = 1e6 * 1e-3  # the target's standard deviation scales every output
    else:
        params = params.copy()
        lo, hi = GROUPS["W2"]
        params[lo:hi:7] = np.nan
        norm[23] = 1e-3
    p, soil, ks, h0 = infiltration(nx, ny, nz, 17, abi.GUESS_LEARNED, picard_max=picard_max)
    pz, _, _, _ = infiltration(nx, ny, nz, 17, abi.GUESS_ZERO, picard_max=picard_max)
    ctx = _ctx(nx, ny, nz, p, soil, ks, h0, params, norm)
    ref = _ctx(nx, ny, nz, pz, soil, ks, h0)
    fallbacks = 0
    for s in range(steps):
        st = ctx.re_step()
        rc = ctx.last_rc
        sr = ref.re_step()
        assert sr.picard_converged  # the first step takes about 40 Picard iterates here
        assert st.solve.converged and st.picard_converged
        fallbacks += st.guess_fallback
        assert rc == (abi.INFO_GUESS_FALLBACK if st.guess_fallback else abi.OK)
    assert fallbacks >= 1
    if broken == "nan_weights":
        assert fallbacks == steps  # every output of every step is NaN
    tol = 1e-7 if picard_max == 1 else 1e-6
    assert np.abs(ctx.re_get("h") - ref.re_get("h")).max() <= tol


def test_learned_start_in_the_theta_r_regime():
    """dt = 10 s on the 40x40x50 infiltration benchmark's soil with the predictor step: the corrector
    drains the top cells to theta_r, and their heads (about -8e5 m) lie far outside any training
    data (standardised, they overflow FP16). From that state a network trained on a milder run still
    gives finite starts (the kernel clamps standardised features): the learned predictor step
    reaches the zero-guess heads, and a learned Picard step (the zero-guess Picard loop does not
    converge from this state within 150 iterates) keeps finite heads and converged solves, and
    reports FREHG_INFO_PICARD_NOT_CONVERGED if it stops at the cap."""
    nx, ny, nz = 8, 6, 20
    say("recovery: learned start from the theta_r regime")
    soil = abi.soil_params(alpha=3.35, n=2.0, theta_s=0.368, theta_r=0.102, Ks_x=9.22e-5, Ks_y=9.22e-5,
                           Ks_z=9.22e-5, Ss=1e-6)

    def params(guess, dt, picard_max):
        return abi.re_params(soil=soil, solver=abi.solver_config(1e-12), bc_type=(0, 0, 0, 0, 3, 1), dx=0.1, dy=0.1,
                             dz=0.02, dt=dt, adaptive_dt=0, use_full3d=1, htop=-0.75, initial_guess=guess,
                             picard_max=picard_max, picard_tol=1e-8)

    h0 = np.full(nx * ny * nz, -10.0)
    ctx = gpu.Context(0)
    ctx.mlp_init(5)
    ctx.mlp_log(stride=1, capacity=200_000)
    ctx.re_init(nx, ny, nz, params(abi.GUESS_ZERO, 1.0, 1))
    ctx.re_set_initial(h=h0)
    for _ in range(6):
        ctx.re_step()
    ctx.mlp_log(stride=0, capacity=0)
    ctx.mlp_train(steps=200, batch=4096, lr=1e-3, seed=3)
    ref = gpu.Context(0)
    for c in (ctx, ref):
        c.re_init(nx, ny, nz, params(abi.GUESS_ZERO, 10.0, 1))
        c.re_set_initial(h=h0)
        for _ in range(2):  # predictor steps into the theta_r regime
            c.re_step()
    assert ref.re_get("h").min() < -1e5, "the case must reach the theta_r regime"
    ctx.re_set_params(params(abi.GUESS_LEARNED, 10.0, 1))
    for s in range(2):
        st, sr = ctx.re_step(), ref.re_step()
        assert st.solve.converged and sr.solve.converged, s
    a, b = ctx.re_get("h"), ref.re_get("h")
    assert np.all(np.isfinite(a)) and b.min() < -1e5
    assert np.max(np.abs(a - b) / (1.0 + np.abs(b))) < 1e-8
    ctx.re_set_params(params(abi.GUESS_LEARNED, 10.0, 150))
    ref.re_set_params(params(abi.GUESS_ZERO, 10.0, 150))
    st, sr = ctx.re_step(), ref.re_step()
    assert st.solve.converged and sr.solve.converged
    assert sr.picard_converged == 0 and sr.picard_iterations == 150
    # a learned start that stops at the cap says so
    assert (ctx.last_rc == abi.INFO_PICARD_NOT_CONVERGED) == (st.picard_converged == 0)
    assert np.all(np.isfinite(ctx.re_get("h")))


def test_stalled_refinement_falls_back_to_fp64():
    """Mixed precision with one refinement allowed and an inner solve that only halves the residual
    (inner_rtol 0.5) cannot reach rtol 1e-13: the solve finishes in FP64 from its iterate
    (solve.fallback = 1) and matches the FP64 oracle."""
    nx, ny, nz = 10, 8, 12
    say("recovery: stalled mixed-precision refinement -> FP64")
    soil = abi.soil_params(alpha=2.0, n=1.8, theta_s=0.40, theta_r=0.05, Ks_x=1e-5, Ks_y=1e-5, Ks_z=1e-5, Ss=1e-5)
    kw = dict(soil=soil, bc_type=(0, 0, 0, 0, 0, 1), dx=0.5, dy=0.5, dz=0.05, dt=20.0, adaptive_dt=0, use_full3d=1,
              htop=-0.3)
    p = abi.re_params(solver=abi.solver_config(1e-13), **kw)
    pm = abi.re_params(solver=abi.solver_config(1e-13, precision=abi.PREC_MIXED, inner_rtol=0.5, max_refine=1), **kw)
    h0 = np.full(nx * ny * nz, -1.5)
    orc, _ = re_pair(nx, ny, nz, p, h0=h0)
    _, ctx = re_pair(nx, ny, nz, pm, h0=h0)
    for s in range(3):
        orc.step()
        st = ctx.re_step()
        assert st.solve.converged and st.solve.fallback == 1, s
    assert np.abs(ctx.re_get("h") - orc.get("h")).max() <= 1e-8


def test_picard_cap_is_reported():
    """A Picard loop held to 2 iterates on a step that needs many: the step is taken from the last
    iterate as in the oracle (same fields), stats.picard_converged = 0 and the call returns
    FREHG_INFO_PICARD_NOT_CONVERGED; with room to converge the same step reports success."""
    nx, ny, nz = 9, 7, 12
    say("recovery: Picard iteration cap")
    soil = abi.soil_params(alpha=2.0, n=1.8, theta_s=0.40, theta_r=0.05, Ks_x=1e-5, Ks_y=1e-5, Ks_z=1e-5, Ss=1e-5)
    kw = dict(soil=soil, solver=abi.solver_config(1e-13), bc_type=(0, 0, 0, 0, 0, 1), dx=0.5, dy=0.5, dz=0.05,
              dt=60.0, adaptive_dt=0, use_full3d=1, htop=-0.3, picard_tol=1e-9)
    h0 = np.full(nx * ny * nz, -1.5)
    orc, ctx = re_pair(nx, ny, nz, abi.re_params(picard_max=2, **kw), h0=h0)
    so = orc.step()
    sg = ctx.re_step()
    assert ctx.last_rc == abi.INFO_PICARD_NOT_CONVERGED
    assert sg.picard_iterations == 2 and sg.picard_converged == 0
    assert so.picard_iterations == 2 and so.picard_converged == 0
    compare_re(ctx, orc, label="picard cap")
    orc2, ctx2 = re_pair(nx, ny, nz, abi.re_params(picard_max=40, **kw), h0=h0)
    so2, sg2 = orc2.step(), ctx2.re_step()
    assert ctx2.last_rc == abi.OK and sg2.picard_converged == 1 and so2.picard_converged == 1
    assert sg2.picard_iterations == so2.picard_iterations > 2
