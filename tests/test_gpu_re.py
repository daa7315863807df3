"""Richards PCA step: HIP path (through the C ABI) vs the FP64 oracle on identical inputs.

Cases follow the reference's RE tests (tests/re/test_re_3d.cpp, test_re_b2_gw.cpp) plus the
40x40x50 infiltration benchmark column (at reduced lateral size and at its own size), saturation/reallocation-heavy
cases, heterogeneous soil, every bottom/top boundary mode, adaptive dt and the FP32-in-FP64
refinement solver. Tolerance: head 1e-9 m per step for the FP64 path (the stated bar is 1e-6 m),
1e-9 relative for the other fields; mixed precision: head 1e-6 m.
"""
import os

import numpy as np
import pytest

from cases import box3d_params, box3d_soil, box3d_wc, interior3
from frehg_amd import abi
from gpu_common import RE_COMPARE, compare_re, re_pair, say

pytestmark = pytest.mark.gpu


def c1_params(nx, ny, nz, bc4=0, dt=2.0, **kw):
    """The 40x40x50 infiltration benchmark column: alpha 3.35, n 2, theta_s .368, theta_r .102, Ks 9.22e-5,
    Ss 1e-6, dx=dy=0.1, dz=0.02, top Dirichlet -0.75, fixed dt, full 3D. At the benchmark dt of
    10 s the PCA corrector drains the top cell below theta_r in the first step (the explicit
    flux divergence exceeds the cell's water; the clamp then sets h = h(theta_r + 1e-7) ~ -8e5 m),
    exactly as the reference's algorithm does; that regime is covered by a relative-tolerance test."""
    soil = abi.soil_params(alpha=3.35, n=2.0, theta_s=0.368, theta_r=0.102, Ks_x=9.22e-5, Ks_y=9.22e-5,
                           Ks_z=9.22e-5, Ss=1e-6)
    return abi.re_params(soil=soil, solver=abi.solver_config(1e-12), bc_type=(0, 0, 0, 0, bc4, 1), dx=0.1, dy=0.1,
                         dz=0.02, botz=-0.02 * nz, dt=dt, adaptive_dt=0, use_full3d=1, htop=-0.75, **kw)


@pytest.mark.parametrize("full3d", [1, 0])
def test_closed_box(full3d):
    say(f"RE closed box full3d={full3d}")
    soil = box3d_soil()
    orc, ctx = re_pair(4, 4, 8, box3d_params(soil, full3d=bool(full3d)), wc0=box3d_wc())
    t0 = interior3(ctx.re_get("wc"), 4, 4, 8).sum()
    for s in range(40):
        so, sg = orc.step(), ctx.re_step()
        assert sg.solve.converged
        if s in (0, 1, 39):
            compare_re(ctx, orc, label=f"step {s}")
    t1 = interior3(ctx.re_get("wc"), 4, 4, 8).sum()
    assert abs(t1 - t0) / t0 < 1e-10


@pytest.mark.parametrize("bc4", [0, 3, 1, 2])
def test_infiltration_column(bc4):
    nx, ny, nz = 10, 9, 25
    say(f"RE infiltration 10x9x25 bc4={bc4}")
    p = c1_params(nx, ny, nz, bc4=bc4, hbot=-3.0, qbot=2e-6)
    rng = np.random.default_rng(7)
    h0 = -10.0 + 0.5 * rng.standard_normal(nx * ny * nz)
    orc, ctx = re_pair(nx, ny, nz, p, h0=h0)
    for s in range(12):
        orc.step()
        st = ctx.re_step()
        assert st.solve.converged
    compare_re(ctx, orc, label=f"bc4={bc4}")


@pytest.mark.parametrize("full3d", [1, 0])
def test_saturating_top_exercises_reallocation(full3d):
    """Ponded top (htop >= 0): cells saturate, the reallocation fixed point must reproduce the
    reference's sequential i->j->k sweep."""
    nx, ny, nz = 8, 6, 12
    say(f"RE saturating top full3d={full3d}")
    soil = abi.soil_params(alpha=2.0, n=1.8, theta_s=0.40, theta_r=0.05, Ks_x=1e-5, Ks_y=1e-5, Ks_z=1e-5, Ss=1e-5)
    p = abi.re_params(soil=soil, solver=abi.solver_config(1e-13), bc_type=(0, 0, 0, 0, 0, 1), dx=0.5, dy=0.5,
                      dz=0.05, dt=60.0, adaptive_dt=0, use_full3d=full3d, htop=0.05)
    k = np.arange(nz)[:, None, None]
    i = np.arange(nx)[None, None, :]
    h0 = np.broadcast_to(-0.05 - 0.08 * k - 0.01 * i, (nz, ny, nx)).ravel()
    orc, ctx = re_pair(nx, ny, nz, p, h0=h0)
    for s in range(15):
        orc.step()
        ctx.re_step()
        compare_re(ctx, orc, label=f"step {s}", atol_h=1e-8, rtol=1e-8)
    wc = interior3(ctx.re_get("wc"), nx, ny, nz)
    assert (wc >= soil.theta_s).sum() > 0, "case must reach saturation"


@pytest.mark.parametrize("full3d", [1, 0])
def test_incremental_kc_matches_full_reevaluation(full3d):
    """After a step, kr and C are re-evaluated only where the head changed since the post-solve
    evaluation (reallocation); a context whose state is re-set before every step (which forces the
    full re-evaluation) must produce the same fields bit for bit. Also covers a ponded top whose
    reallocation replaces many heads."""
    nx, ny, nz = 8, 6, 12
    say(f"RE incremental kr/C full3d={full3d}")
    soil = abi.soil_params(alpha=2.0, n=1.8, theta_s=0.40, theta_r=0.05, Ks_x=1e-5, Ks_y=1e-5, Ks_z=1e-5, Ss=1e-5)
    p = abi.re_params(soil=soil, solver=abi.solver_config(1e-13), bc_type=(0, 0, 0, 0, 0, 1), dx=0.5, dy=0.5,
                      dz=0.05, dt=60.0, adaptive_dt=0, use_full3d=full3d, htop=0.05)
    k = np.arange(nz)[:, None, None]
    i = np.arange(nx)[None, None, :]
    h0 = np.broadcast_to(-0.05 - 0.08 * k - 0.01 * i, (nz, ny, nx)).ravel()
    # the SpMV autotuner could pick different shapes (different partial sums) for the two contexts
    os.environ["FREHG_SPMV_SHAPE"] = "2:12"
    try:
        _, a = re_pair(nx, ny, nz, p, h0=h0)
        _, b = re_pair(nx, ny, nz, p, h0=h0)
    finally:
        del os.environ["FREHG_SPMV_SHAPE"]
    for s in range(10):
        a.re_step()
        b.re_set("h", b.re_get("h"))  # invalidates the incremental kr / C state
        b.re_step()
        for f in ("h", "wc", "ch", "wch", "hwc", "qz", "Gct", "Grhs"):
            assert np.array_equal(a.re_get(f), b.re_get(f)), f"step {s} field {f}"


@pytest.mark.parametrize("full3d,adaptive,guess", [(1, 0, abi.GUESS_ZERO), (0, 0, abi.GUESS_ZERO),
                                                   (1, 1, abi.GUESS_ZERO), (1, 0, abi.GUESS_EXTRAPOLATED)])
def test_lazy_flux_fields_match_stored(full3d, adaptive, guess):
    """By default the step stores no Kx..qz (only face bits for the reallocation) and rebuilds
    them from the post-solve state when they are read or before the state changes; a context
    that stores them in the step (FREHG_LAZY_FLUX=0) must give every field bit for bit, whether
    the fields are read after every step, skipped for a step, or forced out by set_params. With
    the extrapolated start, both keep the post-solve head for the next step's increment (the
    stored-flux context copies it), so they start every solve from the same head."""
    nx, ny, nz = 9, 7, 12
    say(f"RE lazy flux fields full3d={full3d} adaptive={adaptive} guess={guess}")
    soil = abi.soil_params(alpha=2.0, n=1.8, theta_s=0.40, theta_r=0.05, Ks_x=1e-5, Ks_y=1e-5, Ks_z=1e-5, Ss=1e-5)
    p = abi.re_params(soil=soil, solver=abi.solver_config(1e-13), bc_type=(0, 0, 0, 0, 0, 1), dx=0.5, dy=0.5,
                      dz=0.05, dt=60.0, adaptive_dt=adaptive, dt_max=120.0, dt_min=1.0, use_full3d=full3d,
                      htop=0.05, initial_guess=guess)
    k = np.arange(nz)[:, None, None]
    i = np.arange(nx)[None, None, :]
    h0 = np.broadcast_to(-0.05 - 0.08 * k - 0.01 * i, (nz, ny, nx)).ravel()
    os.environ["FREHG_SPMV_SHAPE"] = "2:12"
    try:
        _, a = re_pair(nx, ny, nz, p, h0=h0)
        os.environ["FREHG_LAZY_FLUX"] = "0"
        _, b = re_pair(nx, ny, nz, p, h0=h0)
    finally:
        del os.environ["FREHG_SPMV_SHAPE"]
        os.environ.pop("FREHG_LAZY_FLUX", None)
    for s in range(9):
        a.re_step()
        b.re_step()
        if s % 3 == 1:
            continue  # the pending fields of this step are superseded by the next one
        if s == 5:
            a.re_set_params(p)  # stores the pending fields before the parameters change
            b.re_set_params(p)
        for f in RE_COMPARE:
            assert np.array_equal(a.re_get(f), b.re_get(f)), f"step {s} field {f}"


@pytest.mark.parametrize("mode", ["fp64_zero", "mixed_zero", "mixed_previous", "fp64_extrapolated",
                                  "mixed_extrapolated"])
def test_picard_outer_loop_matches_oracle(mode):
    """Picard outer loop (picard_max > 1) on an infiltration box at a large step: K and C at each
    iterate, storage terms at h^n, stop at max |dh| <= tol. FP64 from x0 = 0 is the oracle's
    algorithm (every field, same Picard count); FP32-in-FP64 from the previous iterate, and the
    loop started from h^n plus the previous step's solve increment, reach the same heads to
    1e-6 m. (At this step size the increment changes fast from step to step, and the extrapolated
    start can take an iterate more than h^n: only convergence is asserted for it.)"""
    nx, ny, nz = 9, 7, 12
    say(f"RE Picard loop {mode}")
    soil = abi.soil_params(alpha=2.0, n=1.8, theta_s=0.40, theta_r=0.05, Ks_x=1e-5, Ks_y=1e-5, Ks_z=1e-5, Ss=1e-5)
    kw = dict(soil=soil, bc_type=(0, 0, 0, 0, 0, 1), dx=0.5, dy=0.5, dz=0.05, dt=60.0, adaptive_dt=0,
              use_full3d=1, htop=-0.3, picard_max=40, picard_tol=1e-9)
    p = abi.re_params(solver=abi.solver_config(1e-13), **kw)
    k = np.arange(nz)[:, None, None]
    i = np.arange(nx)[None, None, :]
    h0 = np.broadcast_to(-1.5 - 0.05 * k - 0.02 * i, (nz, ny, nx)).ravel()
    orc, _ = re_pair(nx, ny, nz, p, h0=h0)
    if mode == "fp64_zero":
        _, ctx = re_pair(nx, ny, nz, p, h0=h0)
    else:
        prec = abi.PREC_FP64 if mode.startswith("fp64") else abi.PREC_MIXED
        guess = {"zero": abi.GUESS_ZERO, "previous": abi.GUESS_PREVIOUS,
                 "extrapolated": abi.GUESS_EXTRAPOLATED}[mode.split("_")[1]]
        pm = abi.re_params(solver=abi.solver_config(1e-13, precision=prec, inner_rtol=1e-5, max_refine=10),
                           initial_guess=guess, **kw)
        _, ctx = re_pair(nx, ny, nz, pm, h0=h0)
    for s in range(4):
        so = orc.step()
        sg = ctx.re_step()
        assert so.picard_iterations > 1
        if mode == "fp64_zero":
            assert sg.picard_iterations == so.picard_iterations, (s, sg.picard_iterations, so.picard_iterations)
        assert sg.picard_iterations < 40 and sg.picard_converged
        assert abs(sg.mass_residual) <= 1e-10 * sg.mass_before
        if mode == "mixed_zero":
            # every solve starts from zero, so each iterate's refinement runs on its own FP32 operator
            assert sg.solve.fallback == 0
    if mode == "fp64_zero":
        compare_re(ctx, orc, label="picard")
    else:
        assert np.abs(ctx.re_get("h") - orc.get("h")).max() <= 1e-6
        # every field after the step (K, C, theta and fluxes at the post-solve head included), also
        # where a solve from h^m returned it unchanged and the loop ended without another pass
        compare_re(ctx, orc, rtol=1e-6, atol_h=1e-6, label=f"picard {mode}")


@pytest.mark.parametrize("prec,guess,ksp,lazy", [(abi.PREC_FP64, abi.GUESS_PREVIOUS, abi.KSP_CG, "1"),
                                                 (abi.PREC_MIXED, abi.GUESS_EXTRAPOLATED, abi.KSP_CG, "1"),
                                                 (abi.PREC_MIXED, abi.GUESS_EXTRAPOLATED, abi.KSP_CG, "0"),
                                                 (abi.PREC_MIXED, abi.GUESS_PREVIOUS, abi.KSP_BICGSTAB, "1"),
                                                 (abi.PREC_FP64, abi.GUESS_EXTRAPOLATED, abi.KSP_BICGSTAB, "1")])
def test_picard_residual_test_before_assembly_equals_assemble_and_solve(prec, guess, ksp, lazy):
    """A Picard iterate h^m whose residual on the system at h^m already meets the solver's first
    test ends the loop without the system being stored or solved (the default), and the fluxes
    and corrector taken in the same pass as the test are the step's; with FREHG_PICARD_CHECK=0 the
    system is stored and the solve returns h^m unchanged, and the fluxes follow in a pass of their
    own. Every field, the system at the last iterate (stored when read) included, and every count
    of the step must agree bit for bit, also after the dt of an adaptive step has changed and after
    a parameter change forces the pending system out; with lazy = "0" (FREHG_LAZY_FLUX=0) the
    face conductivities and fluxes are stored in the step."""
    nx, ny, nz = 9, 7, 12
    say(f"RE Picard residual test before assembly: precision {prec}, guess {guess}, ksp {ksp}, lazy {lazy}")
    soil = abi.soil_params(alpha=2.0, n=1.8, theta_s=0.40, theta_r=0.05, Ks_x=1e-5, Ks_y=1e-5, Ks_z=1e-5, Ss=1e-5)
    kw = dict(soil=soil, bc_type=(0, 0, 0, 0, 0, 1), dx=0.5, dy=0.5, dz=0.05, dt=60.0, adaptive_dt=1, dt_max=90.0,
              dt_min=1.0, use_full3d=1, htop=-0.3, picard_max=40, picard_tol=1e-9)
    p = abi.re_params(solver=abi.solver_config(1e-13, ksp_type=ksp, precision=prec, inner_rtol=1e-5, max_refine=10),
                      initial_guess=guess, **kw)
    k = np.arange(nz)[:, None, None]
    i = np.arange(nx)[None, None, :]
    h0 = np.broadcast_to(-1.5 - 0.05 * k - 0.02 * i, (nz, ny, nx)).ravel()
    ctxs = []
    os.environ["FREHG_SPMV_SHAPE"] = "2:12"
    os.environ["FREHG_LAZY_FLUX"] = lazy
    try:
        for check in ("0", "1"):
            os.environ["FREHG_PICARD_CHECK"] = check
            ctxs.append(re_pair(nx, ny, nz, p, h0=h0)[1])
    finally:
        del os.environ["FREHG_SPMV_SHAPE"]
        del os.environ["FREHG_PICARD_CHECK"]
        del os.environ["FREHG_LAZY_FLUX"]
    a, b = ctxs
    for s in range(6):
        sa, sb = a.re_step(), b.re_step()
        assert sa.picard_converged and sb.picard_converged
        for fld in ("picard_iterations", "dt_used", "dt_next", "mass_residual", "max_dh"):
            assert getattr(sa, fld) == getattr(sb, fld), (s, fld)
        for fld in ("iterations", "refinements", "fallback"):
            assert getattr(sa.solve, fld) == getattr(sb.solve, fld), (s, fld)
        if s % 2 == 0:
            for f in RE_COMPARE:
                assert np.array_equal(a.re_get(f), b.re_get(f)), f"step {s} field {f}"
    # a parameter change between steps: the pending system is stored with the last step's dt first
    p2 = abi.re_params(solver=abi.solver_config(1e-13, ksp_type=ksp, precision=prec, inner_rtol=1e-5, max_refine=10),
                       initial_guess=guess, **dict(kw, dt=30.0, adaptive_dt=0))
    for c in (a, b):
        c.re_set_params(p2)
    for f in ("Gct", "Grhs"):
        assert np.array_equal(a.re_get(f), b.re_get(f)), f"after set_params: {f}"
    for c in (a, b):
        c.close()


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_MIXED])
def test_extrapolated_start_with_its_post_solve_pass_equals_separate_kernels(prec):
    """The extrapolated Picard start h^n + (last post-solve head - last starting head) is formed in
    the kernel that takes the post-solve pass at it (default); FREHG_FUSED_START=0 forms it and
    then runs the post-solve pass. Every field and count agrees bit for bit."""
    nx, ny, nz = 9, 7, 12
    say(f"RE extrapolated start fused with its post-solve pass: precision {prec}")
    soil = abi.soil_params(alpha=2.0, n=1.8, theta_s=0.40, theta_r=0.05, Ks_x=1e-5, Ks_y=1e-5, Ks_z=1e-5, Ss=1e-5)
    p = abi.re_params(solver=abi.solver_config(1e-13, precision=prec, inner_rtol=1e-5, max_refine=10),
                      initial_guess=abi.GUESS_EXTRAPOLATED, soil=soil, bc_type=(0, 0, 0, 0, 0, 1), dx=0.5, dy=0.5,
                      dz=0.05, dt=60.0, adaptive_dt=0, use_full3d=1, htop=0.05, picard_max=40, picard_tol=1e-9)
    k = np.arange(nz)[:, None, None]
    i = np.arange(nx)[None, None, :]
    h0 = np.broadcast_to(-0.05 - 0.08 * k - 0.01 * i, (nz, ny, nx)).ravel()
    ctxs = []
    os.environ["FREHG_SPMV_SHAPE"] = "2:12"
    try:
        for fused in ("0", "1"):
            os.environ["FREHG_FUSED_START"] = fused
            ctxs.append(re_pair(nx, ny, nz, p, h0=h0)[1])
    finally:
        del os.environ["FREHG_SPMV_SHAPE"]
        del os.environ["FREHG_FUSED_START"]
    a, b = ctxs
    for s in range(6):
        sa, sb = a.re_step(), b.re_step()
        assert sa.picard_converged and sb.picard_converged
        assert (sa.picard_iterations, sa.solve.iterations, sa.solve.refinements, sa.mass_residual) == \
            (sb.picard_iterations, sb.solve.iterations, sb.solve.refinements, sb.mass_residual), s
        for f in RE_COMPARE:
            assert np.array_equal(a.re_get(f), b.re_get(f)), f"step {s} field {f}"
    for c in (a, b):
        c.close()


@pytest.mark.parametrize("guess", [abi.GUESS_PREVIOUS, abi.GUESS_EXTRAPOLATED])
def test_assembly_with_the_first_refinement_sweep_matches_separate_kernels(guess):
    """A mixed-precision Picard solve from a nonzero iterate takes its first refinement sweep
    (r = b - A x in FP64, its FP32 copy, the start of the inner CG) in the assembly kernel, from
    the row in registers (default); FREHG_ASSEMBLE_SWEEP=0 assembles and leaves the sweep to the
    solver. The residual of each cell is the same; only the order of the sweep's partial sums
    differs, so heads agree to 1e-12 m and every count of every step is the same."""
    nx, ny, nz = 9, 7, 12
    say(f"RE assembly + first sweep: guess {guess}")
    soil = abi.soil_params(alpha=2.0, n=1.8, theta_s=0.40, theta_r=0.05, Ks_x=1e-5, Ks_y=1e-5, Ks_z=1e-5, Ss=1e-5)
    p = abi.re_params(solver=abi.solver_config(1e-13, precision=abi.PREC_MIXED, inner_rtol=1e-5, max_refine=10),
                      initial_guess=guess, soil=soil, bc_type=(0, 0, 0, 0, 0, 1), dx=0.5, dy=0.5, dz=0.05,
                      dt=60.0, adaptive_dt=0, use_full3d=1, htop=-0.3, picard_max=40, picard_tol=1e-9)
    k = np.arange(nz)[:, None, None]
    i = np.arange(nx)[None, None, :]
    h0 = np.broadcast_to(-1.5 - 0.05 * k - 0.02 * i, (nz, ny, nx)).ravel()
    ctxs = []
    os.environ["FREHG_SPMV_SHAPE"] = "2:12"
    try:
        for fused in ("0", "1"):
            os.environ["FREHG_ASSEMBLE_SWEEP"] = fused
            ctxs.append(re_pair(nx, ny, nz, p, h0=h0)[1])
    finally:
        del os.environ["FREHG_SPMV_SHAPE"]
        del os.environ["FREHG_ASSEMBLE_SWEEP"]
    a, b = ctxs
    for s in range(5):
        sa, sb = a.re_step(), b.re_step()
        assert sa.picard_converged and sb.picard_converged
        assert (sa.picard_iterations, sa.solve.refinements, sa.solve.fallback) == \
            (sb.picard_iterations, sb.solve.refinements, sb.solve.fallback), s
        assert abs(sa.solve.iterations - sb.solve.iterations) <= 1, s
        assert np.abs(a.re_get("h") - b.re_get("h")).max() <= 1e-12, s
        assert sa.mass_residual == pytest.approx(sb.mass_residual, abs=1e-12 * sa.mass_before)
    assert np.abs(a.re_get("Gct") - b.re_get("Gct")).max() <= 1e-12 * np.abs(b.re_get("Gct")).max()
    for c in (a, b):
        c.close()


def test_heterogeneous_soil_and_top_flux():
    nx, ny, nz = 9, 7, 10
    say("RE heterogeneous soil + per-column top flux")
    rng = np.random.default_rng(11)
    base = abi.soil_params(alpha=2.0, n=1.8, theta_s=0.40, theta_r=0.05, Ks_x=1e-5, Ks_y=1e-5, Ks_z=1e-5, Ss=1e-5)
    other = abi.soil_params(alpha=0.8, n=1.4, theta_s=0.45, theta_r=0.08, Ks_x=2e-6, Ks_y=2e-6, Ks_z=2e-6, Ss=1e-5)
    idx = (np.arange(nz)[:, None, None] >= nz // 2).astype(np.int32) * np.ones((nz, ny, nx), np.int32)
    ks = 1e-5 * np.exp(rng.standard_normal(nx * ny * nz))
    q = -1e-6 * (1 + rng.random(nx * ny))
    p = abi.re_params(soil=base, solver=abi.solver_config(1e-13), bc_type=(0, 0, 0, 0, 3, 2), dx=0.5, dy=0.5, dz=0.05,
                      dt=30.0, adaptive_dt=0, use_full3d=1)
    orc, ctx = re_pair(nx, ny, nz, p, soil_map=([base, other], idx.ravel(), True, ks), top_flux=q,
                       h0=np.full(nx * ny * nz, -2.0))
    for s in range(10):
        orc.step()
        ctx.re_step()
    compare_re(ctx, orc, label="hetero")


def test_adaptive_dt_b2gw_column():
    """b2-gw column (tests/re/test_re_b2_gw.cpp:109-153) with adaptive dt: same dt sequence."""
    say("RE adaptive dt b2-gw column")
    soil = abi.soil_params(alpha=1.43, n=1.56, theta_s=0.33, Ks_z=2.89e-6, Ss=1e-5)
    p = abi.re_params(soil=soil, solver=abi.solver_config(1e-12), bc_type=(0, 0, 0, 0, 0, 1), dz=0.01, dt=1e-4,
                      htop=0.0)
    orc, ctx = re_pair(1, 1, 100, p, wc0=np.full(100, 0.033))
    for s in range(60):
        so, sg = orc.step(), ctx.re_step()
        assert sg.dt_used == pytest.approx(so.dt_used, rel=1e-12), f"step {s}"
    compare_re(ctx, orc, label="b2gw", atol_h=1e-8, rtol=1e-8)
    assert ctx.re_get("wc")[1 + 3 + 9] > 0.033


@pytest.mark.parametrize("case", ["ramp_reduce_hold", "courant_dtmin"])
def test_adaptive_dt_every_branch(case):
    """adaptiveTimeStep (ReSolver.cpp:880-929) on an infiltration box where every branch of the rule
    fires (x0.75, hold and x1.25 in one case; the Courant limit and the dt_min clamp in the other):
    the device's flux-imbalance max and Courant min must give the oracle's dt sequence."""
    soil = abi.soil_params(alpha=3.35, n=2.0, theta_s=0.368, theta_r=0.102, Ks_x=9.22e-5, Ks_y=9.22e-5,
                           Ks_z=9.22e-5, Ss=1e-6)
    nx, ny, nz = 3, 2, 10
    say(f"RE adaptive dt, every branch: {case}")
    kw = dict(dt=5.0, dt_min=1e-3, co_max=1e9) if case == "ramp_reduce_hold" else dict(dt=0.5, dt_min=0.4,
                                                                                        co_max=1e-4)
    p = abi.re_params(soil=soil, solver=abi.solver_config(1e-12), bc_type=(0, 0, 0, 0, 0, 1), htop=-0.1, dx=0.1,
                      dy=0.1, dz=0.02, dt_max=100.0, adaptive_dt=1, use_full3d=1, **kw)
    orc, ctx = re_pair(nx, ny, nz, p, h0=np.full(nx * ny * nz, -3.0))
    ratios = []
    for s in range(8):
        so, sg = orc.step(), ctx.re_step()
        assert sg.dt_used == pytest.approx(so.dt_used, rel=1e-12), f"step {s}"
        assert sg.dt_next == pytest.approx(so.dt_next, rel=1e-12), f"step {s}"
        ratios.append(sg.dt_next / sg.dt_used)
    compare_re(ctx, orc, label=case, atol_h=1e-8, rtol=1e-8)
    r = np.array(ratios)
    if case == "ramp_reduce_hold":
        assert np.isclose(r, 0.75).any() and (r == 1.0).any() and np.isclose(r, 1.25).any(), r
    else:
        # Courant-limited steps take a dt that is none of the rule's ratios; dt_min ends the run
        odd = ~(np.isclose(r, 0.75) | np.isclose(r, 1.25) | (r == 1.0))
        assert odd.any(), r
        assert ctx.re_dt() == 0.4


def test_benchmark_column_free_drainage_all_steps():
    """The 40x40x50 infiltration benchmark as configured: free-drainage bottom (bc4 = 3), dt 10 s,
    all 360 steps, Picard to 1e-8 m (at most 50 iterates), FP64 GPU path vs the oracle. A Picard
    count that differs by one where an iterate's max |dh| lands on the tolerance moves the heads
    by about the tolerance, so the heads are compared at 1e-7 m (the stated bar is 1e-6 m); the
    ledger closes every step."""
    nx, ny, nz = 40, 40, 50
    say("RE 40x40x50 benchmark column, free drainage, Picard, 360 steps")
    p = c1_params(nx, ny, nz, bc4=3, dt=10.0, picard_max=50, picard_tol=1e-8)
    orc, ctx = re_pair(nx, ny, nz, p, h0=np.full(nx * ny * nz, -10.0))
    same = 0
    for s in range(360):
        so = orc.step()
        sg = ctx.re_step()
        same += sg.picard_iterations == so.picard_iterations
        assert abs(sg.mass_residual) <= 1e-10 * sg.mass_before
        assert sg.solve.converged
        if s % 60 == 59:
            say(f"  step {s + 1}: Picard {sg.picard_iterations} (oracle {so.picard_iterations})")
    assert same >= 355, same
    compare_re(ctx, orc, fields=("h", "wc", "wch", "hwc", "ch"), label="40x40x50 free drainage", atol_h=1e-7,
               rtol=1e-6)


def test_benchmark_column_full_size_picard_matches_oracle():
    """The 40x40x50 infiltration benchmark at its own size and dt (10 s) with the Picard loop to
    1e-8 m (at most 50 iterates): the first 60 of its 360 steps, FP64 GPU path vs the oracle, with
    the same Picard count every step (the first step stops at the 50-iterate cap in both), every
    field at the end, and a closed water ledger every step."""
    nx, ny, nz = 40, 40, 50
    say("RE 40x40x50 benchmark column, Picard, 60 steps")
    p = c1_params(nx, ny, nz, dt=10.0, picard_max=50, picard_tol=1e-8)
    orc, ctx = re_pair(nx, ny, nz, p, h0=np.full(nx * ny * nz, -10.0))
    for s in range(60):
        so = orc.step()
        sg = ctx.re_step()
        assert sg.picard_iterations == so.picard_iterations, (s, sg.picard_iterations, so.picard_iterations)
        assert abs(sg.mass_residual) <= 1e-10 * sg.mass_before
    compare_re(ctx, orc, label="40x40x50 Picard")


def test_baseline_dt_collapse_regime():
    """dt = 10 s: top cells hit theta_r (h ~ -8e5 m); parity holds relative to |h|."""
    nx, ny, nz = 6, 5, 25
    say("RE dt=10 collapse regime")
    p = c1_params(nx, ny, nz, dt=10.0)
    orc, ctx = re_pair(nx, ny, nz, p, h0=np.full(nx * ny * nz, -10.0))
    for s in range(6):
        orc.step()
        ctx.re_step()
    a, b = ctx.re_get("h"), orc.get("h")
    assert b.min() < -1e5, "the case must reach the clamped regime"
    assert np.max(np.abs(a - b) / (1.0 + np.abs(b))) < 1e-6


@pytest.mark.parametrize("ksp", [abi.KSP_CG, abi.KSP_BICGSTAB])
def test_mixed_precision_matches_fp64_oracle(ksp):
    nx, ny, nz = 16, 12, 20
    say(f"RE mixed precision vs FP64 oracle, ksp={ksp}")
    p = c1_params(nx, ny, nz, bc4=3)
    p.solver.rtol = 1e-12
    orc, _ = re_pair(nx, ny, nz, p, h0=np.full(nx * ny * nz, -10.0))
    pm = c1_params(nx, ny, nz, bc4=3)
    pm.solver = abi.solver_config(1e-12, precision=abi.PREC_MIXED, inner_rtol=1e-5, max_refine=10, ksp_type=ksp)
    _, ctx = re_pair(nx, ny, nz, pm, h0=np.full(nx * ny * nz, -10.0))
    for s in range(10):
        orc.step()
        st = ctx.re_step()
        assert st.solve.converged and st.solve.refinements >= 1
    err = np.abs(ctx.re_get("h") - orc.get("h")).max()
    assert err <= 1e-6, err


def test_previous_step_guess_converges_to_same_solution():
    nx, ny, nz = 12, 10, 15
    say("RE previous-step initial guess")
    p = c1_params(nx, ny, nz)
    orc, _ = re_pair(nx, ny, nz, p, h0=np.full(nx * ny * nz, -10.0))
    pg = c1_params(nx, ny, nz, initial_guess=abi.GUESS_PREVIOUS)
    _, ctx = re_pair(nx, ny, nz, pg, h0=np.full(nx * ny * nz, -10.0))
    for s in range(8):
        orc.step()
        ctx.re_step()
    assert np.abs(ctx.re_get("h") - orc.get("h")).max() <= 1e-8


def test_bicgstab_fp64_matches_cg_oracle():
    """The GW system solved by BiCGSTAB ("bcgs") instead of CG reaches the same discrete solution."""
    nx, ny, nz = 12, 10, 15
    say("RE BiCGSTAB fp64 vs CG oracle")
    p = c1_params(nx, ny, nz, bc4=3)
    p.solver.rtol = 1e-12
    orc, _ = re_pair(nx, ny, nz, p, h0=np.full(nx * ny * nz, -10.0))
    pb = c1_params(nx, ny, nz, bc4=3)
    pb.solver = abi.solver_config(1e-12, ksp_type=abi.KSP_BICGSTAB)
    _, ctx = re_pair(nx, ny, nz, pb, h0=np.full(nx * ny * nz, -10.0))
    for s in range(8):
        orc.step()
        st = ctx.re_step()
        assert st.solve.converged
    assert np.abs(ctx.re_get("h") - orc.get("h")).max() <= 1e-8


@pytest.mark.parametrize("precision", [abi.PREC_FP64, abi.PREC_MIXED])
def test_water_ledger_matches_oracle_and_closes(precision):
    """Every term of the GPU's FP64 water ledger against the oracle's, and closure to 1e-10 of the
    stored water every step (reallocation active: saturating top, heterogeneous Ks)."""
    nx, ny, nz = 12, 9, 16
    say(f"RE water ledger, precision {precision}")
    rng = np.random.default_rng(41)
    soil = abi.soil_params(alpha=3.35, n=2.0, theta_s=0.368, theta_r=0.102, Ks_x=9.22e-5, Ks_y=9.22e-5,
                           Ks_z=9.22e-5, Ss=1e-5)
    ks = 9.22e-5 * np.exp(0.8 * rng.standard_normal(nx * ny * nz))
    p = abi.re_params(soil=soil, solver=abi.solver_config(1e-12, precision=precision, max_refine=10),
                      bc_type=(0, 0, 0, 0, 3, 1), dx=0.1, dy=0.1, dz=0.02, dt=2.0, adaptive_dt=0, use_full3d=1,
                      htop=0.05)
    orc, ctx = re_pair(nx, ny, nz, p, soil_map=([soil], None, True, ks), h0=np.full(nx * ny * nz, -1.0))
    adj = 0.0
    for s in range(12):
        so = orc.step()
        sg = ctx.re_step()
        scale = so.mass_before
        assert abs(sg.mass_residual) <= 1e-10 * scale, (s, sg.mass_residual, scale)
        for f in ("mass_before", "mass_after", "mass_boundary_in", "mass_storage_ss", "mass_adjust"):
            tol = (1e-12 if precision == abi.PREC_FP64 else 1e-8) * scale
            assert abs(getattr(sg, f) - getattr(so, f)) <= tol, (s, f, getattr(sg, f), getattr(so, f))
        assert sg.max_dh == pytest.approx(so.max_dh, rel=1e-6)
        adj += abs(sg.mass_adjust)
    assert adj > 0, "the case must exercise reallocation / clamp"
