"""The multigrid V-cycle preconditioner of the pressure PCG.

A preconditioner changes how fast PCG converges, not what it converges to:
with either preconditioner the GPU must reach the same FP64 tolerance on the
same system and agree with the oracle's converged FP64 solve (relative L2
<= 1e-5).  The V-cycle must do so in a small number of iterations that
barely grows with the grid; grids of odd and skewed shapes must work; a grid
too skewed for the level limit falls back to the IPP (bit-identical to
forcing it).
"""
import ctypes as C
import sys

import numpy as np
import pytest

from helpers import oracle_state_dict, rel_l2, run_oracle_frames

pytestmark = pytest.mark.gpu

FLUID, SOLID, EMPTY = 0x40, 0x20, 0x10


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_solve(orc, mat, scale, b):
    n_i, n_j = mat.shape
    xo = np.zeros(n_i * n_j)
    rr = C.c_double()
    orc.lib().orc_pcg_converged(np.ascontiguousarray(mat.ravel()), n_i, n_j, scale, b, xo, 400000, 1e-12,
                                C.byref(rr))
    return xo


def _dam_system(fg, orc, n, frames):
    p, st = orc.dam_break(n, ppc=4, tol=1e-8)
    run_oracle_frames(orc, p, st, frames)
    mat = st.material.reshape(n, n).copy()
    b = fg.pressure_rhs(mat, st.u.reshape(n + 1, n), st.v.reshape(n, n + 1), p.dx)
    return mat, b, (1.0 / 60) / (p.fluid_density * p.dx * p.dx)


def _solve_both(ctx, fg, mat, scale, b, tol):
    out = {}
    try:
        for kind in (fg.PRECOND_MG, fg.PRECOND_IPP):
            ctx.set_preconditioner(kind)
            out[kind] = ctx.poisson_solve(mat, scale, b, tol=tol)
    finally:
        ctx.set_preconditioner(fg.PRECOND_MG)
    return out


@pytest.mark.parametrize("n,frames", [(128, 4), (384, 3)])
def test_mg_and_ipp_on_dam_break_systems(fg, orc, ctx, n, frames):
    mat, b, scale = _dam_system(fg, orc, n, frames)
    xo = _oracle_solve(orc, mat, scale, b)
    res = _solve_both(ctx, fg, mat, scale, b, 1e-9)
    (xm, sm), (xi, si) = res[fg.PRECOND_MG], res[fg.PRECOND_IPP]
    em, ei = rel_l2(xm, xo), rel_l2(xi, xo)
    _log(f"dam {n}^2: MG {sm.iterations} iters ({sm.sweeps} sweeps) rel L2 {em:.1e}; "
         f"IPP {si.iterations} iters rel L2 {ei:.1e}")
    assert sm.converged and si.converged
    assert sm.rel_residual <= 1e-9 and si.rel_residual <= 1e-9
    assert em <= 1e-5 and ei <= 1e-5
    assert sm.iterations <= 60
    assert 4 * sm.iterations <= si.iterations


def test_mg_on_fragmented_liquid(fg, orc, ctx):
    n = 512
    mat, u, v, b, scale = fg.isolated_projection_system(n, 4)
    xo = _oracle_solve(orc, mat, scale, b)
    res = _solve_both(ctx, fg, mat, scale, b, 1e-9)
    (xm, sm), (xi, si) = res[fg.PRECOND_MG], res[fg.PRECOND_IPP]
    _log(f"noise {n}^2: MG {sm.iterations} iters, IPP {si.iterations} iters")
    assert sm.converged and rel_l2(xm, xo) <= 1e-5 and rel_l2(xi, xo) <= 1e-5
    assert sm.iterations <= 60 and sm.iterations < si.iterations


def _tank(n_i, n_j, seed):
    """Solid border, liquid in the first two thirds of the rows with a solid
    block inside it, air above: every liquid cell connects to the air, so
    the system is regular."""
    rng = np.random.default_rng(seed)
    mat = np.full((n_i, n_j), EMPTY, np.int8)
    mat[: max(1, (2 * n_i) // 3), :] = FLUID
    mat[n_i // 4: n_i // 3, n_j // 4: n_j // 2] = SOLID
    mat[0, :] = mat[-1, :] = mat[:, 0] = mat[:, -1] = SOLID
    b = np.where(mat.ravel() == FLUID, rng.standard_normal(n_i * n_j), 0.0)
    return mat, b


@pytest.mark.parametrize("shape", [(37, 90), (33, 33), (5, 5), (200, 17), (65, 64), (17, 513)])
def test_mg_odd_shapes(fg, orc, ctx, shape):
    mat, b = _tank(*shape, seed=sum(shape))
    if not (mat == FLUID).any():
        pytest.skip("no liquid")
    xo = _oracle_solve(orc, mat, 1.0, b)
    x, st = ctx.poisson_solve(mat, 1.0, b, tol=1e-9)
    e = rel_l2(x, xo)
    _log(f"shape {shape}: {st.iterations} iters, rel L2 {e:.1e}")
    assert st.converged and st.rel_residual <= 1e-9 and e <= 1e-5


def test_too_skewed_for_levels_uses_ipp(fg, orc, ctx):
    # 4 x 70000: halving the long side to the bottom size needs 13 coarse
    # levels, more than the hierarchy holds; liquid only near one end, air
    # elsewhere
    n_i, n_j = 4, 70000
    mat = np.full((n_i, n_j), EMPTY, np.int8)
    mat[1:3, 1:200] = FLUID
    mat[0, :] = mat[-1, :] = mat[:, 0] = mat[:, -1] = SOLID
    rng = np.random.default_rng(3)
    b = np.where(mat.ravel() == FLUID, rng.standard_normal(n_i * n_j), 0.0)
    res = _solve_both(ctx, fg, mat, 1.0, b, 1e-9)
    (xm, sm), (xi, si) = res[fg.PRECOND_MG], res[fg.PRECOND_IPP]
    _log(f"4 x 70000: default {sm.iterations} iters, IPP {si.iterations} iters")
    assert sm.converged and si.converged
    assert sm.iterations == si.iterations
    assert np.array_equal(xm.view(np.uint64), xi.view(np.uint64))
    assert rel_l2(xm, _oracle_solve(orc, mat, 1.0, b)) <= 1e-5


def test_substep_parity_with_ipp(fg, orc, ctx):
    """The IPP path keeps the substep's parity bar (the MG path is the
    default for every other parity test)."""
    n, ppc = 64, 4
    p, st = orc.dam_break(n, ppc=ppc, tol=1e-12)
    run_oracle_frames(orc, p, st, 4)
    vel = orc.lib().orc_max_particle_velocity(st.ref())
    dt = np.float32(min(5.0 / (vel + 1e-15), 1.0 / 30))
    ctx.set_preconditioner(fg.PRECOND_IPP)
    try:
        sc = fg.Scene(ctx, fg.make_params(n, ppc=ppc, tol=1e-10))
        walls, _ = fg.dam_break_geometry(n)
        sc.set_friction(np.zeros(len(walls), np.float32))
        sc.set_state(oracle_state_dict(st))
        gs = sc.substep(float(dt))
        g = sc.get_state()
        sc.close()
    finally:
        ctx.set_preconditioner(fg.PRECOND_MG)
    ostats = orc.OrcStats()
    orc.lib().orc_substep(C.byref(p), st.ref(), C.c_float(dt), C.byref(ostats))
    N = st.n
    _log(f"IPP substep: {gs.pressure_iters} iters")
    assert gs.pressure_rel_resid <= 1e-10
    assert rel_l2(g["pressure"], st.pressure) <= 1e-5
    assert max(rel_l2(g["vx"], st.vx[:N]), rel_l2(g["vy"], st.vy[:N])) <= 1e-5


@pytest.mark.parametrize("n,frames", [(128, 3), (300, 2), (384, 2)])
def test_fused_vcycle_equals_per_level(fg, orc, ctx, n, frames):
    """The fused V-cycle levels (residual + restriction, prolongation +
    post-smoothing, each one launch over LDS tiles) form every value by the
    per-level kernels' expression: the solves must be bit-identical."""
    mat, b, scale = _dam_system(fg, orc, n, frames)
    out = {}
    try:
        for fused in (True, False):
            ctx.set_vcycle_fused(fused)
            out[fused] = ctx.poisson_solve(mat, scale, b, tol=1e-6)
    finally:
        ctx.set_vcycle_fused(True)
    (x1, s1), (x0, s0) = out[True], out[False]
    _log(f"  {n}^2: fused {s1.iterations} iterations, per-level {s0.iterations}")
    assert s1.iterations == s0.iterations and s1.sweeps == s0.sweeps
    assert np.array_equal(x1.view(np.uint8), x0.view(np.uint8))


def test_fused_vcycle_batch_equals_per_level(fg, ctx):
    specs = fg.mixed_scene_specs(7, master_seed=41, resolutions=(64, 96, 128, 192))
    fields = ("px", "py", "vx", "vy", "pid", "u", "v", "pressure")
    got = {}
    try:
        for fused in (True, False):
            ctx.set_vcycle_fused(fused)
            scenes = [fg.make_scene(ctx, sp, tol=1e-6) for sp in specs]
            for _ in range(2):
                fg.batch_step_frame(ctx, scenes)
            got[fused] = [sc.get_state() for sc in scenes]
            for sc in scenes:
                sc.close()
    finally:
        ctx.set_vcycle_fused(True)
    for k in range(len(specs)):
        for f in fields:
            assert np.array_equal(got[True][k][f].view(np.uint8), got[False][k][f].view(np.uint8)), (k, f)


@pytest.mark.parametrize("shape", [(37, 90), (200, 17), (65, 64), (17, 513), (300, 131)])
def test_fused_vcycle_odd_shapes_equal_per_level(fg, ctx, shape):
    """Odd and skewed grids put the fused kernels' tiles across every kind of
    grid edge (partial tiles, clamped halos, levels of one or two cells a
    side): still bit-identical to the per-level kernels."""
    mat, b = _tank(*shape, seed=3 * sum(shape))
    if not (mat == FLUID).any():
        pytest.skip("no liquid")
    out = {}
    try:
        for fused in (True, False):
            ctx.set_vcycle_fused(fused)
            out[fused] = ctx.poisson_solve(mat, 1.0, b, tol=1e-9)
    finally:
        ctx.set_vcycle_fused(True)
    (x1, s1), (x0, s0) = out[True], out[False]
    _log(f"shape {shape}: {s1.iterations} iterations either way")
    assert s1.iterations == s0.iterations
    assert np.array_equal(x1.view(np.uint8), x0.view(np.uint8))
