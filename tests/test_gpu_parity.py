"""HIP path vs the CPU oracle on identical seeded input state.

Bar: bit-exact for integer/index work and for every float stage upstream of
the pressure solve (the kernels and the oracle are both built with
-ffp-contract=off and use the same canonical particle order); pressure
relative L2 <= 1e-5 against the oracle's FP64 solve converged to 1e-12, and
velocities relative L2 <= 1e-5.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from helpers import oracle_state_dict, rel_l2, run_oracle_frames

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,ppc", [(32, 4), (128, 4), (64, 8)])
def test_init_geometry_matches_oracle(fg, orc, ctx, n, ppc):
    _log(f"init geometry n={n} ppc={ppc}")
    p, st = orc.dam_break(n, ppc=ppc)
    sc = fg.dam_break_scene(ctx, n, ppc=ppc)
    g = sc.get_state()
    sc.close()
    assert np.array_equal(g["material"], st.material)
    assert np.array_equal(g["solid_id"], st.solid_id)
    assert np.array_equal(g["solid_sdf"].view(np.uint32), st.solid_sdf.view(np.uint32))
    assert len(g["px"]) == st.n
    assert np.array_equal(g["pid"], st.pid[: st.n])
    assert np.array_equal(g["px"].view(np.uint32), st.px[: st.n].view(np.uint32))
    assert np.array_equal(g["py"].view(np.uint32), st.py[: st.n].view(np.uint32))


def _prepared(orc, n, ppc, frames):
    p, st = orc.dam_break(n, ppc=ppc, tol=1e-12)
    if frames:
        run_oracle_frames(orc, p, st, frames)
    return p, st


@pytest.mark.parametrize("n,ppc,frames", [(32, 4, 3), (128, 4, 4), (64, 8, 6)])
def test_substep_parity(fg, orc, ctx, n, ppc, frames):
    _log(f"substep parity n={n} ppc={ppc} frames={frames}")
    p, st = _prepared(orc, n, ppc, frames)
    vel = orc.lib().orc_max_particle_velocity(st.ref())
    dt = np.float32(min(5.0 / (vel + 1e-15), 1.0 / 30))
    sc = fg.Scene(ctx, fg.make_params(n, ppc=ppc, tol=1e-10))
    walls, _ = fg.dam_break_geometry(n)
    sc.set_friction(np.zeros(len(walls), np.float32))
    sc.set_state(oracle_state_dict(st))
    gs = sc.substep(float(dt))
    g = sc.get_state()
    sc.close()
    ostats = orc.OrcStats()
    orc.lib().orc_substep(C.byref(p), st.ref(), C.c_float(dt), C.byref(ostats))
    N = st.n
    _log(f"  particles gpu={len(g['px'])} oracle={N} iters={gs.pressure_iters} sweeps={gs.pressure_sweeps} "
         f"resid={gs.pressure_rel_resid:.3e}")
    # integer / index work and all float stages upstream of the solve: bit-exact
    assert len(g["px"]) == N
    assert np.array_equal(g["pid"], st.pid[:N])
    assert np.array_equal(g["px"].view(np.uint32), st.px[:N].view(np.uint32))
    assert np.array_equal(g["py"].view(np.uint32), st.py[:N].view(np.uint32))
    assert np.array_equal(g["material"], st.material)
    assert np.array_equal(g["particle_counts"], st.particle_counts)
    assert np.array_equal(g["fluid_sdf"].view(np.uint32), st.fluid_sdf.view(np.uint32))
    assert np.array_equal(g["saved_u"].view(np.uint32), st.saved_u.view(np.uint32))
    assert np.array_equal(g["saved_v"].view(np.uint32), st.saved_v.view(np.uint32))
    assert np.array_equal(g["rhs"].view(np.uint64), st.rhs.view(np.uint64))
    assert np.array_equal(g["u_valid"], st.u_valid)
    assert np.array_equal(g["v_valid"], st.v_valid)
    # solve and everything downstream of it: tolerance
    assert gs.pressure_rel_resid <= 1e-10
    e_p = rel_l2(g["pressure"], st.pressure)
    e_u = rel_l2(g["u"], st.u)
    e_v = rel_l2(g["v"], st.v)
    e_vx = rel_l2(g["vx"], st.vx[:N])
    e_vy = rel_l2(g["vy"], st.vy[:N])
    _log(f"  rel L2: p={e_p:.2e} u={e_u:.2e} v={e_v:.2e} vx={e_vx:.2e} vy={e_vy:.2e}")
    assert e_p <= 1e-5
    assert max(e_u, e_v, e_vx, e_vy) <= 1e-5


def test_substep_parity_at_1e6(fg, orc, ctx):
    """At the production tolerance (FP64 relative residual 1e-6)."""
    n, ppc = 128, 4
    _log("substep parity at solver tolerance 1e-6")
    p, st = _prepared(orc, n, ppc, 5)
    vel = orc.lib().orc_max_particle_velocity(st.ref())
    dt = np.float32(min(5.0 / (vel + 1e-15), 1.0 / 30))
    sc = fg.Scene(ctx, fg.make_params(n, ppc=ppc, tol=1e-6))
    sc.set_friction(np.zeros(4, np.float32))
    sc.set_state(oracle_state_dict(st))
    gs = sc.substep(float(dt))
    g = sc.get_state()
    sc.close()
    orc.lib().orc_substep(C.byref(p), st.ref(), C.c_float(dt), C.byref(orc.OrcStats()))
    N = st.n
    e_p = rel_l2(g["pressure"], st.pressure)
    e_vx = rel_l2(g["vx"], st.vx[:N])
    e_vy = rel_l2(g["vy"], st.vy[:N])
    _log(f"  resid={gs.pressure_rel_resid:.2e} iters={gs.pressure_iters} p={e_p:.2e} vx={e_vx:.2e} vy={e_vy:.2e}")
    assert gs.pressure_rel_resid <= 1e-6
    assert e_p <= 1e-5
    assert max(e_vx, e_vy) <= 1e-5


def test_warm_started_second_substep(fg, orc, ctx):
    """The second of two substeps starts its pressure solve from the first
    substep's pressure (the oracle, like the reference, starts from zero);
    both solves converge to the same solution of the same system."""
    n, ppc = 64, 4
    _log("two substeps, the second warm-started")
    p, st = _prepared(orc, n, ppc, 4)
    sc = fg.Scene(ctx, fg.make_params(n, ppc=ppc, tol=1e-10))
    walls, _ = fg.dam_break_geometry(n)
    sc.set_friction(np.zeros(len(walls), np.float32))
    sc.set_state(oracle_state_dict(st))
    for k in range(2):
        vel = orc.lib().orc_max_particle_velocity(st.ref())
        dt = np.float32(min(5.0 / (vel + 1e-15), 1.0 / 30))
        gs = sc.substep(float(dt))
        orc.lib().orc_substep(C.byref(p), st.ref(), C.c_float(dt), C.byref(orc.OrcStats()))
        _log(f"  substep {k}: iters={gs.pressure_iters} sweeps={gs.pressure_sweeps} resid={gs.pressure_rel_resid:.2e}")
        assert gs.pressure_rel_resid <= 1e-10
    g = sc.get_state()
    sc.close()
    N = st.n
    assert len(g["px"]) == N
    e_p = rel_l2(g["pressure"], st.pressure)
    e_vx = rel_l2(g["vx"], st.vx[:N])
    e_vy = rel_l2(g["vy"], st.vy[:N])
    _log(f"  rel L2 after two substeps: p={e_p:.2e} vx={e_vx:.2e} vy={e_vy:.2e}")
    assert e_p <= 1e-5
    assert max(e_vx, e_vy) <= 1e-5


def test_poisson_solve_vs_oracle(fg, orc, ctx):
    _log("poisson solve 192x160")
    rng = np.random.default_rng(7)
    I, J = 192, 160
    mat = np.where(rng.random((I, J)) < 0.7, orc.FLUID, orc.EMPTY).astype(np.int8)
    mat[0, :] = mat[-1, :] = mat[:, 0] = mat[:, -1] = orc.SOLID
    mat[rng.random((I, J)) < 0.02] = orc.SOLID
    b = rng.standard_normal(I * J)
    b[mat.ravel() & 0x40 == 0] = 0.0
    scale = 0.37
    x, stt = ctx.poisson_solve(mat, scale, b, tol=1e-9)
    xo = np.zeros(I * J)
    rr = C.c_double()
    orc.lib().orc_pcg_converged(mat.ravel(), I, J, scale, b, xo, 100000, 1e-13, C.byref(rr))
    _log(f"  gpu iters={stt.iterations} sweeps={stt.sweeps} resid={stt.rel_residual:.2e}")
    assert stt.converged and stt.rel_residual <= 1e-9
    assert rel_l2(x, xo) <= 1e-6


def test_frames_run(fg, orc, ctx):
    """Several frames from the dam-break init: live counts and a bounded state."""
    _log("3 frames at 128")
    sc = fg.dam_break_scene(ctx, 128, ppc=4)
    subs = 0
    for _ in range(3):
        s = sc.step_frame()
        subs += s.substeps
        assert s.converged
    g = sc.get_state()
    sc.close()
    assert subs >= 3
    assert np.all(np.isfinite(g["vx"])) and np.all(np.isfinite(g["u"]))
    assert len(g["px"]) > 17000


def test_substep_parity_with_friction(fg, orc, ctx):
    """Nonzero wall friction (updateVelocityFromSolids, flipsolver2d.cpp:1077-1104):
    the GPU's friction branch against the oracle's, after frames that ran
    with it."""
    n, ppc = 64, 4
    fr = np.array([0.3, 0.1, 0.5, 0.2], np.float32)
    p, st = orc.dam_break(n, ppc=ppc, tol=1e-12, friction=tuple(fr))
    run_oracle_frames(orc, p, st, 4)
    vel = orc.lib().orc_max_particle_velocity(st.ref())
    dt = np.float32(min(5.0 / (vel + 1e-15), 1.0 / 30))
    sc = fg.Scene(ctx, fg.make_params(n, ppc=ppc, tol=1e-10))
    sc.set_friction(fr)
    sc.set_state(oracle_state_dict(st))
    gs = sc.substep(float(dt))
    g = sc.get_state()
    sc.close()
    ostats = orc.OrcStats()
    orc.lib().orc_substep(C.byref(p), st.ref(), C.c_float(dt), C.byref(ostats))
    N = st.n
    _log(f"friction substep: {gs.pressure_iters} iters")
    assert len(g["px"]) == N
    assert np.array_equal(g["px"].view(np.uint32), st.px[:N].view(np.uint32))
    assert rel_l2(g["pressure"], st.pressure) <= 1e-5
    assert max(rel_l2(g["u"], st.u), rel_l2(g["v"], st.v)) <= 1e-5
    assert max(rel_l2(g["vx"], st.vx[:N]), rel_l2(g["vy"], st.vy[:N])) <= 1e-5


def test_full_grid_keeps_exact_sdf(fg, orc, ctx):
    """A grid whose every cell holds particles has no cell outside the
    liquid, so extrapolateLevelsetInside keeps every fluid-SDF value: the
    values the fast SDF pass skips (cells holding particles are inside) must
    come back exact from the fallback pass."""
    n, ppc = 24, 4
    p = orc.make_params(n, ppc=ppc, gravity=(0.0, 0.0), tol=1e-12)
    st = orc.State(n, n, n * n * ppc)
    st.material[:] = orc.FLUID
    rng = np.random.default_rng(12)
    k = 0
    for c in range(n * n):
        i, j = divmod(c, n)
        for q in range(ppc):
            st.px[k], st.py[k] = i + 0.5 * ((q >> 1) + rng.random()), j + 0.5 * ((q & 1) + rng.random())
            st.pid[k] = k
            k += 1
    st.n = k
    st.solid_sdf[:] = 1.0
    dt = np.float32(1.0 / 60)
    sc = fg.Scene(ctx, fg.make_params(n, ppc=ppc, gravity=(0.0, 0.0), tol=1e-10))
    sc.set_state(oracle_state_dict(st))
    sc.substep(float(dt))
    g = sc.get_state()
    sc.close()
    orc.lib().orc_substep(C.byref(p), st.ref(), C.c_float(dt), C.byref(orc.OrcStats()))
    assert np.all(st.fluid_sdf < 0)   # no cell is outside
    assert np.array_equal(g["fluid_sdf"].view(np.uint32), st.fluid_sdf.view(np.uint32))
    assert np.array_equal(g["material"], st.material)


def test_post_projection_divergence_vs_oracle(fg, orc, ctx):
    """The projection's own error against the oracle: the GPU runs a
    substep's stages up to the grid it projects (the state's work grid), the
    oracle projects that same grid as step() does (rhs from the current
    materials, the matrix from the materials at the substep's start,
    flipsolver2d.cpp:417 and 446, FP64 PCG to 1e-12, pressure applied), and
    the two post-projection grids and their divergence on fluid cells are
    compared.  With the matrix one material update behind, the projected
    grid is divergence-free only where the materials around a cell did not
    change; the divergence left there is bounded relative to the grid's
    divergence before the projection."""
    n, ppc = 128, 4
    p, st = _prepared(orc, n, ppc, 4)
    mat_a = st.material.copy()   # the materials the substep's matrix is built from
    vel = orc.lib().orc_max_particle_velocity(st.ref())
    dt = np.float32(min(5.0 / (vel + 1e-15), 1.0 / 30))
    sc = fg.Scene(ctx, fg.make_params(n, ppc=ppc, tol=1e-10))
    sc.set_friction(np.zeros(4, np.float32))
    sc.set_state(oracle_state_dict(st))
    for stage in (fg.STAGE_ADVECT, fg.STAGE_REBIN, fg.STAGE_DENSITY, fg.STAGE_P2G, fg.STAGE_SDF,
                  fg.STAGE_LEVELSET, fg.STAGE_EXTRAPOLATE, fg.STAGE_FORCES):
        sc.stage(stage, float(dt))
    pre = sc.get_state()
    sc.stage(fg.STAGE_PROJECT, float(dt))
    post = sc.get_state()
    sc.close()
    L = orc.lib()
    o = orc.State(n, n, 1)
    o.u[:] = pre["work_u"]
    o.v[:] = pre["work_v"]
    o.material[:] = pre["material"]
    rhs = np.zeros(n * n)
    L.orc_pressure_rhs(C.byref(p), o.ref(), rhs)
    assert np.array_equal(rhs.view(np.uint64), post["rhs"].view(np.uint64))
    x = np.zeros(n * n)
    rr = C.c_double()
    scale = float(dt) / (p.fluid_density * p.dx * p.dx)
    L.orc_pcg_converged(mat_a, n, n, scale, rhs, x, 400000, 1e-12, C.byref(rr))
    L.orc_apply_pressure(C.byref(p), o.ref(), C.c_float(dt), x)

    def div(u, v):
        u = np.asarray(u, np.float64).reshape(n + 1, n)
        v = np.asarray(v, np.float64).reshape(n, n + 1)
        return (u[1:, :] - u[:-1, :]) + (v[:, 1:] - v[:, :-1])

    m_now = pre["material"].reshape(n, n)
    fluid = (m_now & 0x40) != 0
    changed = np.pad(m_now != mat_a.reshape(n, n), 1)
    near = np.zeros((n, n), bool)   # a material in the cell's 3x3 neighbourhood changed
    for di in range(3):
        for dj in range(3):
            near |= changed[di:di + n, dj:dj + n]
    steady = fluid & ~near
    d_pre = div(pre["work_u"], pre["work_v"])
    d_gpu, d_orc = div(post["work_u"], post["work_v"]), div(o.u, o.v)
    scale_div = np.abs(d_pre[fluid]).max()
    e_div = np.abs(d_gpu - d_orc)[fluid].max() / scale_div
    e_u, e_v = rel_l2(post["work_u"], o.u), rel_l2(post["work_v"], o.v)
    e_p = rel_l2(post["pressure"], x)
    left_gpu = np.abs(d_gpu[steady]).max() / scale_div
    left_orc = np.abs(d_orc[steady]).max() / scale_div
    _log(f"  {int(fluid.sum())} fluid cells ({int(steady.sum())} with unchanged materials around): max|div| before "
         f"{scale_div:.3e}; left on those, relative: gpu {left_gpu:.2e}, oracle {left_orc:.2e}; gpu - oracle: "
         f"div {e_div:.2e} (max, relative), u {e_u:.2e}, v {e_v:.2e}, p {e_p:.2e} (rel L2)")
    assert steady.sum() > 0.5 * fluid.sum()
    assert left_gpu <= 1e-5 and left_orc <= 1e-5
    assert e_div <= 1e-5
    assert max(e_u, e_v) <= 1e-5 and e_p <= 1e-5


def test_work_grid_follows_the_stages(fg, orc, ctx):
    """work_u / work_v are fluidVelocityGrid as step() holds it after each
    stage hook: the extrapolated grid after FG_STAGE_EXTRAPOLATE (the copy
    the saved grid keeps), saved + gravity after FG_STAGE_FORCES, and the
    substep's final grid once the substep is complete."""
    n, ppc = 64, 4
    p, st = _prepared(orc, n, ppc, 2)
    vel = orc.lib().orc_max_particle_velocity(st.ref())
    dt = np.float32(min(5.0 / (vel + 1e-15), 1.0 / 30))
    sc = fg.Scene(ctx, fg.make_params(n, ppc=ppc, tol=1e-8))
    sc.set_friction(np.zeros(4, np.float32))
    sc.set_state(oracle_state_dict(st))
    g0 = sc.get_state()
    assert np.array_equal(g0["work_u"], g0["u"]) and np.array_equal(g0["work_v"], g0["v"])
    for stage in (fg.STAGE_ADVECT, fg.STAGE_REBIN, fg.STAGE_DENSITY, fg.STAGE_P2G, fg.STAGE_SDF,
                  fg.STAGE_LEVELSET, fg.STAGE_EXTRAPOLATE):
        sc.stage(stage, float(dt))
    g1 = sc.get_state()
    assert np.array_equal(g1["work_u"], g1["saved_u"]) and np.array_equal(g1["work_v"], g1["saved_v"])
    sc.stage(fg.STAGE_FORCES, float(dt))
    g2 = sc.get_state()
    du, dv = g2["work_u"] - g2["saved_u"], g2["work_v"] - g2["saved_v"]
    assert np.ptp(du) <= 1e-3 * max(1.0, abs(du).max()) and abs(du).max() > 0   # gravity along i
    assert abs(dv).max() <= 1e-3 * abs(du).max()
    for stage in (fg.STAGE_PROJECT, fg.STAGE_SOLIDS, fg.STAGE_G2P, fg.STAGE_COUNT):
        sc.stage(stage, float(dt))
    g3 = sc.get_state()
    sc.close()
    assert np.array_equal(g3["work_u"], g3["u"]) and np.array_equal(g3["work_v"], g3["v"])
    assert not np.array_equal(g3["u"], g0["u"])
