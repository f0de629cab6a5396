"""NB-FLIP (NBFlipSolver, nbflipsolver.cpp) on the GPU against the oracle.

The bar is FlipSolver's: bit-exact for integer/index work and for every
float stage upstream of the pressure solve -- the band prune, the Euler pass,
the particle and combined level sets, both level-set extrapolations, the
materials, the right-hand side, the band reseed -- and relative L2 <= 1e-5
for pressure and velocities against the oracle's FP64 solve converged to
1e-12.  Frames run from the same first frame on both sides; the solves
converge (GPU 1e-10, oracle 1e-12), so the runs differ by solver round-off
only, which the frames carry forward.
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


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


SOURCE_BOX = (0.55, 0.75, 0.55, 0.8)
SINK_BOX = (0.05, 0.12, 0.7, 0.9)


def _objects(fg, n):
    walls, fluid = fg.dam_break_geometry(n)
    return walls, [fg.box(*SOURCE_BOX)], [fg.box(*SINK_BOX)], fluid


@pytest.mark.parametrize("n", [64, 97])
def test_nb_init_matches_oracle(fg, orc, ctx, n):
    _log(f"NB-FLIP first frame n={n}")
    p, st = orc.nb_dam_break(n, ppc=8)
    sc = fg.nb_dam_break_scene(ctx, n, ppc=8)
    try:
        g = sc.get_state()
    finally:
        sc.close()
    assert len(g["px"]) == 0 and st.n == 0
    assert np.array_equal(g["material"], st.material)
    assert np.array_equal(g["solid_id"], st.solid_id)
    assert np.array_equal(bits(g["solid_sdf"]), bits(st.solid_sdf))
    assert np.array_equal(bits(g["fluid_sdf"]), bits(st.fluid_sdf))
    assert not g["particle_counts"].any()


def _source_pair(fg, orc, ctx, n, ppc, emitters, seed):
    walls, srcs, sinks, fluids = _objects(fg, n)
    p = orc.make_params(n, ppc=ppc, tol=1e-12, sim=orc.SIM_NBFLIP, emitters=emitters, reseed_seed=seed)
    st = orc.init_nb_scene(p, solids=walls, sources=srcs, sinks=sinks, fluids=fluids)
    sc = fg.Scene(ctx, fg.make_params(n, ppc=ppc, tol=1e-10, sim=fg.SIM_NBFLIP))
    sc.init_objects(solids=walls, sources=srcs, sinks=sinks, fluids=fluids,
                    friction=np.zeros(len(walls), np.float32), emitters=emitters, seed=seed)
    return p, st, sc


def test_nb_init_with_sources_and_sinks(fg, orc, ctx):
    n = 80
    p, st, sc = _source_pair(fg, orc, ctx, n, 8, [(0.5, -0.25, True)], 11)
    try:
        g = sc.get_state()
    finally:
        sc.close()
    assert (st.material == orc.SOURCE).any() and (st.material == orc.SINK).any()
    assert np.array_equal(g["material"], st.material)
    assert np.array_equal(g["emitter_id"], st.emitter_id)
    assert np.array_equal(bits(g["fluid_sdf"]), bits(st.fluid_sdf))


def _check_substep(fg, g, gs, st, tag):
    N = st.n
    _log(f"  {tag}: particles gpu={len(g['px'])} oracle={N} iters={gs.pressure_iters} "
         f"resid={gs.pressure_rel_resid:.2e} reseeded={gs.particles_reseeded} killed={gs.particles_killed}")
    assert len(g["px"]) == N
    assert np.array_equal(g["pid"], st.pid[:N])
    assert np.array_equal(bits(g["px"]), bits(st.px[:N]))
    assert np.array_equal(bits(g["py"]), bits(st.py[:N]))
    assert np.array_equal(g["material"], st.material)
    assert np.array_equal(g["particle_counts"], st.particle_counts)
    assert np.array_equal(bits(g["fluid_sdf"]), bits(st.fluid_sdf))
    assert np.array_equal(bits(g["saved_u"]), bits(st.saved_u))
    assert np.array_equal(bits(g["saved_v"]), bits(st.saved_v))
    assert np.array_equal(bits(g["rhs"]), bits(st.rhs))
    assert np.array_equal(g["u_valid"], st.u_valid)
    assert np.array_equal(g["v_valid"], st.v_valid)
    assert gs.pressure_rel_resid <= 1e-10
    e = {"p": rel_l2(g["pressure"], st.pressure), "u": rel_l2(g["u"], st.u), "v": rel_l2(g["v"], st.v),
         "vx": rel_l2(g["vx"], st.vx[:N]), "vy": rel_l2(g["vy"], st.vy[:N])}
    _log("    rel L2: " + " ".join(f"{k}={v:.2e}" for k, v in e.items()))
    assert max(e.values()) <= 1e-5, e


@pytest.mark.parametrize("n,ppc,frames", [(64, 8, 3), (128, 8, 2), (96, 4, 4)])
def test_nb_substep_parity(fg, orc, ctx, n, ppc, frames):
    """One NBFlipSolver::step from an oracle state handed over mid-run (the
    level set, the counts and the reseed counter with it)."""
    _log(f"NB-FLIP substep parity n={n} ppc={ppc} after {frames} oracle frames")
    p, st = orc.nb_dam_break(n, ppc=ppc, tol=1e-12)
    run_oracle_frames(orc, p, st, frames)
    assert st.n > 0
    vel = orc.lib().orc_max_particle_velocity(st.ref())
    dt = np.float32(min(5.0 / (vel + 1e-15), 1.0 / 30))
    sc = fg.Scene(ctx, fg.make_params(n, ppc=ppc, tol=1e-10, sim=fg.SIM_NBFLIP))
    try:
        sc.set_friction(np.zeros(4, np.float32))
        sc.set_state(oracle_state_dict(st))
        gs = sc.substep(float(dt))
        g = sc.get_state()
    finally:
        sc.close()
    orc.lib().orc_nb_substep(C.byref(p), st.ref(), C.c_float(dt), C.byref(orc.OrcStats()))
    _check_substep(fg, g, gs, st, f"dt={dt:.4e}")


@pytest.mark.parametrize("n,ppc,frames", [(32, 8, 4), (64, 8, 3)])
def test_nb_frames_vs_oracle(fg, orc, ctx, n, ppc, frames):
    """NBFlipSolver::stepFrame from the first frame (no particles, the
    level set of the initial fluid) beside orc_step_frame."""
    _log(f"NB-FLIP frames n={n} ppc={ppc}")
    p, st = orc.nb_dam_break(n, ppc=ppc, tol=1e-12)
    sc = fg.nb_dam_break_scene(ctx, n, ppc=ppc, tol=1e-10)
    L = orc.lib()
    try:
        for f in range(frames):
            gs = sc.step_frame()
            os_ = orc.OrcStats()
            subs = L.orc_step_frame(C.byref(p), st.ref(), C.byref(os_))
            _log(f"  frame {f}: substeps gpu={gs.substeps} oracle={subs} particles gpu={gs.particles} "
                 f"oracle={st.n} vmax gpu={gs.max_velocity:.5e} oracle={os_.max_velocity:.5e}")
            assert gs.converged
            assert gs.substeps == subs
            assert abs(gs.last_dt - os_.last_dt) <= 1e-5 * os_.last_dt
            assert abs(gs.max_velocity - os_.max_velocity) <= 1e-5 * max(os_.max_velocity, 1e-30)
            assert gs.particles == st.n
        g = sc.get_state()
    finally:
        sc.close()
    N = st.n
    assert N > 0 and np.array_equal(g["pid"], st.pid[:N])
    e = {"px": rel_l2(g["px"], st.px[:N]), "py": rel_l2(g["py"], st.py[:N]),
         "vx": rel_l2(g["vx"], st.vx[:N]), "vy": rel_l2(g["vy"], st.vy[:N]),
         "u": rel_l2(g["u"], st.u), "v": rel_l2(g["v"], st.v), "sdf": rel_l2(g["fluid_sdf"], st.fluid_sdf)}
    _log("  after the frames, rel L2: " + " ".join(f"{k}={v:.2e}" for k, v in e.items()))
    assert max(e.values()) <= 1e-5, e
    assert int(np.count_nonzero(g["material"] != st.material)) <= max(2, n * n // 2000)


def test_nb_sources_and_sinks_substeps_vs_oracle(fg, orc, ctx):
    """Sources (updateGridFromSources' distance, the source faces, SOURCE
    cells in the reseed) and a sink, substep by substep from the first frame,
    each substep bit-exact upstream of the solve."""
    n, ppc = 64, 8
    p, st, sc = _source_pair(fg, orc, ctx, n, ppc, [(0.5, -0.25, True)], 7)
    L = orc.lib()
    try:
        for k in range(4):
            vel = L.orc_max_particle_velocity(st.ref())
            dt = np.float32(min(5.0 / (vel + 1e-15), 1.0 / 60))
            sc.set_state(oracle_state_dict(st))
            gs = sc.substep(float(dt))
            g = sc.get_state()
            L.orc_nb_substep(C.byref(p), st.ref(), C.c_float(dt), C.byref(orc.OrcStats()))
            _check_substep(fg, g, gs, st, f"substep {k}")
        assert (st.material == orc.SOURCE).any() and st.n > 0
    finally:
        sc.close()


def test_nb_and_flip_batch_equals_solo(fg, ctx):
    """A batch mixing the two solver classes runs them as two groups: every
    scene ends bit-identical to its solo run."""
    def make():
        return [fg.nb_dam_break_scene(ctx, 48, ppc=8), fg.dam_break_scene(ctx, 40, ppc=4),
                fg.nb_dam_break_scene(ctx, 33, ppc=4, fluid_box=(0.2, 0.9, 0.1, 0.6))]
    fields = ("px", "py", "vx", "vy", "pid", "u", "v", "fluid_sdf", "material", "pressure")
    solo = []
    for sc in make():
        for _ in range(3):
            sc.step_frame()
        solo.append(sc.get_state())
        sc.close()
    batch = make()
    try:
        for _ in range(3):
            fg.batch_step_frame(ctx, batch)
        got = [sc.get_state() for sc in batch]
    finally:
        for sc in batch:
            sc.close()
    for k in range(len(solo)):
        for f in fields:
            assert np.array_equal(bits(got[k][f]), bits(solo[k][f])), (k, f)
    assert len(solo[0]["px"]) > 0 and len(solo[2]["px"]) > 0


def test_nb_refusals(fg, ctx):
    sc = fg.nb_dam_break_scene(ctx, 32)
    try:
        with pytest.raises(fg.FgError) as e:
            sc.stage(fg.STAGE_ADVECT, 0.01)
        assert e.value.code == fg.E_INVALID
        g = sc.get_state()
        g["fluid_sdf"] = None
        with pytest.raises(fg.FgError) as e:
            sc.set_state(g)
        assert e.value.code == fg.E_INVALID
    finally:
        sc.close()
    with pytest.raises(fg.FgError) as e:
        fg.Scene(ctx, fg.make_params(32, sim=fg.SIM_NBFLIP, density_mode=fg.DENSITY_ON))
    assert e.value.code == fg.E_INVALID


# ---------------------------------------------------------------------------
# NB-FLIP with viscosity: the advected viscosity grid, its band combination,
# the sources' viscosity and the reseed's (nbflipsolver.cpp:66-113, 115-201,
# 338-346, 400-504), then the light model and the second projection
# ---------------------------------------------------------------------------
def _nb_visc_pair(fg, orc, ctx, n, ppc, seed=3):
    walls, srcs, sinks, fluids = _objects(fg, n)
    fluids = fluids + [fg.box(0.1, 0.25, 0.1, 0.3)]
    emit = [(0.5, -0.25, True, 7.5)]
    p = orc.make_params(n, ppc=ppc, tol=1e-12, sim=orc.SIM_NBFLIP, emitters=emit, reseed_seed=seed, viscosity=True,
                        fluid_viscosity=[2.0, 20.0])
    st = orc.init_nb_scene(p, solids=walls, sources=srcs, sinks=sinks, fluids=fluids)
    sc = fg.Scene(ctx, fg.make_params(n, ppc=ppc, tol=1e-10, sim=fg.SIM_NBFLIP, viscosity=True))
    sc.init_objects(solids=walls, sources=srcs, sinks=sinks, fluids=fluids, friction=np.zeros(len(walls), np.float32),
                    emitters=emit, seed=seed, fluid_viscosity=[2.0, 20.0])
    return p, st, sc


def test_nb_viscous_substeps_vs_oracle(fg, orc, ctx):
    n, ppc = 64, 8
    p, st, sc = _nb_visc_pair(fg, orc, ctx, n, ppc)
    L = orc.lib()
    try:
        g0 = sc.get_state()
        assert np.array_equal(bits(g0["viscosity"]), bits(st.viscosity)) and st.viscosity.max() == 20.0
        for k in range(4):
            vel = L.orc_max_particle_velocity(st.ref())
            dt = np.float32(min(5.0 / (vel + 1e-15), 1.0 / 60))
            sc.set_state(oracle_state_dict(st))
            gs = sc.substep(float(dt))
            g = sc.get_state()
            L.orc_nb_substep(C.byref(p), st.ref(), C.c_float(dt), C.byref(orc.OrcStats()))
            N = st.n
            _log(f"  NB viscous substep {k}: particles {N}, viscosity iters {gs.viscosity_iters}")
            assert len(g["px"]) == N and np.array_equal(g["pid"], st.pid[:N])
            assert np.array_equal(bits(g["px"]), bits(st.px[:N]))
            assert np.array_equal(bits(g["particle_viscosity"]), bits(st.pvisc[:N]))
            assert np.array_equal(bits(g["viscosity"]), bits(st.viscosity))
            assert np.array_equal(bits(g["fluid_sdf"]), bits(st.fluid_sdf))
            assert np.array_equal(g["material"], st.material)
            e = {"p": rel_l2(g["pressure"], st.pressure), "u": rel_l2(g["u"], st.u), "v": rel_l2(g["v"], st.v),
                 "vx": rel_l2(g["vx"], st.vx[:N]), "vy": rel_l2(g["vy"], st.vy[:N])}
            assert max(e.values()) <= 1e-5, e
        assert st.n > 0 and (st.pvisc[:st.n] == np.float32(7.5)).any()
    finally:
        sc.close()


def test_nb_viscous_frames_vs_oracle(fg, orc, ctx):
    n, ppc = 48, 8
    p, st, sc = _nb_visc_pair(fg, orc, ctx, n, ppc)
    L = orc.lib()
    try:
        for f in range(3):
            gs = sc.step_frame()
            os_ = orc.OrcStats()
            subs = L.orc_step_frame(C.byref(p), st.ref(), C.byref(os_))
            _log(f"  frame {f}: substeps gpu={gs.substeps} oracle={subs} particles gpu={gs.particles} oracle={st.n}")
            assert gs.converged and gs.substeps == subs and gs.particles == st.n
        g = sc.get_state()
    finally:
        sc.close()
    N = st.n
    assert np.array_equal(g["pid"], st.pid[:N])
    e = {"px": rel_l2(g["px"], st.px[:N]), "u": rel_l2(g["u"], st.u), "v": rel_l2(g["v"], st.v),
         "visc": rel_l2(g["viscosity"], st.viscosity), "pvisc": rel_l2(g["particle_viscosity"], st.pvisc[:N])}
    _log("  after the frames, rel L2: " + " ".join(f"{k}={v:.2e}" for k, v in e.items()))
    assert max(e.values()) <= 1e-5, e
