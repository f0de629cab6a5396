"""Viscosity (the light model, viscositymodel.cpp:4-140) on the GPU against
the oracle.

FlipSolver::step with viscosityEnabled (flipsolver2d.cpp:412-462): no
density correction, centeredParamsToGrid's viscosity grid after P2G, and
after the projection and updateVelocityFromSolids the light model on U and
V, then a second projection with the substep's matrix.  The bar: bit-exact
for the particles' viscosity (seeding, sort, reseeding) and the viscosity
grid, and for everything upstream of the first pressure solve; pressure,
velocities and the grid relative L2 <= 1e-5 against the oracle, whose
viscosity CG and pressure solve converge in FP64 (1e-12).
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


def _pair(fg, orc, ctx, n, ppc, visc, emitters=(), sources=(), seed=42):
    walls, fluids = fg.dam_break_geometry(n)
    p = orc.make_params(n, ppc=ppc, tol=1e-12, viscosity=True, fluid_viscosity=[visc], emitters=emitters,
                        reseed_seed=seed)
    st = orc.init_scene_objects(p, solids=walls, sources=list(sources), fluids=fluids, seed=seed)
    sc = fg.Scene(ctx, fg.make_params(n, ppc=ppc, tol=1e-10, viscosity=True))
    sc.init_objects(solids=walls, sources=list(sources), fluids=fluids, friction=np.zeros(len(walls), np.float32),
                    emitters=emitters, seed=seed, fluid_viscosity=[visc])
    return p, st, sc


@pytest.mark.parametrize("n,visc", [(48, 5.0), (96, 0.5)])
def test_viscous_init_matches_oracle(fg, orc, ctx, n, visc):
    p, st, sc = _pair(fg, orc, ctx, n, 4, visc)
    try:
        g = sc.get_state()
    finally:
        sc.close()
    N = st.n
    assert len(g["px"]) == N and N > 0
    assert np.array_equal(bits(g["px"]), bits(st.px[:N]))
    assert np.array_equal(bits(g["particle_viscosity"]), bits(st.pvisc[:N]))
    assert np.array_equal(bits(g["viscosity"]), bits(st.viscosity))
    assert g["particle_viscosity"].max() == np.float32(visc) and g["particle_viscosity"].min() < visc


@pytest.mark.parametrize("n,ppc,visc,frames", [(64, 4, 5.0, 3), (128, 4, 0.5, 2), (64, 8, 100.0, 2)])
def test_viscous_substep_parity(fg, orc, ctx, n, ppc, visc, frames):
    _log(f"viscous substep n={n} ppc={ppc} viscosity={visc} after {frames} oracle frames")
    p, st, sc = _pair(fg, orc, ctx, n, ppc, visc)
    try:
        run_oracle_frames(orc, p, st, frames)
        vel = orc.lib().orc_max_particle_velocity(st.ref())
        dt = np.float32(min(5.0 / (vel + 1e-15), 1.0 / 30))
        sc.set_state(oracle_state_dict(st))
        gs = sc.substep(float(dt))
        g = sc.get_state()
    finally:
        sc.close()
    orc.lib().orc_substep(C.byref(p), st.ref(), C.c_float(dt), C.byref(orc.OrcStats()))
    N = st.n
    _log(f"  particles gpu={len(g['px'])} oracle={N} pressure iters={gs.pressure_iters} "
         f"viscosity iters={gs.viscosity_iters} resid={gs.pressure_rel_resid:.2e}")
    assert gs.converged and gs.viscosity_iters > 0
    assert len(g["px"]) == N
    assert np.array_equal(g["pid"], st.pid[:N])
    assert np.array_equal(bits(g["px"]), bits(st.px[:N]))
    assert np.array_equal(bits(g["py"]), bits(st.py[:N]))
    assert np.array_equal(bits(g["particle_viscosity"]), bits(st.pvisc[:N]))
    assert np.array_equal(bits(g["viscosity"]), bits(st.viscosity))
    assert np.array_equal(g["material"], st.material)
    assert np.array_equal(bits(g["saved_u"]), bits(st.saved_u))
    assert np.array_equal(bits(g["saved_v"]), bits(st.saved_v))
    e = {"p": rel_l2(g["pressure"], st.pressure), "rhs": rel_l2(g["rhs"], st.rhs), "u": rel_l2(g["u"], st.u),
         "v": rel_l2(g["v"], st.v), "vx": rel_l2(g["vx"], st.vx[:N]), "vy": rel_l2(g["vy"], st.vy[:N])}
    _log("  rel L2: " + " ".join(f"{k}={v:.2e}" for k, v in e.items()))
    assert max(e.values()) <= 1e-5, e


def test_viscosity_changes_the_flow(fg, ctx):
    """The same dam break with and without viscosity: the viscous one moves
    differently (the model is not a no-op) and every solve converges."""
    out = {}
    for visc in (False, True):
        walls, fluids = fg.dam_break_geometry(64)
        sc = fg.Scene(ctx, fg.make_params(64, ppc=4, viscosity=visc))
        sc.init_objects(solids=walls, fluids=fluids, fluid_viscosity=[50.0])
        for _ in range(4):
            st = sc.step_frame()
            assert st.converged
        out[visc] = sc.get_state()
        sc.close()
    assert rel_l2(out[True]["u"], out[False]["u"]) > 1e-3


@pytest.mark.parametrize("n,ppc,visc,frames", [(32, 4, 5.0, 4), (64, 4, 10.0, 3)])
def test_viscous_frames_vs_oracle(fg, orc, ctx, n, ppc, visc, frames):
    _log(f"viscous frames n={n} viscosity={visc}")
    p, st, sc = _pair(fg, orc, ctx, n, ppc, visc)
    L = orc.lib()
    try:
        for f in range(frames):
            gs = sc.step_frame()
            os_ = orc.OrcStats()
            subs = L.orc_step_frame(C.byref(p), st.ref(), C.byref(os_))
            _log(f"  frame {f}: substeps gpu={gs.substeps} oracle={subs} vmax gpu={gs.max_velocity:.5e} "
                 f"oracle={os_.max_velocity:.5e} viscosity iters={gs.viscosity_iters}")
            assert gs.converged and gs.substeps == subs
            assert abs(gs.max_velocity - os_.max_velocity) <= 1e-5 * max(os_.max_velocity, 1e-30)
            assert gs.particles == st.n
        g = sc.get_state()
    finally:
        sc.close()
    N = st.n
    assert np.array_equal(g["pid"], st.pid[:N])
    e = {"px": rel_l2(g["px"], st.px[:N]), "vx": rel_l2(g["vx"], st.vx[:N]), "vy": rel_l2(g["vy"], st.vy[:N]),
         "u": rel_l2(g["u"], st.u), "v": rel_l2(g["v"], st.v), "visc": rel_l2(g["viscosity"], st.viscosity)}
    _log("  after the frames, rel L2: " + " ".join(f"{k}={v:.2e}" for k, v in e.items()))
    assert max(e.values()) <= 1e-5, e


def test_viscous_source_reseeds_with_its_viscosity(fg, orc, ctx):
    n = 64
    src = [fg.box(0.55, 0.75, 0.55, 0.8)]
    emit = [(0.5, -0.25, True, 7.5)]
    p, st, sc = _pair(fg, orc, ctx, n, 4, 2.0, emitters=emit, sources=src, seed=5)
    L = orc.lib()
    try:
        for k in range(3):
            vel = L.orc_max_particle_velocity(st.ref())
            dt = np.float32(min(5.0 / (vel + 1e-15), 1.0 / 60))
            sc.set_state(oracle_state_dict(st))
            gs = sc.substep(float(dt))
            g = sc.get_state()
            L.orc_substep(C.byref(p), st.ref(), C.c_float(dt), C.byref(orc.OrcStats()))
            N = st.n
            _log(f"  substep {k}: reseeded {gs.particles_reseeded}")
            assert len(g["px"]) == N
            assert np.array_equal(bits(g["particle_viscosity"]), bits(st.pvisc[:N]))
            assert np.array_equal(bits(g["viscosity"]), bits(st.viscosity))
            assert max(rel_l2(g["u"], st.u), rel_l2(g["v"], st.v)) <= 1e-5
        assert (g["particle_viscosity"] == np.float32(7.5)).any()
        assert (g["viscosity"][g["material"] == fg.SOURCE] == np.float32(7.5)).all()
    finally:
        sc.close()


def test_viscosity_refusals(fg, ctx):
    for kw in ({"density_mode": fg.DENSITY_GATED}, {"sim": fg.SIM_NBFLIP, "density_mode": fg.DENSITY_ON}):
        with pytest.raises(fg.FgError) as e:
            fg.Scene(ctx, fg.make_params(32, viscosity=True, **kw))
        assert e.value.code == fg.E_INVALID


def test_mixed_batch_with_viscous_scenes_equals_solo(fg, ctx):
    """A batch of FLIP, viscous FLIP and viscous NB-FLIP scenes: the viscous
    ones solve as a sub-group inside their class's group; every scene ends
    bit-identical to its solo run."""
    def make():
        out = []
        for k, (n, sim, visc) in enumerate([(48, fg.SIM_FLIP, False), (40, fg.SIM_FLIP, True),
                                            (56, fg.SIM_NBFLIP, True), (33, fg.SIM_FLIP, True)]):
            walls, fluids = fg.dam_break_geometry(n)
            sc = fg.Scene(ctx, fg.make_params(n, ppc=4, sim=sim, viscosity=visc))
            sc.init_objects(solids=walls, fluids=fluids, friction=np.zeros(4, np.float32), seed=k,
                            fluid_viscosity=[3.0 + k])
            out.append(sc)
        return out
    fields = ("px", "py", "vx", "vy", "pid", "u", "v", "viscosity", "particle_viscosity", "pressure")
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
