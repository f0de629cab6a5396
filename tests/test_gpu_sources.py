"""Sources, sinks and reseeding, and first-frame init on the device,
against the CPU oracle.

A box with one-cell walls, an initial pool, two sources (one transferring
its velocity to the grid and to reseeded particles, one not) and a sink.
fg_scene_init_objects runs FlipSolver::firstFrameInit's cell passes and the
seeding on the device: every cell array and every particle must equal the
oracle's bit for bit.  One substep from a state the oracle reached after a
few frames (sources refilling, particles draining into the sink) must meet
the substep bar, reseeded particles included: their count, ids and
positions bit for bit, their velocities (the final grid's) within the
velocity tolerance.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from helpers import assert_substep_parity, frame_dt, oracle_state_dict

pytestmark = pytest.mark.gpu

EMITTERS = [(1.0, 0.5, True), (0.0, -0.5, False)]


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _objects(n, orc):
    d = 1.0 / n
    box = orc.box_polys
    walls = [box(0.0, d, 0.0, 1.0), box(1.0 - d, 1.0, 0.0, 1.0), box(0.0, 1.0, 0.0, d), box(0.0, 1.0, 1.0 - d, 1.0)]
    sources = [box(0.1, 0.2, 0.4, 0.6), box(0.15, 0.3, 0.75, 0.85)]
    sinks = [box(0.88, 0.97, 0.35, 0.5)]
    fluids = [box(0.5, 1.0 - d, d, 0.3)]
    return walls, sources, sinks, fluids


def _oracle_init(orc, n, ppc, seed=7):
    p = orc.make_params(n, ppc=ppc, tol=1e-12, emitters=EMITTERS, reseed_seed=seed)
    walls, sources, sinks, fluids = _objects(n, orc)
    st = orc.init_scene_objects(p, walls, sources, sinks, fluids, seed=seed)
    return p, st


@pytest.mark.parametrize("n,ppc", [(64, 4), (96, 6)])
def test_init_objects_matches_oracle(fg, orc, ctx, n, ppc):
    _log(f"device init with sources and sinks n={n} ppc={ppc}")
    p, st = _oracle_init(orc, n, ppc)
    sc = fg.Scene(ctx, fg.make_params(n, ppc=ppc))
    try:
        walls, sources, sinks, fluids = _objects(n, orc)
        sc.init_objects(walls, sources, sinks, fluids, friction=np.zeros(4, np.float32), emitters=EMITTERS, seed=7)
        g = sc.get_state()
    finally:
        sc.close()
    for k in ("material", "solid_id", "emitter_id"):
        assert np.array_equal(g[k], getattr(st, k)), k
    assert np.array_equal(g["solid_sdf"].view(np.uint32), st.solid_sdf.view(np.uint32))
    N = st.n
    assert len(g["px"]) == N and g["next_pid"] == N and g["rng_epoch"] == 0
    for k in ("px", "py"):
        assert np.array_equal(g[k].view(np.uint32), getattr(st, k)[:N].view(np.uint32)), k
    assert np.array_equal(g["pid"], st.pid[:N])
    assert (st.material == orc.SOURCE).sum() > 10 and (st.material == orc.SINK).sum() > 10


def test_sources_sinks_substep_vs_oracle(fg, orc, ctx):
    n, ppc = 64, 4
    p, st = _oracle_init(orc, n, ppc)
    L = orc.lib()
    for f in range(5):
        os_ = orc.OrcStats()
        L.orc_step_frame(C.byref(p), st.ref(), C.byref(os_))
    _log(f"oracle after 5 frames: {st.n} particles, killed {os_.particles_killed}, epoch {st.c.rng_epoch}, "
         f"next id {st.c.next_pid}")
    sc = fg.Scene(ctx, fg.make_params(n, ppc=ppc, tol=1e-10))
    try:
        sc.set_friction(np.zeros(4, np.float32))
        sc.set_state(oracle_state_dict(st))
        sc.set_emitters(EMITTERS, seed=7)
        vel = L.orc_max_particle_velocity(st.ref())
        dt = frame_dt(orc, vel, p)
        before = st.n
        gs = sc.substep(float(dt))
        g = sc.get_state()
    finally:
        sc.close()
    os_ = orc.OrcStats()
    L.orc_substep(C.byref(p), st.ref(), C.c_float(dt), C.byref(os_))
    _log(f"  substep dt={dt:.4e}: reseeded gpu {gs.particles_reseeded}, killed gpu {gs.particles_killed} "
         f"oracle {os_.particles_killed}; particles {len(g['px'])} / {st.n} (from {before})")
    assert gs.particles_reseeded > 0 and gs.particles_killed == os_.particles_killed
    assert g["next_pid"] == st.c.next_pid and g["rng_epoch"] == st.c.rng_epoch
    assert_substep_parity(g, st, _log, what="sources + sink substep")
