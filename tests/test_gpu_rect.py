"""Rectangular grids (size_i != size_j) on the GPU against the oracle, for
every solver class: FLIP, FLIP with viscosity, NB-FLIP and NB-FLIP with
viscosity (flipsolver2d.cpp:412-462, nbflipsolver.cpp:26-64,
viscositymodel.cpp:4-140).  Both orientations run, so that a swapped I and J
in a face grid's shape, a row stride or a lattice index fails one of them.

Each case is a walled box with an initial fluid rectangle, a source emitting
into the domain and a sink, run by the oracle for a few frames and then handed
to the GPU for one substep from the same state.  The bar is the solver's:
bit-exact for integer/index work and every float stage upstream of the
solves, relative L2 <= 1e-5 for pressure and velocities (and for the
right-hand side with viscosity, which the second projection rebuilds from the
viscous solve's velocities) against the oracle's FP64 solves converged to
1e-12.
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


def _rect_objects(fg, I, J):
    """Walls one cell thick around an (I dx) x (J dx) box, dx = 1 / max(I, J)."""
    dx = 1.0 / max(I, J)
    X, Y = I * dx, J * dx
    walls = [fg.box(0.0, dx, 0.0, Y), fg.box(X - dx, X, 0.0, Y), fg.box(0.0, X, 0.0, dx), fg.box(0.0, X, Y - dx, Y)]
    fluids = [fg.box(0.3 * X, X - dx, dx, 0.45 * Y)]
    sources = [fg.box(0.55 * X, 0.75 * X, 0.55 * Y, 0.8 * Y)]
    sinks = [fg.box(0.06 * X, 0.14 * X, 0.7 * Y, 0.9 * Y)]
    return walls, sources, sinks, fluids


CLASSES = ["flip", "flip_visc", "nbflip", "nbflip_visc"]


def _pair(fg, orc, ctx, I, J, cls, ppc=4, seed=5):
    nb = cls.startswith("nb")
    visc = cls.endswith("visc")
    walls, sources, sinks, fluids = _rect_objects(fg, I, J)
    emit = [(0.5, -0.25, True, 7.5)] if visc else [(0.5, -0.25, True)]
    fv = [4.0] if visc else []
    p = orc.make_params(I, J, ppc=ppc, tol=1e-12, sim=orc.SIM_NBFLIP if nb else orc.SIM_FLIP, emitters=emit,
                        reseed_seed=seed, viscosity=visc, fluid_viscosity=fv)
    if nb:
        st = orc.init_nb_scene(p, solids=walls, sources=sources, sinks=sinks, fluids=fluids)
    else:
        st = orc.init_scene_objects(p, solids=walls, sources=sources, sinks=sinks, fluids=fluids, seed=seed)
    sc = fg.Scene(ctx, fg.make_params(I, J, ppc=ppc, tol=1e-10, sim=fg.SIM_NBFLIP if nb else fg.SIM_FLIP,
                                      viscosity=visc))
    sc.init_objects(solids=walls, sources=sources, sinks=sinks, fluids=fluids,
                    friction=np.zeros(len(walls), np.float32), emitters=emit, seed=seed, fluid_viscosity=fv)
    return p, st, sc


@pytest.mark.parametrize("I,J", [(72, 50), (50, 72)])
@pytest.mark.parametrize("cls", CLASSES)
def test_rectangular_substep_vs_oracle(fg, orc, ctx, cls, I, J):
    _log(f"rectangular {cls} {I}x{J}: first frame, 3 oracle frames, one substep")
    nb, visc = cls.startswith("nb"), cls.endswith("visc")
    p, st, sc = _pair(fg, orc, ctx, I, J, cls)
    L = orc.lib()
    try:
        g0 = sc.get_state()
        assert np.array_equal(g0["material"], st.material)
        assert np.array_equal(bits(g0["px"]), bits(st.px[:st.n]))
        if nb:
            assert np.array_equal(bits(g0["fluid_sdf"]), bits(st.fluid_sdf))
        if visc:
            assert np.array_equal(bits(g0["viscosity"]), bits(st.viscosity))
        run_oracle_frames(orc, p, st, 3)
        assert st.n > 0 and (st.material == orc.SOURCE).any()
        vel = L.orc_max_particle_velocity(st.ref())
        dt = np.float32(min(5.0 / (vel + 1e-15), 1.0 / 60))
        sc.set_state(oracle_state_dict(st))
        gs = sc.substep(float(dt))
        g = sc.get_state()
    finally:
        sc.close()
    step = L.orc_nb_substep if nb else L.orc_substep
    step(C.byref(p), st.ref(), C.c_float(dt), C.byref(orc.OrcStats()))
    N = st.n
    _log(f"  particles gpu={len(g['px'])} oracle={N} pressure iters={gs.pressure_iters} "
         f"viscosity iters={gs.viscosity_iters} resid={gs.pressure_rel_resid:.2e}")
    assert gs.converged
    assert len(g["px"]) == N
    assert np.array_equal(g["pid"], st.pid[:N])
    assert np.array_equal(bits(g["px"]), bits(st.px[:N]))
    assert np.array_equal(bits(g["py"]), bits(st.py[:N]))
    assert np.array_equal(g["material"], st.material)
    assert np.array_equal(g["particle_counts"], st.particle_counts)
    assert np.array_equal(bits(g["fluid_sdf"]), bits(st.fluid_sdf))
    assert np.array_equal(bits(g["saved_u"]), bits(st.saved_u))
    assert np.array_equal(bits(g["saved_v"]), bits(st.saved_v))
    assert np.array_equal(g["u_valid"], st.u_valid)
    assert np.array_equal(g["v_valid"], st.v_valid)
    e = {"p": rel_l2(g["pressure"], st.pressure), "u": rel_l2(g["u"], st.u), "v": rel_l2(g["v"], st.v),
         "vx": rel_l2(g["vx"], st.vx[:N]), "vy": rel_l2(g["vy"], st.vy[:N])}
    if visc:
        assert gs.viscosity_iters > 0
        assert np.array_equal(bits(g["particle_viscosity"]), bits(st.pvisc[:N]))
        assert np.array_equal(bits(g["viscosity"]), bits(st.viscosity))
        e["rhs"] = rel_l2(g["rhs"], st.rhs)
    else:
        assert np.array_equal(bits(g["rhs"]), bits(st.rhs))
    _log("  rel L2: " + " ".join(f"{k}={v:.2e}" for k, v in e.items()))
    assert max(e.values()) <= 1e-5, e


def test_rectangular_mixed_batch_equals_solo(fg, ctx):
    """A batch of all four solver classes on rectangular grids of both
    orientations, with sources and sinks: each class solves as its own group
    (the viscous scenes as a sub-group), and every scene ends three frames
    bit-identical to its solo run."""
    cases = [(72, 50, "flip"), (50, 72, "flip_visc"), (64, 40, "nbflip"), (40, 64, "nbflip_visc"),
             (33, 47, "flip"), (47, 33, "nbflip_visc")]

    def make():
        out = []
        for k, (I, J, cls) in enumerate(cases):
            nb, visc = cls.startswith("nb"), cls.endswith("visc")
            walls, sources, sinks, fluids = _rect_objects(fg, I, J)
            emit = [(0.5, -0.25, True, 2.0 + k)] if visc else [(0.5, -0.25, True)]
            sc = fg.Scene(ctx, fg.make_params(I, J, ppc=4, sim=fg.SIM_NBFLIP if nb else fg.SIM_FLIP,
                                              viscosity=visc))
            sc.init_objects(solids=walls, sources=sources, sinks=sinks, fluids=fluids,
                            friction=np.zeros(len(walls), np.float32), emitters=emit, seed=k,
                            fluid_viscosity=[3.0 + k] if visc else [])
            out.append(sc)
        return out

    _log(f"rectangular mixed batch of {len(cases)} scenes vs solo, 3 frames")
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
    for k, (I, J, cls) in enumerate(cases):
        assert len(solo[k]["px"]) > 0, (k, cls)
        names = fields + (("viscosity", "particle_viscosity") if cls.endswith("visc") else ())
        for f in names:
            assert np.array_equal(bits(got[k][f]), bits(solo[k][f])), (k, cls, I, J, f)
