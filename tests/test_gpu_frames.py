"""Whole frames and the 1024^2 dam break against the CPU oracle.

FlipSolver::stepFrame (flipsolver2d.cpp:464-500) splits a frame into CFL
substeps from the particles' maximum speed.  These tests run the HIP path's
frame loop (fg_scene_step_frame) beside the oracle's (orc_step_frame) from
the same seeded dam break and compare every frame: the substep count, the
last dt and the maximum speed, and the state at the end.  The solves on both
sides converge (the GPU to FP64 relative residual 1e-10, the oracle to
1e-12), so the two runs differ only by solver round-off, which the frames
carry forward; the bar is relative L2 <= 1e-5 for pressure and velocities
and equal particle sets.

The 1024^2 dam break (8 particles per cell, 2,342,752 particles) runs its
frames on the GPU at the production tolerance 1e-6; every solve's FP64
residual is recomputed on the host from the matrix the solve used, and one
substep from the GPU's state is compared with the oracle's converged
substep at the substep bar (tests/helpers.assert_substep_parity).
"""
import ctypes as C
import sys
import time

import numpy as np
import pytest

from helpers import (assert_substep_parity, frame_dt, host_rel_residual, oracle_state_from_gpu, rel_l2)

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,ppc,frames", [(32, 4, 5), (128, 4, 4)])
def test_step_frame_vs_oracle(fg, orc, ctx, n, ppc, frames):
    _log(f"step_frame vs orc_step_frame n={n} ppc={ppc} frames={frames}")
    p, st = orc.dam_break(n, ppc=ppc, tol=1e-12)
    sc = fg.dam_break_scene(ctx, n, ppc=ppc, tol=1e-10)
    L = orc.lib()
    try:
        for f in range(frames):
            gs = sc.step_frame()
            os_ = orc.OrcStats()
            subs = L.orc_step_frame(C.byref(p), st.ref(), C.byref(os_))
            _log(f"  frame {f}: substeps gpu={gs.substeps} oracle={subs} dt gpu={gs.last_dt:.6e} "
                 f"oracle={os_.last_dt:.6e} vmax gpu={gs.max_velocity:.6e} oracle={os_.max_velocity:.6e} "
                 f"resid={gs.pressure_rel_resid:.2e}")
            assert gs.converged and gs.pressure_rel_resid <= 1e-10
            assert gs.substeps == subs, f
            assert abs(gs.last_dt - os_.last_dt) <= 1e-5 * os_.last_dt, f
            assert abs(gs.max_velocity - os_.max_velocity) <= 1e-5 * max(os_.max_velocity, 1e-30), f
            assert gs.particles == st.n, f
        g = sc.get_state()
    finally:
        sc.close()
    N = st.n
    assert len(g["px"]) == N
    assert np.array_equal(g["pid"], st.pid[:N])
    e = {"px": rel_l2(g["px"], st.px[:N]), "py": rel_l2(g["py"], st.py[:N]),
         "vx": rel_l2(g["vx"], st.vx[:N]), "vy": rel_l2(g["vy"], st.vy[:N]),
         "u": rel_l2(g["u"], st.u), "v": rel_l2(g["v"], st.v), "p": rel_l2(g["pressure"], st.pressure)}
    _log("  after the frames, rel L2: " + " ".join(f"{k}={v:.2e}" for k, v in e.items()))
    assert max(e.values()) <= 1e-5, e
    # materials are a function of positions: cells whose class differs can only
    # sit where a particle lies within solver round-off of a cell's SDF zero
    diff = int(np.count_nonzero(g["material"] != st.material))
    _log(f"  material cells that differ: {diff} of {n * n}")
    assert diff <= max(2, n * n // 2000)


def test_dam_break_1024_frames_and_substep_vs_oracle(fg, orc, ctx):
    n, ppc = 1024, 8
    rho, dx = 1000.0, 1.0 / n
    _log(f"dam break {n}^2, {ppc} ppc: GPU frames at tol 1e-6")
    sc = fg.dam_break_scene(ctx, n, ppc=ppc, tol=1e-6)
    tight = None
    try:
        assert sc.particle_count() == 2342752
        t0 = time.time()
        for f in range(3):
            s = sc.step_frame()
            _log(f"  frame {f}: {s.substeps} substeps, {s.particles} particles, iters={s.pressure_iters}, "
                 f"resid={s.pressure_rel_resid:.2e}")
            assert s.converged and s.pressure_rel_resid <= 1e-6
        g = sc.get_state()
        r_host = host_rel_residual(orc, fg, g, n, n, s.last_dt, rho, dx)
        _log(f"  last solve: FP64 residual recomputed on the host {r_host:.3e} (GPU {s.pressure_rel_resid:.3e}); "
             f"{time.time() - t0:.1f} s")
        assert r_host <= 1.001e-6
        assert abs(r_host - s.pressure_rel_resid) <= 1e-3 * max(r_host, 1e-30) + 1e-14
        # the next substep from this state: GPU at 1e-10 and at 1e-6, oracle converged
        p = orc.make_params(n, ppc=ppc, tol=1e-12)
        st = oracle_state_from_gpu(orc, g, n, n, capacity=n * n * ppc)
        vel = orc.lib().orc_max_particle_velocity(st.ref())
        dt = frame_dt(orc, vel, p)
        _log(f"  one substep dt={dt:.6e} (max speed {vel:.4e} cells/s)")
        tight = fg.Scene(ctx, fg.make_params(n, ppc=ppc, tol=1e-10))
        tight.set_friction(np.zeros(4, np.float32))
        tight.set_state(g)
        ts = tight.substep(float(dt))
        gt = tight.get_state()
        ps = sc.substep(float(dt))
        gp = sc.get_state()
        r_prod = host_rel_residual(orc, fg, gp, n, n, dt, rho, dx)
        _log(f"  GPU substep: tol 1e-10 resid {ts.pressure_rel_resid:.2e} ({ts.pressure_iters} iters); "
             f"tol 1e-6 resid {ps.pressure_rel_resid:.2e} host {r_prod:.2e} ({ps.pressure_iters} iters)")
        assert ts.pressure_rel_resid <= 1e-10 and r_prod <= 1.001e-6
    finally:
        sc.close()
        if tight is not None:
            tight.close()
    t0 = time.time()
    orc.lib().orc_substep(C.byref(p), st.ref(), C.c_float(dt), C.byref(orc.OrcStats()))
    _log(f"  oracle substep {time.time() - t0:.1f} s")
    assert_substep_parity(gt, st, _log, what="GPU tol 1e-10 vs oracle")
    # the production solve: the same upstream state bit for bit, the solution
    # within what a 1e-6 residual allows on this system
    assert_substep_parity(gp, st, _log, p_tol=1e-4, v_tol=1e-5, what="GPU tol 1e-6 vs oracle")
