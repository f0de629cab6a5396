"""Density correction on the GPU (FlipSolver::densityCorrection,
flipsolver2d.cpp:164-303, 993-1011) against the oracle.

Bar: the density right-hand side bit-exact (same particles in the same
canonical order, same float sums); the correction solve relative L2 <= 1e-5
against the FP64 oracle solve; the particle displacements it causes relative
L2 <= 1e-5.  A batch mixing density modes equals each scene stepped alone.
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


@pytest.mark.parametrize("n,ppc,frames", [(48, 4, 4), (96, 8, 3)])
def test_density_stage_parity(fg, orc, ctx, n, ppc, frames):
    _log(f"density stage parity n={n} ppc={ppc} frames={frames}")
    p, st = orc.dam_break(n, ppc=ppc, tol=1e-12)
    run_oracle_frames(orc, p, st, frames)
    L = orc.lib()
    vel = L.orc_max_particle_velocity(st.ref())
    dt = np.float32(min(5.0 / (vel + 1e-15), 1.0 / 30))
    sc = fg.Scene(ctx, fg.make_params(n, ppc=ppc, tol=1e-10, density_mode=fg.DENSITY_ON))
    sc.set_friction(np.zeros(4, np.float32))
    sc.set_state(oracle_state_dict(st))
    sc.stage(fg.STAGE_ADVECT, float(dt))
    sc.stage(fg.STAGE_REBIN)
    sc.stage(fg.STAGE_DENSITY)
    g = sc.get_state()
    sc.close()
    # oracle: advect, prune + sort, density grid, rhs, converged solve, adjust, re-sort
    alive = np.zeros(st.n, np.uint8)
    L.orc_advect(C.byref(p), st.ref(), C.c_float(dt), alive)
    cs = np.zeros(n * n + 1, np.int32)
    L.orc_sort_particles(C.byref(p), st.ref(), alive.ctypes.data, cs.ctypes.data)
    mat_a = st.material.copy()
    L.orc_density_grid(C.byref(p), st.ref(), cs)
    rhs = np.zeros(n * n)
    L.orc_density_rhs(C.byref(p), st.ref(), C.c_float(dt), rhs)
    assert np.array_equal(g["rhs"].view(np.uint64), rhs.view(np.uint64))
    assert np.abs(rhs).max() > 0
    sA = float(dt) / (p.fluid_density * p.dx * p.dx)
    x = np.zeros(n * n)
    rr = C.c_double()
    L.orc_pcg_converged(mat_a, n, n, sA, rhs, x, 100000, 1e-13, C.byref(rr))
    e_x = rel_l2(g["pressure"], x)
    before = {int(i): (a, b) for i, a, b in zip(st.pid[: st.n], st.px[: st.n], st.py[: st.n])}
    L.orc_adjust_particles(C.byref(p), st.ref(), C.c_float(dt), x)
    L.orc_sort_particles(C.byref(p), st.ref(), None, cs.ctypes.data)
    N = st.n
    _log(f"  particles gpu={len(g['px'])} oracle={N}, correction solve rel L2 {e_x:.2e}")
    assert e_x <= 1e-5
    assert len(g["px"]) == N
    og = np.argsort(g["pid"], kind="stable")
    oo = np.argsort(st.pid[:N], kind="stable")
    assert np.array_equal(g["pid"][og], st.pid[:N][oo])
    b0 = np.array([before[int(i)] for i in st.pid[:N][oo]], np.float64)
    d_orc = np.stack([st.px[:N][oo], st.py[:N][oo]], 1).astype(np.float64) - b0
    d_gpu = np.stack([g["px"][og], g["py"][og]], 1).astype(np.float64) - b0
    e_d = rel_l2(d_gpu, d_orc)
    _log(f"  displacement rel L2 {e_d:.2e}, max |d| {np.abs(d_orc).max():.3e}")
    assert np.abs(d_orc).max() > 0
    assert e_d <= 1e-5


def test_gated_equals_on_when_converged(fg, orc, ctx):
    n, ppc = 64, 4
    p, st = orc.dam_break(n, ppc=ppc, tol=1e-12)
    run_oracle_frames(orc, p, st, 3)
    states = []
    for mode in (fg.DENSITY_ON, fg.DENSITY_GATED):
        sc = fg.Scene(ctx, fg.make_params(n, ppc=ppc, tol=1e-8, density_mode=mode))
        sc.set_friction(np.zeros(4, np.float32))
        sc.set_state(oracle_state_dict(st))
        s = sc.substep(1.0 / 120)
        assert s.density_applied == 1 and s.density_iters > 0
        states.append(sc.get_state())
        sc.close()
    for k in ("px", "py", "vx", "vy", "pressure"):
        assert np.array_equal(states[0][k].view(np.uint8), states[1][k].view(np.uint8)), k


def test_batch_with_mixed_density_modes(fg, ctx):
    specs = fg.mixed_scene_specs(4, master_seed=31, resolutions=(32, 48, 64))
    modes = [fg.DENSITY_OFF, fg.DENSITY_ON, fg.DENSITY_GATED, fg.DENSITY_OFF]
    solo = [fg.make_scene(ctx, sp, tol=1e-6, density_mode=m) for sp, m in zip(specs, modes)]
    mix = [fg.make_scene(ctx, sp, tol=1e-6, density_mode=m) for sp, m in zip(specs, modes)]
    ran = 0
    for _ in range(2):
        so = [sc.step_frame() for sc in solo]
        st = fg.batch_step_frame(ctx, mix)
        for k, s in enumerate(st):
            assert (s.density_iters, s.density_applied) == (so[k].density_iters, so[k].density_applied), k
            if modes[k] == fg.DENSITY_OFF:
                assert s.density_iters == 0 and s.density_applied == 0
            ran += s.density_iters
    assert ran > 0
    for a, b in zip(solo, mix):
        ga, gb = a.get_state(), b.get_state()
        for k in ("px", "py", "vx", "vy", "pid", "material", "pressure"):
            assert np.array_equal(ga[k].view(np.uint8), gb[k].view(np.uint8)), k
    for sc in solo + mix:
        sc.close()
