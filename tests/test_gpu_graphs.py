"""PCG iterations replayed from HIP graphs against direct launches.

Graph replay is switched off (fg_ctx_set_solver_graphs refuses it): the
comparisons below are kept, skipped, for the day it is switched on again.

With fg_ctx_set_solver_graphs a lane captures each PCG iteration's launches
(direction update, residual update, V-cycle) once per iteration index and
launch shape and replays the graph in later solves.  The kernels and their
arguments are those of the direct path, so every field must come out
bit-identical: over several frames of a single scene (one graph per
iteration index, replayed every substep with the host's polls and the
other stages' launches in between), over a batch whose packed launch shapes
change as scenes converge, and over a standalone Poisson solve, each
replayed more than once.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SWITCHED_OFF = pytest.mark.skip(reason="graph replay of PCG iterations is switched off: a 1024^2 run of it under "
                                       "rocprofv3 ended in a host-side fault whose cause was not found")

FIELDS = ("px", "py", "vx", "vy", "pid", "u", "v", "material", "fluid_sdf", "pressure", "rhs")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _same(a, b, where):
    for k in FIELDS:
        assert a[k].shape == b[k].shape, (where, k)
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), (where, k)


@pytest.fixture
def graphs_off(ctx):
    yield ctx
    ctx.set_solver_graphs(False)
    ctx.set_batch_lanes(1)


@SWITCHED_OFF
def test_single_scene_graph_equals_direct(fg, graphs_off):
    ctx = graphs_off
    sp = fg.mixed_scene_specs(1, master_seed=5, resolutions=(160,))[0]
    _log(f"single scene {sp['n']}^2, 3 frames, graphs off then on")
    direct = fg.make_scene(ctx, sp, tol=1e-6)
    graphed = fg.make_scene(ctx, sp, tol=1e-6)
    try:
        for f in range(3):
            ctx.set_solver_graphs(False)
            a = direct.step_frame()
            ctx.set_solver_graphs(True)
            b = graphed.step_frame()
            assert (a.substeps, a.pressure_iters) == (b.substeps, b.pressure_iters), f
            _same(direct.get_state(), graphed.get_state(), f)
    finally:
        direct.close()
        graphed.close()


@SWITCHED_OFF
@pytest.mark.parametrize("lanes", [1, 2])
def test_batch_graph_equals_direct(fg, graphs_off, lanes):
    ctx = graphs_off
    specs = fg.mixed_scene_specs(9, master_seed=23, resolutions=(32, 48, 64, 96))
    _log(f"batch of {len(specs)} {[s['n'] for s in specs]}, 2 frames, {lanes} lanes, graphs off then on")
    ctx.set_batch_lanes(lanes)
    direct = [fg.make_scene(ctx, sp, tol=1e-6) for sp in specs]
    graphed = [fg.make_scene(ctx, sp, tol=1e-6) for sp in specs]
    try:
        for f in range(2):
            ctx.set_solver_graphs(False)
            a = fg.batch_step_frame(ctx, direct)
            ctx.set_solver_graphs(True)
            b = fg.batch_step_frame(ctx, graphed)
            for k in range(len(specs)):
                assert a[k].pressure_iters == b[k].pressure_iters, (f, k)
                _same(direct[k].get_state(), graphed[k].get_state(), (f, k))
    finally:
        for s in direct + graphed:
            s.close()


@SWITCHED_OFF
def test_poisson_solve_graph_equals_direct(fg, graphs_off):
    ctx = graphs_off
    mat, _, _, b, scale = fg.isolated_projection_system(256, 3)
    ctx.set_solver_graphs(False)
    x0, s0 = ctx.poisson_solve(mat, scale, b, tol=1e-6)
    ctx.set_solver_graphs(True)
    for rep in range(3):   # capture, then replays
        x1, s1 = ctx.poisson_solve(mat, scale, b, tol=1e-6)
        _log(f"  solve {rep}: {s1.iterations} iterations (direct {s0.iterations})")
        assert s1.iterations == s0.iterations
        assert np.array_equal(x0.view(np.uint8), x1.view(np.uint8)), rep


def test_solver_graphs_refused(fg, ctx):
    ctx.set_solver_graphs(False)
    with pytest.raises(fg.FgError) as e:
        ctx.set_solver_graphs(True)
    assert e.value.code == fg.E_INVALID and "switched off" in str(e.value)
