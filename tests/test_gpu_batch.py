"""Batched stepping: many scenes, one launch per phase.

Every scene's kernels read only that scene's descriptor and reduce over that
scene's own blocks in a fixed order, so a scene stepped inside a batch must
give exactly the state it gives when stepped alone.  Together with the
single-scene parity tests against the oracle, this carries oracle parity
over to the batched and served forms.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("px", "py", "vx", "vy", "pid", "u", "v", "u_valid", "v_valid", "material", "fluid_sdf",
          "particle_counts", "pressure", "rhs")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _same(a, b, where):
    for k in FIELDS:
        x, y = a[k], b[k]
        assert x.shape == y.shape, (where, k)
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (where, k)


def _scenes(fg, ctx, specs):
    return [fg.make_scene(ctx, sp, tol=1e-6) for sp in specs]


def test_mixed_specs_are_deterministic(fg):
    a = fg.mixed_scene_specs(16, master_seed=7)
    b = fg.mixed_scene_specs(16, master_seed=7)
    assert a == b
    for sp in a:
        x0, x1, y0, y1 = sp["fluid_box"]
        assert 0.1 <= x0 <= x1 <= 0.6 and 0.1 <= y0 <= y1 <= 0.6
        assert 4 <= sp["ppc"] <= 8 and sp["n"] in fg.MIXED_RESOLUTIONS
        assert abs(np.hypot(*sp["gravity"]) - 9.81) < 1e-9


@pytest.mark.parametrize("count,frames,lanes", [(6, 2, 1), (13, 1, 1), (13, 2, 2), (9, 2, 3), (7, 1, 4)])
def test_batch_equals_solo(fg, ctx, count, frames, lanes):
    specs = fg.mixed_scene_specs(count, master_seed=11 + count, resolutions=(32, 48, 64, 80))
    _log(f"batch vs solo: {count} scenes {[s['n'] for s in specs]}, {frames} frames, {lanes} lanes")
    solo = _scenes(fg, ctx, specs)
    batch = _scenes(fg, ctx, specs)
    ctx.set_batch_lanes(lanes)
    try:
        for f in range(frames):
            s_stats = [sc.step_frame() for sc in solo]
            b_stats = fg.batch_step_frame(ctx, batch)
            for k in range(count):
                assert s_stats[k].substeps == b_stats[k].substeps, (f, k)
                assert s_stats[k].pressure_iters == b_stats[k].pressure_iters, (f, k)
                assert s_stats[k].particle_substeps == b_stats[k].particle_substeps, (f, k)
                assert b_stats[k].converged == 1
    finally:
        ctx.set_batch_lanes(1)
    for k in range(count):
        _same(solo[k].get_state(), batch[k].get_state(), k)
    for sc in solo + batch:
        sc.close()


def test_membership_changes_between_frames(fg, ctx):
    """Scenes join and leave the batch between frames (the serving pattern)."""
    specs = fg.mixed_scene_specs(5, master_seed=99, resolutions=(32, 48, 64))
    _log(f"membership: {[s['n'] for s in specs]}")
    solo = _scenes(fg, ctx, specs)
    mix = _scenes(fg, ctx, specs)
    plan = [[0, 1], [0, 1, 2, 3], [2, 3, 4], [4]]   # frame f steps these scenes
    for members in plan:
        for k in members:
            solo[k].step_frame()
        fg.batch_step_frame(ctx, [mix[k] for k in members])
    for k in range(5):
        _same(solo[k].get_state(), mix[k].get_state(), k)
    for sc in solo + mix:
        sc.close()


def test_batch_rejects_duplicates(fg, ctx):
    sc = fg.dam_break_scene(ctx, 32)
    with pytest.raises(fg.FgError):
        fg.batch_step_frame(ctx, [sc, sc])
    sc.close()


def test_autotune_keeps_results(fg, ctx):
    """The lane autotuner steps real frames; the scenes stay equal to solo."""
    specs = fg.mixed_scene_specs(8, master_seed=77, resolutions=(32, 48, 64, 96))
    solo = _scenes(fg, ctx, specs)
    batch = _scenes(fg, ctx, specs)
    try:
        lanes, rates = fg.batch_autotune(ctx, batch, rounds=1)
        assert 1 <= lanes <= fg.MAX_LANES and ctx.batch_lanes() == lanes
        assert np.all(rates > 0)
        # the most lanes among the rates within 3 % of the fastest
        assert rates[lanes - 1] >= 0.97 * rates.max()
        assert all(rates[l - 1] < 0.97 * rates.max() for l in range(lanes + 1, fg.MAX_LANES + 1))
        for _ in range(fg.MAX_LANES):
            for sc in solo:
                sc.step_frame()
        for k in range(len(specs)):
            _same(solo[k].get_state(), batch[k].get_state(), k)
    finally:
        ctx.set_batch_lanes(1)
    for sc in solo + batch:
        sc.close()


def test_tiling_autotune_sets_the_fastest_and_keeps_batch_equal_to_solo(fg, ctx):
    """The tiling autotuner measures each candidate on live frames of a copy
    and sets the fastest; with that tiling a batch still equals its scenes
    stepped alone."""
    specs = fg.mixed_scene_specs(6, master_seed=91, resolutions=(48, 64, 96))
    copy = _scenes(fg, ctx, specs)
    cand = ((1024, 256), (1024, 1024), (64, 2048))
    try:
        k, rates = fg.batch_autotune_tiling(ctx, copy, rounds=2, candidates=cand)
        _log(f"tiling autotune: {cand[k]}, rates {[round(r / 1e6, 2) for r in rates]} M")
        assert 0 <= k < len(cand) and np.all(rates > 0) and rates[k] == rates.max()
        solo = _scenes(fg, ctx, specs)
        batch = _scenes(fg, ctx, specs)
        for f in range(2):
            s_stats = [sc.step_frame() for sc in solo]
            b_stats = fg.batch_step_frame(ctx, batch)
            for j in range(len(specs)):
                assert s_stats[j].pressure_iters == b_stats[j].pressure_iters, (f, j)
                _same(solo[j].get_state(), batch[j].get_state(), (f, j))
        for sc in solo + batch:
            sc.close()
    finally:
        ctx.set_solver_tiling(1024, 1024)
        for sc in copy:
            sc.close()
    with pytest.raises(fg.FgError):
        fg.batch_autotune_tiling(ctx, [], rounds=1)
