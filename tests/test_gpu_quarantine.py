"""Per-scene failure isolation, cancellation and admission failures.

A scene whose velocities turn NaN/Inf is quarantined: it leaves its launch
group and every other scene of the batch completes its frame exactly as it
would alone (bit for bit).  The server ends such a request as FAILED and
keeps ticking; a cancelled request leaves the queue or the batch; a request
that cannot be admitted (its fluid seeds more particles than its
max_particles) is rejected without stopping the queue.
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
        assert a[k].shape == b[k].shape, (where, k)
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), (where, k)


def _poison(sc, what):
    """NaN: max |v| ignores it, the substep's pressure solve sees it.  Inf:
    the frame's CFL split sees an infinite speed before any substep."""
    g = sc.get_state()
    if what == "nan":
        g["vx"][len(g["vx"]) // 2] = np.nan
    else:
        g["vy"][len(g["vy"]) // 3] = np.inf
    sc.set_state(g)


@pytest.mark.parametrize("what", ["nan", "inf"])
def test_nan_scene_is_quarantined_others_unchanged(fg, ctx, what):
    specs = fg.mixed_scene_specs(6, master_seed=31, resolutions=(32, 48, 64))
    _log(f"quarantine ({what} NaN/Inf): {[s['n'] for s in specs]}")
    solo = [fg.make_scene(ctx, sp, tol=1e-6) for sp in specs]
    batch = [fg.make_scene(ctx, sp, tol=1e-6) for sp in specs]
    bad = int(np.argmax([sc.particle_count() for sc in batch]))   # a scene with particles to poison
    try:
        for sc in solo:
            sc.step_frame()
        fg.batch_step_frame(ctx, batch)
        _poison(batch[bad], what)
        for f in range(2):
            for k, sc in enumerate(solo):
                if k != bad:
                    sc.step_frame()
            stats = fg.batch_step_frame(ctx, batch, allow_quarantine=True)
            _log(f"  frame {f}: status {[s.status for s in stats]}")
            for k, s in enumerate(stats):
                if k == bad:
                    assert s.status == fg.E_NUMERIC and not s.converged
                else:
                    assert s.status == 0 and s.converged
        with pytest.raises(fg.FgError) as ei:
            fg.batch_step_frame(ctx, batch)
        assert ei.value.code == fg.E_NUMERIC
        for k in range(len(specs)):
            if k != bad:
                solo[k].step_frame()
        for k in range(len(specs)):
            if k != bad:
                _same(solo[k].get_state(), batch[k].get_state(), k)
        # a quarantined scene refuses to step until it gets a new state
        with pytest.raises(fg.FgError) as ei:
            batch[bad].step_frame()
        assert ei.value.code == fg.E_NUMERIC
        g = solo[bad].get_state()
        batch[bad].set_state(g)
        s = batch[bad].step_frame()
        assert s.status == 0 and s.converged
    finally:
        for sc in solo + batch:
            sc.close()


def test_single_scene_substep_reports_numeric(fg, ctx):
    sc = fg.dam_break_scene(ctx, 48, tol=1e-6)
    try:
        sc.step_frame()
        _poison(sc, "nan")
        with pytest.raises(fg.FgError) as ei:
            sc.substep(1e-3)
        assert ei.value.code == fg.E_NUMERIC
        with pytest.raises(fg.FgError) as ei:
            sc.substep(1e-3)
        assert ei.value.code == fg.E_NUMERIC
    finally:
        sc.close()


def _solo(fg, ctx, spec, frames):
    sc = fg.make_scene(ctx, spec, tol=1e-6)
    for _ in range(frames):
        sc.step_frame()
    g = sc.get_state()
    sc.close()
    return g


def test_server_fails_quarantined_request_and_keeps_ticking(fg, ctx):
    specs = fg.mixed_scene_specs(4, master_seed=41, resolutions=(32, 48, 64))
    srv = fg.Server(ctx, hbm_budget=8 << 30, max_resident=8)
    try:
        ids = [srv.submit_spec(sp, 3, keep=True, tol=1e-6) for sp in specs]
        srv.tick()
        _poison(srv.scene(ids[1]), "nan")
        for _ in range(10):
            t = srv.tick()
            if t.queued == 0 and t.resident == 0:
                break
        st = [srv.status(r) for r in ids]
        _log(f"  states {[s.state for s in st]} errors {[s.error for s in st]}")
        assert st[1].state == fg.REQ_FAILED and st[1].error == fg.E_NUMERIC
        for k in (0, 2, 3):
            assert st[k].state == fg.REQ_DONE and st[k].frames_done == 3
            _same(srv.scene(ids[k]).get_state(), _solo(fg, ctx, specs[k], 3), k)
        summ = srv.summary()
        assert summ.failed == 1 and summ.completed == 3
        srv.release(ids[1])
    finally:
        srv.close()


def test_server_cancel_queued_and_resident(fg, ctx):
    specs = fg.mixed_scene_specs(4, master_seed=43, resolutions=(32, 48))
    srv = fg.Server(ctx, hbm_budget=8 << 30, max_resident=2)
    try:
        ids = [srv.submit_spec(sp, 4, keep=True, tol=1e-6) for sp in specs]
        srv.tick()   # 0, 1 resident; 2, 3 queued
        assert srv.status(ids[0]).state == fg.REQ_RESIDENT and srv.status(ids[3]).state == fg.REQ_QUEUED
        srv.tick()
        srv.cancel(ids[1])   # resident
        srv.cancel(ids[3])   # queued
        assert srv.status(ids[1]).state == fg.REQ_CANCELLED and srv.status(ids[3]).state == fg.REQ_CANCELLED
        with pytest.raises(fg.FgError):
            srv.scene(ids[1])
        with pytest.raises(fg.FgError):
            srv.cancel(ids[1])
        t = srv.tick()   # 2 is admitted into the freed slot
        assert t.resident == 2 and t.admitted == 1
        for _ in range(10):
            t = srv.tick()
            if t.queued == 0 and t.resident == 0:
                break
        for k in (0, 2):
            assert srv.status(ids[k]).state == fg.REQ_DONE
            _same(srv.scene(ids[k]).get_state(), _solo(fg, ctx, specs[k], 4), k)
        summ = srv.summary()
        assert summ.cancelled == 2 and summ.completed == 2
        srv.cancel(ids[0])   # a finished, kept request: released
        assert srv.status(ids[0]).state == fg.REQ_RELEASED
    finally:
        srv.close()


def test_server_rejects_unadmittable_request_and_moves_on(fg, ctx):
    specs = fg.mixed_scene_specs(2, master_seed=47, resolutions=(32, 48))
    srv = fg.Server(ctx, hbm_budget=8 << 30, max_resident=8)
    try:
        a = srv.submit_spec(specs[0], 2, keep=True, tol=1e-6)
        p = fg.make_params(48, ppc=4, max_particles=100)   # the dam break seeds far more
        walls, fluid = fg.dam_break_geometry(48)
        bad = srv.submit(p, walls, fluid, 2)
        b = srv.submit_spec(specs[1], 2, keep=True, tol=1e-6)
        for _ in range(10):
            t = srv.tick()
            if t.queued == 0 and t.resident == 0:
                break
        assert srv.status(bad).state == fg.REQ_REJECTED and srv.status(bad).error == fg.E_CAPACITY
        for k, rid in ((0, a), (1, b)):
            assert srv.status(rid).state == fg.REQ_DONE
            _same(srv.scene(rid).get_state(), _solo(fg, ctx, specs[k], 2), k)
        assert srv.summary().rejected == 1
    finally:
        srv.close()
