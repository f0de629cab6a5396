"""Serving: continuous batching with HBM-budget admission.

Requests arrive while others are on the device; each one's result must be
what the same scene gives when simulated alone (bit for bit: scenes share
launches, never data or reductions).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("px", "py", "vx", "vy", "pid", "u", "v", "material", "fluid_sdf", "pressure")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _solo_state(fg, ctx, spec, frames):
    sc = fg.make_scene(ctx, spec, tol=1e-6)
    for _ in range(frames):
        sc.step_frame()
    st = sc.get_state()
    sc.close()
    return st


def test_served_results_equal_solo(fg, ctx):
    specs = fg.mixed_scene_specs(7, master_seed=5, resolutions=(32, 48, 64))
    frames = [1, 3, 2, 2, 1, 3, 2]
    _log(f"serve vs solo: {[s['n'] for s in specs]} frames {frames}")
    srv = fg.Server(ctx, hbm_budget=8 << 30, max_resident=4)
    arrivals = {0: [0, 1], 1: [2, 3, 4], 3: [5], 4: [6]}   # tick -> requests submitted before it
    ids = {}
    tick = 0
    while True:
        for k in arrivals.get(tick, []):
            ids[k] = srv.submit_spec(specs[k], frames[k], keep=True, tol=1e-6)
        t = srv.tick()
        assert t.resident <= 4
        tick += 1
        if tick > 4 and t.queued == 0 and t.resident == 0:
            break
        assert tick < 50
    for k, rid in ids.items():
        s = srv.status(rid)
        assert s.state == fg.REQ_DONE and s.frames_done == frames[k]
        got = srv.scene(rid).get_state()
        ref = _solo_state(fg, ctx, specs[k], frames[k])
        for f in FIELDS:
            assert np.array_equal(got[f].view(np.uint8), ref[f].view(np.uint8)), (k, f)
        srv.release(rid)
        with pytest.raises(fg.FgError):
            srv.scene(rid)
    summ = srv.summary()
    assert summ.completed == 7 and summ.frames == sum(frames)
    assert summ.p99_frame_ms >= summ.p50_frame_ms > 0
    srv.close()


def test_budget_gates_admission(fg, ctx):
    specs = fg.mixed_scene_specs(6, master_seed=8, resolutions=(48, 64))
    sizes = []
    for sp in specs:
        p = fg.make_params(sp["n"], ppc=sp["ppc"])
        from ctypes import byref, c_uint64
        b = c_uint64()
        assert fg.lib().fg_scene_bytes(byref(p), byref(b)) == 0
        sizes.append(b.value)
    budget = sorted(sizes)[-1] + sorted(sizes)[-2]    # the two largest fit together; all six do not
    assert sum(sizes) > budget
    _log(f"budget {budget} sizes {sizes}")
    srv = fg.Server(ctx, hbm_budget=budget, max_resident=16)
    ids = [srv.submit_spec(sp, 2) for sp in specs]
    big = fg.make_params(512, ppc=8)
    walls, fluid = fg.dam_break_geometry(512)
    rej = srv.submit(big, walls, fluid, 1)
    assert srv.status(rej).state == fg.REQ_REJECTED
    waited = False
    for _ in range(40):
        t = srv.tick()
        assert t.resident_bytes <= budget
        waited |= t.queued > 0
        if t.queued == 0 and t.resident == 0:
            break
    assert waited
    for rid in ids:
        assert srv.status(rid).state == fg.REQ_DONE
    s = srv.summary()
    assert s.peak_resident_bytes <= budget and s.rejected == 1 and s.completed == 6
    srv.close()


def test_served_nbflip_and_viscous_requests_equal_solo(fg, ctx):
    """Requests of both solver classes, with and without viscosity, served
    together (the classes step as separate groups, the viscous scenes' model
    as a sub-group): each result equals the scene stepped alone."""
    kinds = [(48, fg.SIM_NBFLIP, False), (40, fg.SIM_FLIP, True), (56, fg.SIM_NBFLIP, True), (32, fg.SIM_FLIP, False)]
    frames = [2, 3, 2, 3]

    def params(n, sim, visc):
        return fg.make_params(n, ppc=4, tol=1e-6, sim=sim, viscosity=visc)

    srv = fg.Server(ctx, hbm_budget=8 << 30, max_resident=4)
    ids = {}
    for k, (n, sim, visc) in enumerate(kinds):
        walls, fluid = fg.dam_break_geometry(n)
        ids[k] = srv.submit(params(n, sim, visc), walls, fluid, frames[k], seed=k, friction=np.zeros(4, np.float32),
                            keep=True, fluid_viscosity=[4.0])
    for _ in range(10):
        t = srv.tick()
        if t.queued == 0 and t.resident == 0:
            break
    for k, (n, sim, visc) in enumerate(kinds):
        assert srv.status(ids[k]).state == fg.REQ_DONE
        got = srv.scene(ids[k]).get_state()
        sc = fg.Scene(ctx, params(n, sim, visc))
        walls, fluid = fg.dam_break_geometry(n)
        sc.init_objects(solids=walls, fluids=fluid, friction=np.zeros(4, np.float32), seed=k, fluid_viscosity=[4.0])
        for _ in range(frames[k]):
            sc.step_frame()
        ref = sc.get_state()
        sc.close()
        _log(f"  served {n}^2 sim={sim} viscosity={visc}: {len(ref['px'])} particles")
        for f in FIELDS + ("viscosity", "particle_viscosity"):
            assert np.array_equal(got[f].view(np.uint8), ref[f].view(np.uint8)), (k, f)
        srv.release(ids[k])
