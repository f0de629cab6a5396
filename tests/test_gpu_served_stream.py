"""The served stream against the CPU oracle.

A few seconds' worth of the served workload's Poisson stream (resolutions
log-uniform in [64, 1024], dam-break geometry and gravity angle as in the
batched workload, 4-8 particles per cell, production tolerance 1e-6) with
4-10 frames each, submitted at their arrival times while the server
ticks, under an HBM budget of 1.5x the largest scene, so that admission
waits on it whenever large scenes overlap.  The server must finish every
request within the budget; the requests of the smallest scenes keep their
state, and each is compared with the oracle's frames of the same scene from
the same seeded initial state.
"""
import ctypes as C
import sys
import time

import numpy as np
import pytest

from helpers import rel_l2

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_served_stream_under_binding_budget_vs_oracle(fg, orc, ctx):
    reqs = fg.served_request_specs(rate=150.0, duration=1.0, master_seed=77, frame_range=(4, 10))
    sizes = []
    for r in reqs:
        b = C.c_uint64()
        p = fg.make_params(r["n"], ppc=r["ppc"])
        assert fg.lib().fg_scene_bytes(C.byref(p), C.byref(b)) == 0
        sizes.append(b.value)
    budget = max(sizes) * 3 // 2
    assert max(sizes) < budget < sum(sizes)
    small = sorted(range(len(reqs)), key=lambda k: reqs[k]["n"])[:3]
    _log(f"{len(reqs)} requests, resolutions {min(r['n'] for r in reqs)}-{max(r['n'] for r in reqs)}, "
         f"{sum(sizes) / 2**30:.1f} GiB of scenes under a {budget / 2**30:.2f} GiB budget; "
         f"oracle sample {[reqs[k]['n'] for k in small]}")
    srv = fg.Server(ctx, hbm_budget=budget, max_resident=256)
    ctx.set_batch_lanes(4)
    ids = {}
    try:
        t0 = time.time()
        nxt, ticks = 0, 0
        while True:
            now = time.time() - t0
            while nxt < len(reqs) and reqs[nxt]["t"] <= now:
                ids[nxt] = srv.submit_spec(reqs[nxt], reqs[nxt]["frames"], keep=nxt in small, tol=1e-6)
                nxt += 1
            t = srv.tick()
            ticks += 1
            assert t.resident_bytes <= budget
            if nxt == len(reqs) and t.queued == 0 and t.resident == 0:
                break
            if t.queued == 0 and t.resident == 0:   # idle until the next arrival
                time.sleep(max(0.0, min(reqs[nxt]["t"] - (time.time() - t0), 0.01)))
            assert time.time() - t0 < 90, "the stream did not drain"

        summ = srv.summary()
        _log(f"  {ticks} ticks in {time.time() - t0:.1f} s: completed {summ.completed}, rejected {summ.rejected}, "
             f"peak {summ.peak_resident_bytes / 2**30:.2f} GiB / {summ.peak_resident} scenes, "
             f"p50/p99 frame {summ.p50_frame_ms:.1f}/{summ.p99_frame_ms:.1f} ms, "
             f"p99 admit wait {summ.p99_admit_wait_ms:.0f} ms, {summ.budget_waits} ticks waited on the budget")
        assert summ.completed == len(reqs) and summ.rejected == 0 and summ.failed == 0
        assert summ.peak_resident_bytes <= budget and summ.budget_waits > 0 and summ.cap_waits == 0
        for k, rid in ids.items():
            s = srv.status(rid)
            assert s.state == fg.REQ_DONE and s.frames_done == reqs[k]["frames"], k
        got = {k: srv.scene(ids[k]).get_state() for k in small}
    finally:
        ctx.set_batch_lanes(1)
        srv.close()
    L = orc.lib()
    for k in small:
        r = reqs[k]
        n = r["n"]
        p = orc.make_params(n, ppc=r["ppc"], gravity=r["gravity"], tol=1e-12)
        walls, fluids = orc.dam_break_geometry(n, r["fluid_box"])
        st = orc.init_scene(p, walls, fluids, seed=r["seed"])
        for _ in range(r["frames"]):
            L.orc_step_frame(C.byref(p), st.ref(), C.byref(orc.OrcStats()))
        g = got[k]
        N = st.n
        assert len(g["px"]) == N and np.array_equal(g["pid"], st.pid[:N]), k
        e = {"px": rel_l2(g["px"], st.px[:N]), "py": rel_l2(g["py"], st.py[:N]),
             "vx": rel_l2(g["vx"], st.vx[:N]), "vy": rel_l2(g["vy"], st.vy[:N]),
             "p": rel_l2(g["pressure"], st.pressure)}
        _log(f"  request {k} ({n}^2, {r['frames']} frames) vs oracle, rel L2: "
             + " ".join(f"{a}={b:.2e}" for a, b in e.items()))
        assert max(e["px"], e["py"], e["vx"], e["vy"]) <= 1e-5, (k, e)
        assert e["p"] <= 1e-4, (k, e)
