"""The batched workload itself against the CPU oracle: the 256 mixed
dam-break scenes of the benchmark (mixed_scene_specs(256, master_seed=2024):
resolutions {64, 96, 128, 192, 256}, fluid-rectangle corners in [0.1, 0.6],
gravity 9.81 at +-30 degrees, 4-8 particles per cell) at the production
tolerance 1e-6.

After a few batched frames (so the flow is moving), one batched substep
(fg_batch_substep: every scene in one launch per phase) runs from a state
captured straight from the batch.  Every scene's solve must converge, with
its FP64 residual recomputed on the host from the matrix the solve used; a
sample of scenes, one of each resolution, is compared with the oracle's
converged substep from the same captured state, at the substep bar.
"""
import ctypes as C
import sys
import time

import numpy as np
import pytest

from helpers import assert_substep_parity, frame_dt, host_rel_residual, oracle_state_from_gpu

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_batched_256_mixed_scenes_vs_oracle(fg, orc, ctx):
    specs = fg.mixed_scene_specs(256, master_seed=2024)
    t0 = time.time()
    scenes = [fg.make_scene(ctx, sp, tol=1e-6) for sp in specs]
    _log(f"256 mixed scenes created in {time.time() - t0:.1f} s, "
         f"{sum(sc.particle_count() for sc in scenes)} particles")
    ctx.set_batch_lanes(4)
    try:
        for f in range(4):
            stats = fg.batch_step_frame(ctx, scenes)
            assert all(s.converged for s in stats), f
            _log(f"  frame {f}: max substeps {max(s.substeps for s in stats)}, "
                 f"max resid {max(s.pressure_rel_resid for s in stats):.2e}")
        before = [sc.get_state() for sc in scenes]
        params = [orc.make_params(sp["n"], ppc=sp["ppc"], gravity=sp["gravity"], tol=1e-12) for sp in specs]
        dts = []
        for sp, g, p in zip(specs, before, params):
            st = oracle_state_from_gpu(orc, g, sp["n"], sp["n"])
            dts.append(frame_dt(orc, orc.lib().orc_max_particle_velocity(st.ref()), p))
        t0 = time.time()
        stats = fg.batch_substep(ctx, scenes, dts)
        _log(f"  one batched substep of 256 scenes: {1e3 * (time.time() - t0):.1f} ms")
        after = [sc.get_state() for sc in scenes]
    finally:
        ctx.set_batch_lanes(1)
        for sc in scenes:
            sc.close()
    worst = 0.0
    for k, (sp, s, g) in enumerate(zip(specs, stats, after)):
        n = sp["n"]
        assert s.converged and s.pressure_rel_resid <= 1e-6, k
        r = host_rel_residual(orc, fg, g, n, n, dts[k], 1000.0, 1.0 / n)
        assert r <= 1.001e-6, (k, r)
        worst = max(worst, r)
    _log(f"  every scene converged; worst host-recomputed FP64 residual {worst:.3e}")
    # one scene of each resolution against the oracle
    sample = []
    for res in fg.MIXED_RESOLUTIONS:
        sample.append(next(k for k, sp in enumerate(specs) if sp["n"] == res))
    for k in sample:
        n = specs[k]["n"]
        st = oracle_state_from_gpu(orc, before[k], n, n, capacity=n * n * specs[k]["ppc"])
        orc.lib().orc_substep(C.byref(params[k]), st.ref(), C.c_float(dts[k]), C.byref(orc.OrcStats()))
        assert_substep_parity(after[k], st, _log, what=f"scene {k} ({n}^2, ppc {specs[k]['ppc']})")
