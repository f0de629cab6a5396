"""The particle sort on the GPU against the oracle's, with pile-ups.

The reference rebins particles into 3x3-cell bins after advection
(FlipSolver2dLib/markerparticlesystem.cpp:317-342); both the oracle and the
GPU keep particles in the canonical (cell, stable id, input index) order
instead (the oracle's orc_sort_particles).  The GPU orders each cell's
entries by one of three paths chosen by the cell's length: a register
network for up to 32 entries, a wave-cooperative rank sort up to 512 and an
insertion sort beyond.  Cells of every length class, shuffled ids, repeated
ids and a shuffled input order must come out bit-exact.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from helpers import oracle_state_dict

pytestmark = pytest.mark.gpu

# (cell i, cell j, particles added there): one per path and its borders
# (outside the dam's fluid, so each count is the cell's length exactly)
PILES = [(10, 10, 700), (20, 30, 300), (30, 30, 33), (40, 40, 40), (12, 40, 17), (50, 40, 32), (20, 50, 512),
         (44, 50, 513), (8, 30, 16)]


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _piled_state(orc, n, ppc, repeat_ids, seed):
    p, st = orc.dam_break(n, ppc=ppc, tol=1e-10)
    rng = np.random.default_rng(seed)
    d = oracle_state_dict(st)
    xs, ys = [d["px"]], [d["py"]]
    for (ci, cj, cnt) in PILES:
        xs.append((ci + rng.uniform(0.01, 0.99, cnt)).astype(np.float32))
        ys.append((cj + rng.uniform(0.01, 0.99, cnt)).astype(np.float32))
    px, py = np.concatenate(xs), np.concatenate(ys)
    N = len(px)
    vx = rng.standard_normal(N).astype(np.float32)
    vy = rng.standard_normal(N).astype(np.float32)
    pid = rng.permutation(N).astype(np.uint32)
    if repeat_ids:
        pid //= 3   # every id three times
    order = rng.permutation(N)
    d.update(px=px[order], py=py[order], vx=vx[order], vy=vy[order], pid=pid[order])
    st2 = orc.State(n, n, max(N, len(st.px)))
    for k in ("u", "v", "u_valid", "v_valid", "material", "solid_sdf", "fluid_sdf", "solid_id"):
        getattr(st2, k)[...] = d[k]
    for k in ("px", "py", "vx", "vy", "pid"):
        getattr(st2, k)[:N] = d[k]
    st2.n = N
    return p, st2, d


@pytest.mark.parametrize("repeat_ids", [False, True])
def test_sort_with_pileups_matches_oracle(fg, orc, ctx, repeat_ids):
    n, ppc = 64, 4
    p, st, d = _piled_state(orc, n, ppc, repeat_ids, seed=5 + int(repeat_ids))
    N = st.n
    _log(f"sort with pile-ups: {N} particles, piles {[c for (_, _, c) in PILES]}, repeated ids {repeat_ids}")
    sc = fg.Scene(ctx, fg.make_params(n, ppc=ppc, tol=1e-6, max_particles=N + 64))
    walls, _ = fg.dam_break_geometry(n)
    sc.set_friction(np.zeros(len(walls), np.float32))
    sc.set_state(d)
    dt = 1.0 / 120   # the grid velocity is zero: advection moves nothing
    sc.stage(fg.STAGE_ADVECT, dt)
    sc.stage(fg.STAGE_REBIN)
    g = sc.get_state()
    sc.close()
    L = orc.lib()
    alive = np.zeros(N, np.uint8)
    L.orc_advect(C.byref(p), st.ref(), C.c_float(dt), alive)
    cs = np.zeros(n * n + 1, np.int32)
    L.orc_sort_particles(C.byref(p), st.ref(), alive.ctypes.data, cs.ctypes.data)
    M = st.n
    assert M == N and len(g["px"]) == M
    counts = np.diff(cs)
    for (ci, cj, cnt) in PILES:
        assert counts[ci * n + cj] == cnt
    _log(f"  longest cell {counts.max()}, cells over 32: {(counts > 32).sum()}")
    assert np.array_equal(g["pid"], st.pid[:M])
    for k in ("px", "py", "vx", "vy"):
        assert np.array_equal(g[k].view(np.uint32), getattr(st, k)[:M].view(np.uint32)), k
