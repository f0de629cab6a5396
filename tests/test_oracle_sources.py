"""Pins for the oracle's sources, sinks and reseeding.

  - orc_init_objects (flipsolver2d.cpp:502-590, 788-795): with no sources or
    sinks it equals orc_init_geometry (pinned against the compiled
    Geometry2d); with them, a cell's class follows the reference's pass order
    (sinks, then sources, then solids, then initial fluid, each overwriting
    the last), every membership decided by the compiled polygon SDF at the
    cell centre, and a SOURCE cell names the last source that holds it;
  - orc_after_transfer (390-410): exactly the faces U(i,j), V(i,j) of SOURCE
    cells whose emitter transfers velocity get velocity/dx and become valid;
  - orc_reseed (627-680, 1013-1019): a SOURCE cell holding c < ppc/2
    particles gains ppc/2 - c, inside the cell, appended in cell order with
    consecutive ids; velocities are the compiled StaggeredVelocityGrid's
    velocityAt where the source transfers, else zero; other cells gain none.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O

L = O.lib()
R = O.ref_lib()
needs_ref = pytest.mark.skipif(R is None, reason="reference building blocks not built (no reference tree)")


def box(x0, x1, y0, y1):
    return O.box_polys(x0, x1, y0, y1)


def _init(p, solids=(), sources=(), sinks=(), fluids=()):
    I, J = p.size_i, p.size_j
    st = O.State(I, J, I * J * p.particles_per_cell + 64)
    L.orc_init_objects(C.byref(p), st.ref(), *O._polys(solids), *O._polys(sources), *O._polys(sinks),
                       *O._polys(fluids))
    return st


def test_init_objects_without_sources_equals_init_geometry():
    n = 40
    p = O.make_params(n)
    walls, fluids = O.dam_break_geometry(n)
    a = _init(p, solids=walls, fluids=fluids)
    b = O.State(n, n, 16)
    L.orc_init_geometry(C.byref(p), b.ref(), *O._polys(walls), *O._polys(fluids))
    for k in ("material", "solid_id", "solid_sdf"):
        assert np.array_equal(getattr(a, k).view(np.uint8), getattr(b, k).view(np.uint8)), k
    assert np.all(a.emitter_id == -1)


@needs_ref
def test_init_objects_pass_order_and_membership():
    n = 32
    p = O.make_params(n)
    d = 1.0 / n
    solids = [box(0.0, 1.0, 0.0, 2 * d), box(0.5, 0.7, 0.5, 0.7)]
    sources = [box(0.2, 0.45, 0.2, 0.45), box(0.3, 0.6, 0.3, 0.6)]      # overlapping: the later wins
    sinks = [box(0.55, 0.9, 0.1, 0.35), box(0.1, 0.3, 0.6, 0.95)]
    fluids = [box(0.25, 0.35, 0.65, 0.8), box(0.62, 0.8, 0.62, 0.8)]
    st = _init(p, solids, sources, sinks, fluids)

    def inside(polys, x, y):
        return [R.ref_polygon_sdf(np.ascontiguousarray(q.ravel(), np.float32), len(q), x, y) <= 0 for q in polys]

    dx = np.float32(d)
    hits = {O.SINK: 0, O.SOURCE: 0, O.SOLID: 0, O.FLUID: 0}
    for i in range(n):
        for j in range(n):
            xf, yf = (np.float32(i) + np.float32(0.5)) * dx, (np.float32(j) + np.float32(0.5)) * dx
            xs, ys = np.float32(i) * dx + dx / np.float32(2), np.float32(j) * dx + dx / np.float32(2)
            mat, eid = O.EMPTY, -1
            if any(inside(sinks, xf, yf)):
                mat = O.SINK
            src = inside(sources, xs, ys)
            if any(src):
                mat, eid = O.SOURCE, max(k for k, v in enumerate(src) if v)
            if st.solid_sdf[i * n + j] < 0:
                mat = O.SOLID
            if any(inside(fluids, xf, yf)):
                mat = O.FLUID
            assert st.material[i * n + j] == np.int8(mat), (i, j)
            if mat == O.SOURCE:
                assert st.emitter_id[i * n + j] == eid, (i, j)
            hits[mat] = hits.get(mat, 0) + 1
    assert all(v > 3 for v in hits.values()), hits
    # the overlap of the two sources belongs to the second one
    assert (st.emitter_id == 0).any() and (st.emitter_id == 1).any()


def test_after_transfer_sets_source_faces():
    n = 24
    emitters = [(1.5, -0.75, True), (3.0, 2.0, False)]
    p = O.make_params(n, emitters=emitters)
    st = _init(p, sources=[box(0.2, 0.4, 0.2, 0.4), box(0.6, 0.8, 0.6, 0.8)])
    rng = np.random.default_rng(0)
    st.u[:] = rng.standard_normal(st.u.size).astype(np.float32)
    st.v[:] = rng.standard_normal(st.v.size).astype(np.float32)
    u0, v0 = st.u.copy(), st.v.copy()
    L.orc_after_transfer(C.byref(p), st.ref())
    src0 = (st.material == O.SOURCE) & (st.emitter_id == 0)
    assert src0.sum() > 4 and ((st.material == O.SOURCE) & (st.emitter_id == 1)).sum() > 4
    U = st.u.reshape(n + 1, n)
    V = st.v.reshape(n, n + 1)
    for c in range(n * n):
        i, j = divmod(c, n)
        if src0[c]:
            assert U[i, j] == np.float32(1.5 / p.dx) and V[i, j] == np.float32(-0.75 / p.dx)
            assert st.u_valid[i * n + j] == 1 and st.v_valid[i * (n + 1) + j] == 1
        else:
            assert U[i, j] == u0.reshape(n + 1, n)[i, j] and V[i, j] == v0.reshape(n, n + 1)[i, j]


@needs_ref
@pytest.mark.parametrize("ppc", [4, 7])
def test_reseed_fills_source_cells(ppc):
    n = 20
    emitters = [(0.0, 0.0, True), (0.0, 0.0, False)]
    p = O.make_params(n, ppc=ppc, emitters=emitters, reseed_seed=99)
    st = _init(p, sources=[box(0.2, 0.4, 0.2, 0.5), box(0.6, 0.8, 0.6, 0.8)])
    rng = np.random.default_rng(ppc)
    st.u[:] = rng.standard_normal(st.u.size).astype(np.float32)
    st.v[:] = rng.standard_normal(st.v.size).astype(np.float32)
    counts = rng.integers(0, ppc + 1, n * n).astype(np.int32)
    st.particle_counts[:] = counts
    n0 = 37
    st.n = n0
    st.c.next_pid = 1000
    st.c.rng_epoch = 5
    added = L.orc_reseed(C.byref(p), st.ref())
    half = ppc // 2
    src = st.material == O.SOURCE
    expect = np.where(src, np.maximum(half - counts, 0), 0)
    assert added == expect.sum() > 0 and st.n == n0 + added
    k = n0
    for c in np.nonzero(expect)[0]:
        i, j = divmod(int(c), n)
        m = int(expect[c])
        xs, ys = st.px[k:k + m], st.py[k:k + m]
        assert np.all((xs >= i) & (xs < i + 1) & (ys >= j) & (ys < j + 1)), c
        assert np.array_equal(st.pid[k:k + m], np.arange(1000 + k - n0, 1000 + k - n0 + m, dtype=np.uint32))
        ou, ov = np.zeros(m, np.float32), np.zeros(m, np.float32)
        R.ref_velocity_at(st.u, st.v, n, n, np.ascontiguousarray(xs), np.ascontiguousarray(ys), m, ou, ov)
        if st.emitter_id[c] == 0:
            assert np.array_equal(st.vx[k:k + m], ou) and np.array_equal(st.vy[k:k + m], ov)
        else:
            assert np.all(st.vx[k:k + m] == 0) and np.all(st.vy[k:k + m] == 0)
        k += m
    # a different substep draws different positions; the jitter is uniform
    st2 = _init(p, sources=[box(0.2, 0.4, 0.2, 0.5), box(0.6, 0.8, 0.6, 0.8)])
    st2.particle_counts[:] = counts
    st2.c.rng_epoch = 6
    L.orc_reseed(C.byref(p), st2.ref())
    assert not np.array_equal(st2.px[:added], st.px[n0:n0 + added])
    u = np.array([L.orc_reseed_jitter(7, e, c, q, a) for e in range(4) for c in range(50) for q in range(4)
                  for a in range(2)])
    assert np.all((u >= 0) & (u < 1)) and abs(u.mean() - 0.5) < 0.03


def test_reseed_stops_at_capacity():
    n = 12
    p = O.make_params(n, ppc=8)
    st = O.State(n, n, 10)
    L.orc_init_objects(C.byref(p), st.ref(), *O._polys([]), *O._polys([box(0.1, 0.9, 0.1, 0.9)]),
                       *O._polys([]), *O._polys([]))
    st.particle_counts[:] = 0
    assert L.orc_reseed(C.byref(p), st.ref()) == 10 and st.n == 10
