"""Pins for the oracle's NB-FLIP members (NBFlipSolver, nbflipsolver.cpp).

Each check is chosen so that a plausible restatement mistake fails it:
  - orc_extrapolate_levelset_outside (flipsolver2d.cpp:1496-1557): one known
    cell of value v gives exactly v + d at Chebyshev distance d; random known
    sets match a level-synchronous evaluation written independently (levels
    from a brute-force distance transform, each level averaged from the level
    below in getNeighborhood order, plus 1); a cell holding exactly I*J is not
    known (the test is strict);
  - orc_euler_advect (flipsolver2d.cpp:340-351): with zero velocity the output
    at (i, j) is the input grid lerped at the grid index (i, j) itself (the
    compiled Grid2d::lerp, not the sample's offset position); in a uniform
    flow a linear field comes out shifted upstream by u*dt (a sign, offset or
    axis mistake fails it);
  - orc_nb_prune (nbflipsolver.cpp:228-252): explicit cases of the two rules
    (deep-inside non-source particles die; crowded cells below the resampling
    band or SOURCE cells die), the SDF at each particle from the compiled lerp;
  - orc_nb_combine (nbflipsolver.cpp:320-419): the level set is the minimum of
    the particle SDF, the source distance and advected + 1; on a linear level
    set the faces switch from the advected to the particle value at the face
    index the closed form gives;
  - orc_nb_reseed (nbflipsolver.cpp:115-201): inside the band every cell is
    filled up to ppc with the counter RNG's positions and zero velocity, ids in
    cell order; outside the band (and in SOURCE cells above it) nothing is
    added; candidates whose lerped SDF (compiled lerp) is below the narrow band
    are skipped;
  - orc_nb_init_objects (nbflipsolver.cpp:203-211, 291-349): the fluid SDF is
    the compiled polygon distance at the double-arithmetic cell centre over dx,
    FLUID where negative (even over a solid), the source distance is taken at
    the cell corner, and no particle is seeded.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O

L = O.lib()
R = O.ref_lib()
needs_ref = pytest.mark.skipif(R is None, reason="reference building blocks not built (no reference tree)")

FAR = np.float32(np.sqrt(np.float32(3.4028235e38)) - np.float32(0.8))   # updateSdf with no particle near


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def neighbours(c, I, J):
    i, j = divmod(c, J)
    out = []
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            if di == 0 and dj == 0:
                continue
            a, b = i + di, j + dj
            out.append(a * J + b if 0 <= a < I and 0 <= b < J else -1)
    return out


# ---------------------------------------------------------------------------
# extrapolateLevelsetOutside
# ---------------------------------------------------------------------------
def test_levelset_outside_rings_of_one_seed():
    I, J = 15, 13
    g = np.full(I * J, FAR, np.float32)
    g[5 * J + 7] = -0.25
    L.orc_extrapolate_levelset_outside(g, I, J)
    ii, jj = np.meshgrid(np.arange(I), np.arange(J), indexing="ij")
    d = np.maximum(abs(ii - 5), abs(jj - 7)).ravel()
    assert np.array_equal(bits(g), bits((np.float32(-0.25) + d).astype(np.float32)))


def test_levelset_outside_known_set_is_strict():
    I, J = 6, 5
    g = np.full(I * J, FAR, np.float32)
    g[0] = np.float32(I * J)          # not below I*J: unknown
    g[I * J - 1] = np.float32(I * J - 1)
    L.orc_extrapolate_levelset_outside(g, I, J)
    assert g[I * J - 1] == I * J - 1
    assert g[0] == np.float32(I * J - 1 + max(I - 1, J - 1))


def level_sync_outside(g, I, J):
    g = g.astype(np.float32).copy()
    known = g < np.float32(I * J)
    ks = np.argwhere(known.reshape(I, J))
    if len(ks) == 0:
        return g
    ii, jj = np.meshgrid(np.arange(I), np.arange(J), indexing="ij")
    lvl = np.full((I, J), 1 << 30)
    for a, b in ks:
        lvl = np.minimum(lvl, np.maximum(abs(ii - a), abs(jj - b)))
    lvl = lvl.ravel()
    for L_ in range(1, int(lvl.max()) + 1):
        for c in np.flatnonzero(lvl == L_):
            acc, cnt = 0.0, 0
            for k in neighbours(c, I, J):
                if k >= 0 and lvl[k] < L_:
                    acc += float(g[k])
                    cnt += 1
            g[c] = np.float32(acc / cnt + 1.0)
    return g


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_levelset_outside_matches_level_synchronous(seed):
    rng = np.random.default_rng(seed)
    I, J = 21, 17
    g = np.full(I * J, FAR, np.float32)
    pick = rng.random(I * J) < 0.08
    g[pick] = rng.uniform(-3, 2, pick.sum()).astype(np.float32)
    a = g.copy()
    L.orc_extrapolate_levelset_outside(a, I, J)
    assert np.array_equal(bits(a), bits(level_sync_outside(g, I, J)))


def test_levelset_outside_without_known_cells_changes_nothing():
    g = np.full(7 * 9, FAR, np.float32)
    a = g.copy()
    L.orc_extrapolate_levelset_outside(a, 7, 9)
    assert np.array_equal(bits(a), bits(g))


# ---------------------------------------------------------------------------
# eulerAdvectionThread
# ---------------------------------------------------------------------------
@needs_ref
@pytest.mark.parametrize("si,sj,oi,oj", [(13, 11, 0.5, 0.0), (12, 12, 0.0, 0.5), (12, 11, 0.0, 0.0)])
def test_euler_zero_velocity_samples_the_grid_index(si, sj, oi, oj):
    I, J = 12, 11
    u = np.zeros((I + 1) * J, np.float32)
    v = np.zeros(I * (J + 1), np.float32)
    rng = np.random.default_rng(si + sj)
    g = rng.standard_normal(si * sj).astype(np.float32)
    out = np.zeros(si * sj, np.float32)
    L.orc_euler_advect(u, v, I, J, np.float32(0.01), g, si, sj, oi, oj, out)
    ii, jj = np.meshgrid(np.arange(si, dtype=np.float32), np.arange(sj, dtype=np.float32), indexing="ij")
    ref = np.zeros(si * sj, np.float32)
    R.ref_lerp(g, si, sj, oi, oj, np.ascontiguousarray(ii.ravel()), np.ascontiguousarray(jj.ravel()), si * sj, ref)
    assert np.array_equal(bits(out), bits(ref))


@pytest.mark.parametrize("which", ["u", "v", "sdf"])
def test_euler_uniform_flow_shifts_a_linear_field_upstream(which):
    I, J = 40, 36
    a, b, dt = 3.0, -2.0, 0.5
    u = np.full((I + 1) * J, a, np.float32)
    v = np.full(I * (J + 1), b, np.float32)
    si, sj, oi, oj = {"u": (I + 1, J, 0.5, 0.0), "v": (I, J + 1, 0.0, 0.5), "sdf": (I, J, 0.0, 0.0)}[which]
    al, be, ga = 0.75, -0.5, 2.0
    ii, jj = np.meshgrid(np.arange(si, dtype=np.float64), np.arange(sj, dtype=np.float64), indexing="ij")
    g = (al * ii + be * jj + ga).astype(np.float32).ravel()
    out = np.zeros(si * sj, np.float32)
    L.orc_euler_advect(u, v, I, J, np.float32(dt), g, si, sj, oi, oj, out)
    # a linear grid lerped at x gives its value at x + offset - 1/2
    x, y = ii - a * dt, jj - b * dt
    want = al * (x + oi - 0.5) + be * (y + oj - 0.5) + ga
    inner = (x > 2) & (x < si - 3) & (y > 2) & (y < sj - 3)
    assert inner.sum() > 500
    assert np.max(np.abs(out.reshape(si, sj)[inner] - want[inner])) < 1e-4


# ---------------------------------------------------------------------------
# pruneNarrowBand
# ---------------------------------------------------------------------------
@needs_ref
def test_nb_prune_rules():
    I, J, ppc = 12, 10, 4
    p = O.make_params(I, J, ppc=ppc)
    st = O.State(I, J, 64)
    # level set falling along i: sample k holds 1 - k and sits at k + 1/2, so
    # a particle at x reads 1.5 - x
    ii = np.repeat(np.arange(I, dtype=np.float32), J)
    st.fluid_sdf[:] = 1.0 - ii
    st.material[:] = O.FLUID
    src = 8 * J + 3
    st.material[src] = O.SOURCE
    crowded = [6 * J + 2, 8 * J + 3, 2 * J + 5, 9 * J + 7]
    st.particle_counts[:] = ppc
    st.particle_counts[crowded] = 2 * ppc + 1
    cases = [   # (x, y, alive in, expected alive)
        (9.5, 7.5, 1, 0),   # sdf -8, not a source: below the narrow band
        (9.5, 7.5, 0, 0),   # already dead stays dead
        (8.5, 3.5, 1, 0),   # SOURCE cell (sdf -7 does not kill it), crowded
        (8.6, 4.5, 1, 0),   # sdf -7.1, plain cell
        (6.5, 2.5, 1, 0),   # sdf -5 < -3
        (3.5, 2.5, 1, 1),   # sdf -2: inside the band, count <= 2 ppc
        (3.5, 6.5, 1, 1),
        (2.7, 5.5, 1, 0),   # sdf -1.2 < -1 in a crowded cell
        (2.2, 5.5, 1, 1),   # sdf -0.7 in the same crowded cell: above the resampling band
        (1.5, 5.5, 1, 1),   # sdf 0
        (0.5, 0.5, 1, 1),
    ]
    n = len(cases)
    st.n = n
    st.px[:n] = [c[0] for c in cases]
    st.py[:n] = [c[1] for c in cases]
    alive = np.array([c[2] for c in cases], np.uint8)
    sd = np.zeros(n, np.float32)
    R.ref_lerp(st.fluid_sdf, I, J, 0.0, 0.0, np.ascontiguousarray(st.px[:n]), np.ascontiguousarray(st.py[:n]), n, sd)
    expect = []
    for k, (x, y, a, _) in enumerate(cases):
        cell = int(x) * J + int(y)
        is_src = R.ref_material_bits(int(st.material[cell])) & 16
        dead = (not is_src and sd[k] < -3) or ((is_src or sd[k] < -1) and st.particle_counts[cell] > 2 * ppc)
        expect.append(0 if (not a or dead) else 1)
    hand = [c[3] for c in cases]
    assert expect == hand
    L.orc_nb_prune(C.byref(p), st.ref(), alive)
    assert list(alive) == hand


# ---------------------------------------------------------------------------
# updateGridFromSources + combineAdvectedGrids
# ---------------------------------------------------------------------------
def test_nb_combine_level_set_is_the_minimum():
    I, J = 9, 8
    rng = np.random.default_rng(5)
    p = O.make_params(I, J)
    st = O.State(I, J, 4)
    ps = rng.uniform(-4, 4, I * J).astype(np.float32)
    adv = rng.uniform(-4, 4, I * J).astype(np.float32)
    src = rng.uniform(-4, 4, I * J).astype(np.float32)
    src[rng.random(I * J) < 0.5] = np.float32(3.4028235e38)
    p.source_sdf = src.ctypes.data
    st.fluid_sdf[:] = ps
    au = np.zeros((I + 1) * J, np.float32)
    av = np.zeros(I * (J + 1), np.float32)
    L.orc_nb_combine(C.byref(p), st.ref(), au, av, adv)
    assert np.array_equal(bits(st.fluid_sdf), bits(np.minimum(np.minimum(ps, src), adv + np.float32(1))))
    # no source grid: the same minimum without it
    p.source_sdf = None
    st.fluid_sdf[:] = ps
    L.orc_nb_combine(C.byref(p), st.ref(), au, av, adv)
    assert np.array_equal(bits(st.fluid_sdf), bits(np.minimum(ps, adv + np.float32(1))))


def test_nb_combine_faces_switch_at_the_combination_band():
    I, J = 12, 7
    p = O.make_params(I, J)
    st = O.State(I, J, 4)
    st.fluid_sdf[:] = 1e3
    adv = (np.repeat(np.arange(I, dtype=np.float32), J) - np.float32(6.5)).astype(np.float32)
    st.u[:] = 1.0
    st.v[:] = 2.0
    au = np.full((I + 1) * J, -1.0, np.float32)
    av = np.full(I * (J + 1), -2.0, np.float32)
    L.orc_nb_combine(C.byref(p), st.ref(), au, av, adv)
    # combined level set i - 5.5; U face (i, j+1/2) reads i - 6 (clamped at i = 0),
    # V face (i+1/2, j) reads i - 5.5: above -2 from i = 5 (U) and i = 4 (V)
    u = st.u.reshape(I + 1, J)
    v = st.v.reshape(I, J + 1)
    for i in range(I + 1):
        assert np.all(u[i] == (1.0 if i >= 5 else -1.0)), i
    for i in range(I):
        assert np.all(v[i] == (2.0 if i >= 4 else -2.0)), i


# ---------------------------------------------------------------------------
# reseedParticles
# ---------------------------------------------------------------------------
def _reseed_setup(I, J, ppc, sdf, seed=7, epoch=3):
    p = O.make_params(I, J, ppc=ppc, reseed_seed=seed)
    st = O.State(I, J, I * J * ppc + 8)
    st.fluid_sdf[:] = sdf
    st.material[:] = O.FLUID
    st.c.rng_epoch = epoch
    st.c.next_pid = 1000
    return p, st


def test_nb_reseed_fills_band_cells_to_ppc():
    I, J, ppc = 5, 6, 8
    p, st = _reseed_setup(I, J, ppc, -2.0)
    rng = np.random.default_rng(1)
    st.particle_counts[:] = rng.integers(0, ppc + 3, I * J)
    added = L.orc_nb_reseed(C.byref(p), st.ref())
    want = np.maximum(ppc - st.particle_counts, 0)
    assert added == want.sum() == st.n
    k = 0
    for c in range(I * J):
        i, j = divmod(c, J)
        for q in range(want[c]):
            x = np.float32(i) + np.float32(L.orc_reseed_jitter(7, 3, c, q, 0))
            y = np.float32(j) + np.float32(L.orc_reseed_jitter(7, 3, c, q, 1))
            assert st.px[k] == x and st.py[k] == y and st.vx[k] == 0 and st.vy[k] == 0
            assert st.pid[k] == 1000 + k
            k += 1


@pytest.mark.parametrize("sdf", [-0.5, -1.0, -3.0, -5.0, 0.5])
def test_nb_reseed_outside_the_band_adds_nothing(sdf):
    p, st = _reseed_setup(4, 4, 4, sdf)
    assert L.orc_nb_reseed(C.byref(p), st.ref()) == 0 and st.n == 0


def test_nb_reseed_source_cells():
    # a SOURCE cell is filled below the resampling band, not above it
    p, st = _reseed_setup(3, 3, 4, -0.5)
    st.material[4] = O.SOURCE
    assert L.orc_nb_reseed(C.byref(p), st.ref()) == 0
    p, st = _reseed_setup(3, 3, 4, -6.0)
    st.material[4] = O.SOURCE
    st.fluid_sdf[:] = -2.0
    st.fluid_sdf[4] = -1.5
    st.material[:] = O.SOLID
    st.material[4] = O.SOURCE
    added = L.orc_nb_reseed(C.byref(p), st.ref())
    assert added == 9 * 4   # every cell is in the band (-3, -1)


@needs_ref
def test_nb_reseed_skips_candidates_below_the_narrow_band():
    I, J, ppc = 8, 5, 8
    ii = np.repeat(np.arange(I, dtype=np.float32), J)
    p, st = _reseed_setup(I, J, ppc, 0.0)
    st.fluid_sdf[:] = 1.0 - ii      # cells 3..4 are in (-3, -1); candidates reach past -3 in row 4
    L.orc_nb_reseed(C.byref(p), st.ref())
    got = set()
    for k in range(st.n):
        got.add((float(st.px[k]), float(st.py[k])))
    want = set()
    for c in range(I * J):
        i, j = divmod(c, J)
        s = st.fluid_sdf[c]
        if not (-3 < s < -1):
            continue
        for q in range(ppc):
            x = np.float32(i) + np.float32(L.orc_reseed_jitter(7, 3, c, q, 0))
            y = np.float32(j) + np.float32(L.orc_reseed_jitter(7, 3, c, q, 1))
            ns = np.zeros(1, np.float32)
            R.ref_lerp(st.fluid_sdf, I, J, 0.0, 0.0, np.array([x], np.float32), np.array([y], np.float32), 1, ns)
            if ns[0] >= -3:
                want.add((float(x), float(y)))
    assert got == want
    assert 0 < len(want) < 2 * J * ppc   # some candidates of row 4 fell below the band


# ---------------------------------------------------------------------------
# firstFrameInit
# ---------------------------------------------------------------------------
@needs_ref
def test_nb_init_objects_fluid_and_source_distance():
    n = 24
    p = O.make_params(n, sim=O.SIM_NBFLIP)
    walls, fluids = O.dam_break_geometry(n)
    source = [O.box_polys(0.5, 0.7, 0.6, 0.8)]
    # a fluid box that also covers part of a wall: FLUID wins there
    fluids = fluids + [O.box_polys(0.0, 0.2, 0.0, 0.2)]
    st = O.init_nb_scene(p, solids=walls, sources=source, fluids=fluids)
    assert st.n == 0 and not st.particle_counts.any()
    dx = np.float32(p.dx)
    fsdf = st.fluid_sdf.reshape(n, n)
    ssdf = np.frombuffer((C.c_float * (n * n)).from_address(p.source_sdf), np.float32).reshape(n, n)
    for i in range(n):
        for j in range(n):
            x = np.float32((float(np.float32(i)) + 0.5) * float(dx))
            y = np.float32((float(np.float32(j)) + 0.5) * float(dx))
            d = min(np.float32(R.ref_polygon_sdf(np.ascontiguousarray(f.ravel()), len(f), x, y)) / dx for f in fluids)
            assert fsdf[i, j] == d, (i, j)
            if d < 0:
                assert st.material[i * n + j] == O.FLUID
            s = np.float32(R.ref_polygon_sdf(np.ascontiguousarray(source[0].ravel()), 4, np.float32(i) * dx,
                                             np.float32(j) * dx)) / dx
            assert ssdf[i, j] == s, (i, j)
    assert (st.material == O.SOLID).any() and (st.material == O.SOURCE).any()


# ---------------------------------------------------------------------------
# NB-FLIP with viscosity (nbflipsolver.cpp:66-113, 115-201, 291-349, 400-504)
# ---------------------------------------------------------------------------
@needs_ref
def test_nb_viscosity_grid_of_one_particle():
    """NBFlipSolver::centeredParamsToGridThread: |w| > 1e-6 (FlipSolver's
    uses w > 1e-9), v w / (w + 1e-10) where known, 0 elsewhere."""
    I, J = 9, 10
    p = O.make_params(I, J)
    st = O.State(I, J, 4)
    st.n = 1
    st.px[0], st.py[0] = 4.3, 5.8
    st.pvisc[0] = 2.5
    cs = np.zeros(I * J + 1, np.int32)
    cs[4 * J + 5 + 1:] = 1
    L.orc_nb_viscosity_grid(C.byref(p), st.ref(), cs)
    g = st.viscosity.reshape(I, J)
    for i in range(I):
        for j in range(J):
            w = np.float32(R.ref_quadratic_bspline(np.float32(4.3) - np.float32(i), np.float32(5.8) - np.float32(j)))
            want = np.float32(np.float32(w * np.float32(2.5)) / np.float32(np.float32(1e-10) + w)) if abs(w) > 1e-6 \
                else np.float32(0.0)
            assert g[i, j] == want, (i, j)


def test_nb_init_viscosity_of_the_nearest_fluid():
    n = 32
    p = O.make_params(n, sim=O.SIM_NBFLIP, viscosity=True, fluid_viscosity=[2.0, 9.0])
    walls, _ = O.dam_break_geometry(n)
    fluids = [O.box_polys(0.1, 0.45, 0.1, 0.5), O.box_polys(0.55, 0.9, 0.1, 0.5)]
    st = O.init_nb_scene(p, solids=walls, fluids=fluids)
    g = st.viscosity.reshape(n, n)
    fl = st.fluid_sdf.reshape(n, n) < 0
    xs = (np.arange(n) + 0.5) / n
    left = xs[:, None] < 0.5
    assert np.all(g[fl & left] == 2.0) and np.all(g[fl & ~left] == 9.0) and np.all(g[~fl] == 0.0)


@needs_ref
def test_nb_reseed_viscosity_from_the_grid():
    I, J, ppc = 6, 5, 4
    p, st = _reseed_setup(I, J, ppc, -2.0)
    rng = np.random.default_rng(2)
    st.viscosity[:] = rng.uniform(0, 5, I * J).astype(np.float32)
    st.particle_counts[:] = 1
    n = L.orc_nb_reseed(C.byref(p), st.ref())
    want = np.zeros(n, np.float32)
    R.ref_lerp(st.viscosity, I, J, 0.5, 0.5, np.ascontiguousarray(st.px[:n]), np.ascontiguousarray(st.py[:n]), n, want)
    assert n == I * J * (ppc - 1) and np.array_equal(st.pvisc[:n], want)


def test_nb_combine_viscosity_band_and_sources():
    I, J = 12, 7
    p = O.make_params(I, J, emitters=[(0.0, 0.0, False, 4.0), (0.0, 0.0, False, 6.0)])
    st = O.State(I, J, 4)
    # combined level set i - 5.5 at the lerp's samples: the cell centre
    # (i + 1/2, j + 1/2) reads i - 5.5 -> above -2 from i = 4
    st.fluid_sdf[:] = np.repeat(np.arange(I, dtype=np.float32), J) - np.float32(5.5)
    st.viscosity[:] = 1.0
    adv = np.full(I * J, 3.0, np.float32)
    st.material[:] = O.FLUID
    st.material[2 * J + 3] = O.SOURCE
    st.emitter_id[:] = -1
    st.emitter_id[2 * J + 3] = 0
    sid = np.full(I * J, -1, np.int32)
    sid[9 * J + 1] = 1            # a cell inside source 1 (corner distance < 0)
    p.source_id = sid.ctypes.data
    L.orc_nb_combine_viscosity(C.byref(p), st.ref(), adv)
    g = st.viscosity.reshape(I, J)
    for i in range(I):
        for j in range(J):
            if (i, j) == (9, 1):
                want = 6.0                       # the source's viscosity, to both grids
            elif i >= 4:
                want = 1.0                       # the particle grid above the band
            elif (i, j) == (2, 3):
                want = 3.0                       # SOURCE cell's emitter value, then the advected one below the band
            else:
                want = 3.0
            assert g[i, j] == want, (i, j)
    assert adv[9 * J + 1] == 6.0
