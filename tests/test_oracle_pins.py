"""Pins for the oracle members that no compiled reference block covers alone.

Each check is chosen so that a plausible restatement mistake fails it:
  - orc_velocity_from_solids (flipsolver2d.cpp:766-786, 1077-1104): which faces
    are scaled comes from the reference's own compiled predicates
    (MaterialGrid::u/vSampleAffectedBySolid, materialgrid.cpp:152-166); the
    scale factor is checked in closed form (one obstacle: exactly 1 - f; two
    obstacles: the mean friction of the in-bounds 3x3 solid ids, counted by a
    convolution of the id maps);
  - orc_density_rhs (flipsolver2d.cpp:993-1011): uniform rest density gives 0,
    twice it clamps to -0.5/dt, a quarter clamps to +0.5/dt, a value inside
    the clamp gives (1 - r)/dt, and rows that are not isFluid (compiled
    fluidTest) give 0;
  - orc_advect's kill rule (flipsolver2d.cpp:305-338): a particle dies exactly
    when its advected position's cell (compiled simmath::integr) is out of
    bounds or a SINK cell (compiled sinkTest), and survives otherwise;
  - orc_adjust_particles (flipsolver2d.cpp:261-303): for a pressure field
    p = a i^2 + b j^2 + c i j the x move is the difference across the two cell
    centres that bracket the particle along x, taken in the row of the
    particle's own cell (so a(2 ic + 1) + c j), and likewise along y; the
    move is zero along an axis where either bracketing cell is solid.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O

L = O.lib()
R = O.ref_lib()
needs_ref = pytest.mark.skipif(R is None, reason="reference building blocks not built (no reference tree)")


def _state(I, J, cap=16):
    return O.State(I, J, cap)


# ---------------------------------------------------------------------------
# orc_velocity_from_solids
# ---------------------------------------------------------------------------
def _random_solids(I, J, rng, frac=0.15):
    mat = np.where(rng.random((I, J)) < 0.6, O.FLUID, O.EMPTY).astype(np.int8)
    solid = rng.random((I, J)) < frac
    mat[solid] = O.SOLID
    return mat, solid


def _affected(mat, I, J):
    ua = np.zeros(I * J, np.uint8)
    va = np.zeros(I * J, np.uint8)
    R.ref_sample_affected_by_solid(np.ascontiguousarray(mat.ravel()), I, J, ua, va)
    return ua.reshape(I, J).astype(bool), va.reshape(I, J).astype(bool)


def _window_count(mask):
    """Number of True cells in the in-bounds 3x3 window around every cell."""
    p = np.pad(mask.astype(np.int64), 1)
    I, J = mask.shape
    return sum(p[1 + di:1 + di + I, 1 + dj:1 + dj + J] for di in (-1, 0, 1) for dj in (-1, 0, 1))


@needs_ref
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_velocity_from_solids_one_obstacle(seed):
    """One obstacle of friction f: every affected face with a solid id in its
    3x3 window is scaled by exactly (1 - f); every other face is unchanged."""
    rng = np.random.default_rng(seed)
    I, J = 23, 19
    mat, solid = _random_solids(I, J, rng)
    f = np.float32(0.375)
    p = O.make_params(I, J, friction=(f,))
    st = _state(I, J)
    st.material[:] = mat.ravel()
    st.solid_id[:] = np.where(solid, 0, -1).ravel()
    u0 = rng.uniform(-3, 3, (I + 1) * J).astype(np.float32)
    v0 = rng.uniform(-3, 3, I * (J + 1)).astype(np.float32)
    st.u[:] = u0
    st.v[:] = v0
    L.orc_velocity_from_solids(C.byref(p), st.ref())
    ua, va = _affected(mat, I, J)
    has_id = _window_count(solid) > 0
    u_scaled = np.zeros((I + 1, J), bool)
    u_scaled[:I] = ua & has_id
    v_scaled = np.zeros((I, J + 1), bool)
    v_scaled[:, :J] = va & has_id
    exp_u = np.where(u_scaled.ravel(), u0 * (np.float32(1) - f), u0).astype(np.float32)
    exp_v = np.where(v_scaled.ravel(), v0 * (np.float32(1) - f), v0).astype(np.float32)
    assert np.array_equal(st.u.view(np.uint32), exp_u.view(np.uint32))
    assert np.array_equal(st.v.view(np.uint32), exp_v.view(np.uint32))
    # the last U row and last V column are never visited
    assert np.array_equal(st.u.reshape(I + 1, J)[I], u0.reshape(I + 1, J)[I])
    assert np.array_equal(st.v.reshape(I, J + 1)[:, J], v0.reshape(I, J + 1)[:, J])
    assert u_scaled.any() and (~u_scaled[:I]).any()


@needs_ref
def test_velocity_from_solids_mean_of_window_ids():
    """Two obstacles of friction f1, f2: the factor is 1 - (k1 f1 + k2 f2)/(k1 + k2)
    with k the in-bounds 3x3 counts of each obstacle's solid cells (a cell that
    is not solid but carries an id still counts, as solid_id does there)."""
    rng = np.random.default_rng(5)
    I, J = 24, 24
    mat, solid = _random_solids(I, J, rng, 0.2)
    which = rng.random((I, J)) < 0.5
    ids = np.where(solid, np.where(which, 0, 1), -1).astype(np.int32)
    f1, f2 = 0.25, 0.75
    p = O.make_params(I, J, friction=(f1, f2))
    st = _state(I, J)
    st.material[:] = mat.ravel()
    st.solid_id[:] = ids.ravel()
    u0 = rng.uniform(0.5, 2, (I + 1) * J).astype(np.float32)
    v0 = rng.uniform(0.5, 2, I * (J + 1)).astype(np.float32)
    st.u[:] = u0
    st.v[:] = v0
    L.orc_velocity_from_solids(C.byref(p), st.ref())
    k1 = _window_count(ids == 0)
    k2 = _window_count(ids == 1)
    n = k1 + k2
    avg = np.where(n > 0, (k1 * f1 + k2 * f2) / np.maximum(n, 1), 0.0)
    ua, va = _affected(mat, I, J)
    fu = np.ones((I + 1, J))
    fu[:I] = np.where(ua & (n > 0), 1 - avg, 1.0)
    fv = np.ones((I, J + 1))
    fv[:, :J] = np.where(va & (n > 0), 1 - avg, 1.0)
    np.testing.assert_allclose(st.u, u0 * fu.ravel(), rtol=2e-6, atol=0)
    np.testing.assert_allclose(st.v, v0 * fv.ravel(), rtol=2e-6, atol=0)
    # mixed windows exist, so the mean is exercised, not only 1 - f
    mixed = (k1 > 0) & (k2 > 0) & ua
    assert mixed.sum() > 10


# ---------------------------------------------------------------------------
# orc_density_rhs
# ---------------------------------------------------------------------------
@needs_ref
def test_density_rhs_closed_forms():
    I, J = 8, 10
    rho = 1000.0
    dt = np.float32(0.01)
    p = O.make_params(I, J, density=rho)
    codes = np.int8([O.FLUID, O.SOURCE, O.EMPTY, O.SOLID, O.SINK])
    rng = np.random.default_rng(3)
    mat = codes[rng.integers(0, len(codes), I * J)]
    ratio = np.float32([1.0, 2.0, 0.25, 1.2, 0.8, 0.5, 1.5])
    dens = (ratio[rng.integers(0, len(ratio), I * J)] * np.float32(rho)).astype(np.float32)
    st = _state(I, J)
    st.material[:] = mat
    st.density[:] = dens
    rhs = np.zeros(I * J)
    L.orc_density_rhs(C.byref(p), st.ref(), dt, rhs)
    fluid = np.array([R.ref_material_bits(int(m)) & 1 for m in mat], bool)
    r = dens.astype(np.float64) / rho
    inv = 1.0 / float(dt)
    for c in range(I * J):
        if not fluid[c]:
            assert rhs[c] == 0.0
        elif r[c] >= 1.5:
            assert rhs[c] == pytest.approx(-0.5 * inv, rel=1e-15)
        elif r[c] <= 0.5:
            assert rhs[c] == pytest.approx(0.5 * inv, rel=1e-15)
        else:
            assert rhs[c] == pytest.approx((1.0 - r[c]) * inv, rel=1e-12, abs=1e-12)
    # rest density gives exactly zero; every branch was hit
    assert np.all(rhs[fluid & (dens == np.float32(rho))] == 0.0)
    assert {2.0, 0.25, 1.2}.issubset(set(np.round(r[fluid], 3)))
    assert (~fluid).any() and fluid[mat == O.SOURCE].all()


# ---------------------------------------------------------------------------
# orc_advect: the kill rule
# ---------------------------------------------------------------------------
@needs_ref
@pytest.mark.parametrize("disp", [(2.5, -1.25), (-3.0, 4.0)])
def test_advect_kills_out_of_bounds_and_sink(disp):
    """A uniform flow moves every particle by exactly (a, b) cells (RK4 is exact
    there); no solid anywhere.  Survivors are exactly the particles whose new
    cell is inside the grid and not a SINK cell."""
    I, J = 16, 12
    a, b = disp
    dt = np.float32(0.5)
    p = O.make_params(I, J)
    rng = np.random.default_rng(int(abs(10 * a)) + 1)
    n = 400
    st = _state(I, J, n)
    mat = np.where(rng.random((I, J)) < 0.5, O.FLUID, O.EMPTY).astype(np.int8)
    mat[4:9, 3:7] = O.SINK
    mat[12, 8] = O.SINK
    st.material[:] = mat.ravel()
    st.solid_sdf[:] = 1.0
    st.u[:] = np.float32(a / dt)
    st.v[:] = np.float32(b / dt)
    x = rng.uniform(-1.0, I + 1.0, n).astype(np.float32)
    y = rng.uniform(-1.0, J + 1.0, n).astype(np.float32)
    st.px[:n], st.py[:n] = x, y
    st.n = n
    alive = np.zeros(n, np.uint8)
    L.orc_advect(C.byref(p), st.ref(), dt, alive)
    # positions: the composition with the pinned RK4
    ox = np.zeros(n, np.float32)
    oy = np.zeros(n, np.float32)
    L.orc_rk4(st.u, st.v, I, J, dt, x, y, n, ox, oy)
    assert np.array_equal(st.px[:n].view(np.uint32), ox.view(np.uint32))
    assert np.allclose(ox, x + np.float32(a), atol=1e-5) and np.allclose(oy, y + np.float32(b), atol=1e-5)
    expect = np.zeros(n, np.uint8)
    for k in range(n):
        ci, cj = R.ref_integr(float(ox[k])), R.ref_integr(float(oy[k]))
        inb = 0 <= ci < I and 0 <= cj < J
        sink = inb and (R.ref_material_bits(int(mat[ci, cj])) & 32) != 0
        expect[k] = 1 if (inb and not sink) else 0
    assert np.array_equal(alive, expect)
    # every outcome occurs: out of bounds on each side, sink, survivor
    assert (expect == 1).sum() > 50 and (expect == 0).sum() > 50
    assert ((ox < 0) | (oy < 0)).any() and ((ox >= I) | (oy >= J)).any()
    assert any(0 <= R.ref_integr(float(ox[k])) < I and 0 <= R.ref_integr(float(oy[k])) < J and expect[k] == 0
               for k in range(n))
    # positions in (-1, 0) floor to cell -1 and die (a truncating cast would keep them)
    neg = ((ox > -1) & (ox < 0)) | ((oy > -1) & (oy < 0))
    assert neg.any() and np.all(alive[neg] == 0)


@needs_ref
def test_advect_projects_out_of_solids_then_applies_the_rule():
    """A particle advected into a solid is moved by closestSurfacePoint (the
    compiled SdfGrid) before the kill test."""
    I = J = 16
    p = O.make_params(I, J)
    n = 64
    st = _state(I, J, n)
    ii, jj = np.meshgrid(np.arange(I) + 0.5, np.arange(J) + 0.5, indexing="ij")
    sdf = (np.minimum.reduce([ii - 2.0, I - 2.0 - ii, jj - 2.0, J - 2.0 - jj])).astype(np.float32)
    st.solid_sdf[:] = sdf.ravel()
    st.material[:] = np.where(sdf.ravel() < 0, O.SOLID, O.FLUID).astype(np.int8)
    dt = np.float32(0.25)
    st.u[:] = np.float32(3.0 / dt)
    st.v[:] = np.float32(0.0)
    rng = np.random.default_rng(9)
    x = rng.uniform(6, 12, n).astype(np.float32)
    y = rng.uniform(3, 13, n).astype(np.float32)
    st.px[:n], st.py[:n] = x, y
    st.n = n
    alive = np.zeros(n, np.uint8)
    L.orc_advect(C.byref(p), st.ref(), dt, alive)
    ox = np.zeros(n, np.float32)
    oy = np.zeros(n, np.float32)
    L.orc_rk4(st.u, st.v, I, J, dt, x, y, n, ox, oy)
    lx = np.zeros(n, np.float32)
    L.orc_lerp(st.solid_sdf, I, J, 0.0, 0.0, ox, oy, n, lx)
    inside = lx < 0
    assert inside.sum() > 10
    px = np.zeros(n, np.float32)
    py = np.zeros(n, np.float32)
    R.ref_closest_surface_point(st.solid_sdf, I, J, np.ascontiguousarray(ox[inside]),
                                np.ascontiguousarray(oy[inside]), int(inside.sum()), px[:inside.sum()],
                                py[:inside.sum()])
    assert np.array_equal(st.px[:n][inside].view(np.uint32), px[:inside.sum()].view(np.uint32))
    assert np.array_equal(st.py[:n][inside].view(np.uint32), py[:inside.sum()].view(np.uint32))
    assert np.array_equal(st.px[:n][~inside].view(np.uint32), ox[~inside].view(np.uint32))
    fx, fy = st.px[:n], st.py[:n]
    expect = [1 if (0 <= R.ref_integr(float(fx[k])) < I and 0 <= R.ref_integr(float(fy[k])) < J) else 0
              for k in range(n)]
    assert np.array_equal(alive, np.uint8(expect)) and np.all(alive == 1)


# ---------------------------------------------------------------------------
# orc_adjust_particles
# ---------------------------------------------------------------------------
def test_adjust_particles_bracketing_differences():
    I, J = 20, 18
    a, b, c = 3.0, -2.0, 1.0
    ii, jj = np.meshgrid(np.arange(I), np.arange(J), indexing="ij")
    pr = (a * ii ** 2 + b * jj ** 2 + c * ii * jj).astype(np.float64).ravel()
    mat = np.full((I, J), O.FLUID, np.int8)
    mat[0, :] = mat[-1, :] = mat[:, 0] = mat[:, -1] = O.SOLID
    mat[9, 9] = mat[9, 10] = O.SOLID
    rho, dx = 1000.0, 1.0 / 32
    dt = np.float32(0.02)
    p = O.make_params(I, J, density=rho, dx=dx)
    rng = np.random.default_rng(4)
    n = 600
    st = _state(I, J, n)
    st.material[:] = mat.ravel()
    x = rng.uniform(0.0, I, n).astype(np.float32)
    y = rng.uniform(0.0, J, n).astype(np.float32)
    st.px[:n], st.py[:n] = x, y
    st.n = n
    L.orc_adjust_particles(C.byref(p), st.ref(), dt, pr)
    scale = np.float32((float(dt) * float(dt)) / (rho * dx * dx))

    def bracket(t):
        # the cell whose centre is the last one at or before t (centre k + 0.5);
        # a particle before the first centre uses cell 0 and its neighbour
        return max(int(np.floor(t - 0.5)), 0)

    moved_x = moved_y = 0
    for k in range(n):
        i, j = int(np.floor(x[k])), int(np.floor(y[k]))
        ic, jc = bracket(x[k]), bracket(y[k])
        solid = lambda q, r: mat[min(max(q, 0), I - 1), min(max(r, 0), J - 1)] == O.SOLID  # noqa: E731
        gx = 0.0 if (solid(ic, j) or solid(ic + 1, j)) else a * (2 * ic + 1) + c * j
        gy = 0.0 if (solid(i, jc) or solid(i, jc + 1)) else b * (2 * jc + 1) + c * i
        ex = np.float32(x[k] + np.float32(gx) * scale)
        ey = np.float32(y[k] + np.float32(gy) * scale)
        assert st.px[k] == ex, (k, x[k], y[k], ic, j)
        assert st.py[k] == ey, (k, x[k], y[k], i, jc)
        moved_x += gx != 0
        moved_y += gy != 0
    assert moved_x > n // 2 and moved_y > n // 2
    assert moved_x < n and moved_y < n   # some moves were cancelled by solids
