"""The oracle against the reference's own compiled building blocks.

oracle/_ref/libfsref.so is built from the reference sources where they lie
(oracle/Makefile), IEEE-strict, with the reference's thread pool run inline.
Every comparison here is bit-exact unless stated.  FlipSolver itself
(flipsolver2d.cpp) needs Eigen and is not built; its members are pinned by
test_oracle_math.py.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O

R = O.ref_lib()
L = O.lib()
pytestmark = pytest.mark.skipif(R is None, reason="reference building blocks not built (no reference tree)")


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def test_bspline_and_hats_match_reference():
    rng = np.random.default_rng(1)
    xs = np.concatenate([np.linspace(-2, 2, 4001, dtype=np.float32),
                         np.float32([0.0, -0.0, 0.5, -0.5, 1.5, -1.5, 1.0, -1.0, 1e-30, 1.4999999, 0.49999997]),
                         rng.uniform(-3, 3, 2000).astype(np.float32)])
    for x in xs:
        assert np.float32(L.orc_bspline(x)).view(np.uint32) == np.float32(R.ref_bspline(x)).view(np.uint32), x
        assert np.float32(L.orc_linear_hat(x)).view(np.uint32) == np.float32(R.ref_linear_hat(x)).view(np.uint32), x
    for x, y in rng.uniform(-2, 2, (3000, 2)).astype(np.float32):
        assert np.float32(L.orc_quadratic_bspline(x, y)) == np.float32(R.ref_quadratic_bspline(x, y))


@pytest.mark.parametrize("si,sj,oi,oj", [(9, 7, 0.0, 0.0), (10, 7, 0.5, 0.0), (9, 8, 0.0, 0.5), (2, 3, 0.5, 0.0),
                                         (33, 17, 0.0, 0.0)])
def test_lerp_matches_reference(si, sj, oi, oj):
    rng = np.random.default_rng(si * 100 + sj)
    g = rng.standard_normal(si * sj).astype(np.float32)
    n = 5000
    xi = rng.uniform(-2, si + 2, n).astype(np.float32)
    xj = rng.uniform(-2, sj + 2, n).astype(np.float32)
    # exact half/integer lattice points and clamp edges
    hi, hj = np.meshgrid(np.arange(-1, si + 1, 0.5, dtype=np.float32), np.arange(-1, sj + 1, 0.5, dtype=np.float32))
    xi = np.ascontiguousarray(np.concatenate([xi, hi.ravel()]))
    xj = np.ascontiguousarray(np.concatenate([xj, hj.ravel()]))
    a = np.zeros(len(xi), np.float32)
    b = np.zeros(len(xi), np.float32)
    L.orc_lerp(g, si, sj, oi, oj, xi, xj, len(xi), a)
    R.ref_lerp(g, si, sj, oi, oj, xi, xj, len(xi), b)
    assert np.array_equal(bits(a), bits(b))


def test_velocity_at_matches_reference():
    rng = np.random.default_rng(3)
    I, J = 24, 19
    u = rng.standard_normal((I + 1) * J).astype(np.float32)
    v = rng.standard_normal(I * (J + 1)).astype(np.float32)
    n = 20000
    xi = rng.uniform(-1, I + 1, n).astype(np.float32)
    xj = rng.uniform(-1, J + 1, n).astype(np.float32)
    out = [np.zeros(n, np.float32) for _ in range(4)]
    L.orc_velocity_at(u, v, I, J, xi, xj, n, out[0], out[1])
    R.ref_velocity_at(u, v, I, J, xi, xj, n, out[2], out[3])
    assert np.array_equal(bits(out[0]), bits(out[2]))
    assert np.array_equal(bits(out[1]), bits(out[3]))


def _dam_sdf(n):
    p, st = O.dam_break(n, ppc=1)
    return st.solid_sdf.copy()


@pytest.mark.parametrize("n", [32, 50])
def test_closest_surface_point_matches_reference(n):
    sdf = _dam_sdf(n)
    rng = np.random.default_rng(n)
    # points inside the walls and near them, where advect projects particles
    m = 4000
    xi = np.concatenate([rng.uniform(-0.5, 1.6, m), rng.uniform(n - 1.6, n + 0.5, m), rng.uniform(0, n, m)])
    xj = np.concatenate([rng.uniform(0, n, m), rng.uniform(0, n, m), rng.uniform(-0.5, 1.6, m)])
    xi = xi.astype(np.float32)
    xj = xj.astype(np.float32)
    N = len(xi)
    a = [np.zeros(N, np.float32) for _ in range(4)]
    L.orc_closest_surface_point(sdf, n, n, xi, xj, N, a[0], a[1])
    R.ref_closest_surface_point(sdf, n, n, xi, xj, N, a[2], a[3])
    assert np.array_equal(bits(a[0]), bits(a[2]))
    assert np.array_equal(bits(a[1]), bits(a[3]))


def test_closest_surface_point_random_sdf_matches_reference():
    rng = np.random.default_rng(11)
    I, J = 20, 23
    sdf = (rng.standard_normal(I * J) * 3).astype(np.float32)
    xi = rng.uniform(0, I, 3000).astype(np.float32)
    xj = rng.uniform(0, J, 3000).astype(np.float32)
    a = [np.zeros(3000, np.float32) for _ in range(4)]
    L.orc_closest_surface_point(sdf, I, J, xi, xj, 3000, a[0], a[1])
    R.ref_closest_surface_point(sdf, I, J, xi, xj, 3000, a[2], a[3])
    assert np.array_equal(bits(a[0]), bits(a[2]))
    assert np.array_equal(bits(a[1]), bits(a[3]))


@pytest.mark.parametrize("si,sj,radius,density", [(40, 33, 10, 0.15), (17, 60, 3, 0.05), (64, 64, 10, 0.4),
                                                  (25, 25, 10, 0.0), (12, 9, 1, 0.9)])
def test_extrapolate_matches_reference(si, sj, radius, density):
    rng = np.random.default_rng(si * sj + radius)
    g = rng.standard_normal(si * sj).astype(np.float32)
    valid = (rng.random(si * sj) < density).astype(np.uint8)
    g1, v1 = g.copy(), valid.copy()
    g2, v2 = g.copy(), valid.copy()
    L.orc_extrapolate(g1, v1, si, sj, radius)
    R.ref_extrapolate(g2, v2, si, sj, radius)
    assert np.array_equal(v1, v2)
    assert np.array_equal(bits(g1), bits(g2))


def test_polygon_sdf_matches_reference():
    rng = np.random.default_rng(5)
    polys = [O.box_polys(0.1, 0.4, 0.2, 0.9),
             np.float32([[0, 0], [1, 0], [0.5, 1]]),
             np.float32([[0, 0], [0, 1], [0.3, 0.4], [1, 1], [1, 0]]),    # concave
             np.float32([[0.2, 0.2], [0.8, 0.25], [0.9, 0.7], [0.4, 0.95], [0.1, 0.6]])]
    for poly in polys:
        v = np.ascontiguousarray(poly.ravel(), np.float32)
        for x, y in rng.uniform(-0.3, 1.3, (1500, 2)).astype(np.float32):
            a = np.float32(L.orc_polygon_sdf(v, len(poly), x, y))
            b = np.float32(R.ref_polygon_sdf(v, len(poly), x, y))
            assert a.view(np.uint32) == b.view(np.uint32), (poly, x, y)


def test_material_predicates_match_reference():
    for code in range(-128, 128):
        assert L.orc_material_bits(code) == R.ref_material_bits(code), code


def test_nonsolid_count_lin_matches_reference():
    rng = np.random.default_rng(9)
    for I, J in [(6, 5), (11, 13)]:
        mat = rng.choice(np.int8([O.FLUID, O.SOLID, O.EMPTY, O.SINK, O.SOURCE]), I * J).astype(np.int8)
        for lin in range(-2 * J - 3, I * J + 2 * J + 3):
            assert L.orc_nonsolid_count_lin(mat, I, J, lin) == R.ref_nonsolid_count_lin(mat, I, J, lin), lin


def test_split_range_matches_reference():
    T = R.ref_thread_count()
    rng = np.random.default_rng(2)
    for length in list(range(0, 70)) + list(rng.integers(70, 2_000_000, 60)):
        for mn, jobs in [(1, 1), (1, 8), (1, 16), (3, 2)]:
            s1 = np.zeros(4096, np.int64)
            e1 = np.zeros(4096, np.int64)
            s2 = np.zeros(4096, np.int64)
            e2 = np.zeros(4096, np.int64)
            n1 = L.orc_split_range(int(length), mn, jobs, T, s1, e1)
            n2 = R.ref_split_range(int(length), mn, jobs, s2, e2)
            assert n1 == n2, (length, mn, jobs)
            assert np.array_equal(s1[:n1], s2[:n2]) and np.array_equal(e1[:n1], e2[:n2])


def test_vops_reductions_match_reference():
    T = R.ref_thread_count()
    rng = np.random.default_rng(4)
    for n in [1, 7, 8, 9, 100, 12345]:
        v = rng.standard_normal(n)
        v[rng.random(n) < 0.3] = 0.0
        w = rng.standard_normal(n)
        assert L.orc_maxabs_replica(v, n, T) == R.ref_vops_maxabs(v, n)
        assert L.orc_dot_replica(v, w, n, T) == R.ref_vops_dot(v, w, n)
    # the reference's maxAbs keeps each range's last nonzero |v|, not its max
    v = np.zeros(64)
    v[0], v[1] = 5.0, 0.5
    assert R.ref_vops_maxabs(v, 64) == 0.5


def test_bins_for_cell_window():
    """The SDF window is the union of the reference's 9 bins around the cell's bin."""
    for I, J in [(13, 10), (9, 9), (30, 31)]:
        out = np.zeros(9, np.int64)
        nbJ = (J + 2) // 3
        for i in range(I):
            for j in range(J):
                R.ref_bins_for_cell(I, J, i, j, out)
                cells = set()
                for b in out:
                    if b < 0:
                        continue
                    bi, bj = divmod(int(b), nbJ)
                    for ci in range(3 * bi, min(I, 3 * bi + 3)):
                        for cj in range(3 * bj, min(J, 3 * bj + 3)):
                            cells.add((ci, cj))
                bi, bj = i // 3, j // 3
                mine = {(ci, cj) for ci in range(max(0, 3 * bi - 3), min(I, 3 * bi + 6))
                        for cj in range(max(0, 3 * bj - 3), min(J, 3 * bj + 6))}
                assert cells == mine, (i, j)


@pytest.mark.parametrize("n,seed", [(24, 0), (32, 1), (40, 2)])
def test_pcg_replica_matches_reference(n, seed):
    """LinearSolver::solve with the reference IPP, its max-abs test and cap."""
    T = R.ref_thread_count()
    rng = np.random.default_rng(seed)
    p, st = O.dam_break(n, ppc=2)
    mat = st.material.copy()
    # splash some empty cells into the fluid body
    fl = np.where(mat == O.FLUID)[0]
    mat[rng.choice(fl, len(fl) // 10, replace=False)] = O.EMPTY
    I = J = n
    b = np.zeros(I * J)
    b[mat == O.FLUID] = rng.standard_normal(int((mat == O.FLUID).sum()))
    scale = 0.7
    nn = np.zeros(I * J, np.uint8)
    mask = np.zeros(I * J, np.uint8)
    fluid = np.zeros(I * J, np.uint8)
    L.orc_build_matrix(mat, I, J, nn, mask, fluid)
    for limit, tol in [(5, 1e-2), (20, 1e-2), (200, 1e-2), (200, 1e3)]:
        x1 = np.zeros(I * J)
        x2 = np.zeros(I * J)
        it1 = L.orc_pcg_replica(mat, I, J, scale, b, x1, limit, tol, T, None)
        it2 = R.ref_pcg(mat, fluid, nn, mask, I, J, scale, b, x2, limit, tol)
        assert it1 == it2, (limit, tol)
        assert np.array_equal(x1.view(np.uint64), x2.view(np.uint64)), (limit, tol)
