"""Pins for the oracle's viscosity members (the light model,
viscositymodel.cpp:4-140, with centeredParamsToGrid's viscosity).

Each check is chosen so that a plausible restatement mistake fails it:
  - orc_viscosity_matrix (63-134): the dense matrix the reference's assembly
    loop leaves -- replayed literally here, every coeffRef write in its (i, j)
    order, later writes overwriting earlier ones -- equals the oracle's
    closed form of diagonal and last-writer couplings (a swapped end, a
    dropped overwrite or a sign fails it); the result is symmetric and
    strictly diagonally dominant;
  - orc_viscosity_apply (4-61): U and V after the model equal a dense FP64
    solve of the same system with rhs rho * U (and V), divided by rho; with
    zero viscosity the matrix is the identity and the grid comes back
    unchanged; the faces at i = I (U) and j = J (V) are not touched;
  - orc_viscosity_grid (flipsolver2d.cpp:1394-1431, 396-397): one particle
    gives its viscosity times w / (w + 1e-10) in every cell its quadratic
    B-spline reaches (compiled quadraticBSpline) and 0 elsewhere; uniform
    particle viscosity gives that value wherever particles reach; a SOURCE
    cell takes its emitter's viscosity;
  - seeding (682-707): a particle's viscosity is the init viscosity grid lerped
    at its position with offset (1/2, 1/2) (compiled Grid2d::lerp); the sort
    and countParticles carry it with the particle.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O

L = O.lib()
R = O.ref_lib()
needs_ref = pytest.mark.skipif(R is None, reason="reference building blocks not built (no reference tree)")


def _literal_assembly(mat, visc, I, J, dt):
    """LightViscosityModel::getMatrix as written: coeffRef writes in loop order."""
    n = I * J
    A = np.zeros((n, n))
    scale = float(np.float64(np.float32(dt)))
    solid = lambda i, j: 0 <= i < I and 0 <= j < J and (int(mat[i, j]) & 0x20) != 0
    v = lambda i, j: float(visc[min(max(i, 0), I - 1), min(max(j, 0), J - 1)])
    for i in range(I):
        for j in range(J):
            idx = i * J + j
            if solid(i, j):
                A[idx, idx] = 1.0
                continue
            ip1 = 1.0 * ((v(i, j) if solid(i + 1, j) else v(i + 1, j)) * scale)
            jp1 = 1.0 * ((v(i, j) if solid(i, j + 1) else v(i, j + 1)) * scale)
            im1 = 1.0 * ((v(i, j) if solid(i - 1, j) else v(i - 1, j)) * scale)
            jm1 = 1.0 * ((v(i, j) if solid(i, j - 1) else v(i, j - 1)) * scale)
            diag = 4.0
            diag *= v(i, j) * scale
            diag += 1.0
            A[idx, idx] = diag
            for (a, b, val) in ((i + 1, j, ip1), (i, j + 1, jp1), (i - 1, j, im1), (i, j - 1, jm1)):
                if 0 <= a < I and 0 <= b < J:
                    A[idx, a * J + b] = val
                    A[a * J + b, idx] = val
    return A


def _dense(diag, east, north, I, J):
    n = I * J
    A = np.diag(diag.copy())
    for c in range(n):
        i, j = divmod(c, J)
        if i + 1 < I:
            A[c, c + J] = A[c + J, c] = east[c]
        if j + 1 < J:
            A[c, c + 1] = A[c + 1, c] = north[c]
    return A


def _random_case(I, J, seed):
    rng = np.random.default_rng(seed)
    mat = np.where(rng.random((I, J)) < 0.25, O.SOLID, np.where(rng.random((I, J)) < 0.5, O.FLUID, O.EMPTY))
    mat = mat.astype(np.int8)
    visc = rng.uniform(0.0, 3.0, (I, J)).astype(np.float32)
    return mat, visc


@pytest.mark.parametrize("I,J,seed", [(7, 6, 0), (9, 11, 1), (12, 5, 2)])
def test_viscosity_matrix_equals_literal_assembly(I, J, seed):
    mat, visc = _random_case(I, J, seed)
    dt = np.float32(0.0123)
    d, e, n = (np.zeros(I * J) for _ in range(3))
    L.orc_viscosity_matrix(np.ascontiguousarray(mat.ravel()), np.ascontiguousarray(visc.ravel()), I, J, dt, d, e, n)
    A = _dense(d, e, n, I, J)
    B = _literal_assembly(mat, visc, I, J, dt)
    assert np.array_equal(A, B)
    assert np.array_equal(A, A.T)
    off = np.abs(A).sum(axis=1) - np.abs(np.diag(A))
    assert np.all(np.diag(A) > off)


def _grid_state(I, J, rng, visc_value=None):
    st = O.State(I, J, 4)
    mat, visc = _random_case(I, J, int(rng.integers(1 << 30)))
    st.material[:] = mat.ravel()
    st.viscosity[:] = visc.ravel() if visc_value is None else visc_value
    st.u[:] = rng.standard_normal(st.u.size).astype(np.float32)
    st.v[:] = rng.standard_normal(st.v.size).astype(np.float32)
    return st


@pytest.mark.parametrize("I,J", [(8, 9), (13, 7)])
def test_viscosity_apply_equals_dense_solve(I, J):
    rng = np.random.default_rng(I * J)
    p = O.make_params(I, J, density=1000.0)
    st = _grid_state(I, J, rng)
    u0, v0 = st.u.copy(), st.v.copy()
    dt = np.float32(0.02)
    d, e, n = (np.zeros(I * J) for _ in range(3))
    L.orc_viscosity_matrix(st.material, st.viscosity, I, J, dt, d, e, n)
    A = _dense(d, e, n, I, J)
    it = L.orc_viscosity_apply(C.byref(p), st.ref(), dt, 1e-12)
    assert it > 0
    rho = np.float32(1000.0)
    U0, V0 = u0.reshape(I + 1, J), v0.reshape(I, J + 1)
    xu = np.linalg.solve(A, (rho * U0[:I, :J]).astype(np.float64).ravel())
    xv = np.linalg.solve(A, (rho * V0[:I, :J]).astype(np.float64).ravel())
    U, V = st.u.reshape(I + 1, J), st.v.reshape(I, J + 1)
    assert np.max(np.abs(U[:I, :J].ravel() - xu / rho)) <= 1e-6 * np.max(np.abs(xu / rho))
    assert np.max(np.abs(V[:I, :J].ravel() - xv / rho)) <= 1e-6 * np.max(np.abs(xv / rho))
    assert np.array_equal(U[I], U0[I]) and np.array_equal(V[:, J], V0[:, J])


def test_zero_viscosity_is_the_identity():
    I, J = 10, 8
    rng = np.random.default_rng(4)
    p = O.make_params(I, J, density=1000.0)
    st = _grid_state(I, J, rng, visc_value=0.0)
    u0, v0 = st.u.copy(), st.v.copy()
    L.orc_viscosity_apply(C.byref(p), st.ref(), np.float32(0.03), 1e-12)
    # x = rho u solved exactly; u = x / rho rounds back within an ulp
    assert np.max(np.abs(st.u - u0) / np.maximum(np.abs(u0), 1e-30)) <= 1.2e-7
    assert np.max(np.abs(st.v - v0) / np.maximum(np.abs(v0), 1e-30)) <= 1.2e-7


@needs_ref
def test_viscosity_grid_of_one_particle():
    I, J = 9, 10
    p = O.make_params(I, J)
    st = O.State(I, J, 4)
    st.n = 1
    st.px[0], st.py[0] = 4.3, 5.8
    st.pvisc[0] = 2.5
    cs = np.zeros(I * J + 1, np.int32)
    cs[4 * J + 5 + 1:] = 1
    L.orc_viscosity_grid(C.byref(p), st.ref(), cs)
    g = st.viscosity.reshape(I, J)
    for i in range(I):
        for j in range(J):
            w = np.float32(R.ref_quadratic_bspline(np.float32(4.3) - np.float32(i), np.float32(5.8) - np.float32(j)))
            want = np.float32(np.float32(w * np.float32(2.5)) / np.float32(np.float32(1e-10) + w)) if w > 1e-9 \
                else np.float32(0.0)
            assert g[i, j] == want, (i, j)
    assert (g > 0).sum() == 9


def test_viscosity_grid_uniform_and_source():
    I, J = 12, 12
    p = O.make_params(I, J, emitters=[(0.0, 0.0, True, 7.5)])
    st = O.State(I, J, 1000)
    rng = np.random.default_rng(9)
    pos = []
    for c in range(I * J):
        i, j = divmod(c, J)
        if 3 <= i < 9 and 3 <= j < 9:
            for q in range(4):
                pos.append((i + rng.random(), j + rng.random()))
    n = len(pos)
    st.n = n
    st.px[:n] = [q[0] for q in pos]
    st.py[:n] = [q[1] for q in pos]
    st.pvisc[:n] = 1.75
    st.pid[:n] = np.arange(n)
    st.material[:] = O.EMPTY
    st.material[5 * J + 5] = O.SOURCE
    st.emitter_id[:] = -1
    st.emitter_id[5 * J + 5] = 0
    cs = np.zeros(I * J + 1, np.int32)
    L.orc_sort_particles(C.byref(p), st.ref(), None, cs.ctypes.data)
    L.orc_viscosity_grid(C.byref(p), st.ref(), cs)
    g = st.viscosity.reshape(I, J)
    inner = g[4:9, 4:9].ravel()
    inner = np.delete(inner, 1 * 5 + 1)   # the SOURCE cell
    assert np.max(np.abs(inner - 1.75)) <= 1e-6
    assert g[5, 5] == np.float32(7.5)
    assert g[0, 0] == 0 and g[11, 11] == 0


@needs_ref
def test_seeded_viscosity_and_its_carriage():
    n = 24
    p = O.make_params(n, ppc=4, viscosity=True, fluid_viscosity=[3.0])
    walls, fluids = O.dam_break_geometry(n)
    st = O.init_scene_objects(p, solids=walls, fluids=fluids)
    N = st.n
    grid0 = np.zeros(n * n, np.float32)
    grid0[st.material == O.FLUID] = 3.0
    want = np.zeros(N, np.float32)
    R.ref_lerp(grid0, n, n, 0.5, 0.5, np.ascontiguousarray(st.px[:N]), np.ascontiguousarray(st.py[:N]), N, want)
    assert np.array_equal(st.pvisc[:N], want)
    by_id = dict(zip(st.pid[:N].tolist(), st.pvisc[:N].tolist()))
    # reverse the order, sort back: every particle keeps its viscosity
    for k in ("px", "py", "vx", "vy", "pid", "pvisc"):
        a = getattr(st, k)
        a[:N] = a[:N][::-1].copy()
    cs = np.zeros(n * n + 1, np.int32)
    L.orc_sort_particles(C.byref(p), st.ref(), None, cs.ctypes.data)
    assert all(by_id[int(i)] == float(v) for i, v in zip(st.pid[:N], st.pvisc[:N]))
    # crowd one cell beyond 2 ppc: countParticles erases by id and compacts
    st.px[:20] = 10.5
    st.py[:20] = 10.5
    L.orc_sort_particles(C.byref(p), st.ref(), None, cs.ctypes.data)
    L.orc_count_particles(C.byref(p), st.ref(), cs)
    M = st.n
    assert M == N - 12
    assert all(by_id[int(i)] == float(v) for i, v in zip(st.pid[:M], st.pvisc[:M]))
