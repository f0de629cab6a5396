"""The isolated-projection workload generator against the oracle's
calcPressureRhs restatement (bit-exact), and its mask statistics."""
import ctypes as C

import numpy as np
import pytest

import flipsolver2d_amd as fg
import oracle as O


@pytest.mark.parametrize("n,seed", [(64, 1), (97, 3), (256, 5)])
def test_rhs_matches_oracle(n, seed):
    mat, u, v, b, scale = fg.isolated_projection_system(n, seed)
    L = O.lib()
    p = O.make_params(n)
    st = O.State(n, n, 1)
    st.material[:] = mat.ravel()
    st.u[:] = u.ravel()
    st.v[:] = v.ravel()
    rhs = np.zeros(n * n)
    L.orc_pressure_rhs(C.byref(p), st.ref(), rhs)
    assert np.array_equal(b.view(np.uint64), rhs.view(np.uint64))
    assert scale == pytest.approx((1.0 / 30) / (1000.0 / n / n))


def test_mask_statistics():
    n = 512
    mat, u, v, b, _ = fg.isolated_projection_system(n, 2)
    inner = mat[1:-1, 1:-1]
    liquid = ((inner & 0x40) != 0).mean()
    solid = ((inner & 0x20) != 0).mean()
    assert 0.55 < liquid < 0.62 and 0.015 < solid < 0.025
    assert np.all(mat[0, :] == O.SOLID) and np.all(mat[:, -1] == O.SOLID)
    assert u.dtype == np.float32 and np.abs(u).max() <= 1.0
