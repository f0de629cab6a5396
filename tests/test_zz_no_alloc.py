"""The allocation-free isolated solve (fg_poisson_solve_noalloc) against the
allocating one (fg_poisson_solve), through the C-ABI: two calls in a row,
the second on a smaller system (the scratch keeps the first call's
contents), each bit-identical to the allocating solve of the same system;
a system beyond the scratch is refused before any device work."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _system(orc, I, J, seed):
    rng = np.random.default_rng(seed)
    mat = np.where(rng.random((I, J)) < 0.7, orc.FLUID, orc.EMPTY).astype(np.int8)
    mat[0, :] = mat[-1, :] = mat[:, 0] = mat[:, -1] = orc.SOLID
    mat[rng.random((I, J)) < 0.02] = orc.SOLID
    b = rng.standard_normal(I * J)
    b[(mat.ravel() & 0x40) == 0] = 0.0
    return mat, b


def test_noalloc_solve_equals_allocating_solve(fg, orc):
    with fg.Context(0) as ctx:
        for (I, J, seed) in [(320, 288, 3), (96, 160, 4)]:
            _log(f"noalloc solve {I}x{J}")
            mat, b = _system(orc, I, J, seed)
            x_ref, s_ref = ctx.poisson_solve(mat, 0.37, b, tol=1e-9)
            x_na, s_na = ctx.poisson_solve(mat, 0.37, b, tol=1e-9, noalloc=True)
            _log(f"  iterations {s_ref.iterations} / {s_na.iterations}, resid {s_ref.rel_residual:.2e} / "
                 f"{s_na.rel_residual:.2e}")
            assert s_na.converged and s_na.rel_residual <= 1e-9
            assert s_na.iterations == s_ref.iterations and s_na.sweeps == s_ref.sweeps
            assert np.array_equal(x_na.view(np.uint64), x_ref.view(np.uint64))
            # and the oracle's FP64 solution
            xo = np.zeros(I * J)
            orc.lib().orc_pcg_converged(mat.ravel(), I, J, 0.37, b, xo, 100000, 1e-13, C.byref(C.c_double()))
            assert np.linalg.norm(x_na - xo) <= 1e-6 * np.linalg.norm(xo)
        # beyond the scratch: refused (host-side check, nothing launched)
        big = np.full((1024, 1024), orc.FLUID, np.int8)
        with pytest.raises(fg.FgError) as ei:
            ctx.poisson_solve(big, 0.37, np.zeros(1024 * 1024), noalloc=True)
        assert ei.value.code == fg.E_CAPACITY
