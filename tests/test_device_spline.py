"""The quadratic B-spline as the kernels evaluate it (fg_device.h: bspline),
restated here in float32 numpy step by step, against the oracle's bspline
(the reference's summed-pieces form, mathfuncs.cpp:32-43) and, where the
reference tree was built, the reference's own: bit for bit on dense windows
around the breakpoints 0.5 and 1.5, on random bit patterns of every
magnitude, and on the special values (finite ones: for NaN and beyond
2^64 the reference's form gives NaN and the select +0).  The device form
selects one piece and writes the outer one as (0.5 b) b with
b = max(1.5 - |v|, 0); the GPU parity tests check the kernels themselves
against the oracle."""
import numpy as np

import oracle as O

f32 = np.float32


def device_form(v):
    v = np.asarray(v, f32)
    with np.errstate(over="ignore", invalid="ignore"):
        a = np.abs(v)
        b = np.fmax(f32(1.5) - a, f32(0.0))          # fmaxf: NaN -> the other operand
        t1 = f32(0.75) - a * a
        t2 = (f32(0.5) * b) * b
    return np.where(a < f32(0.5), t1, t2).astype(f32)


def _sample():
    rng = np.random.default_rng(5)
    near = []
    for c in (0.5, 1.5, 1.0, 0.0):
        u = np.arange(-(1 << 15), 1 << 15, dtype=np.int64)
        base = np.int64(np.float32(c).view(np.int32))
        w = (base + u)
        w = w[(w >= 0) & (w < 0x7F800000)].astype(np.uint32).view(f32)
        near.append(w)
    rand = rng.integers(0, 1 << 32, 200_000, dtype=np.uint64).astype(np.uint32).view(f32)
    small = rng.uniform(-2, 2, 100_000).astype(f32)
    specials = f32([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45, 1.17549435e-38, 3.4028235e38, 1.5, -1.5,
                    0.5, -0.5, 2.0**64, 2.0**63])
    xs = np.concatenate(near + [rand, small, specials])
    return np.concatenate([xs, -xs])


def test_device_spline_equals_reference_form():
    xs = _sample()
    got = device_form(xs).view(np.uint32)
    L = O.lib()
    R = O.ref_lib()
    for k, x in enumerate(xs):
        want = np.float32(L.orc_bspline(x))
        # beyond 2^64 value^2 overflows and the reference's form gives NaN,
        # as it does for a NaN argument; the select gives +0 there.  Positions
        # reaching the kernels are finite (advection kills the others).
        if not np.isfinite(x) or abs(float(x)) >= 2.0**64:
            continue
        assert got[k] == want.view(np.uint32), (x, got[k], want)
        if R is not None and k % 7 == 0:
            assert np.float32(R.ref_bspline(x)).view(np.uint32) == want.view(np.uint32), x
