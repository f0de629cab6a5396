"""The reference-named C++ API (flipsolver2d_amd/cpp) end to end.

facade_demo drives it the way the reference's callers do:
  - dam: FlipSolver / Obstacle / Emitter, stepFrame(), the accessors
    (markerParticles() with its 3x3-cell bins, materialGrid(),
    fluidVelocityGrid()); the result must equal the C-ABI path bit for bit;
  - hooks: a subclass that overrides a stage hook, so frames run the
    reference's loop (CFL dt from maxParticleVelocity, step() and its hooks
    one substep at a time); the same bits again;
  - solve: FlipSolver::project's pattern (flipsolver2d.cpp:93-126): the
    matrix as the reference's IndexedPressureParameters unit list, assembled
    from materials, LinearSolver::solve with its iteration limit and
    tolerance (max-abs as the reference reads it, or relative L2), the
    failure test `iters >= limit`; against the FP64 oracle;
  - json: JsonSceneReader::loadJson on one of the reference's scene files
    (sources and sinks included), stepped through the API, against the
    C-ABI path built from the same scene.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

from flipsolver2d_amd import build as fgbuild

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(args):
    r = subprocess.run([fgbuild.DEMO] + args, capture_output=True, text=True, timeout=200)
    _log(r.stdout[-2000:] + r.stderr[-2000:])
    assert r.returncode == 0, r.stderr
    return r.stdout


def _read_state(path, I, J):
    raw = open(path, "rb").read()
    off = 0

    def take(dt, n):
        nonlocal off
        a = np.frombuffer(raw, dt, n, off)
        off += a.nbytes
        return a

    n = int(take(np.int64, 1)[0])
    got = {"px": take(np.float32, n), "py": take(np.float32, n), "vx": take(np.float32, n),
           "vy": take(np.float32, n), "pid": take(np.uint32, n), "material": take(np.int8, I * J),
           "u": take(np.float32, (I + 1) * J), "v": take(np.float32, I * (J + 1))}
    got["substeps"] = int(take(np.int32, 1)[0])
    return got


@pytest.mark.parametrize("mode", ["dam", "hooks"])
def test_facade_stepframe_equals_c_abi(fg, ctx, tmp_path, mode):
    res, ppc, frames = 64, 4, 3
    out = str(tmp_path / "dam.bin")
    _log(f"facade {mode} {res} {ppc} {frames}")
    _run([mode, str(res), str(ppc), str(frames), out])
    got = _read_state(out, res, res)
    sc = fg.dam_break_scene(ctx, res, ppc=ppc)
    for _ in range(frames):
        st = sc.step_frame()
    ref = sc.get_state()
    sc.close()
    assert got["substeps"] == st.substeps
    for k in ("px", "py", "vx", "vy", "pid", "material", "u", "v"):
        assert np.array_equal(got[k].view(np.uint8), ref[k].view(np.uint8)), k


@pytest.mark.parametrize("maxabs", [0, 1])
def test_facade_linear_solver_vs_oracle(fg, orc, ctx, tmp_path, maxabs):
    n = 96
    p, st = orc.dam_break(n, ppc=2)
    mat = st.material.copy()
    rng = np.random.default_rng(3)
    b = np.zeros(n * n)
    fl = (mat & 0x40) != 0
    b[fl] = rng.standard_normal(int(fl.sum()))
    scale = 0.37
    tol = 1e-8 if maxabs else 1e-10      # max-abs: every |b - Ax|_k; relative: ||b - Ax|| / ||b||
    inp = str(tmp_path / "sys.bin")
    with open(inp, "wb") as f:
        f.write(np.int32([n, n]).tobytes())
        f.write(np.float64([scale, tol]).tobytes())
        f.write(np.int32([4000]).tobytes())
        f.write(mat.astype(np.int8).tobytes())
        f.write(b.astype(np.float64).tobytes())
    out = str(tmp_path / "x.bin")
    _run(["solve", inp, out, str(maxabs)])
    raw = open(out, "rb").read()
    ret, it = np.frombuffer(raw, np.int32, 2, 0)
    resid = float(np.frombuffer(raw, np.float64, 1, 8)[0])
    x = np.frombuffer(raw, np.float64, n * n, 16)
    xo = np.zeros(n * n)
    orc.lib().orc_pcg_converged(mat, n, n, scale, b, xo, 100000, 1e-13, C.byref(C.c_double()))
    # the true residual of the returned x, on the oracle's operator
    nn, mask, fluid = (np.zeros(n * n, np.uint8) for _ in range(3))
    orc.lib().orc_build_matrix(mat, n, n, nn, mask, fluid)
    ax = np.zeros(n * n)
    orc.lib().orc_spmv(nn, mask, fluid, n, n, scale, np.ascontiguousarray(x), ax)
    r = b - ax
    _log(f"  returned {ret} ({it} iterations), max |r| {np.abs(r).max():.2e}, rel {resid:.2e}")
    assert 0 < ret < 4000 and it > 0
    if maxabs:
        assert np.abs(r).max() <= tol
    else:
        assert resid <= tol and np.linalg.norm(r) <= 1.01 * tol * np.linalg.norm(b)
    assert np.linalg.norm(x - xo) <= 1e-5 * np.linalg.norm(xo)


@pytest.mark.parametrize("name,frames", [("waterfall.json", 3), ("viscosity_test.json", 3),
                                         ("double_dam_break.json", 2)])
def test_facade_json_scene_equals_c_abi(fg, orc, ctx, tmp_path, name, frames):
    """A scene file of the reference through JsonSceneReader::loadJson and the
    C++ API, against the same objects stepped through the C-ABI:
    waterfall.json (128 x 85 cells, nine obstacles, a source and a sink; an
    "nbflip" scene: the loader builds NBFlipSolver), viscosity_test.json
    (NBFlipSolver with viscosity) and double_dam_break.json (FlipSolver with
    viscosity)."""
    scene = os.path.join(HERE, "golden", "scenes", name)
    out = str(tmp_path / "scene.bin")
    _run(["json", scene, str(frames), out])
    info = {}
    lines = subprocess.run([fgbuild.DEMO, "jsoninfo", scene], capture_output=True, text=True,
                           check=True).stdout.splitlines()
    polys = {"solid": [], "source": [], "sink": [], "fluid": []}
    fric, emit, svisc, fvisc = [], [], [], []
    for line in lines:
        k, *v = line.split()
        if k in polys:
            polys[k].append(np.float32(v).reshape(-1, 2))
        elif k == "friction":
            fric.append(float(v[0]))
        elif k == "source_velocity":
            emit.append((float(v[0]), float(v[1]), bool(int(v[2]))))
        elif k == "source_viscosity":
            svisc.append(float(v[0]))
        elif k == "fluid_viscosity":
            fvisc.append(float(v[0]))
        else:
            info[k] = float(v[0]) if len(v) == 1 else v
    emit = [e + (sv,) for e, sv in zip(emit, svisc)]
    I, J = int(info["gridSizeI"]), int(info["gridSizeJ"])
    got = _read_state(out, I, J)
    g = info["globalAcceleration"]
    visc = bool(int(info["viscosityEnabled"]))
    p = fg.make_params(I, J, dx=info["dx"], ppc=int(info["particlesPerCell"]), density=info["fluidDensity"],
                       gravity=(float(g[0]), float(g[1])), pic=info["picRatio"], cfl=info["cflNumber"],
                       particle_scale=info["particleScale"], fps=int(info["fps"]),
                       max_substeps=int(info["maxSubsteps"]), pcg_iter_limit=0,
                       sim=fg.SIM_NBFLIP if int(info["simulationMethod"]) == 3 else fg.SIM_FLIP, viscosity=visc)
    if name != "double_dam_break.json":
        assert p.sim_method == fg.SIM_NBFLIP
    sc = fg.Scene(ctx, p)
    try:
        sc.init_objects(polys["solid"], polys["source"], polys["sink"], polys["fluid"], friction=fric, emitters=emit,
                        seed=int(info["seed"]), fluid_viscosity=fvisc)
        for _ in range(frames):
            st = sc.step_frame()
        ref = sc.get_state()
    finally:
        sc.close()
    _log(f"  {name} {I}x{J}: {len(ref['px'])} particles after {frames} frames, "
         f"{st.particles_reseeded} reseeded in the last, viscosity iters {st.viscosity_iters}")
    assert len(ref["px"]) > 0
    if visc:
        assert st.viscosity_iters > 0
    assert got["substeps"] == st.substeps
    for k in ("px", "py", "vx", "vy", "pid", "material", "u", "v"):
        assert np.array_equal(got[k].view(np.uint8), ref[k].view(np.uint8)), k