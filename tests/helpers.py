"""Shared helpers for the parity tests."""
import ctypes as C

import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def oracle_state_dict(st):
    n = st.n
    return {
        "px": st.px[:n].copy(), "py": st.py[:n].copy(), "vx": st.vx[:n].copy(), "vy": st.vy[:n].copy(),
        "pid": st.pid[:n].copy(), "u": st.u.copy(), "v": st.v.copy(), "u_valid": st.u_valid.copy(),
        "v_valid": st.v_valid.copy(), "material": st.material.copy(), "solid_sdf": st.solid_sdf.copy(),
        "fluid_sdf": st.fluid_sdf.copy(), "solid_id": st.solid_id.copy(), "emitter_id": st.emitter_id.copy(),
        "particle_counts": st.particle_counts.copy(), "particle_viscosity": st.pvisc[:n].copy(),
        "viscosity": st.viscosity.copy(),
        "rng_epoch": int(st.c.rng_epoch), "next_pid": int(st.c.next_pid),
    }


def run_oracle_frames(orc, params, st, frames):
    L = orc.lib()
    stats = orc.OrcStats()
    for _ in range(frames):
        L.orc_step_frame(C.byref(params), st.ref(), C.byref(stats))
    return stats


def oracle_state_from_gpu(orc, g, I, J, capacity=None):
    """An oracle State holding a GPU get_state() dict (the next substep's input)."""
    n = len(g["px"])
    st = orc.State(I, J, max(capacity or 0, n, 1))
    for k in ("px", "py", "vx", "vy", "pid"):
        getattr(st, k)[:n] = g[k]
    for k in ("u", "v", "u_valid", "v_valid", "material", "solid_sdf", "fluid_sdf", "solid_id", "emitter_id",
              "particle_counts"):
        getattr(st, k)[...] = g[k]
    st.n = n
    st.c.rng_epoch, st.c.next_pid = int(g["rng_epoch"]), int(g["next_pid"])
    return st


def host_rel_residual(orc, fg, g, I, J, dt, rho, dx):
    """FP64 ||b - A x||2 / ||b||2 of a GPU state's last pressure solve,
    recomputed on the host from the matrix bytes that solve used (scale
    s = dt/(rho dx^2) of that substep, dt as the float the substep ran with)."""
    n, mask, fluid = fg.stencil_matrix(g["stencil"])
    s = float(np.float32(dt)) / (rho * dx * dx)
    y = np.zeros(I * J, np.float64)
    orc.lib().orc_spmv(n, mask, fluid, I, J, s, np.ascontiguousarray(g["pressure"]), y)
    b = g["rhs"]
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(b - y) / nb) if nb > 0 else 0.0


def frame_dt(orc, vel, params):
    """The first substep's dt of a frame (FlipSolver::stepFrame, flipsolver2d.cpp:476-490)."""
    fin = C.c_int32()
    return np.float32(orc.lib().orc_frame_dt(C.c_float(vel), C.c_float(params.cfl_number),
                                             C.c_float(1.0 / params.fps), C.c_float(0.0), 0,
                                             params.max_substeps, C.byref(fin)))


def assert_substep_parity(g, st, log, *, p_tol=1e-5, v_tol=1e-5, what=""):
    """The bar of one substep from identical input state: integer/index work
    and every float stage upstream of the solve bit-exact; pressure and
    velocities (grid and particles) within relative L2 tolerances."""
    N = st.n
    assert len(g["px"]) == N, (what, len(g["px"]), N)
    assert np.array_equal(g["pid"], st.pid[:N]), what
    assert np.array_equal(g["px"].view(np.uint32), st.px[:N].view(np.uint32)), what
    assert np.array_equal(g["py"].view(np.uint32), st.py[:N].view(np.uint32)), what
    assert np.array_equal(g["material"], st.material), what
    assert np.array_equal(g["particle_counts"], st.particle_counts), what
    assert np.array_equal(g["fluid_sdf"].view(np.uint32), st.fluid_sdf.view(np.uint32)), what
    assert np.array_equal(g["saved_u"].view(np.uint32), st.saved_u.view(np.uint32)), what
    assert np.array_equal(g["saved_v"].view(np.uint32), st.saved_v.view(np.uint32)), what
    assert np.array_equal(g["rhs"].view(np.uint64), st.rhs.view(np.uint64)), what
    assert np.array_equal(g["u_valid"], st.u_valid), what
    assert np.array_equal(g["v_valid"], st.v_valid), what
    e = {"p": rel_l2(g["pressure"], st.pressure), "u": rel_l2(g["u"], st.u), "v": rel_l2(g["v"], st.v),
         "vx": rel_l2(g["vx"], st.vx[:N]), "vy": rel_l2(g["vy"], st.vy[:N])}
    log(f"  {what} rel L2: " + " ".join(f"{k}={v:.2e}" for k, v in e.items()))
    assert e["p"] <= p_tol, (what, e)
    assert max(e["u"], e["v"], e["vx"], e["vy"]) <= v_tol, (what, e)
    return e
