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
        "fluid_sdf": st.fluid_sdf.copy(), "solid_id": st.solid_id.copy(),
    }


def run_oracle_frames(orc, params, st, frames):
    L = orc.lib()
    stats = orc.OrcStats()
    for _ in range(frames):
        L.orc_step_frame(C.byref(params), st.ref(), C.byref(stats))
    return stats
