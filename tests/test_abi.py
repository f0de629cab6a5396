"""The C-ABI libraries load and export every symbol their headers declare.

No compute call is made: these run without a GPU.
"""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header, prefix):
    text = open(os.path.join(ROOT, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z0-9_]+)\s*\(", text)))


def test_product_library_exports_header():
    import flipsolver2d_amd as fg
    lib = fg.lib()
    names = declared("include/flipgpu.h", "fg_")
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(fg.EXPORTS) | {"fg_abi_version", "fg_last_error"}


def test_abi_version_without_device():
    import flipsolver2d_amd as fg
    lib = fg.lib()
    lib.fg_abi_version.restype = C.c_int32
    hdr = open(os.path.join(ROOT, "include/flipgpu.h")).read()
    ver = int(re.search(r"#define FG_ABI_VERSION (\d+)", hdr).group(1))
    assert lib.fg_abi_version() == ver


def test_invalid_arguments_are_rejected_without_device():
    """Argument validation runs before any HIP call."""
    import flipsolver2d_amd as fg
    lib = fg.lib()
    out = C.c_uint64()
    p = fg.make_params(1, 64)
    assert lib.fg_scene_bytes(C.byref(p), C.byref(out)) == -1
    p = fg.make_params(64, 64, max_particles=64 * 64 * 8)
    assert lib.fg_scene_bytes(C.byref(p), C.byref(out)) == fg.FG_OK
    assert out.value > 64 * 64 * 8 * 20


def test_oracle_library_exports_header():
    import oracle as O
    lib = O.lib()
    for n in declared("oracle/flip_oracle.h", "orc_"):
        assert hasattr(lib, n), n


def test_product_does_not_link_the_oracle():
    import flipsolver2d_amd as fg
    so = fg.LIB_PATH
    data = open(so, "rb").read()
    assert b"liboracle" not in data and b"orc_substep" not in data
    src = os.path.join(ROOT, "flipsolver2d_amd")
    for dirpath, _, files in os.walk(src):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "flip_oracle" not in text, f


def test_cpp_api_builds_and_links():
    """The reference-named C++ API library and its demo program are built and
    resolve their libraries (the demo's usage path makes no HIP call)."""
    import subprocess

    from flipsolver2d_amd import build as b

    assert os.path.exists(b.FACADE) and os.path.exists(b.DEMO)
    r = subprocess.run([b.DEMO], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
