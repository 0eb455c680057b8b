"""The C ABI library loads and exports every entry point include/*.h declares (no GPU work)."""
import ctypes
import os
import re
import subprocess

from vgcut import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(vg_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_declares_the_bound_entry_points():
    assert declared_functions() == set(native.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = native.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", native.LIBRARY], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (vg_[a-z0-9_]+)", out))
    assert declared_functions() <= exported


def test_dag_library_exports_every_declared_symbol():
    """include/vgdag.h (the DAG builder's C ABI) against lib/libvgdag.so and the Python binding."""
    from vgcut import _native, dag

    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "vgdag.h")).read(), flags=re.S)
    declared = set(re.findall(r"\b(vgd_[a-z0-9_]+)\s*\(", text))
    assert declared == set(dag.EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", _native.library_path("libvgdag.so")], capture_output=True, text=True,
                         check=True).stdout
    assert declared <= set(re.findall(r" T (vgd_[a-z0-9_]+)", out))


def test_host_only_context_and_defaults():
    ctx = native.Context(device=-1)
    assert ctx.counter_words == 12 + 8 * 64 + 8
    assert ctx.draw_args_words == 20 * 64
    scene = native.Scene(ctx)
    assert scene.device_bytes == 0
    scene.close()
    ctx.close()
    lib = native.load()
    bad = native.Config(0, 0, (1 << 24) + 1, 1, 64, 1024, 256, -1)
    h = ctypes.c_void_p()
    assert lib.vg_create(ctypes.byref(bad), ctypes.byref(h)) == native.VG_E_INVALID
