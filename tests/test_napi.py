"""The N-API addon (three-virtual-geometry_amd/napi): the binding the TypeScript host adds, driven from Node.

CPU: the pool a Node host builds through the addon is byte-identical to the one the Python host builds
through ctypes (same C ABI). GPU: a cut launched from Node equals the same cut launched from Python.
"""
import ctypes
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from helpers import draw_set, ground_camera, small_ruins
from vgcut import native
from vgcut.virtual_geometry import make_view

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "three-virtual-geometry_amd", "napi", "vgcut.node")
DRIVER = os.path.join(ROOT, "tests", "napi_driver.js")
NODE = shutil.which("node")

needs_addon = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON),
                                 reason="node or the built addon (napi/build.sh, run by build()) is missing")


FLAGS = {"maxDrawDistance": 1e30, "minPixelRadius": 0.0, "enabled": True, "shadow": False, "occlusion": False}


def mesh_json(data, matrices):
    return {k: np.asarray(getattr(data, k)).ravel().tolist() for k in
            ["meshletBounds", "meshletInfo", "levelRanges", "levelErrors", "levelGuardErrors", "boundingSphere",
             "lodBoundsSphere", "cullBoundsSphere"]} | {"matrices": np.asarray(matrices, dtype=np.float32).ravel().tolist(),
                                                        "flags": FLAGS}


def load_python(scene, spec):
    for d, m in zip(spec.meshes, spec.matrices):
        if len(m):
            slot = scene.add_mesh(d)
            scene.set_instances(slot, m)
            scene.set_flags(slot, FLAGS["maxDrawDistance"], FLAGS["minPixelRadius"], FLAGS["enabled"], FLAGS["shadow"],
                            FLAGS["occlusion"])


def run_node(tmp_path, spec):
    inp, out = tmp_path / "in.json", tmp_path / "out.json"
    inp.write_text(json.dumps(spec))
    r = subprocess.run([NODE, DRIVER, ADDON, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + (out.read_text() if out.exists() else "")
    return json.loads(out.read_text())


def small_spec():
    return small_ruins(seed=71, instances=600, meshes=3)


@needs_addon
def test_node_host_builds_the_same_pool(tmp_path):
    spec = small_spec()
    config = {"maxViews": 2, "meshSlots": 8, "device": -1}
    got = run_node(tmp_path, {"config": config,
                              "meshes": [mesh_json(d, m) for d, m in zip(spec.meshes, spec.matrices) if len(m)]})
    ctx = native.Context(max_views=2, mesh_slots=8, device=-1)
    scene = native.Scene(ctx)
    orders = []
    for d, m in zip(spec.meshes, spec.matrices):
        if len(m):
            slot = scene.add_mesh(d)
            orders.append(scene.set_instances(slot, m))
            scene.set_flags(slot, FLAGS["maxDrawDistance"], FLAGS["minPixelRadius"], FLAGS["enabled"], FLAGS["shadow"],
                            FLAGS["occlusion"])
    assert got["slots"] == list(range(len(orders)))
    for a, b in zip(got["orders"], orders):
        assert a == list(np.asarray(b).tolist())
    for buf in range(5):
        assert bytes.fromhex(got["pool"][buf]) == scene.pool(buf).tobytes(), f"pool buffer {buf}"


@pytest.mark.gpu
@needs_addon
def test_node_host_cut_equals_python_host(tmp_path, gpu_available):
    spec = small_spec()
    max_meshlets = 1 << 14
    views = [make_view(ground_camera((-20.0, 3.0, -25.0), (10.0, 0.0, 20.0)), 1080, slot=0),
             make_view(ground_camera((15.0, 12.0, 5.0), (-20.0, 0.0, -10.0)), 720, slot=1, error_threshold=0.5)]
    packed = b"".join(ctypes.string_at(ctypes.addressof(v), ctypes.sizeof(v)) for v in views)
    assert ctypes.sizeof(native.View) == 292
    config = {"maxViews": 2, "meshSlots": 8, "maxMeshlets": max_meshlets, "maxTriangles": 2_000_000,
              "maxShadowTriangles": 1_000_000, "device": 0}
    got = run_node(tmp_path, {"config": config, "meshes": [mesh_json(d, m) for d, m in zip(spec.meshes, spec.matrices) if len(m)],
                              "views": packed.hex(), "viewCount": len(views), "viewSlots": [0, 1], "maxEntries": max_meshlets})
    ctx = native.Context(max_triangles=2_000_000, max_shadow_triangles=1_000_000, max_meshlets=max_meshlets, max_views=2,
                         mesh_slots=8, device=0)
    scene = native.Scene(ctx)
    load_python(scene, spec)
    scene.cut(views)
    scene.sync()
    for v, res in zip(views, got["results"]):
        counters, args, draws, _ = scene.read(v.slot)
        assert bytes.fromhex(res["counters"]) == counters.tobytes(), f"slot {v.slot} counters"
        assert bytes.fromhex(res["args"]) == args.tobytes(), f"slot {v.slot} args"
        node_draws = np.frombuffer(bytes.fromhex(res["list"]), dtype=np.uint32).reshape(2, max_meshlets, 4)
        for cut in (0, 1):
            n = min(int(counters[2 + 5 * cut]), max_meshlets)
            assert np.array_equal(draw_set(node_draws[cut], n), draw_set(draws[cut], n)), f"slot {v.slot} cut {cut}"
