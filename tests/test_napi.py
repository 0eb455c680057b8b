"""The N-API addon (three-virtual-geometry_amd/napi): the binding the TypeScript host adds, driven from Node.

CPU: the pool a Node host builds through the addon is byte-identical to the one the Python host builds
through ctypes (same C ABI), and so is a DAG the addon's buildVirtualMesh makes from raw triangles.
GPU: a cut launched from Node equals the same cut launched from Python.
"""
import ctypes
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from helpers import draw_set, ground_camera, small_ruins
from vgcut import dag, native, scenegen, scenes
from vgcut.virtual_geometry import make_view

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "three-virtual-geometry_amd", "napi", "vgcut.node")
DRIVER = os.path.join(ROOT, "tests", "napi_driver.js")
NODE = shutil.which("node")

needs_addon = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON),
                                 reason="node or the built addon (napi/build.sh, run by build()) is missing")


FLAGS = {"maxDrawDistance": 1e30, "minPixelRadius": 0.0, "enabled": True, "shadow": False, "occlusion": False}


def mesh_json(data, matrices, occlusion=False):
    out = {k: np.asarray(getattr(data, k)).ravel().tolist() for k in
           ["meshletBounds", "meshletInfo", "levelRanges", "levelErrors", "levelGuardErrors", "boundingSphere",
            "lodBoundsSphere", "cullBoundsSphere"]} | {"matrices": np.asarray(matrices, dtype=np.float32).ravel().tolist(),
                                                       "flags": FLAGS | {"occlusion": occlusion}}
    for k, dtype in (("meshletTriangles", np.uint32), ("positions", np.float32), ("meshletVertices", np.uint32)):
        if getattr(data, k, None) is not None:
            out[k] = np.ascontiguousarray(getattr(data, k), dtype=dtype).tobytes().hex()
    return out


def load_python(scene, spec, occlusion=False):
    for d, m in zip(spec.meshes, spec.matrices):
        if len(m):
            slot = scene.add_mesh(d)
            scene.set_instances(slot, m)
            scene.set_flags(slot, FLAGS["maxDrawDistance"], FLAGS["minPixelRadius"], FLAGS["enabled"], FLAGS["shadow"],
                            occlusion)


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
    for buf in range(8):
        assert bytes.fromhex(got["pool"][buf]) == scene.pool(buf).tobytes(), f"pool buffer {buf}"


@needs_addon
def test_node_builds_foliage_with_attributes(tmp_path):
    """buildVirtualMesh(src, options) through the addon with normals and uvs on the leaf cards (aggregate and voxel
    levels, added vertices) equals the Python binding's build bit for bit, attributes and stats included."""
    import foliage

    pos, uv, idx, nrm = foliage.leaf_cards(fronds=6)
    mesh = {"source": {"positions": pos.tobytes().hex(), "indices": idx.tobytes().hex(), "normals": nrm.tobytes().hex(),
                       "uvs": uv.tobytes().hex(), "options": {}},
            "matrices": np.eye(4, dtype=np.float32).ravel().tolist(), "flags": FLAGS}
    got = run_node(tmp_path, {"config": {"maxViews": 2, "meshSlots": 8, "device": -1}, "meshes": [mesh]})
    d = dag.build_virtual_mesh(pos, idx, normals=nrm, uvs=uv)
    b = got["built"][0]
    assert d.aggregateGroups > 0 and d.voxelLevels > 0
    assert b["stats"]["aggregateGroups"] == d.aggregateGroups and b["stats"]["voxelLevels"] == d.voxelLevels
    assert b["colors"] is None and d.colors is None
    for f in ("meshletBounds", "meshletInfo", "meshletVertices", "meshletTriangles", "indices", "positions", "normals", "uvs"):
        assert bytes.fromhex(b[f]) == np.ascontiguousarray(getattr(d, f)).tobytes(), f


@needs_addon
def test_node_host_builds_the_dag(tmp_path):
    """buildVirtualMesh through the addon (buildVirtualMesh.ts:185) equals the Python binding's build bit for bit,
    and the pool a Node host fills from it equals the Python host's."""
    pos, idx = scenegen.triangle_mesh(scenegen.COLUMN, seed=5, size=(0.5, 3.0, 0), noise_amplitude=0.05,
                                      noise_frequency=1.5, target_triangles=30_000)
    rng = np.random.default_rng(3)
    n = 40
    mats = scenes.yaw_scale_translate(rng.uniform(0, 6.28, n), rng.uniform(0.5, 2.0, n),
                                      np.stack([rng.uniform(-30, 30, n), np.zeros(n), rng.uniform(-30, 30, n)], 1))
    options = {"groupSize": 10, "minRootTriangles": 8}
    mesh = {"source": {"positions": pos.tobytes().hex(), "indices": idx.tobytes().hex(), "options": options},
            "matrices": mats.ravel().tolist(), "flags": FLAGS}
    got = run_node(tmp_path, {"config": {"maxViews": 2, "meshSlots": 8, "device": -1}, "meshes": [mesh]})
    d = dag.build_virtual_mesh(pos, idx, **options)
    b = got["built"][0]
    assert b["meshletCount"] == d.meshletCount
    assert b["stats"]["lodLevels"] == d.lodLevels and b["stats"]["rootMeshlets"] == d.rootMeshlets
    assert b["stats"]["leafTriangles"] == d.leafTriangles == len(idx) // 3
    for f in ("meshletBounds", "meshletInfo", "meshletVertices", "meshletTriangles", "levelRanges", "levelErrors",
              "levelGuardErrors", "replacementGroup", "replacementStart", "replacementIndices", "indices", "positions",
              "boundingSphere", "lodBoundsSphere", "cullBoundsSphere"):
        assert bytes.fromhex(b[f]) == np.ascontiguousarray(getattr(d, f)).tobytes(), f
    ctx = native.Context(max_views=2, mesh_slots=8, device=-1)
    scene = native.Scene(ctx)
    slot = scene.add_mesh(d)
    scene.set_instances(slot, mats)
    scene.set_flags(slot, FLAGS["maxDrawDistance"], FLAGS["minPixelRadius"], FLAGS["enabled"], FLAGS["shadow"],
                    FLAGS["occlusion"])
    for buf in range(8):
        assert bytes.fromhex(got["pool"][buf]) == scene.pool(buf).tobytes(), f"pool buffer {buf}"


def camera_json(eye, target, height, slot, fov=60.0, aspect=16 / 9, near=0.1, far=2000.0, **levers):
    return dict(kind="perspective", fov=fov, aspect=aspect, near=near, far=far, eye=list(eye), target=list(target), height=height,
                slot=slot, **levers)


def python_view(c):
    """The vg_view the Python host makes for a camera_json description."""
    from vgcut import camera as cam

    if c["kind"] == "orthographic":
        camera = cam.orthographic_camera(c["left"], c["right"], c["top"], c["bottom"], c["near"], c["far"])
    else:
        camera = cam.perspective_camera(c["fov"], c["aspect"], c["near"], c["far"])
    camera.look_at(c["eye"], c["target"], c.get("up", (0.0, 1.0, 0.0)))
    kw = {k: c[k] for k in ("errorThreshold", "occlusion", "minPixelRadius", "lodBlend", "maxErrorThreshold", "triangleBudget",
                            "shadowErrorScale") if k in c}
    names = dict(errorThreshold="error_threshold", minPixelRadius="min_pixel_radius", lodBlend="lod_blend",
                 maxErrorThreshold="max_error_threshold", triangleBudget="triangle_budget", shadowErrorScale="shadow_error_scale")
    kw = {names.get(k, k): v for k, v in kw.items()}
    if "thresholdRange" in c:
        kw["threshold_range"] = tuple(c["thresholdRange"])
    if "shadowSweep" in c:
        kw["shadow_sweep"] = tuple(c["shadowSweep"])
    if c.get("frustumCulling") is False:
        kw["frustum_culling"] = False
    return make_view(camera, c["height"], slot=c["slot"], **kw)


@pytest.mark.gpu
@needs_addon
def test_node_host_cut_equals_python_host(tmp_path, gpu_available):
    spec = small_spec()
    max_meshlets = 1 << 14
    cams = [camera_json((-20.0, 3.0, -25.0), (10.0, 0.0, 20.0), 1080, 0),
            camera_json((15.0, 12.0, 5.0), (-20.0, 0.0, -10.0), 720, 1, errorThreshold=0.5)]
    views = [python_view(c) for c in cams]
    assert ctypes.sizeof(native.View) == 292
    config = {"maxViews": 2, "meshSlots": 8, "maxMeshlets": max_meshlets, "maxTriangles": 2_000_000,
              "maxShadowTriangles": 1_000_000, "device": 0}
    got = run_node(tmp_path, {"config": config, "meshes": [mesh_json(d, m) for d, m in zip(spec.meshes, spec.matrices) if len(m)],
                              "cameras": cams, "viewSlots": [0, 1], "maxEntries": max_meshlets})
    packed = b"".join(ctypes.string_at(ctypes.addressof(v), ctypes.sizeof(v)) for v in views)
    assert bytes.fromhex(got["views"]) == packed, "the JS host packs other view records than the Python host"
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


@pytest.mark.gpu
@needs_addon
def test_node_host_two_phase_equals_python_host(tmp_path, gpu_available):
    """Two-phase occlusion driven from Node (setHzb, cut, setHzb, cutPhase2, readPhase2) equals the same
    frame driven from Python."""
    from vgcut import camera as cam
    from vgcut import scenegen

    spec = small_ruins(seed=72, instances=2000, meshes=3)
    max_meshlets = 1 << 15
    described = [camera_json((0.0, 2.0, -70.0), (0.0, 0.0, 60.0), 540, 0, occlusion=2),
                 camera_json((50.0, 2.0, 10.0), (-60.0, 2.0, 0.0), 540, 1, occlusion=2)]
    cams = [ground_camera(c["eye"], c["target"]) for c in described]
    views = [python_view(c) for c in described]
    depths = {}
    for phase, seed in ((1, 80), (2, 90)):
        for i, c in enumerate(cams):
            vp = cam.mat_mul(c.projectionMatrix, c.matrixWorldInverse)
            depths[(phase, i)] = scenegen.heightfield_depth(vp, 480, 270, (0.0, 0.0), 120.0, 128, 6.0, 1 / 25.0, seed + i)
    hzb = lambda phase: [{"slot": i, "width": 480, "height": 270, "depth": depths[(phase, i)].astype(np.float32).tobytes().hex()}
                         for i in range(len(cams))]
    config = {"maxViews": 2, "meshSlots": 8, "maxMeshlets": max_meshlets, "maxTriangles": 20_000_000,
              "maxShadowTriangles": 10_000_000, "device": 0}
    got = run_node(tmp_path, {"config": config,
                              "meshes": [mesh_json(d, m, occlusion=True) for d, m in zip(spec.meshes, spec.matrices) if len(m)],
                              "cameras": described, "viewSlots": [0, 1], "maxEntries": max_meshlets,
                              "hzb1": hzb(1), "hzb2": hzb(2)})
    assert "error" not in got, got.get("error")
    ctx = native.Context(max_triangles=20_000_000, max_shadow_triangles=10_000_000, max_meshlets=max_meshlets, max_views=2,
                         mesh_slots=8, device=0)
    scene = native.Scene(ctx)
    load_python(scene, spec, occlusion=True)
    for i in range(len(cams)):
        scene.set_hzb(i, depths[(1, i)])
    scene.cut(views)
    scene.sync()
    first = [scene.read(v.slot) for v in views]
    for i in range(len(cams)):
        scene.set_hzb(i, depths[(2, i)])
    scene.cut_phase2(views)
    scene.sync()
    added = 0
    for v, f, res, p2, after in zip(views, first, got["results"], got["phase2"], got["after"]):
        assert bytes.fromhex(res["counters"]) == f[0].tobytes(), f"slot {v.slot} first-phase counters"
        c2, a2 = scene.read_phase2(v.slot)
        assert bytes.fromhex(p2["counters"]) == c2.tobytes(), f"slot {v.slot} second-phase counters"
        assert bytes.fromhex(p2["args"]) == a2.tobytes(), f"slot {v.slot} second-phase args"
        draws = scene.read(v.slot)[2]
        node_draws = np.frombuffer(bytes.fromhex(after["list"]), dtype=np.uint32).reshape(2, max_meshlets, 4)
        n1, n2 = int(f[0][2]), min(int(c2[2]), max_meshlets)
        assert np.array_equal(draw_set(node_draws[0][n1:], n2 - n1), draw_set(draws[0][n1:], n2 - n1)), f"slot {v.slot} second phase"
        added += n2 - n1
    assert added > 0


GUARDS_JS = r"""
'use strict';
const vg = require(process.argv[2]);
const out = {};
const attempt = (name, fn) => {
  try { fn(); out[name] = 'ok'; } catch (e) { out[name] = (e.constructor && e.constructor.name) + ':' + (e.code || '') + ':' + e.message; }
};
const ctx = vg.createContext({ maxViews: 1, meshSlots: 4, device: -1 });
const scene = vg.createScene(ctx);
const mesh = {
  meshletBounds: new Float32Array(16 * 2 - 1), meshletInfo: new Uint32Array([1, 0, 3, 0, 1, 3, 3, 0]),
  levelRanges: new Uint32Array([0, 2]), levelErrors: new Float32Array([0, 1]), levelGuardErrors: new Float32Array([0, 1]),
  boundingSphere: new Float32Array([0, 0, 0, 1]), lodBoundsSphere: new Float32Array([0, 0, 0, 1]),
  cullBoundsSphere: new Float32Array([0, 0, 0, 1]),
};
attempt('shortBounds', () => vg.addMesh(scene, mesh));
attempt('shortErrors', () => vg.addMesh(scene, Object.assign({}, mesh, { meshletBounds: new Float32Array(32), levelErrors: new Float32Array(1) })));
attempt('boundsAsFloat64', () => vg.addMesh(scene, Object.assign({}, mesh, { meshletBounds: new Float64Array(32) })));
const slot = vg.addMesh(scene, Object.assign({}, mesh, { meshletBounds: new Float32Array(32) }));
attempt('matricesAsUint8', () => vg.setInstances(scene, slot, new Uint8Array(64)));
attempt('matricesRagged', () => vg.setInstances(scene, slot, new Float32Array(17)));
vg.setInstances(scene, slot, new Float32Array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]));
const words = vg.counterWords(ctx), argWords = vg.drawArgsWords(ctx);
attempt('countersAsUint8', () => vg.read(scene, 0, new Uint8Array(words * 4), new Uint32Array(argWords), new Uint32Array(8)));
attempt('argsAsInt16', () => vg.read(scene, 0, new Uint32Array(words), new Int16Array(argWords * 2), new Uint32Array(8)));
attempt('viewProjAsFloat64', () => vg.hzbFromCut(scene, 0, new Float64Array(16), 8, 8));
attempt('rastersNotArray', () => vg.hzbFromCuts(scene, { slot: 0 }));
attempt('rasterViewProjShort', () => vg.hzbFromCuts(scene, [{ slot: 0, viewProj: new Float32Array(15), width: 8, height: 8 }]));
attempt('rastersOnHostOnly', () => vg.hzbFromCuts(scene, [{ slot: 0, viewProj: new Float32Array(16), width: 8, height: 8 }]));
// A cut runs on a worker thread: until its promise settles every call on the context throws.
const promise = vg.cut(scene, new ArrayBuffer(292), 1);
attempt('readWhileCutting', () => vg.read(scene, 0, new Uint32Array(words), new Uint32Array(argWords), new Uint32Array(8)));
attempt('setFlagsWhileCutting', () => vg.setFlags(scene, slot, { enabled: true }));
attempt('secondCutWhileCutting', () => vg.cut(scene, new ArrayBuffer(292), 1));
attempt('sceneWhileCutting', () => vg.createScene(ctx));
promise.then(() => { out.cut = 'resolved'; }, (e) => { out.cut = 'rejected:' + e.message; }).then(() => {
  attempt('setFlagsAfter', () => vg.setFlags(scene, slot, { enabled: true }));
  console.log(JSON.stringify(out));
});
"""


@needs_addon
def test_addon_checks_arrays_and_pending_cuts(tmp_path):
    """Typed arrays of another element type are refused with a TypeError (not read or written by element
    counts of the wrong size), short arrays with a RangeError before the library reads them, and every call
    on a context throws while a cut of it runs on a worker thread; calls work again once it settled. The
    context is host-only, so the cut itself rejects (no device) after the guard was tested."""
    script = tmp_path / "guards.js"
    script.write_text(GUARDS_JS)
    r = subprocess.run([NODE, str(script), ADDON], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["shortBounds"].startswith("RangeError"), out
    assert out["shortErrors"].startswith("RangeError"), out
    for k in ("boundsAsFloat64", "matricesAsUint8", "countersAsUint8", "argsAsInt16", "viewProjAsFloat64", "rastersNotArray",
              "rasterViewProjShort"):
        assert out[k].startswith("TypeError"), (k, out[k])
    assert "host-only" in out["rastersOnHostOnly"], out
    assert out["matricesRagged"].startswith("RangeError"), out
    for k in ("readWhileCutting", "setFlagsWhileCutting", "secondCutWhileCutting", "sceneWhileCutting"):
        assert "VG_BUSY" in out[k], (k, out[k])
    assert out["cut"].startswith("rejected:") and "host-only" in out["cut"], out
    assert out["setFlagsAfter"] == "ok", out
