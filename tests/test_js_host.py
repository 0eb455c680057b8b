"""The Node host (three-virtual-geometry_amd/js/virtual_geometry.js): the reference's VirtualGeometry surface
in JavaScript over the N-API addon.

CPU: the vg_view records it packs equal the Python host's byte for byte (same three.js formulas in the
same operation order) for perspective and orthographic cameras, moved and turned, with every lever;
readStats (VirtualGeometry.ts:591-632) gives hand-computed stats for given counters; the budget glide
(VirtualGeometry.ts:549-588) follows a hand-computed trajectory over a scripted stats sequence, in JS and
in the Python host alike. GPU: update(camera) from Node cuts exactly what a Python vg_cut of the same view
cuts, and a second update of an unchanged camera is skipped (cutInputsChanged).
"""
import ctypes
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from helpers import draw_set, small_ruins
from test_napi import camera_json, mesh_json, python_view
from vgcut import native
from vgcut.virtual_geometry import VirtualGeometry as PyVirtualGeometry
from vgcut.virtual_geometry import VirtualGeometryStats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "three-virtual-geometry_amd", "js", "virtual_geometry.js")
ADDON = os.path.join(ROOT, "three-virtual-geometry_amd", "napi", "vgcut.node")
NODE = shutil.which("node")
needs_node = pytest.mark.skipif(NODE is None, reason="node is missing")
needs_addon = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON), reason="node or the built addon is missing")


def node(script: str, payload, tmp_path):
    src = tmp_path / "script.js"
    src.write_text(script)
    inp = tmp_path / "in.json"
    inp.write_text(json.dumps(payload))
    r = subprocess.run([NODE, str(src), HOST, str(inp), ADDON], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


PACK_JS = r"""
'use strict';
const host = require(process.argv[2]);
const cams = JSON.parse(require('fs').readFileSync(process.argv[3], 'utf8'));
const cameraOf = (c) => (c.kind === 'orthographic' ? host.orthographicCamera(c.left, c.right, c.top, c.bottom, c.near, c.far)
  : host.perspectiveCamera(c.fov, c.aspect, c.near, c.far)).lookAt(c.eye, c.target, c.up);
const buf = host.packViews(cams.map((c) => host.viewRecord(cameraOf(c), c.height, c)));
console.log(JSON.stringify({ hex: Buffer.from(buf).toString('hex') }));
"""


def cameras():
    rng = np.random.default_rng(17)
    out = [camera_json((0.0, 0.0, 0.0), (0.0, 0.0, -1.0), 1080, 0),
           camera_json((0.0, 30.0, 0.0), (0.0, 0.0, 0.0), 720, 1),  # looking straight down: three.js's nudge
           dict(kind="orthographic", left=-60.0, right=60.0, top=60.0, bottom=-60.0, near=0.1, far=500.0, eye=[30.0, 200.0, 20.0],
                target=[0.0, 0.0, 0.0], height=2048, slot=2),
           dict(kind="orthographic", left=-13.0, right=41.0, top=7.5, bottom=-30.25, near=0.5, far=900.0, eye=[-3.0, 50.0, 7.0],
                target=[10.0, 0.0, 2.0], up=[0.0, 0.0, 1.0], height=1024, slot=3, minPixelRadius=0.0)]
    for k in range(12):
        eye = rng.uniform(-3000, 3000, 3)
        target = eye + rng.standard_normal(3) * rng.uniform(1, 500)
        levers = dict(errorThreshold=float(rng.uniform(0.5, 3)), occlusion=int(rng.integers(0, 3)),
                      minPixelRadius=float(rng.uniform(0, 2)), lodBlend=float(rng.uniform(0.5, 3)),
                      maxErrorThreshold=float(rng.uniform(8, 100)), triangleBudget=int(rng.integers(0, 10_000_000)),
                      thresholdRange=[float(rng.uniform(0.3, 1)), float(rng.uniform(1, 4))], shadowSweep=rng.uniform(-300, 300, 3).tolist(),
                      shadowErrorScale=float(rng.uniform(1, 4)))
        if k % 4 == 3:
            levers["frustumCulling"] = False
        out.append(camera_json(eye.tolist(), target.tolist(), int(rng.choice([540, 720, 1080, 2160])), 4 + k,
                               fov=float(rng.uniform(30, 100)), aspect=float(rng.uniform(0.5, 2.5)), near=float(rng.uniform(0.01, 1)),
                               far=float(rng.uniform(100, 30000)), **levers))
    return out


@needs_node
def test_js_views_equal_python_views(tmp_path):
    cams = cameras()
    got = bytes.fromhex(node(PACK_JS, cams, tmp_path)["hex"])
    want = b"".join(ctypes.string_at(ctypes.addressof(v), ctypes.sizeof(v)) for v in map(python_view, cams))
    assert len(got) == len(want) == 292 * len(cams)
    for k in range(len(cams)):
        a, b = got[k * 292:(k + 1) * 292], want[k * 292:(k + 1) * 292]
        assert a == b, f"camera {k}: words {[i for i in range(73) if a[i * 4:i * 4 + 4] != b[i * 4:i * 4 + 4]]} differ"


STATS_JS = r"""
'use strict';
const host = require(process.argv[2]);
const slots = 4;
const words = new Uint32Array(12 + 2 * slots * 4 + 8);
words[0] = 3; words[1] = 17;                        // visible cells, occluded instances
words[2] = 1200; words[3] = 1000; words[4] = 40; words[6] = 30000000;   // camera: selected, drawn, survived, requested indices
words[7] = 90; words[8] = 90; words[9] = 12; words[11] = 6000000;        // shadow
const base = 12 + 2 * slots * 4;
words[base] = 25000; words[base + 1] = 10000; words[base + 2] = 9500; words[base + 3] = 7000;
const meshes = [{ visible: true, fullDetailTriangles: 5000, instanceCount: 10 }, { visible: false, fullDetailTriangles: 7, instanceCount: 1 },
  { visible: true, fullDetailTriangles: 250, instanceCount: 2 }];
const c = host.countersOf(words, slots);
const cap = { triangles: 8000000, shadowTriangles: 4000000, meshlets: 1000 };
const cap2 = { triangles: 20000000, shadowTriangles: 4000000, meshlets: 2000 };
console.log(JSON.stringify({ stats: host.statsFromCounters(c, cap, meshes), roomy: host.statsFromCounters(c, cap2, meshes),
  active: host.controllerActive(c), still: host.controllerActive(host.countersOf(new Uint32Array(words.length).fill(0).map((x, i) => i === base || i === base + 1 ? 10000 : (i === base + 2 ? 8000 : 0)), slots)) }));
"""


@needs_node
def test_js_read_stats_known_answer(tmp_path):
    """Counters -> VirtualGeometryStats as VirtualGeometry.ts:591-632 assembles them, worked out by hand."""
    got = node(STATS_JS, {}, tmp_path)
    s = got["stats"]
    assert s["drawnMeshlets"] == 1000 and s["visibleInstances"] == 40 and s["occludedInstances"] == 17
    assert s["fullDetailTriangles"] == 5250 and s["instances"] == 12            # the hidden mesh does not count
    assert s["requestedTriangles"] == 10_000_000                               # 30M indices / 3
    assert s["drawnTriangles"] == 8_000_000                                    # min(requested, capacity)
    assert s["shadowTriangles"] == 2_000_000                                   # min(6M / 3, 4M)
    assert s["thresholdScale"] == 2.5                                          # camera scale word 25000
    assert s["overflow"] is True                                               # 1200 selected > 1000 meshlets
    assert s["capacityUse"] == 1.25                                            # max(10M/8M, 1200/1000, 2M/4M, 90/1000)
    r = got["roomy"]
    assert r["overflow"] is False and r["capacityUse"] == 0.6 and r["drawnTriangles"] == 10_000_000  # max(0.5, 0.6, 0.5, 0.045)
    assert got["active"] is True    # fill 0.95 > 0.9: still steering
    assert got["still"] is False    # fill 0.8 at scale 1: settled


GLIDE_JS = r"""
'use strict';
const host = require(process.argv[2]);
const script = JSON.parse(require('fs').readFileSync(process.argv[3], 'utf8'));
const fake = { createContext: () => ({}), createScene: () => ({}) };
const vg = new host.VirtualGeometry({ addon: fake });
vg.triangleBudget = script.budget;
const out = [];
for (const frame of script.frames) {
  vg.stepThreshold();
  if (frame.readback) {
    vg._statsThreshold = vg.errorThreshold;
    vg.adaptThreshold(frame.readback);
  }
  out.push(vg.errorThreshold);
}
console.log(JSON.stringify({ thresholds: out }));
"""


def glide_script():
    frames = [{} for _ in range(40)]
    frames[0]["readback"] = {"drawnTriangles": 2_000_000, "capacityUse": 0.4}  # ratio 2: out of band once
    frames[1]["readback"] = {"drawnTriangles": 2_000_000, "capacityUse": 0.4}  # twice: glide to sqrt(2)
    frames[20]["readback"] = {"drawnTriangles": 900_000, "capacityUse": 0.3}   # ratio 0.9: inside the band, resets
    frames[21]["readback"] = {"drawnTriangles": 250_000, "capacityUse": 0.5}   # ratio 0.25, once
    frames[22]["readback"] = {"drawnTriangles": 250_000, "capacityUse": 0.5}   # twice: refine toward t/2, held by refineThreshold
    return {"budget": 1_000_000, "frames": frames}


def glide_known_answer():
    """The glide worked out by hand: target sqrt(2) set at frame 1, reached by x1.03 per frame; at frame 22 the
    target is max(t * sqrt(0.25), t * (0.5 / 0.82) ** 0.25) = t * 0.88369..., approached by /1.01 per frame."""
    t, out = 1.0, []
    target = None
    for f in range(40):
        if target is not None:
            t = min(target, t * 1.03) if target > t else max(target, t / 1.01)
            if t == target:
                target = None
        if f == 1:
            target = min(3.0, max(0.5, 1.0 * 2 ** 0.5))
        if f == 22:
            target = max(max(0.5, t * 0.5), t * (0.5 / 0.82) ** 0.25)
        out.append(t)
    return out


def python_glide(script):
    vg = PyVirtualGeometry(device=-1)
    vg.triangleBudget = script["budget"]
    out = []
    for frame in script["frames"]:
        vg.step_threshold()
        if "readback" in frame:
            vg._stats_threshold = vg.errorThreshold
            s = VirtualGeometryStats()
            s.drawnTriangles = frame["readback"]["drawnTriangles"]
            s.capacityUse = frame["readback"]["capacityUse"]
            vg.adapt_threshold(s)
        out.append(vg.errorThreshold)
    vg.close()
    return out


@needs_node
def test_budget_glide_known_answer_js_and_python(tmp_path):
    script = glide_script()
    js = node(GLIDE_JS, script, tmp_path)["thresholds"]
    py = python_glide(script)
    want = glide_known_answer()
    assert abs(want[13] - 2 ** 0.5) < 1e-15 and want[12] < 2 ** 0.5  # 1.03^12 < sqrt(2) <= 1.03^13
    np.testing.assert_allclose(js, want, rtol=1e-12)
    np.testing.assert_allclose(py, want, rtol=1e-12)
    assert want[-1] < want[22] and want[-1] > 1.2  # refined, but held above t/2 by refineThreshold


UPDATE_JS = r"""
'use strict';
const fs = require('fs');
const host = require(process.argv[2]);
const spec = JSON.parse(fs.readFileSync(process.argv[3], 'utf8'));
const vg = new host.VirtualGeometry(Object.assign({ addonPath: process.argv[4] }, spec.config));
const fromHex = (h, T) => { const b = Buffer.from(h, 'hex'); return new T(b.buffer.slice(b.byteOffset, b.byteOffset + b.length)); };
for (const m of spec.meshes) {
  const data = {
    meshletBounds: new Float32Array(m.meshletBounds), meshletInfo: new Uint32Array(m.meshletInfo), levelRanges: new Uint32Array(m.levelRanges),
    levelErrors: new Float32Array(m.levelErrors), levelGuardErrors: new Float32Array(m.levelGuardErrors),
    boundingSphere: new Float32Array(m.boundingSphere), lodBoundsSphere: new Float32Array(m.lodBoundsSphere),
    cullBoundsSphere: new Float32Array(m.cullBoundsSphere), leafTriangles: 1000,
  };
  if (m.meshletTriangles) data.meshletTriangles = fromHex(m.meshletTriangles, Uint32Array);
  vg.createMesh(data, null, { matrices: new Float32Array(m.matrices) });
}
vg.errorThreshold = spec.errorThreshold;
const c = spec.camera;
const camera = host.perspectiveCamera(c.fov, c.aspect, c.near, c.far).lookAt(c.eye, c.target);
const out = {};
vg.update(camera, c.height).then(() => {
  const words = new Uint32Array(vg.addon.counterWords(vg.ctx));
  const args = new Uint32Array(vg.addon.drawArgsWords(vg.ctx));
  const list = new Uint32Array(8 * spec.maxEntries);
  vg.addon.read(vg.scene, 0, words, args, list);
  out.counters = Buffer.from(words.buffer).toString('hex');
  out.args = Buffer.from(args.buffer).toString('hex');
  out.list = Buffer.from(list.buffer).toString('hex');
  out.stats = vg.readStats(0);
  return vg.update(camera, c.height);
}).then(() => {
  out.skipped = vg.skippedFrames;
  out.frames = vg.frameCount;
  console.log(JSON.stringify(out));
}).catch((e) => { console.error(e); process.exitCode = 3; });
"""


@pytest.mark.gpu
@needs_addon
def test_js_host_update_equals_python_cut(tmp_path, gpu_available):
    spec = small_ruins(seed=73, instances=1500, meshes=3)
    mm = 1 << 15
    meshes = []
    for d, m in zip(spec.meshes, spec.matrices):
        if len(m):
            j = mesh_json(d, m)
            j.pop("flags")
            meshes.append(j)
    cam = camera_json((-30.0, 4.0, -35.0), (15.0, 0.0, 25.0), 1080, 0, errorThreshold=0.75)
    config = {"maxMeshlets": mm, "maxTriangles": 20_000_000, "maxShadowTriangles": 1_000_000, "meshSlots": 8, "device": 0}
    got = node(UPDATE_JS, {"config": config, "meshes": meshes, "camera": cam, "errorThreshold": 0.75, "maxEntries": mm}, tmp_path)
    ctx = native.Context(max_triangles=20_000_000, max_shadow_triangles=1_000_000, max_meshlets=mm, mesh_slots=8, device=0)
    scene = native.Scene(ctx)
    for d, m in zip(spec.meshes, spec.matrices):
        if len(m):
            slot = scene.add_mesh(d)
            scene.set_instances(slot, m)
            scene.set_flags(slot, 1e30, 0.0, True, False, False)  # the JS host's mesh defaults
    view = python_view(cam)
    scene.cut([view])
    scene.sync()
    counters, args, draws, _ = scene.read(0)
    assert bytes.fromhex(got["counters"]) == counters.tobytes()
    assert bytes.fromhex(got["args"]) == args.tobytes()
    node_draws = np.frombuffer(bytes.fromhex(got["list"]), dtype=np.uint32).reshape(2, mm, 4)
    n = min(int(counters[2]), mm)
    assert n > 0 and np.array_equal(draw_set(node_draws[0], n), draw_set(draws[0], n))
    assert got["stats"]["drawnMeshlets"] == int(counters[3]) and got["stats"]["visibleInstances"] == int(counters[4])
    assert got["frames"] == 2 and got["skipped"] == 1, "an unchanged camera must reuse the last cut"
    # The Python host's mirror on the same scene and camera: update() then readStats() give the same stats.
    from vgcut import camera as pycam

    vg = PyVirtualGeometry(maxTriangles=20_000_000, maxShadowTriangles=1_000_000, maxMeshlets=mm, meshSlots=8, device=0)
    for d, m in zip(spec.meshes, spec.matrices):
        if len(m):
            vg.createMesh(d, {"matrices": m})
    vg.errorThreshold = 0.75
    camera = pycam.perspective_camera(cam["fov"], cam["aspect"], cam["near"], cam["far"]).look_at(cam["eye"], cam["target"])
    vg.update(camera, cam["height"])
    s = vg.readStats()
    for k in ("drawnMeshlets", "drawnTriangles", "visibleInstances", "requestedTriangles", "overflow", "capacityUse",
              "occludedInstances", "shadowTriangles", "thresholdScale", "instances"):
        assert getattr(s, k) == got["stats"][k], (k, getattr(s, k), got["stats"][k])
    vg.close()
