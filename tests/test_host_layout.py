"""Host-side pool layout of the C ABI (no GPU): Morton order, cells, mesh records and level tables equal an
independent numpy restatement (oracle/pool_ref.py) bit for bit; cells enclose their instances."""
import math

import numpy as np
import pytest

from oracle import pool_ref
from vgcut import native, scenegen, scenes


@pytest.fixture(scope="module")
def two_meshes():
    rng = np.random.default_rng(5)
    a = scenes.prop_mesh(rng, 20_000, 101, kind=0)
    b = scenes.prop_mesh(rng, 30_000, 102, kind=2)
    n = 777
    m1 = scenes.yaw_scale_translate(rng.uniform(0, 6.3, n), rng.uniform(0.5, 2, n), rng.uniform(-300, 300, (n, 3)))
    m2 = scenes.yaw_scale_translate(rng.uniform(0, 6.3, 300), rng.uniform(0.5, 2, 300), rng.uniform(-50, 50, (300, 3)))
    return [a, b], [m1, m2]


def test_pool_layout_matches_restatement(two_meshes):
    meshes, mats = two_meshes
    ctx = native.Context(device=-1)
    scene = native.Scene(ctx)
    slots = [scene.add_mesh(d) for d in meshes]
    orders = [scene.set_instances(s, m) for s, m in zip(slots, mats)]
    scene.set_flags(slots[1], 250.0, 1.5, True, False, True)
    # Morton order
    for d, m, order in zip(meshes, mats, orders):
        assert np.array_equal(order, pool_ref.morton_order(m, d.boundingSphere))
    # Instances in GPU order with the slot in the last float
    inst = scene.pool(native.VG_POOL_INSTANCES).reshape(-1, 20)
    base = 0
    for slot, (m, order) in enumerate(zip(mats, orders)):
        got = inst[base:base + len(m)]
        assert np.array_equal(got[:, :16], m[order])
        assert np.all(got[:, 19] == slot) and np.all(got[:, 16:19] == 1)
        base += len(m)
    # Cells
    cells = scene.pool(native.VG_POOL_CELLS)
    want = np.concatenate([pool_ref.compute_cells(m[o], d.boundingSphere, s, b) for s, (d, m, o, b) in
                           enumerate(zip(meshes, mats, orders, [0, len(mats[0])]))])
    assert np.array_equal(cells, want)
    # Mesh records and level tables
    table = scene.pool(native.VG_POOL_MESH_TABLE)
    want = pool_ref.mesh_table(meshes, 64, flags={1: (250.0, 1.5, 1, 0, 1)})
    assert np.array_equal(table, want)
    # Meshlet bounds and info, pool-global
    bounds = scene.pool(native.VG_POOL_MESHLET_BOUNDS)
    assert np.array_equal(bounds.reshape(-1, 16)[:, [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 13, 14, 15]],
                          np.concatenate([d.meshletBounds.reshape(-1, 16) for d in meshes])[:, [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 13, 14, 15]])
    info = scene.pool(native.VG_POOL_MESHLET_INFO).reshape(-1, 4)
    assert np.array_equal(info[:, 0], np.concatenate([d.meshletInfo.reshape(-1, 4)[:, 0] for d in meshes]))
    first = meshes[0].meshletInfo.reshape(-1, 4)
    assert np.array_equal(info[len(first):, 1], meshes[1].meshletInfo.reshape(-1, 4)[:, 1] + first[:, 0].sum() * 3)
    scene.close()
    ctx.close()


def test_cells_enclose_their_instances(two_meshes):
    meshes, mats = two_meshes
    for d, m in zip(meshes, mats):
        order = pool_ref.morton_order(m, d.boundingSphere)
        cells = pool_ref.compute_cells(m[order], d.boundingSphere, 0, 0).reshape(-1, 8)
        spheres = pool_ref.instance_spheres(m[order], d.boundingSphere)
        for c, cell in enumerate(cells):
            t = spheres[c * 128:(c + 1) * 128]
            dist = np.linalg.norm(t[:, :3] - cell[:3].astype(np.float64), axis=1) + t[:, 3]
            assert np.all(dist <= cell[3]) and np.all(t[:, 3] <= cell[4])


def test_morton_order_is_a_stable_sort_of_codes():
    """Ties keep caller order (a stable radix sort equals a stable comparison sort, GeometryPool.ts:1286-1289)."""
    m = np.tile(np.eye(4, dtype=np.float32).reshape(1, 16), (300, 1))
    m[::2, 12] = 5.0  # two distinct positions, many ties
    order = pool_ref.morton_order(m, [0, 0, 0, 1])
    assert np.array_equal(order, np.concatenate([np.arange(1, 300, 2), np.arange(0, 300, 2)]))
    ctx = native.Context(device=-1)
    scene = native.Scene(ctx)
    slot = scene.add_mesh(scenegen.reference_terrain())
    assert np.array_equal(scene.set_instances(slot, m), order)


def test_invalid_inputs_are_rejected():
    ctx = native.Context(device=-1)
    scene = native.Scene(ctx)
    d = scenegen.reference_terrain()
    bad = scenegen.VirtualMeshData(**{**d.__dict__, "levelRanges": d.levelRanges.copy()})
    bad.levelRanges[-2] = d.meshletCount  # a level running past the meshlets
    with pytest.raises(native.VgError) as e:
        scene.add_mesh(bad)
    assert e.value.code == native.VG_E_INVALID
    with pytest.raises(native.VgError):
        scene.set_instances(5, np.eye(4, dtype=np.float32).reshape(1, 16))
    with pytest.raises(native.VgError) as e:
        scene.cut([native.View()])
    assert e.value.code == native.VG_E_DEVICE  # host-only context: no cut


def test_mesh_remove_reuses_slots_and_repacks_instances():
    """100 add / set-instances / remove cycles on a 4-slot pool (GeometryPool.add takes the first free slot,
    GeometryPool.ts:303-304; remove frees it, :357-363): every add lands in the lowest free slot, the cells
    are exactly the live meshes' cells (in slot order, over their current instance ranges), each live mesh's
    instances sit at its range in Morton order with its slot, and the instance columns stay within twice the
    live instances (dead ranges are repacked away). A removed slot refuses further instance and flag calls."""
    rng = np.random.default_rng(17)
    kinds = [scenes.prop_mesh(rng, 1_500, 300 + k, kind=k % 3) for k in range(3)]
    ctx = native.Context(device=-1, mesh_slots=4)
    scene = native.Scene(ctx)
    live = {}  # slot -> (mesh data, matrices in GPU order)
    for cycle in range(100):
        if len(live) == 4 or (live and rng.random() < 0.4):
            slot = int(rng.choice(sorted(live)))
            scene.remove_mesh(slot)
            del live[slot]
            with pytest.raises(native.VgError):
                scene.set_flags(slot, 1e30, 0.0, True, True, False)
            continue
        data = kinds[cycle % 3]
        slot = scene.add_mesh(data)
        assert slot == min(set(range(4)) - set(live)), (cycle, slot, sorted(live))
        n = int(rng.integers(1, 400))
        m = scenes.yaw_scale_translate(rng.uniform(0, 6.3, n), rng.uniform(0.5, 2, n), rng.uniform(-200, 200, (n, 3)))
        order = scene.set_instances(slot, m)
        if rng.random() < 0.3:  # a second call with a new count moves the mesh to a fresh range
            n = int(rng.integers(1, 400))
            m = scenes.yaw_scale_translate(rng.uniform(0, 6.3, n), rng.uniform(0.5, 2, n), rng.uniform(-200, 200, (n, 3)))
            order = scene.set_instances(slot, m)
        live[slot] = (data, m[order])
        inst = scene.pool(native.VG_POOL_INSTANCES).reshape(-1, 20)
        total_live = sum(len(g) for _, g in live.values())
        assert len(inst) <= 2 * total_live + 400, (cycle, len(inst), total_live)
        want = []
        for s in sorted(live):
            d, g = live[s]
            base, count = scene.instance_base(s)
            assert count == len(g)
            assert np.array_equal(inst[base:base + count, :16], g) and np.all(inst[base:base + count, 19] == s)
            want.append(pool_ref.compute_cells(g, d.boundingSphere, s, base))
        assert np.array_equal(scene.pool(native.VG_POOL_CELLS), np.concatenate(want)), cycle
    scene.close()
    ctx.close()
