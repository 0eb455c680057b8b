"""The expand pass on the CPU: the oracle's restatement (GeometryPool.ts:997-1021) against hand-computed
index buffers, the DAG producer's meshlet triangles, and the pool's triangle array (host-only context)."""
import numpy as np
import pytest

from oracle import oracle
from vgcut import native, scenegen, scenes

FULL = 255 << 16  # blend bits of a meshlet drawn over every pixel (part 0)


def tiny_pool(cap0, cap1, max_meshlets=4):
    """Two mesh slots; instance 0 is of slot 1, instance 1 of slot 0. Meshlet 0: 2 triangles at index 0,
    meshlet 1: 1 triangle at index 6 (triangle word 2)."""
    instances = np.zeros((2, 20), dtype=np.float32)
    instances[0, 19], instances[1, 19] = 1, 0
    info = np.array([[2, 0, 0, 0], [1, 6, 1, 4]], dtype=np.uint32)
    tri = np.array([0 | 1 << 8 | 2 << 16, 2 | 1 << 8 | 3 << 16, 4 | 5 << 8 | 6 << 16], dtype=np.uint32)
    return oracle.Pool(mesh_table=np.zeros(24 * 2, dtype=np.float32), instances=instances.reshape(-1), cells=np.zeros(0, dtype=np.float32),
                       bounds=np.zeros(32, dtype=np.float32), info=info.reshape(-1), slots=2, mesh_slots=2, max_meshlets=max_meshlets,
                       cut_capacity=(cap0, cap1), triangles=tri)


def counters_for(drawn, starts, slots=2):
    """Counter words with C_DRAWN per cut and region starts {(cut, slot, part): start}."""
    c = np.zeros(12 + 2 * slots * 4, dtype=np.uint32)
    for cut, n in enumerate(drawn):
        c[2 + cut * 5 + 1] = n
    for (cut, slot, part), start in starts.items():
        c[12 + cut * slots * 4 + 1 + slot * 4 + part * 2] = start
    return c


def test_known_answer_camera_cut():
    pool = tiny_pool(30, 30)
    draws = np.zeros((2, 4, 4), dtype=np.uint32)
    draws[0, 0] = (0, 0, 3, 2 | FULL)                 # instance 0 (slot 1), part 0, firstIndex 3
    draws[0, 1] = (1, 1, 0, 1 | 200 << 16 | 7 << 8)   # instance 1 (slot 0), part 1 (blended), firstIndex 0
    draws[0, 2] = (1, 1, 9, 1 | FULL)                 # beyond C_DRAWN: not expanded
    counters = counters_for((2, 0), {(0, 1, 0): 10, (0, 0, 1): 20})
    out, written = oracle.expand(pool, 0, counters, draws)
    want = np.zeros(30, dtype=np.uint32)
    want[13:19] = [0, 1, 2, 2, 1, 3]                  # position 0: base 0
    want[20:23] = [132, 133, 134]                     # position 1: base 1 << 7, local ids 4, 5, 6
    assert written == 9
    assert np.array_equal(out, want)


def test_known_answer_capacity_and_shadow_positions():
    pool = tiny_pool(22, 40)
    draws = np.zeros((2, 4, 4), dtype=np.uint32)
    draws[0, 0] = (0, 0, 3, 2 | FULL)
    draws[0, 1] = (1, 1, 0, 1 | 200 << 16)
    # capacity 22: index 20 + 2 is not < 22, so that triangle is skipped; the others fit
    out, written = oracle.expand(pool, 0, counters_for((2, 0), {(0, 1, 0): 10, (0, 0, 1): 20}), draws)
    assert written == 6 and np.all(out[19:] == 0) and np.array_equal(out[13:19], [0, 1, 2, 2, 1, 3])
    # shadow cut: positions count from max_meshlets (4), so entry 0 has base 4 << 7 = 512
    draws[1, 0] = (1, 0, 0, 2 | FULL)
    out, written = oracle.expand(pool, 1, counters_for((0, 1), {(1, 0, 0): 5}), draws)
    want = np.zeros(40, dtype=np.uint32)
    want[5:11] = [512, 513, 514, 514, 513, 515]
    assert written == 6 and np.array_equal(out, want)
    # positions not written keep what the buffer held
    prev = np.full(40, 7, dtype=np.uint32)
    out, _ = oracle.expand(pool, 1, counters_for((0, 1), {(1, 0, 0): 5}), draws, prev)
    assert np.all(out[:5] == 7) and np.all(out[11:] == 7) and np.array_equal(out[5:11], want[5:11])


def test_planted_mistake_fails():
    """A firstIndex one triangle off moves the written range: the comparison must see it."""
    pool = tiny_pool(30, 30)
    draws = np.zeros((2, 4, 4), dtype=np.uint32)
    draws[0, 0] = (0, 0, 3, 2 | FULL)
    good, _ = oracle.expand(pool, 0, counters_for((1, 0), {(0, 1, 0): 10}), draws)
    draws[0, 0, 2] = 6
    bad, _ = oracle.expand(pool, 0, counters_for((1, 0), {(0, 1, 0): 10}), draws)
    assert not np.array_equal(good, bad)


@pytest.fixture(scope="module")
def mesh():
    rng = np.random.default_rng(3)
    return scenes.prop_mesh(rng, 30_000, 77, kind=1)


def test_producer_triangles(mesh):
    """Every meshlet's triangles: its count, contiguous from firstIndex / 3, local ids below its vertex count."""
    info = mesh.meshletInfo.reshape(-1, 4).astype(np.int64)
    tri = mesh.meshletTriangles
    assert tri is not None and len(tri) == int(info[:, 0].sum())
    assert np.array_equal(info[:, 1], np.concatenate([[0], np.cumsum(info[:-1, 0] * 3)]))
    verts = np.diff(np.concatenate([info[:, 3], [info[-1, 3] + 10**9]]))
    for i in range(0, len(info), max(1, len(info) // 200)):
        t = tri[info[i, 1] // 3:info[i, 1] // 3 + info[i, 0]]
        ids = np.stack([t & 255, (t >> 8) & 255, (t >> 16) & 255], axis=1)
        assert np.all(t >> 24 == 0) and np.all(ids < min(verts[i], 128))
        assert np.all((ids[:, 0] != ids[:, 1]) & (ids[:, 1] != ids[:, 2]) & (ids[:, 0] != ids[:, 2]))


def test_pool_triangles_and_validation(mesh):
    rng = np.random.default_rng(4)
    other = scenes.prop_mesh(rng, 20_000, 78, kind=0)
    ctx = native.Context(device=-1, expand=True)
    scene = native.Scene(ctx)
    scene.add_mesh(mesh)
    scene.add_mesh(other)
    tri = scene.pool(native.VG_POOL_MESHLET_TRIANGLES)
    assert np.array_equal(tri, np.concatenate([mesh.meshletTriangles, other.meshletTriangles]))
    info = scene.pool(native.VG_POOL_MESHLET_INFO).reshape(-1, 4)
    n0 = mesh.meshletCount
    assert np.array_equal(info[n0:, 1], other.meshletInfo.reshape(-1, 4)[:, 1] + 3 * len(mesh.meshletTriangles))
    # the expand pass needs triangles, and they must cover every meshlet's range
    bare = scenegen.VirtualMeshData(**{**mesh.__dict__, "meshletTriangles": None})
    with pytest.raises(native.VgError) as e:
        scene.add_mesh(bare)
    assert e.value.code == native.VG_E_INVALID
    short = scenegen.VirtualMeshData(**{**mesh.__dict__, "meshletTriangles": mesh.meshletTriangles[:-1].copy()})
    with pytest.raises(native.VgError) as e:
        scene.add_mesh(short)
    assert e.value.code == native.VG_E_INVALID
    scene.close()
    ctx.close()
