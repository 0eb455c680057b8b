"""The `.vgeo` container (vgcut/vgeo.py), restating the reference's own test of it (tests/serialize.test.ts:15-120)
for the encodings this reader supports: exact round trips (raw, deflated, indices rebuilt from the meshlet
arrays or kept), the container's layout (magic, version, 4-byte aligned payload), unknown fields (typed
array, null, object with Infinity), decoding from an unaligned view, and rejection of empty, foreign, newer,
corrupted and truncated input; files using meshoptimizer's codecs are refused with a message that says so.
A decoded mesh fills a pool byte-identical to the one its source fills."""
import json
import struct

import numpy as np
import pytest

from vgcut import dag, native, scenegen, vgeo


@pytest.fixture(scope="module")
def record():
    pos, idx = scenegen.triangle_mesh(scenegen.BOX, seed=4, size=(1.0, 0.7, 0.5), noise_amplitude=0.05, noise_frequency=2.0,
                                      target_triangles=20_000)
    d = dag.build_virtual_mesh(pos, idx)
    return d, vgeo.record_of(d)


def identical(a, b):
    if list(a.keys()) != list(b.keys()):
        return f"keys differ: {list(a)} vs {list(b)}"
    for k in a:
        x, y = a[k], b[k]
        if isinstance(x, np.ndarray):
            if not isinstance(y, np.ndarray) or x.dtype != y.dtype or x.tobytes() != y.tobytes():
                return f"field {k} differs"
        elif json.dumps(x) != json.dumps(y):
            return f"field {k} differs"
    return None


def header_of(b):
    return json.loads(b[16:16 + struct.unpack_from("<I", b, 8)[0]])


@pytest.mark.parametrize("options", [{}, {"deflate": False}, {"keep_indices": True}])
def test_round_trip_is_exact(record, options):
    _, rec = record
    b = vgeo.encode_virtual_mesh(rec, **options)
    back = vgeo.decode_virtual_mesh(b)
    assert identical(rec, back) is None, identical(rec, back)
    f = [x for x in header_of(b)["fields"] if x["name"] == "indices"][0]
    if options.get("keep_indices"):
        assert f["encoding"] != "meshlet-indices"
    else:
        assert f["encoding"] == "meshlet-indices" and f["byteLength"] == 0


def test_container_layout(record):
    _, rec = record
    b = vgeo.encode_virtual_mesh(rec)
    assert b[:4] == b"VGEO" and struct.unpack_from("<I", b, 4)[0] == vgeo.VG_FORMAT_VERSION
    h = header_of(b)
    assert h["generator"] == "three-virtual-geometry" and h["keys"] == list(rec.keys())
    assert all(f["offset"] % 4 == 0 for f in h["fields"]) and len(b) % 4 == 0
    raw = sum(v.nbytes for v in rec.values() if isinstance(v, np.ndarray))
    assert len(b) < raw  # deflate and the dropped index list make the file smaller than its arrays
    # The header's padding is spaces, and the stored checksum covers everything after the prefix.
    n = struct.unpack_from("<I", b, 8)[0]
    assert set(b[16 + n:16 + vgeo.align4(n)]) <= {0x20}
    assert struct.unpack_from("<I", b, 12)[0] == vgeo.checksum(b, 16)


def test_unknown_fields_and_unaligned_input(record):
    _, rec = record
    extended = dict(rec)
    extended["uvs"] = (np.arange(len(rec["positions"]) // 3 * 2, dtype=np.float32) * 0.25)
    extended["tangents"] = None
    extended["extra"] = {"scale": float("inf"), "tag": "x"}
    b = vgeo.encode_virtual_mesh(extended)
    back = vgeo.decode_virtual_mesh(b)
    assert back["tangents"] is None and back["extra"]["scale"] == float("inf") and back["extra"]["tag"] == "x"
    assert back["uvs"].dtype == np.float32 and np.array_equal(back["uvs"], extended["uvs"])
    unaligned = bytearray(len(b) + 1)
    unaligned[1:] = b
    assert identical(rec, vgeo.decode_virtual_mesh(vgeo.encode_virtual_mesh(rec))) is None
    assert identical(extended | {"extra": back["extra"]}, vgeo.decode_virtual_mesh(memoryview(unaligned)[1:])) is None


def test_bad_input_is_rejected(record):
    _, rec = record
    b = vgeo.encode_virtual_mesh(rec)
    with pytest.raises(vgeo.VirtualMeshFormatError, match="not a three-virtual-geometry"):
        vgeo.decode_virtual_mesh(b"")
    with pytest.raises(vgeo.VirtualMeshFormatError, match="not a three-virtual-geometry"):
        vgeo.decode_virtual_mesh(b'{"glTF": "not this"}')
    newer = bytearray(b)
    struct.pack_into("<I", newer, 4, vgeo.VG_FORMAT_VERSION + 1)
    with pytest.raises(vgeo.VirtualMeshFormatError, match="unsupported .*version"):
        vgeo.decode_virtual_mesh(bytes(newer))
    flipped = bytearray(b)
    flipped[len(b) - 100] ^= 0x10
    with pytest.raises(vgeo.VirtualMeshFormatError, match="corrupt"):
        vgeo.decode_virtual_mesh(bytes(flipped))
    with pytest.raises(vgeo.VirtualMeshFormatError, match="truncated"):
        vgeo.decode_virtual_mesh(b[:len(b) - 64])


def assemble(header, payload: bytes) -> bytes:
    """A file from a header and payload laid out as serialize.ts:143-157 does."""
    text = json.dumps(header, separators=(",", ":")).encode()
    out = bytearray(16 + vgeo.align4(len(text)) + vgeo.align4(len(payload)))
    out[16:16 + len(text)] = text
    out[16 + len(text):16 + vgeo.align4(len(text))] = b" " * (vgeo.align4(len(text)) - len(text))
    out[16 + vgeo.align4(len(text)):16 + vgeo.align4(len(text)) + len(payload)] = payload
    struct.pack_into("<III", out, 0, vgeo.MAGIC, vgeo.VG_FORMAT_VERSION, len(text))
    struct.pack_into("<I", out, 12, vgeo.checksum(bytes(out), 16))
    return bytes(out)


def test_meshopt_encodings_are_refused():
    blob = np.arange(12, dtype=np.float32).tobytes()
    field = {"name": "positions", "type": "Float32Array", "length": 12, "offset": 0, "byteLength": len(blob)}
    ok = assemble({"generator": "three-virtual-geometry", "keys": ["positions"], "fields": [field | {"encoding": "raw"}],
                   "values": {}}, blob)
    assert np.array_equal(vgeo.decode_virtual_mesh(ok)["positions"], np.arange(12, dtype=np.float32))
    for enc in ("meshopt-vertex", "meshlet-triangles"):
        bad = assemble({"generator": "three-virtual-geometry", "keys": ["positions"],
                        "fields": [field | {"encoding": enc, "stride": 12}], "values": {}}, blob)
        with pytest.raises(vgeo.VirtualMeshFormatError, match="meshoptimizer codec"):
            vgeo.decode_virtual_mesh(bad)


def test_decoded_mesh_fills_the_same_pool(record):
    d, rec = record
    back = vgeo.mesh_of(vgeo.decode_virtual_mesh(vgeo.encode_virtual_mesh(rec)))
    mats = np.tile(np.eye(4, dtype=np.float32).reshape(1, 16), (5, 1))
    mats[:, 12] = np.arange(5) * 3.0
    pools = []
    for data in (d, back):
        ctx = native.Context(max_views=1, mesh_slots=4, device=-1)
        scene = native.Scene(ctx)
        slot = scene.add_mesh(data)
        scene.set_instances(slot, mats)
        pools.append([scene.pool(b).tobytes() for b in range(8)])
    assert pools[0] == pools[1]
