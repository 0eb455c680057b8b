"""The `.vgeo` container (vgcut/vgeo.py), restating the reference's own test of it (tests/serialize.test.ts:15-120):
exact round trips (compressed with the meshoptimizer codecs, without deflate, uncompressed, indices rebuilt from
the meshlet arrays or kept), the container's layout (magic, version, 4-byte aligned payload), the codecs in use and
the file much smaller than its arrays, unknown fields (typed array, null, object with Infinity), decoding from an
unaligned view, and rejection of empty, foreign, newer, corrupted and truncated input. A decoded mesh fills a pool
byte-identical to the one its source fills.

The codecs themselves (csrc/vgcodec.cpp) are pinned by streams assembled by hand from the format's description
(meshoptimizer's vertex codec versions 0 and 1: control bytes, byte-group widths and escapes, literal and zero
planes, the three delta transforms, the tail; its index codec: edge, vertex-FIFO, table, slow, free-index, last +- 1
and reset codes), by their rejection rules and by round trips. Parity unpinned against meshoptimizer itself: the
package is absent here and the reference ships no .vgeo file, so no stream meshoptimizer wrote is decoded."""
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


@pytest.mark.parametrize("options", [{}, {"deflate": False}, {"compress": False}, {"keep_indices": True}])
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


def tail_v1(first: bytes, channels: bytes) -> bytes:
    """A version-1 vertex stream's tail: zeros up to 24 bytes, the first vertex, a channel byte per 4-byte group."""
    return bytes(24 - len(first) - len(channels)) + first + channels


VERTEX_STREAMS = [
    # (stream, count, size, decoded): assembled by hand from the format's description
    (bytes([0xa1, 0xaa]) + tail_v1(bytes([1, 2, 3, 4]), b"\0"), 1, 4, bytes([1, 2, 3, 4])),  # zero planes
    # byte deltas (channel 0), plane 0 with one 4-bit group: deltas 0, +3 (zigzag 6)
    (bytes([0xa1, 0xa8, 0x03, 0x06, 0, 0, 0, 0, 0, 0, 0]) + tail_v1(bytes([10, 0, 0, 0]), b"\0"), 2, 4,
     bytes([10, 0, 0, 0, 13, 0, 0, 0])),
    # a 2-bit group whose value 111 (zigzag of -56) is escaped to a literal byte after the group
    (bytes([0xa1, 0xa8, 0x02, 0x30, 0, 0, 0, 0x6f]) + tail_v1(bytes(4), b"\0"), 2, 4, bytes([0, 0, 0, 0, 200, 0, 0, 0])),
    # a literal plane (control 3)
    (bytes([0xa1, 0xab, 0x00, 0x08]) + tail_v1(bytes([5, 0, 0, 0]), b"\0"), 2, 4, bytes([5, 0, 0, 0, 9, 0, 0, 0])),
    # 16-bit deltas (channel 1): lanes 0x0100 -> 0x00ff (-1, zigzag 1) and 5 -> 7 (+2, zigzag 4)
    (bytes([0xa1, 0x88, 0x02, 0x10, 0, 0, 0, 0x03, 0x04, 0, 0, 0, 0, 0, 0, 0]) + tail_v1(bytes([0, 1, 5, 0]), b"\x01"), 2, 4,
     bytes([0, 1, 5, 0, 0xff, 0, 7, 0])),
    # 32-bit xor (channel 2) with rotation field 8: the stored 0x100 rotates left by 24 to 1
    (bytes([0xa1, 0xa2, 0x02, 0x10, 0, 0, 0]) + tail_v1(bytes([0x78, 0x56, 0x34, 0x12]), b"\x82"), 2, 4,
     bytes([0x78, 0x56, 0x34, 0x12, 0x79, 0x56, 0x34, 0x12])),
    # version 0: no control byte, widths {0, 2, 4, 8}, a 32-byte tail without channels
    (bytes([0xa0, 0x02, 0x06, 0, 0, 0, 0, 0, 0, 0, 0x00, 0x00, 0x00]) + bytes(28) + bytes([10, 0, 0, 0]), 2, 4,
     bytes([10, 0, 0, 0, 13, 0, 0, 0])),
]

TABLE = bytes([0x00, 0x01, 0x10, 0x12, 0x21, 0x11, 0x02, 0x20, 0x13, 0x31, 0x22, 0x23, 0x32, 0x14, 0, 0])
INDEX_STREAMS = [
    (bytes([0xe1, 0xf0, 0x10]) + TABLE, [0, 1, 2, 2, 1, 3]),  # table entry (new, new, new), then edge (2, 1) + next
    (bytes([0xe1, 0xf0, 0x12]) + TABLE, [0, 1, 2, 2, 1, 0]),  # edge (2, 1) + the vertex FIFO's entry 2 back
    (bytes([0xe1, 0xf0, 0x0e]) + TABLE, [0, 1, 2, 0, 2, 1]),  # edge (0, 2) + last + 1 (version 1)
    (bytes([0xe1, 0xff, 0xff, 0x0a, 0x08, 0x03]) + TABLE, [5, 9, 7]),  # three free zigzag deltas: +5, +4, -2
    (bytes([0xe1, 0xfe, 0x0f, 0x0e]) + TABLE, [0, 1, 7]),  # slow path: new, new, free +7
    (bytes([0xe1, 0xf0, 0xfe, 0x00]) + TABLE, [0, 1, 2, 0, 1, 2]),  # auxiliary byte 0 resets the next vertex
]


@pytest.mark.parametrize("case", range(len(VERTEX_STREAMS)))
def test_vertex_codec_known_streams(case):
    stream, count, size, want = VERTEX_STREAMS[case]
    assert dag.decode_vertex_buffer(stream, count, size) == want


@pytest.mark.parametrize("case", range(len(INDEX_STREAMS)))
def test_index_codec_known_streams(case):
    stream, want = INDEX_STREAMS[case]
    assert dag.decode_index_buffer(stream, len(want)).tolist() == want


def test_codecs_reject_malformed_streams():
    good, count, size, _ = VERTEX_STREAMS[1]
    for bad, code in ((bytes([0xb1]) + good[1:], -1), (bytes([0xa2]) + good[1:], -1), (good[:20], -2), (good[:4] + good, -3)):
        with pytest.raises(ValueError, match=f"error {code}"):
            dag.decode_vertex_buffer(bad, count, size)
    with pytest.raises(ValueError, match="error -2"):
        dag.decode_vertex_buffer(good, count + 300, size)
    stream, want = INDEX_STREAMS[3]
    for bad, code in ((bytes([0xd1]) + stream[1:], -1), (bytes([0xe2]) + stream[1:], -1), (stream[:10], -2),
                      (stream[:6] + b"\x00" + stream[6:], -3)):
        with pytest.raises(ValueError, match=f"error {code}"):
            dag.decode_index_buffer(bad, len(want))
    # Version 0 reads codes 13 and 14 as vertex FIFO entries (here never written), not as last - 1 / last + 1.
    v0 = dag.decode_index_buffer(bytes([0xe0]) + INDEX_STREAMS[2][0][1:], 6).tolist()
    assert v0[:5] == [0, 1, 2, 0, 2] and v0[5] == 0xFFFFFFFF


def test_codec_round_trips():
    rng = np.random.default_rng(7)
    for stride in (4, 8, 12, 16, 64, 256):
        for count in (0, 1, 2, 15, 16, 17, 255, 256, 257, 700):
            for data in (rng.integers(0, 256, count * stride, dtype=np.uint8).tobytes(),
                         np.cumsum(rng.normal(size=count * stride // 4)).astype(np.float32).tobytes(), bytes(count * stride)):
                assert dag.decode_vertex_buffer(dag.encode_vertex_buffer(data, count, stride), count, stride) == data
    pos, idx = scenegen.triangle_mesh(scenegen.SPHERE, seed=2, size=(1.0, 0, 0), noise_amplitude=0.05, noise_frequency=2.0,
                                      target_triangles=5_000)
    for tris in (idx, rng.integers(0, 1 << 20, 3 * 500, dtype=np.uint32), np.arange(3 * 300, dtype=np.uint32)[::-1].copy()):
        back = dag.decode_index_buffer(dag.encode_index_buffer(tris), len(tris)).reshape(-1, 3)
        t = tris.reshape(-1, 3)
        rotations = [np.roll(t, -r, axis=1) for r in range(3)]
        assert all(any(np.array_equal(b, rt[i]) for rt in rotations) for i, b in enumerate(back)), "a rotation of each triangle"
    assert len(dag.encode_index_buffer(idx)) < len(idx) * 4 // 2  # a mesh's index list codes below 2 bytes per index


def test_compressed_file_uses_the_codecs(record):
    """serialize.test.ts:95-99: the vertex and triangle codecs are used and the file is much smaller than the arrays;
    a hand-assembled file whose fields use both codecs decodes."""
    _, rec = record
    b = vgeo.encode_virtual_mesh(rec)
    enc = {f["name"]: f["encoding"] for f in header_of(b)["fields"]}
    assert "meshopt-vertex" in enc.values() and enc["meshletTriangles"] == "meshlet-triangles"
    raw = sum(v.nbytes for v in rec.values() if isinstance(v, np.ndarray))
    assert len(b) * 2.5 < raw, (len(b), raw)
    stream, count, size, want = VERTEX_STREAMS[4]
    info = np.array([2, 0, 0, 0], dtype=np.uint32)  # one meshlet of 2 triangles at vertex offset 0
    tri = INDEX_STREAMS[0][0]
    rot = bytes([0b0100])  # the second triangle stored rotated once: (2, 1, 3) was (1, 3, 2)
    tblob = struct.pack("<I", len(tri)) + tri + bytes(vgeo.align4(len(tri)) - len(tri)) + rot
    fields = [{"name": "meshletInfo", "type": "Uint32Array", "length": 4, "encoding": "raw", "offset": 0, "byteLength": 16},
              {"name": "positions", "type": "Float32Array", "length": 2, "encoding": "meshopt-vertex", "stride": 4,
               "offset": 16, "byteLength": len(stream)},
              {"name": "meshletTriangles", "type": "Uint32Array", "length": 2, "encoding": "meshlet-triangles",
               "offset": 16 + vgeo.align4(len(stream)), "byteLength": len(tblob)}]
    payload = info.tobytes() + stream + bytes(vgeo.align4(len(stream)) - len(stream)) + tblob
    f = vgeo.decode_virtual_mesh(assemble({"generator": "three-virtual-geometry", "keys": ["meshletInfo", "positions",
                                                                                          "meshletTriangles"],
                                           "fields": fields, "values": {}}, payload))
    assert f["positions"].tobytes() == want
    assert f["meshletTriangles"].tolist() == [0 | 1 << 8 | 2 << 16, 1 | 3 << 8 | 2 << 16]


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
