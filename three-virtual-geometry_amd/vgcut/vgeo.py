"""The `.vgeo` baked-mesh container (src/core/io/serialize.ts): read and write, without meshoptimizer.

Layout (serialize.ts:1-16, 109-158): little-endian prefix ('VGEO', format version 1, header byte length,
checksum of everything after the 16-byte prefix), the UTF-8 JSON header space-padded to 4 bytes, then one
4-byte aligned blob per typed-array field. The header lists every top-level key in order, the typed-array
fields (name, array type, length, encoding, payload offset and size) and every other value as JSON, with
non-finite numbers written as {"$number": "Infinity"} (serialize.ts:454-468).

Encodings: 'raw', any of them with `deflate` (raw deflate stream, serialize.ts:442-452), 'meshlet-indices'
(`indices` rebuilt from meshletInfo / meshletVertices / meshletTriangles, serialize.ts:383-404),
'meshopt-vertex' (meshoptimizer's vertex codec over `stride`-byte elements, serialize.ts:244-245, 338-341) and
'meshlet-triangles' (meshletTriangles as one index buffer over meshlet-vertex slots through meshoptimizer's
index codec, plus 2 bits per triangle of corner rotation, serialize.ts:257-296, 353-381). The codecs are
csrc/vgcodec.cpp (through vgcut.dag). The writer follows encodeField (serialize.ts:232-262): with `compress`,
each field's candidates (raw, the vertex codec at the field's hinted strides, or meshlet-triangles) are
ordered by size, the best one (and raw, where hinted) is also tried deflated, and the first candidate that
decodes back to the same bytes is stored.
"""
from __future__ import annotations

import json
import math
import struct
import zlib
from typing import Dict, Union

import numpy as np

VG_FORMAT_VERSION = 1  # serialize.ts:22
MAGIC = 0x4F454756  # 'VGEO' as a little-endian u32 (serialize.ts:24)
PREFIX_BYTES = 16
MESHLET_INFO_STRIDE = 4

# Typed-array names of the header and their element types (serialize.ts:73-83).
TYPED = {
    "Int8Array": np.int8, "Uint8Array": np.uint8, "Uint8ClampedArray": np.uint8, "Int16Array": np.int16,
    "Uint16Array": np.uint16, "Int32Array": np.int32, "Uint32Array": np.uint32, "Float32Array": np.float32,
    "Float64Array": np.float64,
}
_NAME_OF = {np.dtype(np.int8): "Int8Array", np.dtype(np.uint8): "Uint8Array", np.dtype(np.int16): "Int16Array",
            np.dtype(np.uint16): "Uint16Array", np.dtype(np.int32): "Int32Array", np.dtype(np.uint32): "Uint32Array",
            np.dtype(np.float32): "Float32Array", np.dtype(np.float64): "Float64Array"}


class VirtualMeshFormatError(ValueError):
    """serialize.ts:98-103."""

    def __init__(self, message: str):
        super().__init__(f"three-virtual-geometry: {message}")


def align4(n: int) -> int:
    return (n + 3) & ~3


def checksum(data: bytes, start: int) -> int:
    """32-bit multiplicative hash over the u32 words from `start` (serialize.ts:431-440)."""
    words = np.frombuffer(data, dtype="<u4", count=(len(data) - start) >> 2, offset=start)
    h = (0x811C9DC5 ^ len(words)) & 0xFFFFFFFF
    for w in words.tolist():
        h = ((h ^ w) * 0x01000193) & 0xFFFFFFFF
        h ^= h >> 15
    return h


def _reviver(obj):
    if len(obj) == 1 and "$number" in obj:
        return float(obj["$number"])
    return obj


def _replace(value):
    if isinstance(value, float) and not math.isfinite(value):
        return {"$number": "Infinity" if value > 0 else ("-Infinity" if value < 0 else "NaN")}
    if isinstance(value, dict):
        return {k: _replace(v) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        return [_replace(v) for v in value]
    if isinstance(value, np.ndarray):
        raise VirtualMeshFormatError("typed array nested inside a field cannot be serialized; store it as a top-level field")
    if isinstance(value, (np.floating,)):
        return _replace(float(value))
    if isinstance(value, (np.integer,)):
        return int(value)
    return value


def rebuild_indices(length: int, info: np.ndarray, meshlet_vertices: np.ndarray, triangles: np.ndarray) -> np.ndarray:
    """indices[firstIndex + 3t + k] = meshletVertices[vertexOffset + local id k of triangle t] (serialize.ts:383-404)."""
    info = np.asarray(info, dtype=np.uint32).reshape(-1, MESHLET_INFO_STRIDE)
    out = np.zeros(length, dtype=np.uint32)
    for count, first, _, voff in info.astype(np.int64):
        if first + count * 3 > length or first // 3 + count > len(triangles):
            raise VirtualMeshFormatError("geometry file is corrupt (meshlet range out of bounds)")
        packed = triangles[first // 3: first // 3 + count].astype(np.int64)
        local = np.stack([packed & 255, (packed >> 8) & 255, (packed >> 16) & 255], 1).reshape(-1)
        v = voff + local
        if len(v) and v.max() >= len(meshlet_vertices):
            raise VirtualMeshFormatError("geometry file is corrupt (meshlet vertex out of bounds)")
        out[first:first + count * 3] = meshlet_vertices[v]
    return out


def decode_virtual_mesh(data: Union[bytes, bytearray, memoryview]) -> Dict[str, object]:
    """decodeVirtualMesh (serialize.ts:160-219): the fields in the file's key order."""
    b = bytes(data)
    if len(b) < PREFIX_BYTES or struct.unpack_from("<I", b, 0)[0] != MAGIC:
        raise VirtualMeshFormatError("not a three-virtual-geometry geometry file (bad magic bytes)")
    version, header_length, stored = struct.unpack_from("<III", b, 4)
    if version != VG_FORMAT_VERSION:
        raise VirtualMeshFormatError(f"unsupported geometry format version {version} (this build reads version "
                                     f"{VG_FORMAT_VERSION}); re-bake the file")
    payload_start = PREFIX_BYTES + align4(header_length)
    if payload_start > len(b) or len(b) % 4 != 0:
        raise VirtualMeshFormatError("geometry file is truncated")
    try:
        header = json.loads(b[PREFIX_BYTES:PREFIX_BYTES + header_length].decode("utf-8"), object_hook=_reviver)
    except (UnicodeDecodeError, json.JSONDecodeError):
        raise VirtualMeshFormatError("geometry file is corrupt or truncated (unreadable header)")
    if not isinstance(header, dict) or not isinstance(header.get("fields"), list) or not isinstance(header.get("keys"), list) \
            or not isinstance(header.get("values"), dict):
        raise VirtualMeshFormatError("geometry file is corrupt (malformed header)")
    payload = b[payload_start:]
    rank = {"meshlet-triangles": 1, "meshlet-indices": 2}
    order = sorted(header["fields"], key=lambda f: rank.get(f.get("encoding"), 0))
    for f in order:
        if f.get("type") not in TYPED or not isinstance(f.get("length"), int) or f["length"] < 0:
            raise VirtualMeshFormatError(f'geometry file is corrupt (field "{f.get("name")}")')
        if f["offset"] % 4 != 0 or f["offset"] + f["byteLength"] > len(payload):
            raise VirtualMeshFormatError(f'geometry file is truncated (field "{f["name"]}")')
    if checksum(b, PREFIX_BYTES) != stored:
        raise VirtualMeshFormatError("geometry file is corrupt (checksum mismatch)")
    arrays: Dict[str, np.ndarray] = {}
    for f in order:
        dtype = np.dtype(TYPED[f["type"]]).newbyteorder("<")
        blob = payload[f["offset"]:f["offset"] + f["byteLength"]]
        try:
            if f.get("deflate"):
                blob = zlib.decompress(blob, -15)
        except zlib.error as e:
            raise VirtualMeshFormatError(f'geometry file is corrupt (field "{f["name"]}": {e})')
        enc = f["encoding"]
        if enc == "raw":
            if len(blob) != f["length"] * dtype.itemsize:
                raise VirtualMeshFormatError(f'geometry file is corrupt (field "{f["name"]}" has the wrong size)')
            arrays[f["name"]] = np.frombuffer(blob, dtype=dtype).astype(dtype.newbyteorder("="), copy=True)
        elif enc == "meshlet-indices":
            need = [arrays.get(k) for k in ("meshletInfo", "meshletVertices", "meshletTriangles")]
            if any(a is None or a.dtype != np.uint32 for a in need):
                raise VirtualMeshFormatError(f'geometry file is corrupt ("{f["name"]}" needs the meshlet arrays)')
            arrays[f["name"]] = rebuild_indices(f["length"], *need)
        elif enc == "meshopt-vertex":
            stride = f.get("stride") or 0
            nbytes = f["length"] * dtype.itemsize
            if not isinstance(stride, int) or stride <= 0 or stride % 4 or nbytes % stride:
                raise VirtualMeshFormatError(f'geometry file is corrupt (field "{f["name"]}" stride)')
            try:
                raw = _dag().decode_vertex_buffer(blob, nbytes // stride, stride)
            except ValueError as e:
                raise VirtualMeshFormatError(f'geometry file is corrupt (field "{f["name"]}": {e})')
            arrays[f["name"]] = np.frombuffer(raw, dtype=dtype).astype(dtype.newbyteorder("="), copy=True)
        elif enc == "meshlet-triangles":
            info = arrays.get("meshletInfo")
            if info is None or info.dtype != np.uint32:
                raise VirtualMeshFormatError(f'geometry file is corrupt ("{f["name"]}" needs "meshletInfo")')
            try:
                arrays[f["name"]] = decode_meshlet_triangles(blob, f["length"], info)
            except ValueError as e:
                raise VirtualMeshFormatError(f'geometry file is corrupt (field "{f["name"]}": {e})')
        else:
            raise VirtualMeshFormatError(f'unknown encoding "{enc}" for field "{f["name"]}"')
    return {k: (arrays[k] if k in arrays else header["values"].get(k)) for k in header["keys"]}


def _dag():
    from . import dag

    return dag


# serialize.ts:92-100 FIELD_HINTS: vertex-codec strides to try, and whether raw data is also deflated.
FIELD_HINTS = {
    "positions": {"strides": [12]}, "normals": {"strides": [12]}, "colors": {"strides": [12]}, "uvs": {"strides": [8]},
    "tangents": {"strides": [16]}, "meshletBounds": {"strides": [64], "deflateRaw": True}, "meshletInfo": {"strides": [16]},
}


def encode_meshlet_triangles(triangles: np.ndarray, info) -> Union[bytes, None]:
    """serialize.ts:257-296 encodeMeshletTriangles: u32 codec length, the index codec over meshlet-vertex slots
    (padded to 4 bytes), then each triangle's corner rotation in 2 bits; None when the meshlets do not cover every
    triangle exactly once with local ids below 256."""
    if not isinstance(info, np.ndarray) or info.dtype != np.uint32:
        return None
    tri = np.asarray(triangles, dtype=np.uint32)
    n = len(tri)
    slots = np.zeros(n * 3, dtype=np.uint32)
    covered = np.zeros(n, dtype=np.uint8)
    for count, first_index, _, voff in info.reshape(-1, MESHLET_INFO_STRIDE).astype(np.int64):
        first = first_index // 3
        if first + count > n:
            return None
        packed = tri[first:first + count].astype(np.int64)
        if np.any(covered[first:first + count]) or np.any(packed >> 24):
            return None
        covered[first:first + count] = 1
        s = slots[first * 3:(first + count) * 3].reshape(-1, 3)
        s[:, 0] = voff + (packed & 255)
        s[:, 1] = voff + ((packed >> 8) & 255)
        s[:, 2] = voff + ((packed >> 16) & 255)
    if not np.all(covered == 1):
        return None
    codec = _dag().encode_index_buffer(slots)
    decoded = _dag().decode_index_buffer(codec, len(slots)).reshape(-1, 3).astype(np.int64)
    a, b, c = (slots.reshape(-1, 3)[:, k].astype(np.int64) for k in range(3))
    x, y, z = decoded[:, 0], decoded[:, 1], decoded[:, 2]
    r0 = (a == x) & (b == y) & (c == z)
    r1 = (a == y) & (b == z) & (c == x)
    r2 = (a == z) & (b == x) & (c == y)
    if not np.all(r0 | r1 | r2):
        return None
    rot = np.where(r0, 0, np.where(r1, 1, 2)).astype(np.uint8)
    rotations = np.zeros((n + 3) // 4, dtype=np.uint8)
    for k in range(4):
        part = rot[k::4]
        rotations[:len(part)] |= (part << (2 * k)).astype(np.uint8)
    pad = align4(len(codec)) - len(codec)
    return struct.pack("<I", len(codec)) + codec + b"\0" * pad + rotations.tobytes()


def decode_meshlet_triangles(blob: bytes, length: int, info: np.ndarray) -> np.ndarray:
    """serialize.ts:353-381 decodeMeshletTriangles."""
    if len(blob) < 4:
        raise ValueError("truncated triangle data")
    codec_length = struct.unpack_from("<I", blob, 0)[0]
    start = 4 + align4(codec_length)
    if start + (length + 3) // 4 > len(blob):
        raise ValueError("truncated triangle data")
    slots = _dag().decode_index_buffer(blob[4:4 + codec_length], length * 3).astype(np.int64)
    rotations = np.frombuffer(blob, dtype=np.uint8, offset=start)
    out = np.zeros(length, dtype=np.uint32)
    for count, first_index, _, voff in info.reshape(-1, MESHLET_INFO_STRIDE).astype(np.int64):
        first = first_index // 3
        if first + count > length:
            raise ValueError("meshlet range out of bounds")
        t = np.arange(first, first + count)
        r = (rotations[t >> 2] >> ((t & 3) * 2)) & 3
        xyz = slots[first * 3:(first + count) * 3].reshape(-1, 3) - voff
        x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
        a = np.where(r == 0, x, np.where(r == 1, y, z))
        b = np.where(r == 0, y, np.where(r == 1, z, x))
        c = np.where(r == 0, z, np.where(r == 1, x, y))
        if np.any(((a | b | c) >> 8) != 0):
            raise ValueError("meshlet-local vertex id out of range")
        out[first:first + count] = (a | (b << 8) | (c << 16)).astype(np.uint32)
    return out


def _deflate(raw: bytes) -> bytes:
    c = zlib.compressobj(9, zlib.DEFLATED, -15)
    return c.compress(raw) + c.flush()


def _decode_blob(entry, blob: bytes, info) -> Union[bytes, None]:
    """The raw bytes an encoded candidate decodes to (None if it does not decode)."""
    try:
        if entry.get("deflate"):
            blob = zlib.decompress(blob, -15)
        enc = entry["encoding"]
        itemsize = np.dtype(TYPED[entry["type"]]).itemsize
        if enc == "raw":
            return blob
        if enc == "meshopt-vertex":
            n = entry["length"] * itemsize
            return _dag().decode_vertex_buffer(blob, n // entry["stride"], entry["stride"])
        if enc == "meshlet-triangles":
            return decode_meshlet_triangles(blob, entry["length"], info).astype("<u4").tobytes()
    except (ValueError, zlib.error):
        return None
    return None


def _encode_field(name: str, type_name: str, arr: np.ndarray, record, use_deflate: bool):
    """serialize.ts:232-262 encodeField: the smallest candidate that decodes back exactly."""
    raw = arr.astype(arr.dtype.newbyteorder("<"), copy=False).tobytes()
    base = {"name": name, "type": type_name, "length": int(arr.size)}
    raw_candidate = ({**base, "encoding": "raw"}, raw)
    candidates = [raw_candidate]
    info = record.get("meshletInfo")
    tri = encode_meshlet_triangles(arr, info) if name == "meshletTriangles" and arr.dtype == np.uint32 else None
    if tri is not None:
        candidates.append(({**base, "encoding": "meshlet-triangles"}, tri))
    elif len(raw) % 4 == 0 and len(raw) > 0:
        hint = FIELD_HINTS.get(name, {})
        strides = [s for s in hint.get("strides", [max(4, arr.dtype.itemsize)]) if s % 4 == 0 and s <= 256 and len(raw) % s == 0]
        for stride in strides:
            blob = _dag().encode_vertex_buffer(raw, len(raw) // stride, stride)
            candidates.append(({**base, "encoding": "meshopt-vertex", "stride": stride}, blob))
    candidates.sort(key=lambda c: len(c[1]))
    if use_deflate and len(raw) >= 64:
        hint = FIELD_HINTS.get(name, {})
        to_deflate = [candidates[0], raw_candidate] if hint.get("deflateRaw") and candidates[0] is not raw_candidate else [candidates[0]]
        for entry, blob in to_deflate:
            candidates.append(({**entry, "deflate": True}, _deflate(blob)))
        candidates.sort(key=lambda c: len(c[1]))
    for entry, blob in candidates:
        if _decode_blob(entry, blob, info) == raw:
            return entry, blob
    return raw_candidate


def encode_virtual_mesh(record: Dict[str, object], compress: bool = True, deflate: bool = True,
                        keep_indices: bool = False) -> bytes:
    """encodeVirtualMesh (serialize.ts:109-158): each typed array (numpy array) through encodeField when
    `compress` (raw otherwise); `indices` becomes 'meshlet-indices' (nothing stored) when the meshlet arrays
    rebuild it exactly, unless `keep_indices`. Other values go into the JSON header."""
    keys = list(record.keys())  # None (null) values are kept, as the reference keeps null
    fields, blobs, values = [], [], {}
    offset = 0
    for name in keys:
        value = record[name]
        if not isinstance(value, np.ndarray):
            values[name] = _replace(value)
            continue
        if value.dtype not in _NAME_OF:
            raise VirtualMeshFormatError(f'cannot serialize field "{name}" of type {value.dtype}')
        arr = np.ascontiguousarray(value.reshape(-1))
        entry = {"name": name, "type": _NAME_OF[arr.dtype], "length": int(arr.size)}
        raw = arr.astype(arr.dtype.newbyteorder("<"), copy=False).tobytes()
        if name == "indices" and not keep_indices and _indices_match(record):
            entry["encoding"] = "meshlet-indices"
            blob = b""
        elif compress:
            entry, blob = _encode_field(name, entry["type"], arr, record, deflate)
            entry = dict(entry)
        else:
            entry["encoding"] = "raw"
            blob = raw
        entry["offset"] = offset
        entry["byteLength"] = len(blob)
        fields.append(entry)
        blobs.append(blob)
        offset += align4(len(blob))
    header = {"generator": "three-virtual-geometry", "keys": keys, "fields": fields, "values": values}
    header_bytes = json.dumps(header, separators=(",", ":"), allow_nan=False).encode("utf-8")
    out = bytearray(PREFIX_BYTES + align4(len(header_bytes)) + offset)
    out[PREFIX_BYTES:PREFIX_BYTES + len(header_bytes)] = header_bytes
    out[PREFIX_BYTES + len(header_bytes):PREFIX_BYTES + align4(len(header_bytes))] = b" " * (align4(len(header_bytes)) - len(header_bytes))
    start = PREFIX_BYTES + align4(len(header_bytes))
    for f, blob in zip(fields, blobs):
        out[start + f["offset"]:start + f["offset"] + len(blob)] = blob
    struct.pack_into("<III", out, 0, MAGIC, VG_FORMAT_VERSION, len(header_bytes))
    struct.pack_into("<I", out, 12, checksum(bytes(out), PREFIX_BYTES))
    return bytes(out)


def _indices_match(record) -> bool:
    """indicesMatchMeshlets (serialize.ts:298-309)."""
    arrs = [record.get(k) for k in ("indices", "meshletInfo", "meshletVertices", "meshletTriangles")]
    if any(not isinstance(a, np.ndarray) or a.dtype != np.uint32 for a in arrs):
        return False
    try:
        return np.array_equal(rebuild_indices(len(arrs[0]), arrs[1], arrs[2], arrs[3]), arrs[0])
    except VirtualMeshFormatError:
        return False


def load_virtual_mesh(source) -> Dict[str, object]:
    """loadVirtualMesh (serialize.ts:221-227) for a file path or bytes."""
    if isinstance(source, (bytes, bytearray, memoryview)):
        return decode_virtual_mesh(source)
    with open(source, "rb") as f:
        return decode_virtual_mesh(f.read())


def record_of(data) -> Dict[str, object]:
    """A builder result (vgcut.dag / scenegen VirtualMeshData) as the reference's VirtualMeshData record:
    typed arrays as numpy arrays, spheres as plain lists (buildVirtualMesh.ts:75-138 key names)."""
    out: Dict[str, object] = {}
    for k in ("positions", "normals", "colors", "uvs", "indices", "meshletCount", "meshletBounds", "meshletInfo",
              "meshletVertices", "meshletTriangles", "levelRanges", "levelErrors", "levelGuardErrors", "lodBoundsSphere",
              "cullBoundsSphere", "replacementGroup", "replacementStart", "replacementIndices", "boundingSphere"):
        v = getattr(data, k, None)
        if v is None:
            continue
        if k.endswith("Sphere"):
            v = [float(x) for x in np.asarray(v).reshape(-1)]
        elif isinstance(v, np.ndarray):
            v = np.ascontiguousarray(v)
        out[k] = v
    out["stats"] = {"leafTriangles": int(data.leafTriangles), "leafMeshlets": int(data.leafMeshlets),
                    "rootTriangles": int(data.rootTriangles), "rootMeshlets": int(data.rootMeshlets),
                    "lodLevels": int(getattr(data, "lodLevels", 0) or 0), "voxelLevels": int(getattr(data, "voxelLevels", 0) or 0),
                    "aggregateGroups": int(getattr(data, "aggregateGroups", 0) or 0)}
    return out


def mesh_of(record: Dict[str, object]):
    """A decoded record as the VirtualMeshData vg_mesh_add takes (scene.add_mesh)."""
    from .scenegen import VirtualMeshData

    stats = record.get("stats") or {}
    return VirtualMeshData(
        meshletCount=int(record["meshletCount"]), meshletBounds=record["meshletBounds"], meshletInfo=record["meshletInfo"],
        levelRanges=record["levelRanges"], levelErrors=record["levelErrors"], levelGuardErrors=record["levelGuardErrors"],
        replacementGroup=record.get("replacementGroup", np.zeros(0, np.int32)),
        replacementStart=record.get("replacementStart", np.zeros(1, np.uint32)),
        replacementIndices=record.get("replacementIndices", np.zeros(0, np.uint32)),
        boundingSphere=np.asarray(record["boundingSphere"], dtype=np.float32),
        lodBoundsSphere=np.asarray(record["lodBoundsSphere"], dtype=np.float32),
        cullBoundsSphere=np.asarray(record["cullBoundsSphere"], dtype=np.float32),
        leafTriangles=int(stats.get("leafTriangles", 0)), leafMeshlets=int(stats.get("leafMeshlets", 0)),
        rootTriangles=int(stats.get("rootTriangles", 0)), rootMeshlets=int(stats.get("rootMeshlets", 0)),
        meshletTriangles=record.get("meshletTriangles"), positions=record.get("positions"), indices=record.get("indices"),
        meshletVertices=record.get("meshletVertices"), lodLevels=int(stats.get("lodLevels", 0)), normals=record.get("normals"),
        colors=record.get("colors"), uvs=record.get("uvs"), voxelLevels=int(stats.get("voxelLevels", 0)),
        aggregateGroups=int(stats.get("aggregateGroups", 0)))
