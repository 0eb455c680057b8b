// meshoptimizer's vertex and index buffer codecs, the byte formats the reference's .vgeo files use for the
// 'meshopt-vertex' and 'meshlet-triangles' encodings (src/core/io/serialize.ts:232-341, through the npm
// package meshoptimizer 1.3.0's MeshoptEncoder / MeshoptDecoder). Host-only C++, part of lib/libvgdag.so.
//
// The formats are restated from meshoptimizer's published description of them (vertexcodec.cpp /
// indexcodec.cpp, MIT licence); the code is written here. Decoding follows the format exactly, including its
// validity checks, so a file the reference wrote decodes here. Encoding writes valid streams of the same
// format with encoder choices of our own (which channel transform, which bit widths, which triangle codes),
// so its bytes differ from meshoptimizer's while decoding to the same data.
//
// Vertex codec (header 0xa0 | version, versions 0 and 1): vertices in blocks of up to 256 (8192 bytes /
// vertex size, rounded down to 16); per block and per 4-byte column group, version 1 first stores one control
// byte (2 bits per byte plane: 0 / 1 = byte groups with bit widths {0,1,2,4} / {1,2,4,8}, 2 = all zero,
// 3 = literal bytes); a plane of deltas is coded in groups of 16 bytes, a 2-bit width selector per group in
// a header, values that do not fit the width escaped to a literal byte after the group. Deltas are per byte
// (zigzag, version 0 and channel 0), per 16-bit lane (zigzag, channel 1), or 32-bit xor with a rotation
// (channel 2). The tail holds the first vertex (the deltas' start) and, in version 1, a channel byte per
// column group, padded with zeros in front to at least 32 (v0) / 24 (v1) bytes.
//
// Index codec (header 0xe0 | version, versions 0 and 1): one code byte per triangle, then variable-length
// data, then a 16-byte table of auxiliary codes. A code below 0xf0 names a recent edge (16-entry edge FIFO)
// and codes the third vertex as the next new vertex, a recent vertex (16-entry vertex FIFO), last +- 1
// (version 1) or a zigzag varint delta from the last free index; codes 0xf0-0xfd take a new first vertex and
// the other two from the table entry; 0xfe / 0xff read the auxiliary byte from the data (first vertex new or
// free). The decoder may rotate a triangle's corners, which is why the .vgeo layer stores each rotation.
#include "../../include/vgdag.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr size_t kGroup = 16;             // bytes per byte group
constexpr size_t kBlockBytes = 8192;      // vertex block budget
constexpr size_t kBlockMax = 256;         // vertices per block at most
constexpr size_t kGroupDecodeLimit = 24;  // bytes a group decode may look ahead
constexpr size_t kTailMinV0 = 32, kTailMinV1 = 24;
const int kBitsV0[4] = {0, 2, 4, 8};
const int kBitsV1[5] = {0, 1, 2, 4, 8};

size_t block_vertices(size_t vertex_size) {
  const size_t n = (kBlockBytes / vertex_size) & ~(kGroup - 1);
  return n < kBlockMax ? n : kBlockMax;
}

size_t tail_padded(size_t vertex_size, int version) {
  const size_t tail = vertex_size + (version == 0 ? 0 : vertex_size / 4);
  const size_t lo = version == 0 ? kTailMinV0 : kTailMinV1;
  return tail < lo ? lo : tail;
}

// ---------------------------------------------------------------- vertex decoding

// One group of 16 bytes at `bits` per value; all-ones values are escapes to a literal after the packed bits.
const uint8_t* decode_group(const uint8_t* data, uint8_t* out, int bits) {
  if (bits == 0) {
    std::memset(out, 0, kGroup);
    return data;
  }
  if (bits == 8) {
    std::memcpy(out, data, kGroup);
    return data + kGroup;
  }
  const uint8_t* extra = data + bits * 2;  // 16 values x bits / 8
  const unsigned sentinel = (1u << bits) - 1;
  const int per_byte = 8 / bits;
  for (int i = 0; i < 16; i += per_byte) {
    unsigned byte = *data++;
    for (int k = 0; k < per_byte; k++) {
      const unsigned enc = (byte >> (8 - bits)) & sentinel;
      byte = (byte << bits) & 0xff;
      out[i + k] = enc == sentinel ? *extra++ : static_cast<uint8_t>(enc);
    }
  }
  return extra;
}

const uint8_t* decode_plane(const uint8_t* data, const uint8_t* end, uint8_t* out, size_t size, const int* bits) {
  const size_t header = (size / kGroup + 3) / 4;
  if (static_cast<size_t>(end - data) < header) return nullptr;
  const uint8_t* sel = data;
  data += header;
  for (size_t i = 0; i < size; i += kGroup) {
    if (static_cast<size_t>(end - data) < kGroupDecodeLimit) return nullptr;
    const size_t g = i / kGroup;
    data = decode_group(data, out + i, bits[(sel[g / 4] >> ((g % 4) * 2)) & 3]);
  }
  return data;
}

inline uint32_t unzigzag(uint32_t v) { return (0u - (v & 1)) ^ (v >> 1); }
inline uint16_t unzigzag16(uint16_t v) { return static_cast<uint16_t>((0u - (v & 1)) ^ (v >> 1)); }
inline uint8_t unzigzag8(uint8_t v) { return static_cast<uint8_t>((0u - (v & 1)) ^ (v >> 1)); }
inline uint32_t rotl(uint32_t v, int r) { return r ? (v << r) | (v >> (32 - r)) : v; }

// Rebuilds the 4 columns of one column group from their 4 byte planes (plane j of the group at j * n).
void undo_deltas(const uint8_t* planes, uint8_t* out, size_t n, size_t stride, uint8_t* last, int channel) {
  const int mode = channel & 3;
  if (mode == 0) {
    for (int j = 0; j < 4; j++) {
      uint8_t p = last[j];
      for (size_t i = 0; i < n; i++) {
        p = static_cast<uint8_t>(unzigzag8(planes[j * n + i]) + p);
        out[i * stride + j] = p;
      }
    }
  } else if (mode == 1) {
    for (int h = 0; h < 2; h++) {
      uint16_t p = static_cast<uint16_t>(last[h * 2] | (last[h * 2 + 1] << 8));
      for (size_t i = 0; i < n; i++) {
        const uint16_t v = static_cast<uint16_t>(planes[(h * 2) * n + i] | (planes[(h * 2 + 1) * n + i] << 8));
        p = static_cast<uint16_t>(unzigzag16(v) + p);
        out[i * stride + h * 2] = static_cast<uint8_t>(p);
        out[i * stride + h * 2 + 1] = static_cast<uint8_t>(p >> 8);
      }
    }
  } else {
    const int rot = (32 - (channel >> 4)) & 31;
    uint32_t p = static_cast<uint32_t>(last[0]) | (static_cast<uint32_t>(last[1]) << 8) | (static_cast<uint32_t>(last[2]) << 16) |
                 (static_cast<uint32_t>(last[3]) << 24);
    for (size_t i = 0; i < n; i++) {
      const uint32_t v = static_cast<uint32_t>(planes[i]) | (static_cast<uint32_t>(planes[n + i]) << 8) |
                         (static_cast<uint32_t>(planes[2 * n + i]) << 16) | (static_cast<uint32_t>(planes[3 * n + i]) << 24);
      p = rotl(v, rot) ^ p;
      for (int j = 0; j < 4; j++) out[i * stride + j] = static_cast<uint8_t>(p >> (8 * j));
    }
  }
}

const uint8_t* decode_block(const uint8_t* data, const uint8_t* end, uint8_t* out, size_t n, size_t vertex_size,
                            uint8_t* last, const uint8_t* channels, int version) {
  const size_t aligned = (n + kGroup - 1) & ~(kGroup - 1);
  const size_t control_size = version == 0 ? 0 : vertex_size / 4;
  if (static_cast<size_t>(end - data) < control_size) return nullptr;
  const uint8_t* control = data;
  data += control_size;
  // Plane j of a column group at j * n; each plane decode writes `aligned` bytes, so the buffer has room.
  std::vector<uint8_t> planes(n * 3 + aligned + kGroup);
  for (size_t k = 0; k < vertex_size; k += 4) {
    const uint8_t ctrl_byte = version == 0 ? 0 : control[k / 4];
    for (int j = 0; j < 4; j++) {
      const int ctrl = (ctrl_byte >> (j * 2)) & 3;
      uint8_t* dst = planes.data() + j * n;
      if (ctrl == 3) {
        if (static_cast<size_t>(end - data) < n) return nullptr;
        std::memcpy(dst, data, n);
        data += n;
      } else if (ctrl == 2) {
        std::memset(dst, 0, n);
      } else {
        data = decode_plane(data, end, dst, aligned, version == 0 ? kBitsV0 : kBitsV1 + ctrl);
        if (!data) return nullptr;
      }
    }
    const int channel = version == 0 ? 0 : channels[k / 4];
    if ((channel & 3) == 3) return nullptr;
    undo_deltas(planes.data(), out + k, n, vertex_size, last + k, channel);
  }
  std::memcpy(last, out + (n - 1) * vertex_size, vertex_size);
  return data;
}

// ---------------------------------------------------------------- vertex encoding

size_t group_cost(const uint8_t* g, int bits) {
  if (bits == 0) {
    for (size_t i = 0; i < kGroup; i++)
      if (g[i]) return SIZE_MAX;
    return 0;
  }
  if (bits == 8) return kGroup;
  size_t size = bits * 2;
  const unsigned sentinel = (1u << bits) - 1;
  for (size_t i = 0; i < kGroup; i++) size += g[i] >= sentinel;
  return size;
}

uint8_t* encode_group(uint8_t* data, const uint8_t* g, int bits) {
  if (bits == 0) return data;
  if (bits == 8) {
    std::memcpy(data, g, kGroup);
    return data + kGroup;
  }
  const unsigned sentinel = (1u << bits) - 1;
  const int per_byte = 8 / bits;
  for (int i = 0; i < 16; i += per_byte) {
    unsigned byte = 0;
    for (int k = 0; k < per_byte; k++) byte = (byte << bits) | std::min<unsigned>(g[i + k], sentinel);
    *data++ = static_cast<uint8_t>(byte);
  }
  for (int i = 0; i < 16; i++)
    if (g[i] >= sentinel) *data++ = g[i];
  return data;
}

// Size of a plane coded with a width table (each group at its cheapest width), or its bytes when `out`.
size_t plane_cost(const uint8_t* plane, size_t size, const int* bits, uint8_t* out) {
  const size_t header = (size / kGroup + 3) / 4;
  size_t total = header;
  uint8_t* sel = out;
  uint8_t* data = out ? out + header : nullptr;
  if (out) std::memset(sel, 0, header);
  for (size_t i = 0; i < size; i += kGroup) {
    int best = 3;
    size_t best_size = group_cost(plane + i, bits[3]);
    for (int b = 0; b < 3; b++) {
      const size_t s = group_cost(plane + i, bits[b]);
      if (s < best_size) {
        best = b;
        best_size = s;
      }
    }
    total += best_size;
    if (out) {
      const size_t g = i / kGroup;
      sel[g / 4] = static_cast<uint8_t>(sel[g / 4] | (best << ((g % 4) * 2)));
      data = encode_group(data, plane + i, bits[best]);
    }
  }
  return total;
}

inline uint8_t zigzag8(uint8_t v) { return static_cast<uint8_t>((v << 1) ^ (0u - (v >> 7))); }
inline uint16_t zigzag16(uint16_t v) { return static_cast<uint16_t>((v << 1) ^ (0u - (v >> 15))); }

// Byte planes (4 per column group, n each) of the deltas of one column group under `channel`.
void make_planes(const uint8_t* vertices, size_t n, size_t stride, const uint8_t* last, int channel, uint8_t* planes) {
  const int mode = channel & 3;
  if (mode == 0) {
    for (int j = 0; j < 4; j++) {
      uint8_t p = last[j];
      for (size_t i = 0; i < n; i++) {
        const uint8_t v = vertices[i * stride + j];
        planes[j * n + i] = zigzag8(static_cast<uint8_t>(v - p));
        p = v;
      }
    }
  } else if (mode == 1) {
    for (int h = 0; h < 2; h++) {
      uint16_t p = static_cast<uint16_t>(last[h * 2] | (last[h * 2 + 1] << 8));
      for (size_t i = 0; i < n; i++) {
        const uint16_t v = static_cast<uint16_t>(vertices[i * stride + h * 2] | (vertices[i * stride + h * 2 + 1] << 8));
        const uint16_t d = zigzag16(static_cast<uint16_t>(v - p));
        planes[(h * 2) * n + i] = static_cast<uint8_t>(d);
        planes[(h * 2 + 1) * n + i] = static_cast<uint8_t>(d >> 8);
        p = v;
      }
    }
  } else {
    const int rot = channel >> 4;  // the decoder rotates left by 32 - rot
    uint32_t p = static_cast<uint32_t>(last[0]) | (static_cast<uint32_t>(last[1]) << 8) | (static_cast<uint32_t>(last[2]) << 16) |
                 (static_cast<uint32_t>(last[3]) << 24);
    for (size_t i = 0; i < n; i++) {
      const uint8_t* q = vertices + i * stride;
      const uint32_t v = static_cast<uint32_t>(q[0]) | (static_cast<uint32_t>(q[1]) << 8) | (static_cast<uint32_t>(q[2]) << 16) |
                         (static_cast<uint32_t>(q[3]) << 24);
      const uint32_t d = rotl(v ^ p, rot);
      for (int j = 0; j < 4; j++) planes[j * n + i] = static_cast<uint8_t>(d >> (8 * j));
      p = v;
    }
  }
}

// Cheapest version-1 coding of one plane: (control, size).
std::pair<int, size_t> plane_choice(const uint8_t* plane, size_t n, size_t aligned) {
  bool zero = true;
  for (size_t i = 0; i < n && zero; i++) zero = plane[i] == 0;
  if (zero) return {2, 0};
  const size_t s0 = plane_cost(plane, aligned, kBitsV1 + 0, nullptr);
  const size_t s1 = plane_cost(plane, aligned, kBitsV1 + 1, nullptr);
  if (std::min(s0, s1) >= n) return {3, n};
  return s0 <= s1 ? std::make_pair(0, s0) : std::make_pair(1, s1);
}

}  // namespace

extern "C" size_t vgd_vertex_encode_bound(size_t vertex_count, size_t vertex_size) {
  if (vertex_size == 0 || vertex_size > 256 || vertex_size % 4) return 0;
  const size_t bs = block_vertices(vertex_size);
  const size_t blocks = (vertex_count + bs - 1) / bs;
  const size_t header = (bs / kGroup + 3) / 4;
  return 1 + blocks * (vertex_size / 4 + vertex_size * (header + bs)) + std::max(tail_padded(vertex_size, 1), kTailMinV0);
}

extern "C" size_t vgd_vertex_encode(uint8_t* out, size_t capacity, const void* vertices_v, size_t vertex_count,
                                    size_t vertex_size) {
  if (vertex_size == 0 || vertex_size > 256 || vertex_size % 4 || capacity < vgd_vertex_encode_bound(vertex_count, vertex_size))
    return 0;
  const uint8_t* vertices = static_cast<const uint8_t*>(vertices_v);
  uint8_t* data = out;
  *data++ = 0xa1;
  uint8_t first[256] = {}, last[256] = {};
  if (vertex_count) std::memcpy(first, vertices, vertex_size);
  std::memcpy(last, first, vertex_size);
  const size_t bs = block_vertices(vertex_size);
  const size_t groups = vertex_size / 4;
  // One channel per column group for the whole buffer: the transform whose planes code smallest.
  std::vector<uint8_t> channels(groups, 0);
  std::vector<uint8_t> planes(4 * bs + kGroup);
  if (vertex_count > 1) {
    for (size_t k = 0; k < groups; k++) {
      size_t best = SIZE_MAX;
      for (int channel : {0, 1, 2}) {
        size_t total = 0;
        uint8_t prev[4];
        std::memcpy(prev, first + k * 4, 4);
        for (size_t at = 0; at < vertex_count; at += bs) {
          const size_t n = std::min(bs, vertex_count - at), aligned = (n + kGroup - 1) & ~(kGroup - 1);
          std::fill(planes.begin(), planes.end(), 0);
          make_planes(vertices + at * vertex_size + k * 4, n, vertex_size, prev, channel, planes.data());
          std::vector<uint8_t> padded(aligned, 0);
          for (int j = 0; j < 4; j++) {
            std::fill(padded.begin(), padded.end(), 0);
            std::memcpy(padded.data(), planes.data() + j * n, n);
            total += plane_choice(padded.data(), n, aligned).second;
          }
          std::memcpy(prev, vertices + (at + n - 1) * vertex_size + k * 4, 4);
        }
        if (total < best) {
          best = total;
          channels[k] = static_cast<uint8_t>(channel);
        }
      }
    }
  }
  for (size_t at = 0; at < vertex_count; at += bs) {
    const size_t n = std::min(bs, vertex_count - at), aligned = (n + kGroup - 1) & ~(kGroup - 1);
    uint8_t* control = data;
    std::memset(control, 0, groups);
    data += groups;
    std::vector<uint8_t> padded(aligned, 0);
    for (size_t k = 0; k < groups; k++) {
      std::fill(planes.begin(), planes.end(), 0);
      make_planes(vertices + at * vertex_size + k * 4, n, vertex_size, last + k * 4, channels[k], planes.data());
      for (int j = 0; j < 4; j++) {
        std::fill(padded.begin(), padded.end(), 0);
        std::memcpy(padded.data(), planes.data() + j * n, n);
        const int ctrl = plane_choice(padded.data(), n, aligned).first;
        control[k] = static_cast<uint8_t>(control[k] | (ctrl << (j * 2)));
        if (ctrl == 3) {
          std::memcpy(data, padded.data(), n);
          data += n;
        } else if (ctrl != 2) {
          const size_t size = plane_cost(padded.data(), aligned, kBitsV1 + ctrl, data);
          data += size;
        }
      }
    }
    std::memcpy(last, vertices + (at + n - 1) * vertex_size, vertex_size);
  }
  const size_t tail = vertex_size + groups, pad = tail_padded(vertex_size, 1);
  std::memset(data, 0, pad - tail);
  data += pad - tail;
  std::memcpy(data, first, vertex_size);
  data += vertex_size;
  std::memcpy(data, channels.data(), groups);
  data += groups;
  return static_cast<size_t>(data - out);
}

// 0 on success; -1 unsupported header or vertex size, -2 truncated or malformed, -3 trailing bytes.
extern "C" int vgd_vertex_decode(void* out_v, size_t vertex_count, size_t vertex_size, const uint8_t* buffer, size_t size) {
  if (vertex_size == 0 || vertex_size > 256 || vertex_size % 4) return -1;
  if (size < 1) return -2;
  const uint8_t header = buffer[0];
  if ((header & 0xf0) != 0xa0) return -1;
  const int version = header & 0x0f;
  if (version > 1) return -1;
  const uint8_t* data = buffer + 1;
  const uint8_t* end = buffer + size;
  const size_t tail = vertex_size + (version == 0 ? 0 : vertex_size / 4);
  const size_t pad = tail_padded(vertex_size, version);
  if (static_cast<size_t>(end - data) < pad) return -2;
  const uint8_t* tail_at = end - tail;
  uint8_t last[256];
  std::memcpy(last, tail_at, vertex_size);
  const uint8_t* channels = version == 0 ? nullptr : tail_at + vertex_size;
  uint8_t* out = static_cast<uint8_t*>(out_v);
  const size_t bs = block_vertices(vertex_size);
  for (size_t at = 0; at < vertex_count; at += bs) {
    const size_t n = std::min(bs, vertex_count - at);
    data = decode_block(data, end, out + at * vertex_size, n, vertex_size, last, channels, version);
    if (!data) return -2;
  }
  if (static_cast<size_t>(end - data) != pad) return -3;
  return 0;
}

// ---------------------------------------------------------------- index codec

namespace {

struct IndexState {
  uint32_t edges[16][2];
  uint32_t verts[16];
  size_t eoff = 0, voff = 0;
  uint32_t next = 0, last = 0;
  IndexState() {
    std::memset(edges, 0xff, sizeof edges);
    std::memset(verts, 0xff, sizeof verts);
  }
  void push_edge(uint32_t a, uint32_t b) {
    edges[eoff][0] = a;
    edges[eoff][1] = b;
    eoff = (eoff + 1) & 15;
  }
  void push_vertex(uint32_t v, bool cond = true) {
    verts[voff] = v;
    voff = (voff + (cond ? 1 : 0)) & 15;
  }
  int find_edge(uint32_t a, uint32_t b) const {
    for (int i = 0; i < 16; i++) {
      const size_t k = (eoff - 1 - i) & 15;
      if (edges[k][0] == a && edges[k][1] == b) return i;
    }
    return -1;
  }
  int find_vertex(uint32_t v) const {
    for (int i = 0; i < 16; i++)
      if (verts[(voff - 1 - i) & 15] == v) return i;
    return -1;
  }
};

uint32_t read_varint(const uint8_t*& data) {
  const uint8_t lead = *data++;
  if (lead < 128) return lead;
  uint32_t result = lead & 127, shift = 7;
  for (int i = 0; i < 4; i++) {
    const uint8_t g = *data++;
    result |= static_cast<uint32_t>(g & 127) << shift;
    shift += 7;
    if (g < 128) break;
  }
  return result;
}

uint32_t read_index(const uint8_t*& data, uint32_t last) {
  const uint32_t v = read_varint(data);
  return last + ((v >> 1) ^ (0u - (v & 1)));
}

void write_index(std::vector<uint8_t>& out, uint32_t v, uint32_t last) {
  const uint32_t d = v - last;
  uint32_t z = (d << 1) ^ (0u - (d >> 31));
  do {
    uint8_t b = z & 127;
    z >>= 7;
    if (z) b |= 128;
    out.push_back(b);
  } while (z);
}

// The auxiliary-code table the encoder writes: (second, third) vertex codes of the 0xf0-0xfd fast path.
// 0 = the next new vertex, f = the vertex FIFO's entry f - 1 back (before the first vertex is pushed).
const uint8_t kAuxTable[16] = {0x00, 0x01, 0x10, 0x12, 0x21, 0x11, 0x02, 0x20, 0x13, 0x31, 0x22, 0x23, 0x32, 0x14, 0, 0};

}  // namespace

extern "C" size_t vgd_index_encode_bound(size_t index_count, size_t vertex_count) {
  // Header, a code byte and at most 3 x 5 varint bytes (plus one auxiliary byte) per triangle, the table.
  (void)vertex_count;
  return 1 + (index_count / 3) * 17 + 16;
}

extern "C" size_t vgd_index_encode(uint8_t* out, size_t capacity, const uint32_t* indices, size_t index_count) {
  if (index_count % 3 || capacity < vgd_index_encode_bound(index_count, 0)) return 0;
  const size_t tris = index_count / 3;
  std::vector<uint8_t> codes, data;
  codes.reserve(tris);
  IndexState s;
  const int fecmax = 13;
  for (size_t t = 0; t < tris; t++) {
    const uint32_t tri[3] = {indices[t * 3], indices[t * 3 + 1], indices[t * 3 + 2]};
    bool done = false;
    // An edge of the FIFO: the rotation whose first edge is stored; the third vertex by the cheapest code.
    for (int r = 0; r < 3 && !done; r++) {
      const uint32_t a = tri[r], b = tri[(r + 1) % 3], c = tri[(r + 2) % 3];
      const int fe = s.find_edge(a, b);
      if (fe < 0 || fe > 14) continue;
      const int fc = s.find_vertex(c);
      int fec;
      if (c == s.next) fec = 0;
      else if (fc >= 1 && fc < fecmax) fec = fc;
      else if (c == s.last - 1) fec = 13;
      else if (c == s.last + 1) fec = 14;
      else fec = 15;
      codes.push_back(static_cast<uint8_t>((fe << 4) | fec));
      if (fec == 15) write_index(data, c, s.last);
      if (fec == 0) s.next++;
      if (fec >= fecmax) s.last = c;
      s.push_vertex(c, fec == 0 || fec >= fecmax);
      s.push_edge(c, b);
      s.push_edge(a, c);
      done = true;
    }
    if (done) continue;
    // No stored edge: first vertex new when a rotation allows it, the others new, from the FIFO or free.
    int r = 0;
    for (int k = 0; k < 3; k++)
      if (tri[k] == s.next) {
        r = k;
        break;
      }
    const uint32_t a = tri[r], b = tri[(r + 1) % 3], c = tri[(r + 2) % 3];
    const bool a_new = a == s.next;
    uint32_t next = s.next + (a_new ? 1 : 0);
    auto code_of = [&](uint32_t v, uint32_t& nx) {
      if (v == nx) {
        nx++;
        return 0;
      }
      const int f = s.find_vertex(v);
      if (f >= 0 && f + 1 <= 14) return f + 1;
      return 15;
    };
    int feb = code_of(b, next);
    int fec = code_of(c, next);
    if (!a_new && feb == 0 && fec == 0) {
      // An auxiliary byte of 0 in the slow path resets the decoder's next vertex: code the third as free.
      fec = 15;
      next--;
    }
    const uint8_t aux = static_cast<uint8_t>((feb << 4) | fec);
    int slot = -1;
    if (a_new && feb != 15 && fec != 15)
      for (int i = 0; i < 14; i++)
        if (kAuxTable[i] == aux) {
          slot = i;
          break;
        }
    if (slot >= 0) {
      codes.push_back(static_cast<uint8_t>(0xf0 | slot));
    } else {
      codes.push_back(a_new ? 0xfe : 0xff);
      data.push_back(aux);
      if (!a_new) write_index(data, a, s.last), s.last = a;
      if (feb == 15) write_index(data, b, s.last), s.last = b;
      if (fec == 15) write_index(data, c, s.last), s.last = c;
    }
    s.next = next;
    s.push_vertex(a);
    s.push_vertex(b, feb == 0 || feb == 15);
    s.push_vertex(c, fec == 0 || fec == 15);
    s.push_edge(b, a);
    s.push_edge(c, b);
    s.push_edge(a, c);
  }
  const size_t total = 1 + codes.size() + data.size() + 16;
  if (total > capacity) return 0;
  uint8_t* p = out;
  *p++ = 0xe1;
  std::memcpy(p, codes.data(), codes.size());
  p += codes.size();
  if (!data.empty()) std::memcpy(p, data.data(), data.size());
  p += data.size();
  std::memcpy(p, kAuxTable, 16);
  return total;
}

// 0 on success; -1 unsupported header, -2 truncated or malformed, -3 trailing bytes.
extern "C" int vgd_index_decode(uint32_t* out, size_t index_count, const uint8_t* buffer, size_t size) {
  if (index_count % 3) return -2;
  if (size < 1 + index_count / 3 + 16) return -2;
  if ((buffer[0] & 0xf0) != 0xe0) return -1;
  const int version = buffer[0] & 0x0f;
  if (version > 1) return -1;
  IndexState s;
  const int fecmax = version >= 1 ? 13 : 15;
  const uint8_t* code = buffer + 1;
  const uint8_t* data = code + index_count / 3;
  const uint8_t* safe_end = buffer + size - 16;
  const uint8_t* table = safe_end;
  for (size_t i = 0; i < index_count; i += 3) {
    if (data > safe_end) return -2;
    const uint8_t ct = *code++;
    uint32_t a, b, c;
    if (ct < 0xf0) {
      const int fe = ct >> 4, fec = ct & 15;
      a = s.edges[(s.eoff - 1 - fe) & 15][0];
      b = s.edges[(s.eoff - 1 - fe) & 15][1];
      if (fec < fecmax) {
        c = fec == 0 ? s.next : s.verts[(s.voff - 1 - fec) & 15];
        if (fec == 0) s.next++;
        s.push_vertex(c, fec == 0);
      } else {
        c = fec != 15 ? s.last + (fec == 13 ? -1u : 1u) : read_index(data, s.last);
        s.last = c;
        s.push_vertex(c);
      }
      s.push_edge(c, b);
      s.push_edge(a, c);
    } else if (ct < 0xfe) {
      const uint8_t aux = table[ct & 15];
      const int feb = aux >> 4, fec = aux & 15;
      a = s.next++;
      b = feb == 0 ? s.next : s.verts[(s.voff - feb) & 15];
      if (feb == 0) s.next++;
      c = fec == 0 ? s.next : s.verts[(s.voff - fec) & 15];
      if (fec == 0) s.next++;
      s.push_vertex(a);
      s.push_vertex(b, feb == 0);
      s.push_vertex(c, fec == 0);
      s.push_edge(b, a);
      s.push_edge(c, b);
      s.push_edge(a, c);
    } else {
      const uint8_t aux = *data++;
      const int fea = ct == 0xfe ? 0 : 15, feb = aux >> 4, fec = aux & 15;
      if (aux == 0) s.next = 0;
      a = fea == 0 ? s.next++ : 0;
      b = feb == 0 ? s.next++ : s.verts[(s.voff - feb) & 15];
      c = fec == 0 ? s.next++ : s.verts[(s.voff - fec) & 15];
      if (fea == 15) s.last = a = read_index(data, s.last);
      if (feb == 15) s.last = b = read_index(data, s.last);
      if (fec == 15) s.last = c = read_index(data, s.last);
      s.push_vertex(a);
      s.push_vertex(b, feb == 0 || feb == 15);
      s.push_vertex(c, fec == 0 || fec == 15);
      s.push_edge(b, a);
      s.push_edge(c, b);
      s.push_edge(a, c);
    }
    out[i] = a;
    out[i + 1] = b;
    out[i + 2] = c;
  }
  if (data != safe_end) return -3;
  return 0;
}
