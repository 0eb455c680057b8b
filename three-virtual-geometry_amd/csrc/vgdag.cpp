// Meshlet LOD DAG builder: the reference's buildVirtualMesh (src/core/preprocess/buildVirtualMesh.ts:185-601)
// and partitionMeshlets (partition.ts:16-136), without meshoptimizer. Host-only C++.
//
// The level loop, the tail, error accumulation, bound spheres, pack() and the partitioner follow the
// reference step by step (cited per function). Where the reference calls meshoptimizer the method is ours:
//  - meshlets (meshopt buildMeshlets, <= 128 vertices and <= 128 triangles): greedy growth from a seed
//    triangle, taking at each step the adjacent triangle that adds the fewest new vertices, then the one
//    whose least-used vertex has the fewest free triangles left, then the one nearest the meshlet's
//    centroid; the next seed is the neighbouring free triangle with the fewest free neighbours, so the
//    remaining region is eaten from its edges; fragments left over are merged into neighbours with room;
//  - simplification (meshopt simplify, LockBorder | ErrorAbsolute): quadric-error half-edge collapses onto
//    existing vertices (so every level indexes the input vertex buffer), area-weighted plane quadrics plus
//    edge planes along open borders, a link-condition and normal-flip check per collapse, vertices shared by
//    several positions' copies (seams) and non-manifold vertices locked, border vertices locked when the
//    border is (groups) and otherwise collapsed only along the border (tail). The returned error is the
//    square root of the largest collapse's mean squared plane distance, in object units;
//  - the tail's fallback (meshopt simplifySloppy): vertex clustering on the finest grid that reaches the
//    target, each cluster snapped to its vertex nearest the cluster mean; error = largest snap distance;
//  - sphere bounds (meshopt computeSphereBounds): an axis-extremal pair grown over every sphere, then made
//    exact by the reference's containment pass (buildVirtualMesh.ts:746-762).
#include "../../include/vgdag.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <queue>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

constexpr uint32_t kMaxVertices = 128;   // MAX_MESHLET_VERTICES, src/core/constants.ts:5
constexpr uint32_t kMaxTriangles = 128;  // MAX_MESHLET_TRIANGLES, src/core/constants.ts:2
constexpr double kErrorInfinity = 1e30;  // ERROR_INFINITY, src/core/constants.ts:8
constexpr size_t kTailStartTriangles = 1024;  // TAIL_START_TRIANGLES, buildVirtualMesh.ts:176

struct Sphere {
  double x = 0, y = 0, z = 0, r = 0;
};

struct Meshlet {
  std::vector<uint32_t> indices;   // global vertex ids, 3 per triangle
  std::vector<uint32_t> vertices;  // unique global ids in meshlet-local order
  std::vector<uint8_t> local;      // 3 per triangle: positions in `vertices`
  std::vector<uint64_t> boundary;  // boundary edge keys over welded ids, first-seen order
  uint32_t level = 0;
  Sphere lod, parent, cull;
  double error = 0, parent_error = kErrorInfinity;
  int32_t replacements = -1;  // list of the meshlets that replace it
  int32_t own_group = -1;     // the group whose simplification produced it (-1: leaf)
  int32_t parent_group = -1;  // the group it was merged into (-1: none, a root)
};

// ------------------------------------------------------------------ small helpers

template <typename F>
void parallel_for(size_t n, unsigned threads, F&& fn) {
  if (n == 0) return;
  threads = static_cast<unsigned>(std::min<size_t>(threads, n));
  if (threads <= 1) {
    for (size_t i = 0; i < n; i++) fn(i, 0u);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < threads; t++)
    pool.emplace_back([&, t] {
      for (size_t i = next++; i < n; i = next++) fn(i, t);
    });
  for (auto& th : pool) th.join();
}

// Order-independent edge key (buildVirtualMesh.ts:684 edgeKey): vertex ids stay below 2^26.
inline uint64_t edge_key(uint32_t a, uint32_t b) {
  return a < b ? (static_cast<uint64_t>(a) << 26) + b : (static_cast<uint64_t>(b) << 26) + a;
}

// A triangle subset re-indexed into a compact local vertex buffer (buildVirtualMesh.ts:637-661 compact).
struct Local {
  std::vector<uint32_t> indices;
  std::vector<float> pos;
  std::vector<uint32_t> to_global;
};

// `scratch` holds -1 for every vertex between calls.
Local compact(const uint32_t* idx, size_t n, const float* positions, std::vector<int32_t>& scratch) {
  Local l;
  l.indices.resize(n);
  for (size_t i = 0; i < n; i++) {
    const uint32_t g = idx[i];
    int32_t s = scratch[g];
    if (s < 0) {
      s = static_cast<int32_t>(l.to_global.size());
      scratch[g] = s;
      l.to_global.push_back(g);
    }
    l.indices[i] = static_cast<uint32_t>(s);
  }
  l.pos.resize(l.to_global.size() * 3);
  for (size_t i = 0; i < l.to_global.size(); i++) {
    const uint32_t g = l.to_global[i];
    scratch[g] = -1;
    std::memcpy(&l.pos[i * 3], positions + static_cast<size_t>(g) * 3, 3 * sizeof(float));
  }
  return l;
}

// AABB-centred sphere over the vertices of a triangle list (buildVirtualMesh.ts:704-731 boundingSphere).
Sphere bounding_sphere(const std::vector<uint32_t>& idx, const float* positions) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i : idx)
    for (int k = 0; k < 3; k++) {
      const double v = positions[static_cast<size_t>(i) * 3 + k];
      lo[k] = std::min(lo[k], v);
      hi[k] = std::max(hi[k], v);
    }
  Sphere s;
  s.x = (lo[0] + hi[0]) / 2;
  s.y = (lo[1] + hi[1]) / 2;
  s.z = (lo[2] + hi[2]) / 2;
  double r2 = 0;
  for (uint32_t i : idx) {
    const double dx = positions[static_cast<size_t>(i) * 3] - s.x, dy = positions[static_cast<size_t>(i) * 3 + 1] - s.y,
                 dz = positions[static_cast<size_t>(i) * 3 + 2] - s.z;
    r2 = std::max(r2, dx * dx + dy * dy + dz * dz);
  }
  s.r = std::sqrt(r2);
  return s;
}

// Sphere enclosing spheres (buildVirtualMesh.ts:733-762 enclosingSphere): one sphere is returned as is;
// otherwise a centre from our sphere-bounds estimate, then the reference's exact containment radius,
// padded by 1.0001.
Sphere enclosing(const std::vector<const Sphere*>& ss) {
  if (ss.size() == 1) return *ss[0];
  // Estimate: the two spheres extremal along the axis of largest spread, then grow over the rest.
  size_t lo_i[3] = {0, 0, 0}, hi_i[3] = {0, 0, 0};
  auto comp = [](const Sphere& s, int k) { return k == 0 ? s.x : k == 1 ? s.y : s.z; };
  for (size_t i = 1; i < ss.size(); i++)
    for (int k = 0; k < 3; k++) {
      if (comp(*ss[i], k) - ss[i]->r < comp(*ss[lo_i[k]], k) - ss[lo_i[k]]->r) lo_i[k] = i;
      if (comp(*ss[i], k) + ss[i]->r > comp(*ss[hi_i[k]], k) + ss[hi_i[k]]->r) hi_i[k] = i;
    }
  int axis = 0;
  double spread = -1;
  for (int k = 0; k < 3; k++) {
    const double d = (comp(*ss[hi_i[k]], k) + ss[hi_i[k]]->r) - (comp(*ss[lo_i[k]], k) - ss[lo_i[k]]->r);
    if (d > spread) {
      spread = d;
      axis = k;
    }
  }
  auto merge = [](Sphere a, const Sphere& b) {
    const double dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
    if (d + b.r <= a.r) return a;
    if (d + a.r <= b.r) return b;
    const double r = (d + a.r + b.r) / 2;
    const double t = d > 0 ? (r - a.r) / d : 0;
    return Sphere{a.x + dx * t, a.y + dy * t, a.z + dz * t, r};
  };
  Sphere c = merge(*ss[lo_i[axis]], *ss[hi_i[axis]]);
  for (const Sphere* s : ss) c = merge(c, *s);
  double radius = c.r;
  for (const Sphere* s : ss) radius = std::max(radius, std::hypot(s->x - c.x, s->y - c.y, s->z - c.z) + s->r);
  return Sphere{c.x, c.y, c.z, radius * 1.0001};
}

// How far the simplified triangles' bounding box retreats from the original's, per side
// (buildVirtualMesh.ts:603-629 extentShrink).
double extent_shrink(const std::vector<uint32_t>& before, const std::vector<uint32_t>& after, const float* positions) {
  auto box = [&](const std::vector<uint32_t>& idx, double* lo, double* hi) {
    for (int k = 0; k < 3; k++) {
      lo[k] = INFINITY;
      hi[k] = -INFINITY;
    }
    for (uint32_t i : idx)
      for (int k = 0; k < 3; k++) {
        const double v = positions[static_cast<size_t>(i) * 3 + k];
        lo[k] = std::min(lo[k], v);
        hi[k] = std::max(hi[k], v);
      }
  };
  double alo[3], ahi[3], blo[3], bhi[3];
  box(before, alo, ahi);
  if (after.empty()) return std::max(ahi[0] - alo[0], std::max(ahi[1] - alo[1], ahi[2] - alo[2])) / 2;
  box(after, blo, bhi);
  double shrink = 0;
  for (int k = 0; k < 3; k++) shrink = std::max(shrink, std::max(blo[k] - alo[k], ahi[k] - bhi[k]));
  return shrink;
}

// Edges used by exactly one triangle, keyed by welded ids, in first-seen order (buildVirtualMesh.ts:663-698).
std::vector<uint64_t> boundary_edges(const std::vector<uint32_t>& idx, const uint32_t* weld) {
  std::vector<std::pair<uint64_t, uint32_t>> e;  // (key, first-seen position)
  e.reserve(idx.size());
  for (size_t t = 0; t < idx.size(); t += 3)
    for (int k = 0; k < 3; k++)
      e.emplace_back(edge_key(weld[idx[t + k]], weld[idx[t + (k + 1) % 3]]), static_cast<uint32_t>(t + k));
  std::stable_sort(e.begin(), e.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  std::vector<std::pair<uint32_t, uint64_t>> once;
  for (size_t i = 0; i < e.size();) {
    size_t j = i;
    while (j < e.size() && e[j].first == e[i].first) j++;
    if (j - i == 1) once.emplace_back(e[i].second, e[i].first);
    i = j;
  }
  std::sort(once.begin(), once.end());
  std::vector<uint64_t> out(once.size());
  for (size_t i = 0; i < once.size(); i++) out[i] = once[i].second;
  return out;
}

// Vertex -> first vertex with the same position bits (meshopt generatePositionRemap's contract).
std::vector<uint32_t> position_remap(const float* pos, uint32_t n) {
  struct Key {
    uint32_t b[3];
    bool operator==(const Key& o) const { return b[0] == o.b[0] && b[1] == o.b[1] && b[2] == o.b[2]; }
  };
  struct Hash {
    size_t operator()(const Key& k) const {
      uint64_t h = k.b[0] * 0x9E3779B97F4A7C15ull;
      h ^= (k.b[1] + 0x7F4A7C15ull) * 0xC2B2AE3D27D4EB4Full;
      h ^= (k.b[2] + 0x165667B1ull) * 0x165667B19E3779F9ull;
      return static_cast<size_t>(h ^ (h >> 29));
    }
  };
  std::unordered_map<Key, uint32_t, Hash> first;
  first.reserve(n * 2);
  std::vector<uint32_t> remap(n);
  for (uint32_t v = 0; v < n; v++) {
    Key k;
    std::memcpy(k.b, pos + static_cast<size_t>(v) * 3, 12);
    auto it = first.emplace(k, v).first;
    remap[v] = it->second;
  }
  return remap;
}

// ------------------------------------------------------------------ meshlets

struct Built {
  std::vector<uint32_t> verts;  // local ids
  std::vector<uint8_t> tris;    // 3 per triangle, into verts
  std::vector<uint32_t> ids;    // the input triangles, in order
};

// Growth can strand small islands of free triangles that become meshlets of a few triangles; groups of
// such fragments cannot simplify under locked borders. Each meshlet under a quarter of the triangle limit,
// smallest first, moves into the vertex-adjacent meshlet with room that shares the most vertices with it.
void merge_fragments(const std::vector<uint32_t>& idx, const std::vector<uint32_t>& off, const std::vector<uint32_t>& adj,
                     std::vector<Built>& ms) {
  const uint32_t nt = static_cast<uint32_t>(idx.size() / 3);
  std::vector<uint32_t> owner(nt);
  for (size_t m = 0; m < ms.size(); m++)
    for (uint32_t t : ms[m].ids) owner[t] = static_cast<uint32_t>(m);
  const uint32_t small = kMaxTriangles / 4;
  std::vector<uint32_t> order(ms.size());
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return ms[a].ids.size() < ms[b].ids.size(); });
  std::vector<uint8_t> gone(ms.size(), 0);
  std::vector<std::pair<uint32_t, uint32_t>> nb;  // (meshlet, shared vertex count)
  for (uint32_t m : order) {
    if (gone[m] || ms[m].ids.empty() || ms[m].ids.size() >= small) continue;
    nb.clear();
    for (uint32_t v : ms[m].verts)
      for (uint32_t x = off[v]; x < off[v + 1]; x++) {
        const uint32_t o = owner[adj[x]];
        if (o == m) continue;
        auto it = std::find_if(nb.begin(), nb.end(), [&](const auto& p) { return p.first == o; });
        if (it == nb.end()) nb.emplace_back(o, 1u);
        else it->second++;
      }
    std::stable_sort(nb.begin(), nb.end(), [](const auto& a, const auto& b) { return a.second > b.second; });
    for (const auto& [o, shared_count] : nb) {
      (void)shared_count;
      if (ms[o].ids.size() + ms[m].ids.size() > kMaxTriangles) continue;
      size_t fresh = 0;
      for (uint32_t v : ms[m].verts)
        if (std::find(ms[o].verts.begin(), ms[o].verts.end(), v) == ms[o].verts.end()) fresh++;
      if (ms[o].verts.size() + fresh > kMaxVertices) continue;
      for (uint32_t t : ms[m].ids) {
        for (int k = 0; k < 3; k++) {
          const uint32_t v = idx[t * 3 + k];
          auto it = std::find(ms[o].verts.begin(), ms[o].verts.end(), v);
          if (it == ms[o].verts.end()) {
            ms[o].verts.push_back(v);
            it = ms[o].verts.end() - 1;
          }
          ms[o].tris.push_back(static_cast<uint8_t>(it - ms[o].verts.begin()));
        }
        owner[t] = o;
        ms[o].ids.push_back(t);
      }
      ms[m].ids.clear();
      gone[m] = 1;
      break;
    }
  }
  size_t w = 0;
  for (size_t m = 0; m < ms.size(); m++) {
    if (gone[m]) continue;
    if (w != m) ms[w] = std::move(ms[m]);
    w++;
  }
  ms.resize(w);
}

std::vector<Built> build_meshlets(const std::vector<uint32_t>& idx, const std::vector<float>& pos) {
  const uint32_t nv = static_cast<uint32_t>(pos.size() / 3);
  const uint32_t nt = static_cast<uint32_t>(idx.size() / 3);
  std::vector<uint32_t> off(nv + 1, 0), adj(idx.size());
  for (uint32_t i : idx) off[i + 1]++;
  for (uint32_t v = 0; v < nv; v++) off[v + 1] += off[v];
  {
    std::vector<uint32_t> fill(off.begin(), off.end() - 1);
    for (uint32_t t = 0; t < nt; t++)
      for (int k = 0; k < 3; k++) adj[fill[idx[t * 3 + k]]++] = t;
  }
  std::vector<uint32_t> live(nv);
  for (uint32_t v = 0; v < nv; v++) live[v] = off[v + 1] - off[v];
  std::vector<float> cen(static_cast<size_t>(nt) * 3);
  for (uint32_t t = 0; t < nt; t++)
    for (int k = 0; k < 3; k++)
      cen[t * 3 + k] = (pos[idx[t * 3] * 3 + k] + pos[idx[t * 3 + 1] * 3 + k] + pos[idx[t * 3 + 2] * 3 + k]) / 3.0f;
  std::vector<uint8_t> used(nt, 0);
  std::vector<uint8_t> slot(nv, 0xff);
  std::vector<uint32_t> stamp(nt, ~0u);
  std::vector<uint32_t> cand;
  std::vector<Built> out;
  uint32_t scan = 0;
  double last[3] = {0, 0, 0};
  auto dist2 = [&](uint32_t t, const double* c) {
    const double dx = cen[t * 3] - c[0], dy = cen[t * 3 + 1] - c[1], dz = cen[t * 3 + 2] - c[2];
    return dx * dx + dy * dy + dz * dz;
  };
  for (uint32_t meshlet = 0;; meshlet++) {
    // Seed: the free neighbour of the last meshlet with the fewest free neighbours, else the next free triangle.
    int64_t seed = -1;
    uint32_t best_live = ~0u;
    double best_d = INFINITY;
    for (uint32_t t : cand) {
      if (used[t]) continue;
      const uint32_t l = live[idx[t * 3]] + live[idx[t * 3 + 1]] + live[idx[t * 3 + 2]];
      const double d = dist2(t, last);
      if (l < best_live || (l == best_live && d < best_d)) {
        seed = t;
        best_live = l;
        best_d = d;
      }
    }
    if (seed < 0) {
      while (scan < nt && used[scan]) scan++;
      if (scan == nt) break;
      seed = scan;
    }
    cand.clear();
    Built b;
    double sum[3] = {0, 0, 0};
    uint32_t count = 0;
    auto add = [&](uint32_t t) {
      used[t] = 1;
      for (int k = 0; k < 3; k++) {
        const uint32_t v = idx[t * 3 + k];
        live[v]--;
        if (slot[v] == 0xff) {
          slot[v] = static_cast<uint8_t>(b.verts.size());
          b.verts.push_back(v);
          for (uint32_t a = off[v]; a < off[v + 1]; a++) {
            const uint32_t n = adj[a];
            if (!used[n] && stamp[n] != meshlet) {
              stamp[n] = meshlet;
              cand.push_back(n);
            }
          }
        }
        b.tris.push_back(slot[v]);
      }
      for (int k = 0; k < 3; k++) sum[k] += cen[t * 3 + k];
      b.ids.push_back(t);
      count++;
    };
    add(static_cast<uint32_t>(seed));
    while (count < kMaxTriangles) {
      const double c[3] = {sum[0] / count, sum[1] / count, sum[2] / count};
      int64_t best = -1;
      uint32_t best_new = 4, best_closing = ~0u;
      double bd = INFINITY;
      size_t w = 0;
      for (size_t i = 0; i < cand.size(); i++) {
        const uint32_t t = cand[i];
        if (used[t]) continue;
        cand[w++] = t;
        const uint32_t a = idx[t * 3], bb = idx[t * 3 + 1], cc = idx[t * 3 + 2];
        uint32_t fresh = (slot[a] == 0xff) + (slot[bb] == 0xff && bb != a) + (slot[cc] == 0xff && cc != a && cc != bb);
        if (b.verts.size() + fresh > kMaxVertices) continue;
        // Fewest new vertices; then the triangle whose least-used vertex has the fewest free triangles
        // left (closing vertices off instead of stranding islands); then the nearest to the centroid.
        const uint32_t closing = std::min(live[a], std::min(live[bb], live[cc]));
        const double d = dist2(t, c);
        if (fresh < best_new || (fresh == best_new && (closing < best_closing || (closing == best_closing && d < bd)))) {
          best = t;
          best_new = fresh;
          best_closing = closing;
          bd = d;
        }
      }
      cand.resize(w);
      if (best < 0) break;
      add(static_cast<uint32_t>(best));
    }
    for (uint32_t v : b.verts) slot[v] = 0xff;
    for (int k = 0; k < 3; k++) last[k] = sum[k] / count;
    out.push_back(std::move(b));
  }
  merge_fragments(idx, off, adj, out);
  return out;
}

// ------------------------------------------------------------------ simplification

struct Quadric {
  double a2 = 0, b2 = 0, c2 = 0, ab = 0, ac = 0, bc = 0, ad = 0, bd = 0, cd = 0, d2 = 0, w = 0;
  void add_plane(double a, double b, double c, double d, double weight) {
    a2 += weight * a * a;
    b2 += weight * b * b;
    c2 += weight * c * c;
    ab += weight * a * b;
    ac += weight * a * c;
    bc += weight * b * c;
    ad += weight * a * d;
    bd += weight * b * d;
    cd += weight * c * d;
    d2 += weight * d * d;
    w += weight;
  }
  void add(const Quadric& q) {
    a2 += q.a2;
    b2 += q.b2;
    c2 += q.c2;
    ab += q.ab;
    ac += q.ac;
    bc += q.bc;
    ad += q.ad;
    bd += q.bd;
    cd += q.cd;
    d2 += q.d2;
    w += q.w;
  }
  double eval(const double* p) const {
    const double x = p[0], y = p[1], z = p[2];
    const double r = a2 * x * x + b2 * y * y + c2 * z * z + 2 * (ab * x * y + ac * x * z + bc * y * z) +
                     2 * (ad * x + bd * y + cd * z) + d2;
    return r > 0 ? r : 0;
  }
};

struct Simplified {
  std::vector<uint32_t> indices;
  double error = 0;
};

enum : uint8_t { kFree = 0, kBorder = 1, kLocked = 2 };

// Quadric half-edge collapses down to `target` indices (see the file comment). Indices are local ids.
// `cls` maps each vertex to the first vertex with the same position.
Simplified simplify(const std::vector<uint32_t>& in, const std::vector<float>& pos_f, const std::vector<uint32_t>& cls,
                    size_t target, bool lock_border, bool prune) {
  const uint32_t nv = static_cast<uint32_t>(pos_f.size() / 3);
  Simplified res;
  std::vector<uint8_t> seam(nv, 0);
  for (uint32_t v = 0; v < nv; v++)
    if (cls[v] != v) seam[v] = seam[cls[v]] = 1;
  // Triangles with repeated positions have no area: dropped.
  std::vector<uint32_t> tri;
  tri.reserve(in.size());
  for (size_t t = 0; t + 2 < in.size(); t += 3) {
    const uint32_t a = in[t], b = in[t + 1], c = in[t + 2];
    if (cls[a] == cls[b] || cls[b] == cls[c] || cls[a] == cls[c]) continue;
    tri.insert(tri.end(), {a, b, c});
  }
  const uint32_t ntri = static_cast<uint32_t>(tri.size() / 3);
  if (tri.size() <= target) {
    res.indices = tri;
    return res;
  }
  // Normalised positions.
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t v = 0; v < nv; v++)
    for (int k = 0; k < 3; k++) {
      lo[k] = std::min(lo[k], static_cast<double>(pos_f[v * 3 + k]));
      hi[k] = std::max(hi[k], static_cast<double>(pos_f[v * 3 + k]));
    }
  const double extent = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], std::max(hi[2] - lo[2], 1e-30)));
  std::vector<double> p(static_cast<size_t>(nv) * 3);
  for (uint32_t v = 0; v < nv; v++)
    for (int k = 0; k < 3; k++) p[v * 3 + k] = (pos_f[v * 3 + k] - lo[k]) / extent;

  // Edge use counts over position classes, per triangle corner edge: 1 = border, > 2 = non-manifold.
  std::vector<uint8_t> uses(tri.size());
  {
    std::vector<std::pair<uint64_t, uint32_t>> e(tri.size());
    for (uint32_t t = 0; t < ntri; t++)
      for (int k = 0; k < 3; k++) e[t * 3 + k] = {edge_key(cls[tri[t * 3 + k]], cls[tri[t * 3 + (k + 1) % 3]]), t * 3 + k};
    std::sort(e.begin(), e.end());
    for (size_t i = 0; i < e.size();) {
      size_t j = i;
      while (j < e.size() && e[j].first == e[i].first) j++;
      for (size_t x = i; x < j; x++) uses[e[x].second] = static_cast<uint8_t>(std::min<size_t>(255, j - i));
      i = j;
    }
  }
  std::vector<uint8_t> kind(nv, kFree);
  std::vector<uint8_t> border_edges(nv, 0);
  for (uint32_t t = 0; t < ntri; t++)
    for (int k = 0; k < 3; k++) {
      const uint32_t a = tri[t * 3 + k], b = tri[t * 3 + (k + 1) % 3];
      const uint32_t u = uses[t * 3 + k];
      if (u > 2) kind[a] = kind[b] = kLocked;
      if (u == 1) {
        border_edges[a] = static_cast<uint8_t>(std::min(255, border_edges[a] + 1));
        border_edges[b] = static_cast<uint8_t>(std::min(255, border_edges[b] + 1));
      }
    }
  for (uint32_t v = 0; v < nv; v++) {
    if (seam[v]) kind[v] = kLocked;
    if (kind[v] == kLocked) continue;
    if (border_edges[v]) kind[v] = (lock_border || border_edges[v] != 2) ? kLocked : kBorder;
  }

  // Quadrics: triangle planes weighted by area, plus planes through border edges normal to their triangle.
  std::vector<Quadric> q(nv);
  for (uint32_t t = 0; t < ntri; t++) {
    const double* A = &p[tri[t * 3] * 3];
    const double* B = &p[tri[t * 3 + 1] * 3];
    const double* C = &p[tri[t * 3 + 2] * 3];
    const double e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]}, e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
    double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (len > 0) {
      for (double& x : n) x /= len;
      const double d = -(n[0] * A[0] + n[1] * A[1] + n[2] * A[2]);
      for (int k = 0; k < 3; k++) q[tri[t * 3 + k]].add_plane(n[0], n[1], n[2], d, len * 0.5);
    }
    for (int k = 0; k < 3; k++) {
      const uint32_t a = tri[t * 3 + k], b = tri[t * 3 + (k + 1) % 3];
      if (uses[t * 3 + k] != 1 || len == 0) continue;
      const double* P = &p[a * 3];
      const double* Q = &p[b * 3];
      const double e[3] = {Q[0] - P[0], Q[1] - P[1], Q[2] - P[2]};
      double m[3] = {e[1] * n[2] - e[2] * n[1], e[2] * n[0] - e[0] * n[2], e[0] * n[1] - e[1] * n[0]};
      const double ml = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
      if (ml == 0) continue;
      for (double& x : m) x /= ml;
      const double d = -(m[0] * P[0] + m[1] * P[1] + m[2] * P[2]);
      const double weight = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
      q[a].add_plane(m[0], m[1], m[2], d, weight);
      q[b].add_plane(m[0], m[1], m[2], d, weight);
    }
  }

  // Vertex -> live triangles.
  std::vector<std::vector<uint32_t>> vt(nv);
  for (uint32_t t = 0; t < ntri; t++)
    for (int k = 0; k < 3; k++) vt[tri[t * 3 + k]].push_back(t);
  std::vector<uint8_t> tri_live(ntri, 1);
  std::vector<uint8_t> dead(nv, 0);
  size_t live_tris = ntri;
  double max_cost = 0;

  // Components with no locked vertex, smallest first, for pruning.
  struct Component {
    double radius2;
    std::vector<uint32_t> tris;
  };
  std::vector<Component> comps;
  size_t next_comp = 0;
  if (prune) {
    std::vector<uint32_t> parent(nv);
    std::iota(parent.begin(), parent.end(), 0u);
    auto find = [&](uint32_t x) {
      while (parent[x] != x) x = parent[x] = parent[parent[x]];
      return x;
    };
    for (uint32_t t = 0; t < ntri; t++)
      for (int k = 1; k < 3; k++) {
        const uint32_t a = find(cls[tri[t * 3]]), b = find(cls[tri[t * 3 + k]]);
        if (a != b) parent[a] = b;
      }
    std::unordered_map<uint32_t, size_t> of;
    std::vector<uint8_t> pinned;
    for (uint32_t t = 0; t < ntri; t++) {
      const uint32_t r = find(cls[tri[t * 3]]);
      auto it = of.emplace(r, comps.size()).first;
      if (it->second == comps.size()) {
        comps.push_back({0, {}});
        pinned.push_back(0);
      }
      comps[it->second].tris.push_back(t);
      for (int k = 0; k < 3; k++)
        if (kind[tri[t * 3 + k]] == kLocked) pinned[it->second] = 1;
    }
    std::vector<Component> free_comps;
    for (size_t c = 0; c < comps.size(); c++) {
      if (pinned[c]) continue;
      double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (uint32_t t : comps[c].tris)
        for (int k = 0; k < 3; k++)
          for (int j = 0; j < 3; j++) {
            clo[j] = std::min(clo[j], p[tri[t * 3 + k] * 3 + j]);
            chi[j] = std::max(chi[j], p[tri[t * 3 + k] * 3 + j]);
          }
      const double dx = chi[0] - clo[0], dy = chi[1] - clo[1], dz = chi[2] - clo[2];
      comps[c].radius2 = (dx * dx + dy * dy + dz * dz) / 4;
      free_comps.push_back(std::move(comps[c]));
    }
    std::stable_sort(free_comps.begin(), free_comps.end(),
                     [](const Component& a, const Component& b) { return a.radius2 < b.radius2; });
    comps = std::move(free_comps);
  }
  auto prune_next = [&]() {
    for (uint32_t t : comps[next_comp].tris)
      if (tri_live[t]) {
        tri_live[t] = 0;
        live_tris--;
      }
    for (uint32_t t : comps[next_comp].tris)
      for (int k = 0; k < 3; k++) dead[tri[t * 3 + k]] = 1;
    max_cost = std::max(max_cost, comps[next_comp].radius2);
    next_comp++;
  };

  auto cost = [&](uint32_t u, uint32_t v) {
    Quadric m = q[u];
    m.add(q[v]);
    return m.w > 0 ? m.eval(&p[v * 3]) / m.w : 0.0;
  };
  // Live triangles containing both u and v.
  auto shared = [&](uint32_t u, uint32_t v) {
    uint32_t n = 0;
    for (uint32_t t : vt[u])
      if (tri[t * 3] == v || tri[t * 3 + 1] == v || tri[t * 3 + 2] == v) n++;
    return n;
  };
  auto allowed = [&](uint32_t u, uint32_t v) {
    if (dead[u] || dead[v] || kind[u] == kLocked) return false;
    if (kind[u] == kBorder) return shared(u, v) == 1;
    return true;
  };
  struct Entry {
    double cost;
    uint32_t u, v;
    uint32_t su, sv;
    bool operator<(const Entry& o) const { return cost > o.cost; }
  };
  std::vector<uint32_t> stamp(nv, 0);
  std::priority_queue<Entry> heap;
  std::vector<uint32_t> ring_u, ring_v;
  auto ring = [&](uint32_t u, std::vector<uint32_t>& out) {
    out.clear();
    for (uint32_t t : vt[u])
      for (int k = 0; k < 3; k++) {
        const uint32_t w = tri[t * 3 + k];
        if (w != u && std::find(out.begin(), out.end(), w) == out.end()) out.push_back(w);
      }
  };
  auto push_edges = [&](uint32_t u) {
    std::vector<uint32_t> r;
    ring(u, r);
    for (uint32_t w : r) {
      if (allowed(u, w)) heap.push({cost(u, w), u, w, stamp[u], stamp[w]});
      if (allowed(w, u)) heap.push({cost(w, u), w, u, stamp[w], stamp[u]});
    }
  };
  // The collapse u -> v keeps the mesh a manifold without fold-overs.
  auto valid = [&](uint32_t u, uint32_t v) {
    ring(u, ring_u);
    ring(v, ring_v);
    uint32_t common = 0;
    for (uint32_t w : ring_u)
      for (uint32_t x : ring_v)
        if (cls[w] == cls[x]) common++;
    const uint32_t s = shared(u, v);
    if (s == 0 || common != s) return false;
    const double* P = &p[v * 3];
    for (uint32_t t : vt[u]) {
      const uint32_t a = tri[t * 3], b = tri[t * 3 + 1], c = tri[t * 3 + 2];
      if (a == v || b == v || c == v) continue;
      uint32_t na = a == u ? v : a, nb = b == u ? v : b, nc = c == u ? v : c;
      if (cls[na] == cls[nb] || cls[nb] == cls[nc] || cls[na] == cls[nc]) return false;
      auto normal = [&](uint32_t x, uint32_t y, uint32_t z, const double* over, uint32_t at, double* n) {
        const double* X = x == at ? over : &p[x * 3];
        const double* Y = y == at ? over : &p[y * 3];
        const double* Z = z == at ? over : &p[z * 3];
        const double e1[3] = {Y[0] - X[0], Y[1] - X[1], Y[2] - X[2]}, e2[3] = {Z[0] - X[0], Z[1] - X[1], Z[2] - X[2]};
        n[0] = e1[1] * e2[2] - e1[2] * e2[1];
        n[1] = e1[2] * e2[0] - e1[0] * e2[2];
        n[2] = e1[0] * e2[1] - e1[1] * e2[0];
      };
      double n0[3], n1[3];
      normal(a, b, c, nullptr, ~0u, n0);
      normal(a, b, c, P, u, n1);
      if (n0[0] * n1[0] + n0[1] * n1[1] + n0[2] * n1[2] <= 0) return false;
    }
    return true;
  };
  auto collapse = [&](uint32_t u, uint32_t v) {
    for (uint32_t t : vt[u]) {
      if (!tri_live[t]) continue;
      uint32_t* T = &tri[t * 3];
      if (T[0] == v || T[1] == v || T[2] == v) {
        tri_live[t] = 0;
        live_tris--;
        for (int k = 0; k < 3; k++) {
          const uint32_t w = T[k];
          if (w == u) continue;
          auto& l = vt[w];
          l.erase(std::remove(l.begin(), l.end(), t), l.end());
        }
      } else {
        for (int k = 0; k < 3; k++)
          if (T[k] == u) T[k] = v;
        vt[v].push_back(t);
      }
    }
    vt[u].clear();
    dead[u] = 1;
    q[v].add(q[u]);
    stamp[v]++;
    push_edges(v);
  };

  for (int pass = 0; pass < 8 && live_tris * 3 > target; pass++) {
    heap = {};
    for (uint32_t v = 0; v < nv; v++)
      if (!dead[v] && !vt[v].empty()) {
        std::vector<uint32_t> r;
        ring(v, r);
        for (uint32_t w : r)
          if (allowed(v, w)) heap.push({cost(v, w), v, w, stamp[v], stamp[w]});
      }
    size_t done = 0;
    while (live_tris * 3 > target && !heap.empty()) {
      const Entry e = heap.top();
      if (next_comp < comps.size() && comps[next_comp].radius2 <= e.cost) {
        prune_next();
        continue;
      }
      heap.pop();
      if (dead[e.u] || dead[e.v] || stamp[e.u] != e.su || stamp[e.v] != e.sv) continue;
      if (!allowed(e.u, e.v) || !valid(e.u, e.v)) continue;
      max_cost = std::max(max_cost, e.cost);
      collapse(e.u, e.v);
      done++;
    }
    if (done == 0) break;
  }
  while (live_tris * 3 > target && next_comp < comps.size()) prune_next();

  for (uint32_t t = 0; t < ntri; t++)
    if (tri_live[t]) res.indices.insert(res.indices.end(), {tri[t * 3], tri[t * 3 + 1], tri[t * 3 + 2]});
  res.error = std::sqrt(max_cost) * extent;
  return res;
}

// Vertex clustering on the finest grid that reaches `target` indices (see the file comment).
Simplified simplify_sloppy(const std::vector<uint32_t>& in, const std::vector<float>& pos, size_t target) {
  const uint32_t nv = static_cast<uint32_t>(pos.size() / 3);
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t v = 0; v < nv; v++)
    for (int k = 0; k < 3; k++) {
      lo[k] = std::min(lo[k], static_cast<double>(pos[v * 3 + k]));
      hi[k] = std::max(hi[k], static_cast<double>(pos[v * 3 + k]));
    }
  const double extent = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], std::max(hi[2] - lo[2], 1e-30)));
  auto run = [&](uint32_t grid, Simplified* out) {
    std::unordered_map<uint64_t, uint32_t> cell_of;
    std::vector<uint64_t> key(nv);
    std::vector<double> sum;
    std::vector<uint32_t> cnt;
    std::vector<uint32_t> cell(nv);
    for (uint32_t v = 0; v < nv; v++) {
      uint64_t c[3];
      for (int k = 0; k < 3; k++)
        c[k] = static_cast<uint64_t>(std::min<double>(grid - 1, std::floor((pos[v * 3 + k] - lo[k]) / extent * grid)));
      key[v] = (c[0] * grid + c[1]) * grid + c[2];
      auto it = cell_of.emplace(key[v], static_cast<uint32_t>(cnt.size())).first;
      if (it->second == cnt.size()) {
        cnt.push_back(0);
        sum.insert(sum.end(), {0.0, 0.0, 0.0});
      }
      cell[v] = it->second;
      cnt[cell[v]]++;
      for (int k = 0; k < 3; k++) sum[cell[v] * 3 + k] += pos[v * 3 + k];
    }
    std::vector<uint32_t> rep(cnt.size(), ~0u);
    std::vector<double> best(cnt.size(), INFINITY);
    for (uint32_t v = 0; v < nv; v++) {
      const uint32_t c = cell[v];
      double d = 0;
      for (int k = 0; k < 3; k++) {
        const double x = pos[v * 3 + k] - sum[c * 3 + k] / cnt[c];
        d += x * x;
      }
      if (d < best[c]) {
        best[c] = d;
        rep[c] = v;
      }
    }
    size_t n = 0;
    std::vector<uint32_t> outi;
    for (size_t t = 0; t + 2 < in.size(); t += 3) {
      const uint32_t a = rep[cell[in[t]]], b = rep[cell[in[t + 1]]], c = rep[cell[in[t + 2]]];
      if (a == b || b == c || a == c) continue;
      outi.insert(outi.end(), {a, b, c});
      n++;
    }
    // Duplicate triangles (same three clusters) are kept once.
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), size_t{0});
    auto sorted = [&](size_t i) {
      std::array<uint32_t, 3> s{outi[i * 3], outi[i * 3 + 1], outi[i * 3 + 2]};
      std::sort(s.begin(), s.end());
      return s;
    };
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return sorted(a) < sorted(b); });
    std::vector<uint8_t> keep(n, 1);
    for (size_t i = 1; i < n; i++)
      if (sorted(order[i]) == sorted(order[i - 1])) keep[order[i]] = 0;
    if (out) {
      out->indices.clear();
      for (size_t i = 0; i < n; i++)
        if (keep[i]) out->indices.insert(out->indices.end(), {outi[i * 3], outi[i * 3 + 1], outi[i * 3 + 2]});
      double err = 0;
      for (uint32_t v = 0; v < nv; v++) {
        const uint32_t r = rep[cell[v]];
        double d = 0;
        for (int k = 0; k < 3; k++) {
          const double x = static_cast<double>(pos[v * 3 + k]) - pos[r * 3 + k];
          d += x * x;
        }
        err = std::max(err, std::sqrt(d));
      }
      out->error = err;
    }
    size_t kept = 0;
    for (size_t i = 0; i < n; i++) kept += keep[i];
    return kept * 3;
  };
  uint32_t lo_g = 1, hi_g = 1024;
  if (run(hi_g, nullptr) <= target) lo_g = hi_g;
  while (hi_g - lo_g > 1) {
    const uint32_t mid = (lo_g + hi_g) / 2;
    if (run(mid, nullptr) <= target) lo_g = mid;
    else hi_g = mid;
  }
  Simplified res;
  run(lo_g, &res);
  return res;
}

// ------------------------------------------------------------------ partition (partition.ts:16-136)

// A map that iterates in insertion order and keeps a key's place when its value changes, as a JS Map does.
struct OrderedCounts {
  std::vector<std::pair<uint32_t, uint32_t>> items;  // (key, value); value ~0u marks a deleted entry
  std::unordered_map<uint32_t, size_t> at;
  void add(uint32_t k, uint32_t v) {
    auto it = at.find(k);
    if (it == at.end()) {
      at.emplace(k, items.size());
      items.emplace_back(k, v);
    } else {
      items[it->second].second += v;
    }
  }
  void erase(uint32_t k) {
    auto it = at.find(k);
    if (it == at.end()) return;
    items[it->second].second = ~0u;
    at.erase(it);
  }
  size_t size() const { return at.size(); }
};

uint32_t part1by2(uint32_t x) {
  x &= 0x3ff;
  x = (x | (x << 16)) & 0x030000ff;
  x = (x | (x << 8)) & 0x0300f00f;
  x = (x | (x << 4)) & 0x030c30c3;
  x = (x | (x << 2)) & 0x09249249;
  return x;
}

// Indices stably sorted along a 30-bit Morton curve (partition.ts:111-125 mortonOrder).
std::vector<uint32_t> morton_order(const std::vector<std::array<double, 3>>& pts) {
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (const auto& p : pts)
    for (int k = 0; k < 3; k++) {
      mn[k] = std::min(mn[k], p[k]);
      mx[k] = std::max(mx[k], p[k]);
    }
  double extent = std::max(mx[0] - mn[0], std::max(mx[1] - mn[1], mx[2] - mn[2]));
  if (!(extent != 0)) extent = 1;  // `|| 1`: zero or NaN
  std::vector<uint32_t> codes(pts.size());
  for (size_t i = 0; i < pts.size(); i++) {
    uint32_t c = 0;
    for (int k = 0; k < 3; k++) {
      double f = std::floor(((pts[i][k] - mn[k]) / extent) * 1023);
      f = std::min(1023.0, std::max(0.0, f));
      c |= part1by2(static_cast<uint32_t>(f)) << k;
    }
    codes[i] = c;
  }
  std::vector<uint32_t> order(pts.size());
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return codes[a] < codes[b]; });
  return order;
}

std::vector<std::vector<uint32_t>> partition(const std::vector<const std::vector<uint64_t>*>& edges,
                                             const std::vector<std::array<double, 3>>& centers, uint32_t group_size) {
  const uint32_t count = static_cast<uint32_t>(edges.size());
  std::vector<std::vector<uint32_t>> groups;
  if (count <= group_size) {
    groups.emplace_back(count);
    std::iota(groups[0].begin(), groups[0].end(), 0u);
    return groups;
  }
  // edge -> meshlets touching it, in the order edges are first seen.
  std::unordered_map<uint64_t, uint32_t> edge_at;
  std::vector<std::vector<uint32_t>> owners;
  for (uint32_t i = 0; i < count; i++)
    for (uint64_t e : *edges[i]) {
      auto it = edge_at.emplace(e, static_cast<uint32_t>(owners.size())).first;
      if (it->second == owners.size()) owners.emplace_back();
      auto& o = owners[it->second];
      if (o.empty() || o.back() != i) o.push_back(i);
    }
  // meshlet -> (neighbour -> shared edge count), in insertion order.
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> adjacency(count);
  auto bump = [&](uint32_t i, uint32_t j) {
    for (auto& pr : adjacency[i])
      if (pr.first == j) {
        pr.second++;
        return;
      }
    adjacency[i].emplace_back(j, 1u);
  };
  for (const auto& o : owners)
    for (size_t a = 0; a < o.size(); a++)
      for (size_t b = a + 1; b < o.size(); b++) {
        bump(o[a], o[b]);
        bump(o[b], o[a]);
      }
  const std::vector<uint32_t> order = morton_order(centers);
  std::vector<uint8_t> assigned(count, 0);
  auto dist2 = [&](uint32_t a, uint32_t b) {
    const double dx = centers[a][0] - centers[b][0], dy = centers[a][1] - centers[b][1],
                 dz = centers[a][2] - centers[b][2];
    return dx * dx + dy * dy + dz * dz;
  };
  for (uint32_t seed : order) {
    if (assigned[seed]) continue;
    std::vector<uint32_t> group{seed};
    assigned[seed] = 1;
    OrderedCounts frontier;
    auto add_neighbours = [&](uint32_t m) {
      for (const auto& [n, shared] : adjacency[m])
        if (!assigned[n]) frontier.add(n, shared);
    };
    add_neighbours(seed);
    while (group.size() < group_size && frontier.size() > 0) {
      int64_t best = -1;
      int64_t best_score = -1;
      double best_dist = INFINITY;
      for (const auto& [n, score] : frontier.items) {
        if (score == ~0u || assigned[n]) continue;
        const double d = dist2(n, seed);
        if (static_cast<int64_t>(score) > best_score || (static_cast<int64_t>(score) == best_score && d < best_dist)) {
          best = n;
          best_score = score;
          best_dist = d;
        }
      }
      if (best < 0) break;
      frontier.erase(static_cast<uint32_t>(best));
      assigned[best] = 1;
      group.push_back(static_cast<uint32_t>(best));
      add_neighbours(static_cast<uint32_t>(best));
    }
    groups.push_back(std::move(group));
  }
  // Undersized groups are merged with the following undersized ones (already in Morton order of their seeds).
  std::vector<std::vector<uint32_t>> result;
  std::vector<uint32_t> pending;
  const size_t min_size = std::max<size_t>(2, group_size / 2);
  for (auto& g : groups) {
    if (g.size() >= min_size) {
      result.push_back(std::move(g));
      continue;
    }
    if (pending.size() + g.size() > group_size) {
      result.push_back(std::move(pending));
      pending.clear();
    }
    pending.insert(pending.end(), g.begin(), g.end());
  }
  if (!pending.empty()) result.push_back(std::move(pending));
  return result;
}

// ------------------------------------------------------------------ the builder

// Position classes of a compacted subset from the global weld: local vertex -> first local vertex with
// the same welded id. `scratch` holds -1 for every vertex between calls.
std::vector<uint32_t> local_classes(const Local& l, const uint32_t* weld, std::vector<int32_t>& scratch) {
  std::vector<uint32_t> cls(l.to_global.size());
  for (size_t i = 0; i < cls.size(); i++) {
    int32_t& s = scratch[weld[l.to_global[i]]];
    if (s < 0) s = static_cast<int32_t>(i);
    cls[i] = static_cast<uint32_t>(s);
  }
  for (size_t i = 0; i < cls.size(); i++) scratch[weld[l.to_global[i]]] = -1;
  return cls;
}

struct Build {
  vgd_options opt;
  unsigned threads;
  std::vector<float> positions;
  std::vector<uint32_t> weld;
  std::vector<Meshlet> all;
  std::vector<std::vector<uint32_t>> lists;  // replacement lists (indices into `all`)
  std::vector<std::vector<int32_t>> scratch; // per worker, compact()'s map
  int32_t groups_linked = 0;

  // buildVirtualMesh.ts:453-482 splitIntoMeshlets: meshlets of a triangle list, not yet added to `all`.
  std::vector<Meshlet> split(const std::vector<uint32_t>& indices, uint32_t level, const Sphere* lod, double error,
                             unsigned worker) {
    Local local = compact(indices.data(), indices.size(), positions.data(), scratch[worker]);
    std::vector<Built> built = build_meshlets(local.indices, local.pos);
    std::vector<Meshlet> out(built.size());
    for (size_t m = 0; m < built.size(); m++) {
      Meshlet& r = out[m];
      r.vertices.resize(built[m].verts.size());
      for (size_t v = 0; v < built[m].verts.size(); v++) r.vertices[v] = local.to_global[built[m].verts[v]];
      r.local = std::move(built[m].tris);
      r.indices.resize(r.local.size());
      for (size_t i = 0; i < r.local.size(); i++) r.indices[i] = r.vertices[r.local[i]];
      r.cull = bounding_sphere(r.indices, positions.data());
      r.level = level;
      r.lod = lod ? *lod : r.cull;
      r.error = error;
      r.parent = Sphere{};
      r.parent_error = kErrorInfinity;
      r.boundary = boundary_edges(r.indices, weld.data());
    }
    return out;
  }

  std::vector<uint32_t> append(std::vector<Meshlet>&& ms) {
    std::vector<uint32_t> ids;
    for (auto& m : ms) {
      ids.push_back(static_cast<uint32_t>(all.size()));
      all.push_back(std::move(m));
    }
    return ids;
  }

  static size_t count_triangles(const std::vector<Meshlet>& all, const std::vector<uint32_t>& ids) {
    size_t n = 0;
    for (uint32_t i : ids) n += all[i].indices.size() / 3;
    return n;
  }

  std::vector<uint32_t> concat(const std::vector<uint32_t>& ids) const {
    std::vector<uint32_t> out;
    for (uint32_t i : ids) out.insert(out.end(), all[i].indices.begin(), all[i].indices.end());
    return out;
  }

  void link(const std::vector<uint32_t>& children, const Sphere& lod, double error, const std::vector<uint32_t>& outputs) {
    const int32_t group = groups_linked++;
    for (uint32_t c : children) {
      all[c].parent = lod;
      all[c].parent_error = error;
      all[c].parent_group = group;
    }
    for (uint32_t o : outputs) all[o].own_group = group;
    if (outputs.empty()) return;
    const int32_t list = static_cast<int32_t>(lists.size());
    lists.push_back(outputs);
    for (uint32_t c : children) all[c].replacements = list;
  }

  // buildVirtualMesh.ts:211-294: the level loop; returns lodLevels.
  uint32_t levels(std::vector<uint32_t>& current, size_t& last_triangles) {
    uint32_t lod_levels = 1;
    for (uint32_t level = 1; level <= opt.max_lod_levels && current.size() > 1; level++) {
      if (opt.tail_simplify && last_triangles <= kTailStartTriangles) break;
      std::vector<const std::vector<uint64_t>*> edges;
      std::vector<std::array<double, 3>> centers;
      for (uint32_t i : current) {
        edges.push_back(&all[i].boundary);
        centers.push_back({all[i].cull.x, all[i].cull.y, all[i].cull.z});
      }
      const auto groups = partition(edges, centers, opt.group_size);
      struct Result {
        bool root = true;
        Sphere lod;
        double error = 0;
        std::vector<Meshlet> outputs;
        std::vector<uint32_t> children;
      };
      std::vector<Result> results(groups.size());
      parallel_for(groups.size(), threads, [&](size_t g, unsigned worker) {
        Result& r = results[g];
        for (uint32_t i : groups[g]) r.children.push_back(current[i]);
        const std::vector<uint32_t> merged = concat(r.children);
        Local local = compact(merged.data(), merged.size(), positions.data(), scratch[worker]);
        const size_t target = std::max<size_t>(3, static_cast<size_t>(std::floor(merged.size() * opt.simplify_ratio / 3)) * 3);
        Simplified s = simplify(local.indices, local.pos, local_classes(local, weld.data(), scratch[worker]), target, true,
                                opt.prune != 0);
        std::vector<uint32_t> global(s.indices.size());
        for (size_t i = 0; i < s.indices.size(); i++) global[i] = local.to_global[s.indices[i]];
        const double shape = std::max(s.error, extent_shrink(local.indices, s.indices, local.pos.data()));
        if (global.size() > merged.size() * opt.max_group_keep_ratio) return;  // its meshlets stay roots
        std::vector<const Sphere*> spheres;
        double child_error = 0;
        for (uint32_t c : r.children) {
          spheres.push_back(&all[c].lod);
          child_error = std::max(child_error, all[c].error);
        }
        r.root = false;
        r.lod = enclosing(spheres);
        r.error = child_error + std::max(shape, 1e-7);
        if (!global.empty()) r.outputs = split(global, level, &r.lod, r.error, worker);
      });
      std::vector<uint32_t> next;
      for (Result& r : results) {
        if (r.root) continue;
        std::vector<uint32_t> ids = append(std::move(r.outputs));
        link(r.children, r.lod, r.error, ids);
        next.insert(next.end(), ids.begin(), ids.end());
      }
      if (next.empty()) break;
      lod_levels = level + 1;
      current = std::move(next);
      const size_t triangles = count_triangles(all, current);
      if (triangles > last_triangles * (1 - opt.min_level_reduction)) break;
      last_triangles = triangles;
    }
    return lod_levels;
  }

  // buildVirtualMesh.ts:345-437 simplifyTail without the voxel candidates.
  uint32_t tail(std::vector<uint32_t> nodes, uint32_t levels) {
    if (nodes.empty()) return levels;
    uint32_t level = levels;
    for (; level < levels + opt.max_lod_levels; level++) {
      const std::vector<uint32_t> merged = concat(nodes);
      if (merged.size() / 3 <= opt.min_root_triangles) break;
      Local local = compact(merged.data(), merged.size(), positions.data(), scratch[0]);
      const size_t target = std::max<size_t>(3, static_cast<size_t>(std::floor(merged.size() / 2.0 / 3)) * 3);
      const double keep_limit = merged.size() * opt.max_group_keep_ratio;
      double child_error = 0;
      for (uint32_t n : nodes) child_error = std::max(child_error, all[n].error);
      Simplified s = simplify(local.indices, local.pos, local_classes(local, weld.data(), scratch[0]), target, false,
                              opt.prune != 0);
      if (s.indices.size() > keep_limit) {
        Simplified sloppy = simplify_sloppy(local.indices, local.pos, target);
        if (sloppy.indices.size() < s.indices.size()) s = std::move(sloppy);
      }
      if (!(s.indices.size() <= keep_limit)) break;  // parentError stays infinite
      const double parent_error =
          child_error + std::max(std::max(s.error, extent_shrink(local.indices, s.indices, local.pos.data())), 1e-7);
      std::vector<uint32_t> global(s.indices.size());
      for (size_t i = 0; i < s.indices.size(); i++) global[i] = local.to_global[s.indices[i]];
      std::vector<const Sphere*> spheres;
      for (uint32_t n : nodes) spheres.push_back(&all[n].lod);
      const Sphere lod = enclosing(spheres);
      if (global.empty()) {
        link(nodes, lod, parent_error, {});
        return level + 1;  // pruned away entirely beyond this distance
      }
      std::vector<uint32_t> outputs = append(split(global, level, &lod, parent_error, 0));
      link(nodes, lod, parent_error, outputs);
      nodes = std::move(outputs);
    }
    return level;
  }
};

}  // namespace

struct vgd_mesh {
  std::vector<float> positions;
  std::vector<uint32_t> indices;
  std::vector<float> bounds;
  std::vector<uint32_t> info;
  std::vector<uint32_t> mverts, mtris;
  std::vector<uint32_t> level_ranges;
  std::vector<float> level_errors, level_guard;
  std::vector<int32_t> rgroup;
  std::vector<uint32_t> rstart, rindices;
  float bounding[4], lodb[4], cullb[4];
  uint64_t leaf_tris = 0;
  uint32_t leaf_meshlets = 0, root_tris = 0, root_meshlets = 0, lod_levels = 0;
};

namespace {

void store(float* dst, const Sphere& s) {
  dst[0] = static_cast<float>(s.x);
  dst[1] = static_cast<float>(s.y);
  dst[2] = static_cast<float>(s.z);
  dst[3] = static_cast<float>(s.r);
}

// buildVirtualMesh.ts:484-601 pack. Meshlets are sorted by level as there; within a level (where the
// reference keeps creation order) they are ordered by the group they merge into, then by the group that
// produced them, so siblings that share both LOD tests (sphere and error, own and parent) are adjacent.
void pack(Build& b, uint32_t lod_levels, vgd_mesh& out) {
  const std::vector<Meshlet>& all = b.all;
  std::vector<uint32_t> order(all.size());
  std::iota(order.begin(), order.end(), 0u);
  auto parent_key = [&](uint32_t i) { return all[i].parent_group < 0 ? INT32_MAX : all[i].parent_group; };
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) {
    if (all[a].level != all[c].level) return all[a].level < all[c].level;
    if (parent_key(a) != parent_key(c)) return parent_key(a) < parent_key(c);
    return all[a].own_group < all[c].own_group;
  });
  std::vector<uint32_t> index_of(all.size());
  for (uint32_t i = 0; i < order.size(); i++) index_of[order[i]] = i;
  const size_t n = all.size();
  out.rgroup.assign(n, -1);
  std::vector<int32_t> group_of_list(b.lists.size(), -1);
  out.rstart = {0};
  std::vector<int32_t> lists_in_order;
  uint32_t total = 0;
  for (uint32_t i = 0; i < n; i++) {
    const Meshlet& m = all[order[i]];
    if (m.replacements < 0) continue;
    int32_t& g = group_of_list[m.replacements];
    if (g < 0) {
      g = static_cast<int32_t>(lists_in_order.size());
      lists_in_order.push_back(m.replacements);
      total += static_cast<uint32_t>(b.lists[m.replacements].size());
      out.rstart.push_back(total);
    }
    out.rgroup[i] = g;
  }
  out.rindices.resize(total);
  for (size_t g = 0; g < lists_in_order.size(); g++) {
    const auto& list = b.lists[lists_in_order[g]];
    for (size_t k = 0; k < list.size(); k++) out.rindices[out.rstart[g] + k] = index_of[list[k]];
  }
  out.bounds.assign(n * 16, 0.0f);
  out.info.assign(n * 4, 0u);
  uint32_t first_index = 0, vertex_offset = 0;
  for (uint32_t i = 0; i < n; i++) {
    const Meshlet& m = all[order[i]];
    out.indices.insert(out.indices.end(), m.indices.begin(), m.indices.end());
    out.mverts.insert(out.mverts.end(), m.vertices.begin(), m.vertices.end());
    for (size_t k = 0; k < m.local.size(); k += 3)
      out.mtris.push_back(m.local[k] | (static_cast<uint32_t>(m.local[k + 1]) << 8) |
                          (static_cast<uint32_t>(m.local[k + 2]) << 16));
    float* bd = &out.bounds[i * 16];
    store(bd, m.lod);
    store(bd + 4, m.parent);
    bd[8] = static_cast<float>(m.error);
    bd[9] = static_cast<float>(m.parent_error);
    store(bd + 12, m.cull);
    const uint32_t tris = static_cast<uint32_t>(m.indices.size() / 3);
    out.info[i * 4 + 0] = tris;
    out.info[i * 4 + 1] = first_index;
    out.info[i * 4 + 2] = m.level;
    out.info[i * 4 + 3] = vertex_offset;
    first_index += static_cast<uint32_t>(m.indices.size());
    vertex_offset += static_cast<uint32_t>(m.vertices.size());
    if (m.level == 0) {
      out.leaf_tris += tris;
      out.leaf_meshlets++;
    }
    if (m.parent_error >= kErrorInfinity) {
      out.root_tris += tris;
      out.root_meshlets++;
    }
  }
  std::vector<const Sphere*> leaf, lod, cull;
  for (uint32_t i = 0; i < n; i++) {
    const Meshlet& m = all[order[i]];
    if (m.level == 0) leaf.push_back(&m.cull);
    cull.push_back(&m.cull);
  }
  for (uint32_t i = 0; i < n; i++) lod.push_back(&all[order[i]].lod);
  for (uint32_t i = 0; i < n; i++)
    if (all[order[i]].parent_error < kErrorInfinity) lod.push_back(&all[order[i]].parent);
  store(out.bounding, enclosing(leaf));
  store(out.lodb, enclosing(lod));
  store(out.cullb, enclosing(cull));
  const uint32_t level_count = n ? all[order[n - 1]].level + 1 : 0;
  out.level_ranges.assign(level_count * 2, 0u);
  out.level_errors.assign(level_count * 2, 0.0f);
  out.level_guard.assign(level_count * 2, 0.0f);
  std::vector<double> le(level_count * 2), lg(level_count * 2);
  for (uint32_t l = 0; l < level_count; l++) {
    le[l * 2] = INFINITY;
    le[l * 2 + 1] = 0;
    lg[l * 2] = 0;
    lg[l * 2 + 1] = INFINITY;
  }
  for (uint32_t i = 0; i < n; i++) {
    const Meshlet& m = all[order[i]];
    const uint32_t l = m.level;
    if (out.level_ranges[l * 2 + 1] == 0) out.level_ranges[l * 2] = i;
    out.level_ranges[l * 2 + 1]++;
    le[l * 2] = std::min(le[l * 2], m.error);
    le[l * 2 + 1] = std::max(le[l * 2 + 1], m.parent_error);
    lg[l * 2] = std::max(lg[l * 2], m.error);
    lg[l * 2 + 1] = std::min(lg[l * 2 + 1], m.parent_error);
  }
  for (uint32_t k = 0; k < level_count * 2; k++) {
    out.level_errors[k] = static_cast<float>(le[k]);
    out.level_guard[k] = static_cast<float>(lg[k]);
  }
  out.lod_levels = lod_levels;
}

}  // namespace

extern "C" void vgd_default_options(vgd_options* o) {
  o->group_size = 12;
  o->max_lod_levels = 24;
  o->max_group_keep_ratio = 0.9;
  o->simplify_ratio = 0.5;
  o->min_level_reduction = 0.03;
  o->prune = 0;
  o->tail_simplify = 1;
  o->min_root_triangles = 4;
  o->threads = 0;
}

extern "C" int vgd_build(const float* positions, uint32_t vertex_count, const uint32_t* indices, uint64_t index_count,
                         const vgd_options* options, vgd_mesh** out) {
  *out = nullptr;
  if (!positions || !indices || index_count == 0 || index_count % 3 != 0 || vertex_count == 0 ||
      vertex_count >= (1u << 26))
    return -1;
  for (uint64_t i = 0; i < index_count; i++)
    if (indices[i] >= vertex_count) return -1;
  try {
    Build b;
    vgd_default_options(&b.opt);
    if (options) b.opt = *options;
    if (b.opt.group_size < 2) return -1;
    b.threads = b.opt.threads ? b.opt.threads : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    b.positions.assign(positions, positions + static_cast<size_t>(vertex_count) * 3);
    b.weld = position_remap(positions, vertex_count);
    b.scratch.assign(b.threads, std::vector<int32_t>(vertex_count, -1));
    // buildVirtualMesh.ts:207-210: the leaf level.
    std::vector<uint32_t> src(indices, indices + index_count);
    std::vector<uint32_t> current = b.append(b.split(src, 0, nullptr, 0.0, 0));
    size_t last_triangles = Build::count_triangles(b.all, current);
    uint32_t lod_levels = b.levels(current, last_triangles);
    if (b.opt.tail_simplify) {
      std::vector<uint32_t> roots;
      for (uint32_t i = 0; i < b.all.size(); i++)
        if (b.all[i].parent_error >= kErrorInfinity) roots.push_back(i);
      lod_levels = b.tail(roots, lod_levels);
    }
    auto m = std::make_unique<vgd_mesh>();
    m->positions = std::move(b.positions);
    pack(b, lod_levels, *m);
    *out = m.release();
    return 0;
  } catch (...) {
    return -2;
  }
}

extern "C" void vgd_mesh_get(const vgd_mesh* m, vgd_mesh_view* v) {
  v->vertex_count = static_cast<uint32_t>(m->positions.size() / 3);
  v->positions = m->positions.data();
  v->index_count = m->indices.size();
  v->indices = m->indices.data();
  v->meshlet_count = static_cast<uint32_t>(m->info.size() / 4);
  v->meshlet_bounds = m->bounds.data();
  v->meshlet_info = m->info.data();
  v->meshlet_vertex_count = m->mverts.size();
  v->meshlet_vertices = m->mverts.data();
  v->meshlet_triangles = m->mtris.data();
  v->level_count = static_cast<uint32_t>(m->level_ranges.size() / 2);
  v->level_ranges = m->level_ranges.data();
  v->level_errors = m->level_errors.data();
  v->level_guard_errors = m->level_guard.data();
  v->replacement_group = m->rgroup.data();
  v->group_count = static_cast<uint32_t>(m->rstart.size() - 1);
  v->replacement_start = m->rstart.data();
  v->replacement_indices = m->rindices.data();
  std::memcpy(v->bounding_sphere, m->bounding, sizeof m->bounding);
  std::memcpy(v->lod_bounds_sphere, m->lodb, sizeof m->lodb);
  std::memcpy(v->cull_bounds_sphere, m->cullb, sizeof m->cullb);
  v->leaf_triangles = m->leaf_tris;
  v->leaf_meshlets = m->leaf_meshlets;
  v->root_triangles = m->root_tris;
  v->root_meshlets = m->root_meshlets;
  v->lod_levels = m->lod_levels;
}

extern "C" void vgd_mesh_free(vgd_mesh* m) { delete m; }

extern "C" uint32_t vgd_partition(uint32_t count, const uint64_t* edges, const uint32_t* edge_offset,
                                  const uint32_t* edge_count, const double* centers, uint32_t group_size,
                                  uint32_t* group_of, uint32_t* members) {
  std::vector<std::vector<uint64_t>> e(count);
  std::vector<const std::vector<uint64_t>*> ep(count);
  std::vector<std::array<double, 3>> c(count);
  for (uint32_t i = 0; i < count; i++) {
    e[i].assign(edges + edge_offset[i], edges + edge_offset[i] + edge_count[i]);
    ep[i] = &e[i];
    c[i] = {centers[i * 3], centers[i * 3 + 1], centers[i * 3 + 2]};
  }
  const auto groups = partition(ep, c, group_size);
  uint32_t k = 0;
  for (uint32_t g = 0; g < groups.size(); g++)
    for (uint32_t i : groups[g]) {
      group_of[i] = g;
      members[k++] = i;
    }
  return static_cast<uint32_t>(groups.size());
}
