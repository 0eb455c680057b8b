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
#include <map>
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

// Nearest free triangle centroid to a point: a static kd-tree over the centroids (median splits, leaves of
// up to 8) whose nodes count their free triangles, so emptied subtrees are skipped. Ties go to the lower
// triangle id.
class CentroidTree {
 public:
  CentroidTree(const std::vector<float>& cen, const std::vector<uint8_t>& used) : cen_(cen) {
    const uint32_t nt = static_cast<uint32_t>(cen.size() / 3);
    items_.resize(nt);
    std::iota(items_.begin(), items_.end(), 0u);
    leaf_of_.assign(nt, 0);
    build(0, nt, -1);
    for (uint32_t t = 0; t < nt; t++)
      if (used[t]) remove(t);
  }
  void remove(uint32_t t) {
    for (int32_t n = leaf_of_[t]; n >= 0; n = nodes_[n].parent) nodes_[n].live--;
  }
  // Returns -1 when every triangle is used.
  int64_t nearest(const double* p, const std::vector<uint8_t>& used) const {
    best_ = -1;
    best_d_ = INFINITY;
    if (!nodes_.empty() && nodes_[0].live) visit(0, p, used);
    return best_;
  }

 private:
  struct Node {
    float lo[3], hi[3];
    uint32_t begin, end, live;
    int32_t left = -1, right = -1, parent = -1;
  };
  const std::vector<float>& cen_;
  std::vector<uint32_t> items_;
  std::vector<int32_t> leaf_of_;
  std::vector<Node> nodes_;
  mutable int64_t best_ = -1;
  mutable double best_d_ = INFINITY;

  int32_t build(uint32_t begin, uint32_t end, int32_t parent) {
    const int32_t id = static_cast<int32_t>(nodes_.size());
    nodes_.push_back(Node{});
    Node nd;
    nd.begin = begin;
    nd.end = end;
    nd.live = end - begin;
    nd.parent = parent;
    for (int k = 0; k < 3; k++) {
      nd.lo[k] = INFINITY;
      nd.hi[k] = -INFINITY;
    }
    for (uint32_t i = begin; i < end; i++)
      for (int k = 0; k < 3; k++) {
        nd.lo[k] = std::min(nd.lo[k], cen_[items_[i] * 3 + k]);
        nd.hi[k] = std::max(nd.hi[k], cen_[items_[i] * 3 + k]);
      }
    if (end - begin > 8) {
      int axis = 0;
      for (int k = 1; k < 3; k++)
        if (nd.hi[k] - nd.lo[k] > nd.hi[axis] - nd.lo[axis]) axis = k;
      const uint32_t mid = begin + (end - begin) / 2;
      std::nth_element(items_.begin() + begin, items_.begin() + mid, items_.begin() + end, [&](uint32_t a, uint32_t b) {
        const float ca = cen_[a * 3 + axis], cb = cen_[b * 3 + axis];
        return ca < cb || (ca == cb && a < b);
      });
      nodes_[id] = nd;
      const int32_t l = build(begin, mid, id);
      const int32_t r = build(mid, end, id);
      nodes_[id].left = l;
      nodes_[id].right = r;
    } else {
      nodes_[id] = nd;
      for (uint32_t i = begin; i < end; i++) leaf_of_[items_[i]] = id;
    }
    return id;
  }
  double box_d2(const Node& nd, const double* p) const {
    double d = 0;
    for (int k = 0; k < 3; k++) {
      const double e = p[k] < nd.lo[k] ? nd.lo[k] - p[k] : p[k] > nd.hi[k] ? p[k] - nd.hi[k] : 0.0;
      d += e * e;
    }
    return d;
  }
  void visit(int32_t n, const double* p, const std::vector<uint8_t>& used) const {
    const Node& nd = nodes_[n];
    if (!nd.live || box_d2(nd, p) > best_d_) return;
    if (nd.left < 0) {
      for (uint32_t i = nd.begin; i < nd.end; i++) {
        const uint32_t t = items_[i];
        if (used[t]) continue;
        const double dx = cen_[t * 3] - p[0], dy = cen_[t * 3 + 1] - p[1], dz = cen_[t * 3 + 2] - p[2];
        const double d = dx * dx + dy * dy + dz * dz;
        if (d < best_d_ || (d == best_d_ && t < best_)) {
          best_d_ = d;
          best_ = t;
        }
      }
      return;
    }
    const int32_t a = nd.left, b = nd.right;
    if (box_d2(nodes_[a], p) <= box_d2(nodes_[b], p)) {
      visit(a, p, used);
      visit(b, p, used);
    } else {
      visit(b, p, used);
      visit(a, p, used);
    }
  }
};

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
  std::unique_ptr<CentroidTree> tree;  // built when a meshlet first runs out of adjacent triangles
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
      if (tree) tree->remove(t);
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
      if (best < 0 && w == 0 && b.verts.size() + 3 <= kMaxVertices) {
        // No free triangle touches the meshlet (a separate piece is done): continue with the free triangle
        // nearest the centroid, as meshoptimizer's builder does, so pieces share meshlets.
        if (!tree) tree = std::make_unique<CentroidTree>(cen, used);
        best = tree->nearest(c, used);
      }
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

// The collapses of one simplification run, in order, so that the result for any target can be replayed:
// the collapse order does not depend on the target (a run to a smaller target passes through the state
// where a run to a larger one stops), so the result for target T is the input triangles with the
// collapses applied up to the first point where at most T indices are live.
struct SimplifyLog {
  std::vector<uint32_t> tri;  // the input triangles without zero-area ones
  std::vector<std::pair<uint32_t, uint32_t>> collapses;  // u -> v
  std::vector<size_t> live_after;
  std::vector<double> max_cost_after;
  double extent = 0;
};

// Quadric half-edge collapses down to `target` indices (see the file comment). Indices are local ids.
// `cls` maps each vertex to the first vertex with the same position. With `log`, the run is recorded.
Simplified simplify(const std::vector<uint32_t>& in, const std::vector<float>& pos_f, const std::vector<uint32_t>& cls,
                    size_t target, bool lock_border, bool prune, SimplifyLog* log = nullptr) {
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
  if (log) log->tri = tri;
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
  if (log) log->extent = extent;
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
      if (log) {
        log->collapses.emplace_back(e.u, e.v);
        log->live_after.push_back(live_tris);
        log->max_cost_after.push_back(max_cost);
      }
    }
    if (done == 0) break;
  }
  while (live_tris * 3 > target && next_comp < comps.size()) prune_next();

  for (uint32_t t = 0; t < ntri; t++)
    if (tri_live[t]) res.indices.insert(res.indices.end(), {tri[t * 3], tri[t * 3 + 1], tri[t * 3 + 2]});
  res.error = std::sqrt(max_cost) * extent;
  return res;
}

// The run a SimplifyLog recorded, stopped where a run to `target` indices stops.
Simplified replay(const SimplifyLog& log, size_t target) {
  Simplified res;
  const size_t ntri = log.tri.size() / 3;
  size_t k = 0;  // collapses applied
  if (ntri * 3 > target) {
    while (k < log.collapses.size() && log.live_after[k] * 3 > target) k++;
    if (k < log.collapses.size()) k++;  // the collapse that reached the target
  }
  uint32_t nv = 0;
  for (uint32_t v : log.tri) nv = std::max(nv, v + 1);
  for (size_t i = 0; i < k; i++) nv = std::max(nv, std::max(log.collapses[i].first, log.collapses[i].second) + 1);
  std::vector<uint32_t> rep(nv);
  std::iota(rep.begin(), rep.end(), 0u);
  for (size_t i = 0; i < k; i++) rep[log.collapses[i].first] = log.collapses[i].second;
  auto resolve = [&](uint32_t v) {
    uint32_t r = v;
    while (rep[r] != r) r = rep[r];
    while (rep[v] != r) {
      const uint32_t next = rep[v];
      rep[v] = r;
      v = next;
    }
    return r;
  };
  for (size_t t = 0; t < ntri; t++) {
    const uint32_t a = resolve(log.tri[t * 3]), b = resolve(log.tri[t * 3 + 1]), c = resolve(log.tri[t * 3 + 2]);
    if (a != b && b != c && a != c) res.indices.insert(res.indices.end(), {a, b, c});
  }
  res.error = k ? std::sqrt(log.max_cost_after[k - 1]) * log.extent : 0.0;
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

// ------------------------------------------------------------------ voxel proxies (voxelProxy.ts)

constexpr double kVoxelErrorCells = 2;  // VOXEL_ERROR_CELLS, voxelProxy.ts:15

// Vertex attributes that travel with positions when levels add vertices (buildVirtualMesh.ts:388-407
// appendVertices): normals 3, colors 3, uvs 2 per vertex; an empty vector is an absent attribute.
struct Attributes {
  std::vector<float> normals, colors, uvs;
};

struct VoxelProxy {
  std::vector<float> positions, normals, colors, uvs;  // colors / uvs empty when the source has none
  std::vector<uint32_t> indices;
  double cell = 0;
};

// voxelProxy.ts:333-354 fillMissing: NaN entries (`size` floats per vertex) take the mean of their
// triangles' known neighbours, over at most 8 passes; what is still unknown becomes 0.5.
void fill_missing(std::vector<float>& values, const std::vector<uint32_t>& indices, int size) {
  const size_t count = values.size() / size;
  for (int pass = 0; pass < 8; pass++) {
    size_t missing = 0;
    std::vector<float> sum(count * size, 0.0f);
    std::vector<uint32_t> hits(count, 0);
    for (size_t t = 0; t + 2 < indices.size(); t += 3)
      for (int k = 0; k < 3; k++) {
        const uint32_t v = indices[t + k];
        if (!std::isnan(values[v * size])) continue;
        for (int m = 1; m < 3; m++) {
          const uint32_t u = indices[t + (k + m) % 3];
          if (std::isnan(values[u * size])) continue;
          for (int c = 0; c < size; c++) sum[v * size + c] = static_cast<float>(double(sum[v * size + c]) + values[u * size + c]);
          hits[v]++;
        }
      }
    for (size_t v = 0; v < count; v++) {
      if (!std::isnan(values[v * size])) continue;
      if (hits[v]) {
        for (int c = 0; c < size; c++) values[v * size + c] = static_cast<float>(double(sum[v * size + c]) / hits[v]);
      } else {
        missing++;
      }
    }
    if (!missing) return;
  }
  for (float& x : values)
    if (std::isnan(x)) x = 0.5f;
}

// voxelProxy.ts:33-281 buildVoxelProxy: occupancy from samples at half-cell spacing over every triangle,
// a 3x3x3 closing, a separable 3-tap box blur of the occupancy, and a surface-nets surface at level 0.5
// (one vertex per cell the level crosses, at the mean of its edge crossings; one quad per crossing sample
// edge). JavaScript numbers are doubles and its typed arrays round to binary32 on store: the occupancy
// field, colour sums and sampled UVs are kept in float here and the arithmetic around them in double, as
// there. Returns false for an empty or flat source.
bool build_voxel_proxy(const float* positions, const float* colors, const float* uvs, const std::vector<uint32_t>& indices,
                       int resolution, VoxelProxy& out) {
  if (indices.empty()) return false;
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i : indices)
    for (int k = 0; k < 3; k++) {
      mn[k] = std::min(mn[k], static_cast<double>(positions[static_cast<size_t>(i) * 3 + k]));
      mx[k] = std::max(mx[k], static_cast<double>(positions[static_cast<size_t>(i) * 3 + k]));
    }
  const double extent = std::max(mx[0] - mn[0], std::max(mx[1] - mn[1], mx[2] - mn[2]));
  if (!(extent > 0)) return false;
  const double cell = extent / resolution;
  constexpr int kPad = 3;  // cells of empty border around the samples
  int64_t n[3];
  double origin[3];
  for (int k = 0; k < 3; k++) {
    n[k] = static_cast<int64_t>(std::ceil((mx[k] - mn[k]) / cell)) + 1 + kPad * 2;
    origin[k] = mn[k] - kPad * cell;
  }
  const int64_t nx = n[0], ny = n[1], nz = n[2];
  const int64_t sx = 1, sy = nx, sz = nx * ny, total = nx * ny * nz;

  // 1. Occupancy (+ colour sums, first UV) from barycentric samples.
  std::vector<uint8_t> occupied(total, 0);
  std::vector<float> color_sum(colors ? total * 3 : 0, 0.0f);
  std::vector<uint32_t> color_count(colors ? total : 0, 0);
  std::vector<float> voxel_uv(uvs ? total * 2 : 0, NAN);
  auto hyp = [&](size_t a, size_t b) {
    return std::hypot(double(positions[a]) - positions[b], double(positions[a + 1]) - positions[b + 1],
                      double(positions[a + 2]) - positions[b + 2]);
  };
  for (size_t t = 0; t + 2 < indices.size(); t += 3) {
    const size_t a = static_cast<size_t>(indices[t]) * 3, b = static_cast<size_t>(indices[t + 1]) * 3,
                 c = static_cast<size_t>(indices[t + 2]) * 3;
    const double longest = std::max(hyp(a, b), std::max(hyp(b, c), hyp(c, a)));
    const int64_t steps = std::max<int64_t>(1, static_cast<int64_t>(std::ceil(longest / (cell * 0.5))));
    for (int64_t i = 0; i <= steps; i++)
      for (int64_t j = 0; j <= steps - i; j++) {
        const double u = double(i) / steps, v = double(j) / steps, w = 1 - u - v;
        double p[3];
        for (int k = 0; k < 3; k++) p[k] = positions[a + k] * w + positions[b + k] * u + positions[c + k] * v;
        const int64_t vox = static_cast<int64_t>(std::floor((p[0] - origin[0]) / cell)) * sx +
                            static_cast<int64_t>(std::floor((p[1] - origin[1]) / cell)) * sy +
                            static_cast<int64_t>(std::floor((p[2] - origin[2]) / cell)) * sz;
        occupied[vox] = 1;
        if (uvs && std::isnan(voxel_uv[vox * 2])) {
          const size_t ta = static_cast<size_t>(indices[t]) * 2, tb = static_cast<size_t>(indices[t + 1]) * 2,
                       tc = static_cast<size_t>(indices[t + 2]) * 2;
          voxel_uv[vox * 2] = static_cast<float>(uvs[ta] * w + uvs[tb] * u + uvs[tc] * v);
          voxel_uv[vox * 2 + 1] = static_cast<float>(uvs[ta + 1] * w + uvs[tb + 1] * u + uvs[tc + 1] * v);
        }
        if (colors) {
          for (int k = 0; k < 3; k++) {
            const double col = colors[a + k] * w + colors[b + k] * u + colors[c + k] * v;
            color_sum[vox * 3 + k] = static_cast<float>(double(color_sum[vox * 3 + k]) + col);
          }
          color_count[vox]++;
        }
      }
  }

  // 2. Closing: 26-neighbourhood dilation, then erosion, each separable per axis with empty borders.
  auto morph = [&](const std::vector<uint8_t>& src, int axis, bool grow) {
    std::vector<uint8_t> dst(src.size());
    const int64_t stride = axis == 0 ? sx : axis == 1 ? sy : sz, size = n[axis];
    for (int64_t z = 0; z < nz; z++)
      for (int64_t y = 0; y < ny; y++)
        for (int64_t x = 0; x < nx; x++) {
          const int64_t i = x + y * nx + z * nx * ny, c = axis == 0 ? x : axis == 1 ? y : z;
          const uint8_t prev = c > 0 ? src[i - stride] : 0, next = c < size - 1 ? src[i + stride] : 0;
          dst[i] = grow ? (src[i] | prev | next) : (src[i] & prev & next);
        }
    return dst;
  };
  std::vector<uint8_t> closed = occupied;
  for (int axis = 0; axis < 3; axis++) closed = morph(closed, axis, true);
  for (int axis = 0; axis < 3; axis++) closed = morph(closed, axis, false);

  // 3. Field: the closed occupancy blurred by a 3-tap box along each axis in turn.
  std::vector<float> field(closed.begin(), closed.end());
  for (int axis = 0; axis < 3; axis++) {
    const int64_t stride = axis == 0 ? sx : axis == 1 ? sy : sz, size = n[axis];
    std::vector<float> line(size);
    const int64_t lines = total / size;
    for (int64_t l = 0; l < lines; l++) {
      const int64_t start = axis == 0 ? l * nx : axis == 1 ? (l % nx) + (l / nx) * nx * ny : l;
      for (int64_t k = 0; k < size; k++) line[k] = field[start + k * stride];
      for (int64_t k = 0; k < size; k++) {
        const double lo = k > 0 ? line[k - 1] : 0.0, hi = k < size - 1 ? line[k + 1] : 0.0;
        field[start + k * stride] = static_cast<float>((lo + line[k] + hi) / 3);
      }
    }
  }
  constexpr double kIso = 0.5;

  // 4. Surface nets.
  const int64_t cx = nx - 1, cy = ny - 1, cz = nz - 1;
  std::vector<int32_t> cell_vertex(static_cast<size_t>(cx * cy * cz), -1);
  std::vector<double> pos;  // doubles until the end: the quads are oriented with them
  std::vector<float> nrm, col, uv;
  auto corner_index = [&](int64_t base, int c) { return base + (c & 1) * sx + ((c >> 1) & 1) * sy + ((c >> 2) & 1) * sz; };
  for (int64_t z = 0; z < cz; z++)
    for (int64_t y = 0; y < cy; y++)
      for (int64_t x = 0; x < cx; x++) {
        const int64_t base = x * sx + y * sy + z * sz;
        double corner[8];
        int inside = 0;
        for (int c = 0; c < 8; c++) {
          corner[c] = field[corner_index(base, c)];
          if (corner[c] >= kIso) inside++;
        }
        if (inside == 0 || inside == 8) continue;
        double px = 0, py = 0, pz = 0;
        int crossings = 0;
        for (int c = 0; c < 8; c++)
          for (int bit : {1, 2, 4}) {
            if (c & bit) continue;
            const double f0 = corner[c], f1 = corner[c | bit];
            if ((f0 >= kIso) == (f1 >= kIso)) continue;
            const double s = (kIso - f0) / (f1 - f0);
            px += (c & 1) + (bit == 1 ? s : 0);
            py += ((c >> 1) & 1) + (bit == 2 ? s : 0);
            pz += ((c >> 2) & 1) + (bit == 4 ? s : 0);
            crossings++;
          }
        cell_vertex[x + y * cx + z * cx * cy] = static_cast<int32_t>(pos.size() / 3);
        pos.push_back(origin[0] + (x + 0.5 + px / crossings) * cell);
        pos.push_back(origin[1] + (y + 0.5 + py / crossings) * cell);
        pos.push_back(origin[2] + (z + 0.5 + pz / crossings) * cell);
        // The field falls outward: the outward normal is minus its gradient.
        const double gx = corner[1] + corner[3] + corner[5] + corner[7] - corner[0] - corner[2] - corner[4] - corner[6];
        const double gy = corner[2] + corner[3] + corner[6] + corner[7] - corner[0] - corner[1] - corner[4] - corner[5];
        const double gz = corner[4] + corner[5] + corner[6] + corner[7] - corner[0] - corner[1] - corner[2] - corner[3];
        double len = std::hypot(gx, gy, gz);
        if (len == 0 || std::isnan(len)) len = 1;
        nrm.insert(nrm.end(), {static_cast<float>(-gx / len), static_cast<float>(-gy / len), static_cast<float>(-gz / len)});
        if (colors) {
          double r = 0, g = 0, bl = 0;
          uint64_t count = 0;
          for (int c = 0; c < 8; c++) {
            const int64_t i = corner_index(base, c);
            if (!color_count[i]) continue;
            r += color_sum[i * 3];
            g += color_sum[i * 3 + 1];
            bl += color_sum[i * 3 + 2];
            count += color_count[i];
          }
          if (count) col.insert(col.end(), {static_cast<float>(r / count), static_cast<float>(g / count), static_cast<float>(bl / count)});
          else col.insert(col.end(), {NAN, NAN, NAN});  // filled from neighbours below
        }
        if (uvs) {
          float u = NAN, v = NAN;
          for (int c = 0; c < 8 && std::isnan(u); c++) {
            const int64_t i = corner_index(base, c);
            u = voxel_uv[i * 2];
            v = voxel_uv[i * 2 + 1];
          }
          uv.insert(uv.end(), {u, v});
        }
      }
  std::vector<uint32_t> tris;
  auto cell_at = [&](int64_t x, int64_t y, int64_t z) { return static_cast<uint32_t>(cell_vertex[x + y * cx + z * cx * cy]); };
  auto quad = [&](uint32_t a, uint32_t b, uint32_t c, uint32_t d, double ox, double oy, double oz) {
    // Orient the two triangles so their normal agrees with the crossing edge's outward direction.
    const double ux = pos[c * 3] - pos[a * 3], uy = pos[c * 3 + 1] - pos[a * 3 + 1], uz = pos[c * 3 + 2] - pos[a * 3 + 2];
    const double vx = pos[d * 3] - pos[b * 3], vy = pos[d * 3 + 1] - pos[b * 3 + 1], vz = pos[d * 3 + 2] - pos[b * 3 + 2];
    const double qx = uy * vz - uz * vy, qy = uz * vx - ux * vz, qz = ux * vy - uy * vx;
    if (qx * ox + qy * oy + qz * oz >= 0) tris.insert(tris.end(), {a, b, c, a, c, d});
    else tris.insert(tris.end(), {a, c, b, a, d, c});
  };
  for (int64_t z = 1; z < nz - 1; z++)
    for (int64_t y = 1; y < ny - 1; y++)
      for (int64_t x = 1; x < nx - 1; x++) {
        const int64_t i = x * sx + y * sy + z * sz;
        const bool here = field[i] >= kIso;
        const double s = here ? 1 : -1;
        if (here != (field[i + sx] >= kIso))
          quad(cell_at(x, y - 1, z - 1), cell_at(x, y, z - 1), cell_at(x, y, z), cell_at(x, y - 1, z), s, 0, 0);
        if (here != (field[i + sy] >= kIso))
          quad(cell_at(x - 1, y, z - 1), cell_at(x, y, z - 1), cell_at(x, y, z), cell_at(x - 1, y, z), 0, s, 0);
        if (here != (field[i + sz] >= kIso))
          quad(cell_at(x - 1, y - 1, z), cell_at(x, y - 1, z), cell_at(x, y, z), cell_at(x - 1, y, z), 0, 0, s);
      }
  if (tris.empty()) return false;
  out.positions.assign(pos.begin(), pos.end());
  out.normals = std::move(nrm);
  out.colors = std::move(col);
  out.uvs = std::move(uv);
  if (colors) fill_missing(out.colors, tris, 3);
  if (uvs) fill_missing(out.uvs, tris, 2);
  out.indices = std::move(tris);
  out.cell = cell;
  return true;
}

// ------------------------------------------------------------------ aggregate geometry (aggregate.ts)

constexpr int kPairWindow = 24;  // PAIR_WINDOW, aggregate.ts:34

struct Aggregate {
  std::vector<uint32_t> indices;  // ids >= first_new refer to new_positions
  double error = 0;
  std::vector<double> new_positions;  // 3 per new vertex (binary32 once appended)
  std::vector<uint32_t> new_sources;  // the vertex each new one takes its other attributes from
  uint32_t first_new = 0;
};

struct Piece {
  double area = 0;
  double center[3] = {0, 0, 0};
  double axes[3][3] = {};  // unit axes, longest first
  double extent[3] = {0, 0, 0};
};

// aggregate.ts:291-338 eigenvectors: cyclic Jacobi rotations (at most 24 sweeps) of a symmetric 3x3 matrix;
// the columns of the accumulated rotation, normalised, by decreasing eigenvalue (stable for ties).
void jacobi_axes(const double* m, double axes[3][3]) {
  double a[9], v[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  std::memcpy(a, m, sizeof a);
  for (int sweep = 0; sweep < 24; sweep++) {
    const double off = std::fabs(a[1]) + std::fabs(a[2]) + std::fabs(a[5]);
    if (off < 1e-12 * (std::fabs(a[0]) + std::fabs(a[4]) + std::fabs(a[8]) + 1e-30)) break;
    for (const auto& pq : {std::pair<int, int>{0, 1}, {0, 2}, {1, 2}}) {
      const int p = pq.first, q = pq.second;
      const double apq = a[p * 3 + q];
      if (std::fabs(apq) < 1e-30) continue;
      const double theta = (a[q * 3 + q] - a[p * 3 + p]) / (2 * apq);
      const double sgn = theta > 0 ? 1.0 : theta < 0 ? -1.0 : 1.0;
      const double t = sgn / (std::fabs(theta) + std::sqrt(theta * theta + 1));
      const double c = 1 / std::sqrt(t * t + 1), s = t * c;
      for (int k = 0; k < 3; k++) {
        const double akp = a[k * 3 + p], akq = a[k * 3 + q];
        a[k * 3 + p] = c * akp - s * akq;
        a[k * 3 + q] = s * akp + c * akq;
      }
      for (int k = 0; k < 3; k++) {
        const double apk = a[p * 3 + k], aqk = a[q * 3 + k];
        a[p * 3 + k] = c * apk - s * aqk;
        a[q * 3 + k] = s * apk + c * aqk;
      }
      for (int k = 0; k < 3; k++) {
        const double vkp = v[k * 3 + p], vkq = v[k * 3 + q];
        v[k * 3 + p] = c * vkp - s * vkq;
        v[k * 3 + q] = s * vkp + c * vkq;
      }
    }
  }
  int order[3] = {0, 1, 2};
  std::stable_sort(order, order + 3, [&](int i, int j) { return a[j * 3 + j] - a[i * 3 + i] < 0; });
  for (int r = 0; r < 3; r++) {
    const int i = order[r];
    const double x = v[i], y = v[3 + i], z = v[6 + i];
    double l = std::hypot(x, y, z);
    if (l == 0) l = 1;
    axes[r][0] = x / l;
    axes[r][1] = y / l;
    axes[r][2] = z / l;
  }
}

// aggregate.ts:256-289 describe: area, vertex centroid, principal axes and extents along them.
Piece describe_piece(const std::vector<uint32_t>& idx, const std::vector<float>& positions) {
  Piece p;
  std::vector<uint32_t> unique;  // first-occurrence order, as a JavaScript Set keeps it
  {
    std::unordered_map<uint32_t, uint8_t> once;
    once.reserve(idx.size() * 2);
    for (uint32_t v : idx)
      if (once.emplace(v, 1).second) unique.push_back(v);
  }
  const double count = static_cast<double>(unique.size());
  for (uint32_t v : unique)
    for (int k = 0; k < 3; k++) p.center[k] += positions[static_cast<size_t>(v) * 3 + k] / count;
  double cov[9] = {};
  for (uint32_t v : unique) {
    double d[3];
    for (int k = 0; k < 3; k++) d[k] = positions[static_cast<size_t>(v) * 3 + k] - p.center[k];
    for (int r = 0; r < 3; r++)
      for (int q = 0; q < 3; q++) cov[r * 3 + q] += d[r] * d[q];
  }
  jacobi_axes(cov, p.axes);
  for (int a = 0; a < 3; a++) {
    double lo = INFINITY, hi = -INFINITY;
    for (uint32_t v : unique) {
      const float* q = &positions[static_cast<size_t>(v) * 3];
      const double along = (q[0] - p.center[0]) * p.axes[a][0] + (q[1] - p.center[1]) * p.axes[a][1] + (q[2] - p.center[2]) * p.axes[a][2];
      lo = std::min(lo, along);
      hi = std::max(hi, along);
    }
    p.extent[a] = hi - lo;
  }
  for (size_t t = 0; t + 2 < idx.size(); t += 3) {
    const float* A = &positions[static_cast<size_t>(idx[t]) * 3];
    const float* B = &positions[static_cast<size_t>(idx[t + 1]) * 3];
    const float* C = &positions[static_cast<size_t>(idx[t + 2]) * 3];
    const double u[3] = {double(B[0]) - A[0], double(B[1]) - A[1], double(B[2]) - A[2]};
    const double w[3] = {double(C[0]) - A[0], double(C[1]) - A[1], double(C[2]) - A[2]};
    p.area += 0.5 * std::hypot(u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]);
  }
  return p;
}

// aggregate.ts:349-358 morton: 10 bits per axis of a point in [0, 1]^3 (clamped).
uint32_t morton10(const double* p) {
  auto spread = [](double x) {
    uint32_t v = static_cast<uint32_t>(std::min(1023.0, std::max(0.0, std::floor(x * 1024)))) & 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
  };
  return spread(p[0]) | (spread(p[1]) << 1) | (spread(p[2]) << 2);
}

// aggregate.ts:225-254 widen: a copy of a piece's vertices spread along its first two axes (across its width
// when it is long and thin, evenly in its plane otherwise) to scale its area, and moved by `shift`. New ids
// start at `first_id`; returns the growth (how far an edge moved).
double widen_piece(std::vector<uint32_t>& idx, const std::vector<float>& positions, const Piece& piece, double factor,
                   const double* shift, uint32_t first_id, Aggregate& out) {
  const bool elongated = piece.extent[0] > piece.extent[1] * 2;
  const double scale[2] = {elongated ? 1.0 : std::sqrt(factor), elongated ? factor : std::sqrt(factor)};
  const double growth = std::max((scale[0] - 1) * piece.extent[0], (scale[1] - 1) * piece.extent[1]) / 2;
  std::unordered_map<uint32_t, uint32_t> copy;
  for (uint32_t& v : idx) {
    auto it = copy.find(v);
    if (it == copy.end()) {
      const uint32_t id = first_id + static_cast<uint32_t>(copy.size());
      it = copy.emplace(v, id).first;
      const float* q = &positions[static_cast<size_t>(v) * 3];
      const double d[3] = {q[0] - piece.center[0], q[1] - piece.center[1], q[2] - piece.center[2]};
      double p[3] = {q[0] + shift[0], q[1] + shift[1], q[2] + shift[2]};
      for (int a = 0; a < 2; a++) {
        const double* axis = piece.axes[a];
        const double along = d[0] * axis[0] + d[1] * axis[1] + d[2] * axis[2];
        for (int k = 0; k < 3; k++) p[k] += (scale[a] - 1) * along * axis[k];
      }
      out.new_positions.insert(out.new_positions.end(), {p[0], p[1], p[2]});
      out.new_sources.push_back(v);
    }
    v = it->second;
  }
  return growth;
}

// aggregate.ts:47-223 simplifyAggregate: a group made of separate pieces (connected components over welded
// vertices) reduced piece by piece toward `target_triangles`. Pieces whose every vertex belongs to this group
// alone (owner[weld[v]] == group) are free: each may be simplified on its own with open edges free, and
// neighbouring free pieces (in Morton order of their centres, within PAIR_WINDOW places and twice the median
// spacing) are thinned in pairs, the larger piece widened over both areas and moved to their area-weighted
// centre. Both orders (simplify first, or thin only) are tried; the more accurate of those that reach 1.25 x
// the target wins, else the one with fewest triangles. New vertex ids start at `first_new`.
bool simplify_aggregate(const std::vector<uint32_t>& merged, const std::vector<float>& positions, const std::vector<uint32_t>& weld,
                        const std::vector<int32_t>& owner, int32_t group, double target_triangles, uint32_t first_new,
                        Aggregate& result) {
  const size_t triangle_count = merged.size() / 3;
  // Pieces: components over welded vertices, listed in order of their first triangle.
  std::unordered_map<uint32_t, uint32_t> parent;
  parent.reserve(merged.size());
  auto find = [&](uint32_t x) {
    uint32_t r = x;
    while (parent[r] != r) r = parent[r];
    while (parent[x] != r) {
      const uint32_t next = parent[x];
      parent[x] = r;
      x = next;
    }
    return r;
  };
  for (uint32_t v : merged) parent.emplace(weld[v], weld[v]);
  for (size_t t = 0; t < merged.size(); t += 3) {
    const uint32_t a = find(weld[merged[t]]);
    for (int k = 1; k < 3; k++) {
      const uint32_t b = find(weld[merged[t + k]]);
      if (a != b) parent[b] = a;
    }
  }
  std::unordered_map<uint32_t, uint32_t> piece_of_root;
  std::vector<std::vector<uint32_t>> piece_tris;
  for (size_t t = 0; t < triangle_count; t++) {
    const uint32_t root = find(weld[merged[t * 3]]);
    auto it = piece_of_root.emplace(root, static_cast<uint32_t>(piece_tris.size())).first;
    if (it->second == piece_tris.size()) piece_tris.emplace_back();
    piece_tris[it->second].push_back(static_cast<uint32_t>(t));
  }
  if (piece_tris.size() < 2) return false;
  struct Part {
    std::vector<uint32_t> indices;
    bool free = true;
  };
  std::vector<Part> initial(piece_tris.size());
  for (size_t p = 0; p < piece_tris.size(); p++) {
    for (uint32_t t : piece_tris[p])
      for (int k = 0; k < 3; k++) {
        const uint32_t v = merged[t * 3 + k];
        initial[p].indices.push_back(v);
        if (owner[weld[v]] != group) initial[p].free = false;
      }
  }

  auto reduce = [&](bool simplify_first, Aggregate& out) {
    double error = 0;
    std::vector<Part> pieces = initial;
    const double ratio = target_triangles / static_cast<double>(triangle_count);
    double total = 0;
    std::vector<int32_t> scratch;
    for (Part& piece : pieces) {
      const size_t target = std::max<size_t>(6, static_cast<size_t>(std::floor(piece.indices.size() * ratio / 3)) * 3);
      if (simplify_first && piece.free && piece.indices.size() > target) {
        // compactLocal (aggregate.ts:340-358): local ids in first-use order
        std::unordered_map<uint32_t, uint32_t> local_of;
        std::vector<uint32_t> to_global, local(piece.indices.size());
        for (size_t i = 0; i < piece.indices.size(); i++) {
          auto it = local_of.emplace(piece.indices[i], static_cast<uint32_t>(to_global.size())).first;
          if (it->second == to_global.size()) to_global.push_back(piece.indices[i]);
          local[i] = it->second;
        }
        std::vector<float> lp(to_global.size() * 3);
        for (size_t i = 0; i < to_global.size(); i++)
          std::memcpy(&lp[i * 3], &positions[static_cast<size_t>(to_global[i]) * 3], 12);
        std::vector<uint32_t> cls = position_remap(lp.data(), static_cast<uint32_t>(to_global.size()));
        Simplified s = simplify(local, lp, cls, target, false, false);
        if (!s.indices.empty() && s.indices.size() < piece.indices.size()) {
          piece.indices.resize(s.indices.size());
          for (size_t i = 0; i < s.indices.size(); i++) piece.indices[i] = to_global[s.indices[i]];
          error = std::max(error, s.error);
        }
      }
      total += piece.indices.size() / 3;
    }
    out.first_new = first_new;
    if (total > target_triangles) {
      std::vector<uint32_t> order;
      std::unordered_map<uint32_t, Piece> info;
      for (uint32_t i = 0; i < pieces.size(); i++)
        if (pieces[i].free) {
          order.push_back(i);
          info.emplace(i, describe_piece(pieces[i].indices, positions));
        }
      double b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
      for (uint32_t i : order)
        for (int k = 0; k < 3; k++) {
          b[k] = std::min(b[k], info[i].center[k]);
          b[k + 3] = std::max(b[k + 3], info[i].center[k]);
        }
      const double span = std::max(std::max(b[3] - b[0], b[4] - b[1]), std::max(b[5] - b[2], 1e-9));
      auto key = [&](uint32_t i) {
        const double q[3] = {(info[i].center[0] - b[0]) / span, (info[i].center[1] - b[1]) / span, (info[i].center[2] - b[2]) / span};
        return morton10(q);
      };
      std::vector<uint32_t> keys(pieces.size(), 0);
      for (uint32_t i : order) keys[i] = key(i);
      std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return keys[x] < keys[y]; });
      auto dist = [&](uint32_t x, uint32_t y) {
        const double* p = info[x].center;
        const double* q = info[y].center;
        return std::hypot(p[0] - q[0], p[1] - q[1], p[2] - q[2]);
      };
      const int64_t count = static_cast<int64_t>(order.size());
      std::vector<double> nearest(count, INFINITY);
      for (int64_t n = 0; n < count; n++)
        for (int64_t m = std::max<int64_t>(0, n - kPairWindow); m < std::min<int64_t>(count, n + 1 + kPairWindow); m++)
          if (m != n) nearest[n] = std::min(nearest[n], dist(order[n], order[m]));
      std::vector<double> sorted_nearest = nearest;
      std::sort(sorted_nearest.begin(), sorted_nearest.end());
      const double spacing = count ? sorted_nearest[count / 2] : INFINITY;
      std::vector<uint8_t> taken(pieces.size(), 0);
      for (int64_t n = 0; n < count && total > target_triangles; n++) {
        const uint32_t a = order[n];
        if (taken[a]) continue;
        int64_t best = -1;
        double best_distance = spacing * 2;
        for (int64_t m = n + 1; m < std::min<int64_t>(count, n + 1 + kPairWindow); m++) {
          const uint32_t c = order[m];
          if (taken[c]) continue;
          const double d = dist(a, c);
          if (d <= best_distance) {
            best_distance = d;
            best = c;
          }
        }
        if (best < 0) continue;
        taken[a] = taken[best] = 1;
        const uint32_t keep = info[a].area >= info[static_cast<uint32_t>(best)].area ? a : static_cast<uint32_t>(best);
        const uint32_t drop = keep == a ? static_cast<uint32_t>(best) : a;
        const Piece& k = info[keep];
        const Piece& dropped = info[drop];
        const double pair_area = std::max(k.area + dropped.area, 1e-12);
        const double factor = pair_area / std::max(k.area, 1e-12);
        const double w = dropped.area / pair_area;
        const double shift[3] = {(dropped.center[0] - k.center[0]) * w, (dropped.center[1] - k.center[1]) * w,
                                 (dropped.center[2] - k.center[2]) * w};
        const uint32_t first_id = first_new + static_cast<uint32_t>(out.new_sources.size());
        const double growth = widen_piece(pieces[keep].indices, positions, k, factor, shift, first_id, out);
        error = std::max(error, std::max(best_distance * std::max(w, 1 - w), growth));
        total -= pieces[drop].indices.size() / 3;
        pieces[drop].indices.clear();
      }
    }
    out.indices.clear();
    for (const Part& p : pieces) out.indices.insert(out.indices.end(), p.indices.begin(), p.indices.end());
    if (out.indices.size() >= merged.size()) return false;
    out.error = error;
    return true;
  };

  Aggregate first, thin;
  const bool have_first = reduce(true, first);
  const bool have_thin = reduce(false, thin);
  std::vector<Aggregate*> candidates;
  if (have_first) candidates.push_back(&first);
  if (have_thin) candidates.push_back(&thin);
  if (candidates.empty()) return false;
  std::vector<Aggregate*> reaches;
  for (Aggregate* c : candidates)
    if (c->indices.size() / 3.0 <= target_triangles * 1.25) reaches.push_back(c);
  std::vector<Aggregate*>& pool = reaches.empty() ? candidates : reaches;
  std::stable_sort(pool.begin(), pool.end(), [&](const Aggregate* x, const Aggregate* y) {
    return reaches.empty() ? x->indices.size() < y->indices.size() : x->error < y->error;
  });
  result = std::move(*pool[0]);
  return true;
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

// buildVirtualMesh.ts:177-181 voxelResolutions: from the finest grid down, halving, to 4 cells.
std::vector<int> voxel_resolutions(uint32_t finest) {
  std::vector<int> list;
  for (int r = std::max(4, static_cast<int>(finest)); r >= 4; r /= 2) list.push_back(r);
  return list;
}

struct Build {
  vgd_options opt;
  unsigned threads;
  std::vector<float> positions;
  Attributes attr;  // grows with positions
  std::vector<uint32_t> weld;
  std::vector<Meshlet> all;
  std::vector<std::vector<uint32_t>> lists;  // replacement lists (indices into `all`)
  std::vector<std::vector<int32_t>> scratch; // per worker, compact()'s map
  int32_t groups_linked = 0;
  uint32_t voxel_levels = 0;      // stats.voxelLevels
  uint32_t aggregate_groups = 0;  // stats.aggregateGroups
  // The full-detail source, for voxel proxies.
  const float* src_positions = nullptr;
  const float* src_colors = nullptr;
  const float* src_uvs = nullptr;
  std::vector<uint32_t> src_indices;
  std::map<int, std::unique_ptr<VoxelProxy>> proxies;  // by resolution; null: none
  std::map<int, SimplifyLog> proxy_runs;                          // by resolution
  std::map<std::pair<int, size_t>, Simplified> proxy_simplified;  // by (resolution, target)

  // buildVirtualMesh.ts:410-414 voxelProxy: the source's proxy at `resolution`, built on first use.
  const VoxelProxy* voxel_proxy(int resolution) {
    auto it = proxies.find(resolution);
    if (it == proxies.end()) {
      auto p = std::make_unique<VoxelProxy>();
      if (!build_voxel_proxy(src_positions, src_colors, src_uvs, src_indices, resolution, *p)) p.reset();
      it = proxies.emplace(resolution, std::move(p)).first;
    }
    return it->second.get();
  }

  // A proxy simplified to `target` indices with borders free (meshopt ErrorAbsolute). One complete run per
  // proxy is recorded and replayed to each target the level loop and the tail ask for.
  const Simplified& simplified_proxy(int resolution, const VoxelProxy& proxy, size_t target) {
    auto key = std::make_pair(resolution, target);
    auto it = proxy_simplified.find(key);
    if (it == proxy_simplified.end()) {
      auto run = proxy_runs.find(resolution);
      if (run == proxy_runs.end()) {
        std::vector<uint32_t> cls = position_remap(proxy.positions.data(), static_cast<uint32_t>(proxy.positions.size() / 3));
        run = proxy_runs.emplace(resolution, SimplifyLog{}).first;
        simplify(proxy.indices, proxy.positions, cls, 0, false, false, &run->second);
      }
      it = proxy_simplified.emplace(key, replay(run->second, target)).first;
    }
    return it->second;
  }

  // The cell size a proxy at `resolution` has (build_voxel_proxy's own arithmetic): a proxy's error is at
  // least VOXEL_ERROR_CELLS cells, so proxies that cannot beat a known error are neither built nor simplified.
  double src_extent = 0;
  double cell_of(int resolution) const { return src_extent / resolution; }

  // buildVirtualMesh.ts:298-309 voxelError: the smallest error of a simplified voxel proxy with at most half
  // of `triangles` (infinite when none gets there). Exact whenever it is at most `cap`; above it, some value
  // above `cap` (proxies whose least possible error exceeds both `cap` and the best so far are skipped).
  double voxel_error(size_t triangles, double cap) {
    const size_t target = std::max<size_t>(3, (triangles / 2) * 3);
    double best = INFINITY;
    for (int resolution : voxel_resolutions(opt.voxel_resolution)) {
      if (kVoxelErrorCells * cell_of(resolution) > std::min(best, cap)) continue;
      const VoxelProxy* proxy = voxel_proxy(resolution);
      if (!proxy) continue;
      const Simplified& s = simplified_proxy(resolution, *proxy, std::min(target, proxy->indices.size()));
      if (!s.indices.empty() && s.indices.size() <= target) best = std::min(best, kVoxelErrorCells * proxy->cell + s.error);
    }
    return best;
  }

  // buildVirtualMesh.ts:388-407 appendVertices: positions (rounded to binary32) and their attributes join
  // the shared buffers; each new vertex is its own weld class. Returns the first new id.
  uint32_t append_vertices(const std::vector<double>& p, const float* normals, const float* colors, const float* uvs) {
    const uint32_t first = static_cast<uint32_t>(positions.size() / 3);
    const size_t count = p.size() / 3;
    for (double x : p) positions.push_back(static_cast<float>(x));
    if (!attr.normals.empty()) attr.normals.insert(attr.normals.end(), normals, normals + count * 3);
    if (!attr.colors.empty()) {
      if (colors) attr.colors.insert(attr.colors.end(), colors, colors + count * 3);
      else attr.colors.insert(attr.colors.end(), count * 3, 1.0f);
    }
    if (!attr.uvs.empty()) {
      if (uvs) attr.uvs.insert(attr.uvs.end(), uvs, uvs + count * 2);
      else attr.uvs.insert(attr.uvs.end(), count * 2, 0.0f);
    }
    for (size_t i = 0; i < count; i++) weld.push_back(first + static_cast<uint32_t>(i));
    for (auto& sc : scratch) sc.resize(positions.size() / 3, -1);
    return first;
  }

  // Attributes of existing vertices, gathered (for the copies a thinned piece adds).
  static std::vector<float> gather(const std::vector<float>& a, const std::vector<uint32_t>& ids, int size) {
    std::vector<float> out;
    if (a.empty()) return out;
    out.reserve(ids.size() * size);
    for (uint32_t v : ids) out.insert(out.end(), a.begin() + static_cast<size_t>(v) * size, a.begin() + static_cast<size_t>(v + 1) * size);
    return out;
  }

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

  // buildVirtualMesh.ts:326-339 groupOwners: per welded vertex, the group whose meshlets use it, -2 when
  // meshlets of several groups do, -1 when unused.
  std::vector<int32_t> group_owners(const std::vector<std::vector<uint32_t>>& groups, const std::vector<uint32_t>& current) const {
    std::vector<int32_t> owner(weld.size(), -1);
    for (size_t g = 0; g < groups.size(); g++)
      for (uint32_t i : groups[g])
        for (uint32_t v : all[current[i]].indices) {
          int32_t& o = owner[weld[v]];
          if (o == -1) o = static_cast<int32_t>(g);
          else if (o != static_cast<int32_t>(g)) o = -2;
        }
    return owner;
  }

  // buildVirtualMesh.ts:211-294: the level loop; returns lodLevels. Groups are simplified in parallel; what
  // depends on their order (vertices a thinned group appends, the voxel limit) is settled after, in group order.
  uint32_t levels(std::vector<uint32_t>& current, size_t& last_triangles) {
    uint32_t lod_levels = 1;
    const bool voxels = opt.tail_simplify && opt.voxel_lods;
    for (uint32_t level = 1; level <= opt.max_lod_levels && current.size() > 1; level++) {
      if (opt.tail_simplify && last_triangles <= kTailStartTriangles) break;
      // Aggregate geometry: once a voxel stand-in of the whole mesh with half this level's triangles is as
      // accurate as the level, the tail takes over (buildVirtualMesh.ts:215-218, 311-316 voxelBeats).
      if (voxels && aggregate_groups > 0) {
        double level_error = 0;
        for (uint32_t i : current) level_error = std::max(level_error, all[i].error);
        if (voxel_error(last_triangles, level_error) <= level_error) break;
      }
      std::vector<const std::vector<uint64_t>*> edges;
      std::vector<std::array<double, 3>> centers;
      for (uint32_t i : current) {
        edges.push_back(&all[i].boundary);
        centers.push_back({all[i].cull.x, all[i].cull.y, all[i].cull.z});
      }
      const auto groups = partition(edges, centers, opt.group_size);
      std::vector<int32_t> owner;
      if (opt.aggregate_lods) owner = group_owners(groups, current);
      const uint32_t level_vertices = static_cast<uint32_t>(positions.size() / 3);
      struct Result {
        std::vector<uint32_t> children, merged, global;
        double shape = 0, child_error = 0;
        size_t simplified = 0;  // index count of the regular simplification
        bool has_aggregate = false;
        Aggregate aggregate;
        bool root = true;
        Sphere lod;
        double error = 0;
        std::vector<Meshlet> outputs;
      };
      std::vector<Result> results(groups.size());
      parallel_for(groups.size(), threads, [&](size_t g, unsigned worker) {
        Result& r = results[g];
        for (uint32_t i : groups[g]) r.children.push_back(current[i]);
        r.merged = concat(r.children);
        Local local = compact(r.merged.data(), r.merged.size(), positions.data(), scratch[worker]);
        const size_t target = std::max<size_t>(3, static_cast<size_t>(std::floor(r.merged.size() * opt.simplify_ratio / 3)) * 3);
        Simplified s = simplify(local.indices, local.pos, local_classes(local, weld.data(), scratch[worker]), target, true,
                                opt.prune != 0);
        r.global.resize(s.indices.size());
        for (size_t i = 0; i < s.indices.size(); i++) r.global[i] = local.to_global[s.indices[i]];
        r.simplified = s.indices.size();
        r.shape = std::max(s.error, extent_shrink(local.indices, s.indices, local.pos.data()));
        for (uint32_t c : r.children) r.child_error = std::max(r.child_error, all[c].error);
        // Stalled well short of the target, as groups of separate pieces do: piece by piece (aggregate.ts).
        if (!owner.empty() && s.indices.size() > target * 1.25)
          r.has_aggregate = simplify_aggregate(r.merged, positions, weld, owner, static_cast<int32_t>(g), target / 3.0,
                                               level_vertices, r.aggregate);
      });
      // The voxel limit is compared with each thinned group's error: it needs to be exact only up to the
      // largest of them.
      double voxel_limit = NAN, largest_thinned = 0;
      for (const Result& r : results)
        if (r.has_aggregate) largest_thinned = std::max(largest_thinned, r.child_error + r.aggregate.error);
      for (size_t g = 0; g < groups.size(); g++) {
        Result& r = results[g];
        if (!r.has_aggregate) continue;
        if (std::isnan(voxel_limit)) voxel_limit = voxels ? voxel_error(last_triangles, largest_thinned) : INFINITY;
        Aggregate& a = r.aggregate;
        if (!(a.indices.size() < r.simplified && r.child_error + a.error < voxel_limit)) continue;
        uint32_t first = static_cast<uint32_t>(positions.size() / 3);
        if (!a.new_sources.empty()) {
          const std::vector<float> n = gather(attr.normals, a.new_sources, 3), c = gather(attr.colors, a.new_sources, 3),
                                   u = gather(attr.uvs, a.new_sources, 2);
          first = append_vertices(a.new_positions, n.data(), c.empty() ? nullptr : c.data(), u.empty() ? nullptr : u.data());
        }
        for (uint32_t& v : a.indices)
          if (v >= level_vertices) v = v - level_vertices + first;
        r.global = std::move(a.indices);
        r.shape = std::max(a.error, extent_shrink(r.merged, r.global, positions.data()));
        aggregate_groups++;
      }
      parallel_for(groups.size(), threads, [&](size_t g, unsigned worker) {
        Result& r = results[g];
        if (r.global.size() > r.merged.size() * opt.max_group_keep_ratio) return;  // its meshlets stay roots
        std::vector<const Sphere*> spheres;
        for (uint32_t c : r.children) spheres.push_back(&all[c].lod);
        r.root = false;
        r.lod = enclosing(spheres);
        r.error = r.child_error + std::max(r.shape, 1e-7);
        if (!r.global.empty()) r.outputs = split(r.global, level, &r.lod, r.error, worker);
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

  // buildVirtualMesh.ts:345-437 simplifyTail: merge every root and keep halving, borders free; a simplified
  // voxel proxy of the full-detail mesh replaces the triangles when its error (measured against the
  // original, so it does not accumulate) is smaller.
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
      double parent_error = INFINITY;
      std::vector<uint32_t> global;
      if (s.indices.size() <= keep_limit) {
        parent_error =
            child_error + std::max(std::max(s.error, extent_shrink(local.indices, s.indices, local.pos.data())), 1e-7);
        global.resize(s.indices.size());
        for (size_t i = 0; i < s.indices.size(); i++) global[i] = local.to_global[s.indices[i]];
      }
      const VoxelProxy* voxel = nullptr;
      const Simplified* voxel_tris = nullptr;
      if (opt.voxel_lods) {
        for (int resolution : voxel_resolutions(opt.voxel_resolution)) {
          if (kVoxelErrorCells * cell_of(resolution) >= parent_error) continue;  // cannot win
          const VoxelProxy* proxy = voxel_proxy(resolution);
          if (!proxy) continue;
          const Simplified& p = simplified_proxy(resolution, *proxy, std::min(target, proxy->indices.size()));
          if (p.indices.empty() || p.indices.size() > keep_limit) continue;
          const double e = std::max(child_error * (1 + 1e-5) + 1e-7, kVoxelErrorCells * proxy->cell + p.error);
          if (e < parent_error) {
            parent_error = e;
            voxel = proxy;
            voxel_tris = &p;
          }
        }
      }
      if (!std::isfinite(parent_error)) break;  // parentError stays infinite
      if (voxel) {
        // The proxy vertices this level uses join the shared buffers, in first-use order.
        std::vector<int32_t> map(voxel->positions.size() / 3, -1);
        std::vector<uint32_t> used;
        global.resize(voxel_tris->indices.size());
        for (size_t i = 0; i < voxel_tris->indices.size(); i++) {
          int32_t& m = map[voxel_tris->indices[i]];
          if (m < 0) {
            m = static_cast<int32_t>(used.size());
            used.push_back(voxel_tris->indices[i]);
          }
          global[i] = static_cast<uint32_t>(m);
        }
        std::vector<double> p;
        for (uint32_t v : used)
          for (int k = 0; k < 3; k++) p.push_back(voxel->positions[static_cast<size_t>(v) * 3 + k]);
        const std::vector<float> n = gather(voxel->normals, used, 3), c = gather(voxel->colors, used, 3),
                                 u = gather(voxel->uvs, used, 2);
        const uint32_t first = append_vertices(p, n.data(), c.empty() ? nullptr : c.data(), u.empty() ? nullptr : u.data());
        for (uint32_t& v : global) v += first;
        voxel_levels++;
      }
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
  std::vector<float> positions, normals, colors, uvs;
  uint32_t voxel_levels = 0, aggregate_groups = 0;
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
  o->voxel_lods = 1;
  o->voxel_resolution = 64;
  o->aggregate_lods = 1;
}

extern "C" int vgd_build_source(const vgd_source* src, const vgd_options* options, vgd_mesh** out) {
  *out = nullptr;
  if (!src || !src->positions || !src->indices || src->index_count == 0 || src->index_count % 3 != 0 ||
      src->vertex_count == 0 || src->vertex_count >= (1u << 26))
    return -1;
  const uint32_t vertex_count = src->vertex_count;
  for (uint64_t i = 0; i < src->index_count; i++)
    if (src->indices[i] >= vertex_count) return -1;
  try {
    Build b;
    vgd_default_options(&b.opt);
    if (options) b.opt = *options;
    if (b.opt.group_size < 2 || (b.opt.voxel_lods && b.opt.voxel_resolution == 0)) return -1;
    b.threads = b.opt.threads ? b.opt.threads : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t nv = vertex_count;
    b.positions.assign(src->positions, src->positions + nv * 3);
    if (src->normals) b.attr.normals.assign(src->normals, src->normals + nv * 3);
    if (src->colors) b.attr.colors.assign(src->colors, src->colors + nv * 3);
    if (src->uvs) b.attr.uvs.assign(src->uvs, src->uvs + nv * 2);
    b.weld = position_remap(src->positions, vertex_count);
    b.scratch.assign(b.threads, std::vector<int32_t>(vertex_count, -1));
    b.src_positions = src->positions;
    b.src_colors = src->colors;
    b.src_uvs = src->uvs;
    b.src_indices.assign(src->indices, src->indices + src->index_count);
    {
      double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (uint32_t i : b.src_indices)
        for (int k = 0; k < 3; k++) {
          lo[k] = std::min(lo[k], static_cast<double>(src->positions[static_cast<size_t>(i) * 3 + k]));
          hi[k] = std::max(hi[k], static_cast<double>(src->positions[static_cast<size_t>(i) * 3 + k]));
        }
      b.src_extent = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    }
    // buildVirtualMesh.ts:207-210: the leaf level.
    std::vector<uint32_t> current = b.append(b.split(b.src_indices, 0, nullptr, 0.0, 0));
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
    m->normals = std::move(b.attr.normals);
    m->colors = std::move(b.attr.colors);
    m->uvs = std::move(b.attr.uvs);
    m->voxel_levels = b.voxel_levels;
    m->aggregate_groups = b.aggregate_groups;
    pack(b, lod_levels, *m);
    *out = m.release();
    return 0;
  } catch (...) {
    return -2;
  }
}

extern "C" int vgd_build(const float* positions, uint32_t vertex_count, const uint32_t* indices, uint64_t index_count,
                         const vgd_options* options, vgd_mesh** out) {
  vgd_source src{vertex_count, positions, nullptr, nullptr, nullptr, index_count, indices};
  return vgd_build_source(&src, options, out);
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
  v->normals = m->normals.empty() ? nullptr : m->normals.data();
  v->colors = m->colors.empty() ? nullptr : m->colors.data();
  v->uvs = m->uvs.empty() ? nullptr : m->uvs.data();
  v->voxel_levels = m->voxel_levels;
  v->aggregate_groups = m->aggregate_groups;
}

extern "C" void vgd_mesh_free(vgd_mesh* m) { delete m; }

struct vgd_proxy {
  VoxelProxy p;
};

extern "C" int vgd_voxel_proxy(const float* positions, uint32_t vertex_count, const float* colors, const float* uvs,
                               const uint32_t* indices, uint64_t index_count, uint32_t resolution, vgd_proxy** out) {
  *out = nullptr;
  if (!positions || (index_count && !indices) || index_count % 3 != 0 || resolution == 0 || resolution > 4096) return -1;
  for (uint64_t i = 0; i < index_count; i++)
    if (indices[i] >= vertex_count) return -1;
  try {
    auto m = std::make_unique<vgd_proxy>();
    std::vector<uint32_t> idx(indices, indices + index_count);
    if (!build_voxel_proxy(positions, colors, uvs, idx, static_cast<int>(resolution), m->p)) return 1;
    *out = m.release();
    return 0;
  } catch (...) {
    return -2;
  }
}

extern "C" void vgd_proxy_get(const vgd_proxy* m, vgd_proxy_view* v) {
  v->vertex_count = static_cast<uint32_t>(m->p.positions.size() / 3);
  v->positions = m->p.positions.data();
  v->normals = m->p.normals.data();
  v->colors = m->p.colors.empty() ? nullptr : m->p.colors.data();
  v->uvs = m->p.uvs.empty() ? nullptr : m->p.uvs.data();
  v->index_count = m->p.indices.size();
  v->indices = m->p.indices.data();
  v->cell_size = m->p.cell;
}

extern "C" void vgd_proxy_free(vgd_proxy* m) { delete m; }

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
