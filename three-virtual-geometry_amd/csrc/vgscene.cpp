// Synthetic scene producer: procedural grid-patch meshes, their meshlet LOD DAGs in the reference's
// VirtualMeshData layout (src/core/preprocess/buildVirtualMesh.ts:75-138), instance placements and
// synthetic occluder depth. Host-only C++; bench and tests drive it through vgscene.h.
//
// DAG construction (our own, meshoptimizer-free):
//  - every mesh is a set of parametric patches P(s, t), (s, t) in [0,1]^2, each a grid of nu x nv quads
//    (multiples of 16), 2 triangles per quad;
//  - a patch is cut into leaf nodes of 16 x 16 quads, each 4 meshlets of 8 x 8 quads (128 triangles,
//    MAX_MESHLET_TRIANGLES, src/core/constants.ts:2);
//  - a kd-tree over the leaf nodes merges two sibling nodes per group: the group's 8 meshlets are
//    replaced by the parent node, re-sampled as 16 x 16 quads over the union (half the triangles,
//    simplifyRatio 0.5, buildVirtualMesh.ts:160-172), split into 4 meshlets;
//  - once every patch is one node, a tail merges all roots into one group per level and halves every
//    patch's grid until 1 x 1 quads (tailSimplify, buildVirtualMesh.ts:323-410), packing small pieces
//    into shared meshlets of <= 128 triangles and <= 128 vertices;
//  - group error = max child error + max(shape error, 1e-7) (buildVirtualMesh.ts:262-270), where the
//    shape error is the largest distance between the children's vertices and the parent's surface at
//    the same parameters; the group's LOD sphere encloses the children's LOD spheres (padded 1.0001,
//    buildVirtualMesh.ts:746-762); leaf LOD spheres are their culling spheres (buildVirtualMesh.ts:456);
//  - pack(): meshlets sorted by level, per-level ranges and errors, bound spheres
//    (buildVirtualMesh.ts:484-601).
#include "vgscene.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <functional>
#include <cstdlib>
#include <memory>
#include <unordered_map>
#include <vector>

namespace {

constexpr float kErrorInfinity = 1e30f;  // src/core/constants.ts:8
constexpr int kLeafQuads = 16;           // leaf node side in quads (4 meshlets of 8 x 8)
constexpr int kTile = 8;                 // meshlet tile side in quads (128 triangles)
constexpr int kMaxMeshletTris = 128;
constexpr int kMaxMeshletVerts = 128;

struct V3 {
  double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double len(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) {
  double l = len(a);
  return l > 0 ? a * (1.0 / l) : a;
}

// ------------------------------------------------------------------ random numbers and noise

struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x1234567ull) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

inline uint32_t hash3(int32_t x, int32_t y, int32_t z, uint32_t seed) {
  uint32_t h = seed * 0x27d4eb2du;
  h ^= static_cast<uint32_t>(x) * 0x85ebca6bu;
  h = (h << 13) | (h >> 19);
  h ^= static_cast<uint32_t>(y) * 0xc2b2ae35u;
  h = (h << 17) | (h >> 15);
  h ^= static_cast<uint32_t>(z) * 0x165667b1u;
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}

inline double lattice(int32_t x, int32_t y, int32_t z, uint32_t seed) {
  return (hash3(x, y, z, seed) >> 8) * (2.0 / 16777216.0) - 1.0;  // [-1, 1)
}

inline double fade(double t) { return t * t * (3.0 - 2.0 * t); }

// Trilinear value noise with smoothstep weights, in [-1, 1].
double value_noise(double x, double y, double z, uint32_t seed) {
  double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
  int32_t ix = static_cast<int32_t>(fx), iy = static_cast<int32_t>(fy), iz = static_cast<int32_t>(fz);
  double u = fade(x - fx), v = fade(y - fy), w = fade(z - fz);
  double c[2][2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int d = 0; d < 2; d++) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
  auto lerp = [](double a, double b, double t) { return a + (b - a) * t; };
  double x00 = lerp(c[0][0][0], c[1][0][0], u), x10 = lerp(c[0][1][0], c[1][1][0], u);
  double x01 = lerp(c[0][0][1], c[1][0][1], u), x11 = lerp(c[0][1][1], c[1][1][1], u);
  return lerp(lerp(x00, x10, v), lerp(x01, x11, v), w);
}

// 3 octaves, lacunarity 2, gain 0.5, normalised to [-1, 1].
double fbm3(V3 p, uint32_t seed) {
  double sum = 0, amp = 1, freq = 1;
  for (int o = 0; o < 3; o++) {
    sum += amp * value_noise(p.x * freq, p.y * freq, p.z * freq, seed + 1013u * o);
    amp *= 0.5;
    freq *= 2.0;
  }
  return sum / 1.75;
}

// ------------------------------------------------------------------ parametric patches

struct Patch {
  std::function<V3(double, double)> eval;
  int nu = 16, nv = 16;      // leaf quads
  double lu = 1, lv = 1;     // approximate side lengths (for resolution selection)
};

int round16(double v) { return std::max(kLeafQuads, static_cast<int>(std::lround(v / kLeafQuads)) * kLeafQuads); }

// Assign patch resolutions so the whole mesh has about `triangles` leaf triangles.
void fit_resolution(std::vector<Patch>& patches, double triangles) {
  double area = 0;
  for (auto& p : patches) area += p.lu * p.lv;
  double density = std::sqrt(std::max(1.0, triangles / 2.0) / area);  // quads per unit length
  for (auto& p : patches) {
    p.nu = round16(p.lu * density);
    p.nv = round16(p.lv * density);
  }
}

std::vector<Patch> make_patches(const vgs_mesh_params& m) {
  std::vector<Patch> patches;
  const uint32_t seed = m.seed;
  const double amp = m.noise_amplitude, freq = m.noise_frequency;
  switch (m.kind) {
    case VGS_SPHERE: {
      // Icosahedron as 10 diamonds (two faces sharing an edge each); a diamond's (s, t) grid split along
      // the anti-diagonal reproduces the geodesic subdivision of both faces at frequency nu.
      const double R = m.size[0];
      const double lat = std::atan(0.5);
      V3 N{0, 1, 0}, S{0, -1, 0}, U[5], L[5];
      for (int i = 0; i < 5; i++) {
        double a = 2 * M_PI * i / 5, b = a + M_PI / 5;
        U[i] = {std::cos(lat) * std::cos(a), std::sin(lat), std::cos(lat) * std::sin(a)};
        L[i] = {std::cos(lat) * std::cos(b), -std::sin(lat), std::cos(lat) * std::sin(b)};
      }
      auto diamond = [&](V3 A, V3 B, V3 C, V3 D) {
        Patch p;
        p.eval = [=](double s, double t) {
          V3 flat = (s + t <= 1.0) ? A + (B - A) * s + (D - A) * t : C + (D - C) * (1 - s) + (B - C) * (1 - t);
          V3 dir = normalize(flat);
          return dir * (R * (1.0 + amp * fbm3(dir * (R * freq), seed)));
        };
        p.lu = p.lv = R;
        return p;
      };
      for (int i = 0; i < 5; i++) {
        int j = (i + 1) % 5;
        patches.push_back(diamond(N, U[i], L[i], U[j]));
        patches.push_back(diamond(U[j], L[i], S, L[j]));
      }
      int f = m.target_triangles ? round16(std::sqrt(m.target_triangles / 20.0)) : round16(m.resolution);
      for (auto& p : patches) p.nu = p.nv = f;
      return patches;
    }
    case VGS_TERRAIN: {
      // The reference tests' terrain (tests/cut.test.ts:14-19): 400 x 400, 96 x 96 quads,
      // y = 30 sin(0.05 x) cos(0.04 z).
      Patch p;
      p.eval = [](double s, double t) {
        double x = -200.0 + 400.0 * s, z = -200.0 + 400.0 * t;
        return V3{x, 30.0 * std::sin(x * 0.05) * std::cos(z * 0.04), z};
      };
      p.nu = p.nv = m.resolution ? round16(m.resolution) : 96;
      patches.push_back(p);
      return patches;
    }
    case VGS_BOX: {
      const double hx = m.size[0], hy = m.size[1], hz = m.size[2];
      const double scale = std::max(hx, std::max(hy, hz));
      struct Face { V3 c, u, v, n; double lu, lv; };
      Face faces[6] = {
          {{hx, 0, 0}, {0, 0, -hz}, {0, hy, 0}, {1, 0, 0}, 2 * hz, 2 * hy},
          {{-hx, 0, 0}, {0, 0, hz}, {0, hy, 0}, {-1, 0, 0}, 2 * hz, 2 * hy},
          {{0, hy, 0}, {hx, 0, 0}, {0, 0, -hz}, {0, 1, 0}, 2 * hx, 2 * hz},
          {{0, -hy, 0}, {hx, 0, 0}, {0, 0, hz}, {0, -1, 0}, 2 * hx, 2 * hz},
          {{0, 0, hz}, {hx, 0, 0}, {0, hy, 0}, {0, 0, 1}, 2 * hx, 2 * hy},
          {{0, 0, -hz}, {-hx, 0, 0}, {0, hy, 0}, {0, 0, -1}, 2 * hx, 2 * hy},
      };
      for (const Face& f : faces) {
        Patch p;
        p.eval = [=](double s, double t) {
          V3 q = f.c + f.u * (2 * s - 1) + f.v * (2 * t - 1);
          return q + f.n * (amp * scale * fbm3(q * freq, seed));
        };
        p.lu = f.lu;
        p.lv = f.lv;
        patches.push_back(p);
      }
      break;
    }
    case VGS_COLUMN: {
      const double R = m.size[0], H = m.size[1];
      const double scale = std::max(R, H * 0.5);
      Patch side;
      side.eval = [=](double s, double t) {
        double a = 2 * M_PI * s;
        V3 n{std::cos(a), 0, std::sin(a)};
        V3 q{R * n.x, H * t, R * n.z};
        return q + n * (amp * scale * fbm3(q * freq, seed));
      };
      side.lu = 2 * M_PI * R;
      side.lv = H;
      patches.push_back(side);
      for (int top = 0; top < 2; top++) {
        Patch cap;
        cap.eval = [=](double s, double t) {
          // Elliptical grid mapping of the square onto the disc.
          double u = 2 * s - 1, v = 2 * t - 1;
          double x = u * std::sqrt(std::max(0.0, 1 - v * v / 2)), z = v * std::sqrt(std::max(0.0, 1 - u * u / 2));
          if (!top) z = -z;
          V3 q{R * x, top ? H : 0.0, R * z};
          V3 n{0, top ? 1.0 : -1.0, 0};
          return q + n * (amp * scale * fbm3(q * freq, seed));
        };
        cap.lu = cap.lv = 1.8 * R;
        patches.push_back(cap);
      }
      break;
    }
    case VGS_ARCH: {
      // A square tube of side T along two pillars joined by a half circle, plus the two pillar feet.
      const double span = m.size[0], Hh = m.size[1], T = m.size[2];
      const double r = span / 2, length = 2 * Hh + M_PI * r;
      const double scale = std::max(span, Hh);
      auto path = [=](double sigma, V3& pos, V3& tangent) {
        if (sigma < Hh) {
          pos = {-r, sigma, 0};
          tangent = {0, 1, 0};
        } else if (sigma < Hh + M_PI * r) {
          double phi = (sigma - Hh) / r;
          pos = {-r * std::cos(phi), Hh + r * std::sin(phi), 0};
          tangent = {std::sin(phi), std::cos(phi), 0};
        } else {
          pos = {r, Hh - (sigma - Hh - M_PI * r), 0};
          tangent = {0, -1, 0};
        }
      };
      Patch tube;
      tube.eval = [=](double s, double t) {
        V3 pos, tan;
        path(s * length, pos, tan);
        V3 n{tan.y, -tan.x, 0}, b{0, 0, 1};
        double e4 = std::min(3.999999, t * 4);
        int e = static_cast<int>(e4);
        double f = e4 - e, a = 0, c = 0;
        V3 out;
        switch (e) {
          case 0: a = -T / 2 + f * T; c = -T / 2; out = b * -1.0; break;
          case 1: a = T / 2; c = -T / 2 + f * T; out = n; break;
          case 2: a = T / 2 - f * T; c = T / 2; out = b; break;
          default: a = -T / 2; c = T / 2 - f * T; out = n * -1.0; break;
        }
        V3 q = pos + n * a + b * c;
        return q + out * (amp * scale * fbm3(q * freq, seed));
      };
      tube.lu = length;
      tube.lv = 4 * T;
      patches.push_back(tube);
      for (int side = 0; side < 2; side++) {
        Patch foot;
        double cx = side ? r : -r;
        foot.eval = [=](double s, double t) {
          V3 q{cx + (s - 0.5) * T, 0, (t - 0.5) * T};
          return q + V3{0, -1, 0} * (amp * scale * fbm3(q * freq, seed));
        };
        foot.lu = foot.lv = T;
        patches.push_back(foot);
      }
      break;
    }
    default:
      return patches;
  }
  if (m.target_triangles) {
    fit_resolution(patches, m.target_triangles);
  } else {
    for (auto& p : patches) {
      p.nu = round16(p.lu * m.resolution);
      p.nv = round16(p.lv * m.resolution);
    }
  }
  return patches;
}

// ------------------------------------------------------------------ DAG

struct Sphere {
  double c[4];
};

// Rounds a sphere to f32 values (what the buffers store).
Sphere to_f32(const Sphere& s) {
  Sphere r;
  for (int k = 0; k < 4; k++) r.c[k] = static_cast<float>(s.c[k]);
  return r;
}

// AABB-centred sphere over points (buildVirtualMesh.ts:704-731 `boundingSphere`).
Sphere bounding_sphere(const std::vector<V3>& pts) {
  V3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
  for (const V3& p : pts) {
    lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
    hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
  }
  V3 c = (lo + hi) * 0.5;
  double r2 = 0;
  for (const V3& p : pts) r2 = std::max(r2, dot(p - c, p - c));
  return to_f32(Sphere{{c.x, c.y, c.z, std::sqrt(r2)}});
}

// Sphere enclosing spheres, padded by 1.0001 for float error (buildVirtualMesh.ts:746-762).
Sphere enclosing(const std::vector<Sphere>& ss) {
  if (ss.size() == 1) return ss[0];
  V3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
  for (const Sphere& s : ss) {
    lo = {std::min(lo.x, s.c[0] - s.c[3]), std::min(lo.y, s.c[1] - s.c[3]), std::min(lo.z, s.c[2] - s.c[3])};
    hi = {std::max(hi.x, s.c[0] + s.c[3]), std::max(hi.y, s.c[1] + s.c[3]), std::max(hi.z, s.c[2] + s.c[3])};
  }
  V3 c = (lo + hi) * 0.5;
  c = {static_cast<float>(c.x), static_cast<float>(c.y), static_cast<float>(c.z)};
  double r = 0;
  for (const Sphere& s : ss) r = std::max(r, std::hypot(s.c[0] - c.x, s.c[1] - c.y, s.c[2] - c.z) + s.c[3]);
  return to_f32(Sphere{{c.x, c.y, c.z, r * 1.0001}});
}

// A patch region sampled as rx x ry quads.
struct Rep {
  double s0, s1, t0, t1;
  int rx, ry;
  std::vector<V3> pos;  // (rx + 1) * (ry + 1), row-major in t
  V3 at(int i, int j) const { return pos[static_cast<size_t>(j) * (rx + 1) + i]; }
};

Rep make_rep(const Patch& p, double s0, double s1, double t0, double t1, int rx, int ry) {
  Rep r{s0, s1, t0, t1, rx, ry, {}};
  r.pos.resize(static_cast<size_t>(rx + 1) * (ry + 1));
  for (int j = 0; j <= ry; j++)
    for (int i = 0; i <= rx; i++)
      r.pos[static_cast<size_t>(j) * (rx + 1) + i] = p.eval(s0 + (s1 - s0) * i / rx, t0 + (t1 - t0) * j / ry);
  return r;
}

// Surface of `r` at parameters (s, t): the triangle of its grid cell (anti-diagonal split).
V3 rep_surface(const Rep& r, double s, double t) {
  double fu = (s - r.s0) / (r.s1 - r.s0) * r.rx, fv = (t - r.t0) / (r.t1 - r.t0) * r.ry;
  int i = std::min(r.rx - 1, std::max(0, static_cast<int>(std::floor(fu))));
  int j = std::min(r.ry - 1, std::max(0, static_cast<int>(std::floor(fv))));
  double u = fu - i, v = fv - j;
  V3 p00 = r.at(i, j), p10 = r.at(i + 1, j), p01 = r.at(i, j + 1), p11 = r.at(i + 1, j + 1);
  if (u + v <= 1) return p00 + (p10 - p00) * u + (p01 - p00) * v;
  return p11 + (p01 - p11) * (1 - u) + (p10 - p11) * (1 - v);
}

// Largest distance from the vertices of `child` to the surface of `parent`.
double shape_error(const Rep& child, const Rep& parent) {
  double worst = 0;
  for (int j = 0; j <= child.ry; j++)
    for (int i = 0; i <= child.rx; i++) {
      double s = child.s0 + (child.s1 - child.s0) * i / child.rx;
      double t = child.t0 + (child.t1 - child.t0) * j / child.ry;
      worst = std::max(worst, len(child.at(i, j) - rep_surface(parent, s, t)));
    }
  return worst;
}

struct WipMeshlet {
  int level;
  Sphere lod, parent, cull;
  float error, parent_error;
  uint32_t tris, verts;
  int group;  // group this meshlet is a child of (-1: root)
  // Triangles as 3 local vertex ids (< verts) packed a | b << 8 | c << 16 (buildVirtualMesh.ts
  // meshletTriangles): two per grid quad, the piece's vertices numbered row by row.
  std::vector<uint32_t> tri;
};

struct Builder {
  std::vector<WipMeshlet> all;
  std::vector<std::vector<int>> group_outputs;

  // Meshlet over a set of grid tiles (pieces) of reps.
  struct Piece {
    const Rep* rep;
    int i0, i1, j0, j1;
    uint32_t tris() const { return 2u * (i1 - i0) * (j1 - j0); }
    uint32_t verts() const { return static_cast<uint32_t>((i1 - i0 + 1) * (j1 - j0 + 1)); }
  };

  int add_meshlet(const std::vector<Piece>& pieces, int level, const Sphere* lod, float error) {
    std::vector<V3> pts;
    std::vector<uint32_t> tri;
    uint32_t tris = 0, verts = 0;
    for (const Piece& pc : pieces) {
      for (int j = pc.j0; j <= pc.j1; j++)
        for (int i = pc.i0; i <= pc.i1; i++) pts.push_back(pc.rep->at(i, j));
      const uint32_t row = static_cast<uint32_t>(pc.i1 - pc.i0 + 1);
      for (int j = pc.j0; j < pc.j1; j++)
        for (int i = pc.i0; i < pc.i1; i++) {
          const uint32_t v00 = verts + static_cast<uint32_t>(j - pc.j0) * row + static_cast<uint32_t>(i - pc.i0);
          const uint32_t v10 = v00 + 1, v01 = v00 + row, v11 = v01 + 1;
          tri.push_back(v00 | (v10 << 8) | (v11 << 16));
          tri.push_back(v00 | (v11 << 8) | (v01 << 16));
        }
      tris += pc.tris();
      verts += pc.verts();
    }
    WipMeshlet m;
    m.tri = std::move(tri);
    m.level = level;
    m.cull = bounding_sphere(pts);
    m.lod = lod ? *lod : m.cull;
    m.error = error;
    m.parent = Sphere{{0, 0, 0, 0}};
    m.parent_error = kErrorInfinity;
    m.tris = tris;
    m.verts = verts;
    m.group = -1;
    all.push_back(m);
    return static_cast<int>(all.size()) - 1;
  }

  // Splits reps into <= 8 x 8 tiles and packs them into meshlets of <= 128 triangles and vertices.
  std::vector<int> emit(const std::vector<const Rep*>& reps, int level, const Sphere* lod, float error) {
    std::vector<Piece> pieces;
    for (const Rep* r : reps)
      for (int j0 = 0; j0 < r->ry; j0 += kTile)
        for (int i0 = 0; i0 < r->rx; i0 += kTile)
          pieces.push_back({r, i0, std::min(r->rx, i0 + kTile), j0, std::min(r->ry, j0 + kTile)});
    std::vector<int> out;
    std::vector<Piece> cur;
    uint32_t tris = 0, verts = 0;
    for (const Piece& pc : pieces) {
      if (!cur.empty() && (tris + pc.tris() > kMaxMeshletTris || verts + pc.verts() > kMaxMeshletVerts)) {
        out.push_back(add_meshlet(cur, level, lod, error));
        cur.clear();
        tris = verts = 0;
      }
      cur.push_back(pc);
      tris += pc.tris();
      verts += pc.verts();
    }
    if (!cur.empty()) out.push_back(add_meshlet(cur, level, lod, error));
    return out;
  }

  // Children are replaced by outputs: record the group, the children's parent sphere and error.
  void link(const std::vector<int>& children, const std::vector<int>& outputs, const Sphere& sphere, float error) {
    int g = static_cast<int>(group_outputs.size());
    group_outputs.push_back(outputs);
    for (int c : children) {
      all[c].parent = sphere;
      all[c].parent_error = error;
      all[c].group = g;
    }
  }
};

struct KdNode {
  int x0, x1, y0, y1;  // leaf-node index range
  int a = -1, b = -1;  // children
  int height = 0;
  std::unique_ptr<Rep> rep;
  std::vector<int> meshlets;
  float error = 0;
};

int build_kd(std::vector<KdNode>& nodes, int x0, int x1, int y0, int y1) {
  int id = static_cast<int>(nodes.size());
  nodes.emplace_back();
  nodes[id].x0 = x0;
  nodes[id].x1 = x1;
  nodes[id].y0 = y0;
  nodes[id].y1 = y1;
  if (x1 - x0 == 1 && y1 - y0 == 1) return id;
  int a, b;
  if (x1 - x0 >= y1 - y0) {
    int mid = x0 + (x1 - x0) / 2;
    a = build_kd(nodes, x0, mid, y0, y1);
    b = build_kd(nodes, mid, x1, y0, y1);
  } else {
    int mid = y0 + (y1 - y0) / 2;
    a = build_kd(nodes, x0, x1, y0, mid);
    b = build_kd(nodes, x0, x1, mid, y1);
  }
  nodes[id].a = a;
  nodes[id].b = b;
  nodes[id].height = 1 + std::max(nodes[a].height, nodes[b].height);
  return id;
}

}  // namespace

struct vgs_mesh {
  std::vector<float> bounds;
  std::vector<uint32_t> info;
  std::vector<uint32_t> level_ranges;
  std::vector<float> level_errors, level_guard;
  std::vector<int32_t> rgroup;
  std::vector<uint32_t> rstart, rindices;
  std::vector<uint32_t> triangles;
  float bounding[4], lodb[4], cullb[4];
  uint64_t leaf_tris = 0;
  uint32_t leaf_meshlets = 0, root_tris = 0, root_meshlets = 0;
};

extern "C" int vgs_mesh_build(const vgs_mesh_params* params, vgs_mesh** out) {
  *out = nullptr;
  std::vector<Patch> patches = make_patches(*params);
  if (patches.empty()) return -1;
  Builder B;

  // Per patch kd-tree over leaf nodes.
  std::vector<std::vector<KdNode>> trees(patches.size());
  std::vector<int> roots(patches.size());
  int max_height = 0;
  for (size_t p = 0; p < patches.size(); p++) {
    int lx = patches[p].nu / kLeafQuads, ly = patches[p].nv / kLeafQuads;
    roots[p] = build_kd(trees[p], 0, lx, 0, ly);
    max_height = std::max(max_height, trees[p][roots[p]].height);
  }
  auto rect = [&](size_t p, const KdNode& n, double& s0, double& s1, double& t0, double& t1) {
    int lx = patches[p].nu / kLeafQuads, ly = patches[p].nv / kLeafQuads;
    s0 = static_cast<double>(n.x0) / lx;
    s1 = static_cast<double>(n.x1) / lx;
    t0 = static_cast<double>(n.y0) / ly;
    t1 = static_cast<double>(n.y1) / ly;
  };

  // Level 0: leaf nodes at full resolution, 4 meshlets each, LOD sphere = culling sphere, error 0.
  for (size_t p = 0; p < patches.size(); p++)
    for (KdNode& n : trees[p]) {
      if (n.height != 0) continue;
      double s0, s1, t0, t1;
      rect(p, n, s0, s1, t0, t1);
      n.rep = std::make_unique<Rep>(make_rep(patches[p], s0, s1, t0, t1, kLeafQuads, kLeafQuads));
      n.meshlets = B.emit({n.rep.get()}, 0, nullptr, 0.0f);
    }

  // Levels 1..max_height: merge sibling nodes.
  for (int h = 1; h <= max_height; h++)
    for (size_t p = 0; p < patches.size(); p++)
      for (KdNode& n : trees[p]) {
        if (n.height != h) continue;
        KdNode& A = trees[p][n.a];
        KdNode& C = trees[p][n.b];
        double s0, s1, t0, t1;
        rect(p, n, s0, s1, t0, t1);
        n.rep = std::make_unique<Rep>(make_rep(patches[p], s0, s1, t0, t1, kLeafQuads, kLeafQuads));
        double shape = std::max(shape_error(*A.rep, *n.rep), shape_error(*C.rep, *n.rep));
        double child_error = std::max(A.error, C.error);
        n.error = static_cast<float>(child_error + std::max(shape, 1e-7));
        std::vector<int> children = A.meshlets;
        children.insert(children.end(), C.meshlets.begin(), C.meshlets.end());
        std::vector<Sphere> spheres;
        for (int c : children) spheres.push_back(B.all[c].lod);
        Sphere lod = enclosing(spheres);
        n.meshlets = B.emit({n.rep.get()}, h, &lod, n.error);
        B.link(children, n.meshlets, lod, n.error);
        A.rep.reset();
        C.rep.reset();
      }

  // Tail: all patch roots form one group per level; every patch's grid is halved (larger side first).
  std::vector<std::unique_ptr<Rep>> reps(patches.size());
  std::vector<int> current;
  float error = 0;
  for (size_t p = 0; p < patches.size(); p++) {
    KdNode& r = trees[p][roots[p]];
    reps[p] = std::move(r.rep);
    current.insert(current.end(), r.meshlets.begin(), r.meshlets.end());
    error = std::max(error, r.error);
  }
  trees.clear();
  int level = max_height + 1;
  while (reps[0]->rx * reps[0]->ry > 1) {
    std::vector<std::unique_ptr<Rep>> next(patches.size());
    double shape = 0;
    for (size_t p = 0; p < patches.size(); p++) {
      const Rep& r = *reps[p];
      int rx = r.rx, ry = r.ry;
      if (rx >= ry && rx > 1) rx /= 2;
      else if (ry > 1) ry /= 2;
      next[p] = std::make_unique<Rep>(make_rep(patches[p], 0, 1, 0, 1, rx, ry));
      shape = std::max(shape, shape_error(r, *next[p]));
    }
    float e = static_cast<float>(static_cast<double>(error) + std::max(shape, 1e-7));
    std::vector<Sphere> spheres;
    for (int c : current) spheres.push_back(B.all[c].lod);
    Sphere lod = enclosing(spheres);
    std::vector<const Rep*> rp;
    for (auto& r : next) rp.push_back(r.get());
    std::vector<int> outputs = B.emit(rp, level, &lod, e);
    B.link(current, outputs, lod, e);
    current = outputs;
    error = e;
    reps = std::move(next);
    level++;
  }

  // pack(): stable sort by level (buildVirtualMesh.ts:486).
  const int n = static_cast<int>(B.all.size());
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return B.all[a].level < B.all[b].level; });
  std::vector<int> index_of(n);
  for (int i = 0; i < n; i++) index_of[order[i]] = i;

  auto* mesh = new vgs_mesh();
  mesh->bounds.assign(static_cast<size_t>(n) * 16, 0.0f);
  mesh->info.assign(static_cast<size_t>(n) * 4, 0u);
  mesh->rgroup.assign(n, -1);
  uint32_t first_index = 0, vertex_offset = 0;
  int levels = B.all[order[n - 1]].level + 1;
  mesh->level_ranges.assign(levels * 2, 0u);
  mesh->level_errors.assign(levels * 2, 0.0f);
  mesh->level_guard.assign(levels * 2, 0.0f);
  for (int l = 0; l < levels; l++) {
    mesh->level_errors[l * 2] = INFINITY;
    mesh->level_errors[l * 2 + 1] = 0;
    mesh->level_guard[l * 2] = 0;
    mesh->level_guard[l * 2 + 1] = INFINITY;
  }
  std::vector<Sphere> leaf_cull, lod_all, cull_all;
  for (int i = 0; i < n; i++) {
    const WipMeshlet& m = B.all[order[i]];
    float* b = &mesh->bounds[static_cast<size_t>(i) * 16];
    for (int k = 0; k < 4; k++) {
      b[k] = static_cast<float>(m.lod.c[k]);
      b[4 + k] = static_cast<float>(m.parent.c[k]);
      b[12 + k] = static_cast<float>(m.cull.c[k]);
    }
    b[8] = m.error;
    b[9] = m.parent_error;
    uint32_t* o = &mesh->info[static_cast<size_t>(i) * 4];
    o[0] = m.tris;
    o[1] = first_index;
    o[2] = static_cast<uint32_t>(m.level);
    o[3] = vertex_offset;
    first_index += m.tris * 3;
    vertex_offset += m.verts;
    mesh->triangles.insert(mesh->triangles.end(), m.tri.begin(), m.tri.end());
    mesh->rgroup[i] = m.group;
    int l = m.level;
    if (mesh->level_ranges[l * 2 + 1] == 0) mesh->level_ranges[l * 2] = static_cast<uint32_t>(i);
    mesh->level_ranges[l * 2 + 1]++;
    mesh->level_errors[l * 2] = std::min(mesh->level_errors[l * 2], m.error);
    mesh->level_errors[l * 2 + 1] = std::max(mesh->level_errors[l * 2 + 1], m.parent_error);
    mesh->level_guard[l * 2] = std::max(mesh->level_guard[l * 2], m.error);
    mesh->level_guard[l * 2 + 1] = std::min(mesh->level_guard[l * 2 + 1], m.parent_error);
    if (m.level == 0) {
      mesh->leaf_tris += m.tris;
      mesh->leaf_meshlets++;
      leaf_cull.push_back(m.cull);
    }
    if (m.parent_error >= kErrorInfinity) {
      mesh->root_tris += m.tris;
      mesh->root_meshlets++;
    }
    lod_all.push_back(m.lod);
    if (m.parent_error < kErrorInfinity) lod_all.push_back(m.parent);
    cull_all.push_back(m.cull);
  }
  // Replacement lists, one per group, in output order (buildVirtualMesh.ts:491-509).
  const int groups = static_cast<int>(B.group_outputs.size());
  mesh->rstart.assign(groups + 1, 0u);
  for (int g = 0; g < groups; g++) mesh->rstart[g + 1] = mesh->rstart[g] + static_cast<uint32_t>(B.group_outputs[g].size());
  mesh->rindices.resize(mesh->rstart[groups]);
  for (int g = 0; g < groups; g++)
    for (size_t k = 0; k < B.group_outputs[g].size(); k++) mesh->rindices[mesh->rstart[g] + k] = index_of[B.group_outputs[g][k]];
  Sphere s = enclosing(leaf_cull);
  Sphere L = enclosing(lod_all);
  Sphere C = enclosing(cull_all);
  for (int k = 0; k < 4; k++) {
    mesh->bounding[k] = static_cast<float>(s.c[k]);
    mesh->lodb[k] = static_cast<float>(L.c[k]);
    mesh->cullb[k] = static_cast<float>(C.c[k]);
  }
  *out = mesh;
  return 0;
}

extern "C" void vgs_mesh_get(const vgs_mesh* m, vgs_mesh_view* v) {
  std::memset(v, 0, sizeof(*v));
  v->meshlet_count = static_cast<uint32_t>(m->info.size() / 4);
  v->meshlet_bounds = m->bounds.data();
  v->meshlet_info = m->info.data();
  v->level_count = static_cast<uint32_t>(m->level_ranges.size() / 2);
  v->level_ranges = m->level_ranges.data();
  v->level_errors = m->level_errors.data();
  v->level_guard_errors = m->level_guard.data();
  v->replacement_group = m->rgroup.data();
  v->group_count = static_cast<uint32_t>(m->rstart.size() - 1);
  v->replacement_start = m->rstart.data();
  v->replacement_indices = m->rindices.data();
  std::memcpy(v->bounding_sphere, m->bounding, sizeof(m->bounding));
  std::memcpy(v->lod_bounds_sphere, m->lodb, sizeof(m->lodb));
  std::memcpy(v->cull_bounds_sphere, m->cullb, sizeof(m->cullb));
  v->leaf_triangles = m->leaf_tris;
  v->leaf_meshlets = m->leaf_meshlets;
  v->root_triangles = m->root_tris;
  v->root_meshlets = m->root_meshlets;
  v->meshlet_triangles = m->triangles.data();
  v->triangle_count = m->triangles.size();
}

extern "C" void vgs_mesh_free(vgs_mesh* m) { delete m; }

// ------------------------------------------------------------------ indexed triangle meshes

namespace {

// Characteristic size the noise amplitude is a fraction of (as the patch displacements use it).
double kind_scale(const vgs_mesh_params& m) {
  switch (m.kind) {
    case VGS_SPHERE: return m.size[0];
    case VGS_BOX: return std::max(m.size[0], std::max(m.size[1], m.size[2]));
    case VGS_COLUMN: return std::max(m.size[0], m.size[1] * 0.5);
    case VGS_ARCH: return std::max(m.size[0], m.size[1]);
    default: return 1.0;
  }
}

// Merges the points flagged in `candidate` that lie closer than `eps` to an earlier flagged point (grid
// hash over cells of side eps, 27-cell search); other points keep their own ids. Returns the new id of
// every input point and compacts `pts`.
std::vector<uint32_t> weld_points(std::vector<V3>& pts, const std::vector<uint8_t>& candidate, double eps) {
  struct CellHash {
    size_t operator()(const std::array<int64_t, 3>& c) const {
      return static_cast<size_t>((c[0] * 73856093) ^ (c[1] * 19349663) ^ (c[2] * 83492791));
    }
  };
  std::unordered_map<std::array<int64_t, 3>, std::vector<uint32_t>, CellHash> cells;
  std::vector<uint32_t> id(pts.size());
  std::vector<V3> out;
  out.reserve(pts.size());
  for (size_t i = 0; i < pts.size(); i++) {
    const V3 p = pts[i];
    if (!candidate[i]) {
      id[i] = static_cast<uint32_t>(out.size());
      out.push_back(p);
      continue;
    }
    const std::array<int64_t, 3> c{static_cast<int64_t>(std::floor(p.x / eps)), static_cast<int64_t>(std::floor(p.y / eps)),
                                   static_cast<int64_t>(std::floor(p.z / eps))};
    int64_t found = -1;
    for (int dx = -1; dx <= 1 && found < 0; dx++)
      for (int dy = -1; dy <= 1 && found < 0; dy++)
        for (int dz = -1; dz <= 1 && found < 0; dz++) {
          auto it = cells.find({c[0] + dx, c[1] + dy, c[2] + dz});
          if (it == cells.end()) continue;
          for (uint32_t q : it->second)
            if (len(out[q] - p) <= eps) {
              found = q;
              break;
            }
        }
    if (found < 0) {
      found = static_cast<int64_t>(out.size());
      out.push_back(p);
      cells[c].push_back(static_cast<uint32_t>(found));
    }
    id[i] = static_cast<uint32_t>(found);
  }
  pts.swap(out);
  return id;
}

// Makes the winding consistent across edge-adjacent triangles, then turns every connected piece so
// that its signed volume is positive (outward normals on closed pieces).
void orient(std::vector<uint32_t>& idx, const std::vector<V3>& pts) {
  const size_t nt = idx.size() / 3;
  struct E {
    uint64_t key;
    uint32_t tri;
    uint8_t dir;
  };
  std::vector<E> edges;
  edges.reserve(idx.size());
  for (size_t t = 0; t < nt; t++)
    for (int k = 0; k < 3; k++) {
      const uint32_t a = idx[t * 3 + k], b = idx[t * 3 + (k + 1) % 3];
      edges.push_back({a < b ? (static_cast<uint64_t>(a) << 32) | b : (static_cast<uint64_t>(b) << 32) | a,
                       static_cast<uint32_t>(t), static_cast<uint8_t>(a < b)});
    }
  std::sort(edges.begin(), edges.end(), [](const E& x, const E& y) { return x.key < y.key || (x.key == y.key && x.tri < y.tri); });
  std::vector<std::vector<std::pair<uint32_t, uint8_t>>> nb(nt);  // (neighbour, my dir ^ its dir)
  for (size_t i = 0; i + 1 < edges.size();) {
    size_t j = i;
    while (j < edges.size() && edges[j].key == edges[i].key) j++;
    if (j - i == 2) {
      nb[edges[i].tri].push_back({edges[i + 1].tri, static_cast<uint8_t>(edges[i].dir ^ edges[i + 1].dir)});
      nb[edges[i + 1].tri].push_back({edges[i].tri, static_cast<uint8_t>(edges[i].dir ^ edges[i + 1].dir)});
    }
    i = j;
  }
  std::vector<int8_t> flip(nt, -1);
  std::vector<uint32_t> queue, piece;
  for (size_t s = 0; s < nt; s++) {
    if (flip[s] >= 0) continue;
    flip[s] = 0;
    queue.assign(1, static_cast<uint32_t>(s));
    piece.clear();
    while (!queue.empty()) {
      const uint32_t t = queue.back();
      queue.pop_back();
      piece.push_back(t);
      for (auto [u, differ] : nb[t]) {
        // Shared edge directions must differ once flips apply: f_u = f_t ^ (d_t ^ d_u) ^ 1.
        const int8_t want = static_cast<int8_t>(flip[t] ^ differ ^ 1);
        if (flip[u] < 0) {
          flip[u] = want;
          queue.push_back(u);
        }
      }
    }
    double volume = 0;
    for (uint32_t t : piece) {
      V3 a = pts[idx[t * 3]], b = pts[idx[t * 3 + 1]], c = pts[idx[t * 3 + 2]];
      if (flip[t]) std::swap(b, c);
      volume += dot(a, V3{b.y * c.z - b.z * c.y, b.z * c.x - b.x * c.z, b.x * c.y - b.y * c.x});
    }
    if (volume < 0)
      for (uint32_t t : piece) flip[t] ^= 1;
  }
  for (size_t t = 0; t < nt; t++)
    if (flip[t]) std::swap(idx[t * 3 + 1], idx[t * 3 + 2]);
}

}  // namespace

extern "C" int vgs_mesh_triangles(const vgs_mesh_params* params, vgs_triangles* out) {
  std::memset(out, 0, sizeof(*out));
  vgs_mesh_params flat = *params;
  flat.noise_amplitude = 0.0;
  std::vector<Patch> patches = make_patches(flat);
  if (patches.empty()) return -1;
  const bool column = params->kind == VGS_COLUMN, arch = params->kind == VGS_ARCH;
  if (arch) {
    // The feet close the tube's ends: their grids match the tube's side resolution.
    const int n = std::max(1, patches[0].nv / 4);
    for (size_t p = 1; p < patches.size(); p++) patches[p].nu = patches[p].nv = n;
  }
  std::vector<V3> pts;
  std::vector<uint8_t> rim;  // on a patch's boundary: may coincide with another patch's point
  std::vector<uint32_t> idx;
  for (size_t p = 0; p < patches.size(); p++) {
    if (column && p > 0) continue;  // the caps are fans over the side's rim (below)
    const Patch& pa = patches[p];
    const uint32_t base = static_cast<uint32_t>(pts.size()), row = static_cast<uint32_t>(pa.nu + 1);
    for (int j = 0; j <= pa.nv; j++)
      for (int i = 0; i <= pa.nu; i++) {
        pts.push_back(pa.eval(static_cast<double>(i) / pa.nu, static_cast<double>(j) / pa.nv));
        rim.push_back(i == 0 || j == 0 || i == pa.nu || j == pa.nv);
      }
    for (int j = 0; j < pa.nv; j++)
      for (int i = 0; i < pa.nu; i++) {
        const uint32_t v00 = base + static_cast<uint32_t>(j) * row + static_cast<uint32_t>(i);
        const uint32_t v10 = v00 + 1, v01 = v00 + row, v11 = v01 + 1;
        idx.insert(idx.end(), {v00, v10, v11, v00, v11, v01});
      }
  }
  if (column) {
    const Patch& side = patches[0];
    const uint32_t row = static_cast<uint32_t>(side.nu + 1);
    for (int top = 0; top < 2; top++) {
      const uint32_t center = static_cast<uint32_t>(pts.size());
      pts.push_back(V3{0, top ? params->size[1] : 0.0, 0});
      rim.push_back(0);
      const uint32_t ring = top ? static_cast<uint32_t>(side.nv) * row : 0;
      for (int i = 0; i < side.nu; i++) idx.insert(idx.end(), {center, ring + i, ring + i + 1});
    }
  }
  const double scale = kind_scale(*params);
  const std::vector<uint32_t> id = weld_points(pts, rim, 1e-6 * std::max(scale, 1e-9));
  std::vector<uint32_t> tri;
  tri.reserve(idx.size());
  for (size_t t = 0; t < idx.size(); t += 3) {
    const uint32_t a = id[idx[t]], b = id[idx[t + 1]], c = id[idx[t + 2]];
    if (a == b || b == c || a == c) continue;
    tri.insert(tri.end(), {a, b, c});
  }
  orient(tri, pts);
  if (params->kind != VGS_TERRAIN && params->noise_amplitude != 0.0) {
    // Displacement along the area-weighted vertex normal: continuous across patch seams and edges.
    std::vector<V3> n(pts.size(), V3{0, 0, 0});
    for (size_t t = 0; t < tri.size(); t += 3) {
      const V3 a = pts[tri[t]], b = pts[tri[t + 1]], c = pts[tri[t + 2]];
      const V3 e1 = b - a, e2 = c - a;
      const V3 f{e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x};
      for (int k = 0; k < 3; k++) n[tri[t + k]] = n[tri[t + k]] + f;
    }
    const double amp = params->noise_amplitude * scale, freq = params->noise_frequency;
    for (size_t v = 0; v < pts.size(); v++) pts[v] = pts[v] + normalize(n[v]) * (amp * fbm3(pts[v] * freq, params->seed));
  }
  out->vertex_count = static_cast<uint32_t>(pts.size());
  out->index_count = tri.size();
  out->positions = static_cast<float*>(std::malloc(pts.size() * 3 * sizeof(float)));
  out->indices = static_cast<uint32_t*>(std::malloc(std::max<size_t>(1, tri.size()) * sizeof(uint32_t)));
  if (!out->positions || !out->indices) {
    vgs_triangles_free(out);
    return -2;
  }
  for (size_t v = 0; v < pts.size(); v++) {
    out->positions[v * 3] = static_cast<float>(pts[v].x);
    out->positions[v * 3 + 1] = static_cast<float>(pts[v].y);
    out->positions[v * 3 + 2] = static_cast<float>(pts[v].z);
  }
  std::memcpy(out->indices, tri.data(), tri.size() * sizeof(uint32_t));
  return 0;
}

extern "C" void vgs_triangles_free(vgs_triangles* t) {
  std::free(t->positions);
  std::free(t->indices);
  t->positions = nullptr;
  t->indices = nullptr;
}

// ------------------------------------------------------------------ placement

extern "C" uint32_t vgs_poisson_disk(double x0, double z0, double w, double h, double radius, uint32_t seed,
                                     uint32_t candidates, uint32_t max_points, float* out_xz) {
  // Bridson, "Fast Poisson disk sampling in arbitrary dimensions" (SIGGRAPH 2007 sketch).
  const double cell = radius / std::sqrt(2.0);
  const int gw = static_cast<int>(std::ceil(w / cell)), gh = static_cast<int>(std::ceil(h / cell));
  std::vector<int32_t> grid(static_cast<size_t>(gw) * gh, -1);
  std::vector<double> pts;
  std::vector<uint32_t> active;
  SplitMix rng(seed);
  auto cell_of = [&](double x, double z, int& i, int& j) {
    i = std::min(gw - 1, std::max(0, static_cast<int>((x - x0) / cell)));
    j = std::min(gh - 1, std::max(0, static_cast<int>((z - z0) / cell)));
  };
  auto add = [&](double x, double z) {
    uint32_t id = static_cast<uint32_t>(pts.size() / 2);
    pts.push_back(x);
    pts.push_back(z);
    int i, j;
    cell_of(x, z, i, j);
    grid[static_cast<size_t>(j) * gw + i] = static_cast<int32_t>(id);
    active.push_back(id);
  };
  add(x0 + rng.uniform() * w, z0 + rng.uniform() * h);
  const double r2 = radius * radius;
  while (!active.empty() && pts.size() / 2 < max_points) {
    size_t k = static_cast<size_t>(rng.uniform() * active.size());
    uint32_t id = active[k];
    double px = pts[id * 2], pz = pts[id * 2 + 1];
    bool placed = false;
    for (uint32_t c = 0; c < candidates && !placed; c++) {
      double a = rng.uniform() * 2 * M_PI, d = radius * (1 + rng.uniform());
      double x = px + d * std::cos(a), z = pz + d * std::sin(a);
      if (x < x0 || x >= x0 + w || z < z0 || z >= z0 + h) continue;
      int i, j;
      cell_of(x, z, i, j);
      bool ok = true;
      for (int dj = -2; dj <= 2 && ok; dj++)
        for (int di = -2; di <= 2 && ok; di++) {
          int ii = i + di, jj = j + dj;
          if (ii < 0 || jj < 0 || ii >= gw || jj >= gh) continue;
          int32_t q = grid[static_cast<size_t>(jj) * gw + ii];
          if (q < 0) continue;
          double dx = pts[q * 2] - x, dz = pts[q * 2 + 1] - z;
          if (dx * dx + dz * dz < r2) ok = false;
        }
      if (ok) {
        add(x, z);
        placed = true;
      }
    }
    if (!placed) {
      active[k] = active.back();
      active.pop_back();
    }
  }
  uint32_t count = static_cast<uint32_t>(pts.size() / 2);
  for (uint32_t i = 0; i < count; i++) {
    out_xz[i * 2] = static_cast<float>(pts[i * 2]);
    out_xz[i * 2 + 1] = static_cast<float>(pts[i * 2 + 1]);
  }
  return count;
}

namespace {
double height_at(double x, double z, double amplitude, double frequency, uint32_t seed) {
  return amplitude * fbm3(V3{x * frequency, 0.37, z * frequency}, seed);
}
}  // namespace

extern "C" void vgs_heightfield(const float* xz, uint32_t n, double amplitude, double frequency, uint32_t seed, float* out_y) {
  for (uint32_t i = 0; i < n; i++) out_y[i] = static_cast<float>(height_at(xz[i * 2], xz[i * 2 + 1], amplitude, frequency, seed));
}

extern "C" void vgs_heightfield_depth(const float* vp, uint32_t width, uint32_t height, double cx, double cz, double extent,
                                      uint32_t grid, double amplitude, double frequency, uint32_t seed, float* out_depth) {
  const size_t pixels = static_cast<size_t>(width) * height;
  for (size_t i = 0; i < pixels; i++) out_depth[i] = 1.0f;
  struct Clip {
    double x, y, z, w;
  };
  const uint32_t nv = grid + 1;
  std::vector<Clip> clip(static_cast<size_t>(nv) * nv);
  for (uint32_t j = 0; j < nv; j++)
    for (uint32_t i = 0; i < nv; i++) {
      double x = cx - extent + 2 * extent * i / grid, z = cz - extent + 2 * extent * j / grid;
      double y = height_at(x, z, amplitude, frequency, seed);
      Clip c;
      c.x = vp[0] * x + vp[4] * y + vp[8] * z + vp[12];
      c.y = vp[1] * x + vp[5] * y + vp[9] * z + vp[13];
      c.z = vp[2] * x + vp[6] * y + vp[10] * z + vp[14];
      c.w = vp[3] * x + vp[7] * y + vp[11] * z + vp[15];
      clip[static_cast<size_t>(j) * nv + i] = c;
    }
  auto raster = [&](const Clip& a, const Clip& b, const Clip& c) {
    // Triangles touching the near side are skipped: missing occluders only cull less.
    if (a.w <= 1e-3 || b.w <= 1e-3 || c.w <= 1e-3) return;
    double sx[3], sy[3], sz[3];
    const Clip* v[3] = {&a, &b, &c};
    for (int k = 0; k < 3; k++) {
      sx[k] = (v[k]->x / v[k]->w * 0.5 + 0.5) * width;
      sy[k] = (0.5 - v[k]->y / v[k]->w * 0.5) * height;
      sz[k] = v[k]->z / v[k]->w;
    }
    double area = (sx[1] - sx[0]) * (sy[2] - sy[0]) - (sx[2] - sx[0]) * (sy[1] - sy[0]);
    if (std::fabs(area) < 1e-12) return;
    int xmin = std::max(0, static_cast<int>(std::floor(std::min({sx[0], sx[1], sx[2]}))));
    int xmax = std::min(static_cast<int>(width) - 1, static_cast<int>(std::ceil(std::max({sx[0], sx[1], sx[2]}))));
    int ymin = std::max(0, static_cast<int>(std::floor(std::min({sy[0], sy[1], sy[2]}))));
    int ymax = std::min(static_cast<int>(height) - 1, static_cast<int>(std::ceil(std::max({sy[0], sy[1], sy[2]}))));
    for (int y = ymin; y <= ymax; y++)
      for (int x = xmin; x <= xmax; x++) {
        double px = x + 0.5, py = y + 0.5;
        double w0 = ((sx[1] - px) * (sy[2] - py) - (sx[2] - px) * (sy[1] - py)) / area;
        double w1 = ((sx[2] - px) * (sy[0] - py) - (sx[0] - px) * (sy[2] - py)) / area;
        double w2 = 1 - w0 - w1;
        if (w0 < 0 || w1 < 0 || w2 < 0) continue;
        double z = w0 * sz[0] + w1 * sz[1] + w2 * sz[2];
        if (z < 0 || z > 1) continue;
        float& d = out_depth[static_cast<size_t>(y) * width + x];
        if (z < d) d = static_cast<float>(z);
      }
  };
  for (uint32_t j = 0; j < grid; j++)
    for (uint32_t i = 0; i < grid; i++) {
      const Clip& p00 = clip[static_cast<size_t>(j) * nv + i];
      const Clip& p10 = clip[static_cast<size_t>(j) * nv + i + 1];
      const Clip& p01 = clip[static_cast<size_t>(j + 1) * nv + i];
      const Clip& p11 = clip[static_cast<size_t>(j + 1) * nv + i + 1];
      raster(p00, p10, p01);
      raster(p11, p01, p10);
    }
}
