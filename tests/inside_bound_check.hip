// Host-side check of the cell pass's whole-cell frustum answers (vg::cell_inside_bits in csrc/vgcut_kernels.h,
// the function k_cells calls, and vgi::cell_bound in csrc/vgcut_cells.h, the host code that builds its per-cell
// bounds, both compiled for the host): over random and adversarial cells, instances and planes, whenever a bit
// says that every instance of the cell passes a frustum test, each instance's own binary32 test (the kernel's
// mat_mul, len3 and frustum_distance, in the reference's expression order) does pass it. Adversarial cases put
// a plane's distance from the cell centre within a few ulp of the bits' threshold, facing the cell's outermost
// instance, with zero radii. Two planted mistakes in the bounds (R_in one part in 10^3 too small, the bounding
// spheres' extent taken for the culling spheres') must each give violations. Prints one JSON line.
// Built and run by tests/test_inside_bound.py.
#include <cstdio>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../three-virtual-geometry_amd/csrc/vgcut_kernels.h"
#include "../three-virtual-geometry_amd/csrc/vgcut_cells.h"

namespace {

struct Rng {
  uint64_t s;
  double u() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return (s >> 11) * (1.0 / 9007199254740992.0);
  }
  double range(double a, double b) { return a + (b - a) * u(); }
  double logr(double a, double b) { return std::exp(range(std::log(a), std::log(b))); }
};

// Column-major rotation about a random unit axis, times a per-axis scale, plus a translation.
void instance_matrix(Rng& g, float m[16], const double at[3], double spread, bool uniform) {
  double ax = g.range(-1, 1), ay = g.range(-1, 1), az = g.range(-1, 1);
  const double n = std::sqrt(ax * ax + ay * ay + az * az) + 1e-12;
  ax /= n, ay /= n, az /= n;
  const double a = g.range(0, 6.283185307179586), c = std::cos(a), s = std::sin(a), t = 1 - c;
  const double R[3][3] = {{t * ax * ax + c, t * ax * ay - s * az, t * ax * az + s * ay},
                          {t * ax * ay + s * az, t * ay * ay + c, t * ay * az - s * ax},
                          {t * ax * az - s * ay, t * ay * az + s * ax, t * az * az + c}};
  const double base = g.logr(0.01, 100.0);
  const double sc[3] = {base, uniform ? base : base * g.range(0.3, 3.0), uniform ? base : base * g.range(0.3, 3.0)};
  for (int col = 0; col < 3; col++)
    for (int row = 0; row < 3; row++) m[col * 4 + row] = static_cast<float>(R[row][col] * sc[col]);
  for (int row = 0; row < 3; row++) m[row * 4 + 3] = 0.0f;
  for (int row = 0; row < 3; row++) m[12 + row] = static_cast<float>(at[row] + g.range(-spread, spread));
  m[15] = 1.0f;
}

struct Verdict {
  bool camera, shadow, cull_camera, cull_shadow;
};

// The instance pass's own binary32 tests of one instance (k_instances: centre and scale as it derives them).
Verdict exact_tests(const vg::ViewDev& v, const float* m, const float obj[4], const float cull[4]) {
  const float4 c0 = make_float4(m[0], m[1], m[2], m[3]), c1 = make_float4(m[4], m[5], m[6], m[7]);
  const float4 c2 = make_float4(m[8], m[9], m[10], m[11]), c3 = make_float4(m[12], m[13], m[14], m[15]);
  const float scale = std::fmax(vg::len3(c0.x, c0.y, c0.z), std::fmax(vg::len3(c1.x, c1.y, c1.z), vg::len3(c2.x, c2.y, c2.z)));
  const float4 center = vg::mat_mul(c0, c1, c2, c3, obj[0], obj[1], obj[2], 1.0f);
  const float4 cc = vg::mat_mul(c0, c1, c2, c3, cull[0], cull[1], cull[2], 1.0f);
  const float radius = obj[3] * scale, cull_radius = cull[3] * scale;
  Verdict r;
  r.camera = vg::frustum_distance(v, center.x, center.y, center.z, 0, false) >= -radius;
  r.shadow = vg::frustum_distance(v, center.x, center.y, center.z, 1, false) >= -radius;
  r.cull_camera = vg::frustum_distance(v, cc.x, cc.y, cc.z, 0, false) >= cull_radius;
  r.cull_shadow = vg::frustum_distance(v, cc.x, cc.y, cc.z, 1, false) >= cull_radius;
  return r;
}

// Unit normal (binary32) and constant of a plane.
void set_plane(vg::ViewDev& v, int p, const double n[3], double constant) {
  const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) + 1e-300;
  for (int k = 0; k < 3; k++) v.planes[p * 4 + k] = static_cast<float>(n[k] / len);
  v.planes[p * 4 + 3] = static_cast<float>(constant);
}

void finish_view(vg::ViewDev& v) {
  for (int p = 0; p < 6; p++) v.sweep_reach[p] = std::fmax(vg::dot3(&v.planes[p * 4], v.sweep[0], v.sweep[1], v.sweep[2]), 0.0f);
  vg::frustum_bounds(v);
}

}  // namespace

int main(int argc, char** argv) {
  const long cases = argc > 1 ? std::atol(argv[1]) : 100000;
  Rng g{0x2545f4914f6cdd1dull};
  long certified[4] = {0, 0, 0, 0}, violations = 0, planted_radius = 0, planted_cull = 0, adversarial = 0, instances_total = 0;
  for (long k = 0; k < cases; k++) {
    const int mode = static_cast<int>(k % 4);  // 0: random planes; 1, 2: a plane at the threshold (bounding, culling); 3: non-uniform scale
    const bool tight = mode == 1 || mode == 2;
    // A cell: 1-128 instances spread around a point near the origin or far out.
    const double far = g.u() < 0.5 ? g.logr(1.0, 3e4) : g.logr(1e-3, 10.0);
    const double at[3] = {g.range(-far, far), g.range(-far, far), g.range(-far, far)};
    const double spread = g.logr(0.01, tight ? 2e3 : 500.0);
    const uint32_t count = 1 + static_cast<uint32_t>(g.u() * vg::kCellSize);
    std::vector<float> mats(static_cast<size_t>(count) * 16);
    for (uint32_t i = 0; i < count; i++) instance_matrix(g, &mats[static_cast<size_t>(i) * 16], at, spread, mode != 3);
    const double zero_r = tight || g.u() < 0.2;
    float obj[4] = {static_cast<float>(g.range(-2, 2)), static_cast<float>(g.range(-2, 2)), static_cast<float>(g.range(-2, 2)),
                    zero_r ? 0.0f : static_cast<float>(g.logr(0.01, 5.0))};
    float cull[4] = {static_cast<float>(obj[0] + g.range(-1, 1)), static_cast<float>(obj[1] + g.range(-1, 1)),
                     static_cast<float>(obj[2] + g.range(-1, 1)), zero_r ? 0.0f : static_cast<float>(g.logr(0.01, 5.0))};
    float4 sphere, rec;
    vgi::compute_cell(mats, 0, count, obj, 0, 0, sphere, rec);
    const float4 bound = vgi::cell_bound(mats, 0, count, obj, cull, sphere);
    // Planted mistakes: R_in one part in 10^3 too small; R_cull taken from the bounding spheres' centres.
    const float4 shrunk = make_float4(bound.x, bound.y * 0.999f, bound.z, 0.0f);
    const float4 cull_as_bounding = make_float4(bound.x, bound.y, bound.y, 0.0f);

    vg::ViewDev v{};
    v.sweep[0] = static_cast<float>(g.u() < 0.3 ? 0.0 : g.range(-300, 300));
    v.sweep[1] = static_cast<float>(g.u() < 0.3 ? 0.0 : g.range(-300, 300));
    v.sweep[2] = static_cast<float>(g.u() < 0.3 ? 0.0 : g.range(-300, 300));
    if (!tight) {
      for (int p = 0; p < 6; p++) {
        const double n[3] = {g.range(-1, 1), g.range(-1, 1), g.range(-1, 1)};
        // Planes that pass at up to a few cell radii from the centre, on either side.
        const double dist = g.range(-0.5, 3.0) * (bound.y + 1.0);
        const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) + 1e-300;
        set_plane(v, p, n, dist - (n[0] * sphere.x + n[1] * sphere.y + n[2] * sphere.z) / len);
      }
      finish_view(v);
    } else {
      // Plane 0 faces the cell's outermost instance (bounding or culling sphere), its distance from the
      // centre a few ulp either side of the bits' threshold; the other planes are far away.
      const bool cull_case = mode == 2;
      double best = -1, u[3] = {1, 0, 0};
      for (uint32_t i = 0; i < count; i++) {
        double s[4];
        vgi::instance_sphere(&mats[static_cast<size_t>(i) * 16], cull_case ? cull : obj, s);
        const double d[3] = {s[0] - sphere.x, s[1] - sphere.y, s[2] - sphere.z};
        const double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (len > best && len > 0) best = len, u[0] = d[0] / len, u[1] = d[1] / len, u[2] = d[2] / len;
      }
      const double n0[3] = {-u[0], -u[1], -u[2]};
      set_plane(v, 0, n0, 0.0);
      for (int p = 1; p < 6; p++) {  // the cell well inside them
        const double n[3] = {g.range(-1, 1), g.range(-1, 1), g.range(-1, 1)};
        const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) + 1e-300;
        set_plane(v, p, n, 4.0 * (bound.y + bound.z + 1.0) - (n[0] * sphere.x + n[1] * sphere.y + n[2] * sphere.z) / len);
      }
      if (g.u() < 0.5) v.sweep[0] = v.sweep[1] = v.sweep[2] = 0.0f;
      finish_view(v);
      // Threshold as cell_inside_bits computes it, then the constant that puts D(C) there (+- a few ulp).
      const float mag = std::fmax(std::fabs(sphere.x), std::fmax(std::fabs(sphere.y), std::fabs(sphere.z))) + std::fmax(bound.y, bound.z);
      const float err = 1e-5f * (2.0f * v.frustum_n2 * (mag + bound.x) + v.frustum_c);
      const float reach = v.frustum_n2 * 1.00001f;
      // Every other such case aims at the planted mistakes' thresholds instead (to show that they fail).
      const bool aim_planted = (k / 4) % 2 == 1;
      const float aimed = cull_case ? (aim_planted ? cull_as_bounding.z : bound.z) : (aim_planted ? shrunk.y : bound.y);
      const float target = reach * aimed + err;
      const float* n = &v.planes[0];
      float c = target - vg::dot3(n, sphere.x, sphere.y, sphere.z);
      const int ulps = static_cast<int>(g.u() * 8) - 4;
      for (int i = 0; i < std::abs(ulps); i++) c = std::nextafter(c, ulps < 0 ? -HUGE_VALF : HUGE_VALF);
      v.planes[3] = c;
      finish_view(v);
      adversarial++;
    }
    const uint32_t bits = vg::cell_inside_bits(v, sphere, bound, true);
    const uint32_t bits_shrunk = vg::cell_inside_bits(v, sphere, shrunk, true);
    const uint32_t bits_cull = vg::cell_inside_bits(v, sphere, cull_as_bounding, true);
    for (int b = 0; b < 4; b++) certified[b] += (bits >> (28 + b)) & 1u;
    bool bad = false, bad_radii = false, bad_cull = false;
    for (uint32_t i = 0; i < count; i++) {
      const Verdict t = exact_tests(v, &mats[static_cast<size_t>(i) * 16], obj, cull);
      const bool ok[4] = {t.camera, t.shadow, t.cull_camera, t.cull_shadow};
      for (int b = 0; b < 4; b++) {
        const uint32_t bit = vg::kQInCamera << b;
        bad = bad || ((bits & bit) && !ok[b]);
        bad_radii = bad_radii || ((bits_shrunk & bit) && !ok[b]);
        bad_cull = bad_cull || ((bits_cull & bit) && !ok[b]);
      }
    }
    instances_total += count;
    planted_radius += bad_radii;
    planted_cull += bad_cull;
    if (bad) {
      violations++;
      if (violations <= 5) std::fprintf(stderr, "violation: case %ld mode %d count %u bits %x\n", k, mode, count, bits);
    }
  }
  std::printf("{\"cases\": %ld, \"instances\": %ld, \"adversarial\": %ld, \"in_camera\": %ld, \"in_shadow\": %ld, \"cull_camera\": %ld, "
              "\"cull_shadow\": %ld, \"violations\": %ld, \"planted_shrunk\": %ld, \"planted_cull\": %ld}\n",
              cases, instances_total, adversarial, certified[0], certified[1], certified[2], certified[3], violations, planted_radius,
              planted_cull);
  return violations ? 1 : 0;
}
