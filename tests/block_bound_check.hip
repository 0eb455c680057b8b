// Host-side check of the meshlet pass's block test (vg::block_out_of_band in csrc/vgcut_kernels.h, the same
// function the kernel calls, compiled for the host): over random and adversarial blocks, instances, views
// and bands, a block the test drops never holds a run whose exact binary32 LOD test (vg::projected_error,
// the kernel's, in the reference's expression order) puts it in the band (own <= high and parent > low).
// Prints one JSON line of counts. Built and run by tests/test_block_bound.py.
#include <cstdio>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../three-virtual-geometry_amd/csrc/vgcut_kernels.h"
#include "../three-virtual-geometry_amd/csrc/vgcut_blocks.h"

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
void instance_matrix(Rng& g, float m[16], bool uniform) {
  double ax = g.range(-1, 1), ay = g.range(-1, 1), az = g.range(-1, 1);
  const double n = std::sqrt(ax * ax + ay * ay + az * az) + 1e-12;
  ax /= n, ay /= n, az /= n;
  const double a = g.range(0, 6.283185307179586), c = std::cos(a), s = std::sin(a), t = 1 - c;
  const double R[3][3] = {{t * ax * ax + c, t * ax * ay - s * az, t * ax * az + s * ay},
                          {t * ax * ay + s * az, t * ay * ay + c, t * ay * az - s * ax},
                          {t * ax * az - s * ay, t * ay * az + s * ax, t * az * az + c}};
  const double base = g.logr(0.01, 100.0);
  const double sc[3] = {base, uniform ? base : base * g.range(0.3, 3.0), uniform ? base : base * g.range(0.3, 3.0)};
  const double far = g.u() < 0.5 ? 1e4 : 30.0;
  for (int col = 0; col < 3; col++)
    for (int row = 0; row < 3; row++) m[col * 4 + row] = static_cast<float>(R[row][col] * sc[col]);
  for (int row = 0; row < 3; row++) m[row * 4 + 3] = 0.0f;
  m[12] = static_cast<float>(g.range(-far, far));
  m[13] = static_cast<float>(g.range(-far, far));
  m[14] = static_cast<float>(g.range(-far, far));
  m[15] = 1.0f;
}

// A camera at `eye` looking at `target` (rigid view matrix), as three.js's lookAt and inverse give it.
void view_matrix(const double eye[3], const double target[3], float v[16]) {
  double z[3] = {eye[0] - target[0], eye[1] - target[1], eye[2] - target[2]};
  double n = std::sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
  for (double& k : z) k /= n;
  double x[3] = {z[2], 0.0, -z[0]};  // up (0, 1, 0) x z
  n = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  if (n < 1e-9) {
    x[0] = 1, x[1] = 0, x[2] = 0;
    n = 1;
  }
  for (double& k : x) k /= n;
  const double y[3] = {z[1] * x[2] - z[2] * x[1], z[2] * x[0] - z[0] * x[2], z[0] * x[1] - z[1] * x[0]};
  const double* b[3] = {x, y, z};
  for (int row = 0; row < 3; row++) {
    for (int col = 0; col < 3; col++) v[col * 4 + row] = static_cast<float>(b[row][col]);
    v[12 + row] = static_cast<float>(-(b[row][0] * eye[0] + b[row][1] * eye[1] + b[row][2] * eye[2]));
    v[row * 4 + 3] = 0.0f;
  }
  v[15] = 1.0f;
}

}  // namespace

int main(int argc, char** argv) {
  const long cases = argc > 1 ? std::atol(argv[1]) : 200000;
  Rng g{0x9e3779b97f4a7c15ull};
  long planted_violations = 0;
  long dropped = 0, violations = 0, in_band_blocks = 0, adversarial = 0, adversarial_dropped = 0, runs_total = 0;
  for (long k = 0; k < cases; k++) {
    const int mode = static_cast<int>(k % 5);  // 0, 1: random bands; 2, 3: bands at a run's own / parent value; 4: non-uniform scale
    float m[16];
    instance_matrix(g, m, mode != 4);
    const float4 c0 = make_float4(m[0], m[1], m[2], m[3]), c1 = make_float4(m[4], m[5], m[6], m[7]);
    const float4 c2 = make_float4(m[8], m[9], m[10], m[11]), c3 = make_float4(m[12], m[13], m[14], m[15]);
    const float s = std::fmax(vg::len3(c0.x, c0.y, c0.z), std::fmax(vg::len3(c1.x, c1.y, c1.z), vg::len3(c2.x, c2.y, c2.z)));
    // A block: runs around an object-space point, LOD radii, errors, parent spheres.
    const int count = 1 + static_cast<int>(g.u() * vg::kBlockRuns);
    const double spread = g.logr(0.001, 5.0), ox = g.range(-3, 3), oy = g.range(-3, 3), oz = g.range(-3, 3);
    std::vector<float4> lod(count), par(count);
    std::vector<float2> err(count);
    const double e0 = g.logr(1e-5, 0.5);
    for (int r = 0; r < count; r++) {
      lod[r] = make_float4(static_cast<float>(ox + g.range(-spread, spread)), static_cast<float>(oy + g.range(-spread, spread)),
                           static_cast<float>(oz + g.range(-spread, spread)), static_cast<float>(g.logr(0.01, 2.0)));
      par[r] = make_float4(static_cast<float>(lod[r].x + g.range(-spread, spread)), static_cast<float>(lod[r].y + g.range(-spread, spread)),
                           static_cast<float>(lod[r].z + g.range(-spread, spread)), static_cast<float>(lod[r].w * g.range(1.0, 2.5)));
      const double e = g.u() < 0.05 ? 0.0 : e0 * g.range(0.3, 3.0);
      err[r] = make_float2(static_cast<float>(e), g.u() < 0.05 ? static_cast<float>(1e30) : static_cast<float>(e * g.range(1.0, 4.0) + 1e-6));
    }
    float4 blod, bpar, berr;
    vgi::block_bounds(lod.data(), par.data(), err.data(), 0, count, &blod, &bpar, &berr);
    // A camera near or far from the instance.
    const double d = g.logr(0.05, 2e4);
    const double tx = c3.x + g.range(-5, 5), ty = c3.y + g.range(-5, 5), tz = c3.z + g.range(-5, 5);
    double dir[3] = {g.range(-1, 1), g.range(-1, 1), g.range(-1, 1)};
    const double dn = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]) + 1e-12;
    const double eye[3] = {tx + d * dir[0] / dn, ty + d * dir[1] / dn, tz + d * dir[2] / dn}, target[3] = {tx, ty, tz};
    vg::ViewDev v{};
    view_matrix(eye, target, v.view);
    v.near_plane = static_cast<float>(g.logr(0.01, 1.0));
    v.proj_scale = static_cast<float>(g.range(100, 3000));
    vg::view_bounds(v);
    // The runs' exact values and a band.
    std::vector<float> own(count), parent(count);
    for (int r = 0; r < count; r++) {
      own[r] = vg::projected_error<true>(v, c0, c1, c2, c3, s, lod[r], err[r].x);
      parent[r] = vg::projected_error<true>(v, c0, c1, c2, c3, s, par[r], err[r].y);
    }
    float low, high;
    if (mode == 2) {  // high a few ulp either side of the smallest own value; parents all above low
      float lo_own = own[0];
      for (float x : own) lo_own = std::fmin(lo_own, x);
      const int ulps = static_cast<int>(g.u() * 8) - 4;
      high = lo_own;
      for (int i = 0; i < std::abs(ulps); i++) high = std::nextafter(high, ulps < 0 ? 0.0f : HUGE_VALF);
      low = std::fmin(high * 0.5f, 1e-30f + lo_own * 0.25f);
      adversarial++;
    } else if (mode == 3) {  // low a few ulp either side of the largest finite parent value; own values below high
      float hi_par = 0;
      for (float x : parent)
        if (x < 1e29f) hi_par = std::fmax(hi_par, x);
      const int ulps = static_cast<int>(g.u() * 8) - 4;
      low = hi_par;
      for (int i = 0; i < std::abs(ulps); i++) low = std::nextafter(low, ulps < 0 ? 0.0f : HUGE_VALF);
      high = 1e30f;
      adversarial++;
    } else {
      low = static_cast<float>(g.logr(0.05, 20.0));
      high = low * static_cast<float>(g.range(1.0, 2.0));
    }
    bool any = false;
    for (int r = 0; r < count; r++) any = any || (own[r] <= high && parent[r] > low);
    const vg::BlockBounds ib = vg::instance_bounds(c0, c1, c2, c3, s);
    const bool drop = vg::block_out_of_band(v, c0, c1, c2, c3, ib, blod, bpar, berr, low, high);
    // A planted mistake on the CPU: the same test ignoring how far the runs' centres spread (no stretch).
    vg::BlockBounds flat = ib;
    flat.norm_m = 0.0f;
    planted_violations += vg::block_out_of_band(v, c0, c1, c2, c3, flat, blod, bpar, berr, low, high) && any;
    runs_total += count;
    in_band_blocks += any;
    dropped += drop;
    if (mode == 2 || mode == 3) adversarial_dropped += drop;
    if (drop && any) {
      violations++;
      if (violations <= 5)
        std::fprintf(stderr, "violation: case %ld mode %d count %d low %.9g high %.9g\n", k, mode, count, low, high);
    }
  }
  std::printf("{\"cases\": %ld, \"runs\": %ld, \"dropped\": %ld, \"blocks_in_band\": %ld, \"adversarial\": %ld, "
              "\"adversarial_dropped\": %ld, \"violations\": %ld, \"planted_violations\": %ld}\n",
              cases, runs_total, dropped, in_band_blocks, adversarial, adversarial_dropped, violations, planted_violations);
  return violations ? 1 : 0;
}
