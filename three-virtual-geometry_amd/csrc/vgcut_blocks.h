// Bounds of an LOD block (up to kBlockRuns consecutive LOD runs of one level), built on the host when a
// mesh is added; the meshlet pass's block test (block_out_of_band) reads them. Shared with the host-side
// check of that test.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

namespace vgi {

// The binary32 value nearest x rounded toward +inf / -inf (for bounds that must hold in real arithmetic).
inline float f32_up(double x) {
  const float f = static_cast<float>(x);
  return static_cast<double>(f) < x ? std::nextafter(f, HUGE_VALF) : f;
}
inline float f32_down(double x) {
  const float f = static_cast<float>(x);
  return static_cast<double>(f) > x ? std::nextafter(f, -HUGE_VALF) : f;
}

// A sphere holding the centres of spheres [first, first + count): its centre the middle of their box
// (rounded to binary32), its radius the largest distance from that centre (fp64, rounded up).
inline float4 centre_bound(const float4* spheres, uint32_t first, uint32_t count) {
  double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (uint32_t r = first; r < first + count; r++) {
    const double v[3] = {spheres[r].x, spheres[r].y, spheres[r].z};
    for (int k = 0; k < 3; k++) {
      lo[k] = std::min(lo[k], v[k]);
      hi[k] = std::max(hi[k], v[k]);
    }
  }
  float4 out = make_float4(static_cast<float>(0.5 * (lo[0] + hi[0])), static_cast<float>(0.5 * (lo[1] + hi[1])),
                           static_cast<float>(0.5 * (lo[2] + hi[2])), 0.0f);
  double rho = 0;
  for (uint32_t r = first; r < first + count; r++) {
    const double dx = static_cast<double>(spheres[r].x) - out.x, dy = static_cast<double>(spheres[r].y) - out.y,
                 dz = static_cast<double>(spheres[r].z) - out.z;
    rho = std::max(rho, std::sqrt(dx * dx + dy * dy + dz * dz));
  }
  out.w = f32_up(rho);
  return out;
}

// The block bounds of runs [first, first + count) (their LOD spheres, parent spheres and errors):
// lod / par = centre_bound of the LOD / parent sphere centres; err = (smallest own error, smallest LOD
// radius, largest parent error or +inf when one is >= 5e29 (projected to 1e30 whatever the distance),
// largest parent radius), each rounded outward.
inline void block_bounds(const float4* r_lod, const float4* r_parent, const float2* r_err, uint32_t first, uint32_t count, float4* lod,
                         float4* par, float4* err) {
  const float inf_half = static_cast<float>(1e30 * 0.5);
  double e_min = HUGE_VAL, r_min = HUGE_VAL, pe_max = 0, rp_max = 0;
  for (uint32_t r = first; r < first + count; r++) {
    e_min = std::min(e_min, static_cast<double>(r_err[r].x));
    r_min = std::min(r_min, static_cast<double>(r_lod[r].w));
    pe_max = r_err[r].y >= inf_half ? HUGE_VAL : std::max(pe_max, static_cast<double>(r_err[r].y));
    rp_max = std::max(rp_max, static_cast<double>(r_parent[r].w));
  }
  *lod = centre_bound(r_lod, first, count);
  *par = centre_bound(r_parent, first, count);
  *err = make_float4(f32_down(e_min), f32_down(r_min), std::isinf(pe_max) ? HUGE_VALF : f32_up(pe_max), f32_up(rp_max));
}

}  // namespace vgi
