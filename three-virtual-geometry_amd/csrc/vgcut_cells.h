// Cells of 128 instances, built on the host when a mesh's instances are set: the reference's cell sphere
// (GeometryPool.ts:519-545) and the bounds behind the kernels' whole-cell frustum answers
// (vg::cell_inside_bits). Shared with the host-side check of those answers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace vgi {

// instanceSphere, GeometryPool.ts:1171-1180 (fp64; column lengths as sqrt of the sum of squares).
inline void instance_sphere(const float* m, const float* sphere, double out[4]) {
  const double sx = sphere[0], sy = sphere[1], sz = sphere[2], sr = sphere[3];
  out[0] = static_cast<double>(m[0]) * sx + static_cast<double>(m[4]) * sy + static_cast<double>(m[8]) * sz + m[12];
  out[1] = static_cast<double>(m[1]) * sx + static_cast<double>(m[5]) * sy + static_cast<double>(m[9]) * sz + m[13];
  out[2] = static_cast<double>(m[2]) * sx + static_cast<double>(m[6]) * sy + static_cast<double>(m[10]) * sz + m[14];
  auto col = [&](int k) {
    const double x = m[k * 4], y = m[k * 4 + 1], z = m[k * 4 + 2];
    return std::sqrt(x * x + y * y + z * z);
  };
  out[3] = sr * std::max(col(0), std::max(col(1), col(2)));
}

// computeCell, GeometryPool.ts:519-545 (fp64, stored as f32).
inline void compute_cell(const std::vector<float>& gpu_matrices, uint32_t first, uint32_t end, const float* sphere, uint32_t slot,
                  uint32_t instance_base, float4& out_sphere, float4& out_rec) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY}, s[4];
  double max_radius = 0;
  for (uint32_t g = first; g < end; g++) {
    instance_sphere(&gpu_matrices[static_cast<size_t>(g) * 16], sphere, s);
    for (int a = 0; a < 3; a++) {
      lo[a] = std::min(lo[a], s[a] - s[3]);
      hi[a] = std::max(hi[a], s[a] + s[3]);
    }
    max_radius = std::max(max_radius, s[3]);
  }
  const double cx = (lo[0] + hi[0]) / 2, cy = (lo[1] + hi[1]) / 2, cz = (lo[2] + hi[2]) / 2;
  double radius = 0;
  for (uint32_t g = first; g < end; g++) {
    instance_sphere(&gpu_matrices[static_cast<size_t>(g) * 16], sphere, s);
    const double dx = s[0] - cx, dy = s[1] - cy, dz = s[2] - cz;
    radius = std::max(radius, std::sqrt(dx * dx + dy * dy + dz * dz) + s[3]);
  }
  out_sphere = make_float4(static_cast<float>(cx), static_cast<float>(cy), static_cast<float>(cz),
                           static_cast<float>(radius * (1 + 1e-4) + 1e-3));
  out_rec = make_float4(static_cast<float>(max_radius * (1 + 1e-4) + 1e-4), static_cast<float>(slot),
                        static_cast<float>(instance_base + first), static_cast<float>(end - first));
}

// Bounds of a cell's instances for the kernels' whole-cell frustum answers (vg::cell_inside_bits), in binary64
// from the binary32 matrices, around the cell centre as stored (binary32): x = the largest |c0_k x| + |c1_k y|
// + |c2_k z| + |c3_k| over instances, components k and both sphere centres; y = R_in >= |c_i - C| (bounding
// sphere centres: the camera test only asks that a sphere reach into the frustum), z = R_cull >= |cc_i - C| +
// cr_i (culling spheres, which the fast path needs wholly inside), each rounded up with room to spare.
// A negative sphere radius, a non-affine or non-finite matrix gives bounds no test passes (+inf).
inline float4 cell_bound(const std::vector<float>& gpu_matrices, uint32_t first, uint32_t end, const float* sphere, const float* cull,
                  const float4& center) {
  const double C[3] = {center.x, center.y, center.z};
  double A = 0, r_in = 0, r_cull = 0;
  bool ok = sphere[3] >= 0 && cull[3] >= 0;
  for (uint32_t g = first; ok && g < end; g++) {
    const float* m = &gpu_matrices[static_cast<size_t>(g) * 16];
    ok = m[3] == 0.0f && m[7] == 0.0f && m[11] == 0.0f && m[15] == 1.0f;
    for (int k = 0; k < 16; k++) ok = ok && std::isfinite(m[k]);
    double s[4], t[4];
    instance_sphere(m, sphere, s);
    instance_sphere(m, cull, t);
    for (int k = 0; k < 3; k++)
      for (const float* p : {sphere, cull})
        A = std::max(A, std::fabs(static_cast<double>(m[k]) * p[0]) + std::fabs(static_cast<double>(m[4 + k]) * p[1]) +
                            std::fabs(static_cast<double>(m[8 + k]) * p[2]) + std::fabs(static_cast<double>(m[12 + k])));
    r_in = std::max(r_in, std::sqrt((s[0] - C[0]) * (s[0] - C[0]) + (s[1] - C[1]) * (s[1] - C[1]) + (s[2] - C[2]) * (s[2] - C[2])));
    r_cull = std::max(r_cull, std::sqrt((t[0] - C[0]) * (t[0] - C[0]) + (t[1] - C[1]) * (t[1] - C[1]) + (t[2] - C[2]) * (t[2] - C[2])) + t[3]);
  }
  if (!ok || !std::isfinite(A) || !std::isfinite(r_in) || !std::isfinite(r_cull)) return make_float4(INFINITY, INFINITY, INFINITY, 0);
  auto up = [](double x) { return static_cast<float>(x * (1 + 1e-6) + 1e-30); };
  return make_float4(up(A), up(r_in), up(r_cull), 0.0f);
}

}  // namespace vgi
