/*
 * Synthetic scene producer for the cut: seeded procedural meshes, their meshlet LOD DAGs in the
 * reference's `VirtualMeshData` layout (src/core/preprocess/buildVirtualMesh.ts:75-138), instance
 * placements and synthetic occluder depth.
 *
 * The reference builds its DAGs with meshoptimizer (buildVirtualMesh.ts:185-297), which is not
 * available here. This builder makes DAGs with the same invariants from parametric grid patches:
 * 128-triangle leaf meshlets, groups simplified to half their triangles per level, monotone errors
 * (error = max child error + max(shape error, 1e-7), buildVirtualMesh.ts:262-270), LOD spheres shared
 * by a group's outputs, enclosing lod/cull bound spheres and per-level guard errors (pack,
 * buildVirtualMesh.ts:484-601).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { VGS_SPHERE = 0, VGS_BOX = 1, VGS_COLUMN = 2, VGS_ARCH = 3, VGS_TERRAIN = 4 };

typedef struct vgs_mesh_params {
  int32_t kind;            /* VGS_* */
  uint32_t seed;           /* noise seed */
  double size[3];          /* sphere: radius; box: half extents; column: radius, height; arch: span, height, thickness */
  double noise_amplitude;  /* displacement amplitude as a fraction of the mesh's characteristic size */
  double noise_frequency;  /* base frequency of the 3-octave value noise, per unit length */
  uint32_t resolution;     /* sphere: frequency per diamond edge (multiple of 16); others: quads per unit (scaled) */
  uint32_t target_triangles; /* 0: use resolution; otherwise pick resolutions near this leaf count */
} vgs_mesh_params;

/* A built mesh: VirtualMeshData fields the cut reads, plus the DAG links the coverage checks walk. */
typedef struct vgs_mesh_view {
  uint32_t meshlet_count;
  const float* meshlet_bounds;      /* 16 per meshlet */
  const uint32_t* meshlet_info;     /* 4 per meshlet */
  uint32_t level_count;
  const uint32_t* level_ranges;     /* 2 per level */
  const float* level_errors;        /* 2 per level */
  const float* level_guard_errors;  /* 2 per level */
  const int32_t* replacement_group; /* per meshlet */
  uint32_t group_count;
  const uint32_t* replacement_start;   /* group_count + 1 */
  const uint32_t* replacement_indices;
  float bounding_sphere[4];
  float lod_bounds_sphere[4];
  float cull_bounds_sphere[4];
  uint64_t leaf_triangles;
  uint32_t leaf_meshlets;
  uint32_t root_triangles;
  uint32_t root_meshlets;
  const uint32_t* meshlet_triangles; /* 3 local vertex ids per triangle, packed a | b << 8 | c << 16 */
  uint64_t triangle_count;
} vgs_mesh_view;

typedef struct vgs_mesh vgs_mesh;

int vgs_mesh_build(const vgs_mesh_params* params, vgs_mesh** out);
void vgs_mesh_get(const vgs_mesh* mesh, vgs_mesh_view* view);
void vgs_mesh_free(vgs_mesh* mesh);

/*
 * The same procedural mesh as one watertight indexed triangle mesh, the input of the DAG builder
 * (vgdag.h): each patch is tessellated at its leaf resolution without displacement, coincident points
 * are welded, windings are made consistent and outward, and every vertex is displaced along its
 * area-weighted normal by amplitude * size * fbm(position * frequency). Column caps are fans over the
 * side's rims; arch feet take the tube's side resolution, so no seam leaves a crack.
 */
typedef struct vgs_triangles {
  uint32_t vertex_count;
  float* positions;  /* 3 per vertex */
  uint64_t index_count;
  uint32_t* indices; /* 3 per triangle */
} vgs_triangles;

int vgs_mesh_triangles(const vgs_mesh_params* params, vgs_triangles* out);
void vgs_triangles_free(vgs_triangles* triangles);

/* Bridson Poisson-disk samples over [x0, x0+w] x [z0, z0+h]; returns the count written (<= max_points). */
uint32_t vgs_poisson_disk(double x0, double z0, double w, double h, double radius, uint32_t seed, uint32_t candidates,
                          uint32_t max_points, float* out_xz);

/* 3-octave value-noise heightfield: y = amplitude * fbm(x * frequency, z * frequency). */
void vgs_heightfield(const float* xz, uint32_t n, double amplitude, double frequency, uint32_t seed, float* out_y);

/*
 * Depth of a noise heightfield as seen by a camera: the heightfield is rasterized as a grid mesh over
 * [cx - extent, cx + extent]^2 with `grid` quads per side, WebGPU depth ([0, 1], 1 = far, cleared to 1),
 * row 0 at the top. `view_proj` is column-major proj * view.
 */
void vgs_heightfield_depth(const float* view_proj, uint32_t width, uint32_t height, double cx, double cz,
                           double extent, uint32_t grid, double amplitude, double frequency, uint32_t seed,
                           float* out_depth);

#ifdef __cplusplus
}
#endif
