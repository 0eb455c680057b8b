// The occluder pass (src/core/runtime/OcclusionCulling.ts:4-21, 244-273) as a software raster: the camera
// entries of a view slot's draw list are drawn depth-only into level 0 of the slot's HZB buffer, which
// k_hzb_pass then reduces. The reference draws the occluders' previous cuts with the hardware raster
// (three.js, depth-only material); here the same draw list is read where the cut left it on the device.
// One launch sequence draws any number of rasters (view slots, each with its camera and image size): the
// kernels take a table of RasterView records and walk all their entries as one task space.
//
//   k_hzb_header  writes the HZB header (level offsets and sizes) for a width x height level 0
//   k_hzb_clear   level 0 <- 1.0 (WebGPU far depth), every other texel of the buffer <- 0
//   k_raster      kSub drawn meshlets per wave, one per group of 64 / kSub lanes: a meshlet's vertices are
//                 transformed once into LDS (instance matrix, then view_proj; ((c0 x + c1 y) + c2 z) + c3
//                 in binary32 without contraction), snapped to 1/256 pixel; each lane sets up its share of
//                 the meshlet's triangles; triangles whose pixel box holds at most kRasterLanePixels pixel
//                 centres are filled by their lane; larger ones are cut into kChunkSide^2 rectangles and
//                 queued (by the whole lane group when the queue is full)
//   k_raster_chunks  one block per queued rectangle, one pixel per thread: a triangle near the camera
//                 covering much of the image is spread over the whole GPU instead of one wave Depth is kept
//                 with atomicMin on the float bits (all values in [0, 1], so integer order is float order),
//                 which makes the result independent of the order of the writes.
#include "vgcut_kernels.h"

namespace vg {

constexpr uint32_t kRasterLanePixels = 64;
constexpr int32_t kRasterNoVertex = INT32_MIN;
constexpr uint32_t kRasterBlocksPerCu = 8;
constexpr uint32_t kRasterSub = 4;  // meshlets per wave
constexpr int32_t kChunkSide = 64;    // large triangles are filled in rectangles of this side, one block each

// One raster of a launch: a view slot's camera entries drawn into its HZB.
struct RasterView {
  float vp[16];           // column-major clip = vp * world
  const uint4* draws;     // the slot's draw list; camera entries from position 0
  const uint32_t* count;  // entries to draw: min(*count, clamp)
  float* hzb;             // the slot's HZB buffer (header, then level 0 at kHzbHeader: float bits, row 0 at the top)
  uint32_t width, height;
  uint32_t pad[2];
};

// Most rasters of one launch (k_raster keeps their task ends in LDS).
constexpr uint32_t kMaxRasters = 256;

struct RasterParams {
  const float4* col[4];   // instance matrix columns
  const uint32_t* first;  // per meshlet: firstIndex (pool-rebased); its triangles are at first / 3
  const uint32_t* tris;   // per meshlet: triangle count
  const uint2* vlist;     // per meshlet: (offset into vpos, vertex count; 0 = no positions)
  const uint32_t* tri;    // packed local vertex ids a | b << 8 | c << 16
  const float4* vpos;     // per meshlet vertex, in vertex-list order: position (x, y, z, 0)
  const RasterView* views;
  uint32_t rasters;       // RasterView records (<= kMaxRasters)
  uint32_t clamp;
  uint32_t max_size;      // hzb_max_size: every HZB buffer has the same layout
  uint32_t clear_texels;  // texels of an HZB buffer after its header
  struct RasterChunk* chunks;  // large triangles, cut into pixel rectangles of at most kChunkSide^2
  uint32_t* chunk_count;
  uint32_t chunk_capacity;
};

__device__ __forceinline__ uint32_t* raster_depth(const RasterView& v) { return reinterpret_cast<uint32_t*>(v.hzb + kHzbHeader); }

struct RasterVertex {
  int32_t x, y;  // screen position in 1/256 pixel (y down); x == kRasterNoVertex: not drawable
  float z;       // clip z / w
};

// Transforms an object-space position to a snapped screen vertex; false when it cannot be drawn.
__device__ __forceinline__ RasterVertex raster_vertex(const RasterView& r, float4 c0, float4 c1, float4 c2, float4 c3, float4 p) {
  const float wx = ((c0.x * p.x + c1.x * p.y) + c2.x * p.z) + c3.x;
  const float wy = ((c0.y * p.x + c1.y * p.y) + c2.y * p.z) + c3.y;
  const float wz = ((c0.z * p.x + c1.z * p.y) + c2.z * p.z) + c3.z;
  const float* m = r.vp;
  const float cx = ((m[0] * wx + m[4] * wy) + m[8] * wz) + m[12];
  const float cy = ((m[1] * wx + m[5] * wy) + m[9] * wz) + m[13];
  const float cz = ((m[2] * wx + m[6] * wy) + m[10] * wz) + m[14];
  const float cw = ((m[3] * wx + m[7] * wy) + m[11] * wz) + m[15];
  RasterVertex v{kRasterNoVertex, 0, 0.0f};
  if (!(cw > 0.0f) || !(cz >= 0.0f)) return v;
  const float sx = ((cx / cw) * 0.5f + 0.5f) * static_cast<float>(r.width);
  const float sy = (0.5f - (cy / cw) * 0.5f) * static_cast<float>(r.height);
  if (!(fabsf(sx) <= 1048576.0f) || !(fabsf(sy) <= 1048576.0f)) return v;
  v.x = static_cast<int32_t>(rintf(sx * 256.0f));
  v.y = static_cast<int32_t>(rintf(sy * 256.0f));
  v.z = cz / cw;
  return v;
}

// A front-facing triangle in edge order (positive area) and the pixel box of its centres.
struct RasterTriangle {
  RasterVertex a, b, c;
  int64_t area;
  int32_t x0, x1, y0, y1;  // inclusive pixel box, clamped to the image; empty when x0 > x1 or y0 > y1
};

// Screen y points down, so front faces (counter-clockwise in clip space) have negative signed area here:
// they are swapped into positive order; back faces and degenerate triangles give an empty box.
__device__ __forceinline__ RasterTriangle raster_setup(RasterVertex a, RasterVertex b, RasterVertex c, uint32_t W, uint32_t H) {
  RasterTriangle t;
  t.x0 = 1;
  t.x1 = 0;
  t.y0 = 1;
  t.y1 = 0;
  t.area = 0;
  if (a.x == kRasterNoVertex || b.x == kRasterNoVertex || c.x == kRasterNoVertex) return t;
  const int64_t area = static_cast<int64_t>(b.x - a.x) * (c.y - a.y) - static_cast<int64_t>(b.y - a.y) * (c.x - a.x);
  if (area >= 0) return t;
  t.a = a;
  t.b = c;
  t.c = b;
  t.area = -area;
  const int32_t minx = min(a.x, min(b.x, c.x)), maxx = max(a.x, max(b.x, c.x));
  const int32_t miny = min(a.y, min(b.y, c.y)), maxy = max(a.y, max(b.y, c.y));
  // Pixel centres sit at (i + 1/2) * 256: the first centre >= min and the last <= max (arithmetic shifts floor).
  t.x0 = max((minx - 128 + 255) >> 8, 0);
  t.x1 = min((maxx - 128) >> 8, static_cast<int32_t>(W) - 1);
  t.y0 = max((miny - 128 + 255) >> 8, 0);
  t.y1 = min((maxy - 128) >> 8, static_cast<int32_t>(H) - 1);
  return t;
}

// The depth of a covered pixel from its edge functions: the barycentric blend of the vertices' depths,
// kept with atomicMin on the float bits (depths in [0, 1]).
__device__ __forceinline__ void raster_shade(uint32_t* depth, uint32_t W, const RasterTriangle& t, int64_t e0, int64_t e1,
                                             int64_t e2, int32_t x, int32_t y) {
  float z = ((static_cast<float>(e0) * t.a.z + static_cast<float>(e1) * t.b.z) + static_cast<float>(e2) * t.c.z) /
            static_cast<float>(t.area);
  z = fminf(fmaxf(z, 0.0f), 1.0f);
  atomicMin(depth + static_cast<size_t>(y) * W + x, __float_as_uint(z));
}

__device__ __forceinline__ void raster_pixel(uint32_t* depth, uint32_t W, const RasterTriangle& t, int32_t x, int32_t y) {
  const int64_t px = static_cast<int64_t>(x) * 256 + 128, py = static_cast<int64_t>(y) * 256 + 128;
  const RasterVertex& a = t.a;
  const RasterVertex& b = t.b;
  const RasterVertex& c = t.c;
  const int64_t e0 = static_cast<int64_t>(c.x - b.x) * (py - b.y) - static_cast<int64_t>(c.y - b.y) * (px - b.x);
  const int64_t e1 = static_cast<int64_t>(a.x - c.x) * (py - c.y) - static_cast<int64_t>(a.y - c.y) * (px - c.x);
  const int64_t e2 = static_cast<int64_t>(b.x - a.x) * (py - a.y) - static_cast<int64_t>(b.y - a.y) * (px - a.x);
  if ((e0 | e1 | e2) < 0) return;
  raster_shade(depth, W, t, e0, e1, e2, x, y);
}

// Fills the pixel box of t, stepping the edge functions instead of recomputing them: they are affine in the
// pixel position, so e(x + 1, y) = e(x, y) + dx and e(x, y + 1) = e(x, y) + dy, exactly (64-bit integers;
// snapped coordinates are below 2^29 in magnitude).
__device__ __forceinline__ void raster_box(uint32_t* depth, uint32_t W, const RasterTriangle& t) {
  if (t.x0 > t.x1 || t.y0 > t.y1) return;
  const RasterVertex& a = t.a;
  const RasterVertex& b = t.b;
  const RasterVertex& c = t.c;
  const int64_t px = static_cast<int64_t>(t.x0) * 256 + 128, py = static_cast<int64_t>(t.y0) * 256 + 128;
  int64_t r0 = static_cast<int64_t>(c.x - b.x) * (py - b.y) - static_cast<int64_t>(c.y - b.y) * (px - b.x);
  int64_t r1 = static_cast<int64_t>(a.x - c.x) * (py - c.y) - static_cast<int64_t>(a.y - c.y) * (px - c.x);
  int64_t r2 = static_cast<int64_t>(b.x - a.x) * (py - a.y) - static_cast<int64_t>(b.y - a.y) * (px - a.x);
  const int64_t dx0 = -static_cast<int64_t>(c.y - b.y) * 256, dy0 = static_cast<int64_t>(c.x - b.x) * 256;
  const int64_t dx1 = -static_cast<int64_t>(a.y - c.y) * 256, dy1 = static_cast<int64_t>(a.x - c.x) * 256;
  const int64_t dx2 = -static_cast<int64_t>(b.y - a.y) * 256, dy2 = static_cast<int64_t>(b.x - a.x) * 256;
  for (int32_t y = t.y0; y <= t.y1; y++, r0 += dy0, r1 += dy1, r2 += dy2) {
    int64_t e0 = r0, e1 = r1, e2 = r2;
    for (int32_t x = t.x0; x <= t.x1; x++, e0 += dx0, e1 += dx1, e2 += dx2)
      if ((e0 | e1 | e2) >= 0) raster_shade(depth, W, t, e0, e1, e2, x, y);
  }
}

// A rectangle of a large triangle's pixel box: filled by one block of k_raster_chunks.
struct RasterChunk {
  RasterVertex a, b, c;
  int32_t area_lo, area_hi;
  int32_t x0, x1, y0, y1;
  uint32_t raster;  // its RasterView
};

__device__ __forceinline__ uint32_t raster_box_pixels(const RasterTriangle& t) {
  if (t.x0 > t.x1 || t.y0 > t.y1) return 0;
  return static_cast<uint32_t>(t.x1 - t.x0 + 1) * static_cast<uint32_t>(t.y1 - t.y0 + 1);
}

// One block per raster.
__global__ __launch_bounds__(64) void k_hzb_header(RasterParams r) {
  const RasterView& v = r.views[blockIdx.x];
  float* hzb = v.hzb;
  const uint32_t max_size = r.max_size, width = v.width, height = v.height;
  for (uint32_t i = threadIdx.x; i < kHzbHeader; i += blockDim.x) hzb[i] = 0.0f;
  __syncthreads();
  if (threadIdx.x != 0) return;
  uint32_t offset = kHzbHeader, dim = max_size, lw = width, lh = height, levels = 0;
  for (uint32_t i = 0; i < kHzbMaxLevels; i++) {
    hzb[i * 4] = static_cast<float>(offset);
    hzb[i * 4 + 1] = static_cast<float>(lw);
    hzb[i * 4 + 2] = static_cast<float>(lh);
    hzb[i * 4 + 3] = 0.0f;
    levels++;
    if ((lw == 1 && lh == 1) || dim == 1) break;
    offset += dim * dim;
    dim = (dim + 1) / 2;
    lw = lw > 1 ? (lw + 1) / 2 : 1;
    lh = lh > 1 ? (lh + 1) / 2 : 1;
  }
  hzb[kHzbLevelCountSlot] = static_cast<float>(levels);
}

// Clears each raster's HZB texel area: the first width x height texels (the depth image about to be drawn)
// to 1.0 and the rest of its clear_texels (every level's area, used or not) to 0, so the buffer holds no
// earlier contents whatever the image size. blockIdx.y is the raster. Also resets the chunk queue.
__global__ __launch_bounds__(256) void k_hzb_clear(RasterParams r) {
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *r.chunk_count = 0;
  const RasterView& v = r.views[blockIdx.y];
  uint32_t* texels = raster_depth(v);
  const uint32_t far = v.width * v.height, n = r.clear_texels;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) texels[i] = i < far ? 0x3f800000u : 0u;
}

// kSub meshlets per wave, each on 64 / kSub consecutive lanes (a lane group): every load of a meshlet (its
// positions, contiguous in vertex-list order, and packed triangles) is issued before any is used, and the
// groups' meshlets are in flight together, so the dependent chain entry -> meshlet record -> positions costs
// one wait per kSub meshlets.
// A wave-task is kSub consecutive entries of one raster (so the raster's camera and image are wave-uniform):
// the tasks of all rasters form one space, walked by the persistent waves; task_end[k] (in LDS) is the
// end of raster k's tasks.
template <uint32_t kSub>
__global__ __launch_bounds__(256) void k_raster(RasterParams r) {
  constexpr uint32_t kLanes = 64 / kSub, kPer = 128 / kLanes;
  __shared__ RasterVertex lds[4][kSub][128];
  __shared__ uint32_t task_end[kMaxRasters];
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6, group = lane / kLanes, gl = lane % kLanes;
  RasterVertex* V = lds[wave][group];
  const uint64_t group_mask = (kSub == 1 ? ~0ull : ((1ull << kLanes) - 1ull)) << (group * kLanes);
  if (wave == 0) {  // the rasters' task ends: a prefix over ceil(min(count, clamp) / kSub)
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < r.rasters; b0 += 64) {
      const uint32_t k = b0 + lane;
      const uint32_t n = k < r.rasters ? min(*r.views[k].count, r.clamp) : 0u;
      const uint32_t incl = carry + wave_scan((n + kSub - 1) / kSub);
      if (k < r.rasters) task_end[k] = incl;
      carry = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 63));
    }
  }
  __syncthreads();
  const uint32_t total = task_end[r.rasters - 1];
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  auto wave_sync = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  uint32_t rk = 0;  // the raster of the current task (tasks only increase, so the search moves forward)
  for (uint32_t t = xcd_wave_rank(); t < total; t += waves) {
    while (task_end[rk] <= t) rk++;  // rk < rasters: t < total = task_end[rasters - 1]
    const uint32_t k = uniform(rk);
    const RasterView& rv = r.views[k];
    const uint32_t n = min(*rv.count, r.clamp);
    const uint32_t e = (t - (k ? task_end[k - 1] : 0u)) * kSub + group;
    uint32_t* depth = raster_depth(rv);
    const uint32_t width = rv.width, height = rv.height;
    uint32_t nv = 0, tris = 0, first = 0, voff = 0, instance = 0;
    if (e < n) {
      const uint4 entry = rv.draws[e];
      instance = entry.x;
      const uint2 vl = r.vlist[entry.y];
      voff = vl.x;
      nv = vl.y;
      tris = r.tris[entry.y];
      first = r.first[entry.y] / 3;
    }
    if (nv == 0) tris = 0;
    // The meshlet's positions are contiguous (vertex-list order): the lane group reads them coalesced, in the
    // same round trip as the packed triangles.
    uint32_t packed[kPer];
    float4 pos[kPer];
#pragma unroll
    for (uint32_t j = 0; j < kPer; j++) {
      const uint32_t k = gl + j * kLanes;
      pos[j] = k < nv ? r.vpos[voff + k] : make_float4(0, 0, 0, 0);
      packed[j] = k < tris ? r.tri[first + k] : 0u;
    }
    if (nv) {
      const float4 c0 = r.col[0][instance], c1 = r.col[1][instance], c2 = r.col[2][instance], c3 = r.col[3][instance];
#pragma unroll
      for (uint32_t j = 0; j < kPer; j++)
        if (gl + j * kLanes < nv) V[gl + j * kLanes] = raster_vertex(rv, c0, c1, c2, c3, pos[j]);
    }
    wave_sync();
    bool big[kPer];
#pragma unroll
    for (uint32_t j = 0; j < kPer; j++) {
      big[j] = false;
      if (gl + j * kLanes >= tris) continue;
      const uint32_t p = packed[j];
      const RasterTriangle tr = raster_setup(V[p & 255u], V[(p >> 8) & 255u], V[(p >> 16) & 255u], width, height);
      const uint32_t pixels = raster_box_pixels(tr);
      if (pixels > kRasterLanePixels) {
        // Large: its kChunkSide^2 rectangles go to the chunk queue (filled by every block of k_raster_chunks);
        // when the queue is full, the lane group fills it below.
        const uint32_t cw = static_cast<uint32_t>(tr.x1 - tr.x0) / kChunkSide + 1, ch = static_cast<uint32_t>(tr.y1 - tr.y0) / kChunkSide + 1;
        const uint32_t at = atomicAdd(r.chunk_count, cw * ch);
        if (at >= r.chunk_capacity || cw * ch > r.chunk_capacity - at) {
          // The reservation runs past the queue: the slots it holds inside the queue (read by
          // k_raster_chunks, which takes min(count, capacity)) get empty rectangles.
          for (uint32_t q = at; q < r.chunk_capacity && q - at < cw * ch; q++) {
            RasterChunk c{};
            c.x0 = c.y0 = 1;
            c.x1 = c.y1 = 0;
            r.chunks[q] = c;
          }
          big[j] = true;
          continue;
        }
        for (uint32_t q = 0; q < cw * ch; q++) {
          RasterChunk c;
          c.a = tr.a;
          c.b = tr.b;
          c.c = tr.c;
          c.area_lo = static_cast<int32_t>(static_cast<uint64_t>(tr.area));
          c.area_hi = static_cast<int32_t>(static_cast<uint64_t>(tr.area) >> 32);
          c.x0 = tr.x0 + static_cast<int32_t>(q % cw) * kChunkSide;
          c.y0 = tr.y0 + static_cast<int32_t>(q / cw) * kChunkSide;
          c.x1 = min(c.x0 + kChunkSide - 1, tr.x1);
          c.y1 = min(c.y0 + kChunkSide - 1, tr.y1);
          c.raster = k;
          r.chunks[at + q] = c;
        }
        continue;
      }
      raster_box(depth, width, tr);
    }
    // Large triangles the queue had no room for: each lane group strides over its meshlet's, one at a time.
#pragma unroll
    for (uint32_t j = 0; j < kPer; j++) {
      uint64_t mask = __ballot(big[j]) & group_mask;
      while (mask) {
        const uint32_t src = static_cast<uint32_t>(__ffsll(static_cast<long long>(mask)) - 1);
        mask &= mask - 1;
        const uint32_t p = static_cast<uint32_t>(__shfl(static_cast<int>(packed[j]), static_cast<int>(src)));
        const RasterTriangle tr = raster_setup(V[p & 255u], V[(p >> 8) & 255u], V[(p >> 16) & 255u], width, height);
        const uint32_t bw = static_cast<uint32_t>(tr.x1 - tr.x0 + 1), pixels = raster_box_pixels(tr);
        for (uint32_t i = gl; i < pixels; i += kLanes)
          raster_pixel(depth, width, tr, tr.x0 + static_cast<int32_t>(i % bw), tr.y0 + static_cast<int32_t>(i / bw));
      }
    }
    wave_sync();
  }
}

// The queued rectangles of large triangles, one block each: 16 x 16 pixel tiles, one pixel per thread.
__global__ __launch_bounds__(256) void k_raster_chunks(RasterParams r) {
  const uint32_t n = min(*r.chunk_count, r.chunk_capacity);
  const int32_t tx = static_cast<int32_t>(threadIdx.x & 15u), ty = static_cast<int32_t>(threadIdx.x >> 4);
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
    const RasterChunk c = r.chunks[i];
    RasterTriangle t;
    t.a = c.a;
    t.b = c.b;
    t.c = c.c;
    t.area = static_cast<int64_t>((static_cast<uint64_t>(static_cast<uint32_t>(c.area_hi)) << 32) | static_cast<uint32_t>(c.area_lo));
    const RasterView& rv = r.views[c.raster];
    uint32_t* depth = raster_depth(rv);
    const uint32_t width = rv.width;
    for (int32_t y0 = c.y0; y0 <= c.y1; y0 += 16)
      for (int32_t x0 = c.x0; x0 <= c.x1; x0 += 16) {
        const int32_t x = x0 + tx, y = y0 + ty;
        if (x <= c.x1 && y <= c.y1) raster_pixel(depth, width, t, x, y);
      }
  }
}

// k_hzb_pass over the launch's rasters: blockIdx.z is the raster, the grid is sized for the largest level.
__global__ __launch_bounds__(256) void k_hzb_pass_rasters(RasterParams r, uint32_t base) { hzb_pass(r.views[blockIdx.z].hzb, base); }

}  // namespace vg
