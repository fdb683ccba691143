// j3dg_internal.h -- context, device records and per-frame parameters of the
// gfx950 ray-casting library (private to j3d_amd/csrc).
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "j3dg.h"
#include "rt_math.h"

namespace j3d {

// Device record of one scene object (scene_object, j3d/scene.h:10-25).
struct DevObject {
  const float4* nodes;     // wide BVH nodes
  const float4* tris;      // triangle records in leaf order
  const float* verts;      // original vertices (shadow-ray origins)
  const uint32_t* idx;     // original triangles
  const float* normals;    // per-triangle normals
  const float4* fin;       // per triangle, by id: (v0, n.x), (v1, n.y), (v2, n.z)
  float cs[16];            // object -> world
  float inv_cs[16];        // world -> object
  uint32_t db_id;
  uint32_t width;
  uint32_t tex_w, tex_h;   // texture size (with `tex`)
  const float* vcol;       // optional rgb per vertex (nullptr = none)
  const float* uv;         // optional (u, v) per triangle corner, 6 floats per triangle
  const uint32_t* tex;     // texture for `uv`, row-major 0xAABBGGRR
  const void* mlp;         // learned proposal: packed network (nullptr = none)
  float mlp_norm[4];       // its input normalisation: box centre (object space), half-diagonal
};

// Frame parameters of the trace kernel (passed by value).
struct FrameParams {
  uint32_t w, h;
  int32_t x0, y0, x1, y1;   // clamped, inclusive
  uint32_t tiles_x, n_tiles;
  float proj_inv[16];
  float near_plane;
  float cs[16], cs_inv[16];
  float diagonal;
  float light[4];           // point light, world (j3d/canvas.cpp:703-705)
  float light_dir[3];
  uint32_t flags;
  uint32_t n_obj;
  uint32_t fused;           // primary + hit completion + shadow ray in one launch
  uint32_t obj_base;        // batches: index of this frame's first object in the batch's object table
  const uint32_t* tile_order;  // nullable: ray id / 64 -> tile (heaviest tiles of the last frame first)
  uint16_t* pixel_cost;        // nullable: loop trips of each pixel, [tile * 64 + lane], for the next order
  j3dg_pixel* gbuf;
  const float2* proposal;   // learned: per-pixel (t bound, hit logit), indexed y * w + x
  float margin;
  const void* top_nodes;    // kModeTop: the objects' top-level BVH (TopNode[], root first)
  float top_delta;          // kModeTop: world-box widening for the top level's box tests
  uint32_t top_count;       // kModeTop: nodes of the top level
  uint32_t top_lds;         // kModeTop: blocks copy the top level into LDS after the stack
};
// Largest top level a block copies into LDS (256 nodes of 64 bytes).
constexpr uint32_t kTopLdsBytes = 16384;

// kModeFrameColor: kModeFrame for frames with surface colours (the hit's
// object index goes to the db_id word for k_color).
enum { kModeFrame = 0, kModeRays = 1, kModeAnyRays = 2, kModeShadow = 3, kModeBatch = 4, kModeFrameColor = 5 };
// Added to kModeFrame, kModeFrameColor, kModeRays or kModeAnyRays: scenes of
// more than J3DG_MAX_OBJECTS objects, with a top-level BVH over the objects
// (FrameParams::top_nodes) and the object table read from global memory.
constexpr int kModeTop = 16;
// Largest object count of one scene with a top level.
constexpr uint32_t kMaxSceneObjects = 1u << 16;

// Many frames in one launch (j3dg_render_batch): per-frame parameters on the
// device and each frame's first tile in the batch's global tile numbering.
struct BatchParams {
  const FrameParams* frames;   // n_frames
  const uint32_t* first_tile;  // n_frames + 1
  uint32_t n_frames;
};

// Ray-buffer queries (j3dg_trace_rays / j3dg_any_hit_rays).
struct RayQuery {
  const float* rays;  // n x 8
  uint32_t* obj;
  uint32_t* tri;
  float* t;
  float* u;
  float* v;
  uint8_t* any;
  const uint32_t* count;  // shadow queue: number of rays (read on the device)
};

struct ShadeParams {
  uint32_t w, h;
  const j3dg_pixel* gbuf;
  uint32_t* rgba;
  const uint32_t* matcap;
  uint32_t mw, mh, cavity;
  uint32_t flags;
  float proj_inv[16];
  float near_plane;
  uint32_t bg_top, bg_bottom;
};

inline void invert_orthonormal(const float* m, float* o) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o[4 * j + i] = m[4 * i + j];
  o[3] = o[7] = o[11] = 0.f;
  for (int i = 0; i < 3; ++i) o[12 + i] = -(m[4 * i] * m[12] + m[4 * i + 1] * m[13] + m[4 * i + 2] * m[14]);
  o[15] = 1.f;
}

struct Mesh {
  bool live = false;
  uint32_t db_id = 0;
  float cs[16], inv_cs[16];
  j3dg_mesh_info info{};
  uint32_t max_stack = 1;
  float4* d_nodes = nullptr;
  float4* d_tris = nullptr;
  float* d_verts = nullptr;
  uint32_t* d_idx = nullptr;
  float* d_normals = nullptr;
  float4* d_fin = nullptr;  // hit completion records (see DevObject::fin)
  // surface colour (j3d/scene.h:16-18): per-vertex rgb, per-corner uv + texture
  float* d_vcol = nullptr;
  float* d_uv = nullptr;
  uint32_t* d_tex = nullptr;
  uint32_t tex_w = 0, tex_h = 0;
  std::vector<float> normals_host;
  // learned proposal (see mlp.hip): fp32 parameters, packed bf16 fragments
  float* d_mlp = nullptr;
  void* d_mlp_packed = nullptr;
  float mlp_norm[4] = {0.f, 0.f, 0.f, 1.f};
  bool has_proposal = false;
  void release() {
    if (d_nodes) hipFree(d_nodes);
    if (d_tris) hipFree(d_tris);
    if (d_verts) hipFree(d_verts);
    if (d_idx) hipFree(d_idx);
    if (d_normals) hipFree(d_normals);
    if (d_fin) hipFree(d_fin);
    if (d_vcol) hipFree(d_vcol);
    if (d_uv) hipFree(d_uv);
    if (d_tex) hipFree(d_tex);
    d_vcol = nullptr, d_uv = nullptr, d_tex = nullptr;
    tex_w = tex_h = 0;
    if (d_mlp) hipFree(d_mlp);
    if (d_mlp_packed) hipFree(d_mlp_packed);
    d_nodes = nullptr, d_tris = nullptr, d_verts = nullptr, d_idx = nullptr, d_normals = nullptr, d_mlp = nullptr;
    d_fin = nullptr;
    d_mlp_packed = nullptr;
    normals_host.clear();
    has_proposal = false;
    live = false;
  }
};

struct Target {
  bool live = false;
  uint32_t w = 0, h = 0;
  j3dg_pixel* gbuf = nullptr;
  uint32_t* rgba = nullptr;
  float2* proposal = nullptr;  // learned: per-pixel (t bound, hit logit), allocated on first use
  // cost-ordered tiles: per-pixel trips of the last frame, per-tile sums, and the
  // order (heaviest first) valid for frames with the same rectangle
  uint16_t* pixel_cost = nullptr;
  uint32_t* tile_cost = nullptr;
  uint32_t* tile_order = nullptr;
  uint32_t order_capacity = 0;  // tiles
  bool order_valid = false;
  int32_t order_rect[4] = {0, 0, -1, -1};
  void release() {
    if (gbuf) hipFree(gbuf);
    if (rgba) hipFree(rgba);
    if (proposal) hipFree(proposal);
    if (pixel_cost) hipFree(pixel_cost);
    if (tile_cost) hipFree(tile_cost);
    if (tile_order) hipFree(tile_order);
    gbuf = nullptr, rgba = nullptr, proposal = nullptr, pixel_cost = nullptr, tile_cost = nullptr, tile_order = nullptr;
    order_capacity = 0;
    order_valid = false;
    live = false;
  }
};

// A point cloud (scene_pointcloud, j3d/scene.h:27-38) on the device: points,
// optional normals and colours (0xAABBGGRR), and its pose.
struct PointCloud {
  bool live = false;
  uint32_t db_id = 0;
  uint64_t n = 0;
  float cs[16];
  float* d_xyz = nullptr;
  float* d_normals = nullptr;
  uint32_t* d_colors = nullptr;
  void release() {
    if (d_xyz) hipFree(d_xyz);
    if (d_normals) hipFree(d_normals);
    if (d_colors) hipFree(d_colors);
    d_xyz = nullptr, d_normals = nullptr, d_colors = nullptr;
    n = 0;
    live = false;
  }
};

struct Ctx {
  std::mutex mu;
  int device = 0;
  int n_cus = 1;
  std::string arch;
  std::string last_error;
  hipStream_t stream = nullptr;
  hipEvent_t ev[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  // per-render HIP events (primary trace, shade, mlp, finish + shadow trace: start/end),
  // kept for every render since the last j3dg_reset_timings
  static constexpr uint32_t kRing = 1024;
  std::vector<hipEvent_t> ring;    // kRing * 8
  uint32_t ring_n = 0;             // renders recorded since reset
  uint32_t ring_flags[kRing];      // bit 0 trace, bit 1 shade, bit 2 mlp
  const hipEvent_t* last_ev = nullptr;
  hipEvent_t* cur_ev() { return ring.empty() ? ev : &ring[8 * (ring_n % kRing)]; }
  bool traced = false, shaded = false, mlp_ran = false;
  hipEvent_t splat_ev[2] = {nullptr, nullptr};  // around each k_splat_depth of j3dg_splat_points
  float splat_ms = 0.f;                          // their sum over the last call's clouds
  bool counting = false;
  j3dg_tuning tuning{16, 1, 4, 256, 16, 0, 1};  // depth, rays per item, coop tail, block, refill, split, order
  std::vector<Mesh> meshes;
  std::vector<Target> targets;
  std::vector<PointCloud> clouds;
  uint32_t* d_queue = nullptr;
  uint64_t* d_counters = nullptr;
  void* noalloc_scratch = nullptr;  // this device's j3dg_trace_rays_device working memory (not owned)
  uint2* spill = nullptr;
  uint64_t spill_entries = 0;
  DevObject* d_objs = nullptr;   // the current object table (not owned: a ring slot or d_obj_big)
  // tables of up to J3DG_MAX_OBJECTS objects go through a ring of device
  // slots filled asynchronously from pinned staging, so that a frame whose
  // objects differ from the last one's does not wait for the GPU; a slot is
  // rewritten only after the event recorded when it was retired has passed
  static constexpr int kObjRing = 8;
  DevObject* d_obj_ring = nullptr;  // kObjRing x J3DG_MAX_OBJECTS
  DevObject* h_obj_ring = nullptr;  // pinned, same shape
  hipEvent_t obj_ring_ev[kObjRing] = {};
  bool obj_ring_used[kObjRing] = {};
  int obj_ring_cur = -1;
  DevObject* d_obj_big = nullptr;   // larger tables (top-level scenes), written synchronously
  uint32_t obj_capacity = 0;        // entries of d_obj_big
  std::vector<uint8_t> obj_cache;
  // top-level BVH of the last object table of more than J3DG_MAX_OBJECTS
  // objects (TopNode[]), keyed by the objects' boxes and poses
  void* d_top = nullptr;
  uint32_t top_capacity = 0;     // TopNode slots allocated
  uint32_t top_depth = 0;        // internal levels: stack entries the top level may hold
  uint32_t top_count = 0;        // nodes of the top level
  float top_radius = 0.f;        // largest |coordinate| of the world boxes, translations and local boxes
  std::vector<float> top_key;
  // device buffers of the all-hits / voxel calls, kept between calls (grown on demand)
  void* scratch[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t scratch_bytes[4] = {0, 0, 0, 0};
  uint32_t* d_matcap = nullptr;
  uint64_t matcap_capacity = 0;
  uint32_t matcap_w = 0, matcap_h = 0, cavity = 0;

  int fail(int code, const std::string& msg);
  int ensure_spill(uint64_t lanes, uint32_t entries);
  int build_object_table(const uint32_t* slots, uint32_t n, uint32_t& max_stack, int& width);
  // the device records of the listed meshes into tab[0..n) (host memory; nothing allocated)
  int fill_object_table(const uint32_t* slots, uint32_t n, DevObject* tab, uint32_t& max_stack, int& width);
  int build_top_level(const uint32_t* slots, uint32_t n);
  // scratch buffer i (0..3) of at least `bytes`, kept for later calls; nullptr when allocation fails
  void* scratch_buffer(int i, size_t bytes) {
    if (bytes <= scratch_bytes[i]) return scratch[i];
    hipStreamSynchronize(stream);
    if (scratch[i]) hipFree(scratch[i]);
    scratch[i] = nullptr;
    scratch_bytes[i] = 0;
    if (hipMalloc(&scratch[i], std::max<size_t>(bytes, 256)) != hipSuccess) {
      scratch[i] = nullptr;
      return nullptr;
    }
    scratch_bytes[i] = std::max<size_t>(bytes, 256);
    return scratch[i];
  }
  // top-level box widening for rays whose origins lie within origin_radius (max-norm) of the world origin
  float top_delta(float origin_radius) const;
  int upload_object_table(const DevObject* tab, uint32_t n);
  // learned proposal for the frame's objects into fp.proposal (mlp.hip)
  // the proposal network over a frame's pixels into fp.proposal; objs: the frame's
  // object table (nullptr: the current single-frame table); the render's mlp
  // events are recorded before the launch when `first`, after it when `last`
  int launch_mlp_frame(const FrameParams& fp, uint32_t n_obj, const DevObject* objs = nullptr, bool first = true,
                       bool last = true);
  int reset_counters();  // work-queue heads and statistics counters
  int launch_trace(const FrameParams& fp, uint32_t max_stack, int width, bool learned, Target* tg = nullptr);
  int launch_trace_kernel(int mode, const FrameParams& fp, const RayQuery& q, uint32_t n_rays, uint32_t n_obj,
                          uint32_t max_stack, int width, uint32_t* qcounter, const BatchParams* bp = nullptr);
  int ensure_shadow_queue(uint64_t rays);
  float4* d_shadow = nullptr;
  uint64_t shadow_capacity = 0;
  uint64_t* d_timeline = nullptr;   // counting mode: per-wave timeline of the primary trace
  uint64_t timeline_capacity = 0;
  uint32_t timeline_waves = 0;
  uint32_t trace_grid(uint32_t n_rays, uint32_t D, uint32_t block, uint32_t extra_lds = 0) const;
  int launch_shade(const ShadeParams& sp);
  // batches (j3dg_render_batch): device tables, and the captured graph of
  // reset + trace + shade replayed while its launch parameters are unchanged
  FrameParams* d_bframes = nullptr;
  uint32_t* d_bfirst = nullptr;
  ShadeParams* d_bshade = nullptr;
  DevObject* d_bobjs = nullptr;
  uint32_t bframes_cap = 0, bobjs_cap = 0;
  std::vector<uint8_t> batch_tables;  // what the device tables hold
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<uint8_t> graph_key;
  uint64_t graph_launches = 0;
  void invalidate_graphs() {
    if (graph_exec) hipGraphExecDestroy(graph_exec);
    if (graph) hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    graph_key.clear();
  }
  void release_buffers() {
    if (d_queue) hipFree(d_queue);
    if (d_counters) hipFree(d_counters);
    if (spill) hipFree(spill);
    if (d_obj_big) hipFree(d_obj_big);
    if (d_obj_ring) hipFree(d_obj_ring);
    if (h_obj_ring) hipHostFree(h_obj_ring);
    for (int i = 0; i < kObjRing; ++i)
      if (obj_ring_ev[i]) hipEventDestroy(obj_ring_ev[i]), obj_ring_ev[i] = nullptr;
    d_obj_big = nullptr, d_obj_ring = nullptr, h_obj_ring = nullptr, obj_capacity = 0, obj_ring_cur = -1;
    for (int i = 0; i < kObjRing; ++i) obj_ring_used[i] = false;
    obj_cache.clear();
    if (d_top) hipFree(d_top);
    for (int i = 0; i < 4; ++i)
      if (scratch[i]) hipFree(scratch[i]), scratch[i] = nullptr, scratch_bytes[i] = 0;
    if (d_matcap) hipFree(d_matcap);
    if (d_shadow) hipFree(d_shadow);
    d_shadow = nullptr;
    if (d_timeline) hipFree(d_timeline);
    d_timeline = nullptr;
    invalidate_graphs();
    if (d_bframes) hipFree(d_bframes);
    if (d_bfirst) hipFree(d_bfirst);
    if (d_bshade) hipFree(d_bshade);
    if (d_bobjs) hipFree(d_bobjs);
    d_bframes = nullptr, d_bfirst = nullptr, d_bshade = nullptr, d_bobjs = nullptr;
    bframes_cap = bobjs_cap = 0;
    batch_tables.clear();
    d_queue = nullptr, d_counters = nullptr, spill = nullptr, d_objs = nullptr, d_matcap = nullptr;
    d_top = nullptr;
  }
};

// Device upload and build of a mesh of 2 .. 2^29-1 triangles (bvh_gpu.hip):
// fills m's device arrays, n_nodes, width, max_depth, device_bytes and
// max_stack; m.info's bounding box must be set.  upload_s receives the time of
// the host -> device copies.  On failure the caller releases m.
int device_build_mesh(Ctx* ctx, Mesh& m, const float* xyz, uint64_t n_vertices, const uint32_t* tris,
                      uint64_t n_triangles, int width, int leaf_size, double* upload_s);

// canvas::update_canvas's frame set-up (j3d/canvas.cpp:679-705) on the host.
inline int make_frame_params(const j3dg_frame_desc& fd, const Target& tg, FrameParams& fp) {
  std::memset(&fp, 0, sizeof(fp));
  const uint32_t w = fd.width, h = fd.height;
  if (w == 0 || h == 0) return J3DG_ERR_INVALID_ARG;
  int x0 = fd.x0, y0 = fd.y0, x1 = fd.x1, y1 = fd.y1;
  if (x0 < 0) x0 = 0;
  if (y0 < 0) y0 = 0;
  if (x1 < 0) x1 = 0;
  if (y1 < 0) y1 = 0;
  if (x0 >= (int)w) x0 = w - 1;
  if (y0 >= (int)h) y0 = h - 1;
  if (x1 >= (int)w) x1 = w - 1;
  if (y1 >= (int)h) y1 = h - 1;
  fp.w = w, fp.h = h;
  fp.x0 = x0, fp.y0 = y0, fp.x1 = x1, fp.y1 = y1;
  // an inverted rectangle renders nothing (the reference's loops do not run)
  uint32_t rw = x1 >= x0 ? (uint32_t)(x1 - x0 + 1) : 0, rh = y1 >= y0 ? (uint32_t)(y1 - y0 + 1) : 0;
  fp.tiles_x = (rw + 7) / 8;
  fp.n_tiles = fp.tiles_x * ((rh + 7) / 8);
  std::memcpy(fp.proj_inv, fd.proj_inv, sizeof(fp.proj_inv));
  fp.near_plane = fd.near_plane;
  std::memcpy(fp.cs, fd.cs, sizeof(fp.cs));
  std::memcpy(fp.cs_inv, fd.cs_inv, sizeof(fp.cs_inv));
  fp.diagonal = fd.diagonal;
  f4 light = {fd.pivot[0] + fd.diagonal * 3.f, fd.pivot[1] + fd.diagonal * 3.f, fd.pivot[2] + fd.diagonal * 3.f, 1.f};
  light = mat_vec(fd.cs, light);
  fp.light[0] = light.x, fp.light[1] = light.y, fp.light[2] = light.z, fp.light[3] = light.w;
  fp.light_dir[0] = fd.light_dir[0], fp.light_dir[1] = fd.light_dir[1], fp.light_dir[2] = fd.light_dir[2];
  fp.flags = fd.flags;
  fp.n_obj = fd.n_objects;
  fp.gbuf = tg.gbuf;
  fp.margin = fd.margin;
  return J3DG_OK;
}

inline void make_shade_params(const j3dg_frame_desc& fd, const Target& tg, const Ctx& ctx, ShadeParams& sp) {
  std::memset(&sp, 0, sizeof(sp));
  sp.w = tg.w, sp.h = tg.h;
  sp.gbuf = tg.gbuf;
  sp.rgba = tg.rgba;
  sp.matcap = ctx.d_matcap;
  sp.mw = ctx.matcap_w, sp.mh = ctx.matcap_h, sp.cavity = ctx.cavity;
  sp.flags = fd.flags;
  std::memcpy(sp.proj_inv, fd.proj_inv, sizeof(sp.proj_inv));
  sp.near_plane = fd.near_plane;
  sp.bg_top = fd.bg_top, sp.bg_bottom = fd.bg_bottom;
}

}  // namespace j3d

// Error plumbing shared by the library's translation units: record the HIP
// error text in the context and return a j3dg status code.
#define HIP_TRY(expr, code)                                                                         \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess) return ctx->fail(code, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define HIP_CALL(expr, code) HIP_TRY(expr, code)
// Entry-point prologue: a valid context, held for the whole call.
#define CHECK_CTX()                      \
  if (!ctx) return J3DG_ERR_INVALID_ARG; \
  std::lock_guard<std::mutex> lock_(ctx->mu)
