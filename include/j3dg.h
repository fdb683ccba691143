/*
 * j3dg.h -- C ABI of the MI355X (gfx950) ray-casting path of j3d.
 *
 * The reference renders a frame on the CPU inside the C++ class `canvas`
 * (j3d/canvas.h:12-102), driven by view::render_scene (j3d/view.cpp:421-430):
 *     update_settings(flags) -> render_scene(&scene) -> get_pixels()
 *     -> canvas_to_image(pixels, matcap)
 * This header is the boundary that a `canvas` implementation (or any FFI)
 * binds to replace that path with HIP kernels on one GPU.  Plain C types
 * only; every call returns an int status (J3DG_OK == 0); no C++ exception
 * crosses it; j3dg_last_error() returns the message of the last failure.
 * A context is not re-entrant: serialise calls per context, as the reference
 * serialises render calls through view::_mut (j3d/view.cpp:1262).
 *
 * Matrices are float[16], column-major, translation in m[12..14]
 * (the jtk float4x4 layout used throughout j3d/canvas.cpp).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define J3DG_ABI_VERSION 2

/* status codes */
#define J3DG_OK 0
#define J3DG_ERR_INVALID_ARG 1
#define J3DG_ERR_NO_DEVICE 2
#define J3DG_ERR_OUT_OF_MEMORY 3
#define J3DG_ERR_LAUNCH 4
#define J3DG_ERR_NOT_FOUND 5
#define J3DG_ERR_UNSUPPORTED 6
#define J3DG_ERR_INTERNAL 7

/* Frame flags.  Bits 0-6 mirror canvas::canvas_settings (j3d/canvas.h:16-25). */
#define J3DG_ONE_BIT (1u << 0)
#define J3DG_SHADOW (1u << 1)
#define J3DG_EDGES (1u << 2)
#define J3DG_WIREFRAME (1u << 3)
#define J3DG_SHADING (1u << 4)
#define J3DG_TEXTURED (1u << 5)
#define J3DG_VERTEXCOLORS (1u << 6)
/* Shadow rays toward the direction light_dir instead of the reference's point
   light at pivot + 3*diagonal (j3d/canvas.cpp:703-705). */
#define J3DG_DIRECTIONAL_LIGHT (1u << 7)
/* render_scene(nullptr) (j3d/canvas.cpp:883-890): only db_id and object_id are reset. */
#define J3DG_NO_SCENE (1u << 8)
/* Learned bound for primary rays: the objects' proposal networks predict each
   pixel's closest-hit distance; the trace first traverses with tmax = that
   prediction * (1 + margin), which is exact because any hit inside the bound
   is the global closest hit, and a ray that finds nothing there falls back to
   the unbounded traversal.  No counterpart in the reference. */
#define J3DG_LEARNED (1u << 9)
/* With J3DG_LEARNED: skip the network and use the target's proposal buffer as
   it stands (from an earlier learned frame or j3dg_write_target_proposal). */
#define J3DG_KEEP_PROPOSAL (1u << 10)

/* Parameters of the proposal network (78-64-64-64-2): W1[64][78], b1[64],
   W2[64][64], b2, W3[64][64], b3, W4[2][64], b4[2]. */
#define J3DG_MLP_PARAMS 13506

/* Render stages */
#define J3DG_STAGE_TRACE 1u /* canvas::render_scene: G-buffer */
#define J3DG_STAGE_SHADE 2u /* canvas::canvas_to_image: RGBA from the G-buffer */
#define J3DG_STAGE_ALL 3u

#define J3DG_MAX_OBJECTS 16

typedef struct j3dg_ctx j3dg_ctx;

/* G-buffer record, byte-identical to `pixel` (j3d/pixel.h:11-27). */
typedef struct j3dg_pixel {
  uint8_t mark, r, g, b;
  float u, v, depth;
  uint32_t object_id;
  float barycentric_u, barycentric_v;
  uint32_t db_id;
} j3dg_pixel;

/* Per-frame description: the state canvas::update_canvas reads
   (j3d/canvas.cpp:677-782) plus canvas_to_image's settings. */
typedef struct j3dg_frame_desc {
  uint32_t width, height;       /* canvas size (im.width(), im.height()) */
  int32_t x0, y0, x1, y1;       /* inclusive update rectangle, clamped as j3d/canvas.cpp:682-698 */
  float proj_inv[16];           /* canvas::projection_matrix_inv */
  float near_plane;             /* camera::nearClippingPlane (screen_pos.z) */
  float cs[16];                 /* scene::coordinate_system */
  float cs_inv[16];             /* scene::coordinate_system_inv */
  float pivot[3];               /* scene::pivot */
  float diagonal;               /* scene::diagonal */
  uint32_t flags;               /* J3DG_* flags */
  float light_dir[3];           /* J3DG_DIRECTIONAL_LIGHT: world direction toward the light */
  uint32_t bg_top, bg_bottom;   /* canvas::set_background_color gradient (j3d/canvas.cpp:55-78) */
  uint32_t n_objects;           /* scene objects in scene order (two-level index order) */
  uint32_t objects[J3DG_MAX_OBJECTS]; /* mesh slots from j3dg_upload_mesh */
  float margin;                 /* J3DG_LEARNED: bounded pass tmax = t_pred * (1 + margin) */
  const uint32_t* object_list;  /* nullable: the n_objects mesh slots, any number up to 65536 (scene.h:49
                                   holds any number of objects); replaces objects[].  Scenes of more than
                                   J3DG_MAX_OBJECTS objects are traversed through a top-level BVH over the
                                   objects (canvas.cpp:762); learned frames take at most J3DG_MAX_OBJECTS. */
} j3dg_frame_desc;

/* BVH builders (j3dg_build_opts.builder). */
#define J3DG_BUILD_HOST 0   /* binned SAH on the host, re-packed and uploaded (default) */
#define J3DG_BUILD_DEVICE 1 /* raw triangles uploaded; normals and a linear (Morton-order) BVH
                               built on the GPU: builds far faster, traverses slower.  Meshes
                               of fewer than two triangles are built on the host. */

/* BVH build options for j3dg_upload_mesh (0 = default). */
typedef struct j3dg_build_opts {
  uint32_t width;     /* 4 or 8 (default) children per node */
  uint32_t leaf_size; /* max triangles per leaf (1..4); 0 = 2 for width 8, 4 for width 4 */
  uint32_t threads;   /* host build threads, 0 = hardware_concurrency */
  uint32_t builder;   /* J3DG_BUILD_HOST or J3DG_BUILD_DEVICE */
} j3dg_build_opts;

typedef struct j3dg_mesh_info {
  uint64_t n_vertices, n_triangles, n_nodes;
  uint32_t width, max_depth;
  float bb_min[3], bb_max[3];   /* object-space bounding box (j3d/mesh.cpp:39-58) */
  double build_seconds;         /* host BVH build + re-pack */
  double upload_seconds;        /* host -> device copies */
  uint64_t device_bytes;
} j3dg_mesh_info;

/* Traversal tuning (chosen by the autotuner; 0 = default). */
/* Training of a mesh's proposal network (j3dg_default_train_opts fills the defaults). */
typedef struct j3dg_train_opts {
  uint64_t n_rays;          /* training rays, labelled by the plain traversal (default 2^24) */
  uint32_t epochs;          /* passes over the rays (default 4) */
  uint32_t batch;           /* rays per Adam step, a multiple of 128 (default 16384) */
  float lr, lr_final;       /* Adam learning rate, cosine-decayed from lr to lr_final */
  uint64_t seed;
  uint32_t rays_per_camera; /* rays sharing one random camera (default 1024) */
  float dist_min, dist_max; /* camera distance from the box centre, in half-diagonals (1.2, 3.5) */
  float fov_deg, aspect;    /* frustum of the random cameras (60 degrees vertical, 16:9) */
  uint32_t n_validation;    /* held-out rays for the report (default 65536) */
  uint32_t warm_start;      /* 1: start from the mesh's current network (resume), else a fresh init */
} j3dg_train_opts;

typedef struct j3dg_train_report {
  float loss_first, loss_last; /* mean loss of the first / last 1% of the steps */
  uint64_t steps;
  float hit_fraction;          /* of the training + validation rays */
  float label_ms, train_ms;    /* ray generation + labelling, optimisation */
  float val_hit_accuracy;      /* held-out rays: hit/miss predicted correctly */
  float val_mean_abs_error;    /* held-out hits: mean |y0 - s| (s: the label, in units of R) */
  float val_resolved[4];       /* held-out hits a bounded pass resolves at margins 0, 0.01, 0.03, 0.1 */
  float norm[4];               /* input normalisation: box centre, half-diagonal R */
} j3dg_train_report;

typedef struct j3dg_tuning {
  uint32_t stack_depth;    /* LDS stack entries per lane: 8, 16 or 32 */
  uint32_t rays_per_item;  /* rays each lane takes per queue fetch: 1, 2 or 4 */
  uint32_t coop_tail;      /* once a wave's rays run out, its last rays get groups of lanes: 0 = off;
                              1 = the last <= 8 rays get 8 lanes each, which split each node's children;
                              2, 3, 4 = from 8, 16 or 32 rays left, each ray gets 64 / (8, 16 or 32)
                              lanes, which traverse parts of its tree, each with a stack of its own;
                              as rays finish the rest regroup into wider groups, up to 8 lanes
                              (scenes with a top level use 1 for any value >= 1); default 4 */
  uint32_t block_size;     /* threads per workgroup: 64, 128 or 256 */
  uint32_t refill_min;     /* idle lanes of a wave that trigger handing out new rays: 1..64 (64 = whole waves); default 16 */
  uint32_t split_passes;   /* 0 = one launch per frame (a lane traces its pixel's shadow ray right after
                              the primary hit); 1 = primary pass, hit completion, shadow pass */
  uint32_t cost_order;     /* 1 = hand out a target's tiles heaviest first, by the work each tile took in
                              the target's previous frame with the same rectangle (results unchanged);
                              rectangles of fewer than 8192 8x8 tiles (about 720x720 pixels) skip it:
                              there the three sorting launches cost more than the shorter tail saves */
} j3dg_tuning;

/* Statistics of the last j3dg_render / j3dg_render_batch (with counting on,
   also of the last j3dg_any_hit_rays: its work in node_visits, tri_tests,
   wave_trips and lane_steps). */
typedef struct j3dg_stats {
  uint64_t primary_rays, shadow_rays;
  uint64_t node_visits, tri_tests;   /* plain-kernel counters (only when counting is on) */
  uint64_t hits, shadowed;
  uint64_t learned_resolved;         /* primary rays resolved by the bounded pass */
  uint64_t learned_fallback;         /* primary rays that needed the plain pass */
  uint64_t stack_overflow;           /* must be 0 */
  uint64_t wave_trips, lane_steps;   /* traversal loop trips per wave and steps per lane (counting on) */
  uint64_t shadow_node_visits, shadow_tri_tests, shadow_wave_trips, shadow_lane_steps;
  uint64_t max_ray_steps, max_shadow_ray_steps;   /* longest ray, in traversal steps (counting on) */
  uint64_t ray_steps_log2[7], shadow_ray_steps_log2[7]; /* rays with floor(log2(steps)) = 0..5, >= 6 */
  float trace_ms;                    /* primary-ray trace kernel (HIP events, last render) */
  float shade_ms;                    /* shading/edge kernel */
  float mlp_ms;                      /* learned proposal inference */
  float shadow_ms;                   /* k_finish + shadow-ray trace */
  float splat_ms;                    /* k_splat_depth of the last j3dg_splat_points, summed over its clouds */
} j3dg_stats;

/* ---- context ---------------------------------------------------------- */
int j3dg_create(int device, j3dg_ctx** out);
void j3dg_destroy(j3dg_ctx* ctx);
const char* j3dg_last_error(const j3dg_ctx* ctx);
int j3dg_abi_version(void);
int j3dg_device_count(int* count);
int j3dg_synchronize(j3dg_ctx* ctx);
int j3dg_set_tuning(j3dg_ctx* ctx, const j3dg_tuning* t);
int j3dg_get_tuning(j3dg_ctx* ctx, j3dg_tuning* t);
/* Per-ray node/triangle counters in the trace kernel (roofline accounting). */
int j3dg_set_counting(j3dg_ctx* ctx, int on);
int j3dg_get_stats(j3dg_ctx* ctx, j3dg_stats* st);
/* Per-render kernel durations (HIP events on the context's stream) of every
   render since the last reset (up to 1024): primary trace, shade, MLP and
   finish + shadow trace, in milliseconds (any pointer may be NULL). */
int j3dg_reset_timings(j3dg_ctx* ctx);
int j3dg_get_timings(j3dg_ctx* ctx, float* trace_ms, float* shade_ms, float* mlp_ms, float* shadow_ms, uint32_t max,
                     uint32_t* n);

/* ---- scene ------------------------------------------------------------ */
/* Replaces scene.cpp:8-25 add_object for a mesh: computes triangle normals
   (scene.cpp:20), the bounding box (scene.cpp:22) and the BVH (scene.cpp:23,
   jtk::qbvh), re-packs it for the GPU and uploads it -- or, with
   J3DG_BUILD_DEVICE, uploads the raw arrays and derives all of it on the
   GPU.  Host arrays are only borrowed for the call.  xyz: n_vertices*3 floats; tris: n_triangles*3. */
int j3dg_upload_mesh(j3dg_ctx* ctx, uint32_t db_id, const float* xyz, uint64_t n_vertices, const uint32_t* tris,
                     uint64_t n_triangles, const float cs[16], const j3dg_build_opts* opts, uint32_t* slot);
/* scene.cpp:43-52 remove_object */
int j3dg_remove_mesh(j3dg_ctx* ctx, uint32_t slot);
int j3dg_set_mesh_cs(j3dg_ctx* ctx, uint32_t slot, const float cs[16]);  /* cs finite, else invalid-arg */
/* A mesh's surface colour, scene_object::p_vertex_colors / p_uv_coordinates /
   p_texture (scene.h:16-18, set by scene.cpp:17-19), used by update_canvas's
   textured and vertex-colour branches (canvas.cpp:803-834) under
   J3DG_TEXTURED / J3DG_VERTEXCOLORS.  vertex_colors: rgb floats per vertex
   (n_vertices*3) or NULL; uv: (u, v) of each triangle's three corners
   (n_triangles*6 floats) or NULL; texture: tex_w*tex_h 0xAABBGGRR texels, row
   by row, required with uv.  NULL arrays remove that colour source (an empty
   vector in the reference). */
int j3dg_set_mesh_colors(j3dg_ctx* ctx, uint32_t slot, const float* vertex_colors, const float* uv,
                         const uint32_t* texture, uint32_t tex_w, uint32_t tex_h);
int j3dg_get_mesh_info(j3dg_ctx* ctx, uint32_t slot, j3dg_mesh_info* info);
/* Copies the per-triangle normals the upload computed (n_triangles*3 floats). */
int j3dg_mesh_normals(j3dg_ctx* ctx, uint32_t slot, float* out);

/* ---- canvas targets ----------------------------------------------------- */
/* matcap::im (w*h 0xAABBGGRR, row-major) and matcap::cavity_clr (j3d/matcap.h:18-24). */
int j3dg_set_matcap(j3dg_ctx* ctx, const uint32_t* im, uint32_t w, uint32_t h, uint32_t cavity);
/* canvas::resize (j3d/canvas.cpp:117-134): a device G-buffer initialised to
   pixel() defaults (object_id = 0xFFFFFFFF, db_id = 0, the rest 0) and an RGBA image. */
int j3dg_create_target(j3dg_ctx* ctx, uint32_t w, uint32_t h, uint32_t* target);
int j3dg_destroy_target(j3dg_ctx* ctx, uint32_t target);
/* Upload a host G-buffer into the target (canvas_to_image with caller pixels). */
int j3dg_write_target_gbuffer(j3dg_ctx* ctx, uint32_t target, const j3dg_pixel* gbuf);
/* Device addresses of a target's buffers (for zero-copy consumers). */
int j3dg_target_device_ptrs(j3dg_ctx* ctx, uint32_t target, void** gbuf, void** rgba);

/* One frame on the target.  stages: J3DG_STAGE_*.  If gbuf_out / rgba_out are
   non-NULL the call copies the results back and returns when they are ready;
   otherwise it returns after enqueueing (j3dg_synchronize to wait). */
int j3dg_render(j3dg_ctx* ctx, const j3dg_frame_desc* frame, uint32_t target, uint32_t stages,
                j3dg_pixel* gbuf_out, uint32_t* rgba_out);
/* Copy a target back: G-buffer and/or RGBA (either may be NULL). Synchronous. */
int j3dg_read_target(j3dg_ctx* ctx, uint32_t target, j3dg_pixel* gbuf, uint32_t* rgba);
/* As j3dg_read_target into images whose rows are `*_stride` elements apart
   (jtk::image<pixel> / jtk::image<uint32_t> rows padded to their stride(), as
   canvas.cpp:964-966 reads them): row y of the target goes to
   gbuf + y * gbuf_stride and rgba + y * rgba_stride; strides >= the width.
   The padding between rows is left untouched. */
int j3dg_read_target_strided(j3dg_ctx* ctx, uint32_t target, j3dg_pixel* gbuf, uint32_t gbuf_stride,
                             uint32_t* rgba, uint32_t rgba_stride);

/* Many frames (cameras x meshes of uneven size) in one persistent launch,
   captured in a HIP graph and replayed while the batch is unchanged.  Each
   frame equals the same frame rendered by j3dg_render into its target.  A
   J3DG_LEARNED frame's proposals come from its meshes' networks, run before
   the launch (or, with J3DG_KEEP_PROPOSAL, from its target as it stands).
   Refused: frames with textures or vertex colours, object lists, or
   J3DG_NO_SCENE; meshes of different BVH widths in one batch. */
int j3dg_render_batch(j3dg_ctx* ctx, const j3dg_frame_desc* frames, const uint32_t* targets, uint32_t n);

/* ---- autotuner (no counterpart in the reference) -------------------------
   Times the trace stages of `frame` (median of `repeats` frames, default 5)
   over LDS stack depth {8,16,32} x rays per work item {1,2,4}, then the refill
   threshold {8,16,32,48} and the tail mode (coop_tail) {1,3,4} at the fastest
   of those, applies the fastest tuning to the context and, for a J3DG_LEARNED
   frame, also picks the
   fastest bound margin of {0, 0.005, 0.01, 0.02, 0.05, 0.1, 0.2, 0.4, 0.8, 1.6, 3.2}
   (returned; the caller puts it in its frames).  The
   report's width is that of the frame's meshes. */
typedef struct j3dg_autotune_report {
  j3dg_tuning best;
  float best_margin;
  float best_ms;      /* trace stages of the frame (or one batch launch) with the chosen settings */
  float baseline_ms;  /* with the tuning in effect before the call */
  uint32_t n_tried;
  uint32_t width;     /* BVH width of the chosen frames (4 or 8) */
} j3dg_autotune_report;
int j3dg_autotune(j3dg_ctx* ctx, const j3dg_frame_desc* frame, uint32_t target, uint32_t repeats,
                  j3dg_autotune_report* report);
/* BVH width {4,8} as well: the same frame twice, its objects uploaded once per
   width (frame_w4 over width-4 slots, frame_w8 over width-8 slots; either may be
   NULL).  Each is tuned as j3dg_autotune; the faster is applied and reported. */
int j3dg_autotune_widths(j3dg_ctx* ctx, const j3dg_frame_desc* frame_w4, const j3dg_frame_desc* frame_w8,
                         uint32_t target, uint32_t repeats, j3dg_autotune_report* report);
/* A batch (j3dg_render_batch) over BVH width {4,8} x LDS stack depth {8,16,32} x
   rays per work item {1,2,4}, then the tail mode {1,3,4} at the fastest: the n
   frames given once per width (either list may be NULL), timed as one graph
   launch (median of `repeats`, default 3); applies the fastest tuning and
   reports its width. */
int j3dg_autotune_batch(j3dg_ctx* ctx, const j3dg_frame_desc* frames_w4, const j3dg_frame_desc* frames_w8,
                        const uint32_t* targets, uint32_t n, uint32_t repeats, j3dg_autotune_report* report);

/* ---- learned proposal (no counterpart in the reference) ---------------------
   Trains the mesh's proposal network on the GPU: random-camera rays around the
   mesh in its own coordinates, labelled by the plain traversal, fitted with
   the bf16-MFMA forward/backward kernels and Adam.  The network then serves
   every J3DG_LEARNED frame that contains the mesh (any placement). */
int j3dg_default_train_opts(j3dg_train_opts* opts);
int j3dg_train_proposal(j3dg_ctx* ctx, uint32_t slot, const j3dg_train_opts* opts, j3dg_train_report* report);
/* Checkpoint: J3DG_MLP_PARAMS floats and the normalisation (centre xyz, R). */
int j3dg_get_proposal(j3dg_ctx* ctx, uint32_t slot, float* params, float norm[4]);
/* params == NULL removes the mesh's network. */
int j3dg_set_proposal(j3dg_ctx* ctx, uint32_t slot, const float* params, const float norm[4]);
/* The target's proposal buffer: w*h pairs (tmax bound, hit logit); a pixel with
   logit <= 0 takes the plain traversal.  For J3DG_KEEP_PROPOSAL frames. */
int j3dg_write_target_proposal(j3dg_ctx* ctx, uint32_t target, const float* pairs);
int j3dg_read_target_proposal(j3dg_ctx* ctx, uint32_t target, float* pairs);
/* The network's kernels on host arrays (rays: n x 8 floats, object space;
   labels: n x 2 ((t |d| - a) / R, hit), a = (c - o) . d / |d|, see mlp_oracle.cpp).  forward -> out: n x 2 (y0, y1);
   gradients -> grads: J3DG_MLP_PARAMS floats of the mean loss
   hit * 10 (y0 - s)^2 + softplus(y1) - hit * y1, and *loss;
   adam_step: one Adam update (beta 0.9 / 0.999, eps 1e-8) of params, m, v. */
int j3dg_mlp_forward(j3dg_ctx* ctx, const float* params, const float norm[4], const float* rays, uint64_t n,
                     float* out);
int j3dg_mlp_gradients(j3dg_ctx* ctx, const float* params, const float norm[4], const float* rays,
                       const float* labels, uint64_t n, float* grads, float* loss);
int j3dg_mlp_adam_step(j3dg_ctx* ctx, float* params, float* m, float* v, const float* grads, uint32_t step,
                       float lr);

/* ---- raw ray queries (world space; used by tests and by the learned proposal) ----
   rays: n x 8 floats (ox,oy,oz,dx,dy,dz,t_near,t_far) in host memory.
   Closest hit over the listed mesh slots: lexicographic min of (t, object index, triangle id).
   A ray whose origin or direction is not finite hits nothing (in every query). */
int j3dg_trace_rays(j3dg_ctx* ctx, const uint32_t* slots, uint32_t n_slots, const float* rays, uint64_t n,
                    uint32_t* obj, uint32_t* tri, float* t, float* u, float* v);
int j3dg_any_hit_rays(j3dg_ctx* ctx, const uint32_t* slots, uint32_t n_slots, const float* rays, uint64_t n,
                      uint8_t* hit);
/* j3dg_trace_rays on DEVICE buffers, allocating nothing.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   rays: n x 8 floats on the context's device; obj, tri, t, u, v: n each on the
   device (none may be NULL); results as j3dg_trace_rays.  At most
   J3DG_MAX_OBJECTS slots.  Its working memory (queue words, counters, object
   table, traversal-stack spill area) is one fixed 16 MiB array per GPU, held
   until the process ends; a mesh set whose stack bound needs more spill than
   it holds is refused (J3DG_ERR_UNSUPPORTED).  Synchronous.  Calls that
   overlap on one GPU (from several contexts or threads) share that array and
   race: serialise them. */
int j3dg_trace_rays_device(j3dg_ctx* ctx, const uint32_t* slots, uint32_t n_slots, const float* rays, uint64_t n,
                           uint32_t* obj, uint32_t* tri, float* t, float* u, float* v);

/* All hits of each ray in one mesh (jtk::qbvh::find_all_triangles, called at j3d/vox.cpp:333): every
   triangle whose test accepts t in [t_near, t_far], in (t, triangle id) order.  counts[i] = the ray's
   total; the first max_hits hits go to [i * max_hits ...] of tri / t / u / v (each nullable; unused
   slots: tri 0xFFFFFFFF, 0.f).  World rays as j3dg_trace_rays, taken to the mesh's frame by its cs.
   n * max_hits <= 2^28, max_hits <= 65536. */
int j3dg_all_hits(j3dg_ctx* ctx, uint32_t slot, const float* rays, uint64_t n, uint32_t max_hits, uint32_t* counts,
                  uint32_t* tri, float* t, float* u, float* v);
/* write_vox's voxel grid of a mesh (j3d/vox.cpp:278-376), in the mesh's own frame: dim from the mesh
   box and max_dim (1 .. 1024) as vox.cpp:284-297; rays along x, y and z through every cell centre;
   every hit's position and colour (texture over vertex colours over white, j3dg_set_mesh_colors) to
   a voxel and a palette index (color_to_index, vox.cpp:154-171; index_to_color, vox.cpp:129-152, is
   the palette).  data[x + (y + z * dim[1]) * dim[0]] (capacity bytes), 0 = empty.  Where several
   hits land in one voxel, the sequentially last (axis, row, column, hit in (t, id) order) wins.
   With data NULL only dim is returned.  Float-to-integer casts saturate into range (NaN -> 0). */
int j3dg_voxelize(j3dg_ctx* ctx, uint32_t slot, uint32_t max_dim, uint32_t dim[3], uint8_t* data, uint64_t capacity);

/* ---- point clouds (canvas::render_pointclouds_on_image, j3d/canvas.cpp:952-1030) ----
   A cloud (scene_pointcloud, j3d/scene.h:27-38): n_points xyz, optional per-point normals (3 floats)
   and colours (0xAABBGGRR), db_id and pose cs (NULL = identity), copied to the device; at most
   2^32 - 2 points. */
int j3dg_upload_points(j3dg_ctx* ctx, uint32_t db_id, const float* xyz, uint64_t n_points, const float* normals,
                       const uint32_t* colors, const float cs[16], uint32_t* cloud);
int j3dg_remove_points(j3dg_ctx* ctx, uint32_t cloud);
int j3dg_set_points_cs(j3dg_ctx* ctx, uint32_t cloud, const float cs[16]);
/* Draws the clouds, in the order given, over a target holding a rendered G-buffer and image
   (after J3DG_STAGE_SHADE, as view::render_scene calls it, j3d/view.cpp:428).  proj = the canvas's
   projection matrix (canvas::projection_matrix, camera.cpp:32-73), cs_inv = scene CS^-1.  Each
   point goes to one pixel; it is drawn where its 1/w beats the z-buffer (1/depth on pixels with
   db_id != 0, else 0; canvas.cpp:963-970) and then sets object_id = its index, depth = 1/(1/w),
   db_id = the cloud's, and the image pixel to its colour: the cloud's, or 0xffffffff without colours
   or with J3DG_ONE_BIT; with J3DG_SHADING and normals, R, G, B scaled by the camera-space normal's z
   over its length, clamped to [0, 1].  On equal 1/w the earlier point (or cloud, or the surface)
   stays.  Projection and shading restate jtk's point renderer, which the reference does not ship. */
int j3dg_splat_points(j3dg_ctx* ctx, uint32_t target, const float proj[16], const float cs_inv[16], uint32_t flags,
                      const uint32_t* clouds, uint32_t n_clouds);

/* ---- measurement ---------------------------------------------------------
   A streaming device-to-device copy of `bytes` (read bytes, write bytes),
   repeated `repeats` times: the reference point that scales profiler byte
   counters to HBM bytes.  *ms receives the mean duration of one copy. */
int j3dg_calibration_copy(j3dg_ctx* ctx, uint64_t bytes, uint32_t repeats, float* ms);
/* With counting on, the primary trace kernel of the last render records one
   row of 8 uint64 per wave: device real-time clock (100 MHz, common to all
   XCDs) at start, when its ray pool ran dry, when it entered the cooperative
   tail mode (0 = never), at exit; loop trips, trips in tail mode, CU id, 0. */
int j3dg_debug_timeline(j3dg_ctx* ctx, uint64_t* out, uint32_t max_waves, uint32_t* n_waves);

#ifdef __cplusplus
}
#endif
