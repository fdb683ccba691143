/*
 * oracle.h -- CPU restatement of j3d's per-frame ray-casting path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path in
 * j3d_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product library never links or calls it.
 *
 * What it restates (reference paths relative to the j3d repository root):
 *   camera model         j3d/camera.cpp:5-73
 *   ray generation       j3d/canvas.cpp:700-705, 773-782
 *   G-buffer write       j3d/canvas.cpp:788-801, 859-866, j3d/pixel.h:11-27
 *   shadow ray           j3d/canvas.cpp:836-857
 *   canvas_to_image      j3d/canvas.cpp:287-446, 449-670
 *   background gradient  j3d/canvas.cpp:55-78
 *   matcap tables        j3d/matcap.cpp:9-268
 *   scene bbox/unzoom    j3d/scene.cpp:50-110, j3d/mesh.cpp:39-58
 * and, because the reference's jtk submodule (qbvh, vec math) is absent, it
 * defines the jtk-owned arithmetic itself:
 *   - float4x4 is column-major, m[12..14] the translation; a matrix-vector
 *     product sums its four terms left to right;
 *   - the ray/triangle test is Moller-Trumbore on (v0, e1=v1-v0, e2=v2-v0),
 *     two-sided, det==0 rejects, inclusive barycentric and [t_near, t_far];
 *   - closest hit is the lexicographic minimum of (t, object index, triangle
 *     id) over all accepted triangles, so it does not depend on BVH topology
 *     or traversal order;
 *   - triangle normal = cross(v1-v0, v2-v0) / |cross(v1-v0, v2-v0)|.
 * Parity with the real jtk numerics is unpinned (jtk is not available); the
 * checks that pin this restatement are listed in tests/test_oracle.py.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* j3d/pixel.h:11-27, 32 bytes */
typedef struct orc_pixel {
  uint8_t mark, r, g, b;
  float u, v, depth;
  uint32_t object_id;
  float barycentric_u, barycentric_v;
  uint32_t db_id;
} orc_pixel;

/* j3d/camera.h:5-16 */
typedef struct orc_camera {
  float focal_length;
  float film_w_inch;
  float film_h_inch;
  float near_plane;
  float far_plane;
  float zoom;
  int32_t fit_film; /* 0 = Fill, 1 = Overscan */
} orc_camera;

/* canvas_settings bits (j3d/canvas.h:16-25) plus the directional-light extension */
enum {
  ORC_ONE_BIT = 1u << 0,
  ORC_SHADOW = 1u << 1,
  ORC_EDGES = 1u << 2,
  ORC_WIREFRAME = 1u << 3,
  ORC_SHADING = 1u << 4,
  ORC_TEXTURED = 1u << 5,
  ORC_VERTEXCOLORS = 1u << 6,
  ORC_DIRECTIONAL_LIGHT = 1u << 7
};

void orc_default_camera(orc_camera* c);
void orc_frustum(float l, float r, float b, float t, float n, float f, float out[16]);
void orc_invert_projection(const float in[16], float out[16]);
void orc_projection(const orc_camera* c, int w, int h, float P[16]);
void orc_invert_orthonormal(const float m[16], float out[16]);
void orc_mat_vec(const float m[16], const float v[4], float out[4]);

/* type: 0 red wax, 1 gray, 2 brown, 3 sketch.  im: 512*512 row-major. */
void orc_matcap(int type, uint32_t* im, uint32_t* cavity);
void orc_fill_background(uint32_t* im, uint32_t w, uint32_t h, uint32_t top, uint32_t bottom);

void orc_triangle_normals(const float* v, const uint32_t* t, uint64_t nt, float* n);
void orc_bbox(const float* v, uint64_t nv, float mn[3], float mx[3]);

/* Scene framing, j3d/scene.cpp:50-110: the (cs, bbox) of the scene's objects, then of its
   point clouds -> diagonal, pivot, camera cs. */
void orc_prepare_and_unzoom(uint32_t nobj, const float* cs16, const float* bbmin3, const float* bbmax3,
                            float* diagonal, float pivot[3], float cs[16], float cs_inv[16]);

typedef struct orc_bvh orc_bvh;
orc_bvh* orc_bvh_build(const float* v, uint64_t nv, const uint32_t* t, uint64_t nt, int threads);
void orc_bvh_free(orc_bvh* b);
uint64_t orc_bvh_node_count(const orc_bvh* b);

typedef struct orc_object {
  const orc_bvh* bvh;      /* may be NULL: brute force over all triangles */
  const float* vertices;   /* xyz float32 */
  const uint32_t* triangles;
  const float* normals;    /* per-triangle xyz */
  uint64_t n_triangles;
  float cs[16];            /* object -> world */
  uint32_t db_id;
  /* optional surface colour (j3d/scene.h:16-18; canvas.cpp:803-834): NULL = none */
  const float* vertex_colors;  /* rgb float per vertex */
  const float* uv;             /* per triangle: (u, v) of its three corners */
  const uint32_t* texture;     /* tex_w x tex_h, row-major, 0xAABBGGRR */
  uint32_t tex_w, tex_h;
} orc_object;

/* The byte a colour channel float c becomes: (uint8_t)(c * 255.f) of
   j3d/canvas.cpp:829-831, defined here as truncation toward zero to int32
   (NaN and values outside the int32 range give 0), then the low 8 bits. */
uint8_t orc_color_byte(float c);
/* Surface colour of a hit (j3d/canvas.cpp:803-834).  Returns 1 and the
   channels when the object has one under `flags`, else 0. */
int orc_hit_color(const orc_object* ob, uint32_t flags, uint32_t tri, float u, float v, uint8_t rgb[3]);

typedef struct orc_frame {
  uint32_t w, h;
  int32_t x0, y0, x1, y1;  /* inclusive rectangle, clamped as canvas.cpp:682-698 */
  float proj_inv[16];
  float near_plane;
  float cs[16], cs_inv[16];
  float pivot[3];
  float diagonal;
  uint32_t flags;
  float light_dir[3];      /* used when ORC_DIRECTIONAL_LIGHT is set */
} orc_frame;

typedef struct orc_stats {
  uint64_t primary_rays, shadow_rays, hits, shadowed;
  uint64_t near_ties;      /* only counted when count_ties != 0 */
} orc_stats;

/* canvas::update_canvas (j3d/canvas.cpp:677-874) on the given G-buffer. */
int orc_update_canvas(const orc_object* objs, uint32_t nobj, const orc_frame* f, orc_pixel* gbuf,
                      int threads, int count_ties, orc_stats* st);

/* canvas::canvas_to_image (j3d/canvas.cpp:582-670).  im must already hold the
   background (canvas::render_scene copies it first, canvas.cpp:895). */
void orc_canvas_to_image(const orc_pixel* gbuf, uint32_t w, uint32_t h, const float proj_inv[16],
                         float near_plane, uint32_t flags, const uint32_t* matcap, uint32_t mw,
                         uint32_t mh, uint32_t cavity, uint32_t* im);

/* Raw closest-hit queries.  rays: n x 8 floats (ox,oy,oz,dx,dy,dz,t_near,t_far) in world space.
   Outputs (any may be NULL): object index (0xFFFFFFFF on miss), triangle id, t, u, v,
   and second-nearest t (FLT_MAX if none) when t2 != NULL. */
void orc_closest_hits(const orc_object* objs, uint32_t nobj, const float* rays, uint64_t n, int threads,
                      uint32_t* obj, uint32_t* tri, float* t, float* u, float* v, float* t2);
/* Any-hit queries: hit[i] = 1 if any triangle is accepted in [t_near, t_far]. */
void orc_any_hits(const orc_object* objs, uint32_t nobj, const float* rays, uint64_t n, int threads,
                  uint8_t* hit);

/* All hits of each ray in one object (jtk::qbvh::find_all_triangles, j3d/vox.cpp:333): every
   triangle the test accepts with t in [t_near, t_far], in (t, triangle id) order; the first
   max_hits are stored at [i * max_hits ...], counts[i] is the ray's total.  World rays, taken to
   the object's frame like the closest-hit queries. */
void orc_all_hits(const orc_object* ob, const float* rays, uint64_t n, uint32_t max_hits, int threads,
                  uint32_t* counts, uint32_t* tri, float* t, float* u, float* v);

/* write_vox's grid size (j3d/vox.cpp:284-297) for a mesh box: dim[k] = (uint32_t)(max_dim *
   extent_k / largest extent) in float, at least 1 (a NaN ratio counts as 0). */
void orc_vox_dims(const float bb_min[3], const float bb_max[3], uint32_t max_dim, uint32_t dim[3]);
/* write_vox's voxel grid (j3d/vox.cpp:278-376) of one mesh in its own frame: rays along x, y and z
   (direction (0,0,2) for z) through every cell centre of the two other axes, every hit's position
   and colour (texture over vertex colours over white) mapped to a voxel and a palette index
   (color_to_index, vox.cpp:154-171).  data[x + (y + z * dim[1]) * dim[0]], 0 = empty.  Where
   several hits land in one voxel, the last in sequential order (axis, row, column, hit in (t, id)
   order) wins.  Float-to-integer conversions saturate into range (NaN -> 0) where the reference's
   casts would be undefined.  The object's cs is not applied. */
void orc_voxelize(const orc_object* ob, const float bb_min[3], const float bb_max[3], const uint32_t dim[3],
                  uint8_t* data);
/* vox.cpp:154-171 */
uint8_t orc_vox_color_index(uint8_t r, uint8_t g, uint8_t b);

/* Point clouds drawn over a G-buffer and image (j3d/canvas.cpp:952-1030): the z-buffer of
   canvas.cpp:963-970, then one call per cloud in scene order, each drawing its points in index
   order (see oracle.cpp for the restated projection and colour). */
void orc_splat_zbuffer(const orc_pixel* gbuf, uint64_t n_pix, float* z);
void orc_splat_cloud(orc_pixel* gbuf, uint32_t* im, uint32_t w, uint32_t h, float* z, const float* proj,
                     const float* cs_inv, const float* cs, uint32_t db_id, const float* xyz, uint64_t n,
                     const float* normals, const uint32_t* colors, int shading, int one_bit);

/* Single ray/triangle test, exposed for known-answer tests. Returns 1 on hit. */
int orc_intersect_triangle(const float o[3], const float d[3], const float v0[3], const float v1[3],
                           const float v2[3], float t_near, float t_far, float* t, float* u, float* v);

#ifdef __cplusplus
}
#endif
