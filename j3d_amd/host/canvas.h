// canvas.h -- j3d's canvas / scene / camera render interface (j3d/canvas.h:12-102,
// j3d/scene.h:10-56, j3d/camera.h) over the gfx950 ray-casting library.
//
// The public methods keep the reference's names, argument meaning and call
// order (view.cpp:421-430: update_settings -> render_scene -> get_pixels ->
// canvas_to_image), so a viewer that drives j3d's canvas drives this one the
// same way; the frame is traced and shaded by the HIP kernels behind
// include/j3dg.h, and `get_image()` / `get_pixels()` hold the same RGBA
// (0xAABBGGRR, row 0 = top) and 32-byte G-buffer records.
//
// Differences a caller sees:
//  * A `device` (one j3dg context on one GPU) is shared by the scene and the
//    canvases that render it; `add_object` uploads the mesh and builds its BVH
//    there (the reference builds a jtk::qbvh on the host, scene.cpp:23).
//  * Mesh arrays are borrowed only during add_object (the device keeps its
//    own copy); the scene_object still points at them, like the reference.
//  * image<T> is a plain row-major image without stride padding.  It carries
//    a generation tag so that a G-buffer handed back to canvas_to_image
//    unmodified is shaded from the device copy instead of being re-uploaded;
//    any non-const element access clears the tag.
//  * Errors from the library are thrown as j3d_amd::gpu_error (the reference
//    returns void; its only throw is blit_onto's alignment check, kept).
#pragma once
#include <stdint.h>

#include <array>
#include <list>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "j3dg.h"
#include "j3dg_host.h"

namespace j3d_amd {

typedef j3dg_pixel pixel;     // j3d/pixel.h:11-27, byte-identical
typedef j3dg_camera camera;   // j3d/camera.h
typedef std::array<float, 16> float4x4;  // column-major, as jtk::float4x4

struct gpu_error : std::runtime_error {
  int code;
  gpu_error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

template <class T>
class image {
 public:
  image() = default;
  image(uint32_t w, uint32_t h) : w_(w), h_(h), px_((size_t)w * h) {}
  uint32_t width() const { return w_; }
  uint32_t height() const { return h_; }
  const T* row(uint32_t y) const { return px_.data() + (size_t)y * w_; }
  T* row(uint32_t y) {
    tag_ = 0;
    return px_.data() + (size_t)y * w_;
  }
  const T& operator()(uint32_t x, uint32_t y) const { return px_[(size_t)y * w_ + x]; }
  T& operator()(uint32_t x, uint32_t y) {
    tag_ = 0;
    return px_[(size_t)y * w_ + x];
  }
  const T* data() const { return px_.data(); }
  T* data() {
    tag_ = 0;
    return px_.data();
  }
  typename std::vector<T>::const_iterator begin() const { return px_.begin(); }
  typename std::vector<T>::const_iterator end() const { return px_.end(); }
  typename std::vector<T>::iterator begin() {
    tag_ = 0;
    return px_.begin();
  }
  typename std::vector<T>::iterator end() {
    tag_ = 0;
    return px_.end();
  }
  uint64_t tag() const { return tag_; }
  void set_tag(uint64_t t) { tag_ = t; }

 private:
  uint32_t w_ = 0, h_ = 0;
  std::vector<T> px_;
  uint64_t tag_ = 0;
};

// matcap (j3d/matcap.h:18-24)
struct matcap {
  image<uint32_t> im;
  uint32_t cavity_clr = 0;
};
// make_matcap_red_wax / gray / brown / sketch (j3d/matcap.cpp)
matcap make_matcap_red_wax();
matcap make_matcap_gray();
matcap make_matcap_brown();
matcap make_matcap_sketch();

// make_default_camera, make_projection_matrix, invert_projection_matrix (j3d/camera.cpp:5-73)
camera make_default_camera();
float4x4 make_projection_matrix(const camera& c, int w, int h);
float4x4 invert_projection_matrix(const float4x4& m);
float4x4 invert_orthonormal(const float4x4& m);
float4x4 get_identity();

// One GPU: the j3dg context that scenes upload to and canvases render with.
class device {
 public:
  explicit device(int gpu = 0);
  ~device();
  device(const device&) = delete;
  device& operator=(const device&) = delete;
  j3dg_ctx* ctx() const { return ctx_; }
  void check(int rc, const char* what) const;

 private:
  j3dg_ctx* ctx_ = nullptr;
};

// scene_object (j3d/scene.h:10-25): vertices xyz and triangles as flat arrays.
struct scene_object {
  uint32_t db_id = 0;
  const std::vector<float>* p_vertices = nullptr;     // 3 floats per vertex
  const std::vector<uint32_t>* p_triangles = nullptr;  // 3 indices per triangle
  float min_bb[3] = {0, 0, 0}, max_bb[3] = {0, 0, 0};
  float4x4 cs = get_identity();
  uint32_t slot = 0;          // device mesh slot
  float4x4 device_cs = get_identity();
};

// scene_pointcloud (j3d/scene.h:27-38): points, optional normals and colours
// (0xAABBGGRR), copied to the device by add_pointcloud.
struct scene_pointcloud {
  uint32_t db_id = 0;
  float min_bb[3] = {0, 0, 0}, max_bb[3] = {0, 0, 0};  // compute_bb of the points (j3d/scene.cpp:34)
  float4x4 cs = get_identity();
  uint32_t cloud = 0;  // device cloud id
  float4x4 device_cs = get_identity();
};

// scene (j3d/scene.h:40-56)
struct scene {
  explicit scene(std::shared_ptr<device> d) : dev(std::move(d)) {}
  ~scene();
  std::shared_ptr<device> dev;
  float4x4 coordinate_system = get_identity(), coordinate_system_inv = get_identity();
  float pivot[3] = {0, 0, 0};
  float min_bb[3] = {0, 0, 0}, max_bb[3] = {0, 0, 0};
  float diagonal = 0.f;
  std::list<scene_object> objects;
  std::list<scene_pointcloud> pointclouds;
};

// add_object (j3d/scene.cpp:8-39) for a mesh held by the caller: normals, bounding
// box and BVH are computed by the device upload.  bvh_width: 4 or 8.
// builder: J3DG_BUILD_HOST (binned SAH on the host) or J3DG_BUILD_DEVICE
// (normals and a linear BVH built on the GPU from the raw arrays).
void add_object(uint32_t id, scene& s, const std::vector<float>& vertices, const std::vector<uint32_t>& triangles,
                const float4x4& cs = get_identity(), uint32_t bvh_width = 8, uint32_t builder = J3DG_BUILD_HOST);
// The surface colour sources of object `id`: scene_object::p_vertex_colors,
// p_uv_coordinates and p_texture (j3d/scene.h:16-18, scene.cpp:17-19), used
// under canvas_settings::vertexcolors / textured (canvas.cpp:803-834).
// vertex_colors: rgb per vertex; uv: (u, v) of each triangle's corners (6
// floats per triangle) with `texture`; null or empty = none.  Borrowed only
// during the call.
void set_object_colors(uint32_t id, scene& s, const std::vector<float>* vertex_colors,
                       const std::vector<float>* uv = nullptr, const image<uint32_t>* texture = nullptr);
// A point cloud of the scene (scene::pointclouds, j3d/scene.h:50): 3 floats per
// point; normals (3 floats per point) and colours may be null or empty.
void add_pointcloud(uint32_t id, scene& s, const std::vector<float>& vertices, const std::vector<float>* normals,
                    const std::vector<uint32_t>* colors, const float4x4& cs = get_identity());
void remove_pointcloud(uint32_t id, scene& s);
// remove_object (j3d/scene.cpp:41-49): the object and the point cloud with this id
void remove_object(uint32_t id, scene& s);
// prepare_scene / unzoom (j3d/scene.cpp:50-110)
void prepare_scene(scene& s);
void unzoom(scene& s);

class canvas {
 public:
  // j3d/canvas.h:16-25
  struct canvas_settings {
    bool one_bit;
    bool shadow;
    bool edges;
    bool wireframe;
    bool shading;
    bool textured;
    bool vertexcolors;
  };

  explicit canvas(std::shared_ptr<device> dev);
  canvas(std::shared_ptr<device> dev, uint32_t w, uint32_t h);
  ~canvas();

  void resize(uint32_t w, uint32_t h);
  void get_pixel(pixel& p, float pos_x, float pos_y, float mouse_offset_x, float mouse_offset_y);
  void update_canvas(image<pixel>& out, int x0, int y0, int x1, int y1, const scene& s);
  void update_canvas(int x0, int y0, int x1, int y1, const scene& s);
  void render_scene(image<pixel>& out, const scene* s);
  void render_scene(const scene* s);
  void set_background_color(uint32_t clr_top = 0xff000000, uint32_t clr_bottom = 0xff404040);
  void blit_onto(image<uint32_t>& screen, int32_t pos_x, int32_t pos_y);

  uint32_t width() const { return w_; }
  uint32_t height() const { return h_; }
  void update_settings(const canvas_settings& s) { _settings = s; }

  void canvas_to_image(const image<pixel>& canvas, const matcap& _matcap);
  void canvas_to_image(const matcap& _matcap);
  // j3d/canvas.cpp:952-1030: the scene's point clouds drawn into the image and
  // the G-buffer, under a z-buffer from `pix` (see j3dg_splat_points).
  void render_pointclouds_on_image(const scene* s, const image<pixel>& pix);

  const float4x4& get_projection_matrix() const { return projection_matrix; }
  const float4x4& get_inverse_projection_matrix() const { return projection_matrix_inv; }
  const camera& get_camera() const { return _camera; }
  const image<pixel>& get_pixels() const;
  const image<uint32_t>& get_image() const { return im; }
  image<uint32_t>& get_image() { return im; }

  // ---- additions of the GPU path (no counterpart in the reference) ----
  // The reference canvas always uses make_default_camera(); this sets another
  // camera model (e.g. a 60 degree field of view) and recomputes the projection.
  void set_camera(const camera& c);
  // Shadow rays toward a direction instead of the point light at pivot + 3 diagonal.
  void set_directional_light(bool on, const float dir[3] = nullptr);
  // Learned tmax bound for primary rays (meshes need a trained proposal network).
  void set_learned(bool on, float margin);
  // Durations of the last frame's kernels (ms): trace, shade, mlp, finish + shadow.
  void last_timings(float* trace_ms, float* shade_ms, float* mlp_ms, float* shadow_ms) const;
  uint32_t target() const { return target_; }

 private:
  j3dg_frame_desc frame_desc(const scene* s, int x0, int y0, int x1, int y1) const;
  void sync_objects(const scene& s);
  void upload_matcap(const matcap& m);
  void pull_pixels() const;

  std::shared_ptr<device> dev_;
  uint32_t target_ = 0xFFFFFFFFu;
  uint32_t w_ = 0, h_ = 0;
  image<uint32_t> im;
  mutable image<pixel> _canvas;
  mutable bool pixels_stale_ = false;  // the device G-buffer is newer than _canvas
  uint64_t generation_ = 0;            // tag of the device G-buffer's contents
  camera _camera;
  float4x4 projection_matrix, projection_matrix_inv;
  canvas_settings _settings{};
  uint32_t bg_top_ = 0, bg_bottom_ = 0;
  bool directional_ = false, learned_ = false;
  // the scene's mesh slots when it has more objects than j3dg_frame_desc::objects holds
  mutable std::vector<uint32_t> object_slots_;
  float light_dir_[3] = {0, 0, 1}, margin_ = 0.f;
  const void* matcap_key_ = nullptr;
  uint32_t matcap_cavity_ = 0, matcap_w_ = 0;
};

}  // namespace j3d_amd
