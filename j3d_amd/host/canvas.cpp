// canvas.cpp -- j3d's canvas / scene interface over the C ABI of libj3dg.so.
// Each method cites the reference method whose behaviour it keeps.
#include "canvas.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>

namespace j3d_amd {

namespace {

std::atomic<uint64_t> g_generation{0};
uint64_t next_generation() { return ++g_generation; }

// jtk-style column-major matrix times (x, y, z, 1), summed left to right.
void transform_point(const float4x4& m, const float p[3], float out[3]) {
  for (int r = 0; r < 3; ++r) out[r] = ((m[r] * p[0] + m[4 + r] * p[1]) + m[8 + r] * p[2]) + m[12 + r];
}

// compute_bb (j3d/mesh.cpp:39-58): std::min / std::max folded from the first point.
void compute_bb(const std::vector<float>& xyz, float lo[3], float hi[3]) {
  const size_t n = xyz.size() / 3;
  for (int k = 0; k < 3; ++k) lo[k] = hi[k] = n ? xyz[k] : 0.f;
  for (size_t i = 1; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], xyz[3 * i + k]);
      hi[k] = std::max(hi[k], xyz[3 * i + k]);
    }
}

matcap make_matcap(int type) {
  matcap m;
  m.im = image<uint32_t>(512, 512);
  if (j3dg_make_matcap(type, m.im.data(), &m.cavity_clr) != 0) throw std::invalid_argument("unknown matcap type");
  return m;
}

}  // namespace

matcap make_matcap_red_wax() { return make_matcap(0); }
matcap make_matcap_gray() { return make_matcap(1); }
matcap make_matcap_brown() { return make_matcap(2); }
matcap make_matcap_sketch() { return make_matcap(3); }

camera make_default_camera() {
  camera c;
  j3dg_default_camera(&c);
  return c;
}

float4x4 make_projection_matrix(const camera& c, int w, int h) {
  float4x4 m;
  j3dg_projection(&c, w, h, m.data());
  return m;
}

float4x4 invert_projection_matrix(const float4x4& m) {
  float4x4 o;
  j3dg_invert_projection(m.data(), o.data());
  return o;
}

float4x4 invert_orthonormal(const float4x4& m) {
  float4x4 o;
  j3dg_invert_orthonormal(m.data(), o.data());
  return o;
}

float4x4 get_identity() {
  float4x4 m{};
  m[0] = m[5] = m[10] = m[15] = 1.f;
  return m;
}

// ------------------------------------------------------------------ device ---

device::device(int gpu) {
  const int rc = j3dg_create(gpu, &ctx_);
  if (rc != J3DG_OK) throw gpu_error(rc, "j3dg_create failed for GPU " + std::to_string(gpu));
}

device::~device() {
  if (ctx_) j3dg_destroy(ctx_);
}

void device::check(int rc, const char* what) const {
  if (rc != J3DG_OK) throw gpu_error(rc, std::string(what) + ": " + j3dg_last_error(ctx_));
}

// ------------------------------------------------------------------- scene ---

scene::~scene() {
  if (!dev) return;
  for (const auto& o : objects) j3dg_remove_mesh(dev->ctx(), o.slot);
  for (const auto& p : pointclouds) j3dg_remove_points(dev->ctx(), p.cloud);
}

void add_pointcloud(uint32_t id, scene& s, const std::vector<float>& vertices, const std::vector<float>* normals,
                    const std::vector<uint32_t>* colors, const float4x4& cs) {
  const size_t n = vertices.size() / 3;
  if (vertices.size() % 3 || (normals && !normals->empty() && normals->size() != vertices.size()) ||
      (colors && !colors->empty() && colors->size() != n))
    throw std::invalid_argument("add_pointcloud: arrays of xyz, normals and colours per point");
  scene_pointcloud pc;
  pc.db_id = id;
  pc.cs = cs;
  pc.device_cs = cs;
  compute_bb(vertices, pc.min_bb, pc.max_bb);  // j3d/scene.cpp:34
  s.dev->check(j3dg_upload_points(s.dev->ctx(), id, vertices.data(), n,
                                  normals && !normals->empty() ? normals->data() : nullptr,
                                  colors && !colors->empty() ? colors->data() : nullptr, cs.data(), &pc.cloud),
               "j3dg_upload_points");
  s.pointclouds.push_back(pc);
}

void remove_pointcloud(uint32_t id, scene& s) {
  auto it = std::find_if(s.pointclouds.begin(), s.pointclouds.end(),
                         [&](const scene_pointcloud& p) { return p.db_id == id; });
  if (it != s.pointclouds.end()) {
    s.dev->check(j3dg_remove_points(s.dev->ctx(), it->cloud), "j3dg_remove_points");
    s.pointclouds.erase(it);
  }
}

// j3d/scene.cpp:8-25
void add_object(uint32_t id, scene& s, const std::vector<float>& vertices, const std::vector<uint32_t>& triangles,
                const float4x4& cs, uint32_t bvh_width, uint32_t builder) {
  if (vertices.size() % 3 || triangles.size() % 3) throw std::invalid_argument("add_object: arrays of xyz / ijk");
  scene_object obj;
  obj.db_id = id;
  obj.p_vertices = &vertices;
  obj.p_triangles = &triangles;
  obj.cs = cs;
  obj.device_cs = cs;
  const size_t nv = vertices.size() / 3;
  compute_bb(vertices, obj.min_bb, obj.max_bb);
  j3dg_build_opts bo{bvh_width, 0, 0, builder};
  s.dev->check(j3dg_upload_mesh(s.dev->ctx(), id, vertices.data(), nv, triangles.data(), triangles.size() / 3,
                                cs.data(), &bo, &obj.slot),
               "j3dg_upload_mesh");
  s.objects.emplace_back(std::move(obj));
}

// j3d/scene.h:16-18, scene.cpp:17-19
void set_object_colors(uint32_t id, scene& s, const std::vector<float>* vertex_colors, const std::vector<float>* uv,
                       const image<uint32_t>* texture) {
  auto it = std::find_if(s.objects.begin(), s.objects.end(), [&](const scene_object& so) { return so.db_id == id; });
  if (it == s.objects.end()) throw std::invalid_argument("set_object_colors: unknown object");
  const size_t nv = it->p_vertices->size() / 3, nt = it->p_triangles->size() / 3;
  const bool has_vc = vertex_colors && !vertex_colors->empty();
  const bool has_uv = uv && !uv->empty();
  if (has_vc && vertex_colors->size() != 3 * nv) throw std::invalid_argument("set_object_colors: 3 floats per vertex");
  if (has_uv && (uv->size() != 6 * nt || !texture || texture->width() == 0 || texture->height() == 0))
    throw std::invalid_argument("set_object_colors: 6 floats per triangle and a texture");
  s.dev->check(j3dg_set_mesh_colors(s.dev->ctx(), it->slot, has_vc ? vertex_colors->data() : nullptr,
                                    has_uv ? uv->data() : nullptr, has_uv ? texture->row(0) : nullptr,
                                    has_uv ? texture->width() : 0, has_uv ? texture->height() : 0),
               "j3dg_set_mesh_colors");
}

// j3d/scene.cpp:41-49: the object and the point cloud with this id
void remove_object(uint32_t id, scene& s) {
  auto it = std::find_if(s.objects.begin(), s.objects.end(), [&](const scene_object& so) { return so.db_id == id; });
  if (it != s.objects.end()) {
    s.dev->check(j3dg_remove_mesh(s.dev->ctx(), it->slot), "j3dg_remove_mesh");
    s.objects.erase(it);
  }
  remove_pointcloud(id, s);
}

// j3d/scene.cpp:50-88: the objects' and then the point clouds' world boxes,
// starting from the first object's (or, without objects, the first cloud's)
void prepare_scene(scene& s) {
  auto box = [](const float4x4& cs, const float* mn, const float* mx, float lo[3], float hi[3]) {
    float a[3], b[3];
    transform_point(cs, mn, a);
    transform_point(cs, mx, b);
    for (int k = 0; k < 3; ++k) lo[k] = std::min(a[k], b[k]), hi[k] = std::max(a[k], b[k]);
  };
  auto grow = [&](const float4x4& cs, const float* mn, const float* mx) {
    float lo[3], hi[3];
    box(cs, mn, mx, lo, hi);
    for (int k = 0; k < 3; ++k) s.min_bb[k] = std::min(s.min_bb[k], lo[k]), s.max_bb[k] = std::max(s.max_bb[k], hi[k]);
  };
  if (!s.objects.empty()) {
    box(s.objects.front().cs, s.objects.front().min_bb, s.objects.front().max_bb, s.min_bb, s.max_bb);
  } else if (!s.pointclouds.empty()) {
    box(s.pointclouds.front().cs, s.pointclouds.front().min_bb, s.pointclouds.front().max_bb, s.min_bb, s.max_bb);
  } else {
    for (int k = 0; k < 3; ++k) s.min_bb[k] = s.max_bb[k] = 0.f;
  }
  for (const auto& o : s.objects) grow(o.cs, o.min_bb, o.max_bb);
  for (const auto& p : s.pointclouds) grow(p.cs, p.min_bb, p.max_bb);
  s.diagonal = s.max_bb[0] - s.min_bb[0];
  s.diagonal = std::max<float>(s.diagonal, s.max_bb[1] - s.min_bb[1]);
  s.diagonal = std::max<float>(s.diagonal, s.max_bb[2] - s.min_bb[2]);
}

// j3d/scene.cpp:90-110 (y axis up)
void unzoom(scene& s) {
  s.coordinate_system = get_identity();
  for (int j = 0; j < 3; ++j) s.pivot[j] = (s.min_bb[j] + s.max_bb[j]) * 0.5f;
  s.coordinate_system[12] = s.pivot[0];
  s.coordinate_system[13] = s.pivot[1];
  s.coordinate_system[14] = s.pivot[2] + s.diagonal * 2.f;
  s.coordinate_system_inv = invert_orthonormal(s.coordinate_system);
}

// ------------------------------------------------------------------ canvas ---

// j3d/canvas.cpp:100-104
canvas::canvas(std::shared_ptr<device> dev) : dev_(std::move(dev)) {}

canvas::canvas(std::shared_ptr<device> dev, uint32_t w, uint32_t h) : dev_(std::move(dev)) { resize(w, h); }

canvas::~canvas() {
  if (target_ != 0xFFFFFFFFu) j3dg_destroy_target(dev_->ctx(), target_);
}

// j3d/canvas.cpp:117-134: fresh pixels (object_id = -1, db_id = 0), default camera
void canvas::resize(uint32_t w, uint32_t h) {
  if (target_ != 0xFFFFFFFFu) {
    j3dg_destroy_target(dev_->ctx(), target_);
    target_ = 0xFFFFFFFFu;
  }
  w_ = w, h_ = h;
  im = image<uint32_t>(w, h);
  _canvas = image<pixel>(w, h);
  for (auto& p : _canvas) {
    std::memset(&p, 0, sizeof(p));
    p.object_id = 0xFFFFFFFFu;
  }
  if (w && h) {
    dev_->check(j3dg_create_target(dev_->ctx(), w, h, &target_), "j3dg_create_target");
    generation_ = next_generation();
    _canvas.set_tag(generation_);  // the new target holds the same default pixels
  }
  pixels_stale_ = false;
  _camera = make_default_camera();
  projection_matrix = make_projection_matrix(_camera, (int)w, (int)h);
  projection_matrix_inv = invert_projection_matrix(projection_matrix);
}

// j3d/canvas.cpp:136-139 (the gradient itself is drawn by the shading kernel)
void canvas::set_background_color(uint32_t clr_top, uint32_t clr_bottom) {
  bg_top_ = clr_top;
  bg_bottom_ = clr_bottom;
}

void canvas::set_camera(const camera& c) {
  _camera = c;
  projection_matrix = make_projection_matrix(_camera, (int)w_, (int)h_);
  projection_matrix_inv = invert_projection_matrix(projection_matrix);
}

void canvas::set_directional_light(bool on, const float dir[3]) {
  directional_ = on;
  if (dir) std::memcpy(light_dir_, dir, sizeof(light_dir_));
}

void canvas::set_learned(bool on, float margin) {
  learned_ = on;
  margin_ = margin;
}

j3dg_frame_desc canvas::frame_desc(const scene* s, int x0, int y0, int x1, int y1) const {
  j3dg_frame_desc d;
  std::memset(&d, 0, sizeof(d));
  d.width = w_, d.height = h_;
  d.x0 = x0, d.y0 = y0, d.x1 = x1, d.y1 = y1;
  std::memcpy(d.proj_inv, projection_matrix_inv.data(), sizeof(d.proj_inv));
  d.near_plane = _camera.near_plane;
  uint32_t f = 0;
  if (_settings.one_bit) f |= J3DG_ONE_BIT;
  if (_settings.shadow) f |= J3DG_SHADOW;
  if (_settings.edges) f |= J3DG_EDGES;
  if (_settings.wireframe) f |= J3DG_WIREFRAME;
  if (_settings.shading) f |= J3DG_SHADING;
  if (_settings.textured) f |= J3DG_TEXTURED;
  if (_settings.vertexcolors) f |= J3DG_VERTEXCOLORS;
  if (directional_) f |= J3DG_DIRECTIONAL_LIGHT;
  if (learned_) f |= J3DG_LEARNED;
  std::memcpy(d.light_dir, light_dir_, sizeof(d.light_dir));
  d.bg_top = bg_top_, d.bg_bottom = bg_bottom_;
  d.margin = margin_;
  if (s) {
    std::memcpy(d.cs, s->coordinate_system.data(), sizeof(d.cs));
    std::memcpy(d.cs_inv, s->coordinate_system_inv.data(), sizeof(d.cs_inv));
    std::memcpy(d.pivot, s->pivot, sizeof(d.pivot));
    d.diagonal = s->diagonal;
    object_slots_.clear();
    for (const auto& o : s->objects) object_slots_.push_back(o.slot);
    d.n_objects = (uint32_t)object_slots_.size();
    if (object_slots_.size() > J3DG_MAX_OBJECTS)
      d.object_list = object_slots_.data();  // valid until the next frame description
    else
      std::copy(object_slots_.begin(), object_slots_.end(), d.objects);
  } else {
    d.flags |= J3DG_NO_SCENE;
  }
  d.flags |= f;
  return d;
}

// Objects whose cs changed since the upload get the new transform.
void canvas::sync_objects(const scene& s) {
  for (const auto& o : s.objects)
    if (o.cs != o.device_cs) {
      dev_->check(j3dg_set_mesh_cs(dev_->ctx(), o.slot, o.cs.data()), "j3dg_set_mesh_cs");
      const_cast<scene_object&>(o).device_cs = o.cs;
    }
}

// j3d/canvas.cpp:677-874 on the rectangle; `out` is brought to the device
// first unless it already holds the device's G-buffer (pixels outside the
// rectangle, and the mark/colour bytes of misses, keep their values).
void canvas::update_canvas(image<pixel>& out, int x0, int y0, int x1, int y1, const scene& s) {
  if (out.width() != w_ || out.height() != h_) throw std::invalid_argument("update_canvas: image size differs");
  sync_objects(s);
  const bool on_device = (&out == &_canvas && pixels_stale_) || (out.tag() != 0 && out.tag() == generation_);
  if (!on_device) {
    if (&out != &_canvas) pull_pixels();  // keep the canvas's own G-buffer before the device copy is replaced
    dev_->check(j3dg_write_target_gbuffer(dev_->ctx(), target_, out.data()), "j3dg_write_target_gbuffer");
  }
  const j3dg_frame_desc d = frame_desc(&s, x0, y0, x1, y1);
  dev_->check(j3dg_render(dev_->ctx(), &d, target_, J3DG_STAGE_TRACE, out.data(), nullptr), "j3dg_render");
  generation_ = next_generation();
  out.set_tag(generation_);
  if (&out == &_canvas) pixels_stale_ = false;
}

void canvas::update_canvas(int x0, int y0, int x1, int y1, const scene& s) { update_canvas(_canvas, x0, y0, x1, y1, s); }

// j3d/canvas.cpp:876-891
void canvas::render_scene(image<pixel>& out, const scene* s) {
  if (out.width() != w_ || out.height() != h_) out = image<pixel>(w_, h_);
  if (&out == &_canvas) {
    // the canvas's own G-buffer stays on the device until someone reads it
    if (!pixels_stale_ && _canvas.tag() != generation_)
      dev_->check(j3dg_write_target_gbuffer(dev_->ctx(), target_, _canvas.data()), "j3dg_write_target_gbuffer");
    if (s) sync_objects(*s);
    const j3dg_frame_desc d = frame_desc(s, 0, 0, (int)w_ - 1, (int)h_ - 1);
    dev_->check(j3dg_render(dev_->ctx(), &d, target_, J3DG_STAGE_TRACE, nullptr, nullptr), "j3dg_render");
    generation_ = next_generation();
    pixels_stale_ = true;
    return;
  }
  if (s) {
    update_canvas(out, 0, 0, (int)w_ - 1, (int)h_ - 1, *s);
  } else {
    for (auto& p : out) {
      p.db_id = 0;
      p.object_id = 0xFFFFFFFFu;
    }
  }
}

// j3d/canvas.cpp:893-898 (the background is drawn by the shading pass)
void canvas::render_scene(const scene* s) { render_scene(_canvas, s); }

void canvas::pull_pixels() const {
  if (!pixels_stale_) return;
  dev_->check(j3dg_read_target(dev_->ctx(), target_, _canvas.data(), nullptr), "j3dg_read_target");
  _canvas.set_tag(generation_);
  pixels_stale_ = false;
}

const image<pixel>& canvas::get_pixels() const {
  pull_pixels();
  return _canvas;
}

// j3d/canvas.cpp:147-152
void canvas::get_pixel(pixel& p, float pos_x, float pos_y, float mouse_offset_x, float mouse_offset_y) {
  pull_pixels();
  const uint32_t X = (uint32_t)(pos_x - mouse_offset_x);
  const uint32_t Y = (uint32_t)(pos_y - mouse_offset_y);
  p = static_cast<const image<pixel>&>(_canvas)(X, Y);
}

void canvas::upload_matcap(const matcap& m) {
  if (matcap_key_ == m.im.data() && matcap_cavity_ == m.cavity_clr && matcap_w_ == m.im.width()) return;
  dev_->check(j3dg_set_matcap(dev_->ctx(), m.im.data(), m.im.width(), m.im.height(), m.cavity_clr), "j3dg_set_matcap");
  matcap_key_ = m.im.data();
  matcap_cavity_ = m.cavity_clr;
  matcap_w_ = m.im.width();
}

// j3d/canvas.cpp:582-670 (edges / plain), 449-579 (wireframe, one-bit).  A
// matcap is identified by its pixel buffer: change its pixels through a new
// matcap object.
void canvas::canvas_to_image(const image<pixel>& pixels, const matcap& _matcap) {
  if (pixels.width() != w_ || pixels.height() != h_) throw std::invalid_argument("canvas_to_image: image size differs");
  upload_matcap(_matcap);
  const bool on_device = (&pixels == &_canvas && pixels_stale_) || (pixels.tag() != 0 && pixels.tag() == generation_);
  if (!on_device) {
    if (&pixels != &_canvas) pull_pixels();
    dev_->check(j3dg_write_target_gbuffer(dev_->ctx(), target_, pixels.data()), "j3dg_write_target_gbuffer");
    generation_ = next_generation();
  }
  const j3dg_frame_desc d = frame_desc(nullptr, 0, 0, (int)w_ - 1, (int)h_ - 1);
  dev_->check(j3dg_render(dev_->ctx(), &d, target_, J3DG_STAGE_SHADE, nullptr, im.data()), "j3dg_render");
}

void canvas::canvas_to_image(const matcap& _matcap) { canvas_to_image(_canvas, _matcap); }

// j3d/canvas.cpp:952-1030.  The device target holds the image canvas_to_image
// made; its G-buffer is set to `pix` (the z-buffer source) unless it already
// is.  Afterwards the canvas's pixels are `pix` with the points drawn: the
// reference writes the drawn pixels into _canvas, and view::render_scene
// passes a copy of _canvas as `pix` (view.cpp:426-428), so the two agree.
void canvas::render_pointclouds_on_image(const scene* s, const image<pixel>& pix) {
  if (!s || s->pointclouds.empty()) return;
  if (pix.width() != w_ || pix.height() != h_)
    throw std::invalid_argument("render_pointclouds_on_image: image size differs");
  const bool on_device = (&pix == &_canvas && pixels_stale_) || (pix.tag() != 0 && pix.tag() == generation_);
  if (!on_device) {
    if (&pix != &_canvas) pull_pixels();
    dev_->check(j3dg_write_target_gbuffer(dev_->ctx(), target_, pix.data()), "j3dg_write_target_gbuffer");
  }
  std::vector<uint32_t> clouds;
  for (auto& p : const_cast<scene*>(s)->pointclouds) {
    if (std::memcmp(p.cs.data(), p.device_cs.data(), sizeof(float) * 16) != 0) {
      dev_->check(j3dg_set_points_cs(dev_->ctx(), p.cloud, p.cs.data()), "j3dg_set_points_cs");
      p.device_cs = p.cs;
    }
    clouds.push_back(p.cloud);
  }
  const j3dg_frame_desc d = frame_desc(nullptr, 0, 0, (int)w_ - 1, (int)h_ - 1);
  dev_->check(j3dg_splat_points(dev_->ctx(), target_, projection_matrix.data(), s->coordinate_system_inv.data(),
                                d.flags, clouds.data(), (uint32_t)clouds.size()),
              "j3dg_splat_points");
  dev_->check(j3dg_read_target(dev_->ctx(), target_, _canvas.data(), im.data()), "j3dg_read_target");
  generation_ = next_generation();
  _canvas.set_tag(generation_);
  pixels_stale_ = false;
}

// j3d/canvas.cpp:900-951: copy into `screen`, flipped vertically; pos_x must be
// a multiple of 4 as in the reference.
void canvas::blit_onto(image<uint32_t>& screen, int32_t pos_x, int32_t pos_y) {
  if (pos_x & 3) throw std::runtime_error("blit_onto: can't use SIMD.");
  uint32_t im_x = 0, im_y = 0;
  if (pos_x < 0) im_x = (uint32_t)-pos_x, pos_x = 0;
  if (pos_y < 0) im_y = (uint32_t)-pos_y, pos_y = 0;
  int32_t szx = (int32_t)im.width() - (int32_t)im_x;
  int32_t szy = (int32_t)im.height() - (int32_t)im_y;
  if (pos_x + szx > (int32_t)screen.width()) szx = (int32_t)screen.width() - pos_x;
  if (pos_y + szy > (int32_t)screen.height()) szy = (int32_t)screen.height() - pos_y;
  if (szx < 0 || szy < 0) return;
  const image<uint32_t>& src = im;
  for (int32_t y = 0; y < szy; ++y) {
    uint32_t* dst = screen.row(screen.height() - 1 - (uint32_t)(y + pos_y)) + pos_x;
    std::memcpy(dst, src.row(im_y + (uint32_t)y) + im_x, (size_t)szx * sizeof(uint32_t));
  }
}

void canvas::last_timings(float* trace_ms, float* shade_ms, float* mlp_ms, float* shadow_ms) const {
  j3dg_stats st;
  dev_->check(j3dg_get_stats(dev_->ctx(), &st), "j3dg_get_stats");
  if (trace_ms) *trace_ms = st.trace_ms;
  if (shade_ms) *shade_ms = st.shade_ms;
  if (mlp_ms) *mlp_ms = st.mlp_ms;
  if (shadow_ms) *shadow_ms = st.shadow_ms;
}

}  // namespace j3d_amd
