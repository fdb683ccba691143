// host_tables.cpp -- camera projection and matcap tables computed on the host
// (j3d/camera.cpp:5-73, j3d/matcap.cpp:9-268).  Compiled with -ffp-contract=off.
#include <cfloat>
#include <cmath>
#include <cstring>

#include "j3dg_host.h"

namespace {

inline uint32_t pack_rgb(unsigned char r, unsigned char g, unsigned char b) {
  return 0xff000000u | ((uint32_t)b << 16) | ((uint32_t)g << 8) | (uint32_t)r;
}

struct V3 {
  float x, y, z;
};

inline V3 normalize(V3 a) {
  float l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  return V3{a.x / l, a.y / l, a.z / l};
}

inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// The gray, brown and red-wax tables share one lighting model and differ in
// their coefficients (j3d/matcap.cpp:9-239).
struct LitMatcap {
  float base[3];
  float c1[3], c2[3], c3[3], c4[3];
  uint32_t cavity;
};

void lit_matcap(const LitMatcap& p, uint32_t* im) {
  const uint32_t w = 512, h = 512;
  const V3 n1 = normalize(V3{0.f, 0.8f, 1.f});
  const V3 n2 = normalize(V3{0.f, 0.4f, 1.f});
  const V3 n3 = normalize(V3{0.f, 0.f, 1.f});
  for (uint32_t y = 0; y < h; ++y) {
    uint32_t* row = im + (size_t)(h - y - 1) * w;  // rows written bottom-up
    for (uint32_t x = 0; x < w; ++x) {
      const float u = (float)x / (float)(w - 1) * 2.f - 1.f;
      const float v = (float)y / (float)(h - 1) * 2.f - 1.f;
      if (!(u * u + v * v <= 1.01f)) {
        row[x] = 0xff000000u;
        continue;
      }
      const V3 sp{u, v, std::sqrt(1.01f - u * u - v * v)};
      const float a1 = dot(sp, n1);
      const float a2 = powf(dot(sp, n2), 3.f);
      const float a3 = powf(dot(sp, n3), 5.f);
      const float a4 = powf(dot(sp, n3), 50.f);
      unsigned char ch[3];
      for (int k = 0; k < 3; ++k) {
        float c = p.base[k] + (p.c1[k] * a1 + p.c2[k] * a2 + p.c3[k] * a3 + p.c4[k] * a4) / 1.f;
        c = c > 255.f ? 255.f : c;
        c = c < 0.f ? 0.f : c;
        ch[k] = (unsigned char)c;
      }
      row[x] = pack_rgb(ch[0], ch[1], ch[2]);
    }
  }
}

}  // namespace

extern "C" {

void j3dg_default_camera(j3dg_camera* c) {
  c->focal_length = 35.f;
  c->film_w_inch = 1.024f;
  c->film_h_inch = 0.768f;
  c->near_plane = 0.1f;
  c->far_plane = FLT_MAX;
  c->fit_film = 1;
  c->zoom = 1.f;
}

void j3dg_frustum(float left, float right, float bottom, float top, float n, float f, float* out) {
  std::memset(out, 0, 16 * sizeof(float));
  out[0] = 2.f * n / (right - left);
  out[5] = -2.f * n / (top - bottom);
  out[8] = (right + left) / (right - left);
  out[9] = -(top + bottom) / (top - bottom);
  out[10] = -(f + n) / (f - n);
  out[11] = -1.f;
  out[14] = -(2.f * f * n) / (f - n);
}

void j3dg_invert_projection(const float* in, float* out) {
  std::memset(out, 0, 16 * sizeof(float));
  out[0] = 1.f / in[0];
  out[5] = 1.f / in[5];
  out[11] = 1.f / in[14];
  out[12] = in[8] / in[0];
  out[13] = in[9] / in[5];
  out[14] = -1.f;
  out[15] = in[10] / in[14];
}

void j3dg_projection(const j3dg_camera* c, int screen_w, int screen_h, float* out) {
  const float inch_to_mm = 25.4f;
  float top = ((c->film_h_inch * inch_to_mm / 2.f) / c->focal_length) * c->near_plane;
  float right = ((c->film_w_inch * inch_to_mm / 2.f) / c->focal_length) * c->near_plane;
  float xscale = c->zoom, yscale = c->zoom;
  const float device_aspect = (float)screen_w / (float)screen_h;
  const float film_aspect = c->film_w_inch / c->film_h_inch;
  if (c->fit_film == 1) {  // Overscan
    if (film_aspect > device_aspect)
      yscale *= film_aspect / device_aspect;
    else
      xscale *= device_aspect / film_aspect;
  } else {  // Fill
    if (film_aspect > device_aspect)
      xscale *= device_aspect / film_aspect;
    else
      yscale *= film_aspect / device_aspect;
  }
  right *= xscale;
  top *= yscale;
  j3dg_frustum(-right, right, -top, top, c->near_plane, c->far_plane, out);
}

void j3dg_invert_orthonormal(const float* m, float* o) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o[4 * j + i] = m[4 * i + j];
  o[3] = o[7] = o[11] = 0.f;
  for (int i = 0; i < 3; ++i) o[12 + i] = -(m[4 * i] * m[12] + m[4 * i + 1] * m[13] + m[4 * i + 2] * m[14]);
  o[15] = 1.f;
}

int j3dg_make_matcap(int type, uint32_t* im, uint32_t* cavity) {
  if (!im || !cavity) return 1;
  if (type == 3) {  // sketch
    *cavity = 0xFF505050u;
    const float thres = 0.4f;
    for (uint32_t y = 0; y < 512; ++y) {
      uint32_t* row = im + (size_t)(511 - y) * 512;
      for (uint32_t x = 0; x < 512; ++x) {
        const float u = (float)x / 511.f * 2.f - 1.f;
        const float v = (float)y / 511.f * 2.f - 1.f;
        const float val = std::fabs(1.f - u * u - v * v);
        if (val < thres) {
          const uint32_t s = (uint32_t)(val / thres * (float)0xe1);
          row[x] = 0xff000000u | (s << 16) | (s << 8) | s;
        } else {
          row[x] = 0xffe1e1e1u;
        }
      }
    }
    return 0;
  }
  LitMatcap p;
  const float c3[3] = {50.f, 50.f, 50.f}, c4[3] = {30.f, 30.f, 30.f};
  std::memcpy(p.c3, c3, sizeof(c3));
  std::memcpy(p.c4, c4, sizeof(c4));
  if (type == 1) {  // gray
    const float b[3] = {32.f, 32.f, 32.f}, c1[3] = {200.f / 4.f, 200.f / 4.f, 200.f / 4.f}, c2[3] = {50.f, 50.f, 50.f};
    std::memcpy(p.base, b, sizeof(b));
    std::memcpy(p.c1, c1, sizeof(c1));
    std::memcpy(p.c2, c2, sizeof(c2));
    p.cavity = 0xff505050u;
  } else if (type == 2) {  // brown
    const float b[3] = {32.f, 20.f, 10.f}, c1[3] = {200.f / 4.f, 180.f / 4.f, 160.f / 4.f}, c2[3] = {50.f, 40.f, 30.f};
    std::memcpy(p.base, b, sizeof(b));
    std::memcpy(p.c1, c1, sizeof(c1));
    std::memcpy(p.c2, c2, sizeof(c2));
    p.cavity = 0xff405060u;
  } else if (type == 0) {  // red wax: no base term on green and blue
    const float b[3] = {32.f, 0.f, 0.f}, c1[3] = {200.f / 1.5f, 200.f / 4.f, 150.f / 4.f}, c2[3] = {30.f, 25.f, 20.f};
    std::memcpy(p.base, b, sizeof(b));
    std::memcpy(p.c1, c1, sizeof(c1));
    std::memcpy(p.c2, c2, sizeof(c2));
    p.cavity = 0xFF7D7DFFu;
  } else {
    return 1;
  }
  lit_matcap(p, im);
  *cavity = p.cavity;
  return 0;
}

}  // extern "C"
