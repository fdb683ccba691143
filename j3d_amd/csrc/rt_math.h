// rt_math.h -- arithmetic of the ray-casting path, shared by the HIP kernels
// and the host side of the library (normals, matcap tables, projection).
//
// Every function rounds in the order written: the library is compiled with
// -ffp-contract=off, so no multiply-add is fused and the device results equal
// the host results bit for bit (division and sqrt are correctly rounded on
// both sides).  Matrices are column-major float[16] with the translation in
// m[12..14], the layout of jtk::float4x4 used by j3d/canvas.cpp.
#pragma once
#include <stdint.h>

#include <cfloat>
#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define J3D_HD __host__ __device__ inline
#else
#define J3D_HD inline
#endif

namespace j3d {

struct f4 {
  float x, y, z, w;
};

// matrix_vector_multiply (j3d/canvas.cpp:774-776, 791-792): four terms, left to right.
J3D_HD f4 mat_vec(const float* m, f4 v) {
  f4 r;
  r.x = m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w;
  r.y = m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w;
  r.z = m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w;
  r.w = m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * v.w;
  return r;
}

J3D_HD float dot4(f4 a, f4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }

J3D_HD uint32_t make_color(uint32_t r, uint32_t g, uint32_t b) { return 0xff000000u | (b << 16) | (g << 8) | r; }

// The byte a colour channel becomes, (uint8_t)(c * 255.f) of
// j3d/canvas.cpp:829-831 and j3d/vox.cpp:374: truncation toward zero to int32
// (NaN and values outside the int32 range give 0), then the low 8 bits.
J3D_HD uint32_t color_byte(float c) {
  const float x = c * 255.f;
  if (!(fabsf(x) < 2147483648.f)) return 0u;
  return (uint32_t)(int32_t)x & 0xFFu;
}

// Point splats (canvas::render_pointclouds_on_image, j3d/canvas.cpp:952-1030).
// The reference binds the scene's CS^-1, the cloud's cs and the canvas
// projection P to jtk's point renderer (absent from the reference tree) and
// keeps a z-buffer of 1/depth.  Defined here: a point p is drawn into one
// pixel, clip = P (CS^-1 (cs (p, 1))), each product four terms left to right;
// it is visible when 0 < clip.w <= FLT_MAX and |x|, |y| <= w (NaN fails).  There is no
// near/far test on clip z: with the reference's default far plane (FLT_MAX,
// j3d/camera.cpp:12) the frustum's z row holds -inf (2 * far overflows,
// camera.cpp:18-24), so such a test would drop every point; depth is judged
// by 1/w alone.  Its pixel is floor((x/w + 1) * 0.5 * W) and
// floor((y/w + 1) * 0.5 * H), rejected when equal to W or H, so that a point on
// the ray-cast surface of pixel (x, y) (canvas.cpp:773) lands in (x, y); its
// z-buffer value is 1/w.
J3D_HD bool splat_project(const float* cs, const float* cs_inv, const float* proj, float px, float py, float pz,
                          uint32_t w, uint32_t h, uint32_t& pixel, float& inv_w) {
  const f4 c = mat_vec(proj, mat_vec(cs_inv, mat_vec(cs, f4{px, py, pz, 1.f})));
  if (!(c.w > 0.f) || !(c.w <= FLT_MAX) || !(fabsf(c.x) <= c.w) || !(fabsf(c.y) <= c.w)) return false;
  const float fx = (c.x / c.w + 1.f) * 0.5f * (float)w;
  const float fy = (c.y / c.w + 1.f) * 0.5f * (float)h;
  const uint32_t x = (uint32_t)fx, y = (uint32_t)fy;  // fx, fy in [0, W], [0, H]
  if (x >= w || y >= h) return false;
  pixel = y * w + x;
  inv_w = 1.f / c.w;
  return true;
}

// The colour a drawn point writes: its own colour (0xAABBGGRR), or 0xffffffff
// (the colour argument of the reference's present call, canvas.cpp:997) when
// the cloud has none or the canvas is one-bit; with shading and normals the
// three colour channels are scaled by l = n'.z / |n'|, n' = CS^-1 (cs (n, 0)),
// clamped to [0, 1] (NaN -> 0), each truncated to an integer.
J3D_HD uint32_t splat_color(uint32_t color, bool shade, const float* cs, const float* cs_inv, float nx, float ny,
                            float nz) {
  if (!shade) return color;
  const f4 n = mat_vec(cs_inv, mat_vec(cs, f4{nx, ny, nz, 0.f}));
  float l = n.z / sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
  if (!(l > 0.f)) l = 0.f;
  if (l > 1.f) l = 1.f;
  const uint32_t r = (uint32_t)((float)(color & 0xFFu) * l);
  const uint32_t g = (uint32_t)((float)((color >> 8) & 0xFFu) * l);
  const uint32_t b = (uint32_t)((float)((color >> 16) & 0xFFu) * l);
  return (color & 0xff000000u) | (b << 16) | (g << 8) | r;
}

// Moller-Trumbore ray/triangle test on (v0, e1 = v1-v0, e2 = v2-v0): two-sided,
// det == 0 rejects, barycentric and [t_near, t_far] tests inclusive, NaN rejects.
// This defines the test behind jtk's find_closest_triangle (j3d/canvas.cpp:786).
J3D_HD bool intersect_triangle(float ox, float oy, float oz, float dx, float dy, float dz, float v0x, float v0y,
                               float v0z, float e1x, float e1y, float e1z, float e2x, float e2y, float e2z,
                               float t_near, float t_far, float& t, float& u, float& v) {
  float px = dy * e2z - dz * e2y;
  float py = dz * e2x - dx * e2z;
  float pz = dx * e2y - dy * e2x;
  float det = e1x * px + e1y * py + e1z * pz;
  if (det == 0.f) return false;
  float inv = 1.f / det;
  float sx = ox - v0x, sy = oy - v0y, sz = oz - v0z;
  u = (sx * px + sy * py + sz * pz) * inv;
  if (!(u >= 0.f && u <= 1.f)) return false;
  float qx = sy * e1z - sz * e1y;
  float qy = sz * e1x - sx * e1z;
  float qz = sx * e1y - sy * e1x;
  v = (dx * qx + dy * qy + dz * qz) * inv;
  if (!(v >= 0.f && u + v <= 1.f)) return false;
  t = (e2x * qx + e2y * qy + e2z * qz) * inv;
  if (!(t >= t_near && t <= t_far)) return false;
  return true;
}

// Per-triangle normal: cross(v1-v0, v2-v0) / |cross| (jtk::compute_triangle_normals, j3d/scene.cpp:20).
J3D_HD void triangle_normal(const float* a, const float* b, const float* c, float* n) {
  float e1x = b[0] - a[0], e1y = b[1] - a[1], e1z = b[2] - a[2];
  float e2x = c[0] - a[0], e2y = c[1] - a[1], e2z = c[2] - a[2];
  float x = e1y * e2z - e1z * e2y;
  float y = e1z * e2x - e1x * e2z;
  float z = e1x * e2y - e1y * e2x;
  float len = sqrtf(x * x + y * y + z * z);
  n[0] = x / len;
  n[1] = y / len;
  n[2] = z / len;
}

// ---- shading (j3d/canvas.cpp:287-446) -------------------------------------

struct MatcapView {
  const uint32_t* im;
  uint32_t w, h;
  uint32_t cavity;
};

J3D_HD uint32_t matcap_U(float u, uint32_t w) { return (uint32_t)floorf(0.5f + (u + 1.f) * (float)(w - 1) * 0.5f); }
J3D_HD uint32_t matcap_V(float v, uint32_t h) { return (uint32_t)floorf(0.5f + (-v + 1.f) * (float)(h - 1) * 0.5f); }

// get_color (canvas.cpp:330-344); the texel index is clamped to the table.
J3D_HD uint32_t matcap_color(const MatcapView& m, uint32_t U, uint32_t V, uint32_t shadow) {
  U = U < m.w ? U : m.w - 1;
  V = V < m.h ? V : m.h - 1;
  uint32_t clr = m.im[(size_t)V * m.w + U];
  if (shadow) clr = make_color((clr & 0xff) >> 2, ((clr >> 8) & 0xff) >> 2, ((clr >> 16) & 0xff) >> 2);
  return clr;
}

// get_angle_color (canvas.cpp:346-389).  acos and the scale expression run in
// double and are rounded to float once.
J3D_HD uint32_t angle_color(float angle, const MatcapView& m, float u, float v, uint32_t mark) {
  uint32_t clr = matcap_color(m, matcap_U(u, m.w), matcap_V(v, m.h), mark);
  if (fabsf(angle) <= 1.f) {
    uint32_t r = clr & 0xff, g = (clr >> 8) & 0xff, b = (clr >> 16) & 0xff;
    const double hp = (double)1.57079632679489f;
    float scale = (float)((hp - fabs(acos((double)angle) - hp)) / hp);
    scale = sqrtf(1.f - scale);
    uint32_t c2 = m.cavity;
    uint32_t r2 = c2 & 0xff, g2 = (c2 >> 8) & 0xff, b2 = (c2 >> 16) & 0xff;
    float k = angle > 0.f ? 1.2f : 0.5f;  // concave : convex
    r = (uint32_t)((float)r * (1.f - scale) + (float)r2 * scale * k);
    g = (uint32_t)((float)g * (1.f - scale) + (float)g2 * scale * k);
    b = (uint32_t)((float)b * (1.f - scale) + (float)b2 * scale * k);
    r = r > 255u ? 255u : r;
    g = g > 255u ? 255u : g;
    b = b > 255u ? 255u : b;
    clr = make_color(r, g, b);
  }
  return clr;
}

// canvas::_get_color (canvas.cpp:410-441).  mark, r, g, b are the G-buffer fields.
J3D_HD uint32_t plain_color(uint32_t mark, uint32_t pr, uint32_t pg, uint32_t pb, const MatcapView& m, bool shading,
                            float u, float v) {
  if (mark & 2) {
    if (shading) {
      float nz = sqrtf(1.f - u * u - v * v);
      float d = u * 0.f + v * 0.f + nz * 1.f + 0.f * 0.f;
      float occ = (mark & 1) ? 0.3f : 1.f;
      float dif = (d < 0.f ? 0.f : (d > 1.f ? 1.f : d)) * occ;
      return 0xff000000u | ((uint32_t)((float)pb * dif) << 16) | ((uint32_t)((float)pg * dif) << 8) |
             (uint32_t)((float)pr * dif);
    }
    if (mark & 1) return make_color(pr >> 2, pg >> 2, pb >> 2);
    return make_color(pr, pg, pb);
  }
  return matcap_color(m, matcap_U(u, m.w), matcap_V(v, m.h), mark);
}

// canvas::compute_convex_cos_angle (canvas.cpp:287-315).
J3D_HD float convex_cos_angle(const float* Pinv, float near_plane, float W, float H, float x1, float y1, float u1,
                              float v1, float depth1, float x2, float y2, float u2, float v2, float depth2) {
  f4 s1 = {2.f * ((x1 + 0.5f) / W) - 1.f, 2.f * ((y1 + 0.5f) / H) - 1.f, near_plane, 1.f};
  f4 d1 = mat_vec(Pinv, s1);
  d1.w = 0.f;
  f4 p1 = {depth1 * d1.x, depth1 * d1.y, depth1 * d1.z, depth1 * d1.w};
  f4 s2 = {2.f * ((x2 + 0.5f) / W) - 1.f, 2.f * ((y2 + 0.5f) / H) - 1.f, near_plane, 1.f};
  f4 d2 = mat_vec(Pinv, s2);
  d2.w = 0.f;
  f4 p2 = {depth2 * d2.x, depth2 * d2.y, depth2 * d2.z, depth2 * d2.w};
  f4 n1 = {u1, v1, sqrtf(1.f - u1 * u1 - v1 * v1), 0.f};
  f4 n2 = {u2, v2, sqrtf(1.f - u2 * u2 - v2 * v2), 0.f};
  if (fabs((double)(dot4(n1, n2) - 1.f)) > 0.0001) {
    f4 pt = {p2.x - p1.x, p2.y - p1.y, p2.z - p1.z, p2.w - p1.w};
    float len = sqrtf(dot4(pt, pt));
    pt.x = pt.x / len;
    pt.y = pt.y / len;
    pt.z = pt.z / len;
    pt.w = pt.w / len;
    return dot4(pt, n1);
  }
  return 0.f;
}

// fill_background row colour (canvas.cpp:55-78).
J3D_HD uint32_t background_color(uint32_t y, uint32_t h, uint32_t top, uint32_t bottom) {
  float scale = (float)y / (float)h;
  uint32_t r = (uint32_t)(scale * (float)(bottom & 0xff) + (1.f - scale) * (float)(top & 0xff));
  uint32_t g = (uint32_t)(scale * (float)((bottom >> 8) & 0xff) + (1.f - scale) * (float)((top >> 8) & 0xff));
  uint32_t b = (uint32_t)(scale * (float)((bottom >> 16) & 0xff) + (1.f - scale) * (float)((top >> 16) & 0xff));
  return make_color((unsigned char)r, (unsigned char)g, (unsigned char)b);
}

}  // namespace j3d
