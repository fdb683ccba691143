/*
 * oracle.cpp -- CPU restatement of j3d's per-frame ray-casting path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never part of it.
 * Built with -ffp-contract=off and without fast-math so every float operation
 * rounds once, in the order written here.  See oracle.h for what is restated.
 */
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

/* ---------------------------------------------------------------- math --- */

struct f4 {
  float x, y, z, w;
};

/* matrix_vector_multiply: column-major, four terms summed left to right. */
inline f4 mv(const float* m, f4 v) {
  f4 r;
  r.x = m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w;
  r.y = m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w;
  r.z = m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w;
  r.w = m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * v.w;
  return r;
}

inline float dot3(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline float dot4(f4 a, f4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }

inline void cross3(const float* a, const float* b, float* r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

inline uint32_t make_color(uint32_t r, uint32_t g, uint32_t b) {
  return 0xff000000u | (b << 16) | (g << 8) | r;
}

template <class F>
void parallel_rows(uint32_t y0, uint32_t y1, int threads, F fn) {
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  if (threads == 1 || y1 - y0 <= 1) {
    for (uint32_t y = y0; y < y1; ++y) fn(y);
    return;
  }
  /* one task per row pulled from a shared counter, as pooled_parallel_for does
     over image rows (j3d/canvas.cpp:765) */
  std::atomic<uint32_t> next(y0);
  std::vector<std::thread> pool;
  for (int i = 0; i < threads; ++i)
    pool.emplace_back([&]() {
      for (;;) {
        uint32_t y = next.fetch_add(1);
        if (y >= y1) break;
        fn(y);
      }
    });
  for (auto& th : pool) th.join();
}

/* --------------------------------------------------- ray / triangle ------ */

/* Moller-Trumbore on (v0, e1, e2).  Two-sided; det == 0 rejects; the
   barycentric and [t_near, t_far] tests are inclusive; NaN rejects. */
inline bool mt_test(const float* o, const float* d, const float* v0, const float* e1, const float* e2,
                    float t_near, float t_far, float& t, float& u, float& v) {
  float p[3];
  cross3(d, e2, p);
  float det = dot3(e1, p);
  if (det == 0.f) return false;
  float inv = 1.f / det;
  float s[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
  u = dot3(s, p) * inv;
  if (!(u >= 0.f && u <= 1.f)) return false;
  float q[3];
  cross3(s, e1, q);
  v = dot3(d, q) * inv;
  if (!(v >= 0.f && u + v <= 1.f)) return false;
  t = dot3(e2, q) * inv;
  if (!(t >= t_near && t <= t_far)) return false;
  return true;
}

/* --------------------------------------------------------------- BVH ----- */

/* Binary BVH over triangle bounds, binned SAH, leaves of <= 4 triangles.
   Only its culling speed depends on it: results are the lexicographic
   minimum over every accepted triangle, and culling is conservative
   (boxes are widened by 2^-20 of their magnitude and the t comparisons
   carry a 2^-16 relative slack). */
struct BNode {
  float lo[3], hi[3];
  uint32_t a, b; /* internal: children a, b ; leaf: start a, count b | 0x80000000 */
};

const uint32_t LEAF_BIT = 0x80000000u;

}  // namespace

struct orc_bvh {
  std::vector<BNode> nodes;
  std::vector<uint32_t> prim;
  /* per primitive in prim order: v0, e1, e2 */
  std::vector<float> tri;
};

namespace {

struct BuildCtx {
  const float* cen;
  const float* blo;
  const float* bhi;
  uint32_t* idx;
  BNode* nodes;
  std::atomic<uint32_t> n_nodes;
  std::atomic<int> spare_threads;
};

void node_bounds(BuildCtx& c, uint32_t begin, uint32_t end, float* lo, float* hi) {
  lo[0] = lo[1] = lo[2] = FLT_MAX;
  hi[0] = hi[1] = hi[2] = -FLT_MAX;
  for (uint32_t i = begin; i < end; ++i) {
    uint32_t p = c.idx[i];
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], c.blo[3 * (size_t)p + k]);
      hi[k] = std::max(hi[k], c.bhi[3 * (size_t)p + k]);
    }
  }
}

void build_rec(BuildCtx& c, uint32_t node, uint32_t begin, uint32_t end, int depth) {
  BNode& nd = c.nodes[node];
  node_bounds(c, begin, end, nd.lo, nd.hi);
  uint32_t n = end - begin;
  if (n <= 4 || depth > 60) {
    nd.a = begin;
    nd.b = n | LEAF_BIT;
    return;
  }
  float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (uint32_t i = begin; i < end; ++i) {
    uint32_t p = c.idx[i];
    for (int k = 0; k < 3; ++k) {
      clo[k] = std::min(clo[k], c.cen[3 * (size_t)p + k]);
      chi[k] = std::max(chi[k], c.cen[3 * (size_t)p + k]);
    }
  }
  int axis = 0;
  for (int k = 1; k < 3; ++k)
    if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
  float ext = chi[axis] - clo[axis];
  uint32_t mid = begin + n / 2;
  bool median = false;
  if (ext > 0.f) {
    const int NB = 16;
    uint32_t cnt[NB] = {0};
    float bl[NB][3], bh[NB][3];
    for (int b = 0; b < NB; ++b)
      for (int k = 0; k < 3; ++k) bl[b][k] = FLT_MAX, bh[b][k] = -FLT_MAX;
    float scale = NB / ext;
    auto bin_of = [&](uint32_t p) {
      int b = (int)((c.cen[3 * (size_t)p + axis] - clo[axis]) * scale);
      return std::min(std::max(b, 0), NB - 1);
    };
    for (uint32_t i = begin; i < end; ++i) {
      uint32_t p = c.idx[i];
      int b = bin_of(p);
      cnt[b]++;
      for (int k = 0; k < 3; ++k) {
        bl[b][k] = std::min(bl[b][k], c.blo[3 * (size_t)p + k]);
        bh[b][k] = std::max(bh[b][k], c.bhi[3 * (size_t)p + k]);
      }
    }
    auto area = [](const float* l, const float* h) {
      float dx = h[0] - l[0], dy = h[1] - l[1], dz = h[2] - l[2];
      if (dx < 0) return 0.f;
      return dx * dy + dy * dz + dz * dx;
    };
    float left_area[NB];
    uint32_t left_cnt[NB];
    float l[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, h[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    uint32_t acc = 0;
    for (int b = 0; b < NB - 1; ++b) {
      for (int k = 0; k < 3; ++k) l[k] = std::min(l[k], bl[b][k]), h[k] = std::max(h[k], bh[b][k]);
      acc += cnt[b];
      left_area[b] = area(l, h);
      left_cnt[b] = acc;
    }
    float best = FLT_MAX;
    int best_b = -1;
    for (int k = 0; k < 3; ++k) l[k] = FLT_MAX, h[k] = -FLT_MAX;
    acc = 0;
    for (int b = NB - 1; b > 0; --b) {
      for (int k = 0; k < 3; ++k) l[k] = std::min(l[k], bl[b][k]), h[k] = std::max(h[k], bh[b][k]);
      acc += cnt[b];
      float cost = left_area[b - 1] * left_cnt[b - 1] + area(l, h) * acc;
      if (left_cnt[b - 1] > 0 && acc > 0 && cost < best) {
        best = cost;
        best_b = b;
      }
    }
    if (best_b < 0) {
      median = true;
    } else {
      uint32_t* m = std::partition(c.idx + begin, c.idx + end, [&](uint32_t p) { return bin_of(p) < best_b; });
      mid = (uint32_t)(m - c.idx);
      if (mid == begin || mid == end) median = true;
    }
  }
  if (median) {
    mid = begin + n / 2;
    std::nth_element(c.idx + begin, c.idx + mid, c.idx + end,
                     [&](uint32_t a, uint32_t b) { return c.cen[3 * (size_t)a + axis] < c.cen[3 * (size_t)b + axis]; });
  }
  uint32_t left = c.n_nodes.fetch_add(2);
  nd.a = left;
  nd.b = left + 1;
  if (n > (1u << 16) && c.spare_threads.fetch_sub(1) > 0) {
    std::thread th([&c, left, begin, mid, depth]() { build_rec(c, left, begin, mid, depth + 1); });
    build_rec(c, left + 1, mid, end, depth + 1);
    th.join();
  } else {
    build_rec(c, left, begin, mid, depth + 1);
    build_rec(c, left + 1, mid, end, depth + 1);
  }
}

const float BOX_WIDEN = 1.0f / 1048576.0f; /* 2^-20 */
const float T_SLACK = 1.0f / 65536.0f;     /* 2^-16 */

struct Ray {
  float o[3], d[3], inv[3];
  float t_near, t_far;
};

inline void set_inv(Ray& r) {
  for (int k = 0; k < 3; ++k) {
    float d = r.d[k];
    if (std::fabs(d) < 1e-30f) d = std::copysign(1e-30f, d);
    r.inv[k] = 1.f / d;
  }
}

/* Conservative slab test: returns entry distance, or +inf on a cull. */
inline float box_entry(const BNode& nd, const Ray& r, float t_hi) {
  float t0x = (nd.lo[0] - r.o[0]) * r.inv[0], t1x = (nd.hi[0] - r.o[0]) * r.inv[0];
  float t0y = (nd.lo[1] - r.o[1]) * r.inv[1], t1y = (nd.hi[1] - r.o[1]) * r.inv[1];
  float t0z = (nd.lo[2] - r.o[2]) * r.inv[2], t1z = (nd.hi[2] - r.o[2]) * r.inv[2];
  float en = std::max(std::max(std::min(t0x, t1x), std::min(t0y, t1y)), std::min(t0z, t1z));
  float ex = std::min(std::min(std::max(t0x, t1x), std::max(t0y, t1y)), std::max(t0z, t1z));
  ex = ex + std::fabs(ex) * T_SLACK;
  float lo = r.t_near - std::fabs(r.t_near) * T_SLACK;
  float hi = t_hi + std::fabs(t_hi) * T_SLACK;
  float a = std::max(en, lo), b = std::min(ex, hi);
  return a <= b ? a : INFINITY;
}

struct Best {
  bool found = false;
  float t = 0, u = 0, v = 0;
  uint32_t obj = 0xFFFFFFFFu, tri = 0xFFFFFFFFu;
  float t2 = FLT_MAX; /* second-nearest accepted t (tie tracking) */
};

inline bool better(float t, uint32_t obj, uint32_t tri, const Best& b) {
  if (!b.found) return true;
  if (t != b.t) return t < b.t;
  if (obj != b.obj) return obj < b.obj;
  return tri < b.tri;
}

/* Closest hit inside one object's BVH, ray already in object space.  With
   track2, the culling bound is the second-nearest distance so that t2 is exact. */
void trace_object(const orc_object& ob, uint32_t oi, const Ray& r, Best& best, bool track2, bool any,
                  bool& any_found) {
  auto consider = [&](uint32_t prim_id, const float* v0, const float* e1, const float* e2) {
    float hi = r.t_far;
    if (best.found) hi = std::min(hi, track2 ? best.t2 : best.t);
    float t, u, v;
    if (!mt_test(r.o, r.d, v0, e1, e2, r.t_near, hi, t, u, v)) return;
    if (any) {
      any_found = true;
      return;
    }
    if (better(t, oi, prim_id, best)) {
      if (best.found) best.t2 = std::min(best.t2, best.t);
      best.found = true;
      best.t = t;
      best.u = u;
      best.v = v;
      best.obj = oi;
      best.tri = prim_id;
    } else {
      best.t2 = std::min(best.t2, t);
    }
  };
  if (!ob.bvh) {
    for (uint64_t i = 0; i < ob.n_triangles; ++i) {
      const uint32_t* tr = ob.triangles + 3 * i;
      const float* a = ob.vertices + 3 * (size_t)tr[0];
      const float* b = ob.vertices + 3 * (size_t)tr[1];
      const float* c = ob.vertices + 3 * (size_t)tr[2];
      float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
      float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
      consider((uint32_t)i, a, e1, e2);
      if (any && any_found) return;
    }
    return;
  }
  const orc_bvh& bvh = *ob.bvh;
  uint32_t stack[256];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    uint32_t ni = stack[--sp];
    const BNode& nd = bvh.nodes[ni];
    float hi = r.t_far;
    if (best.found) hi = std::min(hi, track2 ? best.t2 : best.t);
    if (box_entry(nd, r, hi) == INFINITY) continue;
    if (nd.b & LEAF_BIT) {
      uint32_t cnt = nd.b & ~LEAF_BIT;
      for (uint32_t k = 0; k < cnt; ++k) {
        uint32_t slot = nd.a + k;
        const float* tr = &bvh.tri[9 * (size_t)slot];
        consider(bvh.prim[slot], tr, tr + 3, tr + 6);
        if (any && any_found) return;
      }
    } else {
      float ea = box_entry(bvh.nodes[nd.a], r, hi);
      float eb = box_entry(bvh.nodes[nd.b], r, hi);
      /* push the farther child first so the nearer one is popped next */
      if (ea <= eb) {
        if (eb != INFINITY) stack[sp++] = nd.b;
        if (ea != INFINITY) stack[sp++] = nd.a;
      } else {
        if (ea != INFINITY) stack[sp++] = nd.a;
        if (eb != INFINITY) stack[sp++] = nd.b;
      }
    }
  }
}

inline Ray to_object(const orc_object& ob, const float* inv_cs, const f4& o, const f4& d, float tn, float tf) {
  (void)ob;
  f4 oo = mv(inv_cs, o);
  f4 dd = mv(inv_cs, d);
  Ray r;
  r.o[0] = oo.x, r.o[1] = oo.y, r.o[2] = oo.z;
  r.d[0] = dd.x, r.d[1] = dd.y, r.d[2] = dd.z;
  r.t_near = tn;
  r.t_far = tf;
  set_inv(r);
  return r;
}

/* Closest hit over all objects (two-level: each object is traversed in its
   own space, j3d/canvas.cpp:762,786). */
Best closest(const orc_object* objs, uint32_t nobj, const float* inv_cs, f4 o, f4 d, float tn, float tf,
             bool track2) {
  Best best;
  bool dummy = false;
  for (uint32_t i = 0; i < nobj; ++i) {
    Ray r = to_object(objs[i], inv_cs + 16 * i, o, d, tn, tf);
    trace_object(objs[i], i, r, best, track2, false, dummy);
  }
  return best;
}

bool any_hit(const orc_object* objs, uint32_t nobj, const float* inv_cs, f4 o, f4 d, float tn, float tf) {
  Best best;
  for (uint32_t i = 0; i < nobj; ++i) {
    Ray r = to_object(objs[i], inv_cs + 16 * i, o, d, tn, tf);
    bool found = false;
    trace_object(objs[i], i, r, best, false, true, found);
    if (found) return true;
  }
  return false;
}

/* ------------------------------------------------------------ shading ---- */

struct Matcap {
  const uint32_t* im;
  uint32_t w, h;
  uint32_t cavity;
};

/* canvas.cpp:320-328 */
inline uint32_t get_U(float u, const Matcap& m) {
  return (uint32_t)std::floor(0.5f + (u + 1.f) * (float)(m.w - 1) * 0.5f);
}
inline uint32_t get_V(float v, const Matcap& m) {
  return (uint32_t)std::floor(0.5f + (-v + 1.f) * (float)(m.h - 1) * 0.5f);
}

/* canvas.cpp:330-344; the texel read is clamped to the table */
inline uint32_t get_color(const Matcap& m, uint32_t U, uint32_t V, uint32_t shadow) {
  if (U >= m.w) U = m.w - 1;
  if (V >= m.h) V = m.h - 1;
  uint32_t clr = m.im[(size_t)V * m.w + U];
  if (shadow) {
    uint32_t r = clr & 0xff, g = (clr >> 8) & 0xff, b = (clr >> 16) & 0xff;
    clr = make_color(r >> 2, g >> 2, b >> 2);
  }
  return clr;
}

/* canvas.cpp:346-389.  acos and the scale expression are evaluated in double
   (unqualified acos on a float resolves to the C double function when only
   <cmath> is in scope); the result is rounded to float once. */
inline uint32_t get_angle_color(float angle, const Matcap& m, float u, float v, uint32_t mark) {
  uint32_t U = get_U(u, m), V = get_V(v, m);
  uint32_t clr = get_color(m, U, V, mark);
  if (std::fabs(angle) <= 1.f) {
    uint32_t r = clr & 0xff, g = (clr >> 8) & 0xff, b = (clr >> 16) & 0xff;
    const double hp = (double)1.57079632679489f;
    float scale = (float)((hp - std::fabs(std::acos((double)angle) - hp)) / hp);
    scale = std::sqrt(1.f - scale);
    uint32_t c2 = m.cavity;
    uint32_t r2 = c2 & 0xff, g2 = (c2 >> 8) & 0xff, b2 = (c2 >> 16) & 0xff;
    float k = angle > 0.f ? 1.2f : 0.5f; /* concave : convex */
    r = (uint32_t)((float)r * (1.f - scale) + (float)r2 * scale * k);
    g = (uint32_t)((float)g * (1.f - scale) + (float)g2 * scale * k);
    b = (uint32_t)((float)b * (1.f - scale) + (float)b2 * scale * k);
    if (r > 255) r = 255;
    if (g > 255) g = 255;
    if (b > 255) b = 255;
    clr = make_color(r, g, b);
  }
  return clr;
}

/* canvas.cpp:410-441 */
inline uint32_t shade_plain(const orc_pixel* p, const Matcap& m, uint32_t flags, float u, float v) {
  if (p->mark & 2) {
    if (flags & ORC_SHADING) {
      float nz = std::sqrt(1.f - u * u - v * v);
      float d = u * 0.f + v * 0.f + nz * 1.f + 0.f * 0.f;
      float occ = (p->mark & 1) ? 0.3f : 1.f;
      float dif = (d < 0.f ? 0.f : (d > 1.f ? 1.f : d)) * occ;
      return 0xff000000u | ((uint32_t)((float)p->b * dif) << 16) | ((uint32_t)((float)p->g * dif) << 8) |
             (uint32_t)((float)p->r * dif);
    }
    if (p->mark & 1) return make_color(p->r >> 2, p->g >> 2, p->b >> 2);
    return make_color(p->r, p->g, p->b);
  }
  return get_color(m, get_U(u, m), get_V(v, m), p->mark);
}

/* canvas.cpp:287-315 */
inline float convex_cos_angle(const float* Pinv, float near_plane, float W, float H, float x1, float y1,
                              float u1, float v1, float depth1, float x2, float y2, float u2, float v2,
                              float depth2) {
  f4 s1 = {2.f * ((x1 + 0.5f) / W) - 1.f, 2.f * ((y1 + 0.5f) / H) - 1.f, near_plane, 1.f};
  f4 d1 = mv(Pinv, s1);
  d1.w = 0.f;
  f4 p1 = {depth1 * d1.x, depth1 * d1.y, depth1 * d1.z, depth1 * d1.w};
  f4 s2 = {2.f * ((x2 + 0.5f) / W) - 1.f, 2.f * ((y2 + 0.5f) / H) - 1.f, near_plane, 1.f};
  f4 d2 = mv(Pinv, s2);
  d2.w = 0.f;
  f4 p2 = {depth2 * d2.x, depth2 * d2.y, depth2 * d2.z, depth2 * d2.w};
  f4 n1 = {u1, v1, std::sqrt(1.f - u1 * u1 - v1 * v1), 0.f};
  f4 n2 = {u2, v2, std::sqrt(1.f - u2 * u2 - v2 * v2), 0.f};
  float angle;
  if (std::fabs((double)(dot4(n1, n2) - 1.f)) > 0.0001) {
    f4 pt = {p2.x - p1.x, p2.y - p1.y, p2.z - p1.z, p2.w - p1.w};
    float len = std::sqrt(dot4(pt, pt));
    pt.x = pt.x / len, pt.y = pt.y / len, pt.z = pt.z / len, pt.w = pt.w / len;
    angle = dot4(pt, n1);
  } else {
    angle = 0.f;
  }
  return angle;
}

const uint32_t MISS = 0xFFFFFFFFu;

}  // namespace

/* ============================================================== C API === */

extern "C" {

void orc_default_camera(orc_camera* c) {
  /* j3d/camera.cpp:5-16 */
  c->focal_length = 35.f;
  c->film_w_inch = 1.024f;
  c->film_h_inch = 0.768f;
  c->near_plane = 0.1f;
  c->far_plane = FLT_MAX;
  c->fit_film = 1;
  c->zoom = 1.f;
}

void orc_frustum(float left, float right, float bottom, float top, float n, float f, float* o) {
  /* j3d/camera.cpp:18-24, four columns */
  std::memset(o, 0, 16 * sizeof(float));
  o[0] = 2.f * n / (right - left);
  o[5] = -2.f * n / (top - bottom);
  o[8] = (right + left) / (right - left);
  o[9] = -(top + bottom) / (top - bottom);
  o[10] = -(f + n) / (f - n);
  o[11] = -1.f;
  o[14] = -(2.f * f * n) / (f - n);
}

void orc_invert_projection(const float* in, float* o) {
  /* j3d/camera.cpp:26-30 */
  std::memset(o, 0, 16 * sizeof(float));
  o[0] = 1.f / in[0];
  o[5] = 1.f / in[5];
  o[11] = 1.f / in[14];
  o[12] = in[8] / in[0];
  o[13] = in[9] / in[5];
  o[14] = -1.f;
  o[15] = in[10] / in[14];
}

void orc_projection(const orc_camera* c, int w, int h, float* P) {
  /* j3d/camera.cpp:32-73 */
  const float inch_to_mm = 25.4f;
  float top = ((c->film_h_inch * inch_to_mm / 2.f) / c->focal_length) * c->near_plane;
  float right = ((c->film_w_inch * inch_to_mm / 2.f) / c->focal_length) * c->near_plane;
  float xscale = c->zoom, yscale = c->zoom;
  float device_aspect = (float)w / (float)h;
  float film_aspect = c->film_w_inch / c->film_h_inch;
  if (c->fit_film == 1) { /* Overscan */
    if (film_aspect > device_aspect)
      yscale *= film_aspect / device_aspect;
    else
      xscale *= device_aspect / film_aspect;
  } else { /* Fill */
    if (film_aspect > device_aspect)
      xscale *= device_aspect / film_aspect;
    else
      yscale *= film_aspect / device_aspect;
  }
  right *= xscale;
  top *= yscale;
  orc_frustum(-right, right, -top, top, c->near_plane, c->far_plane, P);
}

void orc_invert_orthonormal(const float* m, float* o) {
  /* rigid inverse: R^T and -R^T t */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o[4 * j + i] = m[4 * i + j];
  o[3] = o[7] = o[11] = 0.f;
  for (int i = 0; i < 3; ++i) o[12 + i] = -(m[4 * i] * m[12] + m[4 * i + 1] * m[13] + m[4 * i + 2] * m[14]);
  o[15] = 1.f;
}

void orc_mat_vec(const float* m, const float* v, float* out) {
  f4 r = mv(m, f4{v[0], v[1], v[2], v[3]});
  out[0] = r.x, out[1] = r.y, out[2] = r.z, out[3] = r.w;
}

void orc_matcap(int type, uint32_t* im, uint32_t* cavity) {
  const uint32_t w = 512, h = 512;
  if (type == 3) { /* sketch, j3d/matcap.cpp:241-268 */
    *cavity = 0xFF505050u;
    const float thres = 0.4f;
    for (uint32_t y = 0; y < h; ++y) {
      uint32_t* row = im + (size_t)(h - y - 1) * w;
      for (uint32_t x = 0; x < w; ++x) {
        float u = (float)x / (float)(w - 1) * 2.f - 1.f;
        float v = (float)y / (float)(h - 1) * 2.f - 1.f;
        float val = std::fabs(1.f - u * u - v * v);
        if (val < thres) {
          float scale = val / thres;
          uint32_t s = (uint32_t)(scale * (float)0xe1);
          row[x] = 0xff000000u | (s << 16) | (s << 8) | s;
        } else {
          row[x] = 0xffe1e1e1u;
        }
      }
    }
    return;
  }
  /* red wax (164-239), gray (9-84), brown (87-162) share one generator */
  float c1[3], c2[3], base[3];
  const float c3[3] = {50.f, 50.f, 50.f}, c4[3] = {30.f, 30.f, 30.f};
  if (type == 1) {
    *cavity = 0xff505050u;
    c1[0] = 200.f / 4.f, c1[1] = 200.f / 4.f, c1[2] = 200.f / 4.f;
    c2[0] = 50.f, c2[1] = 50.f, c2[2] = 50.f;
    base[0] = 32.f, base[1] = 32.f, base[2] = 32.f;
  } else if (type == 2) {
    *cavity = 0xff405060u;
    c1[0] = 200.f / 4.f, c1[1] = 180.f / 4.f, c1[2] = 160.f / 4.f;
    c2[0] = 50.f, c2[1] = 40.f, c2[2] = 30.f;
    base[0] = 32.f, base[1] = 20.f, base[2] = 10.f;
  } else {
    *cavity = 0xFF7D7DFFu;
    c1[0] = 200.f / 1.5f, c1[1] = 200.f / 4.f, c1[2] = 150.f / 4.f;
    c2[0] = 30.f, c2[1] = 25.f, c2[2] = 20.f;
    base[0] = 32.f, base[1] = 0.f, base[2] = 0.f;
  }
  auto normalize = [](float* n) {
    float len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    n[0] = n[0] / len, n[1] = n[1] / len, n[2] = n[2] / len;
  };
  float n1[3] = {0.f, 0.8f, 1.f}, n2[3] = {0.f, 0.4f, 1.f}, n3[3] = {0.f, 0.f, 1.f};
  normalize(n1);
  normalize(n2);
  normalize(n3);
  for (uint32_t y = 0; y < h; ++y) {
    uint32_t* row = im + (size_t)(h - y - 1) * w;
    for (uint32_t x = 0; x < w; ++x) {
      float u = (float)x / (float)(w - 1) * 2.f - 1.f;
      float v = (float)y / (float)(h - 1) * 2.f - 1.f;
      if (u * u + v * v <= 1.01f) {
        float lw = std::sqrt(1.01f - u * u - v * v);
        float sp[3] = {u, v, lw};
        float a1 = dot3(sp, n1);
        float a2 = powf(dot3(sp, n2), 3.f);
        float a3 = powf(dot3(sp, n3), 5.f);
        float a4 = powf(dot3(sp, n3), 50.f);
        float ch[3];
        for (int k = 0; k < 3; ++k) {
          float s = (c1[k] * a1 + c2[k] * a2 + c3[k] * a3 + c4[k] * a4) / 1.f;
          /* red wax adds no base to green and blue: 0 + s == s exactly */
          ch[k] = base[k] + s;
          if (ch[k] > 255.f) ch[k] = 255.f;
          if (ch[k] < 0.f) ch[k] = 0.f;
        }
        row[x] = make_color((unsigned char)ch[0], (unsigned char)ch[1], (unsigned char)ch[2]);
      } else {
        row[x] = 0xff000000u;
      }
    }
  }
}

void orc_fill_background(uint32_t* im, uint32_t w, uint32_t h, uint32_t top, uint32_t bottom) {
  /* j3d/canvas.cpp:55-78 */
  uint32_t tr = top & 0xff, tg = (top >> 8) & 0xff, tb = (top >> 16) & 0xff;
  uint32_t br = bottom & 0xff, bg = (bottom >> 8) & 0xff, bb = (bottom >> 16) & 0xff;
  for (uint32_t y = 0; y < h; ++y) {
    float scale = (float)y / (float)h;
    uint32_t r = (uint32_t)(scale * (float)br + (1.f - scale) * (float)tr);
    uint32_t g = (uint32_t)(scale * (float)bg + (1.f - scale) * (float)tg);
    uint32_t b = (uint32_t)(scale * (float)bb + (1.f - scale) * (float)tb);
    uint32_t clr = make_color((unsigned char)r, (unsigned char)g, (unsigned char)b);
    for (uint32_t x = 0; x < w; ++x) im[(size_t)y * w + x] = clr;
  }
}

void orc_triangle_normals(const float* v, const uint32_t* t, uint64_t nt, float* n) {
  for (uint64_t i = 0; i < nt; ++i) {
    const float* a = v + 3 * (size_t)t[3 * i];
    const float* b = v + 3 * (size_t)t[3 * i + 1];
    const float* c = v + 3 * (size_t)t[3 * i + 2];
    float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    float r[3];
    cross3(e1, e2, r);
    float len = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    n[3 * i] = r[0] / len;
    n[3 * i + 1] = r[1] / len;
    n[3 * i + 2] = r[2] / len;
  }
}

void orc_bbox(const float* v, uint64_t nv, float* mn, float* mx) {
  /* j3d/mesh.cpp:39-58 */
  if (nv == 0) return;
  for (int k = 0; k < 3; ++k) mn[k] = mx[k] = v[k];
  for (uint64_t i = 1; i < nv; ++i)
    for (int k = 0; k < 3; ++k) {
      mn[k] = std::min(mn[k], v[3 * i + k]);
      mx[k] = std::max(mx[k], v[3 * i + k]);
    }
}

void orc_prepare_and_unzoom(uint32_t nobj, const float* cs16, const float* bbmin3, const float* bbmax3,
                            float* diagonal, float* pivot, float* cs, float* cs_inv) {
  /* j3d/scene.cpp:50-89 and unzoom 91-110 (y axis up).  The boxes are the scene's
     objects followed by its point clouds: the first one is the reference's
     starting box (the first object's, or without objects the first cloud's). */
  float smin[3] = {0, 0, 0}, smax[3] = {0, 0, 0};
  for (uint32_t i = 0; i < nobj; ++i) {
    const float* m = cs16 + 16 * i;
    f4 a = mv(m, f4{bbmin3[3 * i], bbmin3[3 * i + 1], bbmin3[3 * i + 2], 1.f});
    f4 b = mv(m, f4{bbmax3[3 * i], bbmax3[3 * i + 1], bbmax3[3 * i + 2], 1.f});
    float lo[3] = {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)};
    float hi[3] = {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)};
    for (int k = 0; k < 3; ++k) {
      if (i == 0) {
        smin[k] = lo[k];
        smax[k] = hi[k];
      }
      smin[k] = std::min(smin[k], lo[k]);
      smax[k] = std::max(smax[k], hi[k]);
    }
  }
  float d = smax[0] - smin[0];
  d = std::max(d, smax[1] - smin[1]);
  d = std::max(d, smax[2] - smin[2]);
  *diagonal = d;
  std::memset(cs, 0, 16 * sizeof(float));
  cs[0] = cs[5] = cs[10] = cs[15] = 1.f;
  for (int k = 0; k < 3; ++k) pivot[k] = (smin[k] + smax[k]) * 0.5f;
  cs[12] = pivot[0];
  cs[13] = pivot[1];
  cs[14] = pivot[2] + d * 2.f;
  orc_invert_orthonormal(cs, cs_inv);
}

orc_bvh* orc_bvh_build(const float* v, uint64_t nv, const uint32_t* t, uint64_t nt, int threads) {
  (void)nv;
  orc_bvh* b = new orc_bvh;
  if (nt == 0) {
    BNode leaf;
    leaf.lo[0] = leaf.lo[1] = leaf.lo[2] = 1.f;
    leaf.hi[0] = leaf.hi[1] = leaf.hi[2] = -1.f; /* empty box: never entered */
    leaf.a = 0;
    leaf.b = LEAF_BIT;
    b->nodes.push_back(leaf);
    return b;
  }
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  std::vector<float> cen(3 * nt), blo(3 * nt), bhi(3 * nt);
  b->prim.resize(nt);
  for (uint64_t i = 0; i < nt; ++i) {
    b->prim[i] = (uint32_t)i;
    for (int k = 0; k < 3; ++k) {
      float a0 = v[3 * (size_t)t[3 * i] + k], a1 = v[3 * (size_t)t[3 * i + 1] + k],
            a2 = v[3 * (size_t)t[3 * i + 2] + k];
      blo[3 * i + k] = std::min(a0, std::min(a1, a2));
      bhi[3 * i + k] = std::max(a0, std::max(a1, a2));
      cen[3 * i + k] = (blo[3 * i + k] + bhi[3 * i + k]) * 0.5f;
    }
  }
  b->nodes.resize(2 * nt + 1);
  BuildCtx c;
  c.cen = cen.data();
  c.blo = blo.data();
  c.bhi = bhi.data();
  c.idx = b->prim.data();
  c.nodes = b->nodes.data();
  c.n_nodes = 1;
  c.spare_threads = threads - 1;
  build_rec(c, 0, 0, (uint32_t)nt, 0);
  b->nodes.resize(c.n_nodes.load());
  for (BNode& nd : b->nodes) {
    for (int k = 0; k < 3; ++k) {
      float m = std::max(std::fabs(nd.lo[k]), std::fabs(nd.hi[k]));
      float e = m * BOX_WIDEN + 1e-30f;
      nd.lo[k] = nd.lo[k] - e;
      nd.hi[k] = nd.hi[k] + e;
    }
  }
  b->tri.resize(9 * nt);
  for (uint64_t s = 0; s < nt; ++s) {
    uint32_t p = b->prim[s];
    const float* a = v + 3 * (size_t)t[3 * (size_t)p];
    const float* bb = v + 3 * (size_t)t[3 * (size_t)p + 1];
    const float* cc = v + 3 * (size_t)t[3 * (size_t)p + 2];
    float* o = &b->tri[9 * s];
    for (int k = 0; k < 3; ++k) {
      o[k] = a[k];
      o[3 + k] = bb[k] - a[k];
      o[6 + k] = cc[k] - a[k];
    }
  }
  return b;
}

void orc_bvh_free(orc_bvh* b) { delete b; }

uint64_t orc_bvh_node_count(const orc_bvh* b) { return b ? b->nodes.size() : 0; }

int orc_intersect_triangle(const float* o, const float* d, const float* v0, const float* v1, const float* v2,
                           float t_near, float t_far, float* t, float* u, float* v) {
  float e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
  float e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
  float tt = 0, uu = 0, vv = 0;
  bool h = mt_test(o, d, v0, e1, e2, t_near, t_far, tt, uu, vv);
  *t = tt, *u = uu, *v = vv;
  return h ? 1 : 0;
}

static void inv_cs_table(const orc_object* objs, uint32_t nobj, std::vector<float>& inv) {
  inv.resize(16 * (size_t)std::max<uint32_t>(nobj, 1));
  for (uint32_t i = 0; i < nobj; ++i) orc_invert_orthonormal(objs[i].cs, &inv[16 * i]);
}

void orc_closest_hits(const orc_object* objs, uint32_t nobj, const float* rays, uint64_t n, int threads,
                      uint32_t* obj, uint32_t* tri, float* t, float* u, float* v, float* t2) {
  std::vector<float> inv;
  inv_cs_table(objs, nobj, inv);
  const uint64_t chunk = 256;
  uint64_t nchunks = (n + chunk - 1) / chunk;
  parallel_rows(0, (uint32_t)nchunks, threads, [&](uint32_t ci) {
    for (uint64_t i = ci * chunk; i < std::min(n, (ci + 1) * chunk); ++i) {
      const float* r = rays + 8 * i;
      Best b = closest(objs, nobj, inv.data(), f4{r[0], r[1], r[2], 1.f}, f4{r[3], r[4], r[5], 0.f}, r[6], r[7],
                       t2 != nullptr);
      if (obj) obj[i] = b.found ? b.obj : MISS;
      if (tri) tri[i] = b.found ? b.tri : MISS;
      if (t) t[i] = b.found ? b.t : FLT_MAX;
      if (u) u[i] = b.found ? b.u : 0.f;
      if (v) v[i] = b.found ? b.v : 0.f;
      if (t2) t2[i] = b.t2;
    }
  });
}

void orc_any_hits(const orc_object* objs, uint32_t nobj, const float* rays, uint64_t n, int threads,
                  uint8_t* hit) {
  std::vector<float> inv;
  inv_cs_table(objs, nobj, inv);
  const uint64_t chunk = 256;
  uint64_t nchunks = (n + chunk - 1) / chunk;
  parallel_rows(0, (uint32_t)nchunks, threads, [&](uint32_t ci) {
    for (uint64_t i = ci * chunk; i < std::min(n, (ci + 1) * chunk); ++i) {
      const float* r = rays + 8 * i;
      hit[i] = any_hit(objs, nobj, inv.data(), f4{r[0], r[1], r[2], 1.f}, f4{r[3], r[4], r[5], 0.f}, r[6], r[7])
                   ? 1
                   : 0;
    }
  });
}

/* ---------------------------------------------- all hits, voxels ---- */

struct AnyHit4 {
  float t, u, v;
  uint32_t tri;
};

/* Every accepted hit of one object-space ray: the BVH (conservative, so the
   same set as brute force) or, without one, every triangle. */
static void all_hits_object(const orc_object& ob, const Ray& r, std::vector<AnyHit4>& out) {
  out.clear();
  auto consider = [&](uint32_t id, const float* v0, const float* e1, const float* e2) {
    float t, u, v;
    if (mt_test(r.o, r.d, v0, e1, e2, r.t_near, r.t_far, t, u, v)) out.push_back(AnyHit4{t, u, v, id});
  };
  if (!ob.bvh) {
    for (uint64_t i = 0; i < ob.n_triangles; ++i) {
      const uint32_t* tr = ob.triangles + 3 * i;
      const float* a = ob.vertices + 3 * (size_t)tr[0];
      const float* b = ob.vertices + 3 * (size_t)tr[1];
      const float* c = ob.vertices + 3 * (size_t)tr[2];
      float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
      float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
      consider((uint32_t)i, a, e1, e2);
    }
  } else {
    const orc_bvh& bvh = *ob.bvh;
    std::vector<uint32_t> stack{0u};
    while (!stack.empty()) {
      const uint32_t ni = stack.back();
      stack.pop_back();
      const BNode& nd = bvh.nodes[ni];
      if (box_entry(nd, r, r.t_far) == INFINITY) continue;
      if (nd.b & LEAF_BIT) {
        for (uint32_t k = 0; k < (nd.b & ~LEAF_BIT); ++k) {
          const float* tr = &bvh.tri[9 * (size_t)(nd.a + k)];
          consider(bvh.prim[nd.a + k], tr, tr + 3, tr + 6);
        }
      } else {
        stack.push_back(nd.a);
        stack.push_back(nd.b);
      }
    }
  }
  std::sort(out.begin(), out.end(),
            [](const AnyHit4& x, const AnyHit4& y) { return x.t < y.t || (x.t == y.t && x.tri < y.tri); });
}

void orc_all_hits(const orc_object* ob, const float* rays, uint64_t n, uint32_t max_hits, int threads,
                  uint32_t* counts, uint32_t* tri, float* t, float* u, float* v) {
  std::vector<float> inv;
  inv_cs_table(ob, 1, inv);
  const uint64_t chunk = 256;
  const uint64_t nchunks = (n + chunk - 1) / chunk;
  parallel_rows(0, (uint32_t)nchunks, threads, [&](uint32_t ci) {
    std::vector<AnyHit4> hits;
    for (uint64_t i = ci * chunk; i < std::min(n, (ci + 1) * chunk); ++i) {
      const float* q = rays + 8 * i;
      const Ray r = to_object(*ob, inv.data(), f4{q[0], q[1], q[2], 1.f}, f4{q[3], q[4], q[5], 0.f}, q[6], q[7]);
      all_hits_object(*ob, r, hits);
      counts[i] = (uint32_t)hits.size();
      for (uint32_t k = 0; k < max_hits; ++k) {
        const size_t o = (size_t)i * max_hits + k;
        const bool has = k < hits.size();
        if (tri) tri[o] = has ? hits[k].tri : MISS;
        if (t) t[o] = has ? hits[k].t : 0.f;
        if (u) u[o] = has ? hits[k].u : 0.f;
        if (v) v[o] = has ? hits[k].v : 0.f;
      }
    }
  });
}

/* (uint32_t)(x) of a float that should lie in [0, n): truncation toward zero,
   saturated into [0, n - 1]; NaN -> 0. */
static uint32_t vox_index(float x, uint32_t n) {
  if (!(x > 0.f)) return 0;
  if (x >= (float)n) return n - 1;
  return std::min((uint32_t)x, n - 1);
}

void orc_vox_dims(const float* mn, const float* mx, uint32_t max_dim, uint32_t dim[3]) {
  int L = 0;
  if ((mx[1] - mn[1]) > (mx[L] - mn[L])) L = 1;
  if ((mx[2] - mn[2]) > (mx[L] - mn[L])) L = 2;
  for (int j = 0; j < 3; ++j) {
    const float q = (float)max_dim * (mx[j] - mn[j]) / (mx[L] - mn[L]);
    uint32_t d = 0;
    if (q > 0.f) d = q >= 4294967040.f ? 0xFFFFFF00u : (uint32_t)q;
    dim[j] = d ? d : 1u;
  }
}

uint8_t orc_vox_color_index(uint8_t r, uint8_t g, uint8_t b) {
  r = r < 16 ? 0 : (uint8_t)(r - 16);
  g = g < 16 ? 0 : (uint8_t)(g - 16);
  b = b < 32 ? 0 : (uint8_t)(b - 32);
  const uint8_t ret = (uint8_t)(((r >> 5) << 5) | ((g >> 5) << 2) | (b >> 6));
  return ret ? ret : 1;
}

void orc_voxelize(const orc_object* ob, const float* mn, const float* mx, const uint32_t* dim, uint8_t* data) {
  const uint64_t nvox = (uint64_t)dim[0] * dim[1] * dim[2];
  std::memset(data, 0, nvox);
  const bool use_texture = ob->uv && ob->texture && ob->tex_w > 0 && ob->tex_h > 0;
  const bool use_vcol = !use_texture && ob->vertex_colors;
  std::vector<AnyHit4> hits;
  for (int dir = 0; dir < 3; ++dir) {
    const int d1a = (dir + 1) % 3, d2a = (dir + 2) % 3;
    Ray r;
    for (int k = 0; k < 3; ++k) r.d[k] = 0.f;
    r.d[dir] = dir == 2 ? 2.f : 1.f;
    r.t_near = 0.f;
    r.t_far = FLT_MAX;
    set_inv(r);
    for (uint32_t d1 = 0; d1 < dim[d1a]; ++d1)
      for (uint32_t d2 = 0; d2 < dim[d2a]; ++d2) {
        r.o[dir] = mn[dir];
        r.o[d1a] = (((float)d1 + 0.5f) / (float)dim[d1a]) * (mx[d1a] - mn[d1a]) + mn[d1a];
        r.o[d2a] = (((float)d2 + 0.5f) / (float)dim[d2a]) * (mx[d2a] - mn[d2a]) + mn[d2a];
        all_hits_object(*ob, r, hits);
        for (const AnyHit4& h : hits) {
          const uint32_t* tr = ob->triangles + 3 * (size_t)h.tri;
          const float* p0 = ob->vertices + 3 * (size_t)tr[0];
          const float* p1 = ob->vertices + 3 * (size_t)tr[1];
          const float* p2 = ob->vertices + 3 * (size_t)tr[2];
          const float w = 1.f - h.u - h.v;
          float pos[3], clr[3] = {1.f, 1.f, 1.f};
          for (int k = 0; k < 3; ++k) pos[k] = p0[k] * w + h.u * p1[k] + h.v * p2[k];
          if (use_texture) {
            const float* uvc = ob->uv + 6 * (size_t)h.tri;
            float c[2];
            for (int k = 0; k < 2; ++k) {
              c[k] = w * uvc[k] + h.u * uvc[2 + k] + h.v * uvc[4 + k];
              c[k] = std::max(std::min(c[k], 1.f), 0.f);
            }
            const uint32_t x = vox_index(c[0] * (float)(ob->tex_w - 1), ob->tex_w);
            const uint32_t y = vox_index(c[1] * (float)(ob->tex_h - 1), ob->tex_h);
            const uint32_t col = ob->texture[(size_t)y * ob->tex_w + x];
            clr[0] = (float)(col & 255) / 255.f;
            clr[1] = (float)((col >> 8) & 255) / 255.f;
            clr[2] = (float)((col >> 16) & 255) / 255.f;
          } else if (use_vcol) {
            const float* c0 = ob->vertex_colors + 3 * (size_t)tr[0];
            const float* c1 = ob->vertex_colors + 3 * (size_t)tr[1];
            const float* c2 = ob->vertex_colors + 3 * (size_t)tr[2];
            for (int k = 0; k < 3; ++k) clr[k] = c0[k] * w + h.u * c1[k] + h.v * c2[k];
          }
          uint32_t X[3];
          for (int k = 0; k < 3; ++k) X[k] = vox_index((pos[k] - mn[k]) / (mx[k] - mn[k]) * (float)dim[k], dim[k]);
          data[X[0] + ((size_t)X[1] + (size_t)X[2] * dim[1]) * dim[0]] =
              orc_vox_color_index(orc_color_byte(clr[0]), orc_color_byte(clr[1]), orc_color_byte(clr[2]));
        }
      }
  }
}

uint8_t orc_color_byte(float c) {
  const float x = c * 255.f;
  if (!(std::fabs(x) < 2147483648.f)) return 0;
  return (uint8_t)((uint32_t)(int32_t)x & 0xFFu);
}

int orc_hit_color(const orc_object* ob, uint32_t flags, uint32_t tri, float u, float v, uint8_t rgb[3]) {
  const float s0 = 1.f - u - v;
  if ((flags & ORC_TEXTURED) && ob->uv) {
    /* j3d/canvas.cpp:803-821: interpolated uv, clamped to [0, 1], nearest texel */
    const float* uv = ob->uv + 6 * (size_t)tri;
    float cx = s0 * uv[0] + u * uv[2] + v * uv[4];
    float cy = s0 * uv[1] + u * uv[3] + v * uv[5];
    cx = std::max(std::min(cx, 1.f), 0.f);
    cy = std::max(std::min(cy, 1.f), 0.f);
    const int w = (int)ob->tex_w, h = (int)ob->tex_h;
    int x = (int)(cx * (float)w);
    int y = (int)(cy * (float)h);
    x = x < 0 ? 0 : x >= w ? w - 1 : x;
    y = y < 0 ? 0 : y >= h ? h - 1 : y;
    const uint32_t color = ob->texture[(size_t)y * ob->tex_w + x];
    rgb[0] = (uint8_t)(color & 0xFFu);
    rgb[1] = (uint8_t)((color >> 8) & 0xFFu);
    rgb[2] = (uint8_t)((color >> 16) & 0xFFu);
    return 1;
  }
  if ((flags & ORC_VERTEXCOLORS) && ob->vertex_colors) {
    /* j3d/canvas.cpp:822-834: barycentric blend of the three vertex colours */
    const uint32_t* tr = ob->triangles + 3 * (size_t)tri;
    const float* c0 = ob->vertex_colors + 3 * (size_t)tr[0];
    const float* c1 = ob->vertex_colors + 3 * (size_t)tr[1];
    const float* c2 = ob->vertex_colors + 3 * (size_t)tr[2];
    for (int k = 0; k < 3; ++k) rgb[k] = orc_color_byte(c0[k] * s0 + u * c1[k] + v * c2[k]);
    return 1;
  }
  return 0;
}

int orc_update_canvas(const orc_object* objs, uint32_t nobj, const orc_frame* f, orc_pixel* gbuf, int threads,
                      int count_ties, orc_stats* st) {
  const uint32_t w = f->w, h = f->h;
  if (w == 0 || h == 0) return 0;
  int x0 = f->x0, y0 = f->y0, x1 = f->x1, y1 = f->y1;
  /* j3d/canvas.cpp:682-698 */
  if (x0 < 0) x0 = 0;
  if (y0 < 0) y0 = 0;
  if (x1 < 0) x1 = 0;
  if (y1 < 0) y1 = 0;
  if (x0 >= (int)w) x0 = w - 1;
  if (y0 >= (int)h) y0 = h - 1;
  if (x1 >= (int)w) x1 = w - 1;
  if (y1 >= (int)h) y1 = h - 1;

  std::atomic<uint64_t> n_prim(0), n_shadow(0), n_hits(0), n_shadowed(0), n_ties(0);
  if (nobj == 0) {
    /* j3d/canvas.cpp:744-760 */
    for (int y = y0; y <= y1; ++y)
      for (int x = x0; x <= x1; ++x) {
        orc_pixel& p = gbuf[(size_t)y * w + x];
        p.db_id = 0;
        p.object_id = MISS;
        p.u = 0.f;
        p.v = 0.f;
        p.depth = FLT_MAX;
      }
    if (st) *st = orc_stats{0, 0, 0, 0, 0};
    return 0;
  }
  std::vector<float> inv;
  inv_cs_table(objs, nobj, inv);
  const f4 origin = mv(f->cs, f4{0.f, 0.f, 0.f, 1.f});
  f4 light = {f->pivot[0] + f->diagonal * 3.f, f->pivot[1] + f->diagonal * 3.f, f->pivot[2] + f->diagonal * 3.f,
              1.f};
  light = mv(f->cs, light);
  const bool shadow = (f->flags & ORC_SHADOW) != 0;
  const bool directional = (f->flags & ORC_DIRECTIONAL_LIGHT) != 0;

  parallel_rows((uint32_t)y0, (uint32_t)y1 + 1, threads, [&](uint32_t y) {
    uint64_t lp = 0, ls = 0, lh = 0, lsh = 0, lt = 0;
    for (int x = x0; x <= x1; ++x) {
      orc_pixel& p = gbuf[(size_t)y * w + x];
      /* j3d/canvas.cpp:773-782 */
      f4 sp = {2.f * (((float)x + 0.5f) / (float)w) - 1.f, 2.f * (((float)y + 0.5f) / (float)h) - 1.f,
               f->near_plane, 1.f};
      f4 dir = mv(f->proj_inv, sp);
      dir.w = 0.f;
      dir = mv(f->cs, dir);
      float t_near = f->diagonal / 100.f;
      ++lp;
      Best b = closest(objs, nobj, inv.data(), origin, dir, t_near, FLT_MAX, count_ties != 0);
      if (b.found) {
        ++lh;
        if (count_ties && b.t2 != FLT_MAX && b.t2 - b.t <= 1e-6f * b.t) ++lt;
        const orc_object& ob = objs[b.obj];
        /* j3d/canvas.cpp:790-801 */
        const float* tn = ob.normals + 3 * (size_t)b.tri;
        f4 n = {tn[0], tn[1], tn[2], 0.f};
        n = mv(f->cs_inv, n);
        n = mv(ob.cs, n);
        p.u = n.x;
        p.v = n.y;
        p.depth = b.t;
        p.object_id = b.tri;
        p.barycentric_u = b.u;
        p.barycentric_v = b.v;
        p.db_id = ob.db_id;
        p.mark = 0;
        uint8_t rgb[3];
        if (orc_hit_color(&ob, f->flags, b.tri, b.u, b.v, rgb)) {
          p.r = rgb[0];
          p.g = rgb[1];
          p.b = rgb[2];
          p.mark |= 2;
        }
        if (shadow) {
          /* j3d/canvas.cpp:836-857 */
          const uint32_t* tr = ob.triangles + 3 * (size_t)b.tri;
          f4 V[3];
          for (int k = 0; k < 3; ++k) {
            const float* vv = ob.vertices + 3 * (size_t)tr[k];
            V[k] = mv(ob.cs, f4{vv[0], vv[1], vv[2], 1.f});
          }
          float s0 = 1.f - b.u - b.v;
          f4 pos = {V[0].x * s0 + b.u * V[1].x + b.v * V[2].x, V[0].y * s0 + b.u * V[1].y + b.v * V[2].y,
                    V[0].z * s0 + b.u * V[1].z + b.v * V[2].z, V[0].w * s0 + b.u * V[1].w + b.v * V[2].w};
          f4 ld;
          if (directional)
            ld = f4{f->light_dir[0], f->light_dir[1], f->light_dir[2], 0.f};
          else
            ld = f4{light.x - pos.x, light.y - pos.y, light.z - pos.z, light.w - pos.w};
          ++ls;
          if (any_hit(objs, nobj, inv.data(), pos, ld, 1e-3f, FLT_MAX)) {
            p.mark |= 1;
            ++lsh;
          }
        }
      } else {
        /* j3d/canvas.cpp:859-866 */
        p.db_id = 0;
        p.object_id = MISS;
        p.u = 0.f;
        p.v = 0.f;
        p.depth = FLT_MAX;
      }
    }
    n_prim += lp;
    n_shadow += ls;
    n_hits += lh;
    n_shadowed += lsh;
    n_ties += lt;
  });
  if (st) {
    st->primary_rays = n_prim;
    st->shadow_rays = n_shadow;
    st->hits = n_hits;
    st->shadowed = n_shadowed;
    st->near_ties = n_ties;
  }
  return 0;
}

void orc_canvas_to_image(const orc_pixel* gbuf, uint32_t w, uint32_t h, const float* Pinv, float near_plane,
                         uint32_t flags, const uint32_t* mim, uint32_t mw, uint32_t mh, uint32_t cavity,
                         uint32_t* im) {
  Matcap m{mim, mw, mh, cavity};
  const float th = 0.001f;
  const float W = (float)w, H = (float)h;
  if (w == 0 || h == 0) return;
  if (flags & ORC_ONE_BIT) {
    /* j3d/canvas.cpp:498-579 */
    for (uint32_t y = 0; y < h; ++y) {
      const orc_pixel* row = gbuf + (size_t)y * w;
      const orc_pixel* up = gbuf + (size_t)(y == 0 ? 0 : y - 1) * w;
      for (uint32_t x = 0; x + 1 < w; ++x) {
        const orc_pixel* p = row + x;
        if (p->object_id == MISS) continue;
        const orc_pixel* r = row + x + 1;
        const orc_pixel* q = up + x;
        float angle = 1.f;
        const orc_pixel* nb = nullptr;
        if (r->object_id != MISS && (std::fabs(p->u - r->u) > th || std::fabs(p->v - r->v) > th))
          nb = r;
        else if (q->object_id != MISS && (std::fabs(p->u - q->u) > th || std::fabs(p->v - q->v) > th))
          nb = q;
        if (nb) {
          float u1 = p->u, v1 = p->v, u2 = nb->u, v2 = nb->v;
          float w1 = std::sqrt(1.f - u1 * u1 - v1 * v1);
          float w2 = std::sqrt(1.f - u2 * u2 - v2 * v2);
          angle = u1 * u2 + v1 * v2 + w1 * w2;
        }
        uint32_t clr = get_color(m, get_U(p->u, m), get_V(p->v, m), p->mark);
        int res = (int)((((clr & 0xff0000) >> 16) + ((clr & 0xff00) >> 8) + (clr & 0xff)) >> 7);
        ++res;
        bool black = ((x % res == 0) && (y % res == 0));
        if (std::fabs(angle) < 0.95f) black = (res == 1) ? !black : true;
        im[(size_t)y * w + x] = black ? 0xff000000u : 0xffffffffu;
      }
      const orc_pixel* p = row + (w - 1);
      if (p->object_id != MISS) {
        uint32_t clr = get_color(m, get_U(p->u, m), get_V(p->v, m), p->mark);
        int res = (int)((((clr & 0xff0000) >> 16) + ((clr & 0xff00) >> 8) + (clr & 0xff)) >> 7);
        ++res;
        im[(size_t)y * w + w - 1] = (((w - 1) % res == 0) && (y % res == 0)) ? 0xff000000u : 0xffffffffu;
      }
    }
    return;
  }
  if (flags & ORC_WIREFRAME) {
    /* j3d/canvas.cpp:449-496 */
    for (uint32_t y = 0; y < h; ++y) {
      const orc_pixel* row = gbuf + (size_t)y * w;
      const orc_pixel* up = gbuf + (size_t)(y == 0 ? 0 : y - 1) * w;
      for (uint32_t x = 0; x + 1 < w; ++x) {
        const orc_pixel* p = row + x;
        if (p->object_id == MISS) continue;
        const orc_pixel* r = row + x + 1;
        const orc_pixel* q = up + x;
        if ((r->object_id != MISS && r->object_id != p->object_id) ||
            (q->object_id != MISS && q->object_id != p->object_id)) {
          const float scale = (p->u * p->u + p->v * p->v) * 0.5f;
          unsigned char c = (unsigned char)(255.f * scale);
          im[(size_t)y * w + x] = make_color(c, c, c);
        } else {
          im[(size_t)y * w + x] = shade_plain(p, m, flags, p->u, p->v);
        }
      }
      const orc_pixel* p = row + (w - 1);
      if (p->object_id != MISS) im[(size_t)y * w + w - 1] = shade_plain(p, m, flags, p->u, p->v);
    }
    return;
  }
  if (flags & ORC_EDGES) {
    /* j3d/canvas.cpp:590-649 */
    for (uint32_t y = 0; y < h; ++y) {
      const orc_pixel* row = gbuf + (size_t)y * w;
      const orc_pixel* up = gbuf + (size_t)(y == 0 ? 0 : y - 1) * w;
      for (uint32_t x = 0; x + 1 < w; ++x) {
        const orc_pixel* p = row + x;
        if (p->object_id == MISS) continue;
        const orc_pixel* r = row + x + 1;
        const orc_pixel* q = up + x;
        uint32_t clr;
        if (r->object_id != MISS && (std::fabs(p->u - r->u) > th || std::fabs(p->v - r->v) > th)) {
          float angle = convex_cos_angle(Pinv, near_plane, W, H, (float)x, (float)y, p->u, p->v, p->depth,
                                         (float)x + 1.f, (float)y, r->u, r->v, r->depth);
          clr = get_angle_color(angle, m, p->u, p->v, p->mark);
        } else if (q->object_id != MISS && (std::fabs(p->u - q->u) > th || std::fabs(p->v - q->v) > th)) {
          float angle = convex_cos_angle(Pinv, near_plane, W, H, (float)x, (float)y, p->u, p->v, p->depth,
                                         (float)x, (float)y - 1.f, q->u, q->v, q->depth);
          clr = get_angle_color(angle, m, p->u, p->v, p->mark);
        } else {
          clr = shade_plain(p, m, flags, p->u, p->v);
        }
        im[(size_t)y * w + x] = clr;
      }
      const orc_pixel* p = row + (w - 1);
      if (p->object_id != MISS) im[(size_t)y * w + w - 1] = shade_plain(p, m, flags, p->u, p->v);
    }
    return;
  }
  /* j3d/canvas.cpp:650-669 */
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t x = 0; x < w; ++x) {
      const orc_pixel* p = gbuf + (size_t)y * w + x;
      if (p->object_id != MISS) im[(size_t)y * w + x] = shade_plain(p, m, flags, p->u, p->v);
    }
}


/* ---- point clouds: canvas::render_pointclouds_on_image (j3d/canvas.cpp:952-1030) ---- */

/* The z-buffer of canvas.cpp:963-970: 1/depth where db_id != 0, else 0. */
void orc_splat_zbuffer(const orc_pixel* gbuf, uint64_t n_pix, float* z) {
  for (uint64_t i = 0; i < n_pix; ++i) z[i] = gbuf[i].db_id != 0 ? 1.f / gbuf[i].depth : 0.f;
}

/* One cloud drawn point after point in index order (canvas.cpp:983-1026) into
   one pixel each.  jtk's point renderer is absent from the reference tree; its
   projection is restated as clip = P (CS^-1 (cs p)), drawn when w is positive
   and finite and -w <= x, y <= w (no test on clip z, which the default far plane FLT_MAX of
   j3d/camera.cpp:12 makes infinite), at pixel floor((x/w + 1) * 0.5 * W), floor((y/w + 1) *
   0.5 * H) (W and H themselves dropped), depth test 1/w > z (strict).  A drawn
   point writes object_id = i, depth = 1/z, db_id (canvas.cpp:1000-1026) and
   its colour (0xffffffff without colours or in one-bit mode, canvas.cpp:995,
   997), with shading and normals each of R, G, B times n'.z / |n'| clamped to
   [0, 1], n' = CS^-1 (cs (n, 0)), truncated. */
void orc_splat_cloud(orc_pixel* gbuf, uint32_t* im, uint32_t w, uint32_t h, float* z, const float* proj,
                     const float* cs_inv, const float* cs, uint32_t db_id, const float* xyz, uint64_t n,
                     const float* normals, const uint32_t* colors, int shading, int one_bit) {
  const float* nrm = shading ? normals : nullptr;
  const uint32_t* col = one_bit ? nullptr : colors;
  for (uint64_t i = 0; i < n; ++i) {
    const f4 c = mv(proj, mv(cs_inv, mv(cs, f4{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 1.f})));
    if (!(c.w > 0.f) || std::isinf(c.w)) continue;
    if (!(c.x >= -c.w && c.x <= c.w && c.y >= -c.w && c.y <= c.w)) continue;
    const float sx = (c.x / c.w + 1.f) * 0.5f * (float)w;
    const float sy = (c.y / c.w + 1.f) * 0.5f * (float)h;
    const uint32_t px = (uint32_t)sx, py = (uint32_t)sy;
    if (px >= w || py >= h) continue;
    const size_t k = (size_t)py * w + px;
    const float iw = 1.f / c.w;
    if (!(iw > z[k])) continue;
    z[k] = iw;
    gbuf[k].object_id = (uint32_t)i;
    gbuf[k].depth = 1.f / z[k];
    gbuf[k].db_id = db_id;
    uint32_t clr = col ? col[i] : 0xffffffffu;
    if (nrm) {
      const f4 nn = mv(cs_inv, mv(cs, f4{nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2], 0.f}));
      const float len = std::sqrt(nn.x * nn.x + nn.y * nn.y + nn.z * nn.z);
      float l = nn.z / len;
      l = l > 0.f ? (l < 1.f ? l : 1.f) : 0.f;
      uint32_t ch[3];
      for (int j = 0; j < 3; ++j) ch[j] = (uint32_t)((float)((clr >> (8 * j)) & 0xFFu) * l);
      clr = (clr & 0xff000000u) | (ch[2] << 16) | (ch[1] << 8) | ch[0];
    }
    im[k] = clr;
  }
}

}  // extern "C"
