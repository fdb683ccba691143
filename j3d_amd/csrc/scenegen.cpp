// Seeded synthetic meshes and cameras (see include/j3dg_scenegen.h).
#include "j3dg_scenegen.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

const double kPi = 3.14159265358979323846;

struct V3 {
  double x, y, z;
};

V3 normalized(V3 a) {
  double l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  return V3{a.x / l, a.y / l, a.z / l};
}

double gaussian(uint64_t* s) {
  double u1 = j3dg_uniform01(s);
  double u2 = j3dg_uniform01(s);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
}

/* Gradient noise on the integer lattice (Perlin's improved noise with a
   seeded permutation).  Returns roughly [-1, 1]. */
struct Noise {
  int perm[512];
  explicit Noise(uint64_t seed) {
    int p[256];
    for (int i = 0; i < 256; ++i) p[i] = i;
    uint64_t s = seed;
    for (int i = 255; i > 0; --i) {
      int j = (int)(j3dg_splitmix64(&s) % (uint64_t)(i + 1));
      std::swap(p[i], p[j]);
    }
    for (int i = 0; i < 512; ++i) perm[i] = p[i & 255];
  }
  static double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
  static double lerp(double t, double a, double b) { return a + t * (b - a); }
  static double grad(int h, double x, double y, double z) {
    h &= 15;
    double u = h < 8 ? x : y;
    double v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
    return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
  }
  double operator()(double x, double y, double z) const {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int X = (int)((int64_t)fx & 255), Y = (int)((int64_t)fy & 255), Z = (int)((int64_t)fz & 255);
    x -= fx;
    y -= fy;
    z -= fz;
    double u = fade(x), v = fade(y), w = fade(z);
    int A = perm[X] + Y, AA = perm[A] + Z, AB = perm[A + 1] + Z;
    int B = perm[X + 1] + Y, BA = perm[B] + Z, BB = perm[B + 1] + Z;
    return lerp(w,
                lerp(v, lerp(u, grad(perm[AA], x, y, z), grad(perm[BA], x - 1, y, z)),
                     lerp(u, grad(perm[AB], x, y - 1, z), grad(perm[BB], x - 1, y - 1, z))),
                lerp(v, lerp(u, grad(perm[AA + 1], x, y, z - 1), grad(perm[BA + 1], x - 1, y, z - 1)),
                     lerp(u, grad(perm[AB + 1], x, y - 1, z - 1), grad(perm[BB + 1], x - 1, y - 1, z - 1))));
  }
};

}  // namespace

extern "C" {

uint64_t j3dg_splitmix64(uint64_t* state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double j3dg_uniform01(uint64_t* state) { return (double)((j3dg_splitmix64(state) >> 11) + 1) * (1.0 / 9007199254740992.0); }

void j3dg_random_hemisphere(uint64_t seed, float out[3]) {
  uint64_t s = seed;
  double z = j3dg_uniform01(&s);            /* cos(theta) uniform in (0,1] */
  double phi = 2.0 * kPi * j3dg_uniform01(&s);
  double rxy = std::sqrt(std::max(0.0, 1.0 - z * z));
  out[0] = (float)(rxy * std::cos(phi));
  out[1] = (float)(rxy * std::sin(phi));
  out[2] = (float)z;
}

uint64_t j3dg_icosphere_vertex_count(uint32_t level) { return 10ull * (1ull << (2 * level)) + 2ull; }
uint64_t j3dg_icosphere_triangle_count(uint32_t level) { return 20ull * (1ull << (2 * level)); }

int j3dg_gen_icosphere(uint32_t level, uint64_t seed, float amplitude, float* xyz, uint32_t* tris) {
  if (level > 12 || !xyz || !tris) return 1;
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  std::vector<V3> p = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                       {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  for (auto& v : p) v = normalized(v);
  std::vector<uint32_t> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4,
                             11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8,
                             3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
  for (uint32_t l = 0; l < level; ++l) {
    std::unordered_map<uint64_t, uint32_t> mid;
    mid.reserve(f.size());
    auto midpoint = [&](uint32_t a, uint32_t b) {
      uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      V3 m{(p[a].x + p[b].x) * 0.5, (p[a].y + p[b].y) * 0.5, (p[a].z + p[b].z) * 0.5};
      p.push_back(normalized(m));
      uint32_t id = (uint32_t)p.size() - 1;
      mid.emplace(key, id);
      return id;
    };
    std::vector<uint32_t> nf;
    nf.reserve(f.size() * 4);
    for (size_t i = 0; i < f.size(); i += 3) {
      uint32_t a = f[i], b = f[i + 1], c = f[i + 2];
      uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      uint32_t q[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
      nf.insert(nf.end(), q, q + 12);
    }
    f.swap(nf);
  }
  const size_t nv = p.size();
  std::vector<double> d(nv);
  uint64_t s = seed;
  for (size_t i = 0; i < nv; ++i) d[i] = gaussian(&s);
  /* 1-ring mean over unique edges */
  std::vector<double> sum(d);
  std::vector<uint32_t> cnt(nv, 1);
  {
    std::vector<uint64_t> edges;
    edges.reserve(f.size());
    for (size_t i = 0; i < f.size(); i += 3)
      for (int k = 0; k < 3; ++k) {
        uint32_t a = f[i + k], b = f[i + (k + 1) % 3];
        edges.push_back(a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a));
      }
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    for (uint64_t e : edges) {
      uint32_t a = (uint32_t)(e >> 32), b = (uint32_t)e;
      sum[a] += d[b];
      sum[b] += d[a];
      cnt[a]++;
      cnt[b]++;
    }
  }
  for (size_t i = 0; i < nv; ++i) {
    double r = 1.0 + (double)amplitude * (sum[i] / cnt[i]);
    xyz[3 * i] = (float)(p[i].x * r);
    xyz[3 * i + 1] = (float)(p[i].y * r);
    xyz[3 * i + 2] = (float)(p[i].z * r);
  }
  std::memcpy(tris, f.data(), f.size() * sizeof(uint32_t));
  return 0;
}

int j3dg_gen_torus(uint32_t nu, uint32_t nv, float R, float r, uint64_t seed, float amplitude, uint32_t octaves,
                   float lacunarity, float gain, float base_freq, int threads, float* xyz, uint32_t* tris) {
  if (nu < 3 || nv < 3 || !xyz || !tris) return 1;
  if ((uint64_t)nu * nv > 0xFFFFFFFFull / 2) return 1;
  Noise noise(seed);
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  auto work = [&](uint32_t i0, uint32_t i1) {
    for (uint32_t i = i0; i < i1; ++i) {
      double u = 2.0 * kPi * (double)i / nu;
      double cu = std::cos(u), su = std::sin(u);
      for (uint32_t j = 0; j < nv; ++j) {
        double v = 2.0 * kPi * (double)j / nv;
        double cv = std::cos(v), sv = std::sin(v);
        V3 n{cu * cv, su * cv, sv};
        V3 c{R * cu, R * su, 0.0};
        V3 p0{c.x + r * n.x, c.y + r * n.y, c.z + r * n.z};
        double h = 0.0, a = amplitude, fr = base_freq;
        for (uint32_t o = 0; o < octaves; ++o) {
          h += a * noise(p0.x * fr, p0.y * fr, p0.z * fr);
          a *= gain;
          fr *= lacunarity;
        }
        double rr = r + h;
        size_t k = (size_t)i * nv + j;
        xyz[3 * k] = (float)(c.x + rr * n.x);
        xyz[3 * k + 1] = (float)(c.y + rr * n.y);
        xyz[3 * k + 2] = (float)(c.z + rr * n.z);
        uint32_t i2 = (i + 1) % nu, j2 = (j + 1) % nv;
        uint32_t A = i * nv + j, B = i2 * nv + j, C = i2 * nv + j2, D = i * nv + j2;
        uint32_t* t = tris + 6 * k;
        t[0] = A, t[1] = B, t[2] = C;
        t[3] = A, t[4] = C, t[5] = D;
      }
    }
  };
  std::vector<std::thread> pool;
  uint32_t chunk = (nu + threads - 1) / threads;
  for (int k = 0; k < threads; ++k) {
    uint32_t a = std::min<uint32_t>(nu, k * chunk), b = std::min<uint32_t>(nu, (k + 1) * chunk);
    if (a < b) pool.emplace_back(work, a, b);
  }
  for (auto& th : pool) th.join();
  return 0;
}

int j3dg_orbit_camera(float dist, float azimuth, float elevation, const float target[3], float cs[16],
                      float cs_inv[16]) {
  double ce = std::cos((double)elevation), se = std::sin((double)elevation);
  double ca = std::cos((double)azimuth), sa = std::sin((double)azimuth);
  V3 back{ce * ca, ce * sa, se}; /* camera +z points away from the target */
  V3 pos{target[0] + dist * back.x, target[1] + dist * back.y, target[2] + dist * back.z};
  V3 up{0, 0, 1};
  V3 right{up.y * back.z - up.z * back.y, up.z * back.x - up.x * back.z, up.x * back.y - up.y * back.x};
  if (std::sqrt(right.x * right.x + right.y * right.y + right.z * right.z) < 1e-12) right = V3{1, 0, 0};
  right = normalized(right);
  V3 cup{back.y * right.z - back.z * right.y, back.z * right.x - back.x * right.z, back.x * right.y - back.y * right.x};
  double m[16] = {right.x, right.y, right.z, 0, cup.x, cup.y, cup.z, 0, back.x, back.y, back.z, 0, pos.x, pos.y, pos.z, 1};
  for (int i = 0; i < 16; ++i) cs[i] = (float)m[i];
  /* inverse of the float matrix: transpose and -R^T t, in double, rounded once */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) cs_inv[4 * j + i] = cs[4 * i + j];
  cs_inv[3] = cs_inv[7] = cs_inv[11] = 0.f;
  for (int i = 0; i < 3; ++i)
    cs_inv[12 + i] = (float)(-((double)cs[4 * i] * cs[12] + (double)cs[4 * i + 1] * cs[13] + (double)cs[4 * i + 2] * cs[14]));
  cs_inv[15] = 1.f;
  return 0;
}

}  // extern "C"
