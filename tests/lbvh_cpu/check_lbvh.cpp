// CPU check of the device BVH builder (j3d_amd/csrc/bvh_gpu.hip): its kernels
// and its host SAH over the top levels, compiled by g++ from the same source
// (lbvh_body.inc, cut out of bvh_gpu.hip by the test) and run serially with
// sanitizers, in the order device_build_mesh runs them (std::stable_sort in
// place of the device radix sort, which is stable too).  Checked on each tree:
// every triangle in exactly one leaf and every node reached once, leaf sizes
// and last-in-leaf flags, triangle records and completion records, decoded
// child boxes containing their subtrees, the depth and the traversal-stack
// bound.  Meshes: random soups, identical triangles (all Morton codes equal),
// a flat strip, soups with NaN vertices, all widths and leaf sizes.
#include <random>
#include <vector>

#include "shim.h"

namespace j3d {
#include "lbvh_body.inc"
}
using namespace j3d;
using namespace j3d::lbvh;

struct Built {
  std::vector<uint32_t> nodes;
  std::vector<float4> tris, fin;
  std::vector<float> normals;
  uint32_t n_wide = 0, levels = 0, max_stack = 0;
};

static int fails = 0;
#define CHECK(c, ...)                                   \
  do {                                                  \
    if (!(c) && fails++ < 10) {                         \
      printf("FAIL line %d: ", __LINE__);               \
      printf(__VA_ARGS__);                              \
      printf("\n");                                     \
    }                                                   \
  } while (0)

static Built build(const std::vector<float>& xyz, const std::vector<uint32_t>& t, int W, int leaf) {
  const uint32_t n = (uint32_t)(t.size() / 3);
  const int words = W == 8 ? 24 : 16;
  Built out;
  std::vector<float> box(6 * (size_t)n);
  std::vector<uint64_t> code(n);
  std::vector<uint32_t> index(n);
  out.normals.resize(3 * (size_t)n);
  out.fin.resize(3 * (size_t)n + 3);
  out.tris.assign(3 * (size_t)n + 3, float4{0, 0, 0, 0});
  LAUNCH(k_normals_fin, (n + 255) / 256, 256, xyz.data(), t.data(), n, out.normals.data(), out.fin.data());
  float lo[3], hi[3];
  for (int k = 0; k < 3; ++k) lo[k] = hi[k] = xyz[k];
  for (size_t i = 0; i < xyz.size() / 3; ++i)
    for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], xyz[3 * i + k]), hi[k] = std::max(hi[k], xyz[3 * i + k]);
  const float3 base = make_float3(lo[0], lo[1], lo[2]);
  const float3 inv = make_float3(hi[0] > lo[0] ? 1.f / (hi[0] - lo[0]) : 0.f, hi[1] > lo[1] ? 1.f / (hi[1] - lo[1]) : 0.f,
                                 hi[2] > lo[2] ? 1.f / (hi[2] - lo[2]) : 0.f);
  LAUNCH(k_prims, (n + 255) / 256, 256, xyz.data(), t.data(), n, base, inv, box.data(), code.data(), index.data());
  std::vector<std::pair<uint64_t, uint32_t>> kv(n);
  for (uint32_t i = 0; i < n; ++i) kv[i] = {code[i], index[i]};
  std::stable_sort(kv.begin(), kv.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  for (uint32_t i = 0; i < n; ++i) code[i] = kv[i].first, index[i] = kv[i].second;
  std::vector<BNode> bn(n - 1);
  LAUNCH(k_radix_tree, (n - 1 + 255) / 256, 256, code.data(), n, bn.data());
  std::vector<uint32_t> order(n - 1, 0xdeadbeefu), level_at = {0u, 1u};
  order[0] = 0;
  for (;;) {
    const uint32_t a = level_at[level_at.size() - 2], b = level_at.back();
    if (a == b) break;
    uint32_t c = 0;
    LAUNCH(k_level_down, (b - a + 255) / 256, 256, bn.data(), order.data() + a, b - a, order.data() + b, n - 1 - b, &c);
    CHECK(c <= n - 1 - b, "level list overflow");
    level_at.push_back(b + c);
  }
  CHECK(level_at.back() == n - 1, "level lists hold %u of %u nodes", level_at.back(), n - 1);
  for (size_t L = level_at.size() - 1; L-- > 0;) {
    const uint32_t a = level_at[L], b = level_at[L + 1];
    if (a != b) LAUNCH(k_level_boxes, (b - a + 255) / 256, 256, bn.data(), box.data(), index.data(), order.data() + a, b - a);
  }
  if (kSahTop && n >= 8192) {  // device_build_mesh's SAH over the top levels
    const uint32_t T = std::max<uint32_t>(n >> 14, 2u);
    std::vector<uint32_t> frontier(n), top(n), cnt2(2, 0);
    LAUNCH(k_frontier, (n - 1 + 255) / 256, 256, bn.data(), n, T, frontier.data(), top.data(), cnt2.data());
    const uint32_t m = cnt2[0], k = cnt2[1];
    CHECK(m >= 2 && k + 1 == m, "frontier %u, rebuilt %u", m, k);
    std::vector<float> fb(8 * (size_t)m);
    LAUNCH(k_frontier_boxes, (m + 255) / 256, 256, bn.data(), box.data(), index.data(), frontier.data(), m, fb.data());
    std::vector<Cluster> cl(m);
    for (uint32_t i = 0; i < m; ++i) {
      const float* f = &fb[8 * (size_t)i];
      for (int a = 0; a < 3; ++a) cl[i].lo[a] = f[a], cl[i].hi[a] = f[3 + a], cl[i].c[a] = 0.5f * f[a] + 0.5f * f[3 + a];
      memcpy(&cl[i].count, f + 6, 4);
      memcpy(&cl[i].ref, f + 7, 4);
    }
    std::vector<SahNode> sn;
    std::vector<float> right_area(m + 1);
    sah_build(cl, 0, m, sn, right_area);
    CHECK(sn.size() == k, "rebuilt %zu nodes for %u", sn.size(), k);
    std::vector<uint32_t> index_of(k);
    for (uint32_t j = 0, next = 0; j < k; ++j) {
      if (j == 0) continue;  // the root stays node 0
      while (top[next] == 0) ++next;
      index_of[j] = top[next++];
    }
    std::vector<BNode> vals(k);
    for (uint32_t j = 0; j < k; ++j) {
      auto resolve = [&](uint32_t r) { return (r & kSahTag) && !(r & kPrimBit) ? index_of[r & ~kSahTag] : r; };
      vals[j].left = resolve(sn[j].left);
      vals[j].right = resolve(sn[j].right);
      vals[j].first = 0;
      vals[j].last = std::max(sn[j].count, (uint32_t)leaf + 1u) - 1u;
      for (int a = 0; a < 3; ++a) vals[j].lo[a] = sn[j].lo[a], vals[j].hi[a] = sn[j].hi[a];
    }
    LAUNCH(k_put_nodes, (k + 255) / 256, 256, bn.data(), index_of.data(), vals.data(), k);
  }
  std::vector<uint8_t> last((n + 3) / 4 * 4, 0);
  std::vector<Item> q0(n), q1(n);
  std::vector<uint32_t> wide((size_t)n * words, 0xcdcdcdcdu);
  uint32_t cnt[4] = {0, 1, 0, 0};
  q0[0] = Item{0, 0, 0};
  uint32_t n_in = 1, levels = 0;
  Item* qin = q0.data();
  Item* qout = q1.data();
  while (n_in && levels < 96) {
    cnt[0] = 0;
    cnt[2] = 0;
    CollapseArgs ca{bn.data(), box.data(), index.data(), (uint32_t)leaf, qin, n_in, qout, cnt, n, wide.data(), last.data()};
    if (W == 8)
      LAUNCH(k_collapse<8>, (n_in + 127) / 128, 128, ca);
    else
      LAUNCH(k_collapse<4>, (n_in + 127) / 128, 128, ca);
    CHECK(cnt[2] == 0, "wide node capacity exceeded");
    ++levels;
    n_in = cnt[0];
    std::swap(qin, qout);
  }
  CHECK(n_in == 0, "collapse did not finish");
  LAUNCH(k_tri_records, (n + 255) / 256, 256, xyz.data(), t.data(), index.data(), last.data(), n, out.tris.data());
  out.n_wide = cnt[1];
  out.levels = levels;
  out.max_stack = cnt[3] + 1;
  out.nodes.assign(wide.begin(), wide.begin() + (size_t)cnt[1] * words);
  return out;
}

struct Box {
  float lo[3], hi[3];
};

static Box check(const Built& B, const std::vector<float>& xyz, const std::vector<uint32_t>& t, int W, int leaf,
                 uint32_t node, std::vector<int>& seen, std::vector<int>& nseen, uint32_t depth, uint32_t stack,
                 uint32_t& maxd, uint32_t& maxs) {
  const int words = W == 8 ? 24 : 16;
  Box all{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
  CHECK(node < B.n_wide, "node %u", node);
  if (node >= B.n_wide) return all;
  nseen[node]++;
  const uint32_t* nd = &B.nodes[(size_t)node * words];
  float p[3];
  memcpy(p, nd, 12);
  const int e[3] = {(int)(nd[3] & 255) - 127, (int)((nd[3] >> 8) & 255) - 127, (int)((nd[3] >> 16) & 255) - 127};
  int nch = 0;
  for (int i = 0; i < W; ++i) nch += nd[(W == 8 ? 16 : 10) + i] != kEmptyRef;
  maxd = std::max(maxd, depth);
  const uint32_t pushed = stack + nch - 1;
  maxs = std::max(maxs, pushed);
  for (int i = 0; i < W; ++i) {
    const uint32_t ref = nd[(W == 8 ? 16 : 10) + i];
    if (ref == kEmptyRef) {
      CHECK(i >= nch, "empty slot before a used one");
      continue;
    }
    float qlo[3], qhi[3];
    for (int k = 0; k < 3; ++k) {
      const uint32_t wl = W == 8 ? nd[4 + 4 * k + i / 4] : nd[4 + 2 * k];
      const uint32_t wh = W == 8 ? nd[6 + 4 * k + i / 4] : nd[5 + 2 * k];
      qlo[k] = (float)((double)p[k] + ((wl >> (8 * (i % 4))) & 255) * ldexp(1.0, e[k]));
      qhi[k] = (float)((double)p[k] + ((wh >> (8 * (i % 4))) & 255) * ldexp(1.0, e[k]));
    }
    Box cb{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
    if (ref & kLeafBit) {
      const uint32_t c = ((ref >> kLeafCountShift) & 3) + 1, f = ref & kLeafIndexMask;
      CHECK(c <= (uint32_t)leaf, "leaf of %u triangles", c);
      for (uint32_t s = f; s < f + c; ++s) {
        CHECK(s < t.size() / 3, "record %u out of range", s);
        if (s >= t.size() / 3) continue;
        const uint32_t id = __float_as_uint(B.tris[3 * s].w);
        CHECK(id < t.size() / 3, "id %u", id);
        if (id >= t.size() / 3) continue;
        seen[id]++;
        CHECK(__float_as_uint(B.tris[3 * s + 1].w) == (s == f + c - 1 ? 1u : 0u), "last flag at %u", s);
        const float* v0 = &xyz[3 * t[3 * id]];
        const float* v1 = &xyz[3 * t[3 * id + 1]];
        const float* v2 = &xyz[3 * t[3 * id + 2]];
        const float r[3][3] = {{B.tris[3 * s].x, B.tris[3 * s].y, B.tris[3 * s].z},
                               {B.tris[3 * s + 1].x, B.tris[3 * s + 1].y, B.tris[3 * s + 1].z},
                               {B.tris[3 * s + 2].x, B.tris[3 * s + 2].y, B.tris[3 * s + 2].z}};
        for (int k = 0; k < 3; ++k) {
          const float e1 = v1[k] - v0[k], e2 = v2[k] - v0[k];
          CHECK(__float_as_uint(r[0][k]) == __float_as_uint(v0[k]) && __float_as_uint(r[1][k]) == __float_as_uint(e1) &&
                    __float_as_uint(r[2][k]) == __float_as_uint(e2),
                "record %u", s);
          for (const float* v : {v0, v1, v2}) cb.lo[k] = std::min(cb.lo[k], v[k]), cb.hi[k] = std::max(cb.hi[k], v[k]);
        }
      }
    } else {
      cb = check(B, xyz, t, W, leaf, ref, seen, nseen, depth + 1, pushed, maxd, maxs);
    }
    for (int k = 0; k < 3; ++k) {
      CHECK(std::isnan(cb.lo[k]) || std::isnan(cb.hi[k]) || cb.lo[k] > cb.hi[k] || (qlo[k] <= cb.lo[k] && qhi[k] >= cb.hi[k]),
            "node %u child %d axis %d: [%g %g] does not hold [%g %g]", node, i, k, qlo[k], qhi[k], cb.lo[k], cb.hi[k]);
      all.lo[k] = std::min(all.lo[k], cb.lo[k]);
      all.hi[k] = std::max(all.hi[k], cb.hi[k]);
    }
  }
  return all;
}

int main() {
  std::mt19937 rng(7);
  std::uniform_real_distribution<float> U(-1.f, 1.f);
  struct Case {
    uint32_t n;
    int kind;  // 0 random soup, 1 identical triangles, 2 flat strip, 3 soup with NaN vertices
  };
  const Case cases[] = {{2, 0}, {3, 0}, {5, 0}, {8, 1}, {17, 0}, {100, 1}, {1000, 0}, {4096, 2},
                        {5000, 3}, {20000, 0}, {20000, 3}, {30000, 2}, {40000, 1}};
  for (const Case cs : cases)
    for (int W : {4, 8})
      for (int leaf = 1; leaf <= 4; ++leaf) {
        const uint32_t n = cs.n;
        std::vector<float> xyz;
        std::vector<uint32_t> t;
        if (cs.kind == 0 || cs.kind == 3) {
          for (uint32_t i = 0; i < 3 * n; ++i) {
            for (int k = 0; k < 3; ++k) xyz.push_back(U(rng));
            if (cs.kind == 3 && i % 7 == 0) xyz[3 * i + (i / 7) % 3] = NAN;
            t.push_back(i);
          }
        } else if (cs.kind == 1) {
          xyz = {0, 0, 0, 1, 0, 0, 0, 1, 0};
          for (uint32_t i = 0; i < n; ++i) t.insert(t.end(), {0u, 1u, 2u});
        } else {
          uint32_t g = 0;
          while (2 * g * g < n) ++g;
          for (uint32_t y = 0; y <= g; ++y)
            for (uint32_t x = 0; x <= g; ++x) xyz.insert(xyz.end(), {x * 1e3f, y * 1e-3f, 0.f});
          for (uint32_t i = 0; i < n; ++i) {
            const uint32_t c = i / 2, x = c % g, y = c / g, a = y * (g + 1) + x;
            if (i & 1)
              t.insert(t.end(), {a, a + 1, a + g + 1});
            else
              t.insert(t.end(), {a + 1, a + g + 2, a + g + 1});
          }
        }
        const Built B = build(xyz, t, W, leaf);
        std::vector<int> seen(n, 0), nseen(B.n_wide, 0);
        uint32_t maxd = 0, maxs = 0;
        check(B, xyz, t, W, leaf, 0, seen, nseen, 1, 0, maxd, maxs);
        for (uint32_t i = 0; i < n; ++i) CHECK(seen[i] == 1, "triangle %u in %d leaves", i, seen[i]);
        for (uint32_t i = 0; i < B.n_wide; ++i) CHECK(nseen[i] == 1, "node %u reached %d times", i, nseen[i]);
        CHECK(maxd == B.levels, "depth %u, build says %u", maxd, B.levels);
        CHECK(maxs + 1 == B.max_stack, "stack bound %u, build says %u", maxs + 1, B.max_stack);
        for (uint32_t i = 0; i < n; ++i)
          for (int k = 0; k < 3; ++k) {
            const float* v = &xyz[3 * t[3 * i + k]];
            const float4 f = B.fin[3 * i + k];
            CHECK(__float_as_uint(f.x) == __float_as_uint(v[0]) && __float_as_uint(f.y) == __float_as_uint(v[1]) &&
                      __float_as_uint(f.z) == __float_as_uint(v[2]) &&
                      __float_as_uint(f.w) == __float_as_uint(B.normals[3 * i + k]),
                  "completion record %u", i);
          }
        printf("n=%u kind=%d W=%d leaf=%d: %u nodes, depth %u, stack %u\n", n, cs.kind, W, leaf, B.n_wide, B.levels,
               B.max_stack);
      }
  printf(fails ? "FAILED %d\n" : "ALL OK\n", fails);
  return fails != 0;
}
