// bvh_build.cpp -- binned-SAH binary BVH, collapsed to 4- or 8-wide nodes and
// re-packed for the GPU (layout in bvh_build.h).
#include "bvh_build.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>

namespace j3d {
namespace {

struct Aabb {
  float lo[3], hi[3];
  void reset() {
    lo[0] = lo[1] = lo[2] = FLT_MAX;
    hi[0] = hi[1] = hi[2] = -FLT_MAX;
  }
  void grow(const Aabb& b) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], b.lo[k]);
      hi[k] = std::max(hi[k], b.hi[k]);
    }
  }
  float area() const {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (!(dx >= 0.f)) return 0.f;
    return dx * dy + dy * dz + dz * dx;
  }
};

struct BinNode {
  Aabb box;
  uint32_t a = 0, b = 0;  // internal: children ; leaf: start, count
  bool leaf = false;
};

constexpr int kBins = 32;
constexpr int kMaxSahDepth = 40;

struct Builder {
  const Aabb* pb;       // per-primitive bounds
  const float* pc;      // per-primitive centroid (3 floats)
  uint32_t* idx;
  BinNode* nodes;
  std::atomic<uint32_t> n_nodes{0};
  std::atomic<int> spare{0};
  int leaf_size = 4;
  int threads = 1;

  struct Bins {
    Aabb box[kBins];
    uint32_t cnt[kBins];
    void reset() {
      for (int i = 0; i < kBins; ++i) box[i].reset(), cnt[i] = 0;
    }
  };

  void range_bounds(uint32_t begin, uint32_t end, Aabb& box, Aabb& cbox) {
    box.reset();
    cbox.reset();
    for (uint32_t i = begin; i < end; ++i) {
      uint32_t p = idx[i];
      box.grow(pb[p]);
      for (int k = 0; k < 3; ++k) {
        cbox.lo[k] = std::min(cbox.lo[k], pc[3 * (size_t)p + k]);
        cbox.hi[k] = std::max(cbox.hi[k], pc[3 * (size_t)p + k]);
      }
    }
  }

  // Parallel over large ranges: each worker reduces a slice, then merge.
  template <class Local, class Fn, class Merge>
  void par_reduce(uint32_t begin, uint32_t end, int nthr, Local* locals, Fn fn, Merge merge) {
    uint32_t n = end - begin;
    uint32_t chunk = (n + nthr - 1) / nthr;
    std::vector<std::thread> pool;
    for (int t = 1; t < nthr; ++t) {
      uint32_t a = begin + std::min(n, t * chunk), b = begin + std::min(n, (t + 1) * chunk);
      pool.emplace_back([&, t, a, b]() { fn(locals[t], a, b); });
    }
    fn(locals[0], begin, begin + std::min(n, chunk));
    for (auto& th : pool) th.join();
    for (int t = 1; t < nthr; ++t) merge(locals[0], locals[t]);
  }

  void build(uint32_t node, uint32_t begin, uint32_t end, int depth) {
    BinNode& nd = nodes[node];
    uint32_t n = end - begin;
    Aabb cbox;
    int par = (n > (1u << 20)) ? std::min(threads, 16) : 1;
    if (par > 1) {
      std::vector<std::pair<Aabb, Aabb>> loc(par);
      for (auto& l : loc) l.first.reset(), l.second.reset();
      par_reduce(
          begin, end, par, loc.data(),
          [&](std::pair<Aabb, Aabb>& l, uint32_t a, uint32_t b) { range_bounds(a, b, l.first, l.second); },
          [](std::pair<Aabb, Aabb>& x, const std::pair<Aabb, Aabb>& y) {
            x.first.grow(y.first);
            x.second.grow(y.second);
          });
      nd.box = loc[0].first;
      cbox = loc[0].second;
    } else {
      range_bounds(begin, end, nd.box, cbox);
    }
    if ((int)n <= leaf_size) {
      nd.leaf = true;
      nd.a = begin;
      nd.b = n;
      return;
    }
    int axis = 0;
    for (int k = 1; k < 3; ++k)
      if (cbox.hi[k] - cbox.lo[k] > cbox.hi[axis] - cbox.lo[axis]) axis = k;
    float ext = cbox.hi[axis] - cbox.lo[axis];
    uint32_t mid = begin + n / 2;
    bool median = true;
    if (ext > 0.f && depth < kMaxSahDepth) {
      float scale = kBins / ext;
      float base = cbox.lo[axis];
      auto bin_of = [&](uint32_t p) {
        int b = (int)((pc[3 * (size_t)p + axis] - base) * scale);
        return b < 0 ? 0 : (b >= kBins ? kBins - 1 : b);
      };
      auto bin_range = [&](Bins& bins, uint32_t a, uint32_t b) {
        for (uint32_t i = a; i < b; ++i) {
          uint32_t p = idx[i];
          int k = bin_of(p);
          bins.cnt[k]++;
          bins.box[k].grow(pb[p]);
        }
      };
      std::vector<Bins> bins(par);
      for (auto& b : bins) b.reset();
      if (par > 1)
        par_reduce(begin, end, par, bins.data(), bin_range, [](Bins& x, const Bins& y) {
          for (int i = 0; i < kBins; ++i) x.cnt[i] += y.cnt[i], x.box[i].grow(y.box[i]);
        });
      else
        bin_range(bins[0], begin, end);
      const Bins& B = bins[0];
      float left_area[kBins];
      uint32_t left_cnt[kBins];
      Aabb acc;
      acc.reset();
      uint32_t c = 0;
      for (int i = 0; i < kBins - 1; ++i) {
        acc.grow(B.box[i]);
        c += B.cnt[i];
        left_area[i] = acc.area();
        left_cnt[i] = c;
      }
      acc.reset();
      c = 0;
      float best = FLT_MAX;
      int best_split = -1;
      for (int i = kBins - 1; i > 0; --i) {
        acc.grow(B.box[i]);
        c += B.cnt[i];
        if (left_cnt[i - 1] == 0 || c == 0) continue;
        float cost = left_area[i - 1] * (float)left_cnt[i - 1] + acc.area() * (float)c;
        if (cost < best) {
          best = cost;
          best_split = i;
        }
      }
      if (best_split > 0) {
        uint32_t* m = std::partition(idx + begin, idx + end, [&](uint32_t p) { return bin_of(p) < best_split; });
        mid = (uint32_t)(m - idx);
        median = (mid == begin || mid == end);
      }
    }
    if (median) {
      mid = begin + n / 2;
      std::nth_element(idx + begin, idx + mid, idx + end,
                       [&](uint32_t x, uint32_t y) { return pc[3 * (size_t)x + axis] < pc[3 * (size_t)y + axis]; });
    }
    uint32_t l = n_nodes.fetch_add(2);
    nd.leaf = false;
    nd.a = l;
    nd.b = l + 1;
    if (n > (1u << 15) && spare.fetch_sub(1) > 0) {
      std::thread th([this, l, begin, mid, depth]() { build(l, begin, mid, depth + 1); });
      build(l + 1, mid, end, depth + 1);
      th.join();
    } else {
      build(l, begin, mid, depth + 1);
      build(l + 1, mid, end, depth + 1);
    }
  }
};

inline void widen(const Aabb& b, float* lo, float* hi) {
  for (int k = 0; k < 3; ++k) {
    float m = std::max(std::fabs(b.lo[k]), std::fabs(b.hi[k]));
    float e = m * (1.0f / 1048576.0f) + 1e-30f;
    lo[k] = b.lo[k] - e;
    hi[k] = b.hi[k] + e;
  }
}

// Writes one quantized wide node (layout in bvh_build.h).  Child boxes are
// given widened; quantization rounds them outward (floor / ceil), so the
// decoded box always contains the widened box.
void write_node(uint32_t* nd, int W, int n, const float (*lo)[3], const float (*hi)[3], const uint32_t* refs) {
  const int words = node_words(W);
  for (int i = 0; i < words; ++i) nd[i] = 0;
  float p[3] = {0.f, 0.f, 0.f}, nh[3] = {0.f, 0.f, 0.f};
  for (int k = 0; k < 3; ++k) {
    p[k] = INFINITY;
    nh[k] = -INFINITY;
    for (int i = 0; i < n; ++i) {
      p[k] = std::min(p[k], lo[i][k]);
      nh[k] = std::max(nh[k], hi[i][k]);
    }
    if (n == 0) p[k] = nh[k] = 0.f;
  }
  uint32_t eb[3];
  uint8_t q[2][3][8];  // [lo/hi][axis][child]
  for (int k = 0; k < 3; ++k) {
    const double ext = (double)nh[k] - (double)p[k];
    int e = -126;
    if (ext > 0.0) {
      e = (int)std::ceil(std::log2(ext / 255.0));
      if (e < -126) e = -126;
      while ((double)p[k] + 255.0 * std::ldexp(1.0, e) < (double)nh[k]) ++e;
    }
    eb[k] = (uint32_t)(e + 127);
    const double sc = std::ldexp(1.0, e);
    for (int i = 0; i < 8; ++i) {
      if (i >= n) {  // empty slot: lo > hi on every axis, never entered
        q[0][k][i] = 255;
        q[1][k][i] = 0;
        continue;
      }
      double ql = std::floor(((double)lo[i][k] - (double)p[k]) / sc);
      double qh = std::ceil(((double)hi[i][k] - (double)p[k]) / sc);
      while (ql > 0 && (double)p[k] + ql * sc > (double)lo[i][k]) ql -= 1;
      while ((double)p[k] + qh * sc < (double)hi[i][k]) qh += 1;
      q[0][k][i] = (uint8_t)std::min(std::max(ql, 0.0), 255.0);
      q[1][k][i] = (uint8_t)std::min(std::max(qh, 0.0), 255.0);
    }
  }
  std::memcpy(&nd[0], &p[0], 4);
  std::memcpy(&nd[1], &p[1], 4);
  std::memcpy(&nd[2], &p[2], 4);
  nd[3] = eb[0] | (eb[1] << 8) | (eb[2] << 16);
  auto pack4 = [&](int side, int k, int first) {
    return (uint32_t)q[side][k][first] | ((uint32_t)q[side][k][first + 1] << 8) |
           ((uint32_t)q[side][k][first + 2] << 16) | ((uint32_t)q[side][k][first + 3] << 24);
  };
  if (W == 4) {
    nd[4] = pack4(0, 0, 0), nd[5] = pack4(1, 0, 0);
    nd[6] = pack4(0, 1, 0), nd[7] = pack4(1, 1, 0);
    nd[8] = pack4(0, 2, 0), nd[9] = pack4(1, 2, 0);
    for (int i = 0; i < 4; ++i) nd[10 + i] = i < n ? refs[i] : kEmptyRef;
  } else {
    for (int k = 0; k < 3; ++k) {
      nd[4 + 4 * k] = pack4(0, k, 0), nd[5 + 4 * k] = pack4(0, k, 4);
      nd[6 + 4 * k] = pack4(1, k, 0), nd[7 + 4 * k] = pack4(1, k, 4);
    }
    for (int i = 0; i < 8; ++i) nd[16 + i] = i < n ? refs[i] : kEmptyRef;
  }
}

struct Collapser {
  const BinNode* bn;
  int W;
  std::vector<uint32_t>& out;
  uint32_t n_wide = 0;
  uint32_t max_depth = 0, max_stack = 0;

  Collapser(const BinNode* b, int w, std::vector<uint32_t>& o) : bn(b), W(w), out(o) {}

  uint32_t alloc(uint32_t count) {
    uint32_t first = n_wide;
    n_wide += count;
    out.resize((size_t)n_wide * node_words(W));
    return first;
  }

  // Children of a binary node opened greedily by largest surface area until W.
  void open(uint32_t b, std::vector<uint32_t>& ch) {
    ch.clear();
    if (bn[b].leaf) {
      ch.push_back(b);
      return;
    }
    ch.push_back(bn[b].a);
    ch.push_back(bn[b].b);
    while ((int)ch.size() < W) {
      int best = -1;
      float best_area = -1.f;
      for (size_t i = 0; i < ch.size(); ++i) {
        const BinNode& c = bn[ch[i]];
        if (c.leaf) continue;
        float a = c.box.area();
        if (a > best_area) best_area = a, best = (int)i;
      }
      if (best < 0) break;
      uint32_t c = ch[best];
      ch[best] = bn[c].a;
      ch.push_back(bn[c].b);
    }
  }

  // Fills wide node `w` from binary node `b`.  depth counts wide levels;
  // stack is the bound on entries pushed by the ancestors.
  void fill(uint32_t w, uint32_t b, uint32_t depth, uint32_t stack) {
    std::vector<uint32_t> ch;
    open(b, ch);
    max_depth = std::max(max_depth, depth);
    uint32_t pushed = stack + (uint32_t)(ch.size() - 1);
    max_stack = std::max(max_stack, pushed);
    uint32_t n_int = 0;
    for (uint32_t c : ch)
      if (!bn[c].leaf) ++n_int;
    uint32_t first = n_int ? alloc(n_int) : 0;
    float lo[8][3], hi[8][3];
    uint32_t refs[8];
    std::vector<std::pair<uint32_t, uint32_t>> todo;
    uint32_t k = 0;
    for (size_t i = 0; i < ch.size(); ++i) {
      const BinNode& c = bn[ch[i]];
      widen(c.box, lo[i], hi[i]);
      if (c.leaf) {
        refs[i] = kLeafBit | ((c.b - 1u) << kLeafCountShift) | c.a;
      } else {
        refs[i] = first + k;
        todo.emplace_back(first + k, ch[i]);
        ++k;
      }
    }
    write_node(&out[(size_t)w * node_words(W)], W, (int)ch.size(), lo, hi, refs);
    for (auto& t : todo) fill(t.first, t.second, depth + 1, pushed);
  }
};

}  // namespace

bool build_bvh(const float* xyz, uint64_t n_vertices, const uint32_t* tris, uint64_t n_tris, const BuildOptions& opt,
               BuiltBVH& out, const char** err) {
  auto fail = [&](const char* m) {
    if (err) *err = m;
    return false;
  };
  if (opt.width != 4 && opt.width != 8) return fail("BVH width must be 4 or 8");
  if (opt.leaf_size < 1 || opt.leaf_size > 4) return fail("leaf size must be in [1, 4]");
  if (n_tris > kLeafIndexMask) return fail("too many triangles (max 2^29-1 per mesh)");
  int threads = opt.threads > 0 ? opt.threads : (int)std::max(1u, std::thread::hardware_concurrency());
  for (uint64_t i = 0; i < 3 * n_tris; ++i)
    if (tris[i] >= n_vertices) return fail("triangle index out of range");
  out = BuiltBVH();
  out.width = opt.width;
  const int W = opt.width;
  const uint32_t nt = (uint32_t)n_tris;

  std::vector<Aabb> pb(nt);
  std::vector<float> pc(3 * (size_t)nt);
  std::vector<uint32_t> idx(nt);
  {
    auto work = [&](uint32_t a, uint32_t b) {
      for (uint32_t i = a; i < b; ++i) {
        idx[i] = i;
        const float* v0 = xyz + 3 * (size_t)tris[3 * (size_t)i];
        const float* v1 = xyz + 3 * (size_t)tris[3 * (size_t)i + 1];
        const float* v2 = xyz + 3 * (size_t)tris[3 * (size_t)i + 2];
        for (int k = 0; k < 3; ++k) {
          pb[i].lo[k] = std::min(v0[k], std::min(v1[k], v2[k]));
          pb[i].hi[k] = std::max(v0[k], std::max(v1[k], v2[k]));
          pc[3 * (size_t)i + k] = (pb[i].lo[k] + pb[i].hi[k]) * 0.5f;
        }
      }
    };
    std::vector<std::thread> pool;
    uint32_t chunk = (nt + threads - 1) / std::max(threads, 1);
    for (int t = 0; t < threads && chunk; ++t) {
      uint32_t a = std::min(nt, t * chunk), b = std::min(nt, (t + 1) * chunk);
      if (a < b) pool.emplace_back(work, a, b);
    }
    for (auto& th : pool) th.join();
  }

  std::vector<BinNode> bn(nt ? 2 * (size_t)nt : 1);
  if (nt) {
    Builder B;
    B.pb = pb.data();
    B.pc = pc.data();
    B.idx = idx.data();
    B.nodes = bn.data();
    B.n_nodes = 1;
    B.spare = threads - 1;
    B.leaf_size = opt.leaf_size;
    B.threads = threads;
    B.build(0, 0, nt, 0);
    bn.resize(B.n_nodes.load());
    for (int k = 0; k < 3; ++k) out.lo[k] = bn[0].box.lo[k], out.hi[k] = bn[0].box.hi[k];
  }

  // wide nodes: root at index 0
  Collapser C(bn.data(), W, out.nodes);
  C.alloc(1);
  if (nt) {
    C.fill(0, 0, 1, 0);
  } else {
    write_node(&out.nodes[0], W, 0, nullptr, nullptr, nullptr);
    C.max_depth = 1;
  }
  out.n_nodes = C.n_wide;
  out.max_depth = C.max_depth;
  out.max_stack = C.max_stack;

  // triangle records in leaf order
  out.tris.assign(12 * (size_t)nt, 0.f);
  std::vector<uint8_t> last(nt, 0);
  for (const BinNode& n : bn)
    if (n.leaf && n.b) last[n.a + n.b - 1] = 1;
  {
    auto work = [&](uint32_t a, uint32_t b) {
      for (uint32_t s = a; s < b; ++s) {
        uint32_t p = idx[s];
        const float* v0 = xyz + 3 * (size_t)tris[3 * (size_t)p];
        const float* v1 = xyz + 3 * (size_t)tris[3 * (size_t)p + 1];
        const float* v2 = xyz + 3 * (size_t)tris[3 * (size_t)p + 2];
        float* r = &out.tris[12 * (size_t)s];
        r[0] = v0[0], r[1] = v0[1], r[2] = v0[2];
        std::memcpy(&r[3], &p, 4);
        r[4] = v1[0] - v0[0], r[5] = v1[1] - v0[1], r[6] = v1[2] - v0[2];
        uint32_t flag = last[s];
        std::memcpy(&r[7], &flag, 4);
        r[8] = v2[0] - v0[0], r[9] = v2[1] - v0[1], r[10] = v2[2] - v0[2];
        r[11] = 0.f;
      }
    };
    std::vector<std::thread> pool;
    uint32_t chunk = (nt + threads - 1) / std::max(threads, 1);
    for (int t = 0; t < threads && chunk; ++t) {
      uint32_t a = std::min(nt, t * chunk), b = std::min(nt, (t + 1) * chunk);
      if (a < b) pool.emplace_back(work, a, b);
    }
    for (auto& th : pool) th.join();
  }
  return true;
}

}  // namespace j3d
