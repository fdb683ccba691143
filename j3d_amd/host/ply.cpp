// ply.cpp -- PLY reading and writing for the canvas layer (see ply.h).
#include "ply.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <sstream>

#include "j3dg_io.h"

namespace j3d_amd {

namespace {

thread_local std::string g_error;

bool fail(const std::string& msg) {
  g_error = msg;
  return false;
}

enum class Ty { i8, u8, i16, u16, i32, u32, f32, f64 };

bool parse_type(const std::string& s, Ty& t) {
  static const struct {
    const char* name;
    Ty ty;
  } names[] = {{"char", Ty::i8},    {"int8", Ty::i8},     {"uchar", Ty::u8},   {"uint8", Ty::u8},
               {"short", Ty::i16},  {"int16", Ty::i16},   {"ushort", Ty::u16}, {"uint16", Ty::u16},
               {"int", Ty::i32},    {"int32", Ty::i32},   {"uint", Ty::u32},   {"uint32", Ty::u32},
               {"float", Ty::f32},  {"float32", Ty::f32}, {"double", Ty::f64}, {"float64", Ty::f64}};
  for (const auto& n : names)
    if (s == n.name) {
      t = n.ty;
      return true;
    }
  return false;
}

size_t type_size(Ty t) {
  switch (t) {
    case Ty::i8:
    case Ty::u8:
      return 1;
    case Ty::i16:
    case Ty::u16:
      return 2;
    case Ty::i32:
    case Ty::u32:
    case Ty::f32:
      return 4;
    case Ty::f64:
      return 8;
  }
  return 0;
}

bool is_float(Ty t) { return t == Ty::f32 || t == Ty::f64; }

struct Prop {
  std::string name;
  bool list = false;
  Ty count_ty = Ty::u8, ty = Ty::f32;
};

struct Elem {
  std::string name;
  uint64_t count = 0;
  std::vector<Prop> props;
};

enum class Fmt { ascii, little, big };

// The data section: typed values, one at a time, from the binary bytes or
// from the whitespace-separated ascii tokens.  Every value is returned as a
// double, which holds each PLY scalar type exactly (a float32 token is parsed
// as float first, so its value is the float the file meant).
class Source {
 public:
  Source(Fmt f, const char* p, const char* end) : fmt_(f), p_(p), end_(end) {}
  bool read(Ty t, double& v) {
    if (fmt_ == Fmt::ascii) return read_ascii(t, v);
    const size_t n = type_size(t);
    if ((size_t)(end_ - p_) < n) return false;
    unsigned char b[8];
    std::memcpy(b, p_, n);
    p_ += n;
    if (fmt_ == Fmt::big) std::reverse(b, b + n);
    switch (t) {
      case Ty::i8: v = (double)(int8_t)b[0]; break;
      case Ty::u8: v = (double)b[0]; break;
      case Ty::i16: { int16_t x; std::memcpy(&x, b, 2); v = x; break; }
      case Ty::u16: { uint16_t x; std::memcpy(&x, b, 2); v = x; break; }
      case Ty::i32: { int32_t x; std::memcpy(&x, b, 4); v = x; break; }
      case Ty::u32: { uint32_t x; std::memcpy(&x, b, 4); v = x; break; }
      case Ty::f32: { float x; std::memcpy(&x, b, 4); v = x; break; }
      case Ty::f64: { double x; std::memcpy(&x, b, 8); v = x; break; }
    }
    return true;
  }
  // skips one property instance
  bool skip(const Prop& p) {
    double v;
    if (!p.list) return read(p.ty, v);
    if (!read(p.count_ty, v) || !(v >= 0) || v > 1e9) return false;
    const uint64_t n = (uint64_t)v;
    if (fmt_ != Fmt::ascii) {
      const uint64_t bytes = n * type_size(p.ty);
      if ((uint64_t)(end_ - p_) < bytes) return false;
      p_ += bytes;
      return true;
    }
    for (uint64_t i = 0; i < n; ++i)
      if (!read(p.ty, v)) return false;
    return true;
  }

 private:
  bool read_ascii(Ty t, double& v) {
    while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
    if (p_ == end_) return false;
    const char* s = p_;
    while (p_ < end_ && !(*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
    const std::string tok(s, p_);
    char* e = nullptr;
    if (t == Ty::f32) {
      v = (double)std::strtof(tok.c_str(), &e);
    } else if (t == Ty::f64) {
      v = std::strtod(tok.c_str(), &e);
    } else {
      const long long x = std::strtoll(tok.c_str(), &e, 10);
      v = (double)x;
    }
    return e && *e == 0;
  }
  Fmt fmt_;
  const char* p_;
  const char* end_;
};

bool read_file(const char* filename, std::vector<char>& buf) {
  FILE* f = std::fopen(filename, "rb");
  if (!f) return fail(std::string("cannot open ") + filename);
  std::unique_ptr<FILE, int (*)(FILE*)> guard(f, &std::fclose);
  if (std::fseek(f, 0, SEEK_END) != 0) return fail("cannot seek");
  const long size = std::ftell(f);
  if (size < 0) return fail("cannot size the file");
  if (std::fseek(f, 0, SEEK_SET) != 0) return fail("cannot seek");
  buf.resize((size_t)size);
  if (size && std::fread(buf.data(), 1, (size_t)size, f) != (size_t)size) return fail("short read");
  return true;
}

int find(const Elem& e, std::initializer_list<const char*> names) {
  for (const char* n : names)
    for (size_t i = 0; i < e.props.size(); ++i)
      if (!e.props[i].list && e.props[i].name == n) return (int)i;
  return -1;
}

int find_list(const Elem& e, std::initializer_list<const char*> names) {
  for (const char* n : names)
    for (size_t i = 0; i < e.props.size(); ++i)
      if (e.props[i].list && e.props[i].name == n) return (int)i;
  return -1;
}

uint32_t channel(double v, Ty t) {
  if (is_float(t)) {
    const double c = std::min(1.0, std::max(0.0, v)) * 255.0 + 0.5;
    return (uint32_t)c;
  }
  return (uint32_t)std::min(255.0, std::max(0.0, v));
}

}  // namespace

const std::string& ply_error() { return g_error; }

bool read_ply(const char* filename, std::vector<float>& vertices, std::vector<float>& normals,
              std::vector<uint32_t>& clrs, std::vector<uint32_t>& triangles, std::vector<float>& uv) {
  g_error.clear();
  vertices.clear(), normals.clear(), clrs.clear(), triangles.clear(), uv.clear();
  std::vector<char> buf;
  if (!read_file(filename, buf)) return false;

  // ---- header
  const char* p = buf.data();
  const char* end = p + buf.size();
  auto next_line = [&](std::string& line) {
    if (p >= end) return false;
    const char* s = p;
    while (p < end && *p != '\n') ++p;
    line.assign(s, p);
    if (p < end) ++p;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    return true;
  };
  std::string line;
  if (!next_line(line) || line != "ply") return fail("not a PLY file (no 'ply' magic)");
  Fmt fmt = Fmt::ascii;
  bool have_format = false, have_end = false;
  std::vector<Elem> elems;
  while (next_line(line)) {
    std::istringstream ls(line);
    std::string kw;
    ls >> kw;
    if (kw.empty() || kw == "comment" || kw == "obj_info") continue;
    if (kw == "format") {
      std::string f, ver;
      ls >> f >> ver;
      if (f == "ascii") fmt = Fmt::ascii;
      else if (f == "binary_little_endian") fmt = Fmt::little;
      else if (f == "binary_big_endian") fmt = Fmt::big;
      else return fail("unknown PLY format '" + f + "'");
      have_format = true;
    } else if (kw == "element") {
      Elem e;
      long long n = -1;
      ls >> e.name >> n;
      if (e.name.empty() || n < 0) return fail("malformed element line: " + line);
      e.count = (uint64_t)n;
      elems.push_back(e);
    } else if (kw == "property") {
      if (elems.empty()) return fail("property before any element");
      Prop pr;
      std::string t;
      ls >> t;
      if (t == "list") {
        std::string ct, it;
        ls >> ct >> it >> pr.name;
        pr.list = true;
        if (!parse_type(ct, pr.count_ty) || !parse_type(it, pr.ty) || is_float(pr.count_ty))
          return fail("bad list property: " + line);
      } else {
        ls >> pr.name;
        if (!parse_type(t, pr.ty)) return fail("bad property type: " + line);
      }
      if (pr.name.empty()) return fail("property without a name: " + line);
      elems.back().props.push_back(pr);
    } else if (kw == "end_header") {
      have_end = true;
      break;
    } else {
      return fail("unknown header line: " + line);
    }
  }
  if (!have_format) return fail("no format line");
  if (!have_end) return fail("no end_header");

  uint64_t n_vertices = 0;
  for (const Elem& e : elems)
    if (e.name == "vertex") n_vertices = e.count;
  if (n_vertices > 0xFFFFFFFFull) return fail("too many vertices");
  // every element instance takes at least one byte per property (binary: its
  // size; ascii: a digit and a separator), so counts the file cannot hold are
  // refused before anything is allocated for them
  {
    const uint64_t left = (uint64_t)(end - p);
    uint64_t need = 0;
    for (const Elem& e : elems) {
      uint64_t rec = 0;
      for (const Prop& pr : e.props)
        rec += fmt == Fmt::ascii ? 2 : pr.list ? type_size(pr.count_ty) : type_size(pr.ty);
      if (rec && e.count > left / rec) return fail("element " + e.name + ": more instances than the file holds");
      need += e.count * rec;
      if (need > left) return fail("the header describes more data than the file holds");
    }
  }

  // ---- data
  Source src(fmt, p, end);
  std::vector<float> vuv;  // per-vertex texture coordinates, when the file has them
  std::vector<double> val;
  for (const Elem& e : elems) {
    if (e.name == "vertex") {
      const int ix = find(e, {"x"}), iy = find(e, {"y"}), iz = find(e, {"z"});
      if (ix < 0 || iy < 0 || iz < 0) return fail("vertex element without x, y, z");
      const int inx = find(e, {"nx"}), iny = find(e, {"ny"}), inz = find(e, {"nz"});
      const bool has_n = inx >= 0 && iny >= 0 && inz >= 0;
      const int ir = find(e, {"red", "r"}), ig = find(e, {"green", "g"}), ib = find(e, {"blue", "b"});
      const int ia = find(e, {"alpha", "a"});
      const bool has_c = ir >= 0 && ig >= 0 && ib >= 0;
      const int iu = find(e, {"s", "u", "texture_u"}), iv = find(e, {"t", "v", "texture_v"});
      const bool has_uv = iu >= 0 && iv >= 0;
      vertices.resize(3 * e.count);
      if (has_n) normals.resize(3 * e.count);
      if (has_c) clrs.resize(e.count);
      if (has_uv) vuv.resize(2 * e.count);
      val.assign(e.props.size(), 0.0);
      for (uint64_t i = 0; i < e.count; ++i) {
        for (size_t k = 0; k < e.props.size(); ++k) {
          if (e.props[k].list) {
            if (!src.skip(e.props[k])) return fail("truncated or malformed vertex data");
          } else if (!src.read(e.props[k].ty, val[k])) {
            return fail("truncated or malformed vertex data");
          }
        }
        vertices[3 * i] = (float)val[ix], vertices[3 * i + 1] = (float)val[iy], vertices[3 * i + 2] = (float)val[iz];
        if (has_n)
          normals[3 * i] = (float)val[inx], normals[3 * i + 1] = (float)val[iny], normals[3 * i + 2] = (float)val[inz];
        if (has_c) {
          const uint32_t a = ia >= 0 ? channel(val[ia], e.props[ia].ty) : 255u;
          clrs[i] = channel(val[ir], e.props[ir].ty) | channel(val[ig], e.props[ig].ty) << 8 |
                    channel(val[ib], e.props[ib].ty) << 16 | a << 24;
        }
        if (has_uv) vuv[2 * i] = (float)val[iu], vuv[2 * i + 1] = (float)val[iv];
      }
    } else if (e.name == "face") {
      const int il = find_list(e, {"vertex_indices", "vertex_index"});
      if (il < 0) return fail("face element without vertex_indices");
      const int it = find_list(e, {"texcoord"});
      std::vector<uint32_t> corners;
      std::vector<float> tc;
      bool face_uv = it >= 0;
      for (uint64_t f = 0; f < e.count; ++f) {
        corners.clear();
        tc.clear();
        for (size_t k = 0; k < e.props.size(); ++k) {
          const Prop& pr = e.props[k];
          if ((int)k == il || (int)k == it) {
            double cnt;
            if (!src.read(pr.count_ty, cnt) || !(cnt >= 0) || cnt > 1e9) return fail("truncated or malformed face data");
            const uint64_t n = (uint64_t)cnt;
            for (uint64_t j = 0; j < n; ++j) {
              double x;
              if (!src.read(pr.ty, x)) return fail("truncated or malformed face data");
              if ((int)k == il) {
                if (!(x >= 0) || x >= (double)n_vertices)
                  return fail("face " + std::to_string(f) + " refers to vertex " + std::to_string((long long)x) +
                              " of " + std::to_string(n_vertices));
                corners.push_back((uint32_t)x);
              } else {
                tc.push_back((float)x);
              }
            }
          } else if (!src.skip(pr)) {
            return fail("truncated or malformed face data");
          }
        }
        if (corners.size() < 3) continue;
        if (face_uv && tc.size() != 2 * corners.size())
          return fail("face " + std::to_string(f) + ": texcoord list of " + std::to_string(tc.size()) +
                      " values for " + std::to_string(corners.size()) + " corners");
        for (size_t j = 1; j + 1 < corners.size(); ++j) {
          triangles.push_back(corners[0]);
          triangles.push_back(corners[j]);
          triangles.push_back(corners[j + 1]);
          if (face_uv) {
            for (size_t c : {(size_t)0, j, j + 1}) {
              uv.push_back(tc[2 * c]);
              uv.push_back(tc[2 * c + 1]);
            }
          }
        }
      }
    } else {
      for (uint64_t i = 0; i < e.count; ++i)
        for (const Prop& pr : e.props)
          if (!src.skip(pr)) return fail("truncated or malformed data in element " + e.name);
    }
  }
  if (uv.empty() && !vuv.empty()) {
    uv.resize(2 * triangles.size());
    for (size_t i = 0; i < triangles.size(); ++i) {
      uv[2 * i] = vuv[2 * (size_t)triangles[i]];
      uv[2 * i + 1] = vuv[2 * (size_t)triangles[i] + 1];
    }
  }
  return true;
}

bool write_ply(const char* filename, const std::vector<float>& vertices, const std::vector<float>& normals,
               const std::vector<uint32_t>& clrs, const std::vector<uint32_t>& triangles,
               const std::vector<float>& uv, bool ascii) {
  g_error.clear();
  if (vertices.size() % 3) return fail("vertices: not 3 floats per vertex");
  const size_t nv = vertices.size() / 3;
  if (!normals.empty() && normals.size() != vertices.size()) return fail("normals: not 3 floats per vertex");
  if (!clrs.empty() && clrs.size() != nv) return fail("colours: not one per vertex");
  if (triangles.size() % 3) return fail("triangles: not 3 indices per triangle");
  const size_t nt = triangles.size() / 3;
  if (!uv.empty() && uv.size() != 6 * nt) return fail("uv: not 6 floats per triangle");
  for (uint32_t i : triangles)
    if (i >= nv) return fail("a triangle refers to vertex " + std::to_string(i) + " of " + std::to_string(nv));
  FILE* f = std::fopen(filename, "wb");
  if (!f) return fail(std::string("cannot create ") + filename);
  std::unique_ptr<FILE, int (*)(FILE*)> guard(f, &std::fclose);
  std::string h = "ply\n";
  h += ascii ? "format ascii 1.0\n" : "format binary_little_endian 1.0\n";
  h += "comment written by j3d_amd write_ply\n";
  h += "element vertex " + std::to_string(nv) + "\n";
  h += "property float x\nproperty float y\nproperty float z\n";
  if (!normals.empty()) h += "property float nx\nproperty float ny\nproperty float nz\n";
  if (!clrs.empty()) h += "property uchar red\nproperty uchar green\nproperty uchar blue\nproperty uchar alpha\n";
  h += "element face " + std::to_string(nt) + "\n";
  h += "property list uchar int vertex_indices\n";
  if (!uv.empty()) h += "property list uchar float texcoord\n";
  h += "end_header\n";
  bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size();
  if (ascii) {
    for (size_t i = 0; i < nv && ok; ++i) {
      ok = std::fprintf(f, "%.9g %.9g %.9g", vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]) > 0;
      if (!normals.empty())
        ok = ok && std::fprintf(f, " %.9g %.9g %.9g", normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]) > 0;
      if (!clrs.empty())
        ok = ok && std::fprintf(f, " %u %u %u %u", clrs[i] & 255u, clrs[i] >> 8 & 255u, clrs[i] >> 16 & 255u,
                                clrs[i] >> 24) > 0;
      ok = ok && std::fputc('\n', f) != EOF;
    }
    for (size_t t = 0; t < nt && ok; ++t) {
      ok = std::fprintf(f, "3 %u %u %u", triangles[3 * t], triangles[3 * t + 1], triangles[3 * t + 2]) > 0;
      if (!uv.empty()) {
        ok = ok && std::fprintf(f, " 6") > 0;
        for (int k = 0; k < 6 && ok; ++k) ok = std::fprintf(f, " %.9g", uv[6 * t + k]) > 0;
      }
      ok = ok && std::fputc('\n', f) != EOF;
    }
  } else {
    // little-endian host: the records are written as they lie in memory
    std::vector<char> rec;
    rec.reserve(28);
    for (size_t i = 0; i < nv && ok; ++i) {
      rec.clear();
      auto put = [&](const void* d, size_t n) { rec.insert(rec.end(), (const char*)d, (const char*)d + n); };
      put(&vertices[3 * i], 12);
      if (!normals.empty()) put(&normals[3 * i], 12);
      if (!clrs.empty()) put(&clrs[i], 4);  // bytes r, g, b, a of 0xAABBGGRR
      ok = std::fwrite(rec.data(), 1, rec.size(), f) == rec.size();
    }
    for (size_t t = 0; t < nt && ok; ++t) {
      rec.clear();
      const unsigned char three = 3, six = 6;
      rec.push_back((char)three);
      const int32_t idx[3] = {(int32_t)triangles[3 * t], (int32_t)triangles[3 * t + 1], (int32_t)triangles[3 * t + 2]};
      rec.insert(rec.end(), (const char*)idx, (const char*)idx + 12);
      if (!uv.empty()) {
        rec.push_back((char)six);
        rec.insert(rec.end(), (const char*)&uv[6 * t], (const char*)&uv[6 * t] + 24);
      }
      ok = std::fwrite(rec.data(), 1, rec.size(), f) == rec.size();
    }
  }
  if (!ok) return fail("write failed");
  if (std::fflush(f) != 0) return fail("write failed");
  return true;
}

}  // namespace j3d_amd

// ---- C ABI (include/j3dg_io.h) ----------------------------------------------

namespace {

template <class T>
T* copy_out(const std::vector<T>& v) {
  if (v.empty()) return nullptr;
  T* p = static_cast<T*>(std::malloc(v.size() * sizeof(T)));
  if (p) std::memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

}  // namespace

extern "C" {

int j3dg_read_ply(const char* filename, j3dg_mesh_arrays* out) {
  if (!out) return 1;
  std::memset(out, 0, sizeof(*out));
  if (!filename) {
    j3d_amd::g_error = "null file name";
    return 1;
  }
  std::vector<float> v, n, uv;
  std::vector<uint32_t> c, t;
  try {
    if (!j3d_amd::read_ply(filename, v, n, c, t, uv)) return 1;
  } catch (const std::exception& e) {
    j3d_amd::g_error = std::string("read_ply: ") + e.what();
    return 1;
  }
  out->n_vertices = v.size() / 3;
  out->n_triangles = t.size() / 3;
  out->vertices = copy_out(v);
  out->normals = copy_out(n);
  out->colors = copy_out(c);
  out->triangles = copy_out(t);
  out->uv = copy_out(uv);
  if ((!v.empty() && !out->vertices) || (!n.empty() && !out->normals) || (!c.empty() && !out->colors) ||
      (!t.empty() && !out->triangles) || (!uv.empty() && !out->uv)) {
    j3dg_free_mesh_arrays(out);
    j3d_amd::g_error = "out of memory";
    return 1;
  }
  return 0;
}

void j3dg_free_mesh_arrays(j3dg_mesh_arrays* m) {
  if (!m) return;
  std::free(m->vertices);
  std::free(m->normals);
  std::free(m->colors);
  std::free(m->triangles);
  std::free(m->uv);
  std::memset(m, 0, sizeof(*m));
}

int j3dg_write_ply(const char* filename, const j3dg_mesh_arrays* m, int ascii) {
  if (!filename || !m) {
    j3d_amd::g_error = "null argument";
    return 1;
  }
  if ((m->n_vertices && !m->vertices) || (m->n_triangles && !m->triangles)) {
    j3d_amd::g_error = "null vertex or triangle array";
    return 1;
  }
  const size_t nv = (size_t)m->n_vertices, nt = (size_t)m->n_triangles;
  try {
    std::vector<float> v(m->vertices, m->vertices + 3 * nv);
    std::vector<float> n = m->normals ? std::vector<float>(m->normals, m->normals + 3 * nv) : std::vector<float>();
    std::vector<uint32_t> c = m->colors ? std::vector<uint32_t>(m->colors, m->colors + nv) : std::vector<uint32_t>();
    std::vector<uint32_t> t(m->triangles, m->triangles + 3 * nt);
    std::vector<float> uv = m->uv ? std::vector<float>(m->uv, m->uv + 6 * nt) : std::vector<float>();
    return j3d_amd::write_ply(filename, v, n, c, t, uv, ascii != 0) ? 0 : 1;
  } catch (const std::exception& e) {
    j3d_amd::g_error = std::string("write_ply: ") + e.what();
    return 1;
  }
}

const char* j3dg_ply_error(void) { return j3d_amd::g_error.c_str(); }

}  // extern "C"
