// render_frame.cpp -- headless driver of the canvas interface: builds a seeded
// scene, renders it the way j3d's view does (view.cpp:421-430) and writes the
// G-buffer (32 B per pixel) and RGBA image as raw little-endian files.
//
//   j3d_render_frame --scene icosphere LEVEL SEED AMP | --scene torus NU NV SEED | --scene ply FILE
//                    --size W H [--focal MM] [--cs 16 floats] [--flags edges,shadow,...]
//                    [--light X Y Z] [--matcap 0..3] [--bg TOP BOTTOM] [--rect X0 Y0 X1 Y1]
//                    [--frames N] [--learned MARGIN] [--copies N SPACING] [--points N SEED] [--device-build] [--out PREFIX]
// --scene ply FILE: the mesh read by read_ply (j3d/io.cpp:733-736), as a viewer loads one
// --copies: N objects of the mesh, object k translated by k * SPACING along x
// --points N SEED: a point cloud of N points near the mesh, drawn after the shading
// (db_id 2^29 + k).
// Prints one JSON line (sizes, times).  Without --cs the scene is framed by
// prepare_scene + unzoom (j3d/scene.cpp:50-110).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "canvas.h"
#include "j3dg_scenegen.h"
#include "ply.h"

using namespace j3d_amd;

namespace {

[[noreturn]] void usage(const char* msg) {
  std::fprintf(stderr, "j3d_render_frame: %s\n", msg);
  std::exit(2);
}

bool has_flag(const std::string& list, const char* name) {
  size_t pos = 0;
  const std::string n(name);
  while (pos <= list.size()) {
    size_t e = list.find(',', pos);
    if (e == std::string::npos) e = list.size();
    if (list.compare(pos, e - pos, n) == 0) return true;
    pos = e + 1;
  }
  return false;
}

bool write_file(const std::string& path, const void* data, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const bool ok = std::fwrite(data, 1, bytes, f) == bytes;
  return std::fclose(f) == 0 && ok;
}

}  // namespace

int main(int argc, char** argv) {
  std::string kind, flags = "edges,shading", out, ply_file;
  uint32_t level = 7, nu = 0, nv = 0, w = 0, h = 0, frames = 1, mc = 0, copies = 1, n_points = 0;
  uint64_t point_seed = 0;
  float spacing = 0.f;
  uint32_t builder = J3DG_BUILD_HOST;
  uint64_t seed = 1;
  float amp = 0.05f, focal = 0.f, margin = -1.f;
  float cs[16];
  bool have_cs = false, have_light = false, have_rect = false;
  float light[3] = {0, 0, 1};
  uint32_t bg[2] = {0xff000000u, 0xff404040u};
  int rect[4] = {0, 0, 0, 0};
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto need = [&](int n) {
      if (i + n >= argc) usage(("missing values after " + a).c_str());
    };
    if (a == "--scene") {
      need(1);
      kind = argv[++i];
      if (kind == "icosphere") {
        need(3);
        level = (uint32_t)std::strtoul(argv[i + 1], nullptr, 10);
        seed = std::strtoull(argv[i + 2], nullptr, 10);
        amp = std::strtof(argv[i + 3], nullptr);
        i += 3;
      } else if (kind == "torus") {
        need(3);
        nu = (uint32_t)std::strtoul(argv[i + 1], nullptr, 10);
        nv = (uint32_t)std::strtoul(argv[i + 2], nullptr, 10);
        seed = std::strtoull(argv[i + 3], nullptr, 10);
        i += 3;
      } else if (kind == "ply") {
        need(1);
        ply_file = argv[++i];
      } else {
        usage("scene must be icosphere, torus or ply");
      }
    } else if (a == "--size") {
      need(2);
      w = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
      h = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
    } else if (a == "--focal") {
      need(1);
      focal = std::strtof(argv[++i], nullptr);
    } else if (a == "--cs") {
      need(16);
      for (int k = 0; k < 16; ++k) cs[k] = std::strtof(argv[++i], nullptr);
      have_cs = true;
    } else if (a == "--flags") {
      need(1);
      flags = argv[++i];
    } else if (a == "--light") {
      need(3);
      for (int k = 0; k < 3; ++k) light[k] = std::strtof(argv[++i], nullptr);
      have_light = true;
    } else if (a == "--matcap") {
      need(1);
      mc = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
    } else if (a == "--bg") {
      need(2);
      bg[0] = (uint32_t)std::strtoul(argv[++i], nullptr, 0);
      bg[1] = (uint32_t)std::strtoul(argv[++i], nullptr, 0);
    } else if (a == "--rect") {
      need(4);
      for (int k = 0; k < 4; ++k) rect[k] = (int)std::strtol(argv[++i], nullptr, 10);
      have_rect = true;
    } else if (a == "--frames") {
      need(1);
      frames = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
    } else if (a == "--learned") {
      need(1);
      margin = std::strtof(argv[++i], nullptr);
    } else if (a == "--copies") {
      need(2);
      copies = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
      spacing = std::strtof(argv[++i], nullptr);
      if (!copies) usage("--copies needs at least one object");
    } else if (a == "--points") {
      need(2);
      n_points = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
      point_seed = std::strtoull(argv[++i], nullptr, 10);
    } else if (a == "--device-build") {
      builder = J3DG_BUILD_DEVICE;
    } else if (a == "--out") {
      need(1);
      out = argv[++i];
    } else {
      usage(("unknown argument " + a).c_str());
    }
  }
  if (kind.empty() || !w || !h || !frames) usage("need --scene and --size");

  try {
    std::vector<float> xyz;
    std::vector<uint32_t> tri;
    if (kind == "ply") {
      std::vector<float> normals, uv;
      std::vector<uint32_t> clrs;
      if (!read_ply(ply_file.c_str(), xyz, normals, clrs, tri, uv)) usage(("read_ply: " + ply_error()).c_str());
    } else if (kind == "icosphere") {
      xyz.resize(3 * j3dg_icosphere_vertex_count(level));
      tri.resize(3 * j3dg_icosphere_triangle_count(level));
      if (j3dg_gen_icosphere(level, seed, amp, xyz.data(), tri.data())) usage("icosphere generation failed");
    } else {
      xyz.resize(3 * (size_t)nu * nv);
      tri.resize(6 * (size_t)nu * nv);
      if (j3dg_gen_torus(nu, nv, 1.f, 0.35f, seed, 0.02f, 6, 2.f, 0.5f, 4.f, 0, xyz.data(), tri.data()))
        usage("torus generation failed");
    }
    auto dev = std::make_shared<device>(0);
    scene sc(dev);
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t k = 0; k < copies; ++k) {
      float4x4 ocs = get_identity();
      ocs[12] = (float)k * spacing;
      add_object((1u << 29) + k, sc, xyz, tri, ocs, 8, builder);
    }
    // --points: one point cloud of mesh vertices (cycled) pushed off the surface by
    // an LCG of the seed, with normals and colours; written to PREFIX.points
    std::vector<float> pxyz, pnrm;
    std::vector<uint32_t> pcol;
    if (n_points) {
      uint64_t r = point_seed * 6364136223846793005ull + 1442695040888963407ull;
      auto next = [&]() {
        r = r * 6364136223846793005ull + 1442695040888963407ull;
        return (uint32_t)(r >> 33);
      };
      const size_t nvx = xyz.size() / 3;
      for (uint32_t i = 0; i < n_points; ++i) {
        const size_t j = (size_t)next() % nvx;
        for (int k = 0; k < 3; ++k) {
          pxyz.push_back(xyz[3 * j + k] + ((float)(next() & 0xFFFF) / 65535.f - 0.5f) * 0.04f);
          pnrm.push_back((float)(next() & 0xFFFF) / 32767.5f - 1.f);
        }
        pcol.push_back(next() * 2654435761u);
      }
      add_pointcloud((2u << 29) + 1, sc, pxyz, &pnrm, &pcol);
      if (!out.empty()) {
        std::vector<float> rec(pxyz);
        rec.insert(rec.end(), pnrm.begin(), pnrm.end());
        if (!write_file(out + ".points", rec.data(), rec.size() * sizeof(float)) ||
            !write_file(out + ".colors", pcol.data(), pcol.size() * sizeof(uint32_t)))
          usage("cannot write point files");
      }
    }
    const auto t1 = std::chrono::steady_clock::now();
    prepare_scene(sc);
    unzoom(sc);
    if (have_cs) {
      std::memcpy(sc.coordinate_system.data(), cs, sizeof(cs));
      sc.coordinate_system_inv = invert_orthonormal(sc.coordinate_system);
    }
    canvas cv(dev, w, h);
    if (focal > 0.f) {
      camera c = make_default_camera();
      c.focal_length = focal;
      cv.set_camera(c);
    }
    canvas::canvas_settings st{};
    st.one_bit = has_flag(flags, "one_bit");
    st.shadow = has_flag(flags, "shadow");
    st.edges = has_flag(flags, "edges");
    st.wireframe = has_flag(flags, "wireframe");
    st.shading = has_flag(flags, "shading");
    cv.update_settings(st);
    if (has_flag(flags, "directional") || have_light) cv.set_directional_light(true, light);
    cv.set_background_color(bg[0], bg[1]);  // as view.cpp:121 does at start-up
    if (margin >= 0.f) {
      j3dg_train_report rep;
      dev->check(j3dg_train_proposal(dev->ctx(), sc.objects.front().slot, nullptr, &rep), "j3dg_train_proposal");
      cv.set_learned(true, margin);
    }
    const matcap m = mc == 1 ? make_matcap_gray() : mc == 2 ? make_matcap_brown() : mc == 3 ? make_matcap_sketch()
                                                                                     : make_matcap_red_wax();
    // view::render_scene (j3d/view.cpp:421-430), `frames` times
    double wall = 0.0;
    image<pixel> pixels;
    for (uint32_t f = 0; f < frames; ++f) {
      const auto a = std::chrono::steady_clock::now();
      if (have_rect) {
        cv.update_canvas(rect[0], rect[1], rect[2], rect[3], sc);
      } else {
        cv.render_scene(&sc);
      }
      pixels = cv.get_pixels();
      cv.canvas_to_image(pixels, m);
      cv.render_pointclouds_on_image(&sc, pixels);
      wall += std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
    }
    float tr = 0, sh = 0, ml = 0, sd = 0;
    cv.last_timings(&tr, &sh, &ml, &sd);
    if (!out.empty()) {
      const auto& g = cv.get_pixels();
      const auto& im = cv.get_image();
      if (!write_file(out + ".gbuf", g.data(), (size_t)w * h * sizeof(pixel)) ||
          !write_file(out + ".rgba", im.data(), (size_t)w * h * sizeof(uint32_t)))
        usage("cannot write output files");
    }
    std::printf(
        "{\"width\": %u, \"height\": %u, \"triangles\": %zu, \"frames\": %u, \"upload_s\": %.3f, "
        "\"ms_per_frame_with_readback\": %.3f, \"trace_ms\": %.4f, \"shadow_ms\": %.4f, \"mlp_ms\": %.4f, "
        "\"shade_ms\": %.4f, \"diagonal\": %.9g, \"pivot\": [%.9g, %.9g, %.9g]}\n",
        w, h, tri.size() / 3, frames, std::chrono::duration<double>(t1 - t0).count(), wall / frames * 1e3, tr, sd, ml,
        sh, sc.diagonal, sc.pivot[0], sc.pivot[1], sc.pivot[2]);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "j3d_render_frame: %s\n", e.what());
    return 1;
  }
  return 0;
}
