/*
 * j3dg_io.h -- mesh files for the canvas layer (libj3d_canvas.so): the PLY
 * calls of j3d's io layer, read_ply / write_ply (j3d/io.h, j3d/io.cpp:733-741,
 * which forward to jtk's PLY code).  A viewer loads a mesh file here, then
 * hands the arrays to add_object (j3d/scene.cpp:8-39) as the reference's
 * mesh loader does.  The C++ calls are in j3d_amd/host/ply.h.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A mesh as the reference's io layer returns it, as flat arrays:
   vertices 3 floats per vertex; normals 3 per vertex or NULL; colors one
   0xAABBGGRR per vertex or NULL; triangles 3 indices per triangle; uv 6
   floats per triangle (u, v of each corner) or NULL. */
typedef struct j3dg_mesh_arrays {
  uint64_t n_vertices, n_triangles;
  float* vertices;
  float* normals;
  uint32_t* colors;
  uint32_t* triangles;
  float* uv;
} j3dg_mesh_arrays;

/* read_ply (j3d/io.cpp:733-736): 0 on success, with the arrays allocated by
   the library (free them with j3dg_free_mesh_arrays); 1 on failure (the
   reference's `false`), with j3dg_ply_error() saying why.  ascii,
   binary_little_endian and binary_big_endian; polygons are split into
   triangle fans. */
int j3dg_read_ply(const char* filename, j3dg_mesh_arrays* out);
void j3dg_free_mesh_arrays(j3dg_mesh_arrays* m);
/* write_ply (j3d/io.cpp:738-741): binary_little_endian, or ascii when ascii != 0.
   0 on success, 1 on failure (j3dg_ply_error()). */
int j3dg_write_ply(const char* filename, const j3dg_mesh_arrays* m, int ascii);
/* why the calling thread's last j3dg_read_ply / j3dg_write_ply failed */
const char* j3dg_ply_error(void);

#ifdef __cplusplus
}
#endif
