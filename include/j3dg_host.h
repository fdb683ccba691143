/*
 * j3dg_host.h -- host-side inputs of the ray-casting path that the reference
 * computes on the CPU before a frame: the camera projection (j3d/camera.cpp)
 * and the procedural matcap tables (j3d/matcap.cpp).  Exported by libj3dg.so
 * next to the render ABI in j3dg.h.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* camera (j3d/camera.h:5-16) */
typedef struct j3dg_camera {
  float focal_length;   /* mm */
  float film_w_inch;
  float film_h_inch;
  float near_plane;
  float far_plane;
  float zoom;
  int32_t fit_film;     /* 0 = Fill, 1 = Overscan */
} j3dg_camera;

/* make_default_camera (j3d/camera.cpp:5-16) */
void j3dg_default_camera(j3dg_camera* c);
/* frustum (j3d/camera.cpp:18-24) */
void j3dg_frustum(float left, float right, float bottom, float top, float near_plane, float far_plane, float out[16]);
/* invert_projection_matrix (j3d/camera.cpp:26-30) */
void j3dg_invert_projection(const float in[16], float out[16]);
/* make_projection_matrix (j3d/camera.cpp:32-73) */
void j3dg_projection(const j3dg_camera* c, int screen_w, int screen_h, float out[16]);
/* rigid inverse (jtk::invert_orthonormal as used by j3d/canvas.cpp:210, j3d/scene.cpp:109) */
void j3dg_invert_orthonormal(const float m[16], float out[16]);

/* matcap tables, 512 x 512, row-major 0xAABBGGRR, and their cavity colour:
   type 0 red wax (j3d/matcap.cpp:164-239, the default), 1 gray (9-84),
   2 brown (87-162), 3 sketch (241-268).  im must hold 512*512 entries. */
int j3dg_make_matcap(int type, uint32_t* im, uint32_t* cavity);

#ifdef __cplusplus
}
#endif
