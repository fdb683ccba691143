/*
 * j3dg_scenegen.h -- seeded synthetic meshes and cameras for the ray-casting
 * workloads (no counterpart in the reference, which loads meshes from files
 * through j3d/mesh.cpp:76 read_from_file).  Deterministic for a given seed:
 * every computation runs in double precision on the host and is rounded to
 * float32 once.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Icosahedron subdivided `level` times by edge midpoints re-projected to the
   unit sphere: 10*4^level+2 vertices, 20*4^level triangles. */
uint64_t j3dg_icosphere_vertex_count(uint32_t level);
uint64_t j3dg_icosphere_triangle_count(uint32_t level);
/* Each vertex is displaced radially: r = 1 + amplitude * mean(d_i, d_j for j in
   the 1-ring), d ~ N(0,1) from SplitMix64(seed) and Box-Muller. */
int j3dg_gen_icosphere(uint32_t level, uint64_t seed, float amplitude, float* xyz, uint32_t* tris);

/* Torus around the z axis, major radius R, tube radius r, an nu x nv vertex
   grid (2*nu*nv triangles), displaced along the tube normal by fBm gradient
   noise: `octaves` octaves, first amplitude `amplitude`, base frequency
   `base_freq`, frequency x lacunarity and amplitude x gain per octave. */
int j3dg_gen_torus(uint32_t nu, uint32_t nv, float R, float r, uint64_t seed, float amplitude, uint32_t octaves,
                   float lacunarity, float gain, float base_freq, int threads, float* xyz, uint32_t* tris);

/* Rigid camera-to-world matrix (scene::coordinate_system) of a camera at
   `target + dist * (cos(el)cos(az), cos(el)sin(az), sin(el))` looking at
   `target`, with +z as world up, and its inverse. */
int j3dg_orbit_camera(float dist, float azimuth, float elevation, const float target[3], float cs[16],
                      float cs_inv[16]);

/* SplitMix64 stream helpers (deterministic seeds for cameras, lights, sizes). */
uint64_t j3dg_splitmix64(uint64_t* state);
double j3dg_uniform01(uint64_t* state); /* in (0, 1] */
/* Unit vector with z >= 0, uniformly distributed on the upper hemisphere. */
void j3dg_random_hemisphere(uint64_t seed, float out[3]);

#ifdef __cplusplus
}
#endif
