/*
 * rast_scene.h -- deterministic synthetic scenes, textures and camera paths.
 *
 * Host-only helpers that generate the benchmark inputs (the reference loads
 * OBJ/PNG assets through its app shell instead: CCR/ObjParser.cpp,
 * CCR/CUDATexture.cpp:24-72).  Every generator uses only IEEE-754 basic
 * operations and square roots plus its own polynomial sine/cosine, compiled
 * without FMA contraction, so the same call yields the same bytes on any x86
 * host.  Vertices use the 56-byte IVertex layout of rast.h.
 *
 * Calling convention: pass NULL output pointers to query the sizes first.
 */
#ifndef RAST_SCENE_H
#define RAST_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* UV sphere with `slices` x `stacks` segments: a one-triangle-per-slice fan at
 * the top pole and two triangles per quad in every other stack (the bottom
 * stack's pole triangles are zero-area and are culled), giving
 * slices * (2 * stacks - 1) triangles; 64 x 32 -> 4,032. */
int rsc_uv_sphere(uint32_t slices, uint32_t stacks, float radius,
                  float *verts, uint32_t *num_verts, uint32_t *indices, uint32_t *num_indices);

/* Icosphere of `subdiv` midpoint subdivisions with shared vertices:
 * 20 * 4^s triangles, 10 * 4^s + 2 vertices (s = 5 -> 20,480 / 10,242). */
int rsc_icosphere(uint32_t subdiv, float radius,
                  float *verts, uint32_t *num_verts, uint32_t *indices, uint32_t *num_indices);

/* Heightmap terrain of n x n quads (2 n^2 triangles, (n+1)^2 vertices) over
 * [-size/2, size/2]^2 in x/z, heights = amplitude * fBm(octaves) value noise. */
int rsc_terrain(uint32_t n, float size, float amplitude, uint32_t octaves, uint32_t seed,
                float uv_tiling, float *verts, uint32_t *num_verts,
                uint32_t *indices, uint32_t *num_indices);

/* RGBA8 checker (memory order R,G,B,A). */
int rsc_texture_checker(uint32_t w, uint32_t h, uint32_t square, uint32_t rgba0, uint32_t rgba1,
                        uint32_t *texels);

/* Value-noise maps: kind 0 diffuse (colour), 1 tangent-space normal map,
 * 2 specular (grey), 3 gloss (grey). */
int rsc_texture_noise(uint32_t w, uint32_t h, uint32_t seed, int kind, uint32_t *texels);

/* World matrix (reference layout data[col][row]) = T(t) * R(euler zyx) * S(s). */
void rsc_world_matrix(const float translate[3], const float euler_rad[3], float scale, float out[16]);

/* Seeded uniform float in [lo, hi) from (seed, index). */
float rsc_uniform(uint32_t seed, uint32_t index, float lo, float hi);

/* Deterministic sine / cosine (double precision polynomial). */
double rsc_sin(double x);
double rsc_cos(double x);

/* Camera path over a terrain of the given size: per frame position[3] and
 * forward[3] (reference convention: forward points away from the view
 * direction).  Pitch is -pitch_deg (+- jitter), yaw and position drift with
 * a seeded smooth random walk. */
int rsc_camera_path(uint32_t seed, uint32_t frames, float terrain_size, float height,
                    float pitch_deg, float *pos_fwd6);

/* Eight large triangles facing the camera at view distances d0, d0+step, ...
 * whose screen footprint covers the whole view (3 vertices each, 24 vertices,
 * 24 indices), placed for the given camera and vertical FOV/aspect. */
int rsc_backdrop(const float cam_pos[3], const float cam_forward[3], float fov_deg, float aspect,
                 float d0, float step, float *verts, uint32_t *indices);

#ifdef __cplusplus
}
#endif
#endif
