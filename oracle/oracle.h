/*
 * CPU oracle for the sparse-voxel-octree hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker the HIP path is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library (libsvo.so) never links, loads or calls it.
 *
 * Every function restates the reference's algorithm (AdamYuan/SparseVoxelOctree,
 * GLSL shaders and host C++) in plain C++ with the fixed floating-point rules:
 *   - IEEE-754 binary32, round to nearest, no contraction (-ffp-contract=off);
 *   - correctly rounded '/' and sqrtf;
 *   - normalize(v) = v * (1.0f / sqrtf((v.x*v.x + v.y*v.y) + v.z*v.z));
 *   - GLSL min(x,y) = (y < x) ? y : x, max(x,y) = (x < y) ? y : x;
 *   - float -> uint conversion saturates (negative / NaN -> 0);
 *   - findMSB(0) = 0xFFFFFFFF;
 *   - the ESVO traversal stack starts zeroed, and reads of never-written
 *     entries are counted.
 * Canonical orders (the reference's atomics make both nondeterministic):
 *   - leaf colours are folded in ascending (fragment.x, fragment.y) order;
 *   - octree blocks are laid out by a serial allocator in word-index order.
 */
#ifndef SVO_ORACLE_H
#define SVO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same 32-byte layout as svo_hit in include/svo.h. */
typedef struct orc_hit {
	float pos[3];      /* o_pos of Octree_RayMarchLeaf */
	uint32_t iter;     /* o_iter (0 for the 5-argument variant) */
	uint32_t vxy;      /* hit cube coordinate x | y << 16, 0xFFFFFFFF on a miss */
	uint32_t vz_flags; /* z | depth << 16 | normal code << 24 | hit << 31 */
	uint32_t leaf;     /* node word at the hit (colour + count for leaves) */
	uint32_t nfetch;   /* executed octree loads */
} orc_hit;

/* The path tracer's sampler (shader/sobol.comp:78-86, shader/sobol.glsl:9-10, src/Sobol.cpp:52-62):
 * the Sobol state after n advances from zero, dims words.  mat = the 64 x 32 generator table. */
void orc_sobol_state(const uint32_t *mat, uint32_t dims, uint32_t n, uint32_t *x);

/* ConstColorPathTrace with running means (shader/path_tracer.comp:32-60,111-132) for sample index
 * spp (the Sobol counter): the primary ray through pixel + Sobol_GetVec2(0), bounce cur sampled by
 * fract(Sobol_GetVec2(cur) + blue noise of pixel & 255).  color: float4 per pixel; albedo / normal:
 * RGBA8 UNORM / RGBA8 SNORM per pixel (either may be NULL); spp = samples already accumulated.
 * sobol_mat: 64 x 32 uint32; noise: 256 x 256 x 2 bytes. */
void orc_path_trace(const uint32_t *words, uint64_t W, uint32_t level, const float *cam, uint32_t width,
                    uint32_t height, uint32_t bounce, uint32_t spp, const float *sky, const uint32_t *sobol_mat,
                    const uint8_t *noise, float *color, uint32_t *albedo, uint32_t *normal, int nthreads);

/* Environment-map tables (src/EnvironmentMap.cpp:104-188): texels 4 floats (rgb, pdf) and alias
 * {prob, alias} per texel from rgba (4 floats per texel); 0 when the map has no light. */
int orc_envmap_tables(uint32_t w, uint32_t h, const float *rgba, float *texels, uint32_t *alias);
/* EnvMapPathTrace with the same sampler and buffers as orc_path_trace. */
void orc_path_trace_env(const uint32_t *words, uint64_t W, uint32_t level, const float *cam, uint32_t width,
                        uint32_t height, uint32_t bounce, uint32_t spp, const float *texels, const uint32_t *alias,
                        uint32_t ew, uint32_t eh, float rotation, float multiplier, const uint32_t *sobol_mat,
                        const uint8_t *noise, float *color, uint32_t *albedo, uint32_t *normal, int nthreads);
/* EnvMap_Sample (direction, radiance, pdf) and EnvMap_Radiance + EnvMap_PDF of one direction. */
void orc_env_sample(const float *texels, const uint32_t *alias, uint32_t w, uint32_t h, float rotation, float u,
                    float v, float *dir, float *light, float *pdf);
void orc_env_lookup(const float *texels, uint32_t w, uint32_t h, float rotation, const float *dir, float *out,
                    float *pdf);
float orc_atan2(float y, float x);
float orc_acos(float x);

/* Scene::load_meshes normalisation to [-1,1]^3 (src/Scene.cpp:62-99). In place on xyz[3*nverts]. */
void orc_normalize(float *xyz, uint64_t nverts);

/* Voxelizer (shader/voxelizer.geom:15-54, shader/voxelizer.frag:17-44, conservative rule).
 * tri: 9 floats per triangle (already normalised); rgb: colour per triangle (R in the low byte).
 * Returns the fragment count F.  When frags != NULL, writes min(F, cap) fragments (2 uint32 each)
 * in triangle order, row-major (j outer, i inner) inside a triangle. */
uint64_t orc_voxelize(const float *tri, const uint32_t *rgb, uint64_t T, uint32_t level, uint32_t *frags,
                      uint64_t cap, int nthreads);

/* OctreeBuilder (shader/octree_{init,tag,alloc}_node.comp, octree_modify_arg.comp,
 * src/OctreeBuilder.cpp:142-214).  Returns W = (counter + 1) * 8 words; writes min(W, cap) words. */
uint64_t orc_build(const uint32_t *frags, uint64_t F, uint32_t level, uint32_t *words, uint64_t cap);

/* Breadth-first relayout of any octree in the reference node format into the canonical block order.
 * Returns the canonical word count (0 on a malformed octree); writes min(result, cap) words. */
uint64_t orc_canonicalize(const uint32_t *words, uint64_t W, uint32_t level, uint32_t *out, uint64_t cap);

/* Camera_GenRay (shader/camera.glsl:11-14).  cam = {pos, look, side, up} as 4 x vec4. */
void orc_gen_ray(const float *cam, float cx, float cy, float *d_out);

/* Beam pre-pass (shader/octree_tracer_beam.frag:12-29): bw x bh = (ceil(W/8)+1) x (ceil(H/8)+1) floats. */
void orc_beam(const uint32_t *words, uint64_t W, uint32_t level, const float *cam, uint32_t width,
              uint32_t height, float *beam, uint32_t *beam_nfetch, int nthreads);

/* Primary-ray frame (shader/octree_tracer.frag:23-48 + Octree_RayMarchLeaf 6-arg).
 * beam may be NULL (beam disabled).  hits[width*height], row-major. */
void orc_march_frame(const uint32_t *words, uint64_t W, uint32_t level, const float *cam, uint32_t width,
                     uint32_t height, const float *beam, orc_hit *hits, int nthreads);

/* Rows [y0, y1) of the same frame only; hits[width*(y1-y0)]. */
void orc_march_frame_rows(const uint32_t *words, uint64_t W, uint32_t level, const float *cam, uint32_t width,
                          uint32_t height, uint32_t y0, uint32_t y1, const float *beam, orc_hit *hits, int nthreads);

/* Explicit rays: rays[6*R] = (o.xyz, d.xyz).  variant 0 = Octree_RayMarchLeaf 6-arg
 * (shader/octree.glsl:179-340), 1 = 5-arg (shader/octree.glsl:342-502), 2 = Octree_RayMarchOcclude
 * (shader/octree.glsl:504-629; hit flag only, other fields as documented in include/svo.h). */
void orc_march_rays(const uint32_t *words, uint64_t W, uint32_t level, const float *rays, uint64_t R, int variant,
                    orc_hit *hits, int nthreads);

/* Bounce cur of sample spp from the primary hits of a width x height frame (shader/path_tracer.comp:
 * 42-44, shader/diffuse.glsl:7-14, shader/util.glsl:5-9): o = o_pos, d = Diffuse_Sample(normal,
 * fract(Sobol_GetVec2(cur) + noise(pixel & 255))).  rays[6*width*height]; a miss gets o = 0, d = -1
 * and mask[i] = 0 (mask may be NULL). */
void orc_bounce_rays(const orc_hit *hits, uint32_t width, uint32_t height, uint32_t spp, uint32_t cur,
                     const uint32_t *sobol_mat, const uint8_t *noise, float *rays, uint8_t *mask);

/* sin / cos of phi used by the bounce sampler (own polynomial; identical on the device). */
void orc_sincos(float phi, float *s, float *c);

/* The tracer's shading (shader/octree_tracer.frag:18-47) of a frame of hits into RGBA8 UNORM texels;
 * view 0..3 = diffuse, normal, position, iteration heat; misses take light[3], or the environment
 * map's radiance (texels != NULL, rotation; no multiplier).  pow is the host libm's. */
void orc_shade(const orc_hit *hits, const float *cam, uint32_t width, uint32_t height, int view, const float *light,
               const float *texels, uint32_t ew, uint32_t eh, float rotation, uint32_t *out);

/* Number of stack reads of never-written entries seen since the last reset (must stay 0). */
uint64_t orc_stack_uninit_reads(void);
void orc_reset_counters(void);

#ifdef __cplusplus
}
#endif

#endif
