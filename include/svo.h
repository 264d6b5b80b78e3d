/*
 * svo.h -- C ABI of the MI355X (gfx950) sparse-voxel-octree hot path.
 *
 * This is the boundary the reference's hot-path classes would call instead of recording Vulkan
 * command buffers.  Every entry point names the reference interface it replaces
 * (AdamYuan/SparseVoxelOctree, paths relative to that repository).  Plain C types only: pointers,
 * sizes and opaque handles.  A context owns one HIP stream on one device; calls on one context are
 * ordered on that stream and must not be made from two threads at once; distinct contexts are
 * independent.  "Enqueue" functions return before the GPU finishes; functions that return a size
 * decided on the device (fragment count, octree words) synchronise the context's stream.
 *
 * Data formats (byte-identical to the reference):
 *   fragment   uint32[2]  .x = x | y << 12 | (z & 0xff) << 24,  .y = (z >> 8) << 28 | rgb24
 *                         (shader/voxelizer.frag:38-42)
 *   node word  uint32     0 = empty; bit31 = non-empty; bit30 = leaf.  Leaf: bits29..24 = fragment
 *                         count (saturates at 63), bits23..0 = RGB (R in the low byte).  Internal:
 *                         bits29..0 = word index of the first of its 8 children.  The root is
 *                         implicit; its children are words 0..7; child slot = x | y << 1 | z << 2
 *                         (shader/octree_tag_node.comp:10-60, shader/octree_alloc_node.comp:21)
 *   octree space [1,2]^3 (shader/octree.glsl:52-53); voxel v of level L spans [1+v/2^L, 1+(v+1)/2^L].
 *   Node order: canonical breadth-first order (child blocks in ascending parent word index); leaf
 *   colours folded in ascending (fragment.x, fragment.y) order.
 */
#ifndef SVO_H
#define SVO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------------------------
 * Status codes.  The reference has none (Vulkan results are mostly unchecked, Scene::Create
 * returns nullptr, src/Scene.cpp:398-401); every entry point here returns one of these.
 * ------------------------------------------------------------------------------------------- */
#define SVO_OK 0
#define SVO_ERR_INVALID_ARGUMENT 1 /* null pointer, bad variant, size mismatch */
#define SVO_ERR_BAD_LEVEL 2        /* level outside [1, 12] (src/Config.hpp:18-19) */
#define SVO_ERR_OUT_OF_MEMORY 3    /* device allocation failed */
#define SVO_ERR_HIP 4              /* HIP runtime error (text in svo_last_error) */
#define SVO_ERR_OVERFLOW 5         /* more than 2^32-1 fragments, or an octree above 2^30 words */
#define SVO_ERR_NO_DEVICE 6        /* no HIP device with that index */
#define SVO_ERR_MALFORMED 7        /* octree violates the node format during marching */

/* Opaque handles. */
typedef struct svo_ctx svo_ctx;
typedef struct svo_scene svo_scene;       /* triangles resident in HBM (Scene) */
typedef struct svo_fraglist svo_fraglist; /* voxel fragment list in HBM (Voxelizer) */
typedef struct svo_octree svo_octree;     /* octree words in HBM (OctreeBuilder / Octree) */
typedef struct svo_envmap svo_envmap;     /* environment map + sampling tables in HBM (EnvironmentMap) */

/* Camera uniform block {vec4 pos, look, side, up} (src/Camera.hpp:28-30, src/Camera.cpp:77-95). */
typedef struct svo_camera {
	float pos[4];
	float look[4];
	float side[4];
	float up[4];
} svo_camera;

/* Per-ray result, 32 bytes.  The reference only shades it into RGBA8
 * (shader/octree_tracer.frag:37-47); this record keeps the raw outputs of Octree_RayMarchLeaf. */
typedef struct svo_hit {
	float pos[3];      /* o_pos */
	uint32_t iter;     /* o_iter (0 for the 5-argument marcher) */
	uint32_t vxy;      /* hit cube x | y << 16 (units of the hit cube size); 0xFFFFFFFF on a miss */
	uint32_t vz_flags; /* z | depth << 16 | normal << 24 | hit << 31; normal = 2*axis + (sign > 0).
	                      A miss stores 0xFFFF | normal << 24 */
	uint32_t leaf;     /* node word at the hit (o_color = its low 24 bits) */
	uint32_t nfetch;   /* octree words loaded by this ray */
} svo_hit;

#define SVO_MARCH_LEAF_ITER 0 /* Octree_RayMarchLeaf, 6 arguments (shader/octree.glsl:179-340) */
#define SVO_MARCH_LEAF 1      /* Octree_RayMarchLeaf, 5 arguments (shader/octree.glsl:342-502) */
#define SVO_MARCH_OCCLUDE 2   /* Octree_RayMarchOcclude (shader/octree.glsl:504-629): hit bit only */

/* Device time of the last voxelize / build / march on a context, in milliseconds. */
typedef struct svo_timing {
	float normalize_ms;  /* A1 bounding box + transform */
	float vox_count_ms;  /* count pass + scan (Voxelizer::count_and_create_fragment_list) */
	float vox_fill_ms;   /* fill pass (Voxelizer::CmdVoxelize) */
	float build_ms;      /* all level passes incl. the leaf fold (OctreeBuilder::CmdBuild) */
	float beam_ms;       /* beam pre-pass (OctreeTracer::CmdBeamRenderPass) */
	float march_ms;      /* primary march (OctreeTracer::CmdDrawPipeline) */
	float rays_ms;       /* last svo_march_rays */
	float bounce_ms;     /* last svo_bounce_rays */
} svo_timing;

/* ---- context (replaces myvk::Device / Queue / CommandPool set-up, src/Application.cpp:161-293) */
int svo_ctx_create(int device, svo_ctx **out);
int svo_ctx_destroy(svo_ctx *ctx);
const char *svo_last_error(const svo_ctx *ctx);
int svo_ctx_synchronize(svo_ctx *ctx);
/* Returns the context's hipStream_t, for callers that order their own work against it. */
void *svo_ctx_stream(svo_ctx *ctx);
/* Device-side error bits raised by kernels since the last call (and cleared): 1 = march
 * iteration cap, 2 = octree read out of range, 4 = traversal deeper than the stack. */
int svo_ctx_take_device_errors(svo_ctx *ctx, uint32_t *bits);
/* Release cached device memory held by the context's allocator. */
int svo_ctx_trim(svo_ctx *ctx);
/* Bytes currently held by the context's allocator: in use and cached. */
int svo_ctx_memory(svo_ctx *ctx, uint64_t *in_use, uint64_t *cached, uint64_t *peak_in_use);

/* ---- device memory helpers, for callers that keep ray batches and hit buffers in HBM */
int svo_malloc(svo_ctx *ctx, uint64_t bytes, void **dptr);
int svo_free(svo_ctx *ctx, void *dptr);
/* Device-to-device copy enqueued on the context's stream by a plain copy kernel (16-byte vectors
 * when both pointers and the size are 16-byte aligned).  No reference counterpart (Vulkan
 * CmdCopy); used to grow the node buffer and to calibrate HBM byte counters. */
int svo_copy_device(svo_ctx *ctx, void *dst_dev, const void *src_dev, uint64_t bytes);
int svo_upload(svo_ctx *ctx, void *dst_dev, const void *src_host, uint64_t bytes);   /* synchronous */
int svo_download(svo_ctx *ctx, void *dst_host, const void *src_dev, uint64_t bytes); /* synchronous */

/* ---- per-kernel profiler: when on, every kernel launch on the context is bracketed by HIP events
 * on its stream.  svo_prof_get fills name / launches / total device ms of entry i. */
int svo_prof_enable(svo_ctx *ctx, int on);
int svo_prof_reset(svo_ctx *ctx);
int svo_prof_count(svo_ctx *ctx, int *n);
int svo_prof_get(svo_ctx *ctx, int i, const char **name, uint64_t *launches, double *total_ms);

int svo_last_timing(svo_ctx *ctx, svo_timing *out); /* synchronises */

/* ---- Scene (Scene::Create / load_meshes, src/Scene.cpp:36-147,387-417).
 * tri_xyz: 9 floats per triangle (triangle soup); tri_rgb: RGB8 colour per triangle (R in the low
 * byte; stands in for packUnorm4x8(Kd) of the triangle's material, src/Scene.cpp:164).
 * normalize != 0 applies the reference's normalisation to [-1,1]^3 (src/Scene.cpp:90-99) on the
 * device.  The _device variant copies from device arrays. */
int svo_scene_create(svo_ctx *ctx, const float *tri_xyz, const uint32_t *tri_rgb, uint64_t T, int normalize,
                     svo_scene **out);
int svo_scene_create_device(svo_ctx *ctx, const float *tri_xyz_dev, const uint32_t *tri_rgb_dev, uint64_t T,
                            int normalize, svo_scene **out);
/* New triangles for an existing scene, the same count T (new; the reference loads a scene once):
 * the steady-state loop of svo_build_plan_run rebuilds from them.  Same arguments as create. */
int svo_scene_update(svo_ctx *ctx, svo_scene *scene, const float *tri_xyz, const uint32_t *tri_rgb, int normalize);
int svo_scene_update_device(svo_ctx *ctx, svo_scene *scene, const float *tri_xyz_dev, const uint32_t *tri_rgb_dev,
                            int normalize);
int svo_scene_info(const svo_scene *scene, uint64_t *T, const float **tri_xyz_dev, const uint32_t **tri_rgb_dev);
int svo_scene_download(svo_ctx *ctx, const svo_scene *scene, float *tri_xyz);
int svo_scene_destroy(svo_scene *scene);

/* ---- Voxelizer (Voxelizer::Create + CmdVoxelize, src/Voxelizer.cpp:5-26,134-179).
 * svo_voxelize_count = the count-only pass, a read-back of F and the allocation of the list
 * (Voxelizer::count_and_create_fragment_list); svo_voxelize_fill enqueues the fill pass
 * (Voxelizer::CmdVoxelize); svo_voxelize does both.  Fragments are stored in triangle order and,
 * inside a triangle, row by row of the projection plane. */
int svo_voxelize_count(svo_ctx *ctx, const svo_scene *scene, uint32_t level, svo_fraglist **out);
int svo_voxelize_fill(svo_ctx *ctx, svo_fraglist *frags);
int svo_voxelize(svo_ctx *ctx, const svo_scene *scene, uint32_t level, svo_fraglist **out);
/* n scenes at once (scene i at levels[i]): one count launch, one read-back of every scene's
 * fragment count, one fill launch; out[i] receives scene i's list, identical to svo_voxelize's.
 * The reference voxelizes one scene per Voxelizer::Create + CmdVoxelize
 * (src/Voxelizer.cpp:5-26,166-179); this batched form is new.  n <= 1024. */
int svo_voxelize_batch(svo_ctx *ctx, const svo_scene *const *scenes, const uint32_t *levels, uint32_t n,
                       svo_fraglist **out);
/* A fragment list from the host (e.g. one read back from the reference) for svo_build. */
int svo_fraglist_create(svo_ctx *ctx, const uint32_t *frags, uint64_t F, uint32_t level, svo_fraglist **out);
/* Voxelizer::GetLevel / GetVoxelFragmentCount / GetVoxelFragmentList (src/Voxelizer.hpp:45-52). */
int svo_fraglist_info(const svo_fraglist *frags, uint32_t *level, uint64_t *F, const uint32_t **frags_dev);
int svo_fraglist_download(svo_ctx *ctx, const svo_fraglist *frags, uint32_t *out, uint64_t cap_fragments);
int svo_fraglist_destroy(svo_fraglist *frags);

/* ---- OctreeBuilder (OctreeBuilder::Create / CmdBuild / GetOctreeRange, src/OctreeBuilder.cpp:8-214).
 * svo_build_enqueue (CmdBuild) runs the structure phase (occupancy of every level), reads back the
 * exact node count once, allocates exactly W words and enqueues the node phase (pointers, leaf
 * fold); svo_build_finish waits and returns the used range (GetOctreeRange:
 * W = (counter + 1) * 8 words); svo_build does both.  More than 2^30 words -> SVO_ERR_OVERFLOW. */
int svo_build_enqueue(svo_ctx *ctx, const svo_fraglist *frags, svo_octree **out);
int svo_build_finish(svo_ctx *ctx, svo_octree *oct, uint64_t *W);
int svo_build(svo_ctx *ctx, const svo_fraglist *frags, svo_octree **out);
/* n octrees at once: one launch per build pass for all of them and one read-back of every
 * octree's size; out[i] is identical to svo_build(frags[i]) and is finished by svo_build_finish
 * or svo_ctx_synchronize.  The reference builds one octree per OctreeBuilder::CmdBuild
 * (src/OctreeBuilder.cpp:142-210); this batched form is new.  n <= 1024, at most 2^32-2
 * fragments in total. */
int svo_build_batch(svo_ctx *ctx, const svo_fraglist *const *frags, uint32_t n, svo_octree **out);
/* Captured voxelize + build (new; the reference voxelizes and builds each scene once, reading the
 * fragment count and the octree size back to the host in between, src/Voxelizer.cpp:134-179,
 * src/OctreeBuilder.cpp:142-214).  A plan over n scenes at fixed levels allocates each scene's
 * fragment list and octree at a capacity: frag_caps[i] fragments and word_caps[i] words (either
 * array may be NULL: then one plain voxelize + build of the scenes as they are now sizes it, plus
 * a quarter, words at most 2^30).  Every count stays on the device, so no pass waits for the host, and the passes are
 * captured once as a HIP graph.  svo_build_plan_run replays them on the context's stream; the
 * scenes' triangles may change between runs (svo_scene_update), their number may not.  A run whose
 * counts pass a capacity writes nothing past it.  svo_build_plan_finish waits for the last run,
 * fills F[i] and W[i] (either may be NULL; the counts the scenes needed) and run_ms (may be NULL;
 * device time of the run), and returns SVO_ERR_OVERFLOW when a count passed its capacity: the
 * lists and octrees then hold no valid data until a run fits.  After a run that fits they are
 * identical to svo_voxelize_batch + svo_build_batch of the same scenes.  They belong to the plan
 * (svo_build_plan_get).  The single-scene marches read an octree's size from the host, so they
 * follow svo_build_plan_finish; a march plan made once over the plan's octrees stays valid across
 * runs and needs no finish in between (both are ordered on the context's stream), so a
 * steady-state loop can replay both graphs: build plan run, march plan run.  use_graph(plan, 0)
 * enqueues the same passes as plain launches.
 * n <= 1024; the scenes must outlive the plan. */
typedef struct svo_build_plan svo_build_plan;
int svo_build_plan_create(svo_ctx *ctx, const svo_scene *const *scenes, const uint32_t *levels, uint32_t n,
                          const uint64_t *frag_caps, const uint64_t *word_caps, svo_build_plan **out);
int svo_build_plan_use_graph(svo_build_plan *plan, int on);
int svo_build_plan_run(svo_build_plan *plan);
int svo_build_plan_finish(svo_build_plan *plan, uint64_t *F, uint64_t *W, double *run_ms);
int svo_build_plan_get(const svo_build_plan *plan, uint32_t i, const svo_fraglist **frags, const svo_octree **oct);
int svo_build_plan_destroy(svo_build_plan *plan);
/* An octree from host words (any node order), e.g. one read back from the reference. */
int svo_octree_create(svo_ctx *ctx, const uint32_t *words, uint64_t W, uint32_t level, svo_octree **out);
/* Octree::GetLevel / GetRange / GetBuffer (src/Octree.hpp:18-33). */
int svo_octree_info(const svo_octree *oct, uint32_t *level, uint64_t *W, const uint32_t **words_dev);
int svo_octree_download(svo_ctx *ctx, const svo_octree *oct, uint32_t *out, uint64_t cap_words);
int svo_octree_destroy(svo_octree *oct);
/* Hand-off of a built octree to another context's stream (OctreeBuilder::CmdTransferOctreeOwnership
 * and Octree::CmdTransferOwnership, src/OctreeBuilder.hpp:45, src/Octree.hpp:31; used by
 * src/LoaderThread.cpp:78-82,99-105 to pass the octree from the loader queue to the main queue).
 * Records a HIP event after the work enqueued so far on the context that built `oct` and makes
 * dst's stream wait for it, so work enqueued on dst afterwards sees the finished octree; the host
 * does not wait.  Both contexts must be on one device; the octree's memory stays with the context
 * that built it (free it there). */
int svo_octree_transfer(const svo_octree *oct, svo_ctx *dst);
/* Octree files (the reference keeps octrees only in device memory; this is new): a 24-byte header
 * {char magic[8] = "SVOCTRE1", uint32 level, uint32 0, uint64 W} followed by W little-endian
 * words in canonical order.  svo_octree_load returns an octree ready to march. */
int svo_octree_save(svo_ctx *ctx, const svo_octree *oct, const char *path);
int svo_octree_load(svo_ctx *ctx, const char *path, svo_octree **out);

/* ---- OctreeTracer (CmdBeamRenderPass / CmdDrawPipeline, src/OctreeTracer.cpp:183-246).
 * Primary rays through every pixel of a width x height frame from the camera; beam != 0 runs the
 * 8x8 beam pre-pass first (m_beam_enable).  hits_dev: width*height svo_hit, row-major (y = 0 is
 * the first row, as gl_FragCoord in Vulkan).  beam_dev (optional, may be NULL): receives the
 * (ceil(W/8)+1) x (ceil(H/8)+1) beam image. */
int svo_march(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t width, uint32_t height,
              int beam, svo_hit *hits_dev, float *beam_dev);
/* The beam pre-pass alone (OctreeTracer::CmdBeamRenderPass): beam_dev receives the
 * (ceil(W/8)+1) x (ceil(H/8)+1) image of max(0, t - size) or 1e10; nfetch_dev (optional) the octree
 * words each beam cell loaded. */
int svo_march_beam(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t width, uint32_t height,
                   float *beam_dev, uint32_t *nfetch_dev);
/* Allocation-free primary march.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same records as svo_march(ctx, oct, cam, width, height, 1, hits_dev, NULL), from the
 * same kernels, with the beam image kept in a static device array of 2^20 cells (4 MiB per device,
 * held until the process ends) instead of pool memory: frames need (ceil(W/8)+1)*(ceil(H/8)+1) <=
 * 1048576 beam cells (up to 7680x4320, for example); larger frames return SVO_ERR_INVALID_ARGUMENT.
 * The launch schedule is fixed (svo_march_mode's, else one ray per lane in 16x16 tiles; the autotuner
 * is not consulted) and no stage timer runs.  With the profiler off (svo_prof_enable) it calls no
 * allocating routine.  Every call on one device shares the array, which keeps the last call's beam
 * image: calls that overlap on one GPU, from any context, race; serialise them. */
int svo_march_noalloc(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t width,
                      uint32_t height, svo_hit *hits_dev);
/* The tracer's shaded image (shader/octree_tracer.frag:18-47, OctreeTracer::m_view_type,
 * src/OctreeTracer.hpp:15) from a frame of hit records (svo_march's hits_dev) and the camera they
 * were marched with.  view_type: 0 = diffuse, pow(albedo, 1/2.2); 1 = normal * 0.5 + 0.5;
 * 2 = position, pos - 1; 3 = iterations, Heat(iter / 128) = sin(clamp(x,0,1) * 3 - (1,2,3)) * 0.5 +
 * 0.5.  A miss has pos = 1, normal = 0 and, as colour, the light in the ray's direction: the
 * constant colour light[3], or env's radiance turned by env_rotation when env is not NULL (without
 * the multiplier, as the reference's Light()).  rgba8_dev: width*height uint32, RGBA8 UNORM (R in the
 * low byte, alpha 255), each channel clamped to [0, 1] and rounded to nearest. */
int svo_shade(svo_ctx *ctx, const svo_hit *hits_dev, const svo_camera *cam, uint32_t width, uint32_t height,
              int view_type, const float light[3], const svo_envmap *env, float env_rotation, uint32_t *rgba8_dev);
/* Explicit rays (path-tracer bounces, shader/path_tracer.comp:32-60): rays_dev = R x {o.xyz, d.xyz}
 * floats in octree space; variant = SVO_MARCH_*. */
int svo_march_rays(svo_ctx *ctx, const svo_octree *oct, const float *rays_dev, uint64_t R, int variant,
                   svo_hit *hits_dev);
/* Batched tracing (new; the reference traces one octree per OctreeTracer::CmdDrawPipeline,
 * src/OctreeTracer.cpp:226-246).  svo_march_batch: primary rays of n octrees, camera cams[i], a
 * width x height frame each, in one beam launch and one march launch; hits_dev holds n frames of
 * width*height records, octree i's at i * width * height, each identical to svo_march's.  */
int svo_march_batch(svo_ctx *ctx, const svo_octree *const *octs, const svo_camera *cams, uint32_t n,
                    uint32_t width, uint32_t height, int beam, svo_hit *hits_dev);
/* Reusable form: the octree and camera table and the beam scratch stay on the device; cameras may
 * change between runs; with svo_march_plan_use_graph(plan, 1) the launches are captured once as a
 * HIP graph and replayed by every later run into the same hit buffer. */
typedef struct svo_march_plan svo_march_plan;
int svo_march_plan_create(svo_ctx *ctx, const svo_octree *const *octs, const svo_camera *cams, uint32_t n,
                          uint32_t width, uint32_t height, int beam, svo_march_plan **out);
int svo_march_plan_set_cameras(svo_ctx *ctx, svo_march_plan *plan, const svo_camera *cams);
int svo_march_plan_use_graph(svo_march_plan *plan, int on);
/* Workgroup tile of the march launch: 0 = 16x16 pixels (4 waves), 1 = 8x8 (1 wave), 2 = 16x8,
 * 3 = 32x8; -1 = the default, or the autotuner's choice when svo_autotune is on. */
int svo_march_plan_set_shape(svo_march_plan *plan, int shape);
int svo_march_plan_run(svo_ctx *ctx, svo_march_plan *plan, svo_hit *hits_dev);
int svo_march_plan_destroy(svo_march_plan *plan);
/* Explicit rays against n octrees in one launch: rays [ray_begin[i], ray_begin[i+1]) of rays_dev go
 * to octs[i] (ray_begin: host array of n + 1 non-decreasing entries, ray_begin[0] = 0). */
int svo_march_rays_batch(svo_ctx *ctx, const svo_octree *const *octs, uint32_t n, const uint64_t *ray_begin,
                         const float *rays_dev, int variant, svo_hit *hits_dev);
/* Launch-shape autotuner for the batched marches (off by default).  When on, the first launch of a
 * new shape class (march batch: n, width, height, beam; ray batch: variant, log2 R) times every
 * candidate workgroup/tile shape once and keeps the fastest; the report lists the decisions. */
int svo_autotune(svo_ctx *ctx, int on);
/* Schedule of svo_march, svo_march_rays, svo_march_rays_batch and the server's ray batches:
 * 1 = a fixed grid of waves whose lanes take a new ray as soon as theirs has ended (rays dealt to
 * the waves in 64-ray chunks; wins when ray lengths diverge); 0 = one ray per lane (wins
 * for coherent rays), frames in 16x16-pixel blocks; 2 = as 0 with one-wave 8x8-pixel blocks for
 * frames; -1 (default) = chosen per shape class by the autotuner when svo_autotune is on, else 0.
 * Results are identical. */
int svo_march_mode(svo_ctx *ctx, int mode);
int svo_autotune_report(svo_ctx *ctx, char *buf, uint64_t cap, uint64_t *len);

/* Bounce `cur` (1..31) of path-tracer sample `spp` for a width x height frame of primary hits (ray
 * k = pixel (k % width, k / width)), as shader/path_tracer.comp:42-44 draws it: o = o_pos,
 * d = Diffuse_Sample(normal, fract(Sobol_GetVec2(cur) + noise)) (shader/diffuse.glsl:7-14,
 * shader/util.glsl:5-9), where Sobol_GetVec2 reads the Sobol state after spp advances
 * (shader/sobol.comp, src/Sobol.cpp:45-62) and noise is the pixel's texel of the 256x256 blue-noise
 * texture (src/Noise.inl, fetched at pixel & 255).  rays_dev = width*height x 6 floats; a ray whose
 * hit bit is clear gets o = 0, d = -1. */
int svo_bounce_rays(svo_ctx *ctx, const svo_hit *hits_dev, uint32_t width, uint32_t height, uint32_t spp,
                    uint32_t cur, float *rays_dev);

/* ---- serving loop (new; the reference's nearest analogue is its loader thread, which builds one
 * octree on a side queue and hands it to the renderer, src/LoaderThread.cpp:15-116,
 * src/Octree.cpp:22-35).  Build and ray-query requests for many scenes are accepted from any thread
 * at any time.  Two worker threads, each with its own context (HIP stream) on the device, batch
 * them: every pending ray query is marched in one launch per variant, pending builds are voxelized
 * and built together, and every built octree stays resident, admitted by its exact size against
 * budget_bytes and max_octrees (least-recently-used idle octrees are evicted to make room; a
 * build that still does not fit completes with SVO_ERR_OUT_OF_MEMORY).
 *   submit_build: the octree of `scene` at `level`; its id is the request id.  The scene must stay
 *                 alive until the request completes.
 *   submit_rays:  R rays (rays_dev, 6 floats each) against resident octree `octree_id`; hits_dev
 *                 receives R svo_hit.  Both buffers must stay valid until completion.  A query
 *                 for an octree that is not resident completes with SVO_ERR_INVALID_ARGUMENT.
 *   wait:         blocks until the request completes; status = its result; latency_ms = submit to
 *                 completion on the host clock.  Each request id is waited for at most once. */
typedef struct svo_server svo_server;
typedef struct svo_server_stats {
	uint64_t builds_done, builds_rejected, ray_queries_done, ray_queries_failed, rays_traced, evictions, batches;
	uint64_t resident_octrees, resident_bytes, budget_bytes;
	double build_p50_ms, build_p99_ms, rays_p50_ms, rays_p99_ms;
	double busy_build_ms, busy_rays_ms, elapsed_ms;
} svo_server_stats;
int svo_server_create(int device, uint64_t budget_bytes, uint32_t max_octrees, svo_server **out);
int svo_server_destroy(svo_server *srv); /* drains pending requests, frees the resident octrees */
int svo_server_submit_build(svo_server *srv, const svo_scene *scene, uint32_t level, uint64_t *req_id);
int svo_server_submit_rays(svo_server *srv, uint64_t octree_id, const float *rays_dev, uint64_t R, int variant,
                           svo_hit *hits_dev, uint64_t *req_id);
int svo_server_wait(svo_server *srv, uint64_t req_id, int *status, double *latency_ms);
/* A resident octree's word count and handle (for inspection; owned by the server). */
int svo_server_resident(svo_server *srv, uint64_t octree_id, uint64_t *W, const svo_octree **oct);
int svo_server_get_stats(svo_server *srv, svo_server_stats *out);
int svo_server_reset_stats(svo_server *srv);

/* ---- camera helper (Camera::fetch_uniform_data, src/Camera.cpp:77-95). */
int svo_camera_from_yaw_pitch(const float position[3], float yaw, float pitch, float fov, float aspect,
                              svo_camera *out);

/* ---- progressive path tracer (PathTracer::CmdRender, src/PathTracer.cpp:176-215;
 * ConstColorPathTrace and main of shader/path_tracer.comp:32-60,111-132).  Sample `spp` of every
 * pixel, with the reference's sampler: the Sobol state after spp advances of shader/sobol.comp
 * (2 * (bounce + 1) dimensions, src/PathTracer.cpp:184), the primary ray through
 * pixel + Sobol_GetVec2(0) (the pixel corner at spp 0), bounce cur sampled by
 * fract(Sobol_GetVec2(cur) + noise(pixel & 255)) from the 256x256 blue-noise texture; up to
 * bounce-1 cosine-weighted diffuse bounces (Octree_RayMarchLeaf, 5 arguments); a path that escapes
 * adds sky (linear RGB) times the product of the leaf albedos it met.  The buffers hold the
 * running means after spp + 1 samples, in the reference's image formats (src/PathTracer.cpp:8-16):
 *   color_dev   width*height float4 (RGBA32F, a = 1)
 *   albedo_dev  width*height uint32, RGBA8 UNORM (R in the low byte, a = 255): first-hit albedo
 *   normal_dev  width*height uint32, RGBA8 SNORM (a = 127): first-hit face normal
 * imageLoad / imageStore of the 8-bit formats: byte / 255 (UNORM), max(signed byte / 127, -1)
 * (SNORM); stores round the clamped value times 255 (127) to nearest, halves up.  spp = samples
 * already in the buffers; spp 0 starts them, as after the reference's Reset clears its images.
 * albedo_dev and normal_dev may be NULL.  bounce: 1..31 (the Sobol table has 64 dimensions; the
 * reference's UI allows 2..16, src/Config.hpp:24-26). */
int svo_path_trace(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t width, uint32_t height,
                   uint32_t bounce, uint32_t spp, const float sky[3], float *color_dev, uint32_t *albedo_dev,
                   uint32_t *normal_dev);

/* ---- environment-map light (EnvironmentMap, src/EnvironmentMap.cpp:55-190; EnvMapPathTrace,
 * shader/path_tracer.comp:62-106).  svo_envmap_create: width x height texels of linear RGB (4
 * floats each, the 4th ignored, row 0 = the +y pole), as an equirectangular map.  The texel pdf
 * (luminance * sin(theta), mean 1) and the alias table are built on the host, as the reference does,
 * and kept in HBM.  A map with zero luminance everywhere is rejected.  svo_envmap_download returns
 * the device tables: texels (4 floats: rgb, pdf) and alias ({prob, alias} per texel); either may be
 * NULL. */
int svo_envmap_create(svo_ctx *ctx, uint32_t width, uint32_t height, const float *rgba, svo_envmap **out);
int svo_envmap_download(svo_ctx *ctx, const svo_envmap *env, float *texels, uint32_t *alias);
int svo_envmap_destroy(svo_envmap *env);
/* svo_path_trace with the environment-map light (EnvMapPathTrace, shader/path_tracer.comp:62-106): a
 * path that misses at once takes the map's radiance; every bounce adds a light sample (alias-table
 * texel, occlusion ray) and, when its diffuse ray escapes, the map's radiance, both weighted by the
 * power heuristic and scaled by multiplier; rotation turns the map about +y (radians).  The light
 * sample and the diffuse direction use the same 2-D sample, as the reference does.  Same buffers
 * and sampler as svo_path_trace.  Texture lookups are bilinear with repeat addressing, in fp32. */
int svo_path_trace_env(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t width, uint32_t height,
                       uint32_t bounce, uint32_t spp, const svo_envmap *env, float rotation, float multiplier,
                       float *color_dev, uint32_t *albedo_dev, uint32_t *normal_dev);

/* ---- OBJ + MTL files (Scene::load_meshes, src/Scene.cpp:36-147, over tinyobj::LoadObj): the
 * triangle soup for svo_scene_create.  *xyz: 9 floats per triangle, the file's positions (not yet
 * normalised: pass normalize = 1 to svo_scene_create, as Scene::Create does); *rgb: per triangle,
 * packUnorm4x8(vec4(Kd, 0)) & 0xffffff of its material (src/Scene.cpp:164).  Faces are grouped per
 * material, largest group first; polygons are ear-clipped as tinyobj does.  Both arrays are
 * allocated here and released with svo_obj_free.  Errors (unreadable file, no material, a face
 * without a material, an index of 0 or out of range, no faces) return SVO_ERR_INVALID_ARGUMENT with
 * the reason in msg (msgcap bytes, may be 0). */
int svo_obj_load(const char *path, float **xyz, uint32_t **rgb, uint64_t *T, char *msg, uint64_t msgcap);
void svo_obj_free(float *xyz, uint32_t *rgb);

/* ---- synthetic scenes used by the tests and the benchmark (seeded, deterministic).
 * kind: 1 = UV spheres (64x32, 4032 triangles each), 2 = box room + rotated boxes + spheres,
 * 3 = spheres + thin slivers.  With tri_xyz == NULL only *T is returned (the triangle count). */
int svo_gen_scene(int kind, uint64_t seed, uint64_t param, float *tri_xyz, uint32_t *tri_rgb, uint64_t *T);

#ifdef __cplusplus
}
#endif

#endif /* SVO_H */
