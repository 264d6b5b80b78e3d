// Scene batches: every voxelize and build kernel works on a batch of S scenes (S = 1 for the
// single-scene entry points).  Items of all scenes (triangles, fragments, bitmap words, leaf blocks)
// form one index space; a scene's items are contiguous, and begin[s] (non-decreasing, begin[0] = 0,
// S + 1 entries) is the first item of scene s.  Per-scene parameters live in a device table of
// segment descriptors, so one launch covers the whole batch whatever the scenes' sizes and levels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "svo_runtime.hpp"

namespace svo {

constexpr uint32_t kMaxBatch = 1024; // scenes per batch call
constexpr uint32_t kMaxLevel = 12;

// Scene of item i: the last s with begin[s] <= i (scenes without items are skipped).
__device__ __forceinline__ uint32_t seg_find(const uint64_t *__restrict__ begin, uint32_t S, uint64_t i) {
	uint32_t lo = 0, hi = S - 1u;
	while (lo < hi) {
		const uint32_t mid = (lo + hi + 1u) >> 1;
		if (begin[mid] <= i)
			lo = mid;
		else
			hi = mid - 1u;
	}
	return lo;
}

// Scene of item i knowing a scene s at or before it (a loop's previous scene): s itself when i is
// still inside it, else a binary search over the scenes after s.
__device__ __forceinline__ uint32_t seg_advance(const uint64_t *__restrict__ begin, uint32_t S, uint32_t s, uint64_t i) {
	if (s + 1u >= S || begin[s + 1u] > i)
		return s;
	uint32_t lo = s + 1u, hi = S - 1u;
	while (lo < hi) {
		const uint32_t mid = (lo + hi + 1u) >> 1;
		if (begin[mid] <= i)
			lo = mid;
		else
			hi = mid - 1u;
	}
	return lo;
}

// Same, starting from a scene s at or before item i's (the scene of the block's first item).
__device__ __forceinline__ uint32_t seg_walk(const uint64_t *__restrict__ begin, uint32_t S, uint32_t s, uint64_t i) {
	while (s + 1u < S && begin[s + 1u] <= i)
		++s;
	return s;
}

// Calls body(s) with the scene s of item i, inside a loop step of a block that covers the items
// [base, base + kT).  hint (0 before the loop) tracks the scene of base.  When all kT items of the
// step lie in one scene, s is passed wave-uniform (readfirstlane), so the body's loads of that
// scene's descriptor are scalar; otherwise each lane walks to its own scene.
template <uint32_t kT, class Body>
__device__ __forceinline__ void with_seg(const uint64_t *__restrict__ begin, uint32_t S, uint32_t &hint, uint64_t base,
                                         uint64_t i, Body body) {
	if (S == 1u) {
		body(0u);
		return;
	}
	hint = seg_advance(begin, S, hint, base);
	if (base + kT <= begin[hint + 1u])
		body((uint32_t)__builtin_amdgcn_readfirstlane(hint));
	else
		body(seg_walk(begin, S, hint, i));
}

// Camera of one march (the uniform block of src/Camera.cpp:77-95 without padding).
struct CamArgs {
	float px, py, pz, lx, ly, lz, sx, sy, sz, ux, uy, uz;
};

// One octree of a march batch.  Explicit-ray batches: the segment's rays and hit records.
struct MarchSeg {
	const uint32_t *oct;
	uint64_t W;
	CamArgs cam;
	float orig_sz; // exp2(1 - level), the beam's cone origin size (src/OctreeTracer.cpp:206)
	uint32_t level;
	const float *rays; // 6 floats per ray
	uint4 *hits;       // 2 uint4 per ray
};

// One scene of a voxelize batch.  frag_begin / F are written by the count pass.
struct VoxSeg {
	const float *xyz;
	const uint32_t *rgb;
	uint2 *frags;
	uint64_t T;
	uint64_t frag_begin; // global index of the scene's first fragment
	uint64_t F;
	uint32_t level;
	uint32_t pad;
};

// One scene of a build batch.  W, nbins, overflow and fb are written by the structure phase.
struct BuildSeg {
	const uint2 *frags;
	uint32_t *words;     // the scene's octree (node phase)
	uint64_t F;
	uint64_t bm;         // global word offset of the scene's bitmaps (B[1] first)
	uint64_t bin_begin;  // global index of the scene's first leaf block (node phase)
	uint64_t W;          // octree words = (counter + 1) * 8
	uint64_t nbins;      // leaf blocks = nodes of depth L-1 (the root when L = 1)
	uint32_t L;
	uint32_t overflow;   // W > 2^30
	uint64_t off[14];    // word offset of B[d] relative to bm, d = 1..L; off[L] = the scene's bitmap words
	uint32_t fb[16];     // first block of every depth: fb[0] = 0 (root block), fb[1] = 1
};

// Voxelize batch state, from the count pass to the fill.
struct VoxBatch {
	uint32_t S = 0;
	uint64_t Ttot = 0;
	std::vector<VoxSeg> h_segs;
	std::vector<uint64_t> h_tbeg;
	VoxSeg *segs = nullptr;     // device table
	uint64_t *tbeg = nullptr;   // S + 1 triangle begins
	uint64_t *fend = nullptr;   // the fill's {fragment total, limit}, after tbeg
	uint64_t *offsets = nullptr; // Ttot exclusive fragment offsets (global)
	uint32_t *big = nullptr;    // big-triangle list: count, then global triangle indices
};

// Build batch state, from the structure phase to the node phase.
struct BuildBatch {
	uint32_t S = 0;
	uint32_t maxL = 0;
	uint64_t Ftot = 0, TW = 0, nbins = 0;
	std::vector<BuildSeg> h_segs;
	// begin tables, S + 1 entries each: fbeg (fragments), bbeg (leaf blocks), dw[d] at 2 + d (words of
	// B[d] of the scenes with L > d), dr[d] at 3 + kMaxLevel + d (same, scenes with L > d + 1); then
	// pw (pw_at): the block writer's segments
	std::vector<uint64_t> h_index;
	BuildSeg *segs = nullptr;
	uint64_t *index = nullptr;
	uint32_t *bm = nullptr, *rk = nullptr; // bitmaps (TW words), ranks of 8-word chunks (TW / 8 + 1)
	uint32_t *cnt = nullptr, *aux = nullptr; // per leaf block (nbins)
	uint32_t *node = nullptr;   // 2 * Ftot: {leaf block, place in its bucket}
	uint32_t *bucket = nullptr; // 2 * Ftot
	uint32_t *list = nullptr;   // 2 + 2 * cap: large leaf blocks
	const uint64_t *fbeg() const { return index; }
	const uint64_t *bbeg() const { return index + (S + 1); }
	const uint64_t *dw(uint32_t d) const { return index + (2 + d) * (uint64_t)(S + 1); }
	uint64_t h_dw(uint32_t d, uint32_t s) const { return h_index[(2 + d) * (uint64_t)(S + 1) + s]; }
	const uint64_t *dr(uint32_t d) const { return index + (3 + kMaxLevel + d) * (uint64_t)(S + 1); }
	uint64_t h_dr(uint32_t d, uint32_t s) const { return h_index[(3 + kMaxLevel + d) * (uint64_t)(S + 1) + s]; }
	// block writer's parent words, after the per-depth tables: (kMaxLevel - 2) S segments + 1 entries
	uint64_t pw_at() const { return (3 + 2 * (uint64_t)kMaxLevel + 1) * (S + 1); }
	uint32_t pw_segs() const { return (kMaxLevel - 2u) * S; }
	const uint64_t *pw() const { return index + pw_at(); }
};

void vox_batch_release(svo_ctx *ctx, VoxBatch *vb);
void build_batch_release(svo_ctx *ctx, BuildBatch *bb);

// Voxelize: count pass over the batch (one read-back of every scene's F), then the fill.
int voxelize_count_batch(svo_ctx *ctx, svo_fraglist *const *fls, uint32_t S, VoxBatch *vb);
int voxelize_fill_batch(svo_ctx *ctx, VoxBatch *vb);

// Build: structure phase (one read-back of every scene's W), octree allocation, node phase.
int build_batch(svo_ctx *ctx, svo_octree *const *octs, uint32_t S);

// Voxelize passes of a captured plan: tables over lists the plan allocated, then the count pass
// (fend[0] = total) and the fill (limited by fend[1]) with no read-back between them.
int voxelize_plan_init(svo_ctx *ctx, VoxBatch *vb, svo_fraglist *const *fls, uint32_t S);
hipError_t voxelize_plan_count(svo_ctx *ctx, VoxBatch *vb);
hipError_t voxelize_plan_fill(svo_ctx *ctx, VoxBatch *vb);

// A voxelize + build of a fixed set of scenes whose fragment lists and octrees are allocated at
// capacities, so no pass waits for the host: the counts stay on the device, the grids and scratch
// are sized by the capacities, and the whole sequence is captured once as a HIP graph and
// replayed.  A count above its capacity empties the rest of the run and is reported by finish.
struct BuildPlan {
	uint32_t S = 0;
	VoxBatch vb;
	BuildBatch bb; // h_segs[s].F holds the fragment capacities
	std::vector<svo_fraglist *> fls;
	std::vector<svo_octree *> octs; // cap = word capacity
	uint64_t *caps = nullptr;       // device: fragment capacities [S], word capacities [S]
	uint64_t *out = nullptr;        // device: status (bit 1 fragments, bit 2 words), F[S], W[S]
	std::vector<uint64_t> h_out;
	uint64_t pw_value = 0; // the block writer's parent-word total (fixed by the levels)
	bool use_graph = true, pending = false;
	hipGraphExec_t exec = nullptr;
	StageTimer t_run;
};
int build_plan_init(svo_ctx *ctx, BuildPlan *p, const svo_scene *const *scenes, const uint32_t *levels, uint32_t S,
                    const uint64_t *fcap, const uint64_t *wcap);
int build_plan_run(svo_ctx *ctx, BuildPlan *p);
int build_plan_finish(svo_ctx *ctx, BuildPlan *p, uint64_t *F, uint64_t *W, double *ms);
void build_plan_release(svo_ctx *ctx, BuildPlan *p);

// A reusable batch march of S octrees, one camera each, width x height primary rays each: device
// table of octrees and cameras, beam scratch, the launch shape (default or autotuned), and
// optionally a captured HIP graph of the launches, replayed while the hit buffer stays the same.
struct MarchPlan {
	uint32_t S = 0, width = 0, height = 0, bw = 0, bh = 0;
	int beam = 0;
	std::vector<MarchSeg> h_segs;
	MarchSeg *segs = nullptr;
	float *beam_buf = nullptr;
	int shape = -1;
	bool use_graph = false;
	hipGraphExec_t exec = nullptr;
	const void *exec_hits = nullptr;
};
int march_plan_init(svo_ctx *ctx, MarchPlan *p, const svo_octree *const *octs, const svo_camera *cams, uint32_t S,
                    uint32_t width, uint32_t height, int beam);
int march_plan_set_cameras(svo_ctx *ctx, MarchPlan *p, const svo_camera *cams);
int march_plan_run(svo_ctx *ctx, MarchPlan *p, svo_hit *hits);
void march_plan_release(svo_ctx *ctx, MarchPlan *p);
// Explicit rays against S octrees in one launch; ray_begin (host, S + 1 entries) splits rays.
int march_rays_batch(svo_ctx *ctx, const svo_octree *const *octs, uint32_t S, const uint64_t *ray_begin,
                     const float *rays, int variant, svo_hit *hits);
// Same with separate buffers per octree: rays[s] (count[s] rays) -> hits[s].
int march_rays_segments(svo_ctx *ctx, const svo_octree *const *octs, const float *const *rays, svo_hit *const *hits,
                        const uint64_t *count, uint32_t S, int variant);

} // namespace svo

struct svo_march_plan {
	svo_ctx *ctx;
	svo::MarchPlan p;
};

struct svo_build_plan {
	svo_ctx *ctx;
	svo::BuildPlan p;
};
