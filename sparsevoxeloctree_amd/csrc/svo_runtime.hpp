// Host runtime of the SVO library: context (one HIP stream on one device), caching device
// allocator, per-kernel profiler and stage timers.  Replaces the Vulkan device / queue / command
// pool plumbing of the reference (src/Application.cpp:161-293, dep/MyVK).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/svo.h"

namespace svo {

// Device-side build state (the reference's build-info + counter buffers,
// src/OctreeBuilder.cpp:24-39).  Slot (L & 1) holds the leaf level's range {uAllocBegin, uAllocNum}
// and the block counter uCounter after the structure phase; need_words = W.
struct BuildState {
	uint32_t begin[2];   // uAllocBegin: first word of the level's range
	uint32_t num[2];     // uAllocNum: words in the level's range
	uint32_t counter[2]; // uCounter: 8-word blocks allocated below the root block
	uint32_t overflow;   // non-zero when W exceeds 2^30 words
	uint32_t pad;
	uint64_t need_words; // W
};

// Error bits raised by kernels into the context's error word.
enum DeviceError : uint32_t {
	kErrIterCap = 1u,
	kErrOutOfRange = 2u,
	kErrTooDeep = 4u,
};

class Pool {
public:
	explicit Pool(int /*device: allocations follow the caller's current device*/) {}
	~Pool();
	hipError_t alloc(void **p, uint64_t bytes);
	void free(void *p);
	void trim();
	uint64_t in_use() const { return m_in_use; }
	uint64_t cached() const { return m_cached; }
	uint64_t peak() const { return m_peak; }

private:
	static uint64_t round_size(uint64_t bytes);
	std::multimap<uint64_t, void *> m_free;
	std::unordered_map<void *, uint64_t> m_live;
	uint64_t m_in_use = 0, m_cached = 0, m_peak = 0;
};

struct ProfEntry {
	std::string name;
	uint64_t launches = 0;
	double total_ms = 0.0;
};

class Profiler {
public:
	bool enabled = false;
	void begin(hipStream_t s, const char *name);
	void end(hipStream_t s);
	void resolve(); // waits for pending events
	void reset();
	std::vector<ProfEntry> entries;

private:
	struct Pending {
		int entry;
		hipEvent_t a, b;
	};
	std::vector<Pending> m_pending;
	std::vector<hipEvent_t> m_free_events;
	int m_open = -1;
	hipEvent_t m_open_ev = nullptr;
	hipEvent_t get_event();
};

// Pinned host staging for small tables copied to or from the device in stream order (segment
// descriptors, per-scene results).  A ring of slots, each guarded by an event recorded after its
// last copy, so a slot is reused only once that copy has executed.
class Staging {
public:
	static constexpr uint32_t kSlots = 8;
	static constexpr size_t kSlotBytes = size_t(1) << 20; // a build batch of 1024 scenes: 0.55 MB
	hipError_t init();
	void destroy();
	// Next free slot (waits for its previous copy); null when bytes exceed a slot.
	void *acquire(size_t bytes);
	// Marks the slot returned by the last acquire as in use by work enqueued on s so far.
	hipError_t release(hipStream_t s);

private:
	uint8_t *m_buf = nullptr;
	hipEvent_t m_ev[kSlots] = {};
	bool m_armed[kSlots] = {};
	uint32_t m_next = 0, m_cur = 0;
};

// Stage timer: a pair of events around a stage, read lazily.
struct StageTimer {
	hipEvent_t a = nullptr, b = nullptr;
	bool armed = false;
	void start(hipStream_t s);
	void stop(hipStream_t s);
	float read(); // ms, 0 when never armed
	void destroy();
};

} // namespace svo

struct svo_ctx {
	int device = 0;
	hipStream_t stream = nullptr;
	std::string last_error;
	svo::Pool *pool = nullptr;
	svo::Profiler prof;
	// small device scratch: scan partials, build state, error word
	uint64_t *d_scan_parts = nullptr; // 2 * kScanBlocks uint64
	svo::BuildState *d_state = nullptr;
	uint32_t *d_err = nullptr;
	uint32_t *d_sobol = nullptr; // Sobol state of the path tracer's current sample: 64 words + counter
	float *noalloc_beam = nullptr; // device address of this device's static beam array (svo_march_noalloc)
	svo::BuildState *h_state = nullptr; // pinned mirror for read-back
	svo::StageTimer t_normalize, t_count, t_fill, t_build, t_build2, t_beam, t_march, t_rays, t_bounce, t_path;
	svo::Staging stage;
	int cu_count = 256;
	// launch-shape autotuner: off by default; when on, a batched launch of a new shape class times
	// its candidate shapes once and keeps the fastest
	bool autotune = false;
	// march schedule: 1 = persistent waves with per-lane ray refill, 0 = one ray per lane,
	// -1 = chosen per shape class by the autotuner when it is on, else one ray per lane
	int march_mode = -1;
	std::map<std::string, int> tune_choice;
	std::string tune_log;
};

namespace svo {
struct VoxBatch;
}

struct svo_scene {
	svo_ctx *ctx;
	uint64_t T;
	float *xyz;
	uint32_t *rgb;
};

struct svo_fraglist {
	svo_ctx *ctx;
	const svo_scene *scene; // null when created from host fragments
	uint32_t level;
	uint64_t F;
	uint32_t *frags;   // 2 * F words
	svo::VoxBatch *vb; // count-pass state until the fill (single-scene voxelize), or null
	bool filled;
};

struct svo_octree {
	svo_ctx *ctx;
	uint32_t level;
	uint64_t W;        // used words
	uint64_t cap;      // words allocated
	uint32_t *words;
	const svo_fraglist *src; // fragment list of an enqueued build, or null
	bool pending;            // build enqueued, not yet finished
};

// An environment map (EnvironmentMap, src/EnvironmentMap.cpp:57-70): texels rgb + sampling pdf
// (the reference keeps the pdf in alpha) and the alias table, in HBM.
struct svo_envmap {
	svo_ctx *ctx;
	uint32_t width, height;
	float *texels;    // 4 floats per texel
	uint32_t *alias;  // {prob, alias} per texel
};

namespace svo {

// Shared helpers for the entry points.
int set_error(svo_ctx *ctx, int code, const std::string &msg);
int hip_check(svo_ctx *ctx, hipError_t e, const char *what);

constexpr int kScanBlocks = 1024;  // partial sums per device-wide scan
constexpr int kScanThreads = 256;

// Launch helper: brackets the launch with profiler events when profiling is on.
template <class Kernel, class... Args>
inline hipError_t launch(svo_ctx *ctx, const char *name, Kernel k, dim3 grid, dim3 block, size_t lds,
                         Args... args) {
	if (ctx->prof.enabled)
		ctx->prof.begin(ctx->stream, name);
	hipLaunchKernelGGL(k, grid, block, lds, ctx->stream, args...);
	hipError_t e = hipGetLastError();
	if (ctx->prof.enabled)
		ctx->prof.end(ctx->stream);
	return e;
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Stream-ordered copy of a small host table into device memory through the staging ring.
int upload_table(svo_ctx *ctx, void *dst, const void *src, size_t bytes);
// Copy of a small device table to the host; waits for the stream.
int download_table(svo_ctx *ctx, void *dst, const void *src, size_t bytes);

// Entry points implemented per stage.
int normalize_scene(svo_ctx *ctx, svo_scene *scene);
uint64_t build_bitmap_words(uint32_t level);
int march_frame(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t w, uint32_t h, int beam,
                svo_hit *hits, float *beam_out);
int march_beam(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t w, uint32_t h, float *beam,
               uint32_t *nfetch);
// Allocation-free primary march: noalloc_init looks up the device address of the static beam array
// for the context's device (once per device; the device must be current), march_frame_noalloc
// marches with the beam image there.
int noalloc_init(svo_ctx *ctx);
int march_frame_noalloc(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t w, uint32_t h,
                        svo_hit *hits);
int march_rays(svo_ctx *ctx, const svo_octree *oct, const float *rays, uint64_t R, int variant, svo_hit *hits);
int bounce_rays(svo_ctx *ctx, const svo_hit *hits, uint32_t width, uint32_t height, uint32_t spp, uint32_t cur,
                float *rays);
int path_trace(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t w, uint32_t h, uint32_t bounce,
               uint32_t spp, const float sky[3], float *color, uint32_t *albedo, uint32_t *normal);
int path_trace_env(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t w, uint32_t h,
                   uint32_t bounce, uint32_t spp, const svo_envmap *env, float rotation, float multiplier,
                   float *color, uint32_t *albedo, uint32_t *normal);
int shade_frame(svo_ctx *ctx, const svo_hit *hits, const svo_camera *cam, uint32_t w, uint32_t h, uint32_t view,
                const float light[3], const svo_envmap *env, float rotation, uint32_t *out);
int copy_device(svo_ctx *ctx, void *dst, const void *src, uint64_t bytes);

// Device-wide exclusive scan of n (read from the device) uint32 counts; see svo_scan.hpp.
} // namespace svo
