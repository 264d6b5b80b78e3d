// C ABI of the SVO library (include/svo.h) and the host runtime pieces (allocator, profiler,
// stage timers).  Compiled by hipcc with the kernels into libsvo.so.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include "svo_batch.hpp"
#include "svo_runtime.hpp"

namespace svo {

int set_error(svo_ctx *ctx, int code, const std::string &msg) {
	if (ctx)
		ctx->last_error = msg;
	return code;
}

int hip_check(svo_ctx *ctx, hipError_t e, const char *what) {
	if (e == hipSuccess)
		return SVO_OK;
	std::string m = std::string(what) + ": " + hipGetErrorString(e);
	return set_error(ctx, e == hipErrorOutOfMemory ? SVO_ERR_OUT_OF_MEMORY : SVO_ERR_HIP, m);
}

// ---- Pool: size-class caching allocator.  Blocks are reused only by later work on the owning
// context's stream, so reuse is ordered after every earlier use without extra synchronisation.
Pool::~Pool() {
	for (auto &kv : m_free)
		(void)hipFree(kv.second);
	for (auto &kv : m_live)
		(void)hipFree(kv.first);
}

uint64_t Pool::round_size(uint64_t bytes) {
	if (bytes < 256)
		bytes = 256;
	if (bytes <= (64ull << 20)) {
		uint64_t p = 256;
		while (p < bytes)
			p <<= 1;
		return p;
	}
	const uint64_t g = 64ull << 20;
	return (bytes + g - 1) / g * g;
}

hipError_t Pool::alloc(void **p, uint64_t bytes) {
	uint64_t sz = round_size(bytes);
	auto it = m_free.find(sz);
	// a large request takes the smallest cached block of up to 1.5x its size: batches of varying
	// composition ask for slightly different sizes, and a fresh hipMalloc / hipFree of gigabytes
	// costs milliseconds (hipFree also waits for the device)
	if (it == m_free.end() && sz > (64ull << 20)) {
		it = m_free.lower_bound(sz);
		if (it != m_free.end() && it->first > sz + sz / 2)
			it = m_free.end();
	}
	if (it != m_free.end()) {
		sz = it->first;
		*p = it->second;
		m_free.erase(it);
		m_cached -= sz;
	} else {
		hipError_t e = hipMalloc(p, sz);
		if (e != hipSuccess) {
			(void)hipGetLastError();
			trim(); // give cached blocks back and retry once
			e = hipMalloc(p, sz);
			if (e != hipSuccess) {
				(void)hipGetLastError();
				*p = nullptr;
				return e;
			}
		}
	}
	m_live[*p] = sz;
	m_in_use += sz;
	m_peak = std::max(m_peak, m_in_use);
	return hipSuccess;
}

void Pool::free(void *p) {
	if (!p)
		return;
	auto it = m_live.find(p);
	if (it == m_live.end())
		return;
	m_in_use -= it->second;
	m_cached += it->second;
	m_free.emplace(it->second, p);
	m_live.erase(it);
}

void Pool::trim() {
	for (auto &kv : m_free)
		(void)hipFree(kv.second);
	m_free.clear();
	m_cached = 0;
}

// ---- Profiler
hipEvent_t Profiler::get_event() {
	if (!m_free_events.empty()) {
		hipEvent_t e = m_free_events.back();
		m_free_events.pop_back();
		return e;
	}
	hipEvent_t e = nullptr;
	(void)hipEventCreate(&e);
	return e;
}

void Profiler::begin(hipStream_t s, const char *name) {
	int idx = -1;
	for (size_t i = 0; i < entries.size(); ++i)
		if (entries[i].name == name) {
			idx = (int)i;
			break;
		}
	if (idx < 0) {
		entries.push_back(ProfEntry{name, 0, 0.0});
		idx = (int)entries.size() - 1;
	}
	m_open = idx;
	m_open_ev = get_event();
	(void)hipEventRecord(m_open_ev, s);
}

void Profiler::end(hipStream_t s) {
	if (m_open < 0)
		return;
	hipEvent_t b = get_event();
	(void)hipEventRecord(b, s);
	m_pending.push_back(Pending{m_open, m_open_ev, b});
	m_open = -1;
	m_open_ev = nullptr;
	if (m_pending.size() > 4096)
		resolve();
}

void Profiler::resolve() {
	for (auto &p : m_pending) {
		(void)hipEventSynchronize(p.b);
		float ms = 0.0f;
		if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
			entries[p.entry].launches += 1;
			entries[p.entry].total_ms += ms;
		}
		m_free_events.push_back(p.a);
		m_free_events.push_back(p.b);
	}
	m_pending.clear();
}

void Profiler::reset() {
	resolve();
	entries.clear();
}

void StageTimer::start(hipStream_t s) {
	if (!a) {
		(void)hipEventCreate(&a);
		(void)hipEventCreate(&b);
	}
	(void)hipEventRecord(a, s);
	armed = false;
}
void StageTimer::stop(hipStream_t s) {
	if (!b)
		return;
	(void)hipEventRecord(b, s);
	armed = true;
}
float StageTimer::read() {
	if (!armed)
		return 0.0f;
	(void)hipEventSynchronize(b);
	float ms = 0.0f;
	if (hipEventElapsedTime(&ms, a, b) != hipSuccess)
		return 0.0f;
	return ms;
}
void StageTimer::destroy() {
	if (a)
		(void)hipEventDestroy(a);
	if (b)
		(void)hipEventDestroy(b);
	a = b = nullptr;
}

// ---- Staging: pinned ring for small stream-ordered table copies
hipError_t Staging::init() {
	hipError_t e = hipHostMalloc((void **)&m_buf, kSlots * kSlotBytes, hipHostMallocDefault);
	if (e != hipSuccess)
		return e;
	for (uint32_t i = 0; i < kSlots; ++i)
		if ((e = hipEventCreateWithFlags(&m_ev[i], hipEventDisableTiming)) != hipSuccess)
			return e;
	return hipSuccess;
}

void Staging::destroy() {
	for (uint32_t i = 0; i < kSlots; ++i)
		if (m_ev[i]) {
			(void)hipEventSynchronize(m_ev[i]);
			(void)hipEventDestroy(m_ev[i]);
			m_ev[i] = nullptr;
		}
	if (m_buf)
		(void)hipHostFree(m_buf);
	m_buf = nullptr;
}

void *Staging::acquire(size_t bytes) {
	if (!m_buf || bytes > kSlotBytes)
		return nullptr;
	m_cur = m_next;
	m_next = (m_next + 1) % kSlots;
	if (m_armed[m_cur]) {
		(void)hipEventSynchronize(m_ev[m_cur]);
		m_armed[m_cur] = false;
	}
	return m_buf + m_cur * kSlotBytes;
}

hipError_t Staging::release(hipStream_t s) {
	hipError_t e = hipEventRecord(m_ev[m_cur], s);
	m_armed[m_cur] = e == hipSuccess;
	return e;
}

int upload_table(svo_ctx *ctx, void *dst, const void *src, size_t bytes) {
	if (bytes == 0)
		return SVO_OK;
	void *slot = ctx->stage.acquire(bytes);
	if (!slot)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "table larger than the staging slot");
	std::memcpy(slot, src, bytes);
	hipError_t e = hipMemcpyAsync(dst, slot, bytes, hipMemcpyHostToDevice, ctx->stream);
	if (e == hipSuccess)
		e = ctx->stage.release(ctx->stream);
	return hip_check(ctx, e, "table upload");
}

int download_table(svo_ctx *ctx, void *dst, const void *src, size_t bytes) {
	if (bytes == 0)
		return SVO_OK;
	void *slot = ctx->stage.acquire(bytes);
	if (!slot)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "table larger than the staging slot");
	hipError_t e = hipMemcpyAsync(slot, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
	if (e == hipSuccess)
		e = hipStreamSynchronize(ctx->stream);
	if (e != hipSuccess)
		return hip_check(ctx, e, "table read-back");
	std::memcpy(dst, slot, bytes);
	return SVO_OK;
}

} // namespace svo

using namespace svo;

namespace {

struct DeviceGuard { // makes the context's device current for the call
	int prev = -1;
	explicit DeviceGuard(const svo_ctx *ctx) {
		if (hipGetDevice(&prev) != hipSuccess)
			prev = -1;
		if (ctx && prev != ctx->device)
			(void)hipSetDevice(ctx->device);
	}
	~DeviceGuard() {
		int cur = -1;
		if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev)
			(void)hipSetDevice(prev);
	}
};

} // namespace

extern "C" {

int svo_ctx_create(int device, svo_ctx **out) {
	if (!out)
		return SVO_ERR_INVALID_ARGUMENT;
	*out = nullptr;
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
		return SVO_ERR_NO_DEVICE;
	svo_ctx *ctx = new (std::nothrow) svo_ctx();
	if (!ctx)
		return SVO_ERR_OUT_OF_MEMORY;
	ctx->device = device;
	DeviceGuard g(ctx);
	hipError_t e;
	if ((e = hipSetDevice(device)) != hipSuccess || (e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess) {
		delete ctx;
		return SVO_ERR_HIP;
	}
	hipDeviceProp_t prop;
	if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
		ctx->cu_count = prop.multiProcessorCount;
	ctx->pool = new Pool(device);
	if (hipMalloc((void **)&ctx->d_scan_parts, sizeof(uint64_t) * (2 * kScanBlocks + 8)) != hipSuccess ||
	    hipMalloc((void **)&ctx->d_state, sizeof(BuildState)) != hipSuccess ||
	    hipMalloc((void **)&ctx->d_err, sizeof(uint32_t)) != hipSuccess ||
	    hipMalloc((void **)&ctx->d_sobol, sizeof(uint32_t) * 65) != hipSuccess ||
	    hipHostMalloc((void **)&ctx->h_state, sizeof(BuildState) + 64, hipHostMallocDefault) != hipSuccess ||
	    ctx->stage.init() != hipSuccess) {
		svo_ctx_destroy(ctx);
		return SVO_ERR_OUT_OF_MEMORY;
	}
	if (noalloc_init(ctx) != SVO_OK) {
		svo_ctx_destroy(ctx);
		return SVO_ERR_HIP;
	}
	(void)hipMemsetAsync(ctx->d_err, 0, sizeof(uint32_t), ctx->stream);
	(void)hipMemsetAsync(ctx->d_state, 0, sizeof(BuildState), ctx->stream);
	if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
		svo_ctx_destroy(ctx);
		return SVO_ERR_HIP;
	}
	*out = ctx;
	return SVO_OK;
}

int svo_ctx_destroy(svo_ctx *ctx) {
	if (!ctx)
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	if (ctx->stream)
		(void)hipStreamSynchronize(ctx->stream);
	ctx->prof.reset();
	for (StageTimer *t : {&ctx->t_normalize, &ctx->t_count, &ctx->t_fill, &ctx->t_build, &ctx->t_build2, &ctx->t_beam, &ctx->t_march,
	                      &ctx->t_rays, &ctx->t_bounce, &ctx->t_path})
		t->destroy();
	if (ctx->d_scan_parts)
		(void)hipFree(ctx->d_scan_parts);
	if (ctx->d_state)
		(void)hipFree(ctx->d_state);
	if (ctx->d_err)
		(void)hipFree(ctx->d_err);
	if (ctx->d_sobol)
		(void)hipFree(ctx->d_sobol);
	if (ctx->h_state)
		(void)hipHostFree(ctx->h_state);
	ctx->stage.destroy();
	delete ctx->pool;
	if (ctx->stream)
		(void)hipStreamDestroy(ctx->stream);
	delete ctx;
	return SVO_OK;
}

const char *svo_last_error(const svo_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int svo_ctx_synchronize(svo_ctx *ctx) {
	if (!ctx)
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	return hip_check(ctx, hipStreamSynchronize(ctx->stream), "synchronize");
}

void *svo_ctx_stream(svo_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int svo_ctx_take_device_errors(svo_ctx *ctx, uint32_t *bits) {
	if (!ctx || !bits)
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	hipError_t e;
	if ((e = hipMemcpyAsync(&ctx->h_state->pad, ctx->d_err, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream)) !=
	        hipSuccess ||
	    (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
		return hip_check(ctx, e, "read error word");
	*bits = ctx->h_state->pad;
	if ((e = hipMemsetAsync(ctx->d_err, 0, sizeof(uint32_t), ctx->stream)) != hipSuccess)
		return hip_check(ctx, e, "clear error word");
	return hip_check(ctx, hipStreamSynchronize(ctx->stream), "clear error word");
}

int svo_ctx_trim(svo_ctx *ctx) {
	if (!ctx)
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	(void)hipStreamSynchronize(ctx->stream);
	ctx->pool->trim();
	return SVO_OK;
}

int svo_ctx_memory(svo_ctx *ctx, uint64_t *in_use, uint64_t *cached, uint64_t *peak) {
	if (!ctx)
		return SVO_ERR_INVALID_ARGUMENT;
	if (in_use)
		*in_use = ctx->pool->in_use();
	if (cached)
		*cached = ctx->pool->cached();
	if (peak)
		*peak = ctx->pool->peak();
	return SVO_OK;
}

int svo_malloc(svo_ctx *ctx, uint64_t bytes, void **dptr) {
	if (!ctx || !dptr)
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	if (ctx->pool->alloc(dptr, bytes) != hipSuccess)
		return set_error(ctx, SVO_ERR_OUT_OF_MEMORY, "svo_malloc failed");
	return SVO_OK;
}

int svo_free(svo_ctx *ctx, void *dptr) {
	if (!ctx)
		return SVO_ERR_INVALID_ARGUMENT;
	ctx->pool->free(dptr);
	return SVO_OK;
}

int svo_copy_device(svo_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
	if (!ctx || (bytes && (!dst || !src)))
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	return copy_device(ctx, dst, src, bytes);
}

int svo_upload(svo_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
	if (!ctx || (bytes && (!dst || !src)))
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
	if (e == hipSuccess)
		e = hipStreamSynchronize(ctx->stream);
	return hip_check(ctx, e, "upload");
}

int svo_download(svo_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
	if (!ctx || (bytes && (!dst || !src)))
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
	if (e == hipSuccess)
		e = hipStreamSynchronize(ctx->stream);
	return hip_check(ctx, e, "download");
}

int svo_prof_enable(svo_ctx *ctx, int on) {
	if (!ctx)
		return SVO_ERR_INVALID_ARGUMENT;
	ctx->prof.enabled = on != 0;
	return SVO_OK;
}
int svo_prof_reset(svo_ctx *ctx) {
	if (!ctx)
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	ctx->prof.reset();
	return SVO_OK;
}
int svo_prof_count(svo_ctx *ctx, int *n) {
	if (!ctx || !n)
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	ctx->prof.resolve();
	*n = (int)ctx->prof.entries.size();
	return SVO_OK;
}
int svo_prof_get(svo_ctx *ctx, int i, const char **name, uint64_t *launches, double *total_ms) {
	if (!ctx || i < 0 || i >= (int)ctx->prof.entries.size())
		return SVO_ERR_INVALID_ARGUMENT;
	const ProfEntry &p = ctx->prof.entries[i];
	if (name)
		*name = p.name.c_str();
	if (launches)
		*launches = p.launches;
	if (total_ms)
		*total_ms = p.total_ms;
	return SVO_OK;
}

int svo_last_timing(svo_ctx *ctx, svo_timing *out) {
	if (!ctx || !out)
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	out->normalize_ms = ctx->t_normalize.read();
	out->vox_count_ms = ctx->t_count.read();
	out->vox_fill_ms = ctx->t_fill.read();
	out->build_ms = ctx->t_build.read() + ctx->t_build2.read(); // structure + node phases
	out->beam_ms = ctx->t_beam.read();
	out->march_ms = ctx->t_march.read();
	out->rays_ms = ctx->t_rays.read();
	out->bounce_ms = ctx->t_bounce.read();
	return SVO_OK;
}

// ---- Scene
static int scene_make(svo_ctx *ctx, uint64_t T, svo_scene **out, svo_scene *&s) {
	s = new (std::nothrow) svo_scene();
	if (!s)
		return SVO_ERR_OUT_OF_MEMORY;
	s->ctx = ctx;
	s->T = T;
	if (ctx->pool->alloc((void **)&s->xyz, sizeof(float) * 9 * (T ? T : 1)) != hipSuccess ||
	    ctx->pool->alloc((void **)&s->rgb, sizeof(uint32_t) * (T ? T : 1)) != hipSuccess) {
		svo_scene_destroy(s);
		return set_error(ctx, SVO_ERR_OUT_OF_MEMORY, "scene allocation failed");
	}
	(void)out;
	return SVO_OK;
}

int svo_scene_create(svo_ctx *ctx, const float *tri_xyz, const uint32_t *tri_rgb, uint64_t T, int normalize,
                     svo_scene **out) {
	if (!ctx || !out || (T && (!tri_xyz || !tri_rgb)))
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	*out = nullptr;
	svo_scene *s;
	int rc = scene_make(ctx, T, out, s);
	if (rc)
		return rc;
	hipError_t e = hipSuccess;
	if (T) {
		if ((e = hipMemcpyAsync(s->xyz, tri_xyz, sizeof(float) * 9 * T, hipMemcpyHostToDevice, ctx->stream)) ==
		    hipSuccess)
			e = hipMemcpyAsync(s->rgb, tri_rgb, sizeof(uint32_t) * T, hipMemcpyHostToDevice, ctx->stream);
	}
	if (e == hipSuccess && normalize && T) {
		rc = normalize_scene(ctx, s);
		if (rc) {
			svo_scene_destroy(s);
			return rc;
		}
	}
	if (e == hipSuccess)
		e = hipStreamSynchronize(ctx->stream);
	if (e != hipSuccess) {
		svo_scene_destroy(s);
		return hip_check(ctx, e, "scene upload");
	}
	*out = s;
	return SVO_OK;
}

int svo_scene_create_device(svo_ctx *ctx, const float *tri_xyz_dev, const uint32_t *tri_rgb_dev, uint64_t T,
                            int normalize, svo_scene **out) {
	if (!ctx || !out || (T && (!tri_xyz_dev || !tri_rgb_dev)))
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	*out = nullptr;
	svo_scene *s;
	int rc = scene_make(ctx, T, out, s);
	if (rc)
		return rc;
	hipError_t e = hipSuccess;
	if (T) {
		if ((e = hipMemcpyAsync(s->xyz, tri_xyz_dev, sizeof(float) * 9 * T, hipMemcpyDeviceToDevice, ctx->stream)) ==
		    hipSuccess)
			e = hipMemcpyAsync(s->rgb, tri_rgb_dev, sizeof(uint32_t) * T, hipMemcpyDeviceToDevice, ctx->stream);
	}
	if (e != hipSuccess) {
		svo_scene_destroy(s);
		return hip_check(ctx, e, "scene copy");
	}
	if (normalize && T) {
		rc = normalize_scene(ctx, s);
		if (rc) {
			svo_scene_destroy(s);
			return rc;
		}
	}
	*out = s;
	return SVO_OK;
}

static int scene_update(svo_ctx *ctx, svo_scene *s, const float *xyz, const uint32_t *rgb, int normalize,
                        hipMemcpyKind kind) {
	if (!ctx || !s || s->ctx != ctx || (s->T && (!xyz || !rgb)))
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	const uint64_t T = s->T;
	hipError_t e = hipSuccess;
	if (T && (e = hipMemcpyAsync(s->xyz, xyz, sizeof(float) * 9 * T, kind, ctx->stream)) == hipSuccess)
		e = hipMemcpyAsync(s->rgb, rgb, sizeof(uint32_t) * T, kind, ctx->stream);
	if (e != hipSuccess)
		return hip_check(ctx, e, "scene update");
	if (normalize && T)
		if (int rc = normalize_scene(ctx, s))
			return rc;
	// host arrays may be reused once the call returns
	if (kind == hipMemcpyHostToDevice && (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
		return hip_check(ctx, e, "scene update");
	return SVO_OK;
}

int svo_scene_update(svo_ctx *ctx, svo_scene *scene, const float *tri_xyz, const uint32_t *tri_rgb, int normalize) {
	return scene_update(ctx, scene, tri_xyz, tri_rgb, normalize, hipMemcpyHostToDevice);
}

int svo_scene_update_device(svo_ctx *ctx, svo_scene *scene, const float *tri_xyz_dev, const uint32_t *tri_rgb_dev,
                            int normalize) {
	return scene_update(ctx, scene, tri_xyz_dev, tri_rgb_dev, normalize, hipMemcpyDeviceToDevice);
}

int svo_scene_info(const svo_scene *s, uint64_t *T, const float **xyz, const uint32_t **rgb) {
	if (!s)
		return SVO_ERR_INVALID_ARGUMENT;
	if (T)
		*T = s->T;
	if (xyz)
		*xyz = s->xyz;
	if (rgb)
		*rgb = s->rgb;
	return SVO_OK;
}

int svo_scene_download(svo_ctx *ctx, const svo_scene *s, float *tri_xyz) {
	if (!ctx || !s || (s->T && !tri_xyz))
		return SVO_ERR_INVALID_ARGUMENT;
	return svo_download(ctx, tri_xyz, s->xyz, sizeof(float) * 9 * s->T);
}

int svo_scene_destroy(svo_scene *s) {
	if (!s)
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(s->ctx);
	s->ctx->pool->free(s->xyz);
	s->ctx->pool->free(s->rgb);
	delete s;
	return SVO_OK;
}

// ---- Voxelizer
int svo_voxelize_count(svo_ctx *ctx, const svo_scene *scene, uint32_t level, svo_fraglist **out) {
	if (!ctx || !scene || !out)
		return SVO_ERR_INVALID_ARGUMENT;
	*out = nullptr;
	if (level < 1 || level > 12)
		return set_error(ctx, SVO_ERR_BAD_LEVEL, "level must be in [1, 12]");
	DeviceGuard g(ctx);
	svo_fraglist *fl = new (std::nothrow) svo_fraglist();
	VoxBatch *vb = new (std::nothrow) VoxBatch();
	if (!fl || !vb) {
		delete fl;
		delete vb;
		return SVO_ERR_OUT_OF_MEMORY;
	}
	fl->ctx = ctx;
	fl->scene = scene;
	fl->level = level;
	fl->vb = vb;
	int rc = voxelize_count_batch(ctx, &fl, 1, vb);
	if (rc) {
		svo_fraglist_destroy(fl);
		return rc;
	}
	*out = fl;
	return SVO_OK;
}

int svo_voxelize_fill(svo_ctx *ctx, svo_fraglist *fl) {
	if (!ctx || !fl || !fl->scene || !fl->vb)
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	int rc = voxelize_fill_batch(ctx, fl->vb);
	vox_batch_release(ctx, fl->vb); // stream-ordered: later work only
	delete fl->vb;
	fl->vb = nullptr;
	if (!rc)
		fl->filled = true;
	return rc;
}

int svo_voxelize(svo_ctx *ctx, const svo_scene *scene, uint32_t level, svo_fraglist **out) {
	int rc = svo_voxelize_count(ctx, scene, level, out);
	if (rc)
		return rc;
	rc = svo_voxelize_fill(ctx, *out);
	if (rc) {
		svo_fraglist_destroy(*out);
		*out = nullptr;
	}
	return rc;
}

int svo_voxelize_batch(svo_ctx *ctx, const svo_scene *const *scenes, const uint32_t *levels, uint32_t n,
                       svo_fraglist **out) {
	if (!ctx || (n && (!scenes || !levels || !out)))
		return SVO_ERR_INVALID_ARGUMENT;
	if (n > kMaxBatch)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "batch larger than 1024 scenes");
	for (uint32_t i = 0; i < n; ++i) {
		out[i] = nullptr;
		if (!scenes[i])
			return SVO_ERR_INVALID_ARGUMENT;
		if (levels[i] < 1 || levels[i] > 12)
			return set_error(ctx, SVO_ERR_BAD_LEVEL, "level must be in [1, 12]");
	}
	if (n == 0)
		return SVO_OK;
	DeviceGuard g(ctx);
	std::vector<svo_fraglist *> fls(n, nullptr);
	int rc = SVO_OK;
	for (uint32_t i = 0; i < n && !rc; ++i) {
		fls[i] = new (std::nothrow) svo_fraglist();
		if (!fls[i]) {
			rc = SVO_ERR_OUT_OF_MEMORY;
			break;
		}
		fls[i]->ctx = ctx;
		fls[i]->scene = scenes[i];
		fls[i]->level = levels[i];
	}
	VoxBatch vb;
	if (!rc)
		rc = voxelize_count_batch(ctx, fls.data(), n, &vb);
	if (!rc)
		rc = voxelize_fill_batch(ctx, &vb);
	vox_batch_release(ctx, &vb);
	for (uint32_t i = 0; i < n; ++i) {
		if (!fls[i])
			continue;
		if (rc) {
			svo_fraglist_destroy(fls[i]);
		} else {
			fls[i]->filled = true;
			out[i] = fls[i];
		}
	}
	return rc;
}

int svo_fraglist_create(svo_ctx *ctx, const uint32_t *frags, uint64_t F, uint32_t level, svo_fraglist **out) {
	if (!ctx || !out || (F && !frags))
		return SVO_ERR_INVALID_ARGUMENT;
	*out = nullptr;
	if (level < 1 || level > 12)
		return set_error(ctx, SVO_ERR_BAD_LEVEL, "level must be in [1, 12]");
	if (F > 0xFFFFFFFFull)
		return set_error(ctx, SVO_ERR_OVERFLOW, "more than 2^32-1 fragments");
	DeviceGuard g(ctx);
	svo_fraglist *fl = new (std::nothrow) svo_fraglist();
	if (!fl)
		return SVO_ERR_OUT_OF_MEMORY;
	fl->ctx = ctx;
	fl->level = level;
	fl->F = F;
	if (ctx->pool->alloc((void **)&fl->frags, sizeof(uint32_t) * 2 * (F ? F : 1)) != hipSuccess) {
		svo_fraglist_destroy(fl);
		return set_error(ctx, SVO_ERR_OUT_OF_MEMORY, "fragment list allocation failed");
	}
	int rc = svo_upload(ctx, fl->frags, frags, sizeof(uint32_t) * 2 * F);
	if (rc) {
		svo_fraglist_destroy(fl);
		return rc;
	}
	fl->filled = true;
	*out = fl;
	return SVO_OK;
}

int svo_fraglist_info(const svo_fraglist *fl, uint32_t *level, uint64_t *F, const uint32_t **dev) {
	if (!fl)
		return SVO_ERR_INVALID_ARGUMENT;
	if (level)
		*level = fl->level;
	if (F)
		*F = fl->F;
	if (dev)
		*dev = fl->frags;
	return SVO_OK;
}

int svo_fraglist_download(svo_ctx *ctx, const svo_fraglist *fl, uint32_t *out, uint64_t cap) {
	if (!ctx || !fl || (fl->F && !out))
		return SVO_ERR_INVALID_ARGUMENT;
	return svo_download(ctx, out, fl->frags, sizeof(uint32_t) * 2 * std::min(cap, fl->F));
}

int svo_fraglist_destroy(svo_fraglist *fl) {
	if (!fl)
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(fl->ctx);
	fl->ctx->pool->free(fl->frags);
	if (fl->vb) {
		vox_batch_release(fl->ctx, fl->vb);
		delete fl->vb;
	}
	delete fl;
	return SVO_OK;
}

// ---- OctreeBuilder
int svo_build_enqueue(svo_ctx *ctx, const svo_fraglist *fl, svo_octree **out) {
	if (!ctx || !fl || !out)
		return SVO_ERR_INVALID_ARGUMENT;
	*out = nullptr;
	if (!fl->filled)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "fragment list not filled (call svo_voxelize_fill)");
	DeviceGuard g(ctx);
	svo_octree *o = new (std::nothrow) svo_octree();
	if (!o)
		return SVO_ERR_OUT_OF_MEMORY;
	o->ctx = ctx;
	o->level = fl->level;
	o->src = fl;
	int rc = build_batch(ctx, &o, 1);
	if (rc) {
		svo_octree_destroy(o);
		return rc;
	}
	o->pending = true;
	*out = o;
	return SVO_OK;
}

int svo_build_finish(svo_ctx *ctx, svo_octree *o, uint64_t *W) {
	if (!ctx || !o)
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	if (o->pending) {
		hipError_t e = hipStreamSynchronize(ctx->stream);
		if (e != hipSuccess)
			return hip_check(ctx, e, "octree build");
		o->pending = false;
		o->src = nullptr;
	}
	if (W)
		*W = o->W;
	return SVO_OK;
}

int svo_build(svo_ctx *ctx, const svo_fraglist *fl, svo_octree **out) {
	int rc = svo_build_enqueue(ctx, fl, out);
	if (rc)
		return rc;
	rc = svo_build_finish(ctx, *out, nullptr);
	if (rc) {
		svo_octree_destroy(*out);
		*out = nullptr;
	}
	return rc;
}

int svo_build_batch(svo_ctx *ctx, const svo_fraglist *const *fls, uint32_t n, svo_octree **out) {
	if (!ctx || (n && (!fls || !out)))
		return SVO_ERR_INVALID_ARGUMENT;
	if (n > kMaxBatch)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "batch larger than 1024 scenes");
	for (uint32_t i = 0; i < n; ++i) {
		out[i] = nullptr;
		if (!fls[i])
			return SVO_ERR_INVALID_ARGUMENT;
		if (!fls[i]->filled)
			return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "fragment list not filled (call svo_voxelize_fill)");
	}
	if (n == 0)
		return SVO_OK;
	DeviceGuard g(ctx);
	std::vector<svo_octree *> octs(n, nullptr);
	int rc = SVO_OK;
	for (uint32_t i = 0; i < n; ++i) {
		octs[i] = new (std::nothrow) svo_octree();
		if (!octs[i]) {
			rc = SVO_ERR_OUT_OF_MEMORY;
			break;
		}
		octs[i]->ctx = ctx;
		octs[i]->level = fls[i]->level;
		octs[i]->src = fls[i];
	}
	if (!rc)
		rc = build_batch(ctx, octs.data(), n);
	for (uint32_t i = 0; i < n; ++i) {
		if (!octs[i])
			continue;
		if (rc) {
			svo_octree_destroy(octs[i]);
		} else {
			octs[i]->pending = true;
			out[i] = octs[i];
		}
	}
	return rc;
}

// ---- captured voxelize + build
int svo_build_plan_create(svo_ctx *ctx, const svo_scene *const *scenes, const uint32_t *levels, uint32_t n,
                          const uint64_t *frag_caps, const uint64_t *word_caps, svo_build_plan **out) {
	if (!ctx || !out || !n || !scenes || !levels)
		return SVO_ERR_INVALID_ARGUMENT;
	*out = nullptr;
	if (n > kMaxBatch)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "plan larger than 1024 scenes");
	for (uint32_t i = 0; i < n; ++i) {
		if (!scenes[i] || scenes[i]->ctx != ctx)
			return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "every scene of a plan must live on the plan's context");
		if (levels[i] < 1 || levels[i] > 12)
			return set_error(ctx, SVO_ERR_BAD_LEVEL, "level must be in [1, 12]");
	}
	DeviceGuard g(ctx);
	std::vector<uint64_t> fc(n), wc(n);
	if (!frag_caps || !word_caps) {
		// capacities from one plain voxelize + build of the scenes as they are now, plus a quarter
		std::vector<svo_fraglist *> fl(n, nullptr);
		std::vector<svo_octree *> oc(n, nullptr);
		int rc = svo_voxelize_batch(ctx, scenes, levels, n, fl.data());
		if (!rc)
			rc = svo_build_batch(ctx, fl.data(), n, oc.data());
		for (uint32_t i = 0; i < n; ++i) {
			if (!rc) {
				fc[i] = fl[i]->F + fl[i]->F / 4u + 64u;
				wc[i] = std::min<uint64_t>(oc[i]->W + oc[i]->W / 4u + 64u, 1ull << 30); // 30-bit child pointers
			}
			if (oc[i])
				svo_octree_destroy(oc[i]);
			if (fl[i])
				svo_fraglist_destroy(fl[i]);
		}
		if (rc)
			return rc;
	}
	for (uint32_t i = 0; i < n; ++i) {
		if (frag_caps)
			fc[i] = frag_caps[i];
		if (word_caps)
			wc[i] = word_caps[i];
	}
	svo_build_plan *p = new (std::nothrow) svo_build_plan();
	if (!p)
		return SVO_ERR_OUT_OF_MEMORY;
	p->ctx = ctx;
	int rc = build_plan_init(ctx, &p->p, scenes, levels, n, fc.data(), wc.data());
	if (rc) {
		svo_build_plan_destroy(p);
		return rc;
	}
	*out = p;
	return SVO_OK;
}

int svo_build_plan_use_graph(svo_build_plan *plan, int on) {
	if (!plan)
		return SVO_ERR_INVALID_ARGUMENT;
	plan->p.use_graph = on != 0;
	return SVO_OK;
}

int svo_build_plan_run(svo_build_plan *plan) {
	if (!plan)
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(plan->ctx);
	return build_plan_run(plan->ctx, &plan->p);
}

int svo_build_plan_finish(svo_build_plan *plan, uint64_t *F, uint64_t *W, double *run_ms) {
	if (!plan)
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(plan->ctx);
	return build_plan_finish(plan->ctx, &plan->p, F, W, run_ms);
}

int svo_build_plan_get(const svo_build_plan *plan, uint32_t i, const svo_fraglist **frags, const svo_octree **oct) {
	if (!plan || i >= plan->p.S)
		return SVO_ERR_INVALID_ARGUMENT;
	if (frags)
		*frags = plan->p.fls[i];
	if (oct)
		*oct = plan->p.octs[i];
	return SVO_OK;
}

int svo_build_plan_destroy(svo_build_plan *plan) {
	if (!plan)
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(plan->ctx);
	build_plan_release(plan->ctx, &plan->p);
	delete plan;
	return SVO_OK;
}

int svo_octree_create(svo_ctx *ctx, const uint32_t *words, uint64_t W, uint32_t level, svo_octree **out) {
	if (!ctx || !out || !words || W < 8)
		return SVO_ERR_INVALID_ARGUMENT;
	*out = nullptr;
	if (level < 1 || level > 12)
		return set_error(ctx, SVO_ERR_BAD_LEVEL, "level must be in [1, 12]");
	if (W > (1ull << 30))
		return set_error(ctx, SVO_ERR_OVERFLOW, "octree above 2^30 words");
	DeviceGuard g(ctx);
	svo_octree *o = new (std::nothrow) svo_octree();
	if (!o)
		return SVO_ERR_OUT_OF_MEMORY;
	o->ctx = ctx;
	o->level = level;
	o->W = o->cap = W;
	if (ctx->pool->alloc((void **)&o->words, sizeof(uint32_t) * W) != hipSuccess) {
		svo_octree_destroy(o);
		return set_error(ctx, SVO_ERR_OUT_OF_MEMORY, "octree allocation failed");
	}
	int rc = svo_upload(ctx, o->words, words, sizeof(uint32_t) * W);
	if (rc) {
		svo_octree_destroy(o);
		return rc;
	}
	*out = o;
	return SVO_OK;
}

int svo_octree_info(const svo_octree *o, uint32_t *level, uint64_t *W, const uint32_t **dev) {
	if (!o)
		return SVO_ERR_INVALID_ARGUMENT;
	if (level)
		*level = o->level;
	if (W)
		*W = o->W;
	if (dev)
		*dev = o->words;
	return o->pending ? SVO_ERR_INVALID_ARGUMENT : SVO_OK;
}

int svo_octree_download(svo_ctx *ctx, const svo_octree *o, uint32_t *out, uint64_t cap) {
	if (!ctx || !o || !out || o->pending)
		return SVO_ERR_INVALID_ARGUMENT;
	return svo_download(ctx, out, o->words, sizeof(uint32_t) * std::min(cap, o->W));
}

int svo_octree_destroy(svo_octree *o) {
	if (!o)
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(o->ctx);
	if (o->pending)
		(void)hipStreamSynchronize(o->ctx->stream);
	o->ctx->pool->free(o->words);
	delete o;
	return SVO_OK;
}

namespace {
struct OctreeFileHeader {
	char magic[8];
	uint32_t level;
	uint32_t zero;
	uint64_t W;
};
static_assert(sizeof(OctreeFileHeader) == 24, "octree file header is 24 bytes");
const char kOctreeMagic[8] = {'S', 'V', 'O', 'C', 'T', 'R', 'E', '1'};
} // namespace

int svo_octree_transfer(const svo_octree *oct, svo_ctx *dst) {
	if (!oct || !dst || !oct->ctx)
		return SVO_ERR_INVALID_ARGUMENT;
	svo_ctx *src = oct->ctx;
	if (src->device != dst->device)
		return set_error(dst, SVO_ERR_INVALID_ARGUMENT, "octree is on another device");
	if (src == dst)
		return SVO_OK; // one stream: already ordered
	DeviceGuard g(src);
	hipEvent_t ev = nullptr;
	hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
	if (e == hipSuccess)
		e = hipEventRecord(ev, src->stream); // release: after every kernel enqueued on the builder's stream
	if (e == hipSuccess)
		e = hipStreamWaitEvent(dst->stream, ev, 0); // acquire: later work on dst waits for it
	if (ev)
		(void)hipEventDestroy(ev); // the wait keeps the recorded state
	if (e != hipSuccess)
		return hip_check(dst, e, "octree transfer");
	return SVO_OK;
}

int svo_octree_save(svo_ctx *ctx, const svo_octree *o, const char *path) {
	if (!ctx || !o || !path || o->pending)
		return SVO_ERR_INVALID_ARGUMENT;
	std::vector<uint32_t> words(o->W);
	int rc = svo_octree_download(ctx, o, words.data(), o->W);
	if (rc)
		return rc;
	FILE *f = std::fopen(path, "wb");
	if (!f)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
	OctreeFileHeader h{};
	std::memcpy(h.magic, kOctreeMagic, 8);
	h.level = o->level;
	h.W = o->W;
	bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 && std::fwrite(words.data(), 4, words.size(), f) == words.size();
	ok = (std::fclose(f) == 0) && ok;
	return ok ? SVO_OK : set_error(ctx, SVO_ERR_INVALID_ARGUMENT, std::string("write failed: ") + path);
}

int svo_octree_load(svo_ctx *ctx, const char *path, svo_octree **out) {
	if (!ctx || !path || !out)
		return SVO_ERR_INVALID_ARGUMENT;
	*out = nullptr;
	FILE *f = std::fopen(path, "rb");
	if (!f)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
	OctreeFileHeader h{};
	std::vector<uint32_t> words;
	bool ok = std::fread(&h, sizeof(h), 1, f) == 1 && std::memcmp(h.magic, kOctreeMagic, 8) == 0 && h.W >= 8 &&
	          h.W <= (1ull << 30);
	if (ok) {
		words.resize(h.W);
		ok = std::fread(words.data(), 4, h.W, f) == h.W;
	}
	std::fclose(f);
	if (!ok)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, std::string("not an octree file: ") + path);
	return svo_octree_create(ctx, words.data(), h.W, h.level, out);
}

// ---- OctreeTracer
int svo_march(svo_ctx *ctx, const svo_octree *o, const svo_camera *cam, uint32_t width, uint32_t height, int beam,
              svo_hit *hits_dev, float *beam_dev) {
	if (!ctx || !o || !cam || !hits_dev)
		return SVO_ERR_INVALID_ARGUMENT;
	if (width == 0 || height == 0)
		return SVO_OK;
	if ((uint64_t)width * height > 0xFFFFFFFFull)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "frame too large");
	DeviceGuard g(ctx);
	return march_frame(ctx, o, cam, width, height, beam, hits_dev, beam_dev);
}

int svo_march_noalloc(svo_ctx *ctx, const svo_octree *o, const svo_camera *cam, uint32_t width, uint32_t height,
                      svo_hit *hits_dev) {
	if (!ctx || !o || !cam || !hits_dev)
		return SVO_ERR_INVALID_ARGUMENT;
	if (width == 0 || height == 0)
		return SVO_OK;
	DeviceGuard g(ctx);
	return march_frame_noalloc(ctx, o, cam, width, height, hits_dev);
}

int svo_march_beam(svo_ctx *ctx, const svo_octree *o, const svo_camera *cam, uint32_t width, uint32_t height,
                   float *beam_dev, uint32_t *nfetch_dev) {
	if (!ctx || !o || !cam || !beam_dev)
		return SVO_ERR_INVALID_ARGUMENT;
	if (width == 0 || height == 0)
		return SVO_OK;
	if ((uint64_t)width * height > 0xFFFFFFFFull)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "frame too large");
	DeviceGuard g(ctx);
	return march_beam(ctx, o, cam, width, height, beam_dev, nfetch_dev);
}

int svo_shade(svo_ctx *ctx, const svo_hit *hits_dev, const svo_camera *cam, uint32_t width, uint32_t height,
              int view_type, const float light[3], const svo_envmap *env, float env_rotation, uint32_t *rgba8_dev) {
	if (!ctx || !cam || !light || ((uint64_t)width * height && (!hits_dev || !rgba8_dev)))
		return SVO_ERR_INVALID_ARGUMENT;
	if (view_type < 0 || view_type > 3)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "view type must be 0..3 (diffuse, normal, position, iteration)");
	if ((uint64_t)width * height > 0xFFFFFFFFull)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "frame too large");
	if (env && env->ctx->device != ctx->device)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "environment map is on another device");
	DeviceGuard g(ctx);
	return shade_frame(ctx, hits_dev, cam, width, height, (uint32_t)view_type, light, env, env_rotation, rgba8_dev);
}

int svo_march_rays(svo_ctx *ctx, const svo_octree *o, const float *rays_dev, uint64_t R, int variant,
                   svo_hit *hits_dev) {
	if (!ctx || !o || (R && (!rays_dev || !hits_dev)) || variant < 0 || variant > 2)
		return SVO_ERR_INVALID_ARGUMENT;
	if (R > (uint64_t)0xFFFFFFFFull * 256u)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "too many rays in one call");
	DeviceGuard g(ctx);
	return march_rays(ctx, o, rays_dev, R, variant, hits_dev);
}

int svo_bounce_rays(svo_ctx *ctx, const svo_hit *hits_dev, uint32_t width, uint32_t height, uint32_t spp,
                    uint32_t cur, float *rays_dev) {
	const uint64_t R = (uint64_t)width * height;
	if (!ctx || (R && (!hits_dev || !rays_dev)))
		return SVO_ERR_INVALID_ARGUMENT;
	if (cur < 1 || cur > 31)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "bounce index must be 1..31 (64 Sobol dimensions)");
	if (R > 0xFFFFFFFFull || spp == 0xFFFFFFFFu)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "frame or sample index too large");
	DeviceGuard g(ctx);
	return bounce_rays(ctx, hits_dev, width, height, spp, cur, rays_dev);
}

int svo_path_trace(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t width, uint32_t height,
                   uint32_t bounce, uint32_t spp, const float sky[3], float *color_dev, uint32_t *albedo_dev,
                   uint32_t *normal_dev) {
	if (!ctx || !oct || !cam || !sky || !color_dev || width == 0 || height == 0)
		return SVO_ERR_INVALID_ARGUMENT;
	if (bounce < 1 || bounce > 31)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "bounce must be 1..31 (64 Sobol dimensions)");
	if ((uint64_t)width * height > 0xFFFFFFFFull || spp == 0xFFFFFFFFu)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "frame or sample count too large");
	if (!oct->words || oct->W < 8)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "octree is empty");
	DeviceGuard g(ctx);
	return path_trace(ctx, oct, cam, width, height, bounce, spp, sky, color_dev, albedo_dev, normal_dev);
}

// ---- environment maps (EnvironmentMap::Reset, src/EnvironmentMap.cpp:55-190) ----------------------
namespace {
// weigh_hdr_image (src/EnvironmentMap.cpp:104-128): luminance * sin(theta) per texel, normalised to
// a mean of 1 (the texel's pdf, kept in alpha); generate_alias_table (:130-188): the two-stack
// alias construction, probabilities as 32-bit fractions.  false if the map has no light.
bool envmap_tables(uint32_t w, uint32_t h, const float *rgba, std::vector<float> &texels,
                   std::vector<uint32_t> &alias) {
	const uint64_t n = (uint64_t)w * h;
	std::vector<double> wt(n);
	double sum = 0.0;
	for (uint32_t j = 0; j < h; ++j) {
		const double sin_theta = std::sin(M_PI * ((double)j + 0.5) / (double)h);
		for (uint32_t i = 0; i < w; ++i) {
			const float *c = rgba + 4 * ((uint64_t)j * w + i);
			const float lum = (0.2126f * c[0] + 0.7152f * c[1]) + 0.0722f * c[2];
			wt[(uint64_t)j * w + i] = (double)lum * sin_theta;
			sum += wt[(uint64_t)j * w + i];
		}
	}
	if (!(sum > 0.0))
		return false;
	const double mul = (double)n / sum;
	texels.resize(4 * n);
	for (uint64_t i = 0; i < n; ++i) {
		wt[i] *= mul;
		texels[4 * i + 0] = rgba[4 * i + 0];
		texels[4 * i + 1] = rgba[4 * i + 1];
		texels[4 * i + 2] = rgba[4 * i + 2];
		texels[4 * i + 3] = (float)wt[i];
	}
	alias.assign(2 * n, 0u);
	std::vector<uint32_t> stack(n); // small entries grow from the front, large ones from the back
	uint64_t small = 0, large = n;
	for (uint64_t i = 0; i < n; ++i) {
		if (wt[i] < 1.0)
			stack[small++] = (uint32_t)i;
		else
			stack[--large] = (uint32_t)i;
	}
	while (small > 0 && large < n) {
		const uint32_t s = stack[--small], l = stack[large];
		const double p = wt[s] * 4294967296.0;
		alias[2 * s] = (uint32_t)(p < 0.0 ? 0.0 : p > 4294967295.0 ? 4294967295.0 : p);
		alias[2 * s + 1] = l;
		wt[l] -= 1.0 - wt[s];
		if (wt[l] < 1.0) {
			++large;
			stack[small++] = l;
		}
	}
	while (large < n) {
		const uint32_t l = stack[large++];
		alias[2 * l] = alias[2 * l + 1] = 0xFFFFFFFFu;
	}
	while (small > 0) {
		const uint32_t t = stack[--small];
		alias[2 * t] = alias[2 * t + 1] = 0xFFFFFFFFu;
	}
	return true;
}
} // namespace

int svo_envmap_create(svo_ctx *ctx, uint32_t width, uint32_t height, const float *rgba, svo_envmap **out) {
	if (!ctx || !rgba || !out || width == 0 || height == 0)
		return SVO_ERR_INVALID_ARGUMENT;
	*out = nullptr;
	if ((uint64_t)width * height > 0xFFFFFFFFull / 16u)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "environment map too large");
	std::vector<float> texels;
	std::vector<uint32_t> alias;
	if (!envmap_tables(width, height, rgba, texels, alias))
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "environment map has no light (zero luminance)");
	DeviceGuard g(ctx);
	svo_envmap *e = new (std::nothrow) svo_envmap();
	if (!e)
		return SVO_ERR_OUT_OF_MEMORY;
	e->ctx = ctx;
	e->width = width;
	e->height = height;
	if (ctx->pool->alloc((void **)&e->texels, sizeof(float) * texels.size()) != hipSuccess ||
	    ctx->pool->alloc((void **)&e->alias, sizeof(uint32_t) * alias.size()) != hipSuccess) {
		svo_envmap_destroy(e);
		return set_error(ctx, SVO_ERR_OUT_OF_MEMORY, "environment map allocation failed");
	}
	int rc = svo_upload(ctx, e->texels, texels.data(), sizeof(float) * texels.size());
	if (!rc)
		rc = svo_upload(ctx, e->alias, alias.data(), sizeof(uint32_t) * alias.size());
	if (rc) {
		svo_envmap_destroy(e);
		return rc;
	}
	*out = e;
	return SVO_OK;
}

int svo_envmap_download(svo_ctx *ctx, const svo_envmap *env, float *texels, uint32_t *alias) {
	if (!ctx || !env || (!texels && !alias))
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	const uint64_t n = (uint64_t)env->width * env->height;
	int rc = texels ? svo_download(ctx, texels, env->texels, sizeof(float) * 4 * n) : SVO_OK;
	if (!rc && alias)
		rc = svo_download(ctx, alias, env->alias, sizeof(uint32_t) * 2 * n);
	return rc;
}

int svo_envmap_destroy(svo_envmap *env) {
	if (!env)
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(env->ctx);
	(void)hipStreamSynchronize(env->ctx->stream); // paths in flight may still read it
	env->ctx->pool->free(env->texels);
	env->ctx->pool->free(env->alias);
	delete env;
	return SVO_OK;
}

int svo_path_trace_env(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t width, uint32_t height,
                       uint32_t bounce, uint32_t spp, const svo_envmap *env, float rotation, float multiplier,
                       float *color_dev, uint32_t *albedo_dev, uint32_t *normal_dev) {
	if (!ctx || !oct || !cam || !env || !color_dev || width == 0 || height == 0)
		return SVO_ERR_INVALID_ARGUMENT;
	if (bounce < 1 || bounce > 31)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "bounce must be 1..31 (64 Sobol dimensions)");
	if ((uint64_t)width * height > 0xFFFFFFFFull || spp == 0xFFFFFFFFu)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "frame or sample count too large");
	if (!oct->words || oct->W < 8)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "octree is empty");
	if (env->ctx != ctx)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "environment map belongs to another context");
	DeviceGuard g(ctx);
	return path_trace_env(ctx, oct, cam, width, height, bounce, spp, env, rotation, multiplier, color_dev, albedo_dev,
	                      normal_dev);
}

// ---- batched tracing
namespace {
int check_frame_batch(svo_ctx *ctx, const svo_octree *const *octs, const svo_camera *cams, uint32_t n,
                      uint32_t width, uint32_t height) {
	if (n && (!octs || !cams))
		return SVO_ERR_INVALID_ARGUMENT;
	if (n > kMaxBatch)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "batch larger than 1024 octrees");
	for (uint32_t i = 0; i < n; ++i)
		if (!octs[i] || !octs[i]->words)
			return SVO_ERR_INVALID_ARGUMENT;
	if ((uint64_t)width * height * (n ? n : 1) > 0xFFFFFFFFull)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "more than 2^32-1 rays in one batch");
	return SVO_OK;
}
} // namespace

int svo_march_plan_create(svo_ctx *ctx, const svo_octree *const *octs, const svo_camera *cams, uint32_t n,
                          uint32_t width, uint32_t height, int beam, svo_march_plan **out) {
	if (!ctx || !out)
		return SVO_ERR_INVALID_ARGUMENT;
	*out = nullptr;
	int rc = check_frame_batch(ctx, octs, cams, n, width, height);
	if (rc)
		return rc;
	DeviceGuard g(ctx);
	svo_march_plan *p = new (std::nothrow) svo_march_plan();
	if (!p)
		return SVO_ERR_OUT_OF_MEMORY;
	p->ctx = ctx;
	rc = march_plan_init(ctx, &p->p, octs, cams, n, width, height, beam);
	if (rc) {
		svo_march_plan_destroy(p);
		return rc;
	}
	*out = p;
	return SVO_OK;
}

int svo_march_plan_set_cameras(svo_ctx *ctx, svo_march_plan *plan, const svo_camera *cams) {
	if (!ctx || !plan || plan->ctx != ctx || (plan->p.S && !cams))
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	return march_plan_set_cameras(ctx, &plan->p, cams);
}

int svo_march_plan_use_graph(svo_march_plan *plan, int on) {
	if (!plan)
		return SVO_ERR_INVALID_ARGUMENT;
	plan->p.use_graph = on != 0;
	return SVO_OK;
}

int svo_march_plan_set_shape(svo_march_plan *plan, int shape) {
	if (!plan || shape < -1 || shape > 3)
		return SVO_ERR_INVALID_ARGUMENT;
	if (plan->p.exec) {
		(void)hipGraphExecDestroy(plan->p.exec); // recaptured with the new shape
		plan->p.exec = nullptr;
	}
	plan->p.shape = shape;
	return SVO_OK;
}

int svo_march_plan_run(svo_ctx *ctx, svo_march_plan *plan, svo_hit *hits_dev) {
	if (!ctx || !plan || plan->ctx != ctx || (plan->p.S && !hits_dev))
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(ctx);
	return march_plan_run(ctx, &plan->p, hits_dev);
}

int svo_march_plan_destroy(svo_march_plan *plan) {
	if (!plan)
		return SVO_ERR_INVALID_ARGUMENT;
	DeviceGuard g(plan->ctx);
	march_plan_release(plan->ctx, &plan->p);
	delete plan;
	return SVO_OK;
}

int svo_march_batch(svo_ctx *ctx, const svo_octree *const *octs, const svo_camera *cams, uint32_t n,
                    uint32_t width, uint32_t height, int beam, svo_hit *hits_dev) {
	if (!ctx || (n && !hits_dev))
		return SVO_ERR_INVALID_ARGUMENT;
	svo_march_plan *plan = nullptr;
	int rc = svo_march_plan_create(ctx, octs, cams, n, width, height, beam, &plan);
	if (rc)
		return rc;
	rc = svo_march_plan_run(ctx, plan, hits_dev);
	svo_march_plan_destroy(plan); // stream-ordered pool: later work only
	return rc;
}

int svo_march_rays_batch(svo_ctx *ctx, const svo_octree *const *octs, uint32_t n, const uint64_t *ray_begin,
                         const float *rays_dev, int variant, svo_hit *hits_dev) {
	if (!ctx || !ray_begin || (n && !octs) || variant < 0 || variant > 2)
		return SVO_ERR_INVALID_ARGUMENT;
	if (n > kMaxBatch)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "batch larger than 1024 octrees");
	if (ray_begin[0] != 0)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "ray_begin[0] must be 0");
	for (uint32_t i = 0; i < n; ++i) {
		if (!octs[i] || !octs[i]->words || ray_begin[i + 1] < ray_begin[i])
			return SVO_ERR_INVALID_ARGUMENT;
	}
	const uint64_t R = ray_begin[n];
	if (R && (!rays_dev || !hits_dev))
		return SVO_ERR_INVALID_ARGUMENT;
	if (R > 0xFFFFFFFFull * 64u)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "too many rays in one call");
	DeviceGuard g(ctx);
	return march_rays_batch(ctx, octs, n, ray_begin, rays_dev, variant, hits_dev);
}

int svo_march_mode(svo_ctx *ctx, int mode) {
	if (!ctx || mode < -1 || mode > 2)
		return SVO_ERR_INVALID_ARGUMENT;
	ctx->march_mode = mode;
	return SVO_OK;
}

int svo_autotune(svo_ctx *ctx, int on) {
	if (!ctx)
		return SVO_ERR_INVALID_ARGUMENT;
	ctx->autotune = on != 0;
	return SVO_OK;
}

int svo_autotune_report(svo_ctx *ctx, char *buf, uint64_t cap, uint64_t *len) {
	if (!ctx)
		return SVO_ERR_INVALID_ARGUMENT;
	const std::string &s = ctx->tune_log;
	if (len)
		*len = s.size();
	if (buf && cap) {
		const uint64_t n = std::min<uint64_t>(cap - 1, s.size());
		std::memcpy(buf, s.data(), n);
		buf[n] = 0;
	}
	return SVO_OK;
}

// Camera::fetch_uniform_data (src/Camera.cpp:77-95): yaw about +y, then pitch about -x;
// look = R * (0,0,1), side = normalize(R * (1,0,0)) * tan(fov/2) * aspect,
// up = normalize(cross(look, side)) * tan(fov/2).
int svo_camera_from_yaw_pitch(const float position[3], float yaw, float pitch, float fov, float aspect,
                              svo_camera *out) {
	if (!position || !out)
		return SVO_ERR_INVALID_ARGUMENT;
	const float cy = std::cos(yaw), sy = std::sin(yaw), cp = std::cos(pitch), sp = std::sin(pitch);
	// rotate(yaw, +y) * rotate(pitch, -x) applied to the basis vectors
	float look[3] = {sy * cp, sp, cy * cp};
	float side[3] = {cy, 0.0f, -sy};
	auto nrm = [](float *v) {
		float s = 1.0f / std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
		v[0] *= s, v[1] *= s, v[2] *= s;
	};
	nrm(look);
	nrm(side);
	const float tg = std::tan(fov * 0.5f);
	for (int c = 0; c < 3; ++c)
		side[c] = side[c] * tg * aspect;
	float up[3] = {look[1] * side[2] - side[1] * look[2], look[2] * side[0] - side[2] * look[0],
	               look[0] * side[1] - side[0] * look[1]};
	nrm(up);
	for (int c = 0; c < 3; ++c)
		up[c] *= tg;
	for (int c = 0; c < 3; ++c) {
		out->pos[c] = position[c];
		out->look[c] = look[c];
		out->side[c] = side[c];
		out->up[c] = up[c];
	}
	out->pos[3] = out->look[3] = out->side[3] = out->up[3] = 1.0f;
	return SVO_OK;
}

} // extern "C"
