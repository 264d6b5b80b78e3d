// Serving loop: build and ray-query requests for many scenes arrive at any time and are batched by
// two worker threads with their own contexts (HIP streams) on the same device, so ray queries
// never wait behind a build:
//   ray worker    continuous batching over kRaySlots streams: whenever a stream is free and queries
//                 wait, it takes every waiting query and marches them all with one launch per march
//                 variant (per-request ray and hit buffers, no copies), without waiting for the
//                 batches already on the GPU; completions are polled by event, so a query that
//                 arrives while batches run starts on the next free stream, and batches overlap on
//                 the device.  A query pins its octree at submission (it cannot be evicted while
//                 queued or marching); a query for an octree that is not resident is refused at
//                 submission;
//   build worker  drains pending builds (up to kMaxBuildBatch), voxelizes them with one count and
//                 one fill launch, builds them with one launch per pass (in groups whose transient
//                 scratch fits the budget), then admits each octree
//                 into the resident set by its exact size: least-recently-used unpinned octrees are
//                 evicted until it fits the byte budget and the octree-count limit, else the build
//                 is rejected (SVO_ERR_OUT_OF_MEMORY).
// Latency is measured on the host from submission to the completion of the request's batch.
// The reference's nearest analogue is its loader thread, which builds one octree on a side queue
// and hands it to the renderer (src/LoaderThread.cpp:15-116, src/Octree.cpp:22-35).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "svo_batch.hpp"
#include "svo_runtime.hpp"

namespace {

using Clock = std::chrono::steady_clock;
constexpr uint32_t kMaxBuildBatch = 64;
constexpr int kRaySlots = 2; // ray batches in flight at once, one stream each

struct Request {
	uint64_t id = 0;
	bool build = false;
	const svo_scene *scene = nullptr;
	uint32_t level = 0;
	uint64_t octree = 0;
	const float *rays = nullptr;
	uint64_t R = 0;
	int variant = 0;
	svo_hit *hits = nullptr;
	Clock::time_point t_submit, t_done;
	int status = SVO_OK;
	bool done = false;
};

struct Resident {
	svo_octree *oct = nullptr;
	uint64_t bytes = 0;
	uint64_t last_use = 0;
	uint32_t pins = 0;
};

double ms_between(Clock::time_point a, Clock::time_point b) {
	return std::chrono::duration<double, std::milli>(b - a).count();
}

double percentile(std::vector<double> v, double p) {
	if (v.empty())
		return 0.0;
	const size_t k = std::min(v.size() - 1, (size_t)(p * (double)(v.size() - 1) + 0.5));
	std::nth_element(v.begin(), v.begin() + (long)k, v.end());
	return v[k];
}

} // namespace

struct svo_server {
	int device = 0;
	svo_ctx *cb = nullptr;            // build worker's context
	svo_ctx *cr[kRaySlots] = {};      // ray worker's contexts, one per slot
	hipEvent_t ev[kRaySlots] = {};    // completion of each slot's batch
	uint64_t budget = 0;
	uint32_t max_octrees = 0;
	std::mutex mu;
	std::condition_variable cv_build, cv_rays, cv_done;
	std::deque<Request *> qb, qr;
	std::unordered_map<uint64_t, Request *> reqs;
	std::unordered_map<uint64_t, Resident> res;
	uint64_t resident_bytes = 0, use_clock = 0, next_id = 1;
	bool stop = false;
	std::thread tb, tr;
	// statistics
	std::vector<double> lat_build, lat_rays;
	uint64_t builds_rejected = 0, rays_failed = 0, rays_traced = 0, evictions = 0, batches = 0;
	double busy_build_ms = 0.0, busy_rays_ms = 0.0;
	Clock::time_point t_stats = Clock::now();

	void finish(Request *r, int status, Clock::time_point t) { // mu held
		r->status = status;
		r->t_done = t;
		r->done = true;
		const double ms = ms_between(r->t_submit, t);
		if (r->build) {
			if (status == SVO_OK)
				lat_build.push_back(ms);
			else
				++builds_rejected;
		} else {
			if (status == SVO_OK) {
				lat_rays.push_back(ms);
				rays_traced += r->R;
			} else {
				++rays_failed;
			}
		}
	}

	// Evicts least-recently-used unpinned octrees until `need` more bytes (and, with slot, one more
	// octree) fit.  mu held; only the build worker destroys octrees (they belong to its context).
	bool make_room(uint64_t need, bool slot) {
		while (resident_bytes + need > budget || (slot && res.size() + 1 > max_octrees)) {
			auto victim = res.end();
			for (auto it = res.begin(); it != res.end(); ++it)
				if (it->second.pins == 0 && (victim == res.end() || it->second.last_use < victim->second.last_use))
					victim = it;
			if (victim == res.end())
				return false;
			resident_bytes -= victim->second.bytes;
			svo_octree_destroy(victim->second.oct);
			res.erase(victim);
			++evictions;
		}
		return true;
	}

	void ray_loop();
	void build_loop();
	void build_batch_requests(std::vector<Request *> &batch);
};

void svo_server::ray_loop() {
	(void)hipSetDevice(device);
	struct Flight {
		std::vector<Request *> reqs;
		bool busy = false;
	};
	Flight slot[kRaySlots];
	int inflight = 0;
	Clock::time_point busy_since{};
	std::unique_lock<std::mutex> lk(mu);
	for (;;) {
		// completed batches: finish their queries and unpin their octrees
		for (int k = 0; k < kRaySlots; ++k) {
			if (!slot[k].busy)
				continue;
			const hipError_t q = hipEventQuery(ev[k]);
			if (q == hipErrorNotReady)
				continue;
			const Clock::time_point t1 = Clock::now();
			for (Request *r : slot[k].reqs) {
				auto it = res.find(r->octree);
				if (it != res.end())
					it->second.pins--;
				finish(r, q == hipSuccess ? r->status : SVO_ERR_HIP, t1);
			}
			slot[k].reqs.clear();
			slot[k].busy = false;
			if (--inflight == 0)
				busy_rays_ms += ms_between(busy_since, t1);
			cv_done.notify_all();
		}
		int free_slot = -1;
		for (int k = 0; k < kRaySlots && free_slot < 0; ++k)
			if (!slot[k].busy)
				free_slot = k;
		if (free_slot >= 0 && !qr.empty()) {
			// every waiting query goes into this batch; their octrees were pinned at submission
			std::vector<Request *> run(qr.begin(), qr.end());
			qr.clear();
			std::vector<const svo_octree *> octs;
			for (Request *r : run)
				octs.push_back(res.at(r->octree).oct);
			++batches;
			if (inflight++ == 0)
				busy_since = Clock::now();
			slot[free_slot].busy = true;
			lk.unlock();
			svo_ctx *c = cr[free_slot];
			int rc = SVO_OK;
			for (int variant = 0; variant < 3 && !rc; ++variant) {
				std::vector<const svo_octree *> o;
				std::vector<const float *> rays;
				std::vector<svo_hit *> hits;
				std::vector<uint64_t> count;
				for (size_t i = 0; i < run.size(); ++i)
					if (run[i]->variant == variant) {
						o.push_back(octs[i]);
						rays.push_back(run[i]->rays);
						hits.push_back(run[i]->hits);
						count.push_back(run[i]->R);
					}
				for (size_t b = 0; b < o.size() && !rc; b += svo::kMaxBatch) {
					const uint32_t n = (uint32_t)std::min<size_t>(svo::kMaxBatch, o.size() - b);
					rc = svo::march_rays_segments(c, o.data() + b, rays.data() + b, hits.data() + b, count.data() + b,
					                              n, variant);
				}
			}
			if (hipEventRecord(ev[free_slot], c->stream) != hipSuccess) {
				rc = rc ? rc : SVO_ERR_HIP;
				(void)hipStreamSynchronize(c->stream); // no event: complete this batch here
				(void)hipEventRecord(ev[free_slot], nullptr);
			}
			for (Request *r : run)
				r->status = rc;
			lk.lock();
			slot[free_slot].reqs = std::move(run);
			continue;
		}
		if (stop && qr.empty() && inflight == 0)
			break; // stop requested and drained
		if (inflight == 0)
			cv_rays.wait(lk, [&] { return stop || !qr.empty(); });
		else // poll the batches on the GPU; a new query wakes the loop at once
			cv_rays.wait_for(lk, std::chrono::microseconds(40));
	}
}

// Voxelize + build of a batch of build requests.  The batch is voxelized once; it is then built in
// consecutive groups whose transient build scratch (20 B per fragment plus bitmaps and ranks) fits
// the byte budget after evicting idle octrees: a group that does not fit is halved, and a single
// scene that does not fit even alone is rejected (SVO_ERR_OUT_OF_MEMORY).  A voxelize failure of
// the whole batch retries each request alone, so that one bad scene cannot fail the others.
void svo_server::build_batch_requests(std::vector<Request *> &batch) {
	const uint32_t n = (uint32_t)batch.size();
	std::vector<const svo_scene *> scenes(n);
	std::vector<uint32_t> levels(n);
	for (uint32_t i = 0; i < n; ++i) {
		scenes[i] = batch[i]->scene;
		levels[i] = batch[i]->level;
	}
	std::vector<svo_fraglist *> fls(n, nullptr);
	int rc = svo_voxelize_batch(cb, scenes.data(), levels.data(), n, fls.data());
	if (rc) {
		for (uint32_t i = 0; i < n; ++i)
			if (fls[i])
				svo_fraglist_destroy(fls[i]);
		if (n > 1) {
			for (uint32_t i = 0; i < n; ++i) {
				std::vector<Request *> one{batch[i]};
				build_batch_requests(one);
			}
			return;
		}
		std::unique_lock<std::mutex> lk(mu);
		finish(batch[0], rc, Clock::now());
		return;
	}
	std::vector<uint64_t> scratch(n);
	for (uint32_t i = 0; i < n; ++i) {
		uint64_t F = 0;
		svo_fraglist_info(fls[i], nullptr, &F, nullptr);
		scratch[i] = 20 * F + 8 * svo::build_bitmap_words(levels[i]);
	}
	uint32_t single_until = 0; // after a failed group, its scenes are built one at a time
	for (uint32_t i = 0; i < n;) {
		// the largest group [i, j) whose scratch fits
		uint32_t j = i < single_until ? i + 1 : n;
		bool fits = false;
		for (;;) {
			uint64_t need = 0;
			for (uint32_t k = i; k < j; ++k)
				need += scratch[k];
			{
				std::unique_lock<std::mutex> lk(mu);
				fits = make_room(need, false);
			}
			if (fits || j - i == 1)
				break;
			j = i + (j - i) / 2;
		}
		std::vector<svo_octree *> octs(j - i, nullptr);
		int grc = fits ? svo_build_batch(cb, (const svo_fraglist *const *)fls.data() + i, j - i, octs.data())
		               : SVO_ERR_OUT_OF_MEMORY;
		if (!grc && hipStreamSynchronize(cb->stream) != hipSuccess)
			grc = SVO_ERR_HIP;
		if (grc && j - i > 1) { // retry the group's scenes alone: one bad scene must not fail the others
			for (svo_octree *o : octs)
				if (o)
					svo_octree_destroy(o);
			single_until = j;
			continue;
		}
		const Clock::time_point t1 = Clock::now();
		std::unique_lock<std::mutex> lk(mu);
		for (uint32_t k = i; k < j; ++k) {
			Request *r = batch[k];
			svo_octree *o = octs[k - i];
			svo_fraglist_destroy(fls[k]);
			fls[k] = nullptr;
			if (grc) {
				if (o)
					svo_octree_destroy(o);
				finish(r, grc, t1);
				continue;
			}
			svo_build_finish(cb, o, nullptr);
			uint64_t W = 0;
			svo_octree_info(o, nullptr, &W, nullptr);
			const uint64_t bytes = 4 * W;
			if (!make_room(bytes, true)) {
				svo_octree_destroy(o);
				finish(r, SVO_ERR_OUT_OF_MEMORY, t1); // rejected by admission
				continue;
			}
			Resident e;
			e.oct = o;
			e.bytes = bytes;
			e.last_use = ++use_clock;
			res[r->id] = e;
			resident_bytes += bytes;
			finish(r, SVO_OK, t1);
		}
		i = j;
	}
}

void svo_server::build_loop() {
	(void)hipSetDevice(device);
	std::unique_lock<std::mutex> lk(mu);
	for (;;) {
		cv_build.wait(lk, [&] { return stop || !qb.empty(); });
		if (qb.empty())
			break;
		std::vector<Request *> batch;
		while (!qb.empty() && batch.size() < kMaxBuildBatch) {
			batch.push_back(qb.front());
			qb.pop_front();
		}
		++batches;
		lk.unlock();
		const Clock::time_point t0 = Clock::now();
		build_batch_requests(batch);
		const Clock::time_point t1 = Clock::now();
		lk.lock();
		busy_build_ms += ms_between(t0, t1);
		cv_done.notify_all();
	}
}

extern "C" {

int svo_server_create(int device, uint64_t budget_bytes, uint32_t max_octrees, svo_server **out) {
	if (!out || max_octrees == 0)
		return SVO_ERR_INVALID_ARGUMENT;
	*out = nullptr;
	svo_server *s = new (std::nothrow) svo_server();
	if (!s)
		return SVO_ERR_OUT_OF_MEMORY;
	s->device = device;
	s->budget = budget_bytes;
	s->max_octrees = max_octrees;
	int rc = svo_ctx_create(device, &s->cb);
	for (int k = 0; k < kRaySlots && !rc; ++k) {
		rc = svo_ctx_create(device, &s->cr[k]);
		if (!rc && hipEventCreateWithFlags(&s->ev[k], hipEventDisableTiming) != hipSuccess)
			rc = SVO_ERR_HIP;
	}
	if (rc) {
		for (int k = 0; k < kRaySlots; ++k) {
			if (s->ev[k])
				(void)hipEventDestroy(s->ev[k]);
			if (s->cr[k])
				svo_ctx_destroy(s->cr[k]);
		}
		if (s->cb)
			svo_ctx_destroy(s->cb);
		delete s;
		return rc;
	}
	s->tb = std::thread([s] { s->build_loop(); });
	s->tr = std::thread([s] { s->ray_loop(); });
	*out = s;
	return SVO_OK;
}

int svo_server_destroy(svo_server *s) {
	if (!s)
		return SVO_ERR_INVALID_ARGUMENT;
	{
		std::lock_guard<std::mutex> lk(s->mu);
		s->stop = true;
	}
	s->cv_build.notify_all();
	s->cv_rays.notify_all();
	s->tb.join();
	s->tr.join();
	for (auto &kv : s->res)
		svo_octree_destroy(kv.second.oct);
	for (auto &kv : s->reqs)
		delete kv.second;
	for (int k = 0; k < kRaySlots; ++k) {
		(void)hipEventDestroy(s->ev[k]);
		svo_ctx_destroy(s->cr[k]);
	}
	svo_ctx_destroy(s->cb);
	delete s;
	return SVO_OK;
}

int svo_server_submit_build(svo_server *s, const svo_scene *scene, uint32_t level, uint64_t *req_id) {
	if (!s || !scene || !req_id)
		return SVO_ERR_INVALID_ARGUMENT;
	if (level < 1 || level > 12)
		return SVO_ERR_BAD_LEVEL;
	Request *r = new (std::nothrow) Request();
	if (!r)
		return SVO_ERR_OUT_OF_MEMORY;
	r->build = true;
	r->scene = scene;
	r->level = level;
	r->t_submit = Clock::now();
	{
		std::lock_guard<std::mutex> lk(s->mu);
		if (s->stop) {
			delete r;
			return SVO_ERR_INVALID_ARGUMENT;
		}
		r->id = s->next_id++;
		s->reqs[r->id] = r;
		s->qb.push_back(r);
		*req_id = r->id;
	}
	s->cv_build.notify_one();
	return SVO_OK;
}

int svo_server_submit_rays(svo_server *s, uint64_t octree_id, const float *rays_dev, uint64_t R, int variant,
                           svo_hit *hits_dev, uint64_t *req_id) {
	if (!s || !req_id || (R && (!rays_dev || !hits_dev)) || variant < 0 || variant > 2)
		return SVO_ERR_INVALID_ARGUMENT;
	if (R > 0xFFFFFFFFull)
		return SVO_ERR_INVALID_ARGUMENT;
	Request *r = new (std::nothrow) Request();
	if (!r)
		return SVO_ERR_OUT_OF_MEMORY;
	r->octree = octree_id;
	r->rays = rays_dev;
	r->R = R;
	r->variant = variant;
	r->hits = hits_dev;
	r->t_submit = Clock::now();
	{
		std::lock_guard<std::mutex> lk(s->mu);
		auto it = s->res.find(octree_id);
		if (s->stop || it == s->res.end()) { // not resident (never built, rejected or evicted): no request
			delete r;
			return SVO_ERR_INVALID_ARGUMENT;
		}
		it->second.pins++; // held until the query completes: the octree cannot be evicted meanwhile
		it->second.last_use = ++s->use_clock;
		r->id = s->next_id++;
		s->reqs[r->id] = r;
		s->qr.push_back(r);
		*req_id = r->id;
	}
	s->cv_rays.notify_one();
	return SVO_OK;
}

int svo_server_wait(svo_server *s, uint64_t req_id, int *status, double *latency_ms) {
	if (!s)
		return SVO_ERR_INVALID_ARGUMENT;
	std::unique_lock<std::mutex> lk(s->mu);
	auto it = s->reqs.find(req_id);
	if (it == s->reqs.end())
		return SVO_ERR_INVALID_ARGUMENT;
	Request *r = it->second;
	s->cv_done.wait(lk, [&] { return r->done; });
	if (status)
		*status = r->status;
	if (latency_ms)
		*latency_ms = ms_between(r->t_submit, r->t_done);
	s->reqs.erase(it);
	delete r;
	return SVO_OK;
}

int svo_server_resident(svo_server *s, uint64_t octree_id, uint64_t *W, const svo_octree **oct) {
	if (!s)
		return SVO_ERR_INVALID_ARGUMENT;
	std::lock_guard<std::mutex> lk(s->mu);
	auto it = s->res.find(octree_id);
	if (it == s->res.end())
		return SVO_ERR_INVALID_ARGUMENT;
	if (W)
		*W = it->second.bytes / 4;
	if (oct)
		*oct = it->second.oct;
	return SVO_OK;
}

int svo_server_get_stats(svo_server *s, svo_server_stats *out) {
	if (!s || !out)
		return SVO_ERR_INVALID_ARGUMENT;
	std::lock_guard<std::mutex> lk(s->mu);
	out->builds_done = s->lat_build.size();
	out->builds_rejected = s->builds_rejected;
	out->ray_queries_done = s->lat_rays.size();
	out->ray_queries_failed = s->rays_failed;
	out->rays_traced = s->rays_traced;
	out->evictions = s->evictions;
	out->batches = s->batches;
	out->resident_octrees = s->res.size();
	out->resident_bytes = s->resident_bytes;
	out->budget_bytes = s->budget;
	out->build_p50_ms = percentile(s->lat_build, 0.50);
	out->build_p99_ms = percentile(s->lat_build, 0.99);
	out->rays_p50_ms = percentile(s->lat_rays, 0.50);
	out->rays_p99_ms = percentile(s->lat_rays, 0.99);
	out->busy_build_ms = s->busy_build_ms;
	out->busy_rays_ms = s->busy_rays_ms;
	out->elapsed_ms = ms_between(s->t_stats, Clock::now());
	return SVO_OK;
}

int svo_server_reset_stats(svo_server *s) {
	if (!s)
		return SVO_ERR_INVALID_ARGUMENT;
	std::lock_guard<std::mutex> lk(s->mu);
	s->lat_build.clear();
	s->lat_rays.clear();
	s->builds_rejected = s->rays_failed = s->rays_traced = s->evictions = s->batches = 0;
	s->busy_build_ms = s->busy_rays_ms = 0.0;
	s->t_stats = Clock::now();
	return SVO_OK;
}

} // extern "C"
