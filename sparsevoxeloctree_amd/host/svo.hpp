// C++ class surface of the SVO hot path over the C ABI (include/svo.h).
//
// Same class names, factory style (static Create returning std::shared_ptr, nullptr on failure)
// and method roles as the reference's hot-path classes, re-based on a Context (one HIP stream)
// instead of Vulkan command pools and command buffers; "Cmd*" methods enqueue on the stream:
//
//   Scene          src/Scene.hpp        Create (+ normalisation, src/Scene.cpp:90-99)
//   Voxelizer      src/Voxelizer.hpp:43-52
//   OctreeBuilder  src/OctreeBuilder.hpp:36-48
//   Octree         src/Octree.hpp:18-33
//   Camera         src/Camera.hpp:14-38 (yaw / pitch / fov / aspect -> uniform block)
//   OctreeTracer   src/OctreeTracer.hpp:15-70
//   EnvironmentMap src/EnvironmentMap.hpp:20-33 (tables from linear RGBA texels; no EXR reader)
//   PathTracer     src/PathTracer.hpp (Reset / CmdRender, m_bounce; light from Lighting)
//
// Header-only; link with libsvo.so.
#pragma once

#include <cmath>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/svo.h"

namespace svo {

class Context {
public:
	svo_ctx *h = nullptr;
	std::string error;

	static std::shared_ptr<Context> Create(int device = 0) {
		auto c = std::shared_ptr<Context>(new Context());
		if (svo_ctx_create(device, &c->h) != SVO_OK)
			return nullptr;
		return c;
	}
	~Context() {
		if (h)
			svo_ctx_destroy(h);
	}
	bool Check(int rc) {
		if (rc != SVO_OK)
			error = svo_last_error(h);
		return rc == SVO_OK;
	}
	void Synchronize() { Check(svo_ctx_synchronize(h)); }
	svo_timing Timing() {
		svo_timing t{};
		Check(svo_last_timing(h, &t));
		return t;
	}

private:
	Context() = default;
};

class Scene {
public:
	static std::shared_ptr<Scene> Create(const std::shared_ptr<Context> &ctx, const float *tri_xyz,
	                                     const uint32_t *tri_rgb, uint64_t T, bool normalize = true) {
		auto s = std::shared_ptr<Scene>(new Scene(ctx));
		if (!ctx->Check(svo_scene_create(ctx->h, tri_xyz, tri_rgb, T, normalize ? 1 : 0, &s->m_h)))
			return nullptr;
		s->m_T = T;
		return s;
	}
	~Scene() {
		if (m_h)
			svo_scene_destroy(m_h);
	}
	uint64_t GetTriangleCount() const { return m_T; }
	const svo_scene *Handle() const { return m_h; }
	const std::shared_ptr<Context> &GetContextPtr() const { return m_ctx; }

private:
	explicit Scene(std::shared_ptr<Context> ctx) : m_ctx(std::move(ctx)) {}
	std::shared_ptr<Context> m_ctx;
	svo_scene *m_h = nullptr;
	uint64_t m_T = 0;
};

// Voxelizer::Create runs the count-only pass and allocates the fragment list
// (src/Voxelizer.cpp:5-26,134-165); CmdVoxelize enqueues the fill pass (:167-179).
class Voxelizer {
public:
	static std::shared_ptr<Voxelizer> Create(const std::shared_ptr<Scene> &scene, uint32_t octree_level) {
		auto v = std::shared_ptr<Voxelizer>(new Voxelizer(scene->GetContextPtr()));
		v->m_scene = scene;
		v->m_level = octree_level;
		if (!v->m_ctx->Check(svo_voxelize_count(v->m_ctx->h, scene->Handle(), octree_level, &v->m_h)))
			return nullptr;
		return v;
	}
	// A fragment list produced elsewhere (e.g. read back from the reference), host memory.
	static std::shared_ptr<Voxelizer> CreateFromFragments(const std::shared_ptr<Context> &ctx, const uint32_t *frags,
	                                                      uint64_t F, uint32_t octree_level) {
		auto v = std::shared_ptr<Voxelizer>(new Voxelizer(ctx));
		v->m_level = octree_level;
		if (!ctx->Check(svo_fraglist_create(ctx->h, frags, F, octree_level, &v->m_h)))
			return nullptr;
		return v;
	}
	~Voxelizer() {
		if (m_h)
			svo_fraglist_destroy(m_h);
	}
	bool CmdVoxelize() const { return m_ctx->Check(svo_voxelize_fill(m_ctx->h, m_h)); }
	const std::shared_ptr<Scene> &GetScenePtr() const { return m_scene; }
	uint32_t GetLevel() const { return m_level; }
	uint32_t GetVoxelResolution() const { return 1u << m_level; }
	uint32_t GetVoxelFragmentCount() const {
		uint64_t F = 0;
		svo_fraglist_info(m_h, nullptr, &F, nullptr);
		return (uint32_t)F;
	}
	// Device pointer to the fragment list (uvec2 per fragment).
	const uint32_t *GetVoxelFragmentList() const {
		const uint32_t *p = nullptr;
		svo_fraglist_info(m_h, nullptr, nullptr, &p);
		return p;
	}
	std::vector<uint32_t> DownloadFragmentList() const {
		std::vector<uint32_t> out(2 * (size_t)GetVoxelFragmentCount());
		if (!out.empty())
			m_ctx->Check(svo_fraglist_download(m_ctx->h, m_h, out.data(), out.size() / 2));
		return out;
	}
	const svo_fraglist *Handle() const { return m_h; }
	const std::shared_ptr<Context> &GetContextPtr() const { return m_ctx; }

private:
	explicit Voxelizer(std::shared_ptr<Context> ctx) : m_ctx(std::move(ctx)) {}
	std::shared_ptr<Context> m_ctx;
	std::shared_ptr<Scene> m_scene;
	svo_fraglist *m_h = nullptr;
	uint32_t m_level = 0;
};

// OctreeBuilder::CmdBuild enqueues every level pass (src/OctreeBuilder.cpp:142-210);
// GetOctreeRange waits and returns the used bytes, (counter + 1) * 8 words (:212-214).
class OctreeBuilder {
public:
	static std::shared_ptr<OctreeBuilder> Create(const std::shared_ptr<Voxelizer> &voxelizer) {
		auto b = std::shared_ptr<OctreeBuilder>(new OctreeBuilder());
		b->m_voxelizer = voxelizer;
		return b;
	}
	~OctreeBuilder() {
		if (m_h)
			svo_octree_destroy(m_h);
	}
	const std::shared_ptr<Voxelizer> &GetVoxelizerPtr() const { return m_voxelizer; }
	uint32_t GetLevel() const { return m_voxelizer->GetLevel(); }
	bool CmdBuild() {
		auto &ctx = m_voxelizer->GetContextPtr();
		if (m_h) {
			svo_octree_destroy(m_h);
			m_h = nullptr;
		}
		m_W = 0;
		return ctx->Check(svo_build_enqueue(ctx->h, m_voxelizer->Handle(), &m_h));
	}
	uint64_t GetOctreeRange() {
		if (!m_W && m_h)
			m_voxelizer->GetContextPtr()->Check(svo_build_finish(m_voxelizer->GetContextPtr()->h, m_h, &m_W));
		return m_W * sizeof(uint32_t);
	}
	const svo_octree *GetOctree() const { return m_h; }
	// The loader-to-main hand-off (src/OctreeBuilder.cpp:215-220, src/LoaderThread.cpp:77-105): work
	// enqueued on dst afterwards runs after this build; the host does not wait.
	bool CmdTransferOctreeOwnership(const std::shared_ptr<Context> &dst) const {
		return dst->Check(svo_octree_transfer(m_h, dst->h));
	}

private:
	OctreeBuilder() = default;
	std::shared_ptr<Voxelizer> m_voxelizer;
	svo_octree *m_h = nullptr;
	uint64_t m_W = 0;
};

// Octree::Update takes the builder's buffer and used range (src/Octree.cpp:22-35).
class Octree {
public:
	static std::shared_ptr<Octree> Create(const std::shared_ptr<Context> &ctx) {
		auto o = std::shared_ptr<Octree>(new Octree());
		o->m_ctx = ctx;
		return o;
	}
	void Update(const std::shared_ptr<OctreeBuilder> &builder) {
		m_range = builder->GetOctreeRange();
		m_builder = builder;
		m_h = builder->GetOctree();
		m_level = builder->GetLevel();
	}
	bool Empty() const { return m_h == nullptr; }
	const uint32_t *GetBuffer() const {
		const uint32_t *p = nullptr;
		svo_octree_info(m_h, nullptr, nullptr, &p);
		return p;
	}
	uint32_t GetLevel() const { return m_level; }
	uint64_t GetRange() const { return m_range; }
	const svo_octree *Handle() const { return m_h; }
	const std::shared_ptr<Context> &GetContextPtr() const { return m_ctx; }
	// Octree::CmdTransferOwnership (src/Octree.cpp:37-43): orders dst's later work after the build.
	bool CmdTransferOwnership(const std::shared_ptr<Context> &dst) const {
		return dst->Check(svo_octree_transfer(m_h, dst->h));
	}
	std::vector<uint32_t> Download() const {
		std::vector<uint32_t> out(m_range / sizeof(uint32_t));
		m_ctx->Check(svo_octree_download(m_ctx->h, m_h, out.data(), out.size()));
		return out;
	}

private:
	Octree() = default;
	std::shared_ptr<Context> m_ctx;
	std::shared_ptr<OctreeBuilder> m_builder; // keeps the node buffer alive
	const svo_octree *m_h = nullptr;
	uint32_t m_level = 0;
	uint64_t m_range = 0;
};

// Camera: m_position in octree space ([1,2]^3), yaw about +y, pitch about -x (src/Camera.cpp:77-95).
class Camera {
public:
	float m_position[3] = {1.5f, 1.5f, 1.5f};
	float m_yaw = 0.0f, m_pitch = 0.0f, m_fov = 3.14159265358979323846f / 3.0f, m_aspect_ratio = 1280.0f / 720.0f;

	static std::shared_ptr<Camera> Create() { return std::make_shared<Camera>(); }
	svo_camera fetch_uniform_data() const {
		svo_camera c{};
		svo_camera_from_yaw_pitch(m_position, m_yaw, m_pitch, m_fov, m_aspect_ratio, &c);
		return c;
	}
};

class EnvironmentMap;

// OctreeTracer: CmdBeamRenderPass (beam pre-pass) + CmdDrawPipeline (primary march, then the
// shading of shader/octree_tracer.frag:37-47 by m_view_type into an RGBA8 image), into device
// buffers (src/OctreeTracer.cpp:183-246).  The hit records stay available beside the image.
class OctreeTracer {
public:
	enum class ViewTypes { kDiffuse = 0, kNormal, kPosition, kIteration } m_view_type = ViewTypes::kDiffuse;
	bool m_beam_enable{true};
	float m_light[3] = {5.0f, 5.0f, 5.0f};     // constant light (kDefaultConstantColor, src/Config.hpp:28)
	std::shared_ptr<EnvironmentMap> m_env_map; // when set, misses show its radiance (Light(), :18-20)

	static std::shared_ptr<OctreeTracer> Create(const std::shared_ptr<Octree> &octree,
	                                            const std::shared_ptr<Camera> &camera) {
		auto t = std::shared_ptr<OctreeTracer>(new OctreeTracer());
		t->m_octree = octree;
		t->m_camera = camera;
		t->Resize(1280, 720);
		return t;
	}
	~OctreeTracer() { release(); }
	bool Resize(uint32_t width, uint32_t height) {
		release();
		m_width = width;
		m_height = height;
		auto &ctx = m_octree->GetContextPtr();
		const uint64_t bw = (width + 7) / 8 + 1, bh = (height + 7) / 8 + 1;
		return ctx->Check(svo_malloc(ctx->h, (uint64_t)width * height * sizeof(svo_hit), (void **)&m_hits)) &&
		       ctx->Check(svo_malloc(ctx->h, (uint64_t)width * height * sizeof(uint32_t), (void **)&m_image)) &&
		       ctx->Check(svo_malloc(ctx->h, bw * bh * sizeof(float), (void **)&m_beam));
	}
	bool CmdBeamRenderPass() const {
		if (!m_beam_enable)
			return true;
		auto &ctx = m_octree->GetContextPtr();
		svo_camera c = m_camera->fetch_uniform_data();
		return ctx->Check(svo_march_beam(ctx->h, m_octree->Handle(), &c, m_width, m_height, m_beam, nullptr));
	}
	// Runs the beam pre-pass itself when m_beam_enable is set (so the pair can be called in order),
	// then shades the hits by m_view_type.
	bool CmdDrawPipeline() const;
	svo_hit *GetHits() const { return m_hits; }
	uint32_t *GetImage() const { return m_image; } // width * height RGBA8 UNORM texels
	std::vector<svo_hit> DownloadHits() const {
		std::vector<svo_hit> out((size_t)m_width * m_height);
		auto &ctx = m_octree->GetContextPtr();
		ctx->Check(svo_download(ctx->h, out.data(), m_hits, out.size() * sizeof(svo_hit)));
		return out;
	}
	std::vector<uint32_t> DownloadImage() const {
		std::vector<uint32_t> out((size_t)m_width * m_height);
		auto &ctx = m_octree->GetContextPtr();
		ctx->Check(svo_download(ctx->h, out.data(), m_image, out.size() * sizeof(uint32_t)));
		return out;
	}
	uint32_t GetWidth() const { return m_width; }
	uint32_t GetHeight() const { return m_height; }

private:
	OctreeTracer() = default;
	void release() {
		if (!m_octree)
			return;
		auto &ctx = m_octree->GetContextPtr();
		for (void *b : {(void *)m_hits, (void *)m_image, (void *)m_beam})
			if (b)
				svo_free(ctx->h, b);
		m_hits = nullptr;
		m_image = nullptr;
		m_beam = nullptr;
	}
	std::shared_ptr<Octree> m_octree;
	std::shared_ptr<Camera> m_camera;
	uint32_t m_width = 0, m_height = 0;
	svo_hit *m_hits = nullptr;
	uint32_t *m_image = nullptr;
	float *m_beam = nullptr;
};

// EnvironmentMap: an equirectangular map of linear RGBA texels with its sampling tables in HBM
// (src/EnvironmentMap.cpp:99-188); m_rotation and m_multiplier as in the reference.
class EnvironmentMap {
public:
	float m_rotation = 0.0f, m_multiplier = 1.0f;

	static std::shared_ptr<EnvironmentMap> Create(const std::shared_ptr<Context> &ctx, uint32_t width, uint32_t height,
	                                              const float *rgba) {
		auto e = std::shared_ptr<EnvironmentMap>(new EnvironmentMap());
		e->m_ctx = ctx;
		if (!ctx->Check(svo_envmap_create(ctx->h, width, height, rgba, &e->m_h)))
			return nullptr;
		return e;
	}
	~EnvironmentMap() {
		if (m_h)
			svo_envmap_destroy(m_h);
	}
	const svo_envmap *Handle() const { return m_h; }

private:
	EnvironmentMap() = default;
	std::shared_ptr<Context> m_ctx;
	svo_envmap *m_h = nullptr;
};

inline bool OctreeTracer::CmdDrawPipeline() const {
	auto &ctx = m_octree->GetContextPtr();
	svo_camera c = m_camera->fetch_uniform_data();
	if (!ctx->Check(svo_march(ctx->h, m_octree->Handle(), &c, m_width, m_height, m_beam_enable ? 1 : 0, m_hits,
	                          m_beam_enable ? m_beam : nullptr)))
		return false;
	return ctx->Check(svo_shade(ctx->h, m_hits, &c, m_width, m_height, (int)m_view_type, m_light,
	                            m_env_map ? m_env_map->Handle() : nullptr, m_env_map ? m_env_map->m_rotation : 0.0f,
	                            m_image));
}

// PathTracer: progressive samples into device running means (PathTracer::Reset / CmdRender,
// src/PathTracer.cpp:176-215) in the reference's image formats: colour RGBA32F (float4 per pixel),
// first-hit albedo RGBA8 UNORM and first-hit normal RGBA8 SNORM (one uint32 per pixel).  Sample n
// uses the reference's sampler (Sobol state after n advances, blue noise).  The light is m_env_map
// when set (EnvMapPathTrace), else the constant m_sun_radiance.
class PathTracer {
public:
	uint32_t m_bounce = 4;                      // kDefaultBounce, src/Config.hpp:25
	float m_sun_radiance[3] = {5.0f, 5.0f, 5.0f}; // kDefaultConstantColor, src/Config.hpp:28
	std::shared_ptr<EnvironmentMap> m_env_map;

	static std::shared_ptr<PathTracer> Create(const std::shared_ptr<Octree> &octree,
	                                          const std::shared_ptr<Camera> &camera) {
		auto t = std::shared_ptr<PathTracer>(new PathTracer());
		t->m_octree = octree;
		t->m_camera = camera;
		return t->Resize(1280, 720) ? t : nullptr;
	}
	~PathTracer() { release(); }
	bool Resize(uint32_t width, uint32_t height) {
		release();
		m_width = width;
		m_height = height;
		m_spp = 0;
		auto &ctx = m_octree->GetContextPtr();
		const uint64_t px = (uint64_t)width * height;
		return ctx->Check(svo_malloc(ctx->h, px * 4 * sizeof(float), (void **)&m_color)) &&
		       ctx->Check(svo_malloc(ctx->h, px * sizeof(uint32_t), (void **)&m_albedo)) &&
		       ctx->Check(svo_malloc(ctx->h, px * sizeof(uint32_t), (void **)&m_normal));
	}
	void Reset() { m_spp = 0; } // the next sample overwrites the means, as the cleared images
	bool CmdRender() {
		auto &ctx = m_octree->GetContextPtr();
		svo_camera c = m_camera->fetch_uniform_data();
		const int rc =
		    m_env_map ? svo_path_trace_env(ctx->h, m_octree->Handle(), &c, m_width, m_height, m_bounce, m_spp,
		                                   m_env_map->Handle(), m_env_map->m_rotation, m_env_map->m_multiplier,
		                                   m_color, m_albedo, m_normal)
		              : svo_path_trace(ctx->h, m_octree->Handle(), &c, m_width, m_height, m_bounce, m_spp,
		                               m_sun_radiance, m_color, m_albedo, m_normal);
		if (!ctx->Check(rc))
			return false;
		++m_spp; // Sobol::CmdNext after the dispatch (src/PathTracer.cpp:214)
		return true;
	}
	uint32_t GetSampleCount() const { return m_spp; }
	std::vector<float> DownloadColor() const { return download<float>(m_color, 4); }  // width * height float4
	std::vector<uint32_t> DownloadAlbedo() const { return download<uint32_t>(m_albedo, 1); } // RGBA8 UNORM
	std::vector<uint32_t> DownloadNormal() const { return download<uint32_t>(m_normal, 1); } // RGBA8 SNORM

private:
	PathTracer() = default;
	template <class T> std::vector<T> download(const void *src, size_t per_pixel) const {
		std::vector<T> out((size_t)m_width * m_height * per_pixel);
		auto &ctx = m_octree->GetContextPtr();
		ctx->Check(svo_download(ctx->h, out.data(), src, out.size() * sizeof(T)));
		return out;
	}
	void release() {
		if (!m_octree)
			return;
		auto &ctx = m_octree->GetContextPtr();
		for (void *b : {(void *)m_color, (void *)m_albedo, (void *)m_normal})
			if (b)
				svo_free(ctx->h, b);
		m_color = nullptr;
		m_albedo = m_normal = nullptr;
	}
	std::shared_ptr<Octree> m_octree;
	std::shared_ptr<Camera> m_camera;
	uint32_t m_width = 0, m_height = 0, m_spp = 0;
	float *m_color = nullptr;
	uint32_t *m_albedo = nullptr, *m_normal = nullptr;
};

} // namespace svo
