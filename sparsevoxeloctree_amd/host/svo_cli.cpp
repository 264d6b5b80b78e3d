// svo_cli: the reference's load path (LoaderThread::thread_func, src/LoaderThread.cpp:43-116) on the
// C++ class surface: Scene -> Voxelizer (count pass) -> timed {CmdVoxelize, CmdBuild} ->
// Octree::Update, then optionally one primary frame through OctreeTracer.  As in the reference, the
// scene is loaded and built on a loader context (its own HIP stream, the loader queue) and handed to
// the main context with OctreeBuilder::CmdTransferOctreeOwnership (src/LoaderThread.cpp:77-105);
// the tracer and path tracer run on the main context.
//
//   svo_cli -scene KIND[,SEED[,PARAM]] -lvl N [-frame W H] [-no-beam] [-view V [-image-out FILE]]
//           [-save FILE] [-pt SAMPLES BOUNCE [-env FILE] [-pt-out FILE]]
//   svo_cli -tris FILE.bin -lvl N ...      (FILE.bin: uint64 T, T*9 float32 xyz, T uint32 rgb)
//
// Prints the reference's timing line, "FINISHED in {total} ms (Voxelize {v} ms, Octree building
// {b} ms)" (src/LoaderThread.cpp:92-97), measured on the device, then the fragment count and the
// octree's used range.  -save writes the octree with svo_octree_save.  -pt renders SAMPLES
// progressive path-traced samples of the -frame view through PathTracer (constant sky, or the
// environment map in FILE: uint32 width, height, then width*height float32 RGBA) and -pt-out
// writes the colour means (width*height float32 RGBA).  -view picks OctreeTracer::m_view_type (0
// diffuse, 1 normal, 2 position, 3 iterations) and -image-out writes the shaded frame (width*height
// uint32 RGBA8).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "svo.hpp"

static int usage() {
	std::fprintf(stderr, "usage: svo_cli (-scene KIND[,SEED[,PARAM]] | -tris FILE | -obj FILE.obj) -lvl N "
	                     "[-frame W H] [-no-beam] [-view V [-image-out FILE]] [-save FILE] "
	                     "[-pt SAMPLES BOUNCE [-env FILE] [-pt-out FILE]]\n");
	return 2;
}

int main(int argc, char **argv) {
	int kind = 0;
	uint64_t seed = 1, param = 0;
	std::string tris_path, obj_path, save_path, env_path, pt_out, image_out;
	uint32_t level = 0, fw = 0, fh = 0, pt_samples = 0, pt_bounce = 4, view = 0;
	bool beam = true;
	for (int i = 1; i < argc; ++i) {
		if (!std::strcmp(argv[i], "-scene") && i + 1 < argc) {
			unsigned long long s = 1, p = 0;
			int n = std::sscanf(argv[++i], "%d,%llu,%llu", &kind, &s, &p);
			if (n < 1)
				return usage();
			seed = s, param = p;
		} else if (!std::strcmp(argv[i], "-tris") && i + 1 < argc) {
			tris_path = argv[++i];
		} else if (!std::strcmp(argv[i], "-obj") && i + 1 < argc) {
			obj_path = argv[++i]; // the reference's LoaderThread input (src/LoaderThread.cpp:43-56)
		} else if (!std::strcmp(argv[i], "-lvl") && i + 1 < argc) {
			level = (uint32_t)std::atoi(argv[++i]);
		} else if (!std::strcmp(argv[i], "-frame") && i + 2 < argc) {
			fw = (uint32_t)std::atoi(argv[++i]);
			fh = (uint32_t)std::atoi(argv[++i]);
		} else if (!std::strcmp(argv[i], "-no-beam")) {
			beam = false;
		} else if (!std::strcmp(argv[i], "-save") && i + 1 < argc) {
			save_path = argv[++i];
		} else if (!std::strcmp(argv[i], "-pt") && i + 2 < argc) {
			pt_samples = (uint32_t)std::atoi(argv[++i]);
			pt_bounce = (uint32_t)std::atoi(argv[++i]);
		} else if (!std::strcmp(argv[i], "-env") && i + 1 < argc) {
			env_path = argv[++i];
		} else if (!std::strcmp(argv[i], "-pt-out") && i + 1 < argc) {
			pt_out = argv[++i];
		} else if (!std::strcmp(argv[i], "-view") && i + 1 < argc) {
			view = (uint32_t)std::atoi(argv[++i]);
		} else if (!std::strcmp(argv[i], "-image-out") && i + 1 < argc) {
			image_out = argv[++i];
		} else
			return usage();
	}
	if ((kind != 0) + !tris_path.empty() + !obj_path.empty() != 1 || level == 0)
		return usage();

	std::vector<float> xyz;
	std::vector<uint32_t> rgb;
	uint64_t T = 0;
	if (kind) {
		if (svo_gen_scene(kind, seed, param, nullptr, nullptr, &T) != SVO_OK)
			return usage();
		xyz.resize(9 * T);
		rgb.resize(T);
		svo_gen_scene(kind, seed, param, xyz.data(), rgb.data(), &T);
	} else if (!obj_path.empty()) {
		float *px = nullptr;
		uint32_t *pc = nullptr;
		char msg[512] = {0};
		if (svo_obj_load(obj_path.c_str(), &px, &pc, &T, msg, sizeof msg) != SVO_OK) {
			std::fprintf(stderr, "%s: %s\n", obj_path.c_str(), msg);
			return 1;
		}
		xyz.assign(px, px + 9 * T);
		rgb.assign(pc, pc + T);
		svo_obj_free(px, pc);
	} else {
		FILE *f = std::fopen(tris_path.c_str(), "rb");
		if (!f || std::fread(&T, 8, 1, f) != 1) {
			std::fprintf(stderr, "cannot read %s\n", tris_path.c_str());
			return 1;
		}
		xyz.resize(9 * T);
		rgb.resize(T);
		if (std::fread(xyz.data(), 4, 9 * T, f) != 9 * T || std::fread(rgb.data(), 4, T, f) != T) {
			std::fprintf(stderr, "truncated %s\n", tris_path.c_str());
			return 1;
		}
		std::fclose(f);
	}

	auto ctx = svo::Context::Create(0);    // the main queue
	auto loader = svo::Context::Create(0); // the loader queue (src/LoaderThread.cpp:43-48)
	if (!ctx || !loader) {
		std::fprintf(stderr, "no HIP device\n");
		return 1;
	}
	auto scene = svo::Scene::Create(loader, xyz.data(), rgb.data(), T);
	if (!scene) {
		std::fprintf(stderr, "scene: %s\n", loader->error.c_str());
		return 1;
	}
	auto voxelizer = svo::Voxelizer::Create(scene, level);
	if (!voxelizer) {
		std::fprintf(stderr, "voxelizer: %s\n", loader->error.c_str());
		return 1;
	}
	auto builder = svo::OctreeBuilder::Create(voxelizer);
	// src/LoaderThread.cpp:57-75: the timed span is the fill draw plus all octree passes; then the
	// octree is handed to the main queue (:77-82)
	if (!voxelizer->CmdVoxelize() || !builder->CmdBuild() || !builder->CmdTransferOctreeOwnership(ctx)) {
		std::fprintf(stderr, "build: %s %s\n", loader->error.c_str(), ctx->error.c_str());
		return 1;
	}
	auto octree = svo::Octree::Create(ctx);
	octree->Update(builder); // waits for the build (GetOctreeRange)
	if (octree->GetRange() == 0) {
		std::fprintf(stderr, "build: %s\n", loader->error.c_str());
		return 1;
	}
	svo_timing t = loader->Timing();
	std::printf("FINISHED in %.3f ms (Voxelize %.3f ms, Octree building %.3f ms)\n", t.vox_fill_ms + t.build_ms,
	            t.vox_fill_ms, t.build_ms);
	std::printf("triangles %llu fragments %u octree_words %llu level %u\n", (unsigned long long)T,
	            voxelizer->GetVoxelFragmentCount(), (unsigned long long)(octree->GetRange() / 4), level);

	if (!save_path.empty()) {
		if (!ctx->Check(svo_octree_save(ctx->h, octree->Handle(), save_path.c_str()))) {
			std::fprintf(stderr, "save: %s\n", ctx->error.c_str());
			return 1;
		}
		std::printf("saved %s\n", save_path.c_str());
	}
	if (fw && fh) {
		auto camera = svo::Camera::Create();
		camera->m_aspect_ratio = (float)fw / (float)fh;
		camera->m_position[2] = 1.04f; // near the -z face, looking along +z
		auto tracer = svo::OctreeTracer::Create(octree, camera);
		tracer->m_beam_enable = beam;
		tracer->m_view_type = (svo::OctreeTracer::ViewTypes)(view & 3u);
		if (!tracer->Resize(fw, fh) || !tracer->CmdBeamRenderPass() || !tracer->CmdDrawPipeline()) {
			std::fprintf(stderr, "tracer: %s\n", ctx->error.c_str());
			return 1;
		}
		std::vector<svo_hit> hits = tracer->DownloadHits();
		uint64_t nhit = 0, iters = 0;
		for (const svo_hit &h : hits) {
			nhit += h.vz_flags >> 31;
			iters += h.iter;
		}
		svo_timing t2 = ctx->Timing();
		std::printf("frame %ux%u hits %llu mean_iter %.2f beam %.3f ms march %.3f ms\n", fw, fh,
		            (unsigned long long)nhit, (double)iters / hits.size(), beam ? t2.beam_ms : 0.0, t2.march_ms);
		if (!image_out.empty()) {
			std::vector<uint32_t> img = tracer->DownloadImage();
			FILE *f = std::fopen(image_out.c_str(), "wb");
			if (!f || std::fwrite(img.data(), sizeof(uint32_t), img.size(), f) != img.size()) {
				std::fprintf(stderr, "cannot write %s\n", image_out.c_str());
				return 1;
			}
			std::fclose(f);
			std::printf("image view %u written to %s\n", view & 3u, image_out.c_str());
		}
		if (pt_samples) {
			auto pt = svo::PathTracer::Create(octree, camera);
			if (!pt || !pt->Resize(fw, fh)) {
				std::fprintf(stderr, "path tracer: %s\n", ctx->error.c_str());
				return 1;
			}
			pt->m_bounce = pt_bounce;
			if (!env_path.empty()) {
				FILE *f = std::fopen(env_path.c_str(), "rb");
				uint32_t wh[2] = {0, 0};
				std::vector<float> rgba;
				bool ok = f && std::fread(wh, sizeof(uint32_t), 2, f) == 2 && wh[0] && wh[1] &&
				          (uint64_t)wh[0] * wh[1] <= (1ull << 28);
				if (ok) {
					rgba.resize((size_t)wh[0] * wh[1] * 4);
					ok = std::fread(rgba.data(), sizeof(float), rgba.size(), f) == rgba.size();
				}
				if (f)
					std::fclose(f);
				if (!ok) {
					std::fprintf(stderr, "cannot read environment map %s\n", env_path.c_str());
					return 1;
				}
				pt->m_env_map = svo::EnvironmentMap::Create(ctx, wh[0], wh[1], rgba.data());
				if (!pt->m_env_map) {
					std::fprintf(stderr, "environment map: %s\n", ctx->error.c_str());
					return 1;
				}
			}
			const auto t0 = std::chrono::steady_clock::now();
			for (uint32_t s = 0; s < pt_samples; ++s)
				if (!pt->CmdRender()) {
					std::fprintf(stderr, "path tracer: %s\n", ctx->error.c_str());
					return 1;
				}
			ctx->Synchronize();
			const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
			std::vector<float> col = pt->DownloadColor();
			double mean[3] = {0.0, 0.0, 0.0};
			for (size_t k = 0; k < col.size(); k += 4)
				for (int c = 0; c < 3; ++c)
					mean[c] += col[k + c];
			const double npx = (double)fw * fh;
			std::printf("path_trace %s samples %u bounce %u mean %.6f %.6f %.6f %.3f ms/sample\n",
			            pt->m_env_map ? "env" : "sky", pt->GetSampleCount(), pt_bounce, mean[0] / npx, mean[1] / npx,
			            mean[2] / npx, ms / pt_samples);
			if (!pt_out.empty()) {
				FILE *f = std::fopen(pt_out.c_str(), "wb");
				const bool ok = f && std::fwrite(col.data(), sizeof(float), col.size(), f) == col.size();
				if (f)
					std::fclose(f);
				if (!ok) {
					std::fprintf(stderr, "cannot write %s\n", pt_out.c_str());
					return 1;
				}
			}
		}
	}
	return 0;
}
