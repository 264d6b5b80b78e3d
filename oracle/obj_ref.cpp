// Test infrastructure: the reference's OBJ path as a command-line checker.  Built by
// oracle/build_oracle.py against the reference's own tinyobj reader and glm (dep/ under the
// reference tree, compiled where they lie) into oracle/_ref/objref.  It loads a file with
// tinyobj::LoadObj exactly as Scene::load_meshes does (src/Scene.cpp:38-57: default triangulation,
// base directory of the file, at least one material), groups the triangles per material, orders the
// groups by size, largest first, with equal sizes in material order (src/Scene.cpp:118-126 uses an
// unspecified order there), drops empty groups, and prints one line per triangle: the 9 position
// floats as hexadecimal bit patterns, then glm::packUnorm4x8(vec4(Kd, 0)) & 0xffffff
// (src/Scene.cpp:164).  Positions are not normalised.  Exit status 1 on a load error.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "glm/glm.hpp"
#include "glm/gtc/packing.hpp"
#include "tiny_obj_loader.h"

int main(int argc, char **argv) {
	if (argc != 2) {
		std::fprintf(stderr, "usage: objref file.obj\n");
		return 2;
	}
	const std::string path = argv[1];
	const size_t slash = path.find_last_of("/\\");
	const std::string dir = slash == std::string::npos ? std::string() : path.substr(0, slash + 1);
	tinyobj::attrib_t attrib;
	std::vector<tinyobj::shape_t> shapes;
	std::vector<tinyobj::material_t> materials;
	std::string warn, err;
	if (!tinyobj::LoadObj(&attrib, &shapes, &materials, &warn, &err, path.c_str(), dir.c_str()) || materials.empty() ||
	    !err.empty()) {
		std::printf("error\n");
		return 1;
	}
	std::vector<std::vector<float>> per(materials.size());
	for (const tinyobj::shape_t &sh : shapes) {
		size_t off = 0;
		for (size_t f = 0; f < sh.mesh.num_face_vertices.size(); ++f) {
			const int m = sh.mesh.material_ids[f];
			const size_t nv = sh.mesh.num_face_vertices[f];
			if (m < 0 || (size_t)m >= materials.size()) {
				std::printf("error\n"); // the reference indexes meshes->at(m) here
				return 1;
			}
			for (size_t v = 0; v < nv; ++v) {
				const int vi = sh.mesh.indices[off + v].vertex_index;
				for (int a = 0; a < 3; ++a)
					per[m].push_back(attrib.vertices[3 * (size_t)vi + a]);
			}
			off += nv;
		}
	}
	std::vector<size_t> order;
	for (size_t m = 0; m < per.size(); ++m)
		if (!per[m].empty())
			order.push_back(m);
	if (order.empty()) {
		std::printf("error\n");
		return 1;
	}
	std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return per[a].size() > per[b].size(); });
	for (size_t m : order) {
		const glm::vec3 kd(materials[m].diffuse[0], materials[m].diffuse[1], materials[m].diffuse[2]);
		const uint32_t colour = glm::packUnorm4x8(glm::vec4(kd, 0.0f)) & 0xffffffu;
		const std::vector<float> &p = per[m];
		for (size_t t = 0; t + 9 <= p.size(); t += 9) {
			for (int k = 0; k < 9; ++k) {
				uint32_t bits;
				std::memcpy(&bits, &p[t + k], 4);
				std::printf("%08x ", bits);
			}
			std::printf("%06x\n", colour);
		}
	}
	return 0;
}
