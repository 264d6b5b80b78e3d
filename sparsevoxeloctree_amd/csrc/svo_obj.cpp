// Wavefront OBJ + MTL ingestion (Scene::load_meshes, src/Scene.cpp:36-147, which reads files with
// tinyobj::LoadObj, dep/tiny_obj_loader.h): the triangle soup and per-triangle colour that
// svo_scene_create takes.  Host code; the parse follows the reader the reference links, so that
// the same file gives the same vertex bits and the same triangles:
//   - numbers: decimal mantissa accumulated in double (integer digits times 10, fraction digits
//     times 10^-k), a decimal exponent applied as ldexp(m * 5^e, e), then rounded to float
//     (dep/tiny_obj_loader.h:837-960); a token that is not a number reads as 0;
//   - face indices: 1-based, negative = relative to the vertices read so far, 0 is an error;
//     texture and normal indices are checked the same way and then ignored;
//   - polygons with more than 3 corners: ear clipping in the plane of the two axes the first
//     non-degenerate corner does not face (dep/tiny_obj_loader.h:1391-1590), point-in-triangle by
//     the even-odd crossing rule;
//   - usemtl selects the material of later faces; mtllib files are read relative to the OBJ file.
// Then, as Scene::load_meshes does, the faces are grouped per material, the groups ordered by size
// (largest first; equal sizes keep material order, where the reference's std::sort leaves them
// unspecified) and the empty ones dropped.  A triangle's colour is packUnorm4x8(vec4(Kd, 0)) &
// 0xffffff of its material, packed on the host by glm (src/Scene.cpp:164): round(clamp(c, 0, 1) *
// 255) in float, halfway cases away from zero.  Diffuse textures are not read: a textured material
// contributes its Kd.
//
// The number parser and the polygon triangulation restate steps of tinyobjloader, which carries
// this notice:
//
//   The MIT License (MIT)
//
//   Copyright (c) 2012-2018 Syoyo Fujita and many contributors.
//
//   Permission is hereby granted, free of charge, to any person obtaining a copy
//   of this software and associated documentation files (the "Software"), to deal
//   in the Software without restriction, including without limitation the rights
//   to use, copy, modify, merge, publish, distribute, sublicense, and/or sell
//   copies of the Software, and to permit persons to whom the Software is
//   furnished to do so, subject to the following conditions:
//
//   The above copyright notice and this permission notice shall be included in
//   all copies or substantial portions of the Software.
//
//   THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS OR
//   IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF MERCHANTABILITY,
//   FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT. IN NO EVENT SHALL THE
//   AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM, DAMAGES OR OTHER
//   LIABILITY, WHETHER IN AN ACTION OF CONTRACT, TORT OR OTHERWISE, ARISING FROM,
//   OUT OF OR IN CONNECTION WITH THE SOFTWARE OR THE USE OR OTHER DEALINGS IN
//   THE SOFTWARE.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <limits>
#include <map>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "svo.h"

namespace {

bool is_digit(char c) { return c >= '0' && c <= '9'; }

// The number in [s, e), by the reference reader's accumulation; false if it is not one.
bool parse_number(const char *s, const char *e, double *out) {
	if (s >= e)
		return false;
	double m = 0.0;
	int exponent = 0;
	bool negative = false, exp_negative = false;
	const char *c = s;
	bool leading_dot = false;
	if (*c == '+' || *c == '-') {
		negative = *c == '-';
		++c;
		if (c != e && *c == '.')
			leading_dot = true;
	} else if (*c == '.') {
		leading_dot = true;
	} else if (!is_digit(*c)) {
		return false;
	}
	if (!leading_dot) {
		int digits = 0;
		while (c != e && is_digit(*c)) {
			m = m * 10 + (double)(*c - '0');
			++c;
			++digits;
		}
		if (digits == 0)
			return false;
	}
	if (c != e) {
		if (*c == '.') {
			++c;
			static const double kTenth[8] = {1.0, 0.1, 0.01, 0.001, 0.0001, 0.00001, 0.000001, 0.0000001};
			for (int k = 1; c != e && is_digit(*c); ++k, ++c)
				m += (double)(*c - '0') * (k < 8 ? kTenth[k] : std::pow(10.0, -k));
		}
		if (c != e && (*c == 'e' || *c == 'E')) {
			++c;
			if (c != e && (*c == '+' || *c == '-')) {
				exp_negative = *c == '-';
				++c;
			} else if (c == e || !is_digit(*c)) {
				return false;
			}
			int digits = 0;
			while (c != e && is_digit(*c)) {
				exponent = exponent * 10 + (*c - '0');
				++c;
				++digits;
			}
			if (digits == 0)
				return false;
			if (exp_negative)
				exponent = -exponent;
		}
	}
	const double v = exponent ? std::ldexp(m * std::pow(5.0, exponent), exponent) : m;
	*out = negative ? -v : v;
	return true;
}

const char *skip_blank(const char *p) {
	while (*p == ' ' || *p == '\t')
		++p;
	return p;
}

// Next whitespace-delimited token as a float (0 when it is not a number).
float next_real(const char *&p) {
	p = skip_blank(p);
	const char *e = p + std::strcspn(p, " \t\r");
	double v = 0.0;
	parse_number(p, e, &v);
	p = e;
	return (float)v;
}

std::string next_name(const char *&p) {
	p = skip_blank(p);
	const size_t n = std::strcspn(p, " \t\r");
	std::string s(p, n);
	p += n;
	return s;
}

// 1-based or negative (relative) index -> 0-based; false for 0.
bool resolve_index(int idx, int n, int *out) {
	if (idx > 0) {
		*out = idx - 1;
		return true;
	}
	if (idx < 0) {
		*out = n + idx;
		return true;
	}
	return false;
}

struct Face {
	std::vector<int> v; // 0-based vertex indices
	int material;
};

struct Parsed {
	std::vector<float> pos; // 3 per vertex
	std::vector<Face> faces;
	std::vector<std::string> names;
	std::vector<float> kd; // 3 per material
	std::map<std::string, int> material_of;
};

bool read_lines(const std::string &path, std::vector<std::string> &lines) {
	std::ifstream f(path, std::ios::binary);
	if (!f)
		return false;
	// a line ends at "\n", "\r\n" or a lone "\r", as the reference reader's safeGetline
	// (dep/tiny_obj_loader.h:731-760)
	std::string data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
	std::string line;
	for (size_t i = 0; i < data.size(); ++i) {
		const char c = data[i];
		if (c == '\n' || c == '\r') {
			if (c == '\r' && i + 1 < data.size() && data[i + 1] == '\n')
				++i;
			lines.push_back(line);
			line.clear();
		} else {
			line.push_back(c);
		}
	}
	if (!line.empty())
		lines.push_back(line);
	return true;
}

// Materials of one MTL file; false if it cannot be read.  A material's name is the rest of its
// newmtl line; a repeated name adds a material that usemtl never selects (the first one keeps the
// name), as in the reference reader.
bool parse_mtl(const std::string &path, Parsed &ps) {
	std::vector<std::string> lines;
	if (!read_lines(path, lines))
		return false;
	int cur = -1;
	for (const std::string &l : lines) {
		const char *p = skip_blank(l.c_str());
		if (std::strncmp(p, "newmtl", 6) == 0 && (p[6] == ' ' || p[6] == '\t')) {
			std::string name = skip_blank(p + 7);
			while (!name.empty() && (name.back() == ' ' || name.back() == '\t'))
				name.pop_back();
			cur = (int)ps.names.size();
			ps.material_of.insert({name, cur});
			ps.names.push_back(name);
			ps.kd.insert(ps.kd.end(), {0.0f, 0.0f, 0.0f});
		} else if (p[0] == 'K' && p[1] == 'd' && (p[2] == ' ' || p[2] == '\t') && cur >= 0) {
			p += 2;
			for (int k = 0; k < 3; ++k)
				ps.kd[3 * cur + k] = next_real(p);
		}
	}
	return true;
}

bool parse_obj(const std::string &path, Parsed &ps, std::string &err) {
	std::vector<std::string> lines;
	if (!read_lines(path, lines)) {
		err = "cannot read " + path;
		return false;
	}
	const size_t slash = path.find_last_of("/\\");
	const std::string dir = slash == std::string::npos ? std::string() : path.substr(0, slash + 1);
	int material = -1;
	size_t vt = 0, vn = 0;
	for (size_t ln = 0; ln < lines.size(); ++ln) {
		const char *p = skip_blank(lines[ln].c_str());
		if (p[0] == 'v' && (p[1] == ' ' || p[1] == '\t')) {
			p += 2;
			for (int k = 0; k < 3; ++k)
				ps.pos.push_back(next_real(p));
		} else if (p[0] == 'v' && p[1] == 't' && (p[2] == ' ' || p[2] == '\t')) {
			++vt;
		} else if (p[0] == 'v' && p[1] == 'n' && (p[2] == ' ' || p[2] == '\t')) {
			++vn;
		} else if (p[0] == 'f' && (p[1] == ' ' || p[1] == '\t')) {
			p = skip_blank(p + 2);
			Face f;
			f.material = material;
			const int nv = (int)(ps.pos.size() / 3);
			while (*p) {
				int v, t, n;
				const size_t len = std::strcspn(p, "/ \t\r");
				if (len == 0) { // not an index token (lines hold no '\r': they end there)
					err = "bad face token on line " + std::to_string(ln + 1);
					return false;
				}
				bool ok = resolve_index(std::atoi(p), nv, &v);
				p += len;
				if (ok && *p == '/') {
					++p;
					if (*p == '/') { // i//k
						++p;
						ok = resolve_index(std::atoi(p), (int)vn, &n);
						p += std::strcspn(p, "/ \t\r");
					} else { // i/j or i/j/k
						ok = resolve_index(std::atoi(p), (int)vt, &t);
						p += std::strcspn(p, "/ \t\r");
						if (ok && *p == '/') {
							++p;
							ok = resolve_index(std::atoi(p), (int)vn, &n);
							p += std::strcspn(p, "/ \t\r");
						}
					}
				}
				if (!ok) {
					err = "bad face index on line " + std::to_string(ln + 1) + " (0 is not an index)";
					return false;
				}
				if (v < 0 || v >= nv) {
					err = "face index out of range on line " + std::to_string(ln + 1);
					return false;
				}
				f.v.push_back(v);
				p = skip_blank(p);
			}
			if (f.v.size() >= 3)
				ps.faces.push_back(std::move(f));
		} else if (std::strncmp(p, "usemtl", 6) == 0 && (p[6] == ' ' || p[6] == '\t')) {
			p += 6;
			auto it = ps.material_of.find(next_name(p));
			material = it == ps.material_of.end() ? -1 : it->second;
		} else if (std::strncmp(p, "mtllib", 6) == 0 && (p[6] == ' ' || p[6] == '\t')) {
			p += 6;
			for (;;) { // one or more file names: the first that can be read is used
				const std::string name = next_name(p);
				if (name.empty() || parse_mtl(dir + name, ps))
					break;
			}
		}
	}
	return true;
}

// Even-odd crossing test of (tx, ty) against the triangle (x[3], y[3]) (W. R. Franklin's pnpoly).
bool inside_triangle(const float *x, const float *y, float tx, float ty) {
	bool c = false;
	for (int i = 0, j = 2; i < 3; j = i++)
		if (((y[i] > ty) != (y[j] > ty)) && (tx < (x[j] - x[i]) * (ty - y[i]) / (y[j] - y[i]) + x[i]))
			c = !c;
	return c;
}

// Triangles of one polygon, appended as vertex index triples.
void triangulate(const std::vector<float> &v, const std::vector<int> &poly, std::vector<int> &out) {
	const size_t n0 = poly.size();
	if (n0 == 3) {
		out.insert(out.end(), poly.begin(), poly.end());
		return;
	}
	// the plane: drop the axis of the largest normal component at the first corner that has one
	int ax0 = 1, ax1 = 2;
	for (size_t k = 0; k < n0; ++k) {
		const float *a = &v[3 * poly[k]], *b = &v[3 * poly[(k + 1) % n0]], *c = &v[3 * poly[(k + 2) % n0]];
		const float e0x = b[0] - a[0], e0y = b[1] - a[1], e0z = b[2] - a[2];
		const float e1x = c[0] - b[0], e1y = c[1] - b[1], e1z = c[2] - b[2];
		const float cx = std::fabs(e0y * e1z - e0z * e1y), cy = std::fabs(e0z * e1x - e0x * e1z),
		            cz = std::fabs(e0x * e1y - e0y * e1x);
		const float eps = std::numeric_limits<float>::epsilon();
		if (cx > eps || cy > eps || cz > eps) {
			if (!(cx > cy && cx > cz)) {
				ax0 = 0;
				if (cz > cx && cz > cy)
					ax1 = 1;
			}
			break;
		}
	}
	float area = 0.0f; // signed area in that plane: its sign gives the winding
	for (size_t k = 0; k < n0; ++k) {
		const float *a = &v[3 * poly[k]], *b = &v[3 * poly[(k + 1) % n0]];
		area += (a[ax0] * b[ax1] - a[ax1] * b[ax0]) * 0.5f;
	}
	std::vector<int> rem = poly;
	size_t guess = 0, tries = rem.size(), last_size = rem.size();
	while (rem.size() > 3 && tries > 0) {
		const size_t n = rem.size();
		if (guess >= n)
			guess -= n;
		if (last_size != n) {
			last_size = n;
			tries = n;
		} else {
			--tries;
		}
		int ind[3];
		float x[3], y[3];
		for (int k = 0; k < 3; ++k) {
			ind[k] = rem[(guess + k) % n];
			x[k] = v[3 * ind[k] + ax0];
			y[k] = v[3 * ind[k] + ax1];
		}
		const float e0x = x[1] - x[0], e0y = y[1] - y[0], e1x = x[2] - x[1], e1y = y[2] - y[1];
		const float cross = e0x * e1y - e0y * e1x;
		if (cross * area < 0.0f) { // a reflex corner
			guess += 1;
			continue;
		}
		bool overlap = false;
		for (size_t o = 3; o < n; ++o) {
			const int vi = rem[(guess + o) % n];
			if (inside_triangle(x, y, v[3 * vi + ax0], v[3 * vi + ax1])) {
				overlap = true;
				break;
			}
		}
		if (overlap) {
			guess += 1;
			continue;
		}
		out.insert(out.end(), {ind[0], ind[1], ind[2]});
		rem.erase(rem.begin() + (long)((guess + 1) % n)); // clip the ear's middle corner
	}
	if (rem.size() == 3)
		out.insert(out.end(), rem.begin(), rem.end());
}

// glm::packUnorm4x8's channel: round(clamp(c, 0, 1) * 255) in float, halfway cases away from zero.
uint32_t unorm8(float c) {
	c = c < 0.0f ? 0.0f : c > 1.0f ? 1.0f : c;
	return (uint32_t)std::round(c * 255.0f);
}

} // namespace

namespace {
int obj_load(const char *path, float **xyz, uint32_t **rgb, uint64_t *T, char *msg, uint64_t msgcap) {
	auto fail = [&](int code, const std::string &m) {
		if (msg && msgcap) {
			std::snprintf(msg, msgcap, "%s", m.c_str());
		}
		return code;
	};
	if (!path || !xyz || !rgb || !T)
		return fail(SVO_ERR_INVALID_ARGUMENT, "null argument");
	*xyz = nullptr;
	*rgb = nullptr;
	*T = 0;
	Parsed ps;
	std::string err;
	if (!parse_obj(path, ps, err))
		return fail(SVO_ERR_INVALID_ARGUMENT, err);
	if (ps.names.empty())
		return fail(SVO_ERR_INVALID_ARGUMENT, "no material found"); // src/Scene.cpp:47-50
	// triangles per material, in file order
	std::vector<std::vector<int>> per(ps.names.size());
	std::vector<int> tri;
	for (const Face &f : ps.faces) {
		if (f.material < 0)
			return fail(SVO_ERR_INVALID_ARGUMENT, "a face has no material (usemtl missing or unknown)");
		tri.clear();
		triangulate(ps.pos, f.v, tri);
		per[f.material].insert(per[f.material].end(), tri.begin(), tri.end());
	}
	std::vector<size_t> order;
	for (size_t m = 0; m < per.size(); ++m)
		if (!per[m].empty())
			order.push_back(m);
	if (order.empty())
		return fail(SVO_ERR_INVALID_ARGUMENT, "empty mesh"); // src/Scene.cpp:128-131
	// largest group first (src/Scene.cpp:118-126), equal sizes in material order
	for (size_t i = 1; i < order.size(); ++i)
		for (size_t j = i; j > 0 && per[order[j]].size() > per[order[j - 1]].size(); --j)
			std::swap(order[j], order[j - 1]);
	uint64_t n = 0;
	for (size_t m : order)
		n += per[m].size() / 3;
	float *X = (float *)std::malloc(sizeof(float) * 9 * n);
	uint32_t *C = (uint32_t *)std::malloc(sizeof(uint32_t) * (n ? n : 1));
	if (!X || !C) {
		std::free(X);
		std::free(C);
		return fail(SVO_ERR_OUT_OF_MEMORY, "out of host memory");
	}
	uint64_t t = 0;
	for (size_t m : order) {
		const uint32_t colour =
		    unorm8(ps.kd[3 * m]) | (unorm8(ps.kd[3 * m + 1]) << 8) | (unorm8(ps.kd[3 * m + 2]) << 16);
		const std::vector<int> &idx = per[m];
		for (size_t k = 0; k < idx.size(); k += 3, ++t) {
			for (int c = 0; c < 3; ++c)
				for (int a = 0; a < 3; ++a)
					X[9 * t + 3 * c + a] = ps.pos[3 * idx[k + c] + a];
			C[t] = colour;
		}
	}
	*xyz = X;
	*rgb = C;
	*T = n;
	return SVO_OK;
}
} // namespace

extern "C" int svo_obj_load(const char *path, float **xyz, uint32_t **rgb, uint64_t *T, char *msg, uint64_t msgcap) {
	try {
		return obj_load(path, xyz, rgb, T, msg, msgcap);
	} catch (const std::bad_alloc &) { // a file too large for host memory: no exception crosses the C ABI
		if (msg && msgcap)
			std::snprintf(msg, msgcap, "out of host memory");
		return SVO_ERR_OUT_OF_MEMORY;
	}
}

extern "C" void svo_obj_free(float *xyz, uint32_t *rgb) {
	std::free(xyz);
	std::free(rgb);
}
