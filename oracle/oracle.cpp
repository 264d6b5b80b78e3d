/*
 * CPU oracle for the sparse-voxel-octree hot path -- TEST INFRASTRUCTURE ONLY.
 * The checker the HIP path is compared against; never linked into the product.
 * Floating-point rules and canonical orders are listed in oracle.h.
 *
 * Build: g++ -O2 -ffp-contract=off -fno-fast-math -std=c++17 -shared -fPIC (oracle/build_oracle.py).
 *
 * The ray-marching part restates the ESVO traversal of shader/octree.glsl:43-629, which carries
 * this notice:
 *
 *  Copyright (c) 2009-2011, NVIDIA Corporation
 *  All rights reserved.
 *
 *  Redistribution and use in source and binary forms, with or without
 *  modification, are permitted provided that the following conditions are met:
 *      * Redistributions of source code must retain the above copyright
 *        notice, this list of conditions and the following disclaimer.
 *      * Redistributions in binary form must reproduce the above copyright
 *        notice, this list of conditions and the following disclaimer in the
 *        documentation and/or other materials provided with the distribution.
 *      * Neither the name of NVIDIA Corporation nor the
 *        names of its contributors may be used to endorse or promote products
 *        derived from this software without specific prior written permission.
 *
 *  THIS SOFTWARE IS PROVIDED BY THE COPYRIGHT HOLDERS AND CONTRIBUTORS "AS IS"
 *  AND ANY EXPRESS OR IMPLIED WARRANTIES, INCLUDING, BUT NOT LIMITED TO, THE
 *  IMPLIED WARRANTIES OF MERCHANTABILITY AND FITNESS FOR A PARTICULAR PURPOSE
 *  ARE DISCLAIMED. IN NO EVENT SHALL <COPYRIGHT HOLDER> BE LIABLE FOR ANY
 *  DIRECT, INDIRECT, INCIDENTAL, SPECIAL, EXEMPLARY, OR CONSEQUENTIAL DAMAGES
 *  (INCLUDING, BUT NOT LIMITED TO, PROCUREMENT OF SUBSTITUTE GOODS OR SERVICES;
 *  LOSS OF USE, DATA, OR PROFITS; OR BUSINESS INTERRUPTION) HOWEVER CAUSED AND
 *  ON ANY THEORY OF LIABILITY, WHETHER IN CONTRACT, STRICT LIABILITY, OR TORT
 *  (INCLUDING NEGLIGENCE OR OTHERWISE) ARISING IN ANY WAY OUT OF THE USE OF
 *  THIS SOFTWARE, EVEN IF ADVISED OF THE POSSIBILITY OF SUCH DAMAGE.
 */
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------------------------
// GLSL semantics under the fixed FP rules.
// ---------------------------------------------------------------------------------------------
inline float gmin(float x, float y) { return y < x ? y : x; }
inline float gmax(float x, float y) { return x < y ? y : x; }
inline uint32_t umin(uint32_t a, uint32_t b) { return b < a ? b : a; }
inline uint32_t umax(uint32_t a, uint32_t b) { return a < b ? b : a; }
inline uint32_t f2u(float f) { // uint(float), saturating; NaN and negatives -> 0
	if (!(f > 0.0f))
		return 0u;
	if (f >= 4294967296.0f)
		return 0xFFFFFFFFu;
	return (uint32_t)f;
}
inline uint32_t fbits(float f) {
	uint32_t u;
	std::memcpy(&u, &f, 4);
	return u;
}
inline float bitsf(uint32_t u) {
	float f;
	std::memcpy(&f, &u, 4);
	return f;
}
inline uint32_t find_msb(uint32_t x) { return x ? 31u - (uint32_t)__builtin_clz(x) : 0xFFFFFFFFu; }

struct V3 {
	float x, y, z;
};
inline V3 cross(V3 a, V3 b) { // GLSL cross
	return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 normalize(V3 v) {
	float s = 1.0f / std::sqrt((v.x * v.x + v.y * v.y) + v.z * v.z);
	return {v.x * s, v.y * s, v.z * s};
}

template <class F> void parallel_for(uint64_t n, int nthreads, F &&fn) {
	if (nthreads <= 1 || n < 1024) {
		fn(0, n);
		return;
	}
	std::vector<std::thread> th;
	uint64_t chunk = (n + nthreads - 1) / nthreads;
	for (int t = 0; t < nthreads; ++t) {
		uint64_t b = chunk * t, e = std::min(n, b + chunk);
		if (b >= e)
			break;
		th.emplace_back([&, b, e] { fn(b, e); });
	}
	for (auto &t : th)
		t.join();
}

std::atomic<uint64_t> g_uninit_reads{0};

// ---------------------------------------------------------------------------------------------
// Voxelizer: shader/voxelizer.geom:15-54 (projection, AABB, depth range), shader/voxelizer.frag:17-44
// (voxel position, clamps, packing); coverage = conservative rule on 1/256-pixel snapped vertices.
// ---------------------------------------------------------------------------------------------
struct Tri {
	bool degenerate;
	uint32_t axis;
	uint32_t aabb[4];  // gAABB
	uint32_t dr[2];    // gDepthRange
	float px[3], py[3], pz[3];
	int64_t X[3], Y[3]; // px, py * 256 rounded to nearest even
};

inline V3 project(V3 v, uint32_t axis) { // voxelizer.geom:15-19
	V3 r = axis == 0 ? V3{v.y, v.z, v.x} : (axis == 1 ? V3{v.z, v.x, v.y} : v);
	r.z = (r.z + 1.0f) * 0.5f;
	return r;
}

void tri_setup(const float *t, uint32_t res, Tri &s) {
	V3 p0{t[0], t[1], t[2]}, p1{t[3], t[4], t[5]}, p2{t[6], t[7], t[8]};
	V3 e1{p1.x - p0.x, p1.y - p0.y, p1.z - p0.z}, e2{p2.x - p0.x, p2.y - p0.y, p2.z - p0.z};
	V3 n = cross(e1, e2); // voxelizer.geom:29
	V3 w{std::fabs(n.x), std::fabs(n.y), std::fabs(n.z)};
	s.axis = (w.x > w.y && w.x > w.z) ? 0u : (w.y > w.z ? 1u : 2u); // voxelizer.geom:30-33
	s.degenerate = (n.x == 0.0f && n.y == 0.0f && n.z == 0.0f);
	V3 q[3] = {project(p0, s.axis), project(p1, s.axis), project(p2, s.axis)};
	float fres = (float)res;
	// voxelizer.geom:40-43
	float mnx = gmin(q[0].x, gmin(q[1].x, q[2].x)), mny = gmin(q[0].y, gmin(q[1].y, q[2].y));
	float mxx = gmax(q[0].x, gmax(q[1].x, q[2].x)), mxy = gmax(q[0].y, gmax(q[1].y, q[2].y));
	s.aabb[0] = f2u((mnx + 1.0f) * 0.5f * fres);
	s.aabb[1] = f2u((mny + 1.0f) * 0.5f * fres);
	s.aabb[2] = f2u((mxx + 1.0f) * 0.5f * fres);
	s.aabb[3] = f2u((mxy + 1.0f) * 0.5f * fres);
	s.dr[0] = f2u(gmin(q[0].z, gmin(q[1].z, q[2].z)) * fres);
	s.dr[1] = f2u(gmax(q[0].z, gmax(q[1].z, q[2].z)) * fres);
	for (int k = 0; k < 3; ++k) {
		s.px[k] = (q[k].x + 1.0f) * 0.5f * fres;
		s.py[k] = (q[k].y + 1.0f) * 0.5f * fres;
		s.pz[k] = q[k].z;
		s.X[k] = (int64_t)std::nearbyint(s.px[k] * 256.0f);
		s.Y[k] = (int64_t)std::nearbyint(s.py[k] * 256.0f);
	}
}

// Does the closed (snapped) triangle meet the half-open pixel square [i,i+1) x [j,j+1)?
// Separating-axis test over x, y and the three edge normals, exact in int64; an end of the
// square's projection is open when it needs x = i+1 or y = j+1.
bool covers(const Tri &s, int64_t i, int64_t j) {
	const int64_t x0 = i * 256, x1 = x0 + 256, y0 = j * 256, y1 = y0 + 256;
	int64_t xmn = std::min(s.X[0], std::min(s.X[1], s.X[2])), xmx = std::max(s.X[0], std::max(s.X[1], s.X[2]));
	int64_t ymn = std::min(s.Y[0], std::min(s.Y[1], s.Y[2])), ymx = std::max(s.Y[0], std::max(s.Y[1], s.Y[2]));
	if (xmx < x0 || xmn >= x1 || ymx < y0 || ymn >= y1)
		return false;
	for (int k = 0; k < 3; ++k) {
		int a = k, b = (k + 1) % 3, c = (k + 2) % 3;
		int64_t A = s.Y[b] - s.Y[a], B = s.X[a] - s.X[b];
		int64_t ce = A * s.X[a] + B * s.Y[a], cc = A * s.X[c] + B * s.Y[c];
		int64_t tlo = std::min(ce, cc), thi = std::max(ce, cc);
		int64_t slo = A * (A >= 0 ? x0 : x1) + B * (B >= 0 ? y0 : y1);
		int64_t shi = A * (A >= 0 ? x1 : x0) + B * (B >= 0 ? y1 : y0);
		bool lo_att = A >= 0 && B >= 0, hi_att = A <= 0 && B <= 0;
		if (thi < slo || (thi == slo && !lo_att))
			return false;
		if (tlo > shi || (tlo == shi && !hi_att))
			return false;
	}
	return true;
}

// Depth at the pixel centre: plane of the projected triangle (pixel units, z in [0,1]) through
// vertex 0 with gradients dz/dx = -nx/nz, dz/dy = -ny/nz (both 0 when the projection is degenerate).
float depth_at(const Tri &s, uint32_t i, uint32_t j) {
	float ax = s.px[1] - s.px[0], ay = s.py[1] - s.py[0], az = s.pz[1] - s.pz[0];
	float bx = s.px[2] - s.px[0], by = s.py[2] - s.py[0], bz = s.pz[2] - s.pz[0];
	float nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
	float dzdx = 0.0f, dzdy = 0.0f;
	if (nz != 0.0f) {
		dzdx = -nx / nz;
		dzdy = -ny / nz;
	}
	float cx = (float)i + 0.5f, cy = (float)j + 0.5f;
	return s.pz[0] + (dzdx * (cx - s.px[0]) + dzdy * (cy - s.py[0]));
}

template <class Emit> uint64_t raster(const Tri &s, uint32_t res, uint32_t colour, Emit &&emit) {
	if (s.degenerate)
		return 0;
	uint32_t ilo = s.aabb[0], ihi = umin(s.aabb[2], res - 1), jlo = s.aabb[1], jhi = umin(s.aabb[3], res - 1);
	uint64_t cnt = 0;
	for (uint32_t j = jlo; j <= jhi && j < res; ++j)
		for (uint32_t i = ilo; i <= ihi && i < res; ++i) {
			if (!covers(s, i, j))
				continue;
			// voxelizer.frag:17-25
			float vz = depth_at(s, i, j) * (float)res;
			uint32_t uz = umin(f2u(vz), res - 1);
			uz = umin(umax(uz, s.dr[0]), s.dr[1]);
			uint32_t vx, vy, vzz;
			if (s.axis == 0)
				vx = uz, vy = i, vzz = j;
			else if (s.axis == 1)
				vx = j, vy = uz, vzz = i;
			else
				vx = i, vy = j, vzz = uz;
			// voxelizer.frag:38-42
			uint32_t fx = vx | (vy << 12u) | ((vzz & 0xffu) << 24u);
			uint32_t fy = ((vzz >> 8u) << 28u) | (colour & 0x00ffffffu);
			emit(fx, fy);
			++cnt;
		}
	return cnt;
}

// ---------------------------------------------------------------------------------------------
// ESVO traversal (shader/octree.glsl).
// ---------------------------------------------------------------------------------------------
constexpr uint32_t kStackSize = 23;        // octree.glsl:43
const float kEps = 3.552713678800501e-15f; // octree.glsl:44, 2^-48

struct Octree {
	const uint32_t *w;
	uint64_t n;
	uint32_t load(uint64_t i, uint32_t &nfetch) const {
		++nfetch;
		return i < n ? w[i] : 0u;
	}
};

struct Stack {
	uint32_t v[kStackSize];
	bool written[kStackSize];
	Stack() {
		std::memset(v, 0, sizeof(v));
		std::memset(written, 0, sizeof(written));
	}
	void put(uint32_t s, uint32_t p) {
		v[s] = p;
		written[s] = true;
	}
	uint32_t get(uint32_t s) {
		if (!written[s])
			g_uninit_reads.fetch_add(1, std::memory_order_relaxed);
		return v[s];
	}
};

struct MarchState {
	V3 t_coef, t_bias;
	uint32_t oct_mask;
	float t_min, t_max, h;
	uint32_t parent, cur, idx, scale;
	V3 pos;
	float scale_exp2;
};

// Ray set-up shared by all variants (octree.glsl:48-83 / 182-217 / 343-377 / 505-539).
void march_init(V3 o, V3 &d, bool ge_eps, MarchState &m) {
	if (ge_eps) {
		d.x = std::fabs(d.x) >= kEps ? d.x : (d.x >= 0 ? kEps : -kEps);
		d.y = std::fabs(d.y) >= kEps ? d.y : (d.y >= 0 ? kEps : -kEps);
		d.z = std::fabs(d.z) >= kEps ? d.z : (d.z >= 0 ? kEps : -kEps);
	} else {
		d.x = std::fabs(d.x) > kEps ? d.x : (d.x >= 0 ? kEps : -kEps);
		d.y = std::fabs(d.y) > kEps ? d.y : (d.y >= 0 ? kEps : -kEps);
		d.z = std::fabs(d.z) > kEps ? d.z : (d.z >= 0 ? kEps : -kEps);
	}
	m.t_coef = {1.0f / -std::fabs(d.x), 1.0f / -std::fabs(d.y), 1.0f / -std::fabs(d.z)};
	m.t_bias = {m.t_coef.x * o.x, m.t_coef.y * o.y, m.t_coef.z * o.z};
	m.oct_mask = 0u;
	if (d.x > 0.0f)
		m.oct_mask ^= 1u, m.t_bias.x = 3.0f * m.t_coef.x - m.t_bias.x;
	if (d.y > 0.0f)
		m.oct_mask ^= 2u, m.t_bias.y = 3.0f * m.t_coef.y - m.t_bias.y;
	if (d.z > 0.0f)
		m.oct_mask ^= 4u, m.t_bias.z = 3.0f * m.t_coef.z - m.t_bias.z;
	m.t_min = gmax(gmax(2.0f * m.t_coef.x - m.t_bias.x, 2.0f * m.t_coef.y - m.t_bias.y),
	               2.0f * m.t_coef.z - m.t_bias.z);
	m.t_max = gmin(gmin(m.t_coef.x - m.t_bias.x, m.t_coef.y - m.t_bias.y), m.t_coef.z - m.t_bias.z);
	m.t_min = gmax(m.t_min, 0.0f);
	m.h = m.t_max;
	m.parent = 0u;
	m.cur = 0u;
	m.pos = {1.0f, 1.0f, 1.0f};
	m.idx = 0u;
	if (1.5f * m.t_coef.x - m.t_bias.x > m.t_min)
		m.idx ^= 1u, m.pos.x = 1.5f;
	if (1.5f * m.t_coef.y - m.t_bias.y > m.t_min)
		m.idx ^= 2u, m.pos.y = 1.5f;
	if (1.5f * m.t_coef.z - m.t_bias.z > m.t_min)
		m.idx ^= 4u, m.pos.z = 1.5f;
	m.scale = kStackSize - 1;
	m.scale_exp2 = 0.5f;
}

// One loop body of the traversal.  Returns 1 to continue, 0 to break (leaf found or stop).
// coarse: Octree_RayMarchCoarse extra stop (octree.glsl:94-96).
enum Step { kContinue, kBreak };
inline Step march_step(const Octree &oc, MarchState &m, Stack &stk, uint32_t &nfetch, bool coarse, float orig_sz,
                       float dir_sz) {
	if (m.cur == 0u)
		m.cur = oc.load((uint64_t)m.parent + (m.idx ^ m.oct_mask), nfetch);
	V3 t_corner{m.pos.x * m.t_coef.x - m.t_bias.x, m.pos.y * m.t_coef.y - m.t_bias.y,
	            m.pos.z * m.t_coef.z - m.t_bias.z};
	float tc_max = gmin(gmin(t_corner.x, t_corner.y), t_corner.z);

	if ((m.cur & 0x80000000u) != 0 && m.t_min <= m.t_max) {
		if (coarse && orig_sz + tc_max * dir_sz >= m.scale_exp2)
			return kBreak;
		// INTERSECT
		float half_scale_exp2 = m.scale_exp2 * 0.5f;
		V3 t_center{half_scale_exp2 * m.t_coef.x + t_corner.x, half_scale_exp2 * m.t_coef.y + t_corner.y,
		            half_scale_exp2 * m.t_coef.z + t_corner.z};
		if ((m.cur & 0x40000000u) != 0)
			return kBreak;
		// PUSH
		if (tc_max < m.h)
			stk.put(m.scale, m.parent);
		m.h = tc_max;
		m.parent = m.cur & 0x3fffffffu;
		m.idx = 0u;
		--m.scale;
		m.scale_exp2 = half_scale_exp2;
		if (t_center.x > m.t_min)
			m.idx ^= 1u, m.pos.x += m.scale_exp2;
		if (t_center.y > m.t_min)
			m.idx ^= 2u, m.pos.y += m.scale_exp2;
		if (t_center.z > m.t_min)
			m.idx ^= 4u, m.pos.z += m.scale_exp2;
		m.cur = 0;
		return kContinue;
	}
	// ADVANCE
	uint32_t step_mask = 0u;
	if (t_corner.x <= tc_max)
		step_mask ^= 1u, m.pos.x -= m.scale_exp2;
	if (t_corner.y <= tc_max)
		step_mask ^= 2u, m.pos.y -= m.scale_exp2;
	if (t_corner.z <= tc_max)
		step_mask ^= 4u, m.pos.z -= m.scale_exp2;
	m.t_min = tc_max;
	m.idx ^= step_mask;
	if ((m.idx & step_mask) != 0) {
		// POP
		uint32_t differing_bits = 0;
		if ((step_mask & 1u) != 0)
			differing_bits |= fbits(m.pos.x) ^ fbits(m.pos.x + m.scale_exp2);
		if ((step_mask & 2u) != 0)
			differing_bits |= fbits(m.pos.y) ^ fbits(m.pos.y + m.scale_exp2);
		if ((step_mask & 4u) != 0)
			differing_bits |= fbits(m.pos.z) ^ fbits(m.pos.z + m.scale_exp2);
		m.scale = find_msb(differing_bits);
		if (m.scale >= kStackSize)
			return kBreak;
		m.scale_exp2 = bitsf((m.scale - kStackSize + 127u) << 23u);
		m.parent = stk.get(m.scale);
		uint32_t shx = fbits(m.pos.x) >> m.scale, shy = fbits(m.pos.y) >> m.scale, shz = fbits(m.pos.z) >> m.scale;
		m.pos.x = bitsf(shx << m.scale);
		m.pos.y = bitsf(shy << m.scale);
		m.pos.z = bitsf(shz << m.scale);
		m.idx = (shx & 1u) | ((shy & 1u) << 1u) | ((shz & 1u) << 2u);
		m.h = 0.0f;
		m.cur = 0;
	}
	return kContinue;
}

// Octree_RayMarchLeaf, 6-arg (variant 0, octree.glsl:179-340) or 5-arg (variant 1, :342-502);
// Octree_RayMarchOcclude (variant 2, :504-629) has the 5-arg loop and only returns the hit flag.
orc_hit march_leaf(const Octree &oc, V3 o, V3 d, int variant) {
	MarchState m;
	Stack stk;
	uint32_t nfetch = 0, iter = 0;
	march_init(o, d, variant != 0, m);
	while (m.scale < kStackSize) {
		if (variant == 0)
			++iter;
		if (march_step(oc, m, stk, nfetch, false, 0.0f, 0.0f) == kBreak)
			break;
	}
	bool hit = m.scale < kStackSize && m.t_min <= m.t_max;
	orc_hit out;
	std::memset(&out, 0, sizeof(out));
	if (variant == 2) {
		out.vxy = 0xFFFFFFFFu;
		out.vz_flags = hit ? 0x80000000u : 0u;
		out.nfetch = nfetch;
		return out;
	}
	// octree.glsl:307-317
	V3 t_corner{m.t_coef.x * (m.pos.x + m.scale_exp2) - m.t_bias.x,
	            m.t_coef.y * (m.pos.y + m.scale_exp2) - m.t_bias.y,
	            m.t_coef.z * (m.pos.z + m.scale_exp2) - m.t_bias.z};
	int axis = (t_corner.x > t_corner.y && t_corner.x > t_corner.z) ? 0 : (t_corner.y > t_corner.z ? 1 : 2);
	V3 norm{0.0f, 0.0f, 0.0f};
	(axis == 0 ? norm.x : axis == 1 ? norm.y : norm.z) = -1.0f;
	if ((m.oct_mask & 1u) == 0u)
		norm.x = -norm.x;
	if ((m.oct_mask & 2u) == 0u)
		norm.y = -norm.y;
	if ((m.oct_mask & 4u) == 0u)
		norm.z = -norm.z;
	// octree.glsl:320-325
	V3 pos = m.pos;
	if ((m.oct_mask & 1u) != 0u)
		pos.x = 3.0f - m.scale_exp2 - pos.x;
	if ((m.oct_mask & 2u) != 0u)
		pos.y = 3.0f - m.scale_exp2 - pos.y;
	if ((m.oct_mask & 4u) != 0u)
		pos.z = 3.0f - m.scale_exp2 - pos.z;
	// octree.glsl:328-334, clamp(x, lo, hi) = min(max(x, lo), hi)
	V3 op{o.x + m.t_min * d.x, o.y + m.t_min * d.y, o.z + m.t_min * d.z};
	op.x = gmin(gmax(op.x, pos.x), pos.x + m.scale_exp2);
	op.y = gmin(gmax(op.y, pos.y), pos.y + m.scale_exp2);
	op.z = gmin(gmax(op.z, pos.z), pos.z + m.scale_exp2);
	if (norm.x != 0)
		op.x = norm.x > 0 ? pos.x + m.scale_exp2 + kEps * 2 : pos.x - kEps;
	if (norm.y != 0)
		op.y = norm.y > 0 ? pos.y + m.scale_exp2 + kEps * 2 : pos.y - kEps;
	if (norm.z != 0)
		op.z = norm.z > 0 ? pos.z + m.scale_exp2 + kEps * 2 : pos.z - kEps;

	float nsign = axis == 0 ? norm.x : axis == 1 ? norm.y : norm.z;
	uint32_t code = (uint32_t)axis * 2u + (nsign > 0 ? 1u : 0u);
	out.pos[0] = op.x;
	out.pos[1] = op.y;
	out.pos[2] = op.z;
	out.iter = iter;
	out.leaf = m.cur;
	out.nfetch = nfetch;
	if (hit) {
		float inv = bitsf((127u + (kStackSize - m.scale)) << 23u); // 2^(23-scale) = 1 / scale_exp2
		uint32_t vx = (uint32_t)((pos.x - 1.0f) * inv), vy = (uint32_t)((pos.y - 1.0f) * inv),
		         vz = (uint32_t)((pos.z - 1.0f) * inv);
		out.vxy = vx | (vy << 16);
		out.vz_flags = vz | ((kStackSize - m.scale) << 16) | (code << 24) | 0x80000000u;
	} else {
		out.vxy = 0xFFFFFFFFu;
		out.vz_flags = 0xFFFFu | (code << 24);
	}
	return out;
}

// Octree_RayMarchCoarse (octree.glsl:47-177).
bool march_coarse(const Octree &oc, V3 o, V3 d, float orig_sz, float dir_sz, float &t, float &size,
                  uint32_t &nfetch) {
	MarchState m;
	Stack stk;
	march_init(o, d, false, m);
	for (;;) {
		if (march_step(oc, m, stk, nfetch, true, orig_sz, dir_sz) == kBreak)
			break;
	}
	t = m.t_min;
	size = m.scale_exp2;
	return m.scale < kStackSize && m.t_min <= m.t_max;
}

V3 cam_vec(const float *cam, int k) { return {cam[4 * k + 0], cam[4 * k + 1], cam[4 * k + 2]}; }

// Camera_GenRay (shader/camera.glsl:11-14)
V3 gen_ray(const float *cam, float cx, float cy) {
	cx = cx * 2.0f - 1.0f;
	cy = cy * 2.0f - 1.0f;
	V3 look = cam_vec(cam, 1), side = cam_vec(cam, 2), up = cam_vec(cam, 3);
	V3 v{look.x - side.x * cx - up.x * cy, look.y - side.y * cx - up.y * cy, look.z - side.z * cx - up.z * cy};
	return normalize(v);
}

uint32_t beam_dim(uint32_t x) { return (x + 7u) / 8u + 1u; } // src/OctreeTracer.cpp:6

// ---------------------------------------------------------------------------------------------
// Bounce sampling.
// ---------------------------------------------------------------------------------------------
// The reference's path-tracer sampler (shader/path_tracer.comp:42,72,113-117, shader/sobol.glsl,
// shader/sobol.comp, src/Sobol.cpp:52-62, src/PathTracer.cpp:24-28).
//   * Sobol state: `dims` 32-bit words, all zero after Sobol::Reset; each CmdNext (after each
//     frame, src/PathTracer.cpp:214) XORs column c = findLSB(~n) of every dimension's matrix into
//     it and increments the counter n (shader/sobol.comp:78-86).  Restated here as that sequential
//     advance (the device computes the same state from the Gray code of n).
//   * Sobol_GetVec2(i) = vec2(x[2i], x[2i+1]) / 4294967296.0.
//   * noise = texelFetch(uNoise, pixel & 0xff).xy of a 256x256 RG8 UNORM texture: byte / 255.
//   * bounce cur's sample = fract(Sobol_GetVec2(cur) + noise); the primary ray goes through
//     pixel + Sobol_GetVec2(0).
struct Sampler {
	uint32_t x[64];
	const uint8_t *noise;
	void init(const uint32_t *mat, const uint8_t *noise_tab, uint32_t dims, uint32_t n) {
		noise = noise_tab;
		for (uint32_t d = 0; d < 64; ++d)
			x[d] = 0u;
		for (uint32_t i = 0; i < n; ++i) { // shader/sobol.comp:82-85
			uint32_t c = 0;
			while ((i >> c) & 1u) // c = findLSB(~i), the lowest zero bit of i (i < 2^32 - 1)
				++c;
			for (uint32_t d = 0; d < dims; ++d)
				x[d] ^= mat[d * 32 + c];
		}
	}
	static float to_unit(uint32_t v) { return (float)v / 4294967296.0f; }
	void vec2(uint32_t i, float &u, float &v) const { u = to_unit(x[2 * i]), v = to_unit(x[2 * i + 1]); }
	void pixel_noise(uint32_t px, uint32_t py, float &nu, float &nv) const {
		const uint8_t *t = noise + (((py & 255u) << 8) | (px & 255u)) * 2u;
		nu = (float)t[0] / 255.0f, nv = (float)t[1] / 255.0f;
	}
	void bounce(uint32_t cur, float nu, float nv, float &u, float &v) const { // fract(Sobol_GetVec2(cur) + noise)
		float a, b;
		vec2(cur, a, b);
		const float s = a + nu, t = b + nv;
		u = s - std::floor(s), v = t - std::floor(t);
	}
};

// RGBA8 UNORM / SNORM texels of the path tracer's albedo and normal images (src/PathTracer.cpp:13-
// 16): imageLoad gives byte / 255 (UNORM) or max(signed byte / 127, -1) (SNORM); imageStore rounds
// the clamped value times 255 (or 127) to nearest, halves up.
uint32_t sat_u(float f) { return f >= 1.0f ? (f < 4294967296.0f ? (uint32_t)f : 0xFFFFFFFFu) : 0u; }
uint32_t unorm8(float f) {
	f = f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f);
	return sat_u(std::floor(f * 255.0f + 0.5f));
}
uint32_t snorm8(float f) {
	f = f < -1.0f ? -1.0f : (f > 1.0f ? 1.0f : f);
	return (uint32_t)((int32_t)std::floor(f * 127.0f + 0.5f) & 0xff);
}
float unorm8_f(uint32_t c) { return (float)(c & 0xffu) / 255.0f; }
float snorm8_f(uint32_t c) {
	const float v = (float)(int8_t)(uint8_t)(c & 0xffu) / 127.0f;
	return v < -1.0f ? -1.0f : v;
}
// Running mean of a pixel (shader/path_tracer.comp:123-131): (old * spp + v) / (spp + 1); spp 0
// starts from the cleared image.
void mean_f4(float *img, uint64_t k, uint32_t spp, const float v[3]) {
	if (!img)
		return;
	float *p = img + 4 * k;
	const float n = (float)spp, n1 = (float)(spp + 1u);
	for (int c = 0; c < 3; ++c)
		p[c] = spp == 0u ? v[c] : (p[c] * n + v[c]) / n1;
	p[3] = 1.0f;
}
void mean_rgba8(uint32_t *img, uint64_t k, uint32_t spp, const float v[3], bool snorm) {
	if (!img)
		return;
	const uint32_t old = spp == 0u ? 0u : img[k];
	const float n = (float)spp, n1 = (float)(spp + 1u);
	uint32_t out = snorm ? 127u << 24 : 255u << 24; // alpha 1
	for (int c = 0; c < 3; ++c) {
		const uint32_t b = old >> (8 * c);
		const float o = snorm ? snorm8_f(b) : unorm8_f(b);
		const float m = (o * n + v[c]) / n1;
		out |= (snorm ? snorm8(m) : unorm8(m)) << (8 * c);
	}
	img[k] = out;
}


// sin/cos: quadrant reduction with a two-constant split of pi/2 and Taylor polynomials on
// [-pi/4, pi/4]; all in binary32 with a fixed operation order.
void sincos_poly(float phi, float &s, float &c) {
	const float two_over_pi = 0x1.45f306p-1f; // 2/pi
	const float pio2_hi = 0x1.921ep+0f;       // pi/2 cut to 16 significant bits: k * pio2_hi is exact
	const float pio2_lo = 0x1.b54442p-16f;    // pi/2 - pio2_hi
	float k = std::nearbyint(phi * two_over_pi);
	float r = (phi - k * pio2_hi) - k * pio2_lo;
	float r2 = r * r;
	// Taylor coefficients (-1)^n / (2n+1)! and (-1)^n / (2n)!, rounded to binary32.
	float sp = r + r * r2 * (-0x1.555556p-3f + r2 * (0x1.111112p-7f + r2 * (-0x1.a01a02p-13f + r2 * 0x1.71de3ap-19f)));
	float cp = 1.0f + r2 * (-0.5f + r2 * (0x1.555556p-5f + r2 * (-0x1.6c16c2p-10f + r2 * 0x1.a01a02p-16f)));
	int q = ((int)k) & 3;
	if (q == 0)
		s = sp, c = cp;
	else if (q == 1)
		s = cp, c = -sp;
	else if (q == 2)
		s = -sp, c = -cp;
	else
		s = -cp, c = sp;
}

} // namespace

// =============================================================================================
extern "C" {

void orc_normalize(float *xyz, uint64_t nverts) {
	// src/Scene.cpp:62-99: glm::min/max over all positions, then (v - center) * inv_extent.
	float mn[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
	float mx[3] = {-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
	for (uint64_t v = 0; v < nverts; ++v)
		for (int c = 0; c < 3; ++c) {
			float p = xyz[3 * v + c];
			mn[c] = mn[c] < p ? mn[c] : p; // glm::min(position, pmin) = (pmin < position) ? pmin : position
			mx[c] = p < mx[c] ? mx[c] : p; // glm::max(position, pmax) = (position < pmax) ? pmax : position
		}
	float ext[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
	float extent = gmax(ext[0], gmax(ext[1], ext[2])) * 0.5f;
	float inv = 1.0f / extent;
	float ctr[3] = {(mx[0] + mn[0]) * 0.5f, (mx[1] + mn[1]) * 0.5f, (mx[2] + mn[2]) * 0.5f};
	for (uint64_t v = 0; v < nverts; ++v)
		for (int c = 0; c < 3; ++c)
			xyz[3 * v + c] = (xyz[3 * v + c] - ctr[c]) * inv;
}

uint64_t orc_voxelize(const float *tri, const uint32_t *rgb, uint64_t T, uint32_t level, uint32_t *frags,
                      uint64_t cap, int nthreads) {
	if (level < 1 || level > 12)
		return 0;
	const uint32_t res = 1u << level;
	std::vector<uint64_t> cnt(T + 1, 0);
	parallel_for(T, nthreads, [&](uint64_t b, uint64_t e) {
		for (uint64_t t = b; t < e; ++t) {
			Tri s;
			tri_setup(tri + 9 * t, res, s);
			cnt[t] = raster(s, res, 0u, [](uint32_t, uint32_t) {});
		}
	});
	uint64_t total = 0;
	for (uint64_t t = 0; t < T; ++t) {
		uint64_t c = cnt[t];
		cnt[t] = total;
		total += c;
	}
	if (!frags)
		return total;
	parallel_for(T, nthreads, [&](uint64_t b, uint64_t e) {
		for (uint64_t t = b; t < e; ++t) {
			Tri s;
			tri_setup(tri + 9 * t, res, s);
			uint64_t k = cnt[t];
			raster(s, res, rgb[t], [&](uint32_t fx, uint32_t fy) {
				if (k < cap) {
					frags[2 * k] = fx;
					frags[2 * k + 1] = fy;
				}
				++k;
			});
		}
	});
	return total;
}

uint64_t orc_build(const uint32_t *frags, uint64_t F, uint32_t level, uint32_t *words, uint64_t cap) {
	if (level < 1 || level > 12)
		return 0;
	const uint32_t res = 1u << level;
	// Canonical fold order: ascending (fragment.x, fragment.y).
	std::vector<uint64_t> key(F);
	for (uint64_t f = 0; f < F; ++f)
		key[f] = ((uint64_t)frags[2 * f] << 32) | frags[2 * f + 1];
	std::sort(key.begin(), key.end());

	std::vector<uint32_t> oct(8, 0u);
	uint64_t begin = 0, num = 8, counter = 0; // src/OctreeBuilder.cpp:27-32
	for (uint32_t i = 1; i <= level; ++i) {
		// octree_init_node.comp:7-11
		if (oct.size() < begin + num)
			oct.resize(begin + num);
		std::fill(oct.begin() + begin, oct.begin() + begin + num, 0u);
		// octree_tag_node.comp:32-60
		for (uint64_t f = 0; f < F; ++f) {
			uint32_t fx = (uint32_t)(key[f] >> 32), fy = (uint32_t)key[f];
			uint32_t p[3] = {fx & 0xfffu, (fx >> 12u) & 0xfffu, (fx >> 24u) | ((fy >> 28u) << 8u)};
			// TraverseOctree (octree_tag_node.comp:18-30)
			uint32_t level_dim = res, idx = 0u, cur = 0u;
			do {
				level_dim >>= 1;
				uint32_t cx = p[0] >= level_dim, cy = p[1] >= level_dim, cz = p[2] >= level_dim;
				idx = cur | cx | (cy << 1u) | (cz << 2u);
				cur = oct[idx] & 0x3fffffffu;
				p[0] -= cx * level_dim;
				p[1] -= cy * level_dim;
				p[2] -= cz * level_dim;
			} while (cur != 0 && level_dim > 1u);
			bool is_leaf = level_dim == 1u;
			if (is_leaf) {
				// octree_tag_node.comp:43-57: the CAS loop, executed serially in canonical order.
				uint32_t prev = oct[idx];
				uint32_t c = fy & 0xffffffu;
				if (prev == 0u) {
					oct[idx] = 0xC1000000u | c;
				} else {
					uint32_t n = (prev >> 24u) & 0x3fu;
					uint32_t r = (prev & 0xffu) * n + (c & 0xffu);
					uint32_t g = ((prev >> 8u) & 0xffu) * n + ((c >> 8u) & 0xffu);
					uint32_t b = ((prev >> 16u) & 0xffu) * n + ((c >> 16u) & 0xffu);
					uint32_t w = n + 1u;
					r /= w, g /= w, b /= w;
					oct[idx] = (umin(w, 0x3fu) << 24u) | (r & 0xffu) | ((g & 0xffu) << 8u) | ((b & 0xffu) << 16u) |
					           0xC0000000u;
				}
			} else
				oct[idx] = 0x80000000u;
		}
		if (i != level) {
			// octree_alloc_node.comp:9-23 with a serial allocator in word-index order
			for (uint64_t k = begin; k < begin + num; ++k)
				if (oct[k] & 0x80000000u) {
					oct[k] = (uint32_t)(((counter + 1u) << 3u) | 0x80000000u);
					++counter;
				}
			// octree_modify_arg.comp:9-13
			begin += num;
			num = (counter << 3u) - begin + 8u;
		}
	}
	uint64_t W = (counter + 1u) * 8u; // src/OctreeBuilder.cpp:212-214
	if (words)
		std::memcpy(words, oct.data(), sizeof(uint32_t) * std::min(W, cap));
	return W;
}

uint64_t orc_canonicalize(const uint32_t *words, uint64_t W, uint32_t level, uint32_t *out, uint64_t cap) {
	if (W < 8)
		return 0;
	std::vector<uint32_t> res;
	std::vector<uint64_t> src_of{0};
	std::vector<uint32_t> depth_of{1};
	res.reserve(W);
	for (uint64_t b = 0; b < src_of.size(); ++b) {
		uint64_t s = src_of[b];
		if (s + 8 > W || depth_of[b] > level)
			return 0;
		for (int k = 0; k < 8; ++k) {
			uint32_t w = words[s + k];
			if ((w & 0x80000000u) && !(w & 0x40000000u)) {
				if (src_of.size() * 8 >= W)
					return 0; // more blocks than words: a cycle
				uint64_t nb = src_of.size();
				src_of.push_back(w & 0x3fffffffu);
				depth_of.push_back(depth_of[b] + 1);
				res.push_back((uint32_t)(0x80000000u | (nb << 3)));
			} else
				res.push_back(w);
		}
	}
	if (out)
		std::memcpy(out, res.data(), sizeof(uint32_t) * std::min<uint64_t>(res.size(), cap));
	return res.size();
}

void orc_gen_ray(const float *cam, float cx, float cy, float *d_out) {
	V3 d = gen_ray(cam, cx, cy);
	d_out[0] = d.x, d_out[1] = d.y, d_out[2] = d.z;
}

void orc_beam(const uint32_t *words, uint64_t W, uint32_t level, const float *cam, uint32_t width, uint32_t height,
              float *beam, uint32_t *beam_nfetch, int nthreads) {
	Octree oc{words, W};
	uint32_t bw = beam_dim(width), bh = beam_dim(height);
	float orig_sz = std::exp2((float)(1.0f - (float)level)); // src/OctreeTracer.cpp:206, a power of two
	float fw = (float)width, fh = (float)height, bs = 8.0f;
	V3 o = cam_vec(cam, 0);
	parallel_for((uint64_t)bw * bh, nthreads, [&](uint64_t b, uint64_t e) {
		for (uint64_t k = b; k < e; ++k) {
			uint32_t bx = (uint32_t)(k % bw), by = (uint32_t)(k / bw);
			float fx = (float)bx, fy = (float)by;
			// octree_tracer_beam.frag:12-15
			V3 d = gen_ray(cam, fx / fw * bs, fy / fh * bs);
			auto tangent = [&](float ox, float oy) {
				V3 y = gen_ray(cam, (fx + ox) / fw * bs, (fy + oy) / fh * bs);
				float c = dot(d, y);
				return std::sqrt(1.0f - c * c) / std::fabs(c);
			};
			float t0 = tangent(0.5f, 0.5f), t1 = tangent(-0.5f, 0.5f), t2 = tangent(0.5f, -0.5f),
			      t3 = tangent(-0.5f, -0.5f);
			float dir_sz = 2.0f * gmax(gmax(t0, t1), gmax(t2, t3));
			float t, size;
			uint32_t nf = 0;
			bool hit = march_coarse(oc, o, d, orig_sz, dir_sz, t, size, nf);
			beam[k] = hit ? gmax(0.0f, t - size) : 1e10f;
			if (beam_nfetch)
				beam_nfetch[k] = nf;
		}
	});
}

void orc_march_frame(const uint32_t *words, uint64_t W, uint32_t level, const float *cam, uint32_t width,
                     uint32_t height, const float *beam, orc_hit *hits, int nthreads) {
	orc_march_frame_rows(words, W, level, cam, width, height, 0, height, beam, hits, nthreads);
}

void orc_march_frame_rows(const uint32_t *words, uint64_t W, uint32_t level, const float *cam, uint32_t width,
                          uint32_t height, uint32_t y0, uint32_t y1, const float *beam, orc_hit *hits, int nthreads) {
	(void)level;
	Octree oc{words, W};
	uint32_t bw = beam_dim(width);
	float fw = (float)width, fh = (float)height;
	parallel_for((uint64_t)width * (y1 - y0), nthreads, [&](uint64_t b, uint64_t e) {
		for (uint64_t kk = b; kk < e; ++kk) {
			uint64_t k = kk; // index into hits (rows y0..y1)
			uint32_t px = (uint32_t)(kk % width), py = y0 + (uint32_t)(kk / width);
			// octree_tracer.frag:23-32
			V3 o = cam_vec(cam, 0);
			V3 d = gen_ray(cam, (float)px / fw, (float)py / fh);
			if (beam) {
				uint32_t bx = px >> 3, by = py >> 3;
				float bm = gmin(gmin(beam[by * bw + bx], beam[by * bw + bx + 1]),
				                gmin(beam[(by + 1) * bw + bx], beam[(by + 1) * bw + bx + 1]));
				o = {o.x + d.x * bm, o.y + d.y * bm, o.z + d.z * bm};
			}
			hits[k] = march_leaf(oc, o, d, 0);
		}
	});
}

void orc_march_rays(const uint32_t *words, uint64_t W, uint32_t level, const float *rays, uint64_t R, int variant,
                    orc_hit *hits, int nthreads) {
	(void)level;
	Octree oc{words, W};
	parallel_for(R, nthreads, [&](uint64_t b, uint64_t e) {
		for (uint64_t k = b; k < e; ++k) {
			const float *r = rays + 6 * k;
			hits[k] = march_leaf(oc, V3{r[0], r[1], r[2]}, V3{r[3], r[4], r[5]}, variant);
		}
	});
}

void orc_sincos(float phi, float *s, float *c) { sincos_poly(phi, *s, *c); }

// Diffuse_Sample + AlignDirection (shader/diffuse.glsl:7-14, shader/util.glsl:5-9) about the axis
// normal of face code `code` (axis * 2 + positive) for the sample (u, v).
static V3 diffuse_dir(uint32_t code, float u, float v) {
	const float kPi = 3.1415926535897932384626433832795f; // shader/util.glsl:4
	V3 nrm{0.0f, 0.0f, 0.0f};
	const float sg = (code & 1u) ? 1.0f : -1.0f;
	(code >> 1 == 0 ? nrm.x : code >> 1 == 1 ? nrm.y : nrm.z) = sg;
	float rr = std::sqrt(u), phi = 2.0f * kPi * v;
	float sn, cs;
	sincos_poly(phi, sn, cs);
	V3 dl{rr * cs, rr * sn, std::sqrt(1.0f - u)};
	V3 ax = std::fabs(nrm.x) > 0.01f ? V3{0.0f, 1.0f, 0.0f} : V3{1.0f, 0.0f, 0.0f};
	V3 uu = normalize(cross(ax, nrm));
	V3 vv = cross(nrm, uu);
	return V3{dl.x * uu.x + dl.y * vv.x + dl.z * nrm.x, dl.x * uu.y + dl.y * vv.y + dl.z * nrm.y,
	          dl.x * uu.z + dl.y * vv.z + dl.z * nrm.z};
}

void orc_sobol_state(const uint32_t *mat, uint32_t dims, uint32_t n, uint32_t *x) {
	Sampler sm;
	sm.init(mat, nullptr, dims, n);
	for (uint32_t d = 0; d < dims; ++d)
		x[d] = sm.x[d];
}

void orc_bounce_rays(const orc_hit *hits, uint32_t width, uint32_t height, uint32_t spp, uint32_t cur,
                     const uint32_t *sobol_mat, const uint8_t *noise, float *rays, uint8_t *mask) {
	Sampler sm;
	sm.init(sobol_mat, noise, 2 * (cur + 1), spp);
	const uint64_t R = (uint64_t)width * height;
	for (uint64_t k = 0; k < R; ++k) {
		float *r = rays + 6 * k;
		const orc_hit &h = hits[k];
		if (!(h.vz_flags & 0x80000000u)) {
			r[0] = r[1] = r[2] = 0.0f;
			r[3] = r[4] = r[5] = -1.0f;
			if (mask)
				mask[k] = 0;
			continue;
		}
		if (mask)
			mask[k] = 1;
		const uint32_t px = (uint32_t)(k % width), py = (uint32_t)(k / width);
		float nu, nv, u, v;
		sm.pixel_noise(px, py, nu, nv);
		sm.bounce(cur, nu, nv, u, v);
		const V3 d = diffuse_dir((h.vz_flags >> 24) & 7u, u, v);
		r[0] = h.pos[0], r[1] = h.pos[1], r[2] = h.pos[2];
		r[3] = d.x, r[4] = d.y, r[5] = d.z;
	}
}

static V3 face_normal(uint32_t code) {
	V3 n{0.0f, 0.0f, 0.0f};
	const float sg = (code & 1u) ? 1.0f : -1.0f;
	(code >> 1 == 0 ? n.x : code >> 1 == 1 ? n.y : n.z) = sg;
	return n;
}
static float leaf_colour(uint32_t w, int c) { return (float)((w >> (8 * c)) & 0xffu) / 255.0f; } // unpackUnorm4x8

// Stores one sample of pixel k (shader/path_tracer.comp:123-131).
static void store_sample(uint64_t k, uint32_t spp, V3 rad, V3 alb, V3 nrm, float *color, uint32_t *albedo,
                         uint32_t *normal) {
	const float c[3] = {rad.x, rad.y, rad.z}, a[3] = {alb.x, alb.y, alb.z}, n[3] = {nrm.x, nrm.y, nrm.z};
	mean_f4(color, k, spp, c);
	mean_rgba8(albedo, k, spp, a, false);
	mean_rgba8(normal, k, spp, n, true);
}

// ConstColorPathTrace + main (shader/path_tracer.comp:32-60,111-132) for sample index spp.
void orc_path_trace(const uint32_t *words, uint64_t W, uint32_t level, const float *cam, uint32_t width,
                    uint32_t height, uint32_t bounce, uint32_t spp, const float *sky, const uint32_t *sobol_mat,
                    const uint8_t *noise, float *color, uint32_t *albedo, uint32_t *normal, int nthreads) {
	(void)level;
	Octree oc{words, W};
	Sampler sm;
	sm.init(sobol_mat, noise, (bounce + 1) * 2, spp); // src/PathTracer.cpp:184
	float jx, jy;
	sm.vec2(0, jx, jy);
	parallel_for((uint64_t)width * height, nthreads, [&](uint64_t b, uint64_t e) {
		for (uint64_t kk = b; kk < e; ++kk) {
			const uint32_t k = (uint32_t)kk, px = k % width, py = k / width;
			float nu, nv;
			sm.pixel_noise(px, py, nu, nv);
			V3 o = cam_vec(cam, 0);
			V3 d = gen_ray(cam, ((float)px + jx) / (float)width, ((float)py + jy) / (float)height);
			orc_hit h = march_leaf(oc, o, d, 1);
			V3 rad{sky[0], sky[1], sky[2]}, alb0{0.0f, 0.0f, 0.0f}, nrm0{0.0f, 0.0f, 0.0f};
			if (h.vz_flags & 0x80000000u) {
				V3 acc{leaf_colour(h.leaf, 0), leaf_colour(h.leaf, 1), leaf_colour(h.leaf, 2)};
				alb0 = acc;
				uint32_t code = (h.vz_flags >> 24) & 7u;
				nrm0 = face_normal(code);
				rad = V3{0.0f, 0.0f, 0.0f};
				for (uint32_t cur = 1; cur < bounce; ++cur) {
					float u, v;
					sm.bounce(cur, nu, nv, u, v);
					o = V3{h.pos[0], h.pos[1], h.pos[2]};
					d = diffuse_dir(code, u, v);
					h = march_leaf(oc, o, d, 1);
					if (!(h.vz_flags & 0x80000000u)) {
						rad = V3{acc.x * sky[0], acc.y * sky[1], acc.z * sky[2]};
						break;
					}
					acc = V3{acc.x * leaf_colour(h.leaf, 0), acc.y * leaf_colour(h.leaf, 1), acc.z * leaf_colour(h.leaf, 2)};
					code = (h.vz_flags >> 24) & 7u;
				}
			}
			store_sample(k, spp, rad, alb0, nrm0, color, albedo, normal);
		}
	});
}

// ---- environment-map light (EnvMapPathTrace, shader/path_tracer.comp:62-106) ----------------------
// weigh_hdr_image + generate_alias_table (src/EnvironmentMap.cpp:104-188); false without light.
int orc_envmap_tables(uint32_t w, uint32_t h, const float *rgba, float *texels, uint32_t *alias) {
	const uint64_t n = (uint64_t)w * h;
	std::vector<double> wt(n);
	double sum = 0.0;
	for (uint32_t j = 0; j < h; ++j) {
		const double sin_theta = std::sin(M_PI * ((double)j + 0.5) / (double)h);
		for (uint32_t i = 0; i < w; ++i) {
			const float *c = rgba + 4 * ((uint64_t)j * w + i);
			wt[(uint64_t)j * w + i] = (double)((0.2126f * c[0] + 0.7152f * c[1]) + 0.0722f * c[2]) * sin_theta;
			sum += wt[(uint64_t)j * w + i];
		}
	}
	if (!(sum > 0.0))
		return 0;
	const double mul = (double)n / sum;
	for (uint64_t i = 0; i < n; ++i) {
		wt[i] *= mul;
		for (int c = 0; c < 3; ++c)
			texels[4 * i + c] = rgba[4 * i + c];
		texels[4 * i + 3] = (float)wt[i];
	}
	std::vector<uint32_t> small, large; // the reference's two stacks
	for (uint64_t i = 0; i < n; ++i)
		(wt[i] < 1.0 ? small : large).push_back((uint32_t)i);
	for (uint64_t i = 0; i < 2 * n; ++i)
		alias[i] = 0u;
	while (!small.empty() && !large.empty()) {
		const uint32_t s = small.back(), l = large.back();
		small.pop_back();
		const double p = std::min(std::max(wt[s] * 4294967296.0, 0.0), 4294967295.0);
		alias[2 * s] = (uint32_t)p;
		alias[2 * s + 1] = l;
		wt[l] -= 1.0 - wt[s];
		if (wt[l] < 1.0) {
			large.pop_back();
			small.push_back(l);
		}
	}
	for (uint32_t l : large)
		alias[2 * l] = alias[2 * l + 1] = 0xFFFFFFFFu;
	for (uint32_t t : small)
		alias[2 * t] = alias[2 * t + 1] = 0xFFFFFFFFu;
	return 1;
}

// atan on [0, 1] by pi/4 + atan((t-1)/(t+1)) above tan(pi/8) and a fitted odd polynomial; the
// product's device function performs the same operations.
static float atan01(float t) {
	const bool big = t > 0.41421356f;
	const float x = big ? (t - 1.0f) / (t + 1.0f) : t, z = x * x;
	float p = -0x1.0a311cp-4f;
	p = p * z + 0x1.b8b1c6p-4f;
	p = p * z + -0x1.24286cp-3f;
	p = p * z + 0x1.999774p-3f;
	p = p * z + -0x1.555554p-2f;
	const float r = x + (x * z) * p;
	return big ? 0x1.921fb6p-1f + r : r;
}
static float atan2_poly(float y, float x) {
	const float ax = std::fabs(x), ay = std::fabs(y), mx = gmax(ax, ay), mn = gmin(ax, ay);
	float r = atan01(mx > 0.0f ? mn / mx : 0.0f);
	r = ay > ax ? 0x1.921fb6p+0f - r : r;
	r = x < 0.0f ? 0x1.921fb6p+1f - r : r;
	return y < 0.0f ? -r : r;
}
static float acos_poly(float x) { return atan2_poly(std::sqrt(gmax(1.0f - x * x, 0.0f)), x); }
float orc_atan2(float y, float x) { return atan2_poly(y, x); }
float orc_acos(float x) { return acos_poly(x); }

namespace {
struct Env {
	const float *tex; // 4 per texel
	const uint32_t *alias;
	uint32_t w, h;
	float rot, mult;
	void texture(float u, float v, float out[4]) const { // bilinear, repeat
		const float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
		const float fx = std::floor(x), fy = std::floor(y), a = x - fx, b = y - fy;
		auto wrap = [](float f, uint32_t n) -> uint32_t {
			if (!(f > -1e9f && f < 1e9f))
				return 0u;
			const int64_t i = (int64_t)f % (int64_t)n;
			return (uint32_t)(i < 0 ? i + n : i);
		};
		const uint32_t i0 = wrap(fx, w), i1 = wrap(fx + 1.0f, w), j0 = wrap(fy, h), j1 = wrap(fy + 1.0f, h);
		const float *t00 = tex + 4 * ((uint64_t)j0 * w + i0), *t10 = tex + 4 * ((uint64_t)j0 * w + i1);
		const float *t01 = tex + 4 * ((uint64_t)j1 * w + i0), *t11 = tex + 4 * ((uint64_t)j1 * w + i1);
		const float a1 = 1.0f - a, b1 = 1.0f - b;
		for (int c = 0; c < 4; ++c)
			out[c] = (t00[c] * a1 + t10[c] * a) * b1 + (t01[c] * a1 + t11[c] * a) * b;
	}
	void lookup(V3 d, float out[4]) const {
		const float kPi = 0x1.921fb6p+1f;
		texture(((atan2_poly(d.x, d.z) + rot) / kPi) * 0.5f, acos_poly(d.y) / kPi, out);
	}
	float pdf(V3 d) const {
		const float kTwoPiPi = (2.0f * 0x1.921fb6p+1f) * 0x1.921fb6p+1f;
		float t[4];
		lookup(d, t);
		return t[3] / (kTwoPiPi * std::sqrt(1.0f - d.y * d.y));
	}
	V3 sample(float su, float sv, V3 &light, float &p) const {
		const float kPi = 0x1.921fb6p+1f, kTwoPiPi = (2.0f * kPi) * kPi;
		const uint32_t sz = w * h;
		uint32_t idx = (uint32_t)(su * (float)sz);
		idx = idx < sz - 1u ? idx : sz - 1u;
		const uint32_t prob = (uint32_t)(sv * 4294967296.0f);
		const uint32_t q = prob > alias[2 * idx] ? alias[2 * idx + 1] : idx;
		const float ux = ((float)(q % w) + 0.5f) / (float)w, uy = ((float)(q / w) + 0.5f) / (float)h;
		const float phi = (2.0f * kPi) * ux - rot, theta = kPi * uy;
		float st, ct, sp, cp;
		sincos_poly(theta, st, ct);
		sincos_poly(phi, sp, cp);
		const float *t = tex + 4 * (uint64_t)q;
		light = V3{t[0], t[1], t[2]};
		p = t[3] / (kTwoPiPi * st);
		return V3{st * sp, ct, st * cp};
	}
};
} // namespace

void orc_env_sample(const float *texels, const uint32_t *alias, uint32_t w, uint32_t h, float rotation, float u,
                    float v, float *dir, float *light, float *pdf) {
	const Env env{texels, alias, w, h, rotation, 1.0f};
	V3 l;
	const V3 d = env.sample(u, v, l, *pdf);
	dir[0] = d.x, dir[1] = d.y, dir[2] = d.z;
	light[0] = l.x, light[1] = l.y, light[2] = l.z;
}
void orc_env_lookup(const float *texels, uint32_t w, uint32_t h, float rotation, const float *dir, float *out,
                    float *pdf) {
	const Env env{texels, nullptr, w, h, rotation, 1.0f};
	const V3 d{dir[0], dir[1], dir[2]};
	env.lookup(d, out);
	*pdf = env.pdf(d);
}

void orc_path_trace_env(const uint32_t *words, uint64_t W, uint32_t level, const float *cam, uint32_t width,
                        uint32_t height, uint32_t bounce, uint32_t spp, const float *texels, const uint32_t *alias,
                        uint32_t ew, uint32_t eh, float rotation, float multiplier, const uint32_t *sobol_mat,
                        const uint8_t *noise, float *color, uint32_t *albedo, uint32_t *normal, int nthreads) {
	(void)level;
	Octree oc{words, W};
	const Env env{texels, alias, ew, eh, rotation, multiplier};
	Sampler sm;
	sm.init(sobol_mat, noise, (bounce + 1) * 2, spp);
	float jx, jy;
	sm.vec2(0, jx, jy);
	const float kPi = 0x1.921fb6p+1f, kBsdf = 1.0f / kPi;
	auto mis = [](float a, float b) {
		const float t = a * a;
		return t / (b * b + t);
	};
	parallel_for((uint64_t)width * height, nthreads, [&](uint64_t b, uint64_t e) {
		for (uint64_t kk = b; kk < e; ++kk) {
			const uint32_t k = (uint32_t)kk, px = k % width, py = k / width;
			float nu, nv;
			sm.pixel_noise(px, py, nu, nv);
			V3 o = cam_vec(cam, 0);
			V3 d = gen_ray(cam, ((float)px + jx) / (float)width, ((float)py + jy) / (float)height);
			orc_hit h = march_leaf(oc, o, d, 1);
			V3 rad{0.0f, 0.0f, 0.0f}, alb0{0.0f, 0.0f, 0.0f}, nrm0{0.0f, 0.0f, 0.0f};
			if (!(h.vz_flags & 0x80000000u)) {
				float t[4];
				env.lookup(d, t);
				rad = V3{t[0], t[1], t[2]};
			} else {
				V3 acc{leaf_colour(h.leaf, 0), leaf_colour(h.leaf, 1), leaf_colour(h.leaf, 2)};
				alb0 = acc;
				uint32_t code = (h.vz_flags >> 24) & 7u;
				nrm0 = face_normal(code);
				for (uint32_t cur = 1; cur < bounce; ++cur) {
					float u, v;
					sm.bounce(cur, nu, nv, u, v);
					o = V3{h.pos[0], h.pos[1], h.pos[2]};
					const V3 nrm = face_normal(code);
					{
						V3 light;
						float lpdf;
						const V3 dl = env.sample(u, v, light, lpdf);
						const float ndd = (nrm.x * dl.x + nrm.y * dl.y) + nrm.z * dl.z;
						if (ndd > 0.0f && !std::isnan(lpdf) && lpdf > 0.0f &&
						    !(march_leaf(oc, o, dl, 2).vz_flags & 0x80000000u)) {
							const float bpdf = ndd / kPi, wgt = mis(lpdf, bpdf);
							rad = V3{rad.x + ((((acc.x * light.x) * multiplier) * wgt) * ndd) * kBsdf / lpdf,
							         rad.y + ((((acc.y * light.y) * multiplier) * wgt) * ndd) * kBsdf / lpdf,
							         rad.z + ((((acc.z * light.z) * multiplier) * wgt) * ndd) * kBsdf / lpdf};
						}
					}
					d = diffuse_dir(code, u, v);
					const float bpdf = std::sqrt(1.0f - u) / kPi;
					h = march_leaf(oc, o, d, 1);
					if (!(h.vz_flags & 0x80000000u)) {
						const float lpdf = env.pdf(d);
						if (!std::isnan(lpdf) && lpdf > 0.0f) {
							float t[4];
							env.lookup(d, t);
							const float wgt = mis(bpdf, lpdf);
							rad = V3{rad.x + ((acc.x * t[0]) * multiplier) * wgt, rad.y + ((acc.y * t[1]) * multiplier) * wgt,
							         rad.z + ((acc.z * t[2]) * multiplier) * wgt};
						}
						break;
					}
					acc = V3{acc.x * leaf_colour(h.leaf, 0), acc.y * leaf_colour(h.leaf, 1), acc.z * leaf_colour(h.leaf, 2)};
					code = (h.vz_flags >> 24) & 7u;
				}
			}
			store_sample(k, spp, rad, alb0, nrm0, color, albedo, normal);
		}
	});
}

// The reference tracer's shading of a primary frame (shader/octree_tracer.frag:18-47), into RGBA8
// UNORM texels.  texels may be NULL (constant light).
void orc_shade(const orc_hit *hits, const float *cam, uint32_t width, uint32_t height, int view, const float *light,
               const float *texels, uint32_t ew, uint32_t eh, float rotation, uint32_t *out) {
	const Env env{texels, nullptr, ew, eh, rotation, 1.0f};
	for (uint64_t k = 0; k < (uint64_t)width * height; ++k) {
		const orc_hit &h = hits[k];
		const bool hit = (h.vz_flags & 0x80000000u) != 0u;
		float c[3];
		if (view == 3) { // Heat(iter / 128.0)
			float x = (float)h.iter / 128.0f;
			x = x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
			for (int i = 0; i < 3; ++i) {
				float sn, cs;
				sincos_poly(x * 3.0f - (float)(i + 1), sn, cs);
				c[i] = sn * 0.5f + 0.5f;
			}
		} else if (view == 2) { // pos - 1.0; a miss has pos = 1
			for (int i = 0; i < 3; ++i)
				c[i] = hit ? h.pos[i] - 1.0f : 0.0f;
		} else if (view == 1) { // normal * 0.5 + 0.5; a miss has normal = 0
			const V3 n = hit ? face_normal((h.vz_flags >> 24) & 7u) : V3{0.0f, 0.0f, 0.0f};
			c[0] = n.x * 0.5f + 0.5f, c[1] = n.y * 0.5f + 0.5f, c[2] = n.z * 0.5f + 0.5f;
		} else { // pow(color, 1 / 2.2); a miss takes Light(d)
			float col[3];
			if (hit) {
				for (int i = 0; i < 3; ++i)
					col[i] = leaf_colour(h.leaf, i);
			} else if (texels) {
				const uint32_t px = (uint32_t)(k % width), py = (uint32_t)(k / width);
				const V3 d = gen_ray(cam, (float)px / (float)width, (float)py / (float)height);
				float t[4];
				env.lookup(d, t);
				col[0] = t[0], col[1] = t[1], col[2] = t[2];
			} else {
				col[0] = light[0], col[1] = light[1], col[2] = light[2];
			}
			for (int i = 0; i < 3; ++i)
				c[i] = std::pow(col[i], 1.0f / 2.2f);
		}
		out[k] = unorm8(c[0]) | unorm8(c[1]) << 8 | unorm8(c[2]) << 16 | 0xFF000000u;
	}
}

uint64_t orc_stack_uninit_reads(void) { return g_uninit_reads.load(); }
void orc_reset_counters(void) { g_uninit_reads.store(0); }

} // extern "C"
