// ESVO ray marching over the octree (OctreeTracer + shader/octree.glsl), one ray per lane.
//
//   march_beam    per 8x8-pixel beam cell: Octree_RayMarchCoarse (shader/octree.glsl:47-177) with
//                 the cone-width stop of shader/octree_tracer_beam.frag:12-29
//   march_frame   per pixel: Camera_GenRay + beam start + Octree_RayMarchLeaf 6-arg
//                 (shader/octree_tracer.frag:23-37, shader/octree.glsl:179-340)
//   march_rays    explicit rays: 6-arg, 5-arg (:342-502) or occlusion (:504-629)
//   bounce_rays   one cosine-weighted diffuse bounce per hit (shader/path_tracer.comp:38-44)
//
// The traversal stack (shader/octree.glsl:45, 23 entries indexed by scale) lives in LDS, one
// column per lane, and only holds the scales an octree of level <= 13 can push (11..22); a deeper
// push is reported as an error instead of indexing out of the array.  Octree loads are bounds
// checked, and the loop has an iteration cap so a malformed octree cannot hang the GPU; both
// raise error bits and never trigger on a well-formed octree.
//
// The traversal restates code that carries this notice:
//
//  Copyright (c) 2009-2011, NVIDIA Corporation
//  All rights reserved.
//
//  Redistribution and use in source and binary forms, with or without
//  modification, are permitted provided that the following conditions are met:
//      * Redistributions of source code must retain the above copyright
//        notice, this list of conditions and the following disclaimer.
//      * Redistributions in binary form must reproduce the above copyright
//        notice, this list of conditions and the following disclaimer in the
//        documentation and/or other materials provided with the distribution.
//      * Neither the name of NVIDIA Corporation nor the
//        names of its contributors may be used to endorse or promote products
//        derived from this software without specific prior written permission.
//
//  THIS SOFTWARE IS PROVIDED BY THE COPYRIGHT HOLDERS AND CONTRIBUTORS "AS IS"
//  AND ANY EXPRESS OR IMPLIED WARRANTIES, INCLUDING, BUT NOT LIMITED TO, THE
//  IMPLIED WARRANTIES OF MERCHANTABILITY AND FITNESS FOR A PARTICULAR PURPOSE
//  ARE DISCLAIMED. IN NO EVENT SHALL <COPYRIGHT HOLDER> BE LIABLE FOR ANY
//  DIRECT, INDIRECT, INCIDENTAL, SPECIAL, EXEMPLARY, OR CONSEQUENTIAL DAMAGES
//  (INCLUDING, BUT NOT LIMITED TO, PROCUREMENT OF SUBSTITUTE GOODS OR SERVICES;
//  LOSS OF USE, DATA, OR PROFITS; OR BUSINESS INTERRUPTION) HOWEVER CAUSED AND
//  ON ANY THEORY OF LIABILITY, WHETHER IN CONTRACT, STRICT LIABILITY, OR TORT
//  (INCLUDING NEGLIGENCE OR OTHERWISE) ARISING IN ANY WAY OUT OF THE USE OF
//  THIS SOFTWARE, EVEN IF ADVISED OF THE POSSIBILITY OF SUCH DAMAGE.
#include "svo_device.hpp"
#include "svo_runtime.hpp"

namespace svo {
namespace {

constexpr uint32_t kStackSize = 23u;                   // octree.glsl:43
constexpr uint32_t kStackSlots = 12u;                  // scales 11..22
constexpr uint32_t kMinScale = kStackSize - kStackSlots;
constexpr uint32_t kMaxIter = 1u << 20;
constexpr int kBlock = 256;
#define SVO_EPS 3.552713678800501e-15f                 // octree.glsl:44, 2^-48

struct March {
	f3 t_coef, t_bias;
	uint32_t oct_mask;
	float t_min, t_max, h;
	uint32_t parent, cur, idx, scale;
	f3 pos;
	float scale_exp2;
};

// Ray set-up (octree.glsl:48-83 and its copies in the other three marchers).
template <bool kGeEps> __device__ __forceinline__ void march_init(f3 o, f3 &d, March &m) {
	if (kGeEps) {
		d.x = fabsf(d.x) >= SVO_EPS ? d.x : (d.x >= 0 ? SVO_EPS : -SVO_EPS);
		d.y = fabsf(d.y) >= SVO_EPS ? d.y : (d.y >= 0 ? SVO_EPS : -SVO_EPS);
		d.z = fabsf(d.z) >= SVO_EPS ? d.z : (d.z >= 0 ? SVO_EPS : -SVO_EPS);
	} else {
		d.x = fabsf(d.x) > SVO_EPS ? d.x : (d.x >= 0 ? SVO_EPS : -SVO_EPS);
		d.y = fabsf(d.y) > SVO_EPS ? d.y : (d.y >= 0 ? SVO_EPS : -SVO_EPS);
		d.z = fabsf(d.z) > SVO_EPS ? d.z : (d.z >= 0 ? SVO_EPS : -SVO_EPS);
	}
	m.t_coef = {1.0f / -fabsf(d.x), 1.0f / -fabsf(d.y), 1.0f / -fabsf(d.z)};
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
	m.scale = kStackSize - 1u;
	m.scale_exp2 = 0.5f;
}

// One trip of the traversal loop (octree.glsl:86-172 / 221-304).  Returns false on break.
// stk: this lane's LDS column (slot s at stk[s * kBlock]).
template <bool kCoarse>
__device__ __forceinline__ bool march_step(const uint32_t *__restrict__ oct, uint64_t W, March &m, uint32_t *stk,
                                           uint32_t &nfetch, uint32_t &err, float orig_sz, float dir_sz) {
	if (m.cur == 0u) {
		uint64_t a = (uint64_t)m.parent + (m.idx ^ m.oct_mask);
		++nfetch;
		if (a < W)
			m.cur = oct[a];
		else
			err |= kErrOutOfRange;
	}
	f3 t_corner{m.pos.x * m.t_coef.x - m.t_bias.x, m.pos.y * m.t_coef.y - m.t_bias.y,
	            m.pos.z * m.t_coef.z - m.t_bias.z};
	float tc_max = gmin(gmin(t_corner.x, t_corner.y), t_corner.z);

	if ((m.cur & 0x80000000u) != 0u && m.t_min <= m.t_max) {
		if (kCoarse && orig_sz + tc_max * dir_sz >= m.scale_exp2)
			return false;
		// INTERSECT
		float half_scale_exp2 = m.scale_exp2 * 0.5f;
		f3 t_center{half_scale_exp2 * m.t_coef.x + t_corner.x, half_scale_exp2 * m.t_coef.y + t_corner.y,
		            half_scale_exp2 * m.t_coef.z + t_corner.z};
		if ((m.cur & 0x40000000u) != 0u)
			return false; // leaf
		// PUSH
		if (m.scale < kMinScale) {
			err |= kErrTooDeep;
			m.scale = kStackSize; // ends the ray as a miss
			return false;
		}
		if (tc_max < m.h)
			stk[(m.scale - kMinScale) * kBlock] = m.parent;
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
		m.cur = 0u;
		return true;
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
	if ((m.idx & step_mask) != 0u) {
		// POP
		uint32_t differing_bits = 0u;
		if ((step_mask & 1u) != 0u)
			differing_bits |= fbits(m.pos.x) ^ fbits(m.pos.x + m.scale_exp2);
		if ((step_mask & 2u) != 0u)
			differing_bits |= fbits(m.pos.y) ^ fbits(m.pos.y + m.scale_exp2);
		if ((step_mask & 4u) != 0u)
			differing_bits |= fbits(m.pos.z) ^ fbits(m.pos.z + m.scale_exp2);
		m.scale = find_msb(differing_bits);
		if (m.scale >= kStackSize)
			return false;
		if (m.scale < kMinScale) {
			err |= kErrTooDeep;
			m.scale = kStackSize;
			return false;
		}
		m.scale_exp2 = bitsf((m.scale - kStackSize + 127u) << 23u);
		m.parent = stk[(m.scale - kMinScale) * kBlock];
		uint32_t shx = fbits(m.pos.x) >> m.scale, shy = fbits(m.pos.y) >> m.scale, shz = fbits(m.pos.z) >> m.scale;
		m.pos.x = bitsf(shx << m.scale);
		m.pos.y = bitsf(shy << m.scale);
		m.pos.z = bitsf(shz << m.scale);
		m.idx = (shx & 1u) | ((shy & 1u) << 1u) | ((shz & 1u) << 2u);
		m.h = 0.0f;
		m.cur = 0u;
	}
	return true;
}

// Octree_RayMarchLeaf (variant 0: 6-arg with iteration count; 1: 5-arg) and
// Octree_RayMarchOcclude (variant 2), with the outputs of octree.glsl:307-339.
template <int kVariant>
__device__ __forceinline__ void march_leaf(const uint32_t *__restrict__ oct, uint64_t W, f3 o, f3 d, uint32_t *stk,
                                           uint32_t &err, uint4 &out0, uint4 &out1) {
	March m;
	uint32_t nfetch = 0u, iter = 0u;
	march_init<kVariant != 0>(o, d, m);
	while (m.scale < kStackSize) {
		++iter;
		if (iter > kMaxIter) {
			err |= kErrIterCap;
			break;
		}
		if (!march_step<false>(oct, W, m, stk, nfetch, err, 0.0f, 0.0f))
			break;
	}
	const bool hit = m.scale < kStackSize && m.t_min <= m.t_max;
	if (kVariant == 2) {
		out0 = make_uint4(0u, 0u, 0u, 0u);
		out1 = make_uint4(0xFFFFFFFFu, hit ? 0x80000000u : 0u, 0u, nfetch);
		return;
	}
	f3 t_corner{m.t_coef.x * (m.pos.x + m.scale_exp2) - m.t_bias.x,
	            m.t_coef.y * (m.pos.y + m.scale_exp2) - m.t_bias.y,
	            m.t_coef.z * (m.pos.z + m.scale_exp2) - m.t_bias.z};
	const uint32_t axis = (t_corner.x > t_corner.y && t_corner.x > t_corner.z) ? 0u : (t_corner.y > t_corner.z ? 1u : 2u);
	// norm = -e_axis, flipped where the ray was not mirrored
	const bool positive = ((m.oct_mask >> axis) & 1u) == 0u;
	f3 pos = m.pos;
	if ((m.oct_mask & 1u) != 0u)
		pos.x = 3.0f - m.scale_exp2 - pos.x;
	if ((m.oct_mask & 2u) != 0u)
		pos.y = 3.0f - m.scale_exp2 - pos.y;
	if ((m.oct_mask & 4u) != 0u)
		pos.z = 3.0f - m.scale_exp2 - pos.z;
	f3 op{o.x + m.t_min * d.x, o.y + m.t_min * d.y, o.z + m.t_min * d.z};
	op.x = gmin(gmax(op.x, pos.x), pos.x + m.scale_exp2);
	op.y = gmin(gmax(op.y, pos.y), pos.y + m.scale_exp2);
	op.z = gmin(gmax(op.z, pos.z), pos.z + m.scale_exp2);
	float face = positive ? ((axis == 0u ? pos.x : axis == 1u ? pos.y : pos.z) + m.scale_exp2) + SVO_EPS * 2.0f
	                      : (axis == 0u ? pos.x : axis == 1u ? pos.y : pos.z) - SVO_EPS;
	if (axis == 0u)
		op.x = face;
	else if (axis == 1u)
		op.y = face;
	else
		op.z = face;
	const uint32_t code = axis * 2u + (positive ? 1u : 0u);
	uint32_t vxy, vzf;
	if (hit) {
		const float inv = bitsf((127u + (kStackSize - m.scale)) << 23u);
		uint32_t vx = (uint32_t)((pos.x - 1.0f) * inv), vy = (uint32_t)((pos.y - 1.0f) * inv),
		         vz = (uint32_t)((pos.z - 1.0f) * inv);
		vxy = vx | (vy << 16);
		vzf = vz | ((kStackSize - m.scale) << 16) | (code << 24) | 0x80000000u;
	} else {
		vxy = 0xFFFFFFFFu;
		vzf = 0xFFFFu | (code << 24);
	}
	out0 = make_uint4(fbits(op.x), fbits(op.y), fbits(op.z), kVariant == 0 ? iter : 0u);
	out1 = make_uint4(vxy, vzf, m.cur, nfetch);
}

struct CamArgs {
	float px, py, pz, lx, ly, lz, sx, sy, sz, ux, uy, uz;
};

// Camera_GenRay (shader/camera.glsl:11-14)
__device__ __forceinline__ f3 gen_ray(const CamArgs &c, float cx, float cy) {
	cx = cx * 2.0f - 1.0f;
	cy = cy * 2.0f - 1.0f;
	f3 v{c.lx - c.sx * cx - c.ux * cy, c.ly - c.sy * cx - c.uy * cy, c.lz - c.sz * cx - c.uz * cy};
	return normalize3(v);
}

// Linear block id -> tile so that the blocks one XCD receives (ids b, b+8, b+16, ...) cover one
// contiguous band of the frame and share its octree working set in that XCD's L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nblocks) {
	const uint32_t per = (nblocks + 7u) / 8u;
	const uint32_t full = nblocks - (per - 1u) * 8u; // XCDs that get `per` blocks
	const uint32_t x = b & 7u, k = b >> 3;
	const uint32_t start = x < full ? x * per : full * per + (x - full) * (per - 1u);
	return start + k;
}

__global__ void __launch_bounds__(kBlock)
    march_beam_kernel(const uint32_t *__restrict__ oct, uint64_t W, CamArgs cam, uint32_t width, uint32_t height,
                      uint32_t bw, uint32_t bh, float orig_sz, float *__restrict__ beam, uint32_t *__restrict__ nf_out,
                      uint32_t *err_out) {
	__shared__ uint32_t s_stack[kStackSlots * kBlock];
	const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
	if (k >= bw * bh)
		return;
	const uint32_t bx = k % bw, by = k / bw;
	const float fx = (float)bx, fy = (float)by, fw = (float)width, fh = (float)height, bs = 8.0f;
	// octree_tracer_beam.frag:12-29
	f3 d = gen_ray(cam, fx / fw * bs, fy / fh * bs);
	float tmax = 0.0f;
	const float ox[4] = {0.5f, -0.5f, 0.5f, -0.5f}, oy[4] = {0.5f, 0.5f, -0.5f, -0.5f};
	float tg[4];
#pragma unroll
	for (int q = 0; q < 4; ++q) {
		f3 y = gen_ray(cam, (fx + ox[q]) / fw * bs, (fy + oy[q]) / fh * bs);
		float c = dot3(d, y);
		tg[q] = __builtin_sqrtf(1.0f - c * c) / fabsf(c);
	}
	tmax = gmax(gmax(tg[0], tg[1]), gmax(tg[2], tg[3]));
	const float dir_sz = 2.0f * tmax;
	uint32_t err = 0u, nfetch = 0u;
	March m;
	f3 o{cam.px, cam.py, cam.pz};
	march_init<false>(o, d, m);
	uint32_t *stk = s_stack + threadIdx.x;
	uint32_t iter = 0u;
	for (;;) {
		if (++iter > kMaxIter) {
			err |= kErrIterCap;
			break;
		}
		if (!march_step<true>(oct, W, m, stk, nfetch, err, orig_sz, dir_sz))
			break;
	}
	const bool hit = m.scale < kStackSize && m.t_min <= m.t_max;
	beam[k] = hit ? gmax(0.0f, m.t_min - m.scale_exp2) : 1e10f;
	if (nf_out)
		nf_out[k] = nfetch;
	if (err)
		atomicOr(err_out, err);
}

__global__ void __launch_bounds__(kBlock)
    march_frame_kernel(const uint32_t *__restrict__ oct, uint64_t W, CamArgs cam, uint32_t width, uint32_t height,
                       const float *__restrict__ beam, uint32_t bw, uint32_t tiles_x, uint4 *__restrict__ hits,
                       uint32_t *err_out) {
	__shared__ uint32_t s_stack[kStackSlots * kBlock];
	// 16x16-pixel tile per block, one 8x8 quad per wave
	const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
	const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
	const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
	const uint32_t px = tx * 16u + (wave & 1u) * 8u + (lane & 7u);
	const uint32_t py = ty * 16u + (wave >> 1) * 8u + (lane >> 3);
	if (px >= width || py >= height)
		return;
	// octree_tracer.frag:23-32
	f3 o{cam.px, cam.py, cam.pz};
	f3 d = gen_ray(cam, (float)px / (float)width, (float)py / (float)height);
	if (beam) {
		const uint32_t bx = px >> 3, by = py >> 3;
		const float *b0 = beam + by * bw + bx, *b1 = b0 + bw;
		float bm = gmin(gmin(b0[0], b0[1]), gmin(b1[0], b1[1]));
		o = f3{o.x + d.x * bm, o.y + d.y * bm, o.z + d.z * bm};
	}
	uint32_t err = 0u;
	uint4 r0, r1;
	march_leaf<0>(oct, W, o, d, s_stack + threadIdx.x, err, r0, r1);
	const uint64_t k = (uint64_t)py * width + px;
	hits[2 * k] = r0;
	hits[2 * k + 1] = r1;
	if (err)
		atomicOr(err_out, err);
}

template <int kVariant>
__global__ void __launch_bounds__(kBlock)
    march_rays_kernel(const uint32_t *__restrict__ oct, uint64_t W, const float *__restrict__ rays, uint64_t R,
                      uint4 *__restrict__ hits, uint32_t *err_out) {
	__shared__ uint32_t s_stack[kStackSlots * kBlock];
	const uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
	if (k >= R)
		return;
	const float *r = rays + 6 * k;
	f3 o{r[0], r[1], r[2]}, d{r[3], r[4], r[5]};
	uint32_t err = 0u;
	uint4 r0, r1;
	march_leaf<kVariant>(oct, W, o, d, s_stack + threadIdx.x, err, r0, r1);
	hits[2 * k] = r0;
	hits[2 * k + 1] = r1;
	if (err)
		atomicOr(err_out, err);
}

// ---- bounce sampling ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t sobol0(uint32_t n) { return __brev(n); }
__device__ __forceinline__ uint32_t sobol1(uint32_t n) { // primitive polynomial x + 1
	uint32_t v = 0x80000000u, r = 0u;
	for (int k = 0; k < 32; ++k) {
		if ((n >> k) & 1u)
			r ^= v;
		v ^= v >> 1;
	}
	return r;
}

// sin/cos: quadrant reduction with a two-constant split of pi/2, Taylor polynomials on [-pi/4, pi/4].
__device__ __forceinline__ void sincos_poly(float phi, float &s, float &c) {
	const float two_over_pi = 0x1.45f306p-1f;
	const float pio2_hi = 0x1.921ep+0f;
	const float pio2_lo = 0x1.b54442p-16f;
	float k = __builtin_rintf(phi * two_over_pi);
	float r = (phi - k * pio2_hi) - k * pio2_lo;
	float r2 = r * r;
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

__global__ void __launch_bounds__(kBlock)
    bounce_kernel(const uint4 *__restrict__ hits, uint64_t R, uint32_t s0, uint32_t s1, float *__restrict__ rays) {
	const uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
	if (k >= R)
		return;
	const uint4 h0 = hits[2 * k], h1 = hits[2 * k + 1];
	float *r = rays + 6 * k;
	if (!(h1.y & 0x80000000u)) {
		r[0] = r[1] = r[2] = 0.0f;
		r[3] = r[4] = r[5] = -1.0f;
		return;
	}
	const float kPi = 3.1415926535897932384626433832795f; // shader/util.glsl:4
	const uint32_t n = (uint32_t)k;
	float u = (float)((sobol0(n) ^ s0) >> 8) * 5.9604644775390625e-08f;
	float v = (float)((sobol1(n) ^ s1) >> 8) * 5.9604644775390625e-08f;
	const uint32_t code = (h1.y >> 24) & 7u;
	const float sg = (code & 1u) ? 1.0f : -1.0f;
	f3 nrm{code >> 1 == 0u ? sg : 0.0f, code >> 1 == 1u ? sg : 0.0f, code >> 1 == 2u ? sg : 0.0f};
	// Diffuse_Sample (shader/diffuse.glsl:7-14)
	float rr = __builtin_sqrtf(u), phi = 2.0f * kPi * v;
	float sn, cs;
	sincos_poly(phi, sn, cs);
	f3 dl{rr * cs, rr * sn, __builtin_sqrtf(1.0f - u)};
	// AlignDirection (shader/util.glsl:5-9)
	f3 ax = fabsf(nrm.x) > 0.01f ? f3{0.0f, 1.0f, 0.0f} : f3{1.0f, 0.0f, 0.0f};
	f3 uu = normalize3(cross3(ax, nrm));
	f3 vv = cross3(nrm, uu);
	r[0] = __uint_as_float(h0.x);
	r[1] = __uint_as_float(h0.y);
	r[2] = __uint_as_float(h0.z);
	r[3] = dl.x * uu.x + dl.y * vv.x + dl.z * nrm.x;
	r[4] = dl.x * uu.y + dl.y * vv.y + dl.z * nrm.y;
	r[5] = dl.x * uu.z + dl.y * vv.z + dl.z * nrm.z;
}

CamArgs cam_args(const svo_camera *c) {
	return CamArgs{c->pos[0],  c->pos[1],  c->pos[2],  c->look[0], c->look[1], c->look[2],
	               c->side[0], c->side[1], c->side[2], c->up[0],   c->up[1],   c->up[2]};
}

uint64_t splitmix64(uint64_t &s) {
	uint64_t z = (s += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

} // namespace

int march_beam(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t w, uint32_t h, float *beam_img,
               uint32_t *nfetch) {
	const CamArgs ca = cam_args(cam);
	const uint32_t bw = (w + 7u) / 8u + 1u, bh = (h + 7u) / 8u + 1u; // src/OctreeTracer.cpp:6
	// exp2f(1.0f - level) (src/OctreeTracer.cpp:206): an exact power of two
	const float orig_sz = ldexpf(1.0f, 1 - (int)oct->level);
	ctx->t_beam.start(ctx->stream);
	hipError_t e = launch(ctx, "march_beam", march_beam_kernel, dim3(div_up((uint64_t)bw * bh, kBlock)), dim3(kBlock),
	                      0, (const uint32_t *)oct->words, oct->W, ca, w, h, bw, bh, orig_sz, beam_img, nfetch,
	                      ctx->d_err);
	ctx->t_beam.stop(ctx->stream);
	if (e != hipSuccess)
		return hip_check(ctx, e, "march_beam");
	return SVO_OK;
}

int march_frame(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t w, uint32_t h, int beam,
                svo_hit *hits, float *beam_out) {
	hipError_t e;
	const CamArgs ca = cam_args(cam);
	const uint32_t bw = (w + 7u) / 8u + 1u, bh = (h + 7u) / 8u + 1u; // src/OctreeTracer.cpp:6
	float *beam_img = nullptr;
	if (beam) {
		beam_img = beam_out;
		if (!beam_img) {
			if ((e = ctx->pool->alloc((void **)&beam_img, sizeof(float) * bw * bh)) != hipSuccess)
				return set_error(ctx, SVO_ERR_OUT_OF_MEMORY, "beam image allocation failed");
		}
		int rc = march_beam(ctx, oct, cam, w, h, beam_img, nullptr);
		if (rc)
			return rc;
	}
	const uint32_t tiles_x = (w + 15u) / 16u, tiles_y = (h + 15u) / 16u;
	ctx->t_march.start(ctx->stream);
	e = launch(ctx, "march_frame", march_frame_kernel, dim3(tiles_x * tiles_y), dim3(kBlock), 0,
	           (const uint32_t *)oct->words, oct->W, ca, w, h, (const float *)beam_img, bw, tiles_x,
	           reinterpret_cast<uint4 *>(hits), ctx->d_err);
	ctx->t_march.stop(ctx->stream);
	if (e != hipSuccess)
		return hip_check(ctx, e, "march_frame");
	if (beam && !beam_out)
		ctx->pool->free(beam_img); // stream-ordered reuse: the pool only hands it to later work on this stream
	return SVO_OK;
}

int march_rays(svo_ctx *ctx, const svo_octree *oct, const float *rays, uint64_t R, int variant, svo_hit *hits) {
	if (R == 0)
		return SVO_OK;
	hipError_t e;
	const dim3 grid(div_up(R, kBlock));
	uint4 *out = reinterpret_cast<uint4 *>(hits);
	ctx->t_rays.start(ctx->stream);
	if (variant == SVO_MARCH_LEAF_ITER)
		e = launch(ctx, "march_rays6", march_rays_kernel<0>, grid, dim3(kBlock), 0, (const uint32_t *)oct->words,
		           oct->W, rays, R, out, ctx->d_err);
	else if (variant == SVO_MARCH_LEAF)
		e = launch(ctx, "march_rays5", march_rays_kernel<1>, grid, dim3(kBlock), 0, (const uint32_t *)oct->words,
		           oct->W, rays, R, out, ctx->d_err);
	else
		e = launch(ctx, "march_occlude", march_rays_kernel<2>, grid, dim3(kBlock), 0, (const uint32_t *)oct->words,
		           oct->W, rays, R, out, ctx->d_err);
	ctx->t_rays.stop(ctx->stream);
	if (e != hipSuccess)
		return hip_check(ctx, e, "march_rays");
	return SVO_OK;
}

int bounce_rays(svo_ctx *ctx, const svo_hit *hits, uint64_t R, uint64_t seed, float *rays) {
	if (R == 0)
		return SVO_OK;
	uint64_t st = seed;
	uint64_t z = splitmix64(st);
	ctx->t_bounce.start(ctx->stream);
	hipError_t e = launch(ctx, "bounce", bounce_kernel, dim3(div_up(R, kBlock)), dim3(kBlock), 0,
	                      reinterpret_cast<const uint4 *>(hits), R, (uint32_t)z, (uint32_t)(z >> 32), rays);
	ctx->t_bounce.stop(ctx->stream);
	if (e != hipSuccess)
		return hip_check(ctx, e, "bounce");
	return SVO_OK;
}

} // namespace svo
