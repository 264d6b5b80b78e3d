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
#include <algorithm>
#include <cstdio>
#include <mutex>
#include <string>

#include "svo_batch.hpp"
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
// stk: this lane's LDS column (slot s at stk[s * kStride], kStride = threads per block).
//
// Node words are addressed with 32-bit offsets: octrees hold at most 2^30 words (30-bit child
// pointers; svo_octree_create and the builder reject larger ones), so an index below W is below
// 2^30 and its byte offset fits in 32 bits.
template <bool kCoarse, uint32_t kStride = kBlock>
__device__ __forceinline__ bool march_step(const uint32_t *__restrict__ oct, uint64_t W, March &m, uint32_t *stk,
                                           uint32_t &nfetch, uint32_t &err, float orig_sz, float dir_sz) {
	const uint32_t W32 = W < (1ull << 30) ? (uint32_t)W : (1u << 30);
	uint32_t *stk_s = stk - kMinScale * kStride; // slot of scale s at stk_s[s * kStride]
	if (m.cur == 0u) {
		const uint32_t a = m.parent + (m.idx ^ m.oct_mask); // parent < 2^30: no wrap-around
		++nfetch;
		if (a < W32)
			m.cur = *reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(oct) + (a << 2u));
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
		if ((m.cur & 0x40000000u) != 0u)
			return false; // leaf
		// PUSH.  A push below the stack's last slot is an error that ends the ray as a miss: it
		// writes no slot, and the scale it leaves (kStackSize) stops the loop before the next trip
		// counts, as an immediate break would.
		const bool too_deep = m.scale < kMinScale;
		err |= too_deep ? kErrTooDeep : 0u;
		const float half_scale_exp2 = m.scale_exp2 * 0.5f;
		f3 t_center{half_scale_exp2 * m.t_coef.x + t_corner.x, half_scale_exp2 * m.t_coef.y + t_corner.y,
		            half_scale_exp2 * m.t_coef.z + t_corner.z};
		if (tc_max < m.h && !too_deep)
			stk_s[m.scale * kStride] = m.parent;
		m.h = tc_max;
		m.parent = m.cur & 0x3fffffffu;
		m.scale = too_deep ? kStackSize : m.scale - 1u;
		m.scale_exp2 = half_scale_exp2;
		const bool cx = t_center.x > m.t_min, cy = t_center.y > m.t_min, cz = t_center.z > m.t_min;
		m.idx = (cx ? 1u : 0u) | (cy ? 2u : 0u) | (cz ? 4u : 0u);
		m.pos.x = cx ? m.pos.x + half_scale_exp2 : m.pos.x;
		m.pos.y = cy ? m.pos.y + half_scale_exp2 : m.pos.y;
		m.pos.z = cz ? m.pos.z + half_scale_exp2 : m.pos.z;
		m.cur = 0u;
		return true;
	}
	// ADVANCE.  pos is a multiple of scale_exp2 in [1, 2), so pos - scale_exp2 is exact and the
	// POP's pos + scale_exp2 of a stepped axis (octree.glsl:275-280) is pos before the step; an axis
	// that does not step contributes no differing bits.
	const bool sx = t_corner.x <= tc_max, sy = t_corner.y <= tc_max, sz = t_corner.z <= tc_max;
	const uint32_t step_mask = (sx ? 1u : 0u) | (sy ? 2u : 0u) | (sz ? 4u : 0u);
	const float nx = sx ? m.pos.x - m.scale_exp2 : m.pos.x, ny = sy ? m.pos.y - m.scale_exp2 : m.pos.y,
	            nz = sz ? m.pos.z - m.scale_exp2 : m.pos.z;
	const uint32_t differing_bits =
	    (fbits(nx) ^ fbits(m.pos.x)) | (fbits(ny) ^ fbits(m.pos.y)) | (fbits(nz) ^ fbits(m.pos.z));
	m.pos = f3{nx, ny, nz};
	m.t_min = tc_max;
	m.idx ^= step_mask;
	if ((m.idx & step_mask) != 0u) {
		// POP.  differing_bits != 0 and its highest bit lies above the bit of the old scale (the
		// step borrowed from it), so the new scale exceeds the old one (>= kMinScale - 1) and its
		// stack slot exists.
		m.scale = 31u - (uint32_t)__builtin_clz(differing_bits);
		if (m.scale >= kStackSize)
			return false;
		m.scale_exp2 = bitsf((m.scale - kStackSize + 127u) << 23u);
		m.parent = stk_s[m.scale * kStride];
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

// Outputs of Octree_RayMarchLeaf / Octree_RayMarchOcclude (octree.glsl:307-339) from the final
// traversal state; d is the direction as march_init left it.
template <int kVariant>
__device__ __forceinline__ void march_finish(const March &m, f3 o, f3 d, uint32_t iter, uint32_t nfetch, uint4 &out0,
                                             uint4 &out1) {
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

// One trip of the leaf march loop (octree.glsl:219-305 `while (scale < STACK_SIZE)`); returns false
// when the loop has ended.
template <uint32_t kStride>
__device__ __forceinline__ bool march_trip(const uint32_t *__restrict__ oct, uint64_t W, March &m, uint32_t *stk,
                                           uint32_t &iter, uint32_t &nfetch, uint32_t &err) {
	if (m.scale >= kStackSize)
		return false;
	if (++iter > kMaxIter) {
		err |= kErrIterCap;
		return false;
	}
	return march_step<false, kStride>(oct, W, m, stk, nfetch, err, 0.0f, 0.0f);
}

// Octree_RayMarchLeaf (variant 0: 6-arg with iteration count; 1: 5-arg) and
// Octree_RayMarchOcclude (variant 2), with the outputs of octree.glsl:307-339.
template <int kVariant, uint32_t kStride = kBlock>
__device__ __forceinline__ void march_leaf(const uint32_t *__restrict__ oct, uint64_t W, f3 o, f3 d, uint32_t *stk,
                                           uint32_t &err, uint4 &out0, uint4 &out1) {
	March m;
	uint32_t nfetch = 0u, iter = 0u;
	march_init<kVariant != 0>(o, d, m);
	while (march_trip<kStride>(oct, W, m, stk, iter, nfetch, err)) {
	}
	march_finish<kVariant>(m, o, d, iter, nfetch, out0, out1);
}

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

// One beam cell (octree_tracer_beam.frag:12-29): the distance every ray of the cell can skip.
template <uint32_t kStride>
__device__ __forceinline__ float beam_cell(const uint32_t *__restrict__ oct, uint64_t W, const CamArgs &cam,
                                           uint32_t width, uint32_t height, uint32_t bx, uint32_t by, float orig_sz,
                                           uint32_t *stk, uint32_t &err, uint32_t &nfetch) {
	const float fx = (float)bx, fy = (float)by, fw = (float)width, fh = (float)height, bs = 8.0f;
	f3 d = gen_ray(cam, fx / fw * bs, fy / fh * bs);
	const float ox[4] = {0.5f, -0.5f, 0.5f, -0.5f}, oy[4] = {0.5f, 0.5f, -0.5f, -0.5f};
	float tg[4];
#pragma unroll
	for (int q = 0; q < 4; ++q) {
		f3 y = gen_ray(cam, (fx + ox[q]) / fw * bs, (fy + oy[q]) / fh * bs);
		float c = dot3(d, y);
		tg[q] = __builtin_sqrtf(1.0f - c * c) / fabsf(c);
	}
	const float tmax = gmax(gmax(tg[0], tg[1]), gmax(tg[2], tg[3]));
	const float dir_sz = 2.0f * tmax;
	March m;
	f3 o{cam.px, cam.py, cam.pz};
	march_init<false>(o, d, m);
	uint32_t iter = 0u;
	for (;;) {
		if (m.scale >= kStackSize) // after a push below the stack's last slot
			break;
		if (++iter > kMaxIter) {
			err |= kErrIterCap;
			break;
		}
		if (!march_step<true, kStride>(oct, W, m, stk, nfetch, err, orig_sz, dir_sz))
			break;
	}
	const bool hit = m.scale < kStackSize && m.t_min <= m.t_max;
	return hit ? gmax(0.0f, m.t_min - m.scale_exp2) : 1e10f;
}

// One pixel of the primary march (octree_tracer.frag:23-37): ray through the pixel corner, start
// moved to the minimum of the 4 beam texels around it, 6-argument leaf march.
// Primary ray through pixel (px, py) (octree_tracer.frag:23-32): through the pixel corner, start
// moved to the minimum of the 4 beam texels around it.
__device__ __forceinline__ void frame_ray(const CamArgs &cam, uint32_t width, uint32_t height,
                                          const float *__restrict__ beam, uint32_t bw, uint32_t px, uint32_t py, f3 &o,
                                          f3 &d) {
	o = f3{cam.px, cam.py, cam.pz};
	d = gen_ray(cam, (float)px / (float)width, (float)py / (float)height);
	if (beam) {
		const uint32_t bx = px >> 3, by = py >> 3;
		const float *b0 = beam + by * bw + bx, *b1 = b0 + bw;
		float bm = gmin(gmin(b0[0], b0[1]), gmin(b1[0], b1[1]));
		o = f3{o.x + d.x * bm, o.y + d.y * bm, o.z + d.z * bm};
	}
}

// One pixel of the primary march (octree_tracer.frag:23-37), 6-argument leaf march.
template <uint32_t kStride>
__device__ __forceinline__ void frame_pixel(const uint32_t *__restrict__ oct, uint64_t W, const CamArgs &cam,
                                            uint32_t width, uint32_t height, const float *__restrict__ beam, uint32_t bw,
                                            uint32_t px, uint32_t py, uint32_t *stk, uint32_t &err, uint4 &r0,
                                            uint4 &r1) {
	f3 o, d;
	frame_ray(cam, width, height, beam, bw, px, py, o, d);
	march_leaf<0, kStride>(oct, W, o, d, stk, err, r0, r1);
}

// 16 beam cells per block (one wave, 16 lanes active): a frame has only (W/8+1)(H/8+1) beam cells
// (33K at 1920x1080), too few to fill the GPU one per lane.  Spreading them thinly gives every SIMD
// several waves to switch between while node loads are in flight, and a wave lasts as long as the
// longest of 16 coarse marches instead of 64.
constexpr uint32_t kBeamBlock = 16;

__global__ void __launch_bounds__(64)
    march_beam_kernel(const uint32_t *__restrict__ oct, uint64_t W, CamArgs cam, uint32_t width, uint32_t height,
                      uint32_t bw, uint32_t bh, float orig_sz, float *__restrict__ beam, uint32_t *__restrict__ nf_out,
                      uint32_t *err_out) {
	__shared__ uint32_t s_stack[kStackSlots * kBeamBlock];
	const uint32_t k = blockIdx.x * kBeamBlock + threadIdx.x;
	if (k >= bw * bh)
		return;
	uint32_t err = 0u, nfetch = 0u;
	beam[k] = beam_cell<kBeamBlock>(oct, W, cam, width, height, k % bw, k / bw, orig_sz, s_stack + threadIdx.x, err,
	                                nfetch);
	if (nf_out)
		nf_out[k] = nfetch;
	if (err)
		atomicOr(err_out, err);
}

// (8 WX) x (8 WY)-pixel tile per block, one 8x8 quad per wave, XCD-banded tile order.
template <uint32_t WX, uint32_t WY>
__global__ void __launch_bounds__(64 * WX * WY)
    march_frame_kernel(const uint32_t *__restrict__ oct, uint64_t W, CamArgs cam, uint32_t width, uint32_t height,
                       const float *__restrict__ beam, uint32_t bw, uint32_t tiles_x, uint4 *__restrict__ hits,
                       uint32_t *err_out) {
	constexpr uint32_t kT = 64 * WX * WY;
	__shared__ uint32_t s_stack[kStackSlots * kT];
	const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
	const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
	const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
	const uint32_t px = tx * (8u * WX) + (wave % WX) * 8u + (lane & 7u);
	const uint32_t py = ty * (8u * WY) + (wave / WX) * 8u + (lane >> 3);
	if (px >= width || py >= height)
		return;
	uint32_t err = 0u;
	uint4 r0, r1;
	frame_pixel<kT>(oct, W, cam, width, height, beam, bw, px, py, s_stack + threadIdx.x, err, r0, r1);
	const uint64_t k = (uint64_t)py * width + px;
	hits[2 * k] = r0;
	hits[2 * k + 1] = r1;
	if (err)
		atomicOr(err_out, err);
}

// ---- batches: many octrees, one launch ----------------------------------------------------------
// Beam cells of S scenes: blocks [s * bps, (s + 1) * bps) belong to scene s.
__global__ void __launch_bounds__(kBlock)
    march_beam_batch(const MarchSeg *__restrict__ segs, uint32_t width, uint32_t height, uint32_t bw, uint32_t bh,
                     uint32_t bps, float *__restrict__ beam, uint32_t *err_out) {
	__shared__ uint32_t s_stack[kStackSlots * kBlock];
	const uint32_t s = blockIdx.x / bps;
	const uint32_t k = (blockIdx.x % bps) * kBlock + threadIdx.x;
	if (k >= bw * bh)
		return;
	const MarchSeg &g = segs[s];
	uint32_t err = 0u, nfetch = 0u;
	beam[(uint64_t)s * bw * bh + k] =
	    beam_cell<kBlock>(g.oct, g.W, g.cam, width, height, k % bw, k / bw, g.orig_sz, s_stack + threadIdx.x, err, nfetch);
	if (err)
		atomicOr(err_out, err);
}

// Frames of S scenes: (8 WX) x (8 WY)-pixel tiles, one 8x8 quad per wave; the tiles of all scenes
// form one XCD-banded sequence, so an XCD works on few scenes at a time.
template <uint32_t WX, uint32_t WY>
__global__ void __launch_bounds__(64 * WX * WY)
    march_frame_batch(const MarchSeg *__restrict__ segs, uint32_t width, uint32_t height,
                      const float *__restrict__ beam, uint32_t bw, uint32_t bh, uint32_t tiles_x,
                      uint32_t tiles_per_scene, uint4 *__restrict__ hits, uint32_t *err_out) {
	constexpr uint32_t kT = 64 * WX * WY;
	__shared__ uint32_t s_stack[kStackSlots * kT];
	const uint32_t tile_g = xcd_tile(blockIdx.x, gridDim.x);
	const uint32_t s = tile_g / tiles_per_scene, tile = tile_g % tiles_per_scene;
	const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
	const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
	const uint32_t px = tx * (8u * WX) + (wave % WX) * 8u + (lane & 7u);
	const uint32_t py = ty * (8u * WY) + (wave / WX) * 8u + (lane >> 3);
	if (px >= width || py >= height)
		return;
	const MarchSeg &g = segs[s];
	uint32_t err = 0u;
	uint4 r0, r1;
	frame_pixel<kT>(g.oct, g.W, g.cam, width, height, beam ? beam + (uint64_t)s * bw * bh : nullptr, bw, px, py,
	                s_stack + threadIdx.x, err, r0, r1);
	const uint64_t k = ((uint64_t)s * height + py) * width + px;
	hits[2 * k] = r0;
	hits[2 * k + 1] = r1;
	if (err)
		atomicOr(err_out, err);
}

// Explicit rays against S octrees: rays [rbeg[s], rbeg[s+1]) go to octree s, read from and written
// to the segment's own buffers.
template <int kVariant, uint32_t kT>
__global__ void __launch_bounds__(kT)
    march_rays_batch(const MarchSeg *__restrict__ segs, const uint64_t *__restrict__ rbeg, uint32_t S, uint64_t R,
                     uint32_t *err_out) {
	__shared__ uint32_t s_stack[kStackSlots * kT];
	const uint64_t base = (uint64_t)blockIdx.x * kT, k = base + threadIdx.x;
	if (k >= R)
		return;
	const uint32_t s = S == 1u ? 0u : seg_walk(rbeg, S, seg_find(rbeg, S, base), k);
	const MarchSeg &g = segs[s];
	const uint64_t local = k - rbeg[s];
	const float *r = g.rays + 6 * local;
	f3 o{r[0], r[1], r[2]}, d{r[3], r[4], r[5]};
	uint32_t err = 0u;
	uint4 r0, r1;
	march_leaf<kVariant, kT>(g.oct, g.W, o, d, s_stack + threadIdx.x, err, r0, r1);
	g.hits[2 * local] = r0;
	g.hits[2 * local + 1] = r1;
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

// ---- persistent march with per-lane ray refill ------------------------------------------------
// Rays of one march diverge: a wave lasts as long as its longest ray while finished lanes idle.
// Here a fixed grid of waves marches all rays; a lane whose ray has ended writes its record and,
// once kRefill lanes of the wave are idle, they take the next rays of the wave's pool.  The rays
// are dealt to the waves in chunks of 64 (an 8x8-pixel tile of a frame), chunk c to wave
// c mod (number of waves), so costly image regions spread over all waves and no atomics are
// needed.  Every ray runs the same per-ray arithmetic as the one-ray-per-lane kernels, so the
// records are identical; only which lane marches which ray changes.
// A refill costs a ray set-up for the whole wave, so checks are spaced: 16 idle lanes and 16 trips
// between checks measured best (4 or 8 trips: 4-8% slower; 32 trips or 32 lanes: slower again).
constexpr uint32_t kRefill = 16; // idle lanes that trigger a refill
constexpr uint32_t kTrips = 16;  // loop trips between refill checks
constexpr uint32_t kChunk = 64;  // rays per chunk

// Ray sources: ray(k, ...) sets up ray k (false: no ray, e.g. a pixel outside the frame) and the
// address of its 32-byte record.
struct FrameRays { // primary rays of one frame, numbered in 8x8-pixel tiles
	const uint32_t *oct;
	uint64_t W;
	CamArgs cam;
	uint32_t width, height, bw, tiles_x;
	const float *beam;
	uint4 *hits;
	__device__ bool ray(uint64_t k, f3 &o, f3 &d, const uint32_t *&oc, uint64_t &w, uint4 *&out) const {
		const uint32_t tile = (uint32_t)(k >> 6), q = (uint32_t)k & 63u;
		const uint32_t px = (tile % tiles_x) * 8u + (q & 7u), py = (tile / tiles_x) * 8u + (q >> 3);
		if (px >= width || py >= height)
			return false;
		frame_ray(cam, width, height, beam, bw, px, py, o, d);
		oc = oct;
		w = W;
		out = hits + 2 * ((uint64_t)py * width + px);
		return true;
	}
};
struct ExplicitRays { // R x {o, d} against one octree
	const uint32_t *oct;
	uint64_t W;
	const float *rays;
	uint4 *hits;
	__device__ bool ray(uint64_t k, f3 &o, f3 &d, const uint32_t *&oc, uint64_t &w, uint4 *&out) const {
		const float *r = rays + 6 * k;
		o = f3{r[0], r[1], r[2]};
		d = f3{r[3], r[4], r[5]};
		oc = oct;
		w = W;
		out = hits + 2 * k;
		return true;
	}
};
struct SegmentRays { // rays [rbeg[s], rbeg[s+1]) against octree s, per-segment buffers
	const MarchSeg *segs;
	const uint64_t *rbeg;
	uint32_t S;
	__device__ bool ray(uint64_t k, f3 &o, f3 &d, const uint32_t *&oc, uint64_t &w, uint4 *&out) const {
		const uint32_t s = S == 1u ? 0u : seg_find(rbeg, S, k);
		const MarchSeg &g = segs[s];
		const uint64_t local = k - rbeg[s];
		const float *r = g.rays + 6 * local;
		o = f3{r[0], r[1], r[2]};
		d = f3{r[3], r[4], r[5]};
		oc = g.oct;
		w = g.W;
		out = g.hits + 2 * local;
		return true;
	}
};

template <int kVariant, uint32_t kT, class Src>
__global__ void __launch_bounds__(kT)
    march_persist(Src src, uint64_t n, uint32_t *err_out) {
	__shared__ uint32_t s_stack[kStackSlots * kT];
	uint32_t *stk = s_stack + threadIdx.x;
	const uint32_t lane = threadIdx.x & 63u;
	const uint64_t wave = (uint64_t)blockIdx.x * (kT / 64u) + (threadIdx.x >> 6);
	const uint64_t nwaves = (uint64_t)gridDim.x * (kT / 64u);
	const uint64_t below = (1ull << lane) - 1ull;
	March m;
	f3 o{0.0f, 0.0f, 0.0f}, d{0.0f, 0.0f, 0.0f};
	const uint32_t *oct = nullptr;
	uint64_t W = 0;
	uint4 *out = nullptr;
	uint32_t iter = 0u, nfetch = 0u, err = 0u;
	bool active = false;
	// wave-uniform: the pool [pool_next, pool_end) and the wave's next chunk
	uint64_t pool_next = 0, pool_end = 0, chunk = wave;
	bool exhausted = false;
	for (;;) {
		const uint64_t idle = __ballot(!active);
		const uint32_t nidle = (uint32_t)__popcll(idle);
		if (nidle >= kRefill && (!exhausted || pool_next < pool_end)) {
			if (pool_next == pool_end) {
				const uint64_t lo = chunk * kChunk;
				if (lo < n) {
					pool_next = lo;
					pool_end = lo + kChunk < n ? lo + kChunk : n;
					chunk += nwaves;
				} else {
					exhausted = true;
				}
			}
			if (pool_next < pool_end) {
				if (!active) {
					const uint64_t k = pool_next + (uint32_t)__popcll(idle & below);
					if (k < pool_end && src.ray(k, o, d, oct, W, out)) {
						march_init<kVariant != 0>(o, d, m);
						iter = 0u;
						nfetch = 0u;
						active = true;
					}
				}
				pool_next = pool_next + nidle < pool_end ? pool_next + nidle : pool_end;
			}
		}
		if (__ballot(active) == 0ull) {
			if (exhausted && pool_next == pool_end)
				break;
			continue;
		}
		// kTrips trips; a lane whose ray ends stops stepping and writes its record after the trips,
		// together with the other lanes that ended in them
		bool marching = active;
#pragma unroll 1
		for (uint32_t t = 0; t < kTrips; ++t)
			if (marching)
				marching = march_trip<kT>(oct, W, m, stk, iter, nfetch, err);
		if (active && !marching) {
			uint4 r0, r1;
			march_finish<kVariant>(m, o, d, iter, nfetch, r0, r1);
			out[0] = r0;
			out[1] = r1;
			active = false;
		}
	}
	if (err)
		atomicOr(err_out, err);
}

// ---- the path tracer's sampler (shader/path_tracer.comp:42,72,113-117, shader/sobol.glsl:9-10) ----
// The Sobol generator matrices of shader/sobol.comp:4-86 (64 dimensions x 32 columns) and the 256x256
// RG8 blue-noise texture of src/Noise.inl, compiled in from the package's data files.
__device__ const uint32_t g_sobol_mat[64 * 32] = {
#include "svo_sobol_matrices.inc"
};
__device__ const uint16_t g_blue_noise[256 * 256] = { // texel (x, y) at y * 256 + x: R | G << 8
#include "svo_blue_noise.inc"
};

// Sobol state of sample index n in all 64 dimensions, into out[0..63] (out[64] = n, the counter of
// shader/sobol.comp).  n advances of sobol.comp from zero XOR column findLSB(~i) in at step i,
// which is the bit where the Gray codes of i and i + 1 differ, so the state is the XOR of the
// columns of the set bits of gray(n) = n ^ (n >> 1).  One block of 64 threads.
__global__ void __launch_bounds__(64) sobol_point_kernel(uint32_t n, uint32_t *__restrict__ out) {
	const uint32_t d = threadIdx.x, g = n ^ (n >> 1);
	uint32_t x = 0u;
	for (uint32_t c = 0; c < 32u; ++c)
		if ((g >> c) & 1u)
			x ^= g_sobol_mat[d * 32u + c];
	out[d] = x;
	if (d == 0u)
		out[64] = n;
}

// Sobol_GetVec2 component: uint / 2^32.
__device__ __forceinline__ float sobol_unit(uint32_t x) { return (float)x / 4294967296.0f; }
// texelFetch(uNoise, pixel & 0xff): both UNORM8 channels of the texel, packed R | G << 8.
__device__ __forceinline__ uint32_t noise_texel(uint32_t px, uint32_t py) {
	return g_blue_noise[((py & 255u) << 8) | (px & 255u)];
}
// Bounce cur's sample: fract(Sobol_GetVec2(cur) + noise) (shader/path_tracer.comp:42,72).
__device__ __forceinline__ void bounce_sample(const uint32_t *__restrict__ sob, uint32_t cur, uint32_t nz, float &u,
                                              float &v) {
	const float s = sobol_unit(sob[2u * cur]) + (float)(nz & 0xffu) / 255.0f;
	const float t = sobol_unit(sob[2u * cur + 1u]) + (float)(nz >> 8) / 255.0f;
	u = s - floorf(s);
	v = t - floorf(t);
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

// Diffuse_Sample + AlignDirection (shader/diffuse.glsl:7-14, shader/util.glsl:5-9) around the axis
// normal of face code `code` (axis * 2 + positive), for the sample (u, v).
__device__ __forceinline__ f3 diffuse_dir(uint32_t code, float u, float v, f3 &nrm) {
	const float kPi = 3.1415926535897932384626433832795f; // shader/util.glsl:4
	const float sg = (code & 1u) ? 1.0f : -1.0f;
	nrm = f3{code >> 1 == 0u ? sg : 0.0f, code >> 1 == 1u ? sg : 0.0f, code >> 1 == 2u ? sg : 0.0f};
	float rr = __builtin_sqrtf(u), phi = 2.0f * kPi * v;
	float sn, cs;
	sincos_poly(phi, sn, cs);
	f3 dl{rr * cs, rr * sn, __builtin_sqrtf(1.0f - u)};
	f3 ax = fabsf(nrm.x) > 0.01f ? f3{0.0f, 1.0f, 0.0f} : f3{1.0f, 0.0f, 0.0f};
	f3 uu = normalize3(cross3(ax, nrm));
	f3 vv = cross3(nrm, uu);
	return f3{dl.x * uu.x + dl.y * vv.x + dl.z * nrm.x, dl.x * uu.y + dl.y * vv.y + dl.z * nrm.y,
	          dl.x * uu.z + dl.y * vv.z + dl.z * nrm.z};
}

// Bounce cur of sample index spp for every primary hit of a width-wide frame (ray k = pixel
// (k % width, k / width)): o = o_pos, d = Diffuse_Sample(normal, fract(Sobol_GetVec2(cur) + noise))
// (shader/path_tracer.comp:42-44).  A miss gets o = 0, d = -1.
__global__ void __launch_bounds__(kBlock) bounce_kernel(const uint4 *__restrict__ hits, uint64_t R, uint32_t width,
                                                        const uint32_t *__restrict__ sob, uint32_t cur,
                                                        float *__restrict__ rays) {
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
	const uint32_t px = (uint32_t)(k % width), py = (uint32_t)(k / width);
	float u, v;
	bounce_sample(sob, cur, noise_texel(px, py), u, v);
	f3 nrm;
	const f3 d = diffuse_dir((h1.y >> 24) & 7u, u, v, nrm);
	r[0] = __uint_as_float(h0.x);
	r[1] = __uint_as_float(h0.y);
	r[2] = __uint_as_float(h0.z);
	r[3] = d.x;
	r[4] = d.y;
	r[5] = d.z;
}

// ---- progressive path tracing (ConstColorPathTrace + accumulation, shader/path_tracer.comp:32-60,
// 111-132) --------------------------------------------------------------------------------------

struct PathArgs {
	CamArgs cam;
	uint32_t width, height, bounce, spp;
	float sky[3];
	const uint32_t *sobol; // Sobol state of sample spp (sobol_point_kernel)
	uint32_t tiles_x;
	float4 *color;             // RGBA32F running mean
	uint32_t *albedo, *normal; // RGBA8 UNORM / RGBA8 SNORM running means
};

// The steps of one path, shared by both schedules so that they give the same bits.
// Primary ray of pixel (px, py): through pixel + Sobol_GetVec2(0) (shader/path_tracer.comp:117).
__device__ __forceinline__ void path_primary(const PathArgs &a, uint32_t px, uint32_t py, f3 &o, f3 &d) {
	const float jx = sobol_unit(a.sobol[0]), jy = sobol_unit(a.sobol[1]);
	o = f3{a.cam.px, a.cam.py, a.cam.pz};
	d = gen_ray(a.cam, ((float)px + jx) / (float)a.width, ((float)py + jy) / (float)a.height);
}
// Bounce cur >= 1 from a hit record: origin o_pos, cosine-weighted direction about the hit face's
// normal for the sample fract(Sobol_GetVec2(cur) + noise); nz = the pixel's noise texel.
__device__ __forceinline__ void path_bounce(const PathArgs &a, uint32_t nz, uint32_t cur, uint4 h0, uint32_t code,
                                            f3 &o, f3 &d) {
	float u, v;
	bounce_sample(a.sobol, cur, nz, u, v);
	f3 nrm;
	o = f3{__uint_as_float(h0.x), __uint_as_float(h0.y), __uint_as_float(h0.z)};
	d = diffuse_dir(code, u, v, nrm);
}
__device__ __forceinline__ f3 leaf_albedo(uint32_t leaf) { // unpackUnorm4x8: c / 255
	const float k255 = 255.0f;
	return f3{(float)(leaf & 0xffu) / k255, (float)((leaf >> 8) & 0xffu) / k255, (float)((leaf >> 16) & 0xffu) / k255};
}
__device__ __forceinline__ f3 face_normal(uint32_t code) {
	const float sg = (code & 1u) ? 1.0f : -1.0f;
	return f3{code >> 1 == 0u ? sg : 0.0f, code >> 1 == 1u ? sg : 0.0f, code >> 1 == 2u ? sg : 0.0f};
}
// Running mean of one pixel (shader/path_tracer.comp:123-131): (old * spp + v) / (spp + 1); spp 0
// starts it (the reference clears its images first).
__device__ __forceinline__ void path_mean(const PathArgs &a, float4 *img, uint32_t k, f3 v) {
	if (!img)
		return;
	if (a.spp == 0u) {
		img[k] = make_float4(v.x, v.y, v.z, 1.0f);
		return;
	}
	const float n = (float)a.spp, n1 = (float)(a.spp + 1u);
	const float4 p = img[k];
	img[k] = make_float4((p.x * n + v.x) / n1, (p.y * n + v.y) / n1, (p.z * n + v.z) / n1, 1.0f);
}
// The same mean through an RGBA8 texel (src/PathTracer.cpp:13-16): imageLoad gives byte / 255
// (UNORM) or max(signed byte / 127, -1) (SNORM); imageStore rounds the clamped value times 255 (127)
// to nearest, halves up; alpha stores 1.
template <bool kSnorm> __device__ __forceinline__ float rgba8_load(uint32_t b) {
	if (kSnorm) {
		const float f = (float)(int32_t)(int8_t)(uint8_t)(b & 0xffu) / 127.0f;
		return f < -1.0f ? -1.0f : f;
	}
	return (float)(b & 0xffu) / 255.0f;
}
template <bool kSnorm> __device__ __forceinline__ uint32_t rgba8_store(float f) {
	if (kSnorm) {
		f = f < -1.0f ? -1.0f : (f > 1.0f ? 1.0f : f);
		return (uint32_t)((int32_t)floorf(f * 127.0f + 0.5f) & 0xff);
	}
	f = f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f);
	return (uint32_t)floorf(f * 255.0f + 0.5f);
}
template <bool kSnorm> __device__ __forceinline__ void path_mean8(const PathArgs &a, uint32_t *img, uint32_t k, f3 v) {
	if (!img)
		return;
	const uint32_t old = a.spp == 0u ? 0u : img[k];
	const float n = (float)a.spp, n1 = (float)(a.spp + 1u);
	const float m0 = (rgba8_load<kSnorm>(old) * n + v.x) / n1, m1 = (rgba8_load<kSnorm>(old >> 8) * n + v.y) / n1,
	            m2 = (rgba8_load<kSnorm>(old >> 16) * n + v.z) / n1;
	img[k] = rgba8_store<kSnorm>(m0) | rgba8_store<kSnorm>(m1) << 8 | rgba8_store<kSnorm>(m2) << 16 |
	         (kSnorm ? 127u : 255u) << 24;
}

// Sample spp of every pixel of a progressive render (ConstColorPathTrace + main,
// shader/path_tracer.comp:32-60,111-132): the camera ray through pixel + Sobol_GetVec2(0) (the pixel
// corner at spp 0), then up to bounce-1 cosine-weighted diffuse bounces sampled by
// fract(Sobol_GetVec2(cur) + blue noise); a path that escapes brings the constant sky colour times
// the product of the albedos it met.  The colour (RGBA32F), first-hit albedo (RGBA8 UNORM) and
// first-hit normal (RGBA8 SNORM) are running means over spp + 1 samples.  This schedule: one path
// per lane, 8x8-pixel tiles.
__global__ void __launch_bounds__(64) path_trace_kernel(const uint32_t *__restrict__ oct, uint64_t W, PathArgs a,
                                                        uint32_t *err_out) {
	__shared__ uint32_t s_stack[kStackSlots * 64];
	const uint32_t tx = blockIdx.x % a.tiles_x, ty = blockIdx.x / a.tiles_x, lane = threadIdx.x;
	const uint32_t px = tx * 8u + (lane & 7u), py = ty * 8u + (lane >> 3);
	if (px >= a.width || py >= a.height)
		return;
	uint32_t err = 0u;
	const uint32_t k = py * a.width + px, nz = noise_texel(px, py);
	f3 o, d;
	path_primary(a, px, py, o, d);
	uint4 h0, h1;
	march_leaf<1, 64>(oct, W, o, d, s_stack + lane, err, h0, h1);
	f3 rad{a.sky[0], a.sky[1], a.sky[2]}, alb0{0.0f, 0.0f, 0.0f}, nrm0{0.0f, 0.0f, 0.0f};
	if (h1.y & 0x80000000u) {
		f3 acc = leaf_albedo(h1.z);
		alb0 = acc;
		uint32_t code = (h1.y >> 24) & 7u;
		nrm0 = face_normal(code);
		rad = f3{0.0f, 0.0f, 0.0f};
		for (uint32_t cur = 1; cur < a.bounce; ++cur) {
			path_bounce(a, nz, cur, h0, code, o, d);
			march_leaf<1, 64>(oct, W, o, d, s_stack + lane, err, h0, h1);
			if (!(h1.y & 0x80000000u)) {
				rad = f3{acc.x * a.sky[0], acc.y * a.sky[1], acc.z * a.sky[2]};
				break;
			}
			const f3 al = leaf_albedo(h1.z);
			acc = f3{acc.x * al.x, acc.y * al.y, acc.z * al.z};
			code = (h1.y >> 24) & 7u;
		}
	}
	path_mean(a, a.color, k, rad);
	path_mean8<false>(a, a.albedo, k, alb0);
	path_mean8<true>(a, a.normal, k, nrm0);
	if (err)
		atomicOr(err_out, err);
}

// ---- environment-map light (EnvMapPathTrace, shader/path_tracer.comp:62-106;
// shader/environment_map.glsl; shader/mis.glsl) -----------------------------------------------
// atan on [0, 1]: atan(t) = pi/4 + atan((t-1)/(t+1)) above tan(pi/8), then an odd polynomial
// x + x^3 P(x^2) fitted (least squares, float coefficients) to atan on [0, tan(pi/8)]; max error
// 9e-8.  Own function, identical on host and device, since libm and device math differ.
__device__ __forceinline__ float atan01(float t) {
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
__device__ __forceinline__ float atan2_poly(float y, float x) { // GLSL atan(y, x); atan2(0, 0) = 0
	const float ax = fabsf(x), ay = fabsf(y), mx = gmax(ax, ay), mn = gmin(ax, ay);
	float r = atan01(mx > 0.0f ? mn / mx : 0.0f);
	r = ay > ax ? 0x1.921fb6p+0f - r : r;
	r = x < 0.0f ? 0x1.921fb6p+1f - r : r;
	return y < 0.0f ? -r : r;
}
__device__ __forceinline__ float acos_poly(float x) { return atan2_poly(__builtin_sqrtf(gmax(1.0f - x * x, 0.0f)), x); }

struct EnvArgs {
	const float4 *texels; // rgb + pdf
	const uint2 *alias;   // {prob, alias}
	uint32_t width, height;
	float rotation, multiplier;
};

// texture(): bilinear, repeat addressing in both directions (src/EnvironmentMap.cpp:10), in fp32.
__device__ __forceinline__ float4 env_texture(const EnvArgs &e, float u, float v) {
	const float x = u * (float)e.width - 0.5f, y = v * (float)e.height - 0.5f;
	const float fx = floorf(x), fy = floorf(y);
	const float a = x - fx, b = y - fy;
	auto wrap = [](float f, uint32_t n) -> uint32_t {
		if (!(f > -1e9f && f < 1e9f))
			return 0u; // not a finite coordinate
		const int64_t i = (int64_t)f % (int64_t)n;
		return (uint32_t)(i < 0 ? i + n : i);
	};
	const uint32_t i0 = wrap(fx, e.width), i1 = wrap(fx + 1.0f, e.width);
	const uint32_t j0 = wrap(fy, e.height), j1 = wrap(fy + 1.0f, e.height);
	const float4 t00 = e.texels[j0 * e.width + i0], t10 = e.texels[j0 * e.width + i1];
	const float4 t01 = e.texels[j1 * e.width + i0], t11 = e.texels[j1 * e.width + i1];
	const float a1 = 1.0f - a, b1 = 1.0f - b;
	auto mix2 = [&](float p00, float p10, float p01, float p11) {
		return (p00 * a1 + p10 * a) * b1 + (p01 * a1 + p11 * a) * b;
	};
	return make_float4(mix2(t00.x, t10.x, t01.x, t11.x), mix2(t00.y, t10.y, t01.y, t11.y),
	                   mix2(t00.z, t10.z, t01.z, t11.z), mix2(t00.w, t10.w, t01.w, t11.w));
}
__device__ __forceinline__ float4 env_lookup(const EnvArgs &e, f3 d) { // EnvMap_Dir2Coord + texture
	const float kPi = 0x1.921fb6p+1f;
	const float u = ((atan2_poly(d.x, d.z) + e.rotation) / kPi) * 0.5f, v = acos_poly(d.y) / kPi;
	return env_texture(e, u, v);
}
__device__ __forceinline__ float env_pdf(const EnvArgs &e, f3 d) { // EnvMap_PDF
	const float kTwoPiPi = (2.0f * 0x1.921fb6p+1f) * 0x1.921fb6p+1f;
	return env_lookup(e, d).w / (kTwoPiPi * __builtin_sqrtf(1.0f - d.y * d.y));
}
// EnvMap_Sample: texel by the alias table, its centre's direction, radiance and pdf.
__device__ __forceinline__ f3 env_sample(const EnvArgs &e, float su, float sv, f3 &light, float &pdf) {
	const float kPi = 0x1.921fb6p+1f, kTwoPiPi = (2.0f * kPi) * kPi;
	const uint32_t sz = e.width * e.height;
	uint32_t idx = (uint32_t)(su * (float)sz);
	idx = idx < sz - 1u ? idx : sz - 1u;
	const uint32_t prob = (uint32_t)(sv * 4294967296.0f);
	const uint2 al = e.alias[idx];
	const uint32_t p = prob > al.x ? al.y : idx;
	const float ux = ((float)(p % e.width) + 0.5f) / (float)e.width, uy = ((float)(p / e.width) + 0.5f) / (float)e.height;
	const float phi = (2.0f * kPi) * ux - e.rotation, theta = kPi * uy;
	float st, ct, sp, cp;
	sincos_poly(theta, st, ct);
	sincos_poly(phi, sp, cp);
	const float4 t = e.texels[p];
	light = f3{t.x, t.y, t.z};
	pdf = t.w / (kTwoPiPi * st);
	return f3{st * sp, ct, st * cp};
}
__device__ __forceinline__ float mis_power(float a, float b) { // MIS_PowerHeuristic
	const float t = a * a;
	return t / (b * b + t);
}

// EnvMapPathTrace per pixel, one path per lane in 8x8 tiles (the same sampling as the constant-sky
// paths; the light sample reuses the bounce's 2-D point, as the reference does).
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8))) path_trace_env_kernel(const uint32_t *__restrict__ oct, uint64_t W, PathArgs a,
                                                            EnvArgs e, uint32_t *err_out) {
	__shared__ uint32_t s_stack[kStackSlots * 64];
	const uint32_t tx = blockIdx.x % a.tiles_x, ty = blockIdx.x / a.tiles_x, lane = threadIdx.x;
	const uint32_t px = tx * 8u + (lane & 7u), py = ty * 8u + (lane >> 3);
	if (px >= a.width || py >= a.height)
		return;
	const float kPi = 0x1.921fb6p+1f, kBsdf = 1.0f / kPi;
	uint32_t err = 0u;
	const uint32_t k = py * a.width + px, nz = noise_texel(px, py);
	f3 o, d;
	path_primary(a, px, py, o, d);
	uint4 h0, h1;
	march_leaf<1, 64>(oct, W, o, d, s_stack + lane, err, h0, h1);
	f3 rad{0.0f, 0.0f, 0.0f}, alb0{0.0f, 0.0f, 0.0f}, nrm0{0.0f, 0.0f, 0.0f};
	if (!(h1.y & 0x80000000u)) {
		const float4 t = env_lookup(e, d);
		rad = f3{t.x, t.y, t.z};
	} else {
		f3 acc = leaf_albedo(h1.z);
		alb0 = acc;
		uint32_t code = (h1.y >> 24) & 7u;
		nrm0 = face_normal(code);
		for (uint32_t cur = 1; cur < a.bounce; ++cur) {
			float u, v;
			bounce_sample(a.sobol, cur, nz, u, v);
			o = f3{__uint_as_float(h0.x), __uint_as_float(h0.y), __uint_as_float(h0.z)};
			const f3 nrm = face_normal(code);
			{ // light sample
				f3 light;
				float lpdf;
				const f3 dl = env_sample(e, u, v, light, lpdf);
				const float ndd = (nrm.x * dl.x + nrm.y * dl.y) + nrm.z * dl.z;
				if (ndd > 0.0f && !__builtin_isnan(lpdf) && lpdf > 0.0f) {
					uint4 q0, q1;
					march_leaf<2, 64>(oct, W, o, dl, s_stack + lane, err, q0, q1);
					if (!(q1.y & 0x80000000u)) {
						const float bpdf = ndd / kPi, w = mis_power(lpdf, bpdf);
						rad = f3{rad.x + ((((acc.x * light.x) * e.multiplier) * w) * ndd) * kBsdf / lpdf,
						         rad.y + ((((acc.y * light.y) * e.multiplier) * w) * ndd) * kBsdf / lpdf,
						         rad.z + ((((acc.z * light.z) * e.multiplier) * w) * ndd) * kBsdf / lpdf};
					}
				}
			}
			f3 n2;
			d = diffuse_dir(code, u, v, n2);
			const float bpdf = __builtin_sqrtf(1.0f - u) / kPi;
			march_leaf<1, 64>(oct, W, o, d, s_stack + lane, err, h0, h1);
			if (!(h1.y & 0x80000000u)) {
				const float lpdf = env_pdf(e, d);
				if (!__builtin_isnan(lpdf) && lpdf > 0.0f) {
					const float4 t = env_lookup(e, d);
					const float w = mis_power(bpdf, lpdf);
					rad = f3{rad.x + ((acc.x * t.x) * e.multiplier) * w, rad.y + ((acc.y * t.y) * e.multiplier) * w,
					         rad.z + ((acc.z * t.z) * e.multiplier) * w};
				}
				break;
			}
			const f3 al = leaf_albedo(h1.z);
			acc = f3{acc.x * al.x, acc.y * al.y, acc.z * al.z};
			code = (h1.y >> 24) & 7u;
		}
	}
	path_mean(a, a.color, k, rad);
	path_mean8<false>(a, a.albedo, k, alb0);
	path_mean8<true>(a, a.normal, k, nrm0);
	if (err)
		atomicOr(err_out, err);
}

// The same paths on a persistent grid (march_persist's refill scheme): a lane whose path ends writes
// its pixel and takes the next pixel of its wave's pool; a path that goes on starts its next ray
// in place.  Pixels are numbered in 8x8 tiles and dealt to the waves in 64-pixel chunks.
template <uint32_t kT>
__global__ void __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(5))) path_persist(const uint32_t *__restrict__ oct, uint64_t W, PathArgs a, uint64_t n,
                                                   uint32_t *err_out) {
	__shared__ uint32_t s_stack[kStackSlots * kT];
	uint32_t *stk = s_stack + threadIdx.x;
	const uint32_t lane = threadIdx.x & 63u;
	const uint64_t wave = (uint64_t)blockIdx.x * (kT / 64u) + (threadIdx.x >> 6);
	const uint64_t nwaves = (uint64_t)gridDim.x * (kT / 64u);
	const uint64_t below = (1ull << lane) - 1ull;
	March m;
	f3 o{0.0f, 0.0f, 0.0f}, d{0.0f, 0.0f, 0.0f}, acc{0.0f, 0.0f, 0.0f};
	uint32_t k = 0u, nz = 0u, cur = 0u, iter = 0u, nfetch = 0u, err = 0u;
	bool active = false;
	uint64_t pool_next = 0, pool_end = 0, chunk = wave;
	bool exhausted = false;
	for (;;) {
		const uint64_t idle = __ballot(!active);
		const uint32_t nidle = (uint32_t)__popcll(idle);
		if (nidle >= kRefill && (!exhausted || pool_next < pool_end)) {
			if (pool_next == pool_end) {
				const uint64_t lo = chunk * kChunk;
				if (lo < n) {
					pool_next = lo;
					pool_end = lo + kChunk < n ? lo + kChunk : n;
					chunk += nwaves;
				} else {
					exhausted = true;
				}
			}
			if (pool_next < pool_end) {
				if (!active) {
					const uint64_t q = pool_next + (uint32_t)__popcll(idle & below);
					const uint32_t tile = (uint32_t)(q >> 6), r = (uint32_t)q & 63u;
					const uint32_t px = (tile % a.tiles_x) * 8u + (r & 7u), py = (tile / a.tiles_x) * 8u + (r >> 3);
					if (q < pool_end && px < a.width && py < a.height) {
						k = py * a.width + px;
						nz = noise_texel(px, py);
						cur = 0u;
						path_primary(a, px, py, o, d);
						march_init<true>(o, d, m);
						iter = nfetch = 0u;
						active = true;
					}
				}
				pool_next = pool_next + nidle < pool_end ? pool_next + nidle : pool_end;
			}
		}
		if (__ballot(active) == 0ull) {
			if (exhausted && pool_next == pool_end)
				break;
			continue;
		}
		bool marching = active;
#pragma unroll 1
		for (uint32_t t = 0; t < kTrips; ++t)
			if (marching)
				marching = march_trip<kT>(oct, W, m, stk, iter, nfetch, err);
		if (active && !marching) {
			uint4 h0, h1;
			march_finish<1>(m, o, d, iter, nfetch, h0, h1);
			const bool hit = (h1.y & 0x80000000u) != 0u;
			bool done = true;
			f3 rad{0.0f, 0.0f, 0.0f};
			if (cur == 0u) {
				f3 alb0{0.0f, 0.0f, 0.0f}, nrm0{0.0f, 0.0f, 0.0f};
				if (hit) {
					acc = alb0 = leaf_albedo(h1.z);
					nrm0 = face_normal((h1.y >> 24) & 7u);
					done = a.bounce <= 1u;
				} else {
					rad = f3{a.sky[0], a.sky[1], a.sky[2]};
				}
				path_mean8<false>(a, a.albedo, k, alb0);
				path_mean8<true>(a, a.normal, k, nrm0);
			} else if (!hit) {
				rad = f3{acc.x * a.sky[0], acc.y * a.sky[1], acc.z * a.sky[2]};
			} else {
				const f3 al = leaf_albedo(h1.z);
				acc = f3{acc.x * al.x, acc.y * al.y, acc.z * al.z};
				done = cur + 1u >= a.bounce;
			}
			if (done) {
				path_mean(a, a.color, k, rad);
				active = false;
			} else { // the next bounce of the same path
				++cur;
				path_bounce(a, nz, cur, h0, (h1.y >> 24) & 7u, o, d);
				march_init<true>(o, d, m);
				iter = nfetch = 0u;
			}
		}
	}
	if (err)
		atomicOr(err_out, err);
}

// The environment-map paths on the persistent grid of path_persist.  A bounce takes up to two
// marches of one lane: the occlusion ray towards the light sample (only when the sample faces the
// surface; its weighted radiance waits in `pend` until the ray is known to escape), then the BSDF
// ray.  Every value is computed with the operations of path_trace_env_kernel, in the same order.
constexpr uint32_t kPersistEnvBlocksPerCu = 4; // 120 VGPRs: 4 waves of 64 per SIMD
template <uint32_t kT>
__global__ void __launch_bounds__(kT) path_persist_env(const uint32_t *__restrict__ oct, uint64_t W, PathArgs a,
                                                       EnvArgs e, uint64_t n, uint32_t *err_out) {
	__shared__ uint32_t s_stack[kStackSlots * kT];
	uint32_t *stk = s_stack + threadIdx.x;
	const uint32_t lane = threadIdx.x & 63u;
	const uint64_t wave = (uint64_t)blockIdx.x * (kT / 64u) + (threadIdx.x >> 6);
	const uint64_t nwaves = (uint64_t)gridDim.x * (kT / 64u);
	const uint64_t below = (1ull << lane) - 1ull;
	const float kPi = 0x1.921fb6p+1f, kBsdf = 1.0f / kPi;
	March m;
	f3 o{0.0f, 0.0f, 0.0f}, d{0.0f, 0.0f, 0.0f}, acc{0.0f, 0.0f, 0.0f}, rad{0.0f, 0.0f, 0.0f}, pend{0.0f, 0.0f, 0.0f};
	float u = 0.0f, v = 0.0f, bpdf = 0.0f;
	uint32_t k = 0u, nz = 0u, cur = 0u, code = 0u, iter = 0u, nfetch = 0u, err = 0u;
	bool active = false, shadow = false;
	uint64_t pool_next = 0, pool_end = 0, chunk = wave;
	bool exhausted = false;
	for (;;) {
		const uint64_t idle = __ballot(!active);
		const uint32_t nidle = (uint32_t)__popcll(idle);
		if (nidle >= kRefill && (!exhausted || pool_next < pool_end)) {
			if (pool_next == pool_end) {
				const uint64_t lo = chunk * kChunk;
				if (lo < n) {
					pool_next = lo;
					pool_end = lo + kChunk < n ? lo + kChunk : n;
					chunk += nwaves;
				} else {
					exhausted = true;
				}
			}
			if (pool_next < pool_end) {
				if (!active) {
					const uint64_t q = pool_next + (uint32_t)__popcll(idle & below);
					const uint32_t tile = (uint32_t)(q >> 6), r = (uint32_t)q & 63u;
					const uint32_t px = (tile % a.tiles_x) * 8u + (r & 7u), py = (tile / a.tiles_x) * 8u + (r >> 3);
					if (q < pool_end && px < a.width && py < a.height) {
						k = py * a.width + px;
						nz = noise_texel(px, py);
						cur = 0u;
						shadow = false;
						rad = f3{0.0f, 0.0f, 0.0f};
						path_primary(a, px, py, o, d);
						march_init<true>(o, d, m);
						iter = nfetch = 0u;
						active = true;
					}
				}
				pool_next = pool_next + nidle < pool_end ? pool_next + nidle : pool_end;
			}
		}
		if (__ballot(active) == 0ull) {
			if (exhausted && pool_next == pool_end)
				break;
			continue;
		}
		bool marching = active;
#pragma unroll 1
		for (uint32_t t = 0; t < kTrips; ++t)
			if (marching)
				marching = march_trip<kT>(oct, W, m, stk, iter, nfetch, err);
		if (!active || marching)
			continue;
		uint4 h0, h1;
		if (shadow) { // occlusion ray done: keep the light term if it escaped, then the BSDF ray
			march_finish<2>(m, o, d, iter, nfetch, h0, h1);
			if (!(h1.y & 0x80000000u))
				rad = f3{rad.x + pend.x, rad.y + pend.y, rad.z + pend.z};
			shadow = false;
			f3 n2;
			d = diffuse_dir(code, u, v, n2);
			bpdf = __builtin_sqrtf(1.0f - u) / kPi;
			march_init<true>(o, d, m);
			iter = nfetch = 0u;
			continue;
		}
		march_finish<1>(m, o, d, iter, nfetch, h0, h1);
		const bool hit = (h1.y & 0x80000000u) != 0u;
		bool done = true;
		if (cur == 0u) {
			f3 alb0{0.0f, 0.0f, 0.0f}, nrm0{0.0f, 0.0f, 0.0f};
			if (hit) {
				acc = alb0 = leaf_albedo(h1.z);
				code = (h1.y >> 24) & 7u;
				nrm0 = face_normal(code);
				done = a.bounce <= 1u;
			} else {
				const float4 t = env_lookup(e, d);
				rad = f3{t.x, t.y, t.z};
			}
			path_mean8<false>(a, a.albedo, k, alb0);
			path_mean8<true>(a, a.normal, k, nrm0);
		} else if (!hit) {
			const float lpdf = env_pdf(e, d);
			if (!__builtin_isnan(lpdf) && lpdf > 0.0f) {
				const float4 t = env_lookup(e, d);
				const float w = mis_power(bpdf, lpdf);
				rad = f3{rad.x + ((acc.x * t.x) * e.multiplier) * w, rad.y + ((acc.y * t.y) * e.multiplier) * w,
				         rad.z + ((acc.z * t.z) * e.multiplier) * w};
			}
		} else {
			const f3 al = leaf_albedo(h1.z);
			acc = f3{acc.x * al.x, acc.y * al.y, acc.z * al.z};
			code = (h1.y >> 24) & 7u;
			done = cur + 1u >= a.bounce;
		}
		if (done) {
			path_mean(a, a.color, k, rad);
			active = false;
			continue;
		}
		// bounce cur + 1 from this hit: the light sample's occlusion ray first when it faces the face
		++cur;
		bounce_sample(a.sobol, cur, nz, u, v);
		o = f3{__uint_as_float(h0.x), __uint_as_float(h0.y), __uint_as_float(h0.z)};
		const f3 nrm = face_normal(code);
		f3 light;
		float lpdf;
		const f3 dl = env_sample(e, u, v, light, lpdf);
		const float ndd = (nrm.x * dl.x + nrm.y * dl.y) + nrm.z * dl.z;
		if (ndd > 0.0f && !__builtin_isnan(lpdf) && lpdf > 0.0f) {
			const float bp = ndd / kPi, w = mis_power(lpdf, bp);
			pend = f3{((((acc.x * light.x) * e.multiplier) * w) * ndd) * kBsdf / lpdf,
			          ((((acc.y * light.y) * e.multiplier) * w) * ndd) * kBsdf / lpdf,
			          ((((acc.z * light.z) * e.multiplier) * w) * ndd) * kBsdf / lpdf};
			d = dl;
			shadow = true;
		} else {
			f3 n2;
			d = diffuse_dir(code, u, v, n2);
			bpdf = __builtin_sqrtf(1.0f - u) / kPi;
		}
		march_init<true>(o, d, m);
		iter = nfetch = 0u;
	}
	if (err)
		atomicOr(err_out, err);
}

// ---- view-type shading of a primary frame (shader/octree_tracer.frag:18-47) ------------------------
// oColor of the reference's tracer from the hit records: view 0 = pow(albedo, 1/2.2), 1 = normal *
// 0.5 + 0.5, 2 = pos - 1, 3 = Heat(iter / 128); a miss has pos = 1, normal = 0 and the light's
// radiance in its direction as colour (the constant colour, or the environment map without its
// multiplier).  Stored as RGBA8 UNORM (alpha 255), rounded to nearest.
struct ShadeArgs {
	CamArgs cam;
	uint32_t width, height, view;
	float light[3];
	const float4 *env_texels; // null: the constant colour
	uint32_t env_w, env_h;
	float env_rotation;
};
__device__ __forceinline__ uint32_t unorm8_store(float f) {
	f = f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f);
	return (uint32_t)floorf(f * 255.0f + 0.5f);
}
__global__ void __launch_bounds__(kBlock) shade_kernel(const uint4 *__restrict__ hits, ShadeArgs a,
                                                       uint32_t *__restrict__ out) {
	const uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
	if (k >= (uint64_t)a.width * a.height)
		return;
	const uint4 h0 = hits[2 * k], h1 = hits[2 * k + 1];
	const bool hit = (h1.y & 0x80000000u) != 0u;
	float c[3];
	if (a.view == 3u) { // Heat(x) = sin(clamp(x, 0, 1) * 3 - vec3(1, 2, 3)) * 0.5 + 0.5
		float x = (float)h0.w / 128.0f;
		x = x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
		for (int i = 0; i < 3; ++i) {
			float sn, cs;
			sincos_poly(x * 3.0f - (float)(i + 1), sn, cs);
			c[i] = sn * 0.5f + 0.5f;
		}
	} else if (a.view == 2u) { // pos - 1.0
		c[0] = hit ? __uint_as_float(h0.x) - 1.0f : 0.0f;
		c[1] = hit ? __uint_as_float(h0.y) - 1.0f : 0.0f;
		c[2] = hit ? __uint_as_float(h0.z) - 1.0f : 0.0f;
	} else if (a.view == 1u) { // normal * 0.5 + 0.5
		const f3 n = hit ? face_normal((h1.y >> 24) & 7u) : f3{0.0f, 0.0f, 0.0f};
		c[0] = n.x * 0.5f + 0.5f, c[1] = n.y * 0.5f + 0.5f, c[2] = n.z * 0.5f + 0.5f;
	} else { // pow(color, 1 / 2.2)
		f3 col;
		if (hit) {
			col = leaf_albedo(h1.z);
		} else if (a.env_texels) {
			const uint32_t px = (uint32_t)(k % a.width), py = (uint32_t)(k / a.width);
			const f3 d = gen_ray(a.cam, (float)px / (float)a.width, (float)py / (float)a.height);
			const EnvArgs e{a.env_texels, nullptr, a.env_w, a.env_h, a.env_rotation, 1.0f};
			const float4 t = env_lookup(e, d);
			col = f3{t.x, t.y, t.z};
		} else {
			col = f3{a.light[0], a.light[1], a.light[2]};
		}
		const float g = 1.0f / 2.2f;
		c[0] = powf(col.x, g), c[1] = powf(col.y, g), c[2] = powf(col.z, g);
	}
	out[k] = unorm8_store(c[0]) | unorm8_store(c[1]) << 8 | unorm8_store(c[2]) << 16 | 0xFF000000u;
}

CamArgs cam_args(const svo_camera *c) {
	return CamArgs{c->pos[0],  c->pos[1],  c->pos[2],  c->look[0], c->look[1], c->look[2],
	               c->side[0], c->side[1], c->side[2], c->up[0],   c->up[1],   c->up[2]};
}

// Times fn(shape) for every candidate (one warm-up, best of two timed runs) and returns the fastest;
// the decision is kept per key and logged.
template <class Fn>
int tune(svo_ctx *ctx, const std::string &key, int nshapes, const char *const *names, Fn fn, int &choice) {
	auto it = ctx->tune_choice.find(key);
	if (it != ctx->tune_choice.end()) {
		choice = it->second;
		return SVO_OK;
	}
	hipEvent_t a = nullptr, b = nullptr;
	hipError_t e = hipEventCreate(&a);
	if (e == hipSuccess)
		e = hipEventCreate(&b);
	float best = 3.4e38f;
	int bi = 0;
	std::string line = key + ":";
	for (int sh = 0; sh < nshapes && e == hipSuccess; ++sh) {
		if ((e = fn(sh)) != hipSuccess)
			break;
		float tmin = 3.4e38f;
		for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
			if ((e = hipEventRecord(a, ctx->stream)) != hipSuccess || (e = fn(sh)) != hipSuccess ||
			    (e = hipEventRecord(b, ctx->stream)) != hipSuccess || (e = hipEventSynchronize(b)) != hipSuccess)
				break;
			float ms = 0.0f;
			if ((e = hipEventElapsedTime(&ms, a, b)) == hipSuccess)
				tmin = std::min(tmin, ms);
		}
		char buf[64];
		std::snprintf(buf, sizeof buf, " %s %.4f ms", names[sh], (double)tmin);
		line += buf;
		if (tmin < best) {
			best = tmin;
			bi = sh;
		}
	}
	if (a)
		(void)hipEventDestroy(a);
	if (b)
		(void)hipEventDestroy(b);
	if (e != hipSuccess)
		return hip_check(ctx, e, "autotune");
	ctx->tune_choice[key] = bi;
	ctx->tune_log += line + " -> " + names[bi] + "\n";
	choice = bi;
	return SVO_OK;
}

const char *const kScheduleName[3] = {"one-ray-per-lane", "persistent-refill", "one-ray-per-lane-8x8"};

// Persistent march of n rays from src; name is the profiler's name of the march.
template <int kVariant, class Src>
hipError_t launch_persist(svo_ctx *ctx, const char *name, const Src &src, uint64_t n) {
	uint64_t grid = (n + kBlock - 1) / kBlock;
	const uint64_t cap = (uint64_t)ctx->cu_count * 6u; // 6 blocks of 4 waves per CU: the occupancy limit
	grid = grid < cap ? grid : cap;
	return launch(ctx, name, march_persist<kVariant, kBlock, Src>, dim3((uint32_t)(grid ? grid : 1)), dim3(kBlock), 0,
	              src, n, ctx->d_err);
}

} // namespace

namespace {

// The beam pre-pass into beam_img; timed: bracket it with the context's stage timer (whose events
// are created on first use).
int march_beam_run(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t w, uint32_t h,
                   float *beam_img, uint32_t *nfetch, bool timed) {
	const CamArgs ca = cam_args(cam);
	const uint32_t bw = (w + 7u) / 8u + 1u, bh = (h + 7u) / 8u + 1u; // src/OctreeTracer.cpp:6
	// exp2f(1.0f - level) (src/OctreeTracer.cpp:206): an exact power of two
	const float orig_sz = ldexpf(1.0f, 1 - (int)oct->level);
	if (timed)
		ctx->t_beam.start(ctx->stream);
	hipError_t e = launch(ctx, "march_beam", march_beam_kernel, dim3(div_up((uint64_t)bw * bh, kBeamBlock)), dim3(kBeamBlock),
	                      0, (const uint32_t *)oct->words, oct->W, ca, w, h, bw, bh, orig_sz, beam_img, nfetch,
	                      ctx->d_err);
	if (timed)
		ctx->t_beam.stop(ctx->stream);
	if (e != hipSuccess)
		return hip_check(ctx, e, "march_beam");
	return SVO_OK;
}

} // namespace

int march_beam(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t w, uint32_t h, float *beam_img,
               uint32_t *nfetch) {
	return march_beam_run(ctx, oct, cam, w, h, beam_img, nfetch, true);
}

namespace {

// The primary march of a frame (and its beam pre-pass when beam_img is not null, written there).
// may_tune: the autotuner may time the candidate schedules of a new shape class; without it the
// schedule is the context's fixed one, or one ray per lane in 16x16 tiles.  timed: bracket the
// passes with the context's stage timers.
int march_frame_into(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t w, uint32_t h,
                     float *beam_img, svo_hit *hits, bool may_tune, bool timed = true) {
	hipError_t e;
	const CamArgs ca = cam_args(cam);
	const uint32_t bw = (w + 7u) / 8u + 1u; // src/OctreeTracer.cpp:6
	const int beam = beam_img != nullptr;
	if (beam) {
		int rc = march_beam_run(ctx, oct, cam, w, h, beam_img, nullptr, timed);
		if (rc)
			return rc;
	}
	// schedules: 0 = one ray per lane in 16x16 tiles, 1 = persistent refill, 2 = one ray per lane in
	// one-wave 8x8 tiles
	auto run = [&](int mode) -> hipError_t {
		if (mode == 1) {
			const uint32_t tx8 = (w + 7u) / 8u, ty8 = (h + 7u) / 8u;
			FrameRays src{(const uint32_t *)oct->words, oct->W, ca, w, h, bw, tx8, beam_img,
			              reinterpret_cast<uint4 *>(hits)};
			return launch_persist<0>(ctx, "march_frame", src, (uint64_t)tx8 * ty8 * 64u);
		}
		const uint32_t WX = mode == 0 ? 2u : 1u, WY = mode == 0 ? 2u : 1u; // 16x16 or 8x8 tiles
		const uint32_t tiles_x = (w + 8u * WX - 1u) / (8u * WX), tiles_y = (h + 8u * WY - 1u) / (8u * WY);
		if (mode == 0)
			return launch(ctx, "march_frame", march_frame_kernel<2, 2>, dim3(tiles_x * tiles_y), dim3(256), 0,
			              (const uint32_t *)oct->words, oct->W, ca, w, h, (const float *)beam_img, bw, tiles_x,
			              reinterpret_cast<uint4 *>(hits), ctx->d_err);
		return launch(ctx, "march_frame", march_frame_kernel<1, 1>, dim3(tiles_x * tiles_y), dim3(64), 0,
		              (const uint32_t *)oct->words, oct->W, ca, w, h, (const float *)beam_img, bw, tiles_x,
		              reinterpret_cast<uint4 *>(hits), ctx->d_err);
	};
	int mode = ctx->march_mode;
	if (mode < 0) {
		mode = 0;
		if (ctx->autotune && may_tune) {
			char key[96];
			std::snprintf(key, sizeof key, "march_frame %ux%u level=%u beam=%d", w, h, oct->level, beam ? 1 : 0);
			const int rc = tune(ctx, key, 3, kScheduleName, run, mode);
			if (rc)
				return rc;
		}
	}
	if (timed)
		ctx->t_march.start(ctx->stream);
	e = run(mode);
	if (timed)
		ctx->t_march.stop(ctx->stream);
	if (e != hipSuccess)
		return hip_check(ctx, e, "march_frame");
	return SVO_OK;
}

// ---- allocation-free primary march ----------------------------------------------------------------
// The beam image of svo_march_noalloc lives in this file-scope device array, one copy per GPU for the
// life of the process.  In host code `scratch` names a host-side stand-in, not device memory: its
// device address comes from hipGetSymbolAddress, taken once per device with that device current
// (noalloc_init, called when a context is created) and kept in g_scratch_dev.
constexpr uint64_t kNoAllocBeamCells = 1ull << 20; // 4 MiB of floats: (ceil(W/8)+1)(ceil(H/8)+1) cells
constexpr int kMaxDevices = 64;
static __device__ __align__(256) float scratch[kNoAllocBeamCells];
std::mutex g_scratch_mu;
float *g_scratch_dev[kMaxDevices] = {};

} // namespace

int noalloc_init(svo_ctx *ctx) {
	if (ctx->device < 0 || ctx->device >= kMaxDevices)
		return set_error(ctx, SVO_ERR_NO_DEVICE, "device index beyond the allocation-free march's table");
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	float *&p = g_scratch_dev[ctx->device];
	if (!p) {
		void *addr = nullptr;
		const hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !addr)
			return hip_check(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
		p = static_cast<float *>(addr);
	}
	ctx->noalloc_beam = p;
	return SVO_OK;
}

int march_frame_noalloc(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t w, uint32_t h,
                        svo_hit *hits) {
	const uint64_t cells = ((uint64_t)(w + 7u) / 8u + 1u) * ((uint64_t)(h + 7u) / 8u + 1u);
	if (cells > kNoAllocBeamCells)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT,
		                 "frame too large for the allocation-free march: (ceil(W/8)+1)*(ceil(H/8)+1) = " +
		                     std::to_string(cells) + " beam cells, at most " + std::to_string(kNoAllocBeamCells));
	if (!ctx->noalloc_beam)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "allocation-free march not set up on this context");
	return march_frame_into(ctx, oct, cam, w, h, ctx->noalloc_beam, hits, false, false);
}

int march_frame(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t w, uint32_t h, int beam,
                svo_hit *hits, float *beam_out) {
	const uint32_t bw = (w + 7u) / 8u + 1u, bh = (h + 7u) / 8u + 1u; // src/OctreeTracer.cpp:6
	float *beam_img = nullptr;
	if (beam) {
		beam_img = beam_out;
		if (!beam_img && ctx->pool->alloc((void **)&beam_img, sizeof(float) * bw * bh) != hipSuccess)
			return set_error(ctx, SVO_ERR_OUT_OF_MEMORY, "beam image allocation failed");
	}
	const int rc = march_frame_into(ctx, oct, cam, w, h, beam_img, hits, true);
	if (beam && !beam_out)
		ctx->pool->free(beam_img); // stream-ordered reuse: the pool only hands it to later work on this stream
	return rc;
}

int march_rays(svo_ctx *ctx, const svo_octree *oct, const float *rays, uint64_t R, int variant, svo_hit *hits) {
	if (R == 0)
		return SVO_OK;
	hipError_t e;
	const dim3 grid(div_up(R, kBlock));
	uint4 *out = reinterpret_cast<uint4 *>(hits);
	const ExplicitRays src{(const uint32_t *)oct->words, oct->W, rays, out};
	auto run = [&](int mode) -> hipError_t {
		if (mode == 1)
			return variant == SVO_MARCH_LEAF_ITER ? launch_persist<0>(ctx, "march_rays6", src, R)
			       : variant == SVO_MARCH_LEAF    ? launch_persist<1>(ctx, "march_rays5", src, R)
			                                      : launch_persist<2>(ctx, "march_occlude", src, R);
		if (variant == SVO_MARCH_LEAF_ITER)
			return launch(ctx, "march_rays6", march_rays_kernel<0>, grid, dim3(kBlock), 0,
			              (const uint32_t *)oct->words, oct->W, rays, R, out, ctx->d_err);
		if (variant == SVO_MARCH_LEAF)
			return launch(ctx, "march_rays5", march_rays_kernel<1>, grid, dim3(kBlock), 0,
			              (const uint32_t *)oct->words, oct->W, rays, R, out, ctx->d_err);
		return launch(ctx, "march_occlude", march_rays_kernel<2>, grid, dim3(kBlock), 0, (const uint32_t *)oct->words,
		              oct->W, rays, R, out, ctx->d_err);
	};
	int mode = ctx->march_mode;
	if (mode < 0) {
		mode = 0;
		if (ctx->autotune) {
			uint32_t lg = 0;
			while ((1ull << (lg + 1)) <= R)
				++lg;
			char key[96];
			std::snprintf(key, sizeof key, "march_rays variant=%d level=%u R~2^%u", variant, oct->level, lg);
			const int rc = tune(ctx, key, 2, kScheduleName, run, mode);
			if (rc)
				return rc;
		}
	}
	ctx->t_rays.start(ctx->stream);
	e = run(mode);
	ctx->t_rays.stop(ctx->stream);
	if (e != hipSuccess)
		return hip_check(ctx, e, "march_rays");
	return SVO_OK;
}

namespace {
// Writes the Sobol state of sample index n into the context's Sobol buffer (stream-ordered).
hipError_t sobol_point(svo_ctx *ctx, uint32_t n) {
	return launch(ctx, "sobol_point", sobol_point_kernel, dim3(1), dim3(64), 0, n, ctx->d_sobol);
}
} // namespace

int bounce_rays(svo_ctx *ctx, const svo_hit *hits, uint32_t width, uint32_t height, uint32_t spp, uint32_t cur,
                float *rays) {
	const uint64_t R = (uint64_t)width * height;
	if (R == 0)
		return SVO_OK;
	hipError_t e = sobol_point(ctx, spp);
	if (e != hipSuccess)
		return hip_check(ctx, e, "sobol_point");
	ctx->t_bounce.start(ctx->stream);
	e = launch(ctx, "bounce", bounce_kernel, dim3(div_up(R, kBlock)), dim3(kBlock), 0, reinterpret_cast<const uint4 *>(hits),
	           R, width, (const uint32_t *)ctx->d_sobol, cur, rays);
	ctx->t_bounce.stop(ctx->stream);
	if (e != hipSuccess)
		return hip_check(ctx, e, "bounce");
	return SVO_OK;
}

namespace {
// Autotuned launch of one progressive sample: the trials render sample 0 into a scratch buffer, so
// the caller's running means see one sample only.  run_with(args, schedule) launches.
template <class Run>
int path_run(svo_ctx *ctx, const char *what, const char *key, const PathArgs &a, Run run_with) {
	int mode = ctx->march_mode == 1 ? 1 : 0;
	if (ctx->march_mode < 0 && ctx->autotune) {
		if (ctx->tune_choice.find(key) == ctx->tune_choice.end()) {
			float4 *scratch = nullptr;
			if (ctx->pool->alloc((void **)&scratch, sizeof(float4) * (uint64_t)a.width * a.height) != hipSuccess)
				return set_error(ctx, SVO_ERR_OUT_OF_MEMORY, "path tracer tuning buffer");
			PathArgs trial = a;
			trial.spp = 0u;
			trial.color = scratch;
			trial.albedo = trial.normal = reinterpret_cast<uint32_t *>(scratch);
			const int rc = tune(ctx, key, 2, kScheduleName, [&](int m) { return run_with(trial, m); }, mode);
			ctx->pool->free(scratch); // stream-ordered reuse
			if (rc)
				return rc;
		} else {
			mode = ctx->tune_choice[key];
		}
	}
	ctx->t_path.start(ctx->stream);
	const hipError_t e = run_with(a, mode);
	ctx->t_path.stop(ctx->stream);
	if (e != hipSuccess)
		return hip_check(ctx, e, what);
	return SVO_OK;
}
} // namespace

// One progressive sample per pixel (PathTracer::CmdRender, src/PathTracer.cpp:196-215): the Sobol
// state of sample spp, then one path per lane in 8x8 tiles (schedule 0) or persistent waves with
// per-lane refill (schedule 1), autotuned like the marches; the results are identical.
int path_trace(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t w, uint32_t h, uint32_t bounce,
               uint32_t spp, const float sky[3], float *color, uint32_t *albedo, uint32_t *normal) {
	const uint32_t tiles_x = (w + 7u) / 8u, tiles_y = (h + 7u) / 8u;
	const PathArgs a{cam_args(cam), w, h, bounce, spp, {sky[0], sky[1], sky[2]}, ctx->d_sobol, tiles_x,
	                 reinterpret_cast<float4 *>(color), albedo, normal};
	hipError_t e = sobol_point(ctx, spp);
	if (e != hipSuccess)
		return hip_check(ctx, e, "sobol_point");
	auto run_with = [&](const PathArgs &pa, int mode) -> hipError_t {
		if (mode == 1) {
			const uint64_t n = (uint64_t)tiles_x * tiles_y * 64u;
			uint64_t grid = (n + kBlock - 1) / kBlock;
			const uint64_t cap = (uint64_t)ctx->cu_count * 5u; // 96 VGPRs: 5 waves per SIMD
			grid = grid < cap ? grid : cap;
			return launch(ctx, "path_trace", path_persist<kBlock>, dim3((uint32_t)(grid ? grid : 1)), dim3(kBlock), 0,
			              (const uint32_t *)oct->words, oct->W, pa, n, ctx->d_err);
		}
		return launch(ctx, "path_trace", path_trace_kernel, dim3(tiles_x * tiles_y), dim3(64), 0,
		              (const uint32_t *)oct->words, oct->W, pa, ctx->d_err);
	};
	char key[96];
	std::snprintf(key, sizeof key, "path_trace %ux%u level=%u bounce=%u", w, h, oct->level, bounce);
	return path_run(ctx, "path_trace", key, a, run_with);
}

// One progressive sample per pixel with the environment-map light (EnvMapPathTrace).
int path_trace_env(svo_ctx *ctx, const svo_octree *oct, const svo_camera *cam, uint32_t w, uint32_t h,
                   uint32_t bounce, uint32_t spp, const svo_envmap *env, float rotation, float multiplier,
                   float *color, uint32_t *albedo, uint32_t *normal) {
	const uint32_t tiles_x = (w + 7u) / 8u, tiles_y = (h + 7u) / 8u;
	const PathArgs a{cam_args(cam), w, h, bounce, spp, {0.0f, 0.0f, 0.0f}, ctx->d_sobol, tiles_x,
	                 reinterpret_cast<float4 *>(color), albedo, normal};
	const EnvArgs e{reinterpret_cast<const float4 *>(env->texels), reinterpret_cast<const uint2 *>(env->alias),
	                env->width, env->height, rotation, multiplier};
	hipError_t err = sobol_point(ctx, spp);
	if (err != hipSuccess)
		return hip_check(ctx, err, "sobol_point");
	auto run_with = [&](const PathArgs &pa, int mode) -> hipError_t {
		if (mode == 1) {
			const uint64_t n = (uint64_t)tiles_x * tiles_y * 64u;
			uint64_t grid = (n + kBlock - 1) / kBlock;
			const uint64_t cap = (uint64_t)ctx->cu_count * kPersistEnvBlocksPerCu;
			grid = grid < cap ? grid : cap;
			return launch(ctx, "path_trace_env", path_persist_env<kBlock>, dim3((uint32_t)(grid ? grid : 1)),
			              dim3(kBlock), 0, (const uint32_t *)oct->words, oct->W, pa, e, n, ctx->d_err);
		}
		return launch(ctx, "path_trace_env", path_trace_env_kernel, dim3(tiles_x * tiles_y), dim3(64), 0,
		              (const uint32_t *)oct->words, oct->W, pa, e, ctx->d_err);
	};
	char key[96];
	std::snprintf(key, sizeof key, "path_trace_env %ux%u level=%u bounce=%u", w, h, oct->level, bounce);
	return path_run(ctx, "path_trace_env", key, a, run_with);
}

// The tracer's shaded image (OctreeTracer::CmdDrawPipeline's colour output, m_view_type).
int shade_frame(svo_ctx *ctx, const svo_hit *hits, const svo_camera *cam, uint32_t w, uint32_t h, uint32_t view,
                const float light[3], const svo_envmap *env, float rotation, uint32_t *out) {
	const uint64_t n = (uint64_t)w * h;
	if (n == 0)
		return SVO_OK;
	const ShadeArgs a{cam_args(cam), w, h, view, {light[0], light[1], light[2]},
	                  env ? reinterpret_cast<const float4 *>(env->texels) : nullptr, env ? env->width : 0u,
	                  env ? env->height : 0u, rotation};
	const hipError_t e = launch(ctx, "shade", shade_kernel, dim3(div_up(n, kBlock)), dim3(kBlock), 0,
	                            reinterpret_cast<const uint4 *>(hits), a, out);
	if (e != hipSuccess)
		return hip_check(ctx, e, "shade");
	return SVO_OK;
}

// ---- batched marches ----------------------------------------------------------------------------
namespace {

constexpr int kFrameShapes = 4;
const char *const kFrameShapeName[kFrameShapes] = {"16x16", "8x8", "16x8", "32x8"};
constexpr uint32_t kShapeWX[kFrameShapes] = {2, 1, 2, 4};
constexpr uint32_t kShapeWY[kFrameShapes] = {2, 1, 1, 1};
// ray-batch schedules: one ray per lane in blocks of 256 / 128 / 64 threads, or persistent refill
constexpr int kRayShapes = 4;
constexpr uint32_t kRayBlock[3] = {256, 128, 64};

MarchSeg march_seg(const svo_octree *o, const svo_camera *cam) {
	MarchSeg g{};
	g.oct = o->words;
	// node loads are bounds-checked against the allocation: an octree of a build plan keeps its
	// buffer while its used size changes from run to run, so a march plan made over it once stays
	// valid (every other octree's allocation is its used size)
	g.W = o->cap > o->W ? o->cap : o->W;
	if (cam)
		g.cam = cam_args(cam);
	g.orig_sz = ldexpf(1.0f, 1 - (int)o->level); // exp2f(1.0f - level), src/OctreeTracer.cpp:206
	g.level = o->level;
	return g;
}

// Enqueues a plan's kernels with frame launch shape `shape`; prof = false bypasses the per-kernel
// profiler (graph capture, tuning).
hipError_t plan_enqueue(svo_ctx *ctx, const MarchPlan &p, int shape, svo_hit *hits, bool prof) {
	const bool saved = ctx->prof.enabled;
	if (!prof)
		ctx->prof.enabled = false;
	hipError_t e = hipSuccess;
	const MarchSeg *segs = p.segs;
	if (p.beam) {
		const uint32_t bps = div_up((uint64_t)p.bw * p.bh, kBlock);
		e = launch(ctx, "march_beam_batch", march_beam_batch, dim3(bps * p.S), dim3(kBlock), 0, segs, p.width,
		           p.height, p.bw, p.bh, bps, p.beam_buf, ctx->d_err);
	}
	if (e == hipSuccess) {
		const uint32_t WX = kShapeWX[shape], WY = kShapeWY[shape];
		const uint32_t tiles_x = div_up(p.width, 8 * WX), tiles_y = div_up(p.height, 8 * WY);
		const uint32_t tps = tiles_x * tiles_y;
		const dim3 grid(tps * p.S), block(64 * WX * WY);
		const float *bb = p.beam ? p.beam_buf : nullptr;
		uint4 *out = reinterpret_cast<uint4 *>(hits);
		switch (shape) {
		case 0:
			e = launch(ctx, "march_frame_batch", march_frame_batch<2, 2>, grid, block, 0, segs, p.width, p.height, bb,
			           p.bw, p.bh, tiles_x, tps, out, ctx->d_err);
			break;
		case 1:
			e = launch(ctx, "march_frame_batch", march_frame_batch<1, 1>, grid, block, 0, segs, p.width, p.height, bb,
			           p.bw, p.bh, tiles_x, tps, out, ctx->d_err);
			break;
		case 2:
			e = launch(ctx, "march_frame_batch", march_frame_batch<2, 1>, grid, block, 0, segs, p.width, p.height, bb,
			           p.bw, p.bh, tiles_x, tps, out, ctx->d_err);
			break;
		default:
			e = launch(ctx, "march_frame_batch", march_frame_batch<4, 1>, grid, block, 0, segs, p.width, p.height, bb,
			           p.bw, p.bh, tiles_x, tps, out, ctx->d_err);
			break;
		}
	}
	ctx->prof.enabled = saved;
	return e;
}

} // namespace

int march_plan_init(svo_ctx *ctx, MarchPlan *p, const svo_octree *const *octs, const svo_camera *cams, uint32_t S,
                    uint32_t width, uint32_t height, int beam) {
	p->S = S;
	p->width = width;
	p->height = height;
	p->beam = beam ? 1 : 0;
	p->bw = (width + 7u) / 8u + 1u; // src/OctreeTracer.cpp:6
	p->bh = (height + 7u) / 8u + 1u;
	p->h_segs.resize(S);
	for (uint32_t s = 0; s < S; ++s)
		p->h_segs[s] = march_seg(octs[s], &cams[s]);
	if (ctx->pool->alloc((void **)&p->segs, sizeof(MarchSeg) * S) != hipSuccess ||
	    (p->beam && ctx->pool->alloc((void **)&p->beam_buf, sizeof(float) * S * (uint64_t)p->bw * p->bh) != hipSuccess))
		return set_error(ctx, SVO_ERR_OUT_OF_MEMORY, "march plan allocation failed");
	return upload_table(ctx, p->segs, p->h_segs.data(), sizeof(MarchSeg) * S);
}

int march_plan_set_cameras(svo_ctx *ctx, MarchPlan *p, const svo_camera *cams) {
	for (uint32_t s = 0; s < p->S; ++s)
		p->h_segs[s].cam = cam_args(&cams[s]);
	return upload_table(ctx, p->segs, p->h_segs.data(), sizeof(MarchSeg) * p->S);
}

int march_plan_run(svo_ctx *ctx, MarchPlan *p, svo_hit *hits) {
	if (p->S == 0 || p->width == 0 || p->height == 0)
		return SVO_OK;
	if (p->shape < 0) {
		p->shape = 0;
		if (ctx->autotune) {
			char key[96];
			std::snprintf(key, sizeof key, "march_frame_batch S=%u %ux%u beam=%d", p->S, p->width, p->height, p->beam);
			int rc = tune(
			    ctx, key, kFrameShapes, kFrameShapeName,
			    [&](int sh) { return plan_enqueue(ctx, *p, sh, hits, false); }, p->shape);
			if (rc)
				return rc;
		}
	}
	hipError_t e = hipSuccess;
	ctx->t_march.start(ctx->stream);
	if (p->use_graph) {
		if (!p->exec || p->exec_hits != hits) {
			if (p->exec) {
				(void)hipGraphExecDestroy(p->exec);
				p->exec = nullptr;
			}
			hipGraph_t g = nullptr;
			e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
			if (e == hipSuccess) {
				const hipError_t e2 = plan_enqueue(ctx, *p, p->shape, hits, false);
				e = hipStreamEndCapture(ctx->stream, &g);
				if (e == hipSuccess)
					e = e2;
			}
			if (e == hipSuccess)
				e = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
			if (g)
				(void)hipGraphDestroy(g);
			p->exec_hits = e == hipSuccess ? hits : nullptr;
		}
		if (e == hipSuccess)
			e = hipGraphLaunch(p->exec, ctx->stream);
	} else {
		e = plan_enqueue(ctx, *p, p->shape, hits, true);
	}
	ctx->t_march.stop(ctx->stream);
	return hip_check(ctx, e, "march batch");
}

void march_plan_release(svo_ctx *ctx, MarchPlan *p) {
	if (p->exec)
		(void)hipGraphExecDestroy(p->exec);
	p->exec = nullptr;
	ctx->pool->free(p->segs);
	ctx->pool->free(p->beam_buf);
	p->segs = nullptr;
	p->beam_buf = nullptr;
}

int march_rays_segments(svo_ctx *ctx, const svo_octree *const *octs, const float *const *rays, svo_hit *const *hits,
                        const uint64_t *count, uint32_t S, int variant) {
	if (S == 0)
		return SVO_OK;
	std::vector<MarchSeg> h(S);
	std::vector<uint64_t> rb(S + 1, 0);
	for (uint32_t s = 0; s < S; ++s) {
		h[s] = march_seg(octs[s], nullptr);
		h[s].rays = rays[s];
		h[s].hits = reinterpret_cast<uint4 *>(hits[s]);
		rb[s + 1] = rb[s] + count[s];
	}
	const uint64_t R = rb[S];
	if (R == 0)
		return SVO_OK;
	MarchSeg *segs = nullptr;
	uint64_t *rbeg = nullptr;
	if (ctx->pool->alloc((void **)&segs, sizeof(MarchSeg) * S) != hipSuccess ||
	    ctx->pool->alloc((void **)&rbeg, sizeof(uint64_t) * (S + 1)) != hipSuccess) {
		ctx->pool->free(segs);
		return set_error(ctx, SVO_ERR_OUT_OF_MEMORY, "ray batch allocation failed");
	}
	int rc = upload_table(ctx, segs, h.data(), sizeof(MarchSeg) * S);
	if (!rc)
		rc = upload_table(ctx, rbeg, rb.data(), sizeof(uint64_t) * (S + 1));
	auto run = [&](int sh) -> hipError_t {
		if (sh == 3) {
			const SegmentRays src{(const MarchSeg *)segs, (const uint64_t *)rbeg, S};
			return variant == SVO_MARCH_LEAF_ITER ? launch_persist<0>(ctx, "march_rays_batch", src, R)
			       : variant == SVO_MARCH_LEAF    ? launch_persist<1>(ctx, "march_rays_batch", src, R)
			                                      : launch_persist<2>(ctx, "march_rays_batch", src, R);
		}
		const uint32_t bs = kRayBlock[sh];
		const dim3 grid(div_up(R, bs)), block(bs);
#define SVO_RAYS(V, T)                                                                                                 \
	launch(ctx, "march_rays_batch", march_rays_batch<V, T>, grid, block, 0, (const MarchSeg *)segs,                    \
	       (const uint64_t *)rbeg, S, R, ctx->d_err)
		if (variant == SVO_MARCH_LEAF_ITER)
			return bs == 256 ? SVO_RAYS(0, 256) : bs == 128 ? SVO_RAYS(0, 128) : SVO_RAYS(0, 64);
		if (variant == SVO_MARCH_LEAF)
			return bs == 256 ? SVO_RAYS(1, 256) : bs == 128 ? SVO_RAYS(1, 128) : SVO_RAYS(1, 64);
		return bs == 256 ? SVO_RAYS(2, 256) : bs == 128 ? SVO_RAYS(2, 128) : SVO_RAYS(2, 64);
#undef SVO_RAYS
	};
	// an explicit schedule (svo_march_mode 1: persistent; 0, 2: one ray per lane in blocks of 256),
	// else the autotuner's choice among the four
	int shape = ctx->march_mode == 1 ? 3 : 0;
	if (!rc && ctx->march_mode < 0 && ctx->autotune) {
		uint32_t lg = 0;
		while ((1ull << (lg + 1)) <= R)
			++lg;
		char key[96];
		std::snprintf(key, sizeof key, "march_rays_batch variant=%d R~2^%u", variant, lg);
		static const char *const names[kRayShapes] = {"block256", "block128", "block64", "persistent-refill"};
		rc = tune(ctx, key, kRayShapes, names, run, shape);
	}
	if (!rc) {
		ctx->t_rays.start(ctx->stream);
		const hipError_t e = run(shape);
		ctx->t_rays.stop(ctx->stream);
		rc = hip_check(ctx, e, "march_rays_batch");
	}
	ctx->pool->free(segs);
	ctx->pool->free(rbeg);
	return rc;
}

int march_rays_batch(svo_ctx *ctx, const svo_octree *const *octs, uint32_t S, const uint64_t *ray_begin,
                     const float *rays, int variant, svo_hit *hits) {
	std::vector<const float *> r(S);
	std::vector<svo_hit *> h(S);
	std::vector<uint64_t> c(S);
	for (uint32_t s = 0; s < S; ++s) {
		r[s] = rays + 6 * ray_begin[s];
		h[s] = hits + ray_begin[s];
		c[s] = ray_begin[s + 1] - ray_begin[s];
	}
	return march_rays_segments(ctx, octs, r.data(), h.data(), c.data(), S, variant);
}

} // namespace svo
