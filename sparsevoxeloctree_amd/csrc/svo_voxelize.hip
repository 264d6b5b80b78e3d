// Mesh voxelization into the reference's fragment list (Voxelizer, src/Voxelizer.cpp:5-179), as
// compute kernels instead of the geometry-shader + conservative-raster pipeline.
//
//   per triangle (shader/voxelizer.geom:21-54): dominant axis of the face normal, swizzled
//     projection, pixel AABB and voxel depth range, exactly as the geometry shader computes them;
//   coverage (conservative rasterization, src/Voxelizer.cpp:112-127): pixel (i,j) is covered when
//     the closed triangle, with vertices snapped to 1/256 pixel, meets the half-open pixel square
//     [i,i+1) x [j,j+1).  Written as three linear constraints lo_k <= Sx_k*i + Sy_k*j <= hi_k
//     (one per edge normal, exact int64), so every row's covered pixels form one span found with
//     two divisions per edge instead of testing pixels one by one;
//   per pixel (shader/voxelizer.frag:17-44): depth of the triangle's plane at the pixel centre,
//     clamps to [0,res-1] and to the depth range, swizzle back, fragment packing.
// Two passes like the reference's count-only + fill draws: vox_count writes the per-triangle count,
// a device scan turns counts into offsets, vox_fill writes each triangle's fragments at its offset
// (triangle order, row-major inside a triangle).  Both passes cover a batch of scenes in one launch
// (segment tables, svo_batch.hpp): the triangles of all scenes form one index space, and every
// scene's fragments land in its own list.
//
// Scene normalisation (src/Scene.cpp:62-99) runs here too: bounding-box reduction + transform.
#include <cstring>
#include <string>
#include <vector>

#include "svo_batch.hpp"
#include "svo_device.hpp"
#include "svo_runtime.hpp"
#include "svo_scan.hpp"

namespace svo {
namespace {

constexpr int kThreads = 256;

struct Edge {
	int64_t sx, sy; // 256 * edge normal
	int64_t lo, hi; // covered iff lo <= sx*i + sy*j <= hi
};

struct Setup {
	bool empty;
	uint32_t axis;
	uint32_t dr0, dr1;
	int32_t ilo, ihi, jlo, jhi; // pixel rectangle: AABB ∩ snapped bbox ∩ [0, res)
	float px0, py0, pz0;
	float dzdx, dzdy; // plane of the projected triangle (pixel units, depth in [0,1])
	Edge e[3];
};

__device__ __forceinline__ f3 project(f3 v, uint32_t axis) { // voxelizer.geom:15-19
	f3 r = axis == 0 ? f3{v.y, v.z, v.x} : (axis == 1 ? f3{v.z, v.x, v.y} : v);
	r.z = (r.z + 1.0f) * 0.5f;
	return r;
}

__device__ __forceinline__ int64_t imin64(int64_t a, int64_t b) { return a < b ? a : b; }
__device__ __forceinline__ int64_t imax64(int64_t a, int64_t b) { return a < b ? b : a; }

__device__ void tri_setup(const float *__restrict__ tri, uint64_t t, uint32_t res, Setup &s) {
	const float *p = tri + 9 * t;
	f3 p0{p[0], p[1], p[2]}, p1{p[3], p[4], p[5]}, p2{p[6], p[7], p[8]};
	f3 e1{p1.x - p0.x, p1.y - p0.y, p1.z - p0.z}, e2{p2.x - p0.x, p2.y - p0.y, p2.z - p0.z};
	f3 n = cross3(e1, e2); // voxelizer.geom:29
	float wx = fabsf(n.x), wy = fabsf(n.y), wz = fabsf(n.z);
	s.axis = (wx > wy && wx > wz) ? 0u : (wy > wz ? 1u : 2u);
	bool degenerate = (n.x == 0.0f && n.y == 0.0f && n.z == 0.0f);
	f3 q0 = project(p0, s.axis), q1 = project(p1, s.axis), q2 = project(p2, s.axis);
	const float fres = (float)res;
	// voxelizer.geom:40-43
	float mnx = gmin(q0.x, gmin(q1.x, q2.x)), mny = gmin(q0.y, gmin(q1.y, q2.y));
	float mxx = gmax(q0.x, gmax(q1.x, q2.x)), mxy = gmax(q0.y, gmax(q1.y, q2.y));
	uint32_t a0 = f2u_sat((mnx + 1.0f) * 0.5f * fres), a1 = f2u_sat((mny + 1.0f) * 0.5f * fres);
	uint32_t a2 = f2u_sat((mxx + 1.0f) * 0.5f * fres), a3 = f2u_sat((mxy + 1.0f) * 0.5f * fres);
	s.dr0 = f2u_sat(gmin(q0.z, gmin(q1.z, q2.z)) * fres);
	s.dr1 = f2u_sat(gmax(q0.z, gmax(q1.z, q2.z)) * fres);

	float px[3] = {(q0.x + 1.0f) * 0.5f * fres, (q1.x + 1.0f) * 0.5f * fres, (q2.x + 1.0f) * 0.5f * fres};
	float py[3] = {(q0.y + 1.0f) * 0.5f * fres, (q1.y + 1.0f) * 0.5f * fres, (q2.y + 1.0f) * 0.5f * fres};
	float pz[3] = {q0.z, q1.z, q2.z};
	int64_t X[3], Y[3];
#pragma unroll
	for (int k = 0; k < 3; ++k) {
		X[k] = (int64_t)__builtin_rintf(px[k] * 256.0f);
		Y[k] = (int64_t)__builtin_rintf(py[k] * 256.0f);
	}
	s.px0 = px[0];
	s.py0 = py[0];
	s.pz0 = pz[0];
	{
		float ax = px[1] - px[0], ay = py[1] - py[0], az = pz[1] - pz[0];
		float bx = px[2] - px[0], by = py[2] - py[0], bz = pz[2] - pz[0];
		float nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
		s.dzdx = 0.0f;
		s.dzdy = 0.0f;
		if (nz != 0.0f) {
			s.dzdx = -nx / nz;
			s.dzdy = -ny / nz;
		}
	}
	int64_t xmn = imin64(X[0], imin64(X[1], X[2])), xmx = imax64(X[0], imax64(X[1], X[2]));
	int64_t ymn = imin64(Y[0], imin64(Y[1], Y[2])), ymx = imax64(Y[0], imax64(Y[1], Y[2]));
	int64_t ilo = imax64((int64_t)a0, xmn >> 8), ihi = imin64(imin64((int64_t)a2, (int64_t)res - 1), xmx >> 8);
	int64_t jlo = imax64((int64_t)a1, ymn >> 8), jhi = imin64(imin64((int64_t)a3, (int64_t)res - 1), ymx >> 8);
	s.ilo = (int32_t)imax64(ilo, 0);
	s.ihi = (int32_t)imax64(ihi, -1);
	s.jlo = (int32_t)imax64(jlo, 0);
	s.jhi = (int32_t)imax64(jhi, -1);
	s.empty = degenerate || s.ilo > s.ihi || s.jlo > s.jhi;
#pragma unroll
	for (int k = 0; k < 3; ++k) {
		const int a = k, b = (k + 1) % 3, c = (k + 2) % 3;
		int64_t A = Y[b] - Y[a], B = X[a] - X[b];
		int64_t ce = A * X[a] + B * Y[a], cc = A * X[c] + B * Y[c];
		int64_t tlo = imin64(ce, cc), thi = imax64(ce, cc);
		bool lo_att = A >= 0 && B >= 0, hi_att = A <= 0 && B <= 0;
		int64_t thi2 = lo_att ? thi : thi - 1, tlo2 = hi_att ? tlo : tlo + 1;
		int64_t sx = A * 256, sy = B * 256;
		int64_t off_lo = (A < 0 ? sx : 0) + (B < 0 ? sy : 0);
		int64_t width = (sx < 0 ? -sx : sx) + (sy < 0 ? -sy : sy);
		s.e[k] = Edge{sx, sy, tlo2 - off_lo - width, thi2 - off_lo};
	}
}

// Covered pixels of row j: [i0, i1] (empty when i0 > i1).  Exact: the double quotient of two
// integers below 2^53 never crosses an integer for |quotient| < 2^25 (|sx| <= 2^29), and bounds
// further out are clamped away by the pixel rectangle.
__device__ __forceinline__ void row_span(const Setup &s, int32_t j, int32_t &i0, int32_t &i1) {
	double lo = (double)s.ilo, hi = (double)s.ihi;
#pragma unroll
	for (int k = 0; k < 3; ++k) {
		const Edge &e = s.e[k];
		int64_t syj = e.sy * (int64_t)j;
		int64_t nlo = e.lo - syj, nhi = e.hi - syj;
		if (e.sx > 0) {
			lo = fmax(lo, ceil((double)nlo / (double)e.sx));
			hi = fmin(hi, floor((double)nhi / (double)e.sx));
		} else if (e.sx < 0) {
			lo = fmax(lo, ceil((double)nhi / (double)e.sx));
			hi = fmin(hi, floor((double)nlo / (double)e.sx));
		} else if (nlo > 0 || nhi < 0) {
			hi = -1.0;
		}
	}
	if (lo > hi) {
		i0 = 1;
		i1 = 0;
	} else {
		i0 = (int32_t)lo;
		i1 = (int32_t)hi;
	}
}

// Exact floor(a / b) for b > 0 and |a| < 2^53: the double quotient is within one of it, and the
// integer remainder corrects it.
__device__ __forceinline__ int64_t floor_div(int64_t a, int64_t b) {
	int64_t q = (int64_t)floor((double)a / (double)b);
	const int64_t r = a - q * b;
	q -= r < 0 ? 1 : 0;
	q += r >= b ? 1 : 0;
	return q;
}

// floor(X_j / p) for X_j = X_0 + j * s, kept as X_j = q p + r (0 <= r < p) and advanced one step at
// a time with integer adds only.  p = 256 |A| and |s| = 256 |B| for snapped edge vectors (A, B)
// below 2^21, so p < 2^29, r fits 32 bits and the per-step quotient |floor(s / p)| < 2^21.
struct FloorTrack {
	int64_t q;
	int32_t dq;
	uint32_t r, dr;
	__device__ void init(int64_t X, int64_t step, uint32_t p) {
		q = floor_div(X, p);
		r = (uint32_t)(X - q * (int64_t)p);
		const int64_t d = floor_div(step, p);
		dq = (int32_t)d;
		dr = (uint32_t)(step - d * (int64_t)p);
	}
	__device__ __forceinline__ void next(uint32_t p) {
		q += dq;
		r += dr;
		if (r >= p) {
			r -= p;
			++q;
		}
	}
};

// The covered spans of consecutive rows, as row_span computes them, with no division per row.  An
// edge with sx = 0 constrains only which rows are covered, once per triangle (rows() narrows the
// triangle's row range).  For the others, with p = |sx| and U_j, V_j the constraint's two sides
// mirrored to a positive divisor (both change by the same step per row), row j's span is
// [max(ilo, ceil(U_j / p)), min(ihi, floor(V_j / p))], and ceil(U / p) = -floor(-U / p).
struct RowSpans {
	FloorTrack lo[3], hi[3];
	uint32_t p[3]; // |sx| per edge
	bool on[3];
	int32_t ilo, ihi;
	// rows [j0, j1] that every sx = 0 edge admits, within the triangle's [jlo, jhi]; then the
	// trackers start at row j0
	__device__ void init(const Setup &s, int32_t &j0, int32_t &j1) {
		ilo = s.ilo;
		ihi = s.ihi;
		int64_t jl = s.jlo, jh = s.jhi;
#pragma unroll
		for (int k = 0; k < 3; ++k) {
			const Edge &e = s.e[k];
			on[k] = e.sx != 0;
			if (e.sx != 0)
				continue;
			// covered iff lo <= sy j <= hi
			if (e.sy == 0) {
				if (e.lo > 0 || e.hi < 0)
					jh = jl - 1;
			} else if (e.sy > 0) {
				jl = imax64(jl, -floor_div(-e.lo, e.sy));
				jh = imin64(jh, floor_div(e.hi, e.sy));
			} else {
				jl = imax64(jl, -floor_div(e.hi, -e.sy));
				jh = imin64(jh, floor_div(-e.lo, -e.sy));
			}
		}
		j0 = (int32_t)jl;
		j1 = (int32_t)imax64(jh, jl - 1);
#pragma unroll
		for (int k = 0; k < 3; ++k) {
			const Edge &e = s.e[k];
			if (!on[k])
				continue;
			const int64_t nlo = e.lo - e.sy * (int64_t)j0, nhi = e.hi - e.sy * (int64_t)j0;
			const uint32_t pk = (uint32_t)(e.sx > 0 ? e.sx : -e.sx);
			p[k] = pk;
			const int64_t step = e.sx > 0 ? -e.sy : e.sy;   // per row, of U and V
			const int64_t U = e.sx > 0 ? nlo : -nhi, V = e.sx > 0 ? nhi : -nlo;
			lo[k].init(-U, -step, pk); // floor(-U_j / p)
			hi[k].init(V, step, pk);   // floor(V_j / p)
		}
	}
	__device__ __forceinline__ void span(int32_t &i0, int32_t &i1) const {
		int64_t a = ilo, b = ihi;
#pragma unroll
		for (int k = 0; k < 3; ++k)
			if (on[k]) {
				a = imax64(a, -lo[k].q);
				b = imin64(b, hi[k].q);
			}
		if (a > b) {
			i0 = 1;
			i1 = 0;
		} else {
			i0 = (int32_t)a;
			i1 = (int32_t)b;
		}
	}
	__device__ __forceinline__ void next() {
#pragma unroll
		for (int k = 0; k < 3; ++k)
			if (on[k]) {
				lo[k].next(p[k]);
				hi[k].next(p[k]);
			}
	}
};

__global__ void vox_zero_counter(uint32_t *c) { c[0] = 0u; }

// Count pass.  Triangles with more than kBigTri fragments are also appended to `big` (big[0] is
// the count) for vox_fill_big; the per-lane fill skips them.
constexpr uint32_t kBigTri = 128;

__global__ void __launch_bounds__(kThreads)
    vox_count(const VoxSeg *__restrict__ segs, const uint64_t *__restrict__ tbeg, uint32_t S, uint64_t T,
              uint64_t *__restrict__ counts, uint32_t *__restrict__ big) {
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	uint32_t hint = 0;
	for (uint64_t base = (uint64_t)blockIdx.x * kThreads; base < T; base += stride) {
		const uint64_t t = base + threadIdx.x;
		if (t >= T)
			continue;
		const float *xyz;
		uint64_t tl;
		uint32_t level;
		with_seg<kThreads>(tbeg, S, hint, base, t, [&](uint32_t sg) {
			xyz = segs[sg].xyz;
			level = segs[sg].level;
			tl = t - tbeg[sg];
		});
		Setup s;
		tri_setup(xyz, tl, 1u << level, s);
		uint32_t c = 0;
		if (!s.empty) {
			RowSpans rs;
			int32_t j0, j1;
			rs.init(s, j0, j1);
			for (int32_t j = j0; j <= j1; ++j) {
				int32_t i0, i1;
				rs.span(i0, i1);
				if (i0 <= i1)
					c += (uint32_t)(i1 - i0 + 1);
				rs.next();
			}
		}
		counts[t] = c;
		const uint32_t q = wave_append(&big[0], c > kBigTri);
		if (c > kBigTri)
			big[1 + q] = (uint32_t)t;
	}
}

// Per scene: first fragment and fragment count from the scanned offsets (total = all fragments).
__global__ void vox_seg_totals(VoxSeg *segs, const uint64_t *__restrict__ tbeg, uint32_t S, uint64_t T,
                               const uint64_t *__restrict__ offsets, const uint64_t *__restrict__ total) {
	const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= S)
		return;
	const uint64_t b = tbeg[s] < T ? offsets[tbeg[s]] : *total;
	const uint64_t e = tbeg[s + 1u] < T ? offsets[tbeg[s + 1u]] : *total;
	segs[s].frag_begin = b;
	segs[s].F = e - b;
}

// One fragment: depth at the pixel centre, clamps, swizzle back (voxelizer.frag:17-25) and packing
// (voxelizer.frag:38-42).  dy = dz/dy * (j + 0.5 - py0) of the row.
__device__ __forceinline__ uint2 make_frag(const Setup &s, int32_t i, int32_t j, float dy, uint32_t res,
                                          uint32_t colour) {
	const float zc = s.pz0 + (s.dzdx * (((float)i + 0.5f) - s.px0) + dy);
	uint32_t uz = umin32(f2u_sat(zc * (float)res), res - 1u);
	uz = umin32(umax32(uz, s.dr0), s.dr1);
	const uint32_t ui = (uint32_t)i, uj = (uint32_t)j;
	uint32_t vx, vy, vz;
	if (s.axis == 0u)
		vx = uz, vy = ui, vz = uj;
	else if (s.axis == 1u)
		vx = uj, vy = uz, vz = ui;
	else
		vx = ui, vy = uj, vz = uz;
	return make_uint2(vx | (vy << 12u) | ((vz & 0xffu) << 24u), ((vz >> 8u) << 28u) | colour);
}

// Fill of the small triangles: each lane walks its triangle's rows and writes its fragments at the
// triangle's offset (triangle order, row-major inside the triangle).  Offsets are global over the
// batch; fend = {the batch's fragment total, the fill limit}: a triangle whose fragments end past
// the limit is not written (the limit is the total, or 0 when a captured plan found a fragment
// count above its list's capacity).
__global__ void __launch_bounds__(kThreads)
    vox_fill(const VoxSeg *__restrict__ segs, const uint64_t *__restrict__ tbeg, uint32_t S, uint64_t T,
             const uint64_t *__restrict__ offsets, const uint64_t *__restrict__ fend) {
	const uint64_t Ftot = fend[0], lim = fend[1];
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	uint32_t hint = 0;
	for (uint64_t base = (uint64_t)blockIdx.x * kThreads; base < T; base += stride) {
		const uint64_t t = base + threadIdx.x;
		if (t >= T)
			continue;
		uint64_t k = offsets[t];
		const uint64_t kend = (t + 1 < T) ? offsets[t + 1] : Ftot;
		if (k == kend || kend - k > kBigTri || kend > lim)
			continue;
		const float *xyz;
		const uint32_t *rgb;
		uint2 *frags;
		uint64_t tl, fb; // fb: offsets are global over the batch
		uint32_t level;
		with_seg<kThreads>(tbeg, S, hint, base, t, [&](uint32_t sg) {
			const VoxSeg &g = segs[sg];
			xyz = g.xyz;
			rgb = g.rgb;
			frags = g.frags;
			fb = g.frag_begin;
			level = g.level;
			tl = t - tbeg[sg];
		});
		const uint32_t res = 1u << level;
		Setup s;
		tri_setup(xyz, tl, res, s);
		const uint32_t colour = rgb[tl] & 0x00ffffffu;
		RowSpans rs;
		int32_t j0, j1;
		rs.init(s, j0, j1);
		for (int32_t j = j0; j <= j1 && k < kend; ++j) {
			int32_t i0, i1;
			rs.span(i0, i1);
			rs.next();
			const float dy = s.dzdy * (((float)j + 0.5f) - s.py0);
			for (int32_t i = i0; i <= i1 && k < kend; ++i)
				frags[(k++) - fb] = make_frag(s, i, j, dy, res, colour);
		}
	}
}

// Fill of the big triangles: one block per triangle.  Rows are taken kRowChunk at a time: one
// thread per row computes the row's span into LDS, a block scan gives each row's first output
// position, then every thread maps output positions p = tid, tid + 256, ... to (row, column) by a
// binary search over the row offsets, so stores are coalesced and the work is even whatever the
// triangle's shape.
constexpr int kRowChunk = 1024;

__global__ void __launch_bounds__(kThreads)
    vox_fill_big(const VoxSeg *__restrict__ segs, const uint64_t *__restrict__ tbeg, uint32_t S, uint64_t T,
                 const uint64_t *__restrict__ offsets, const uint32_t *__restrict__ big,
                 const uint64_t *__restrict__ fend) {
	__shared__ int32_t s_i0[kRowChunk];
	__shared__ uint32_t s_off[kRowChunk + 1];
	__shared__ uint32_t s_wave[kThreads / 64];
	const uint64_t Ftot = fend[0], lim = fend[1];
	const uint32_t nbig = big[0];
	for (uint32_t b = blockIdx.x; b < nbig; b += gridDim.x) {
		const uint64_t t = big[1 + b];
		if (t >= T)
			continue;
		uint64_t k = offsets[t];
		const uint64_t kend = (t + 1 < T) ? offsets[t + 1] : Ftot;
		if (kend > lim) // block-uniform: t is the block's triangle
			continue;
		const uint32_t sg = S == 1u ? 0u : seg_find(tbeg, S, t);
		const VoxSeg &g = segs[sg];
		const uint64_t tl = t - tbeg[sg];
		const uint32_t res = 1u << g.level;
		uint2 *frags = g.frags;
		const uint64_t fb = g.frag_begin; // offsets are global over the batch
		Setup s;
		tri_setup(g.xyz, tl, res, s);
		const uint32_t colour = g.rgb[tl] & 0x00ffffffu;
		for (int32_t r0 = s.jlo; r0 <= s.jhi && k < kend; r0 += kRowChunk) {
			const int32_t nrows = (s.jhi - r0 + 1) < kRowChunk ? (s.jhi - r0 + 1) : kRowChunk;
			// spans and counts, kRowChunk / kThreads rows per thread
			uint32_t local = 0;
			int32_t cnt[kRowChunk / kThreads];
#pragma unroll
			for (int q = 0; q < kRowChunk / kThreads; ++q) {
				const int32_t r = (int32_t)threadIdx.x * (kRowChunk / kThreads) + q;
				int32_t i0 = 1, i1 = 0;
				if (r < nrows)
					row_span(s, r0 + r, i0, i1);
				cnt[q] = i0 <= i1 ? i1 - i0 + 1 : 0;
				if (r < nrows)
					s_i0[r] = i0;
				local += (uint32_t)cnt[q];
			}
			// block exclusive scan of the per-thread totals
			const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
			uint32_t incl = local;
#pragma unroll
			for (int d = 1; d < 64; d <<= 1) {
				uint32_t o = __shfl_up(incl, (unsigned)d, 64);
				if ((int)lane >= d)
					incl += o;
			}
			if (lane == 63u)
				s_wave[wave] = incl;
			__syncthreads();
			uint32_t base = 0, total = 0;
#pragma unroll
			for (int w = 0; w < kThreads / 64; ++w) {
				base += (w < (int)wave) ? s_wave[w] : 0u;
				total += s_wave[w];
			}
			uint32_t run = base + incl - local;
#pragma unroll
			for (int q = 0; q < kRowChunk / kThreads; ++q) {
				const int32_t r = (int32_t)threadIdx.x * (kRowChunk / kThreads) + q;
				s_off[r] = run; // rows past nrows get the chunk total and are never selected
				run += (uint32_t)cnt[q];
			}
			if (threadIdx.x == 0)
				s_off[kRowChunk] = total;
			__syncthreads();
			const uint32_t avail = (uint32_t)((kend - k) < (uint64_t)total ? (kend - k) : total);
			for (uint32_t p = threadIdx.x; p < avail; p += kThreads) {
				// last row with s_off[row] <= p
				int32_t lo = 0, hi = nrows - 1;
				while (lo < hi) {
					const int32_t mid = (lo + hi + 1) >> 1;
					if (s_off[mid] <= p)
						lo = mid;
					else
						hi = mid - 1;
				}
				const int32_t j = r0 + lo;
				const int32_t i = s_i0[lo] + (int32_t)(p - s_off[lo]);
				const float dy = s.dzdy * (((float)j + 0.5f) - s.py0);
				frags[k + p - fb] = make_frag(s, i, j, dy, res, colour);
			}
			k += total;
			__syncthreads();
		}
	}
}

struct CountSrc {
	static constexpr int kItems = 4;
	const uint64_t *counts;
	uint64_t n;
	__device__ uint64_t size() const { return n; }
	__device__ uint32_t operator()(uint64_t i) const { return (uint32_t)counts[i]; }
};
struct TotalCommit {
	uint64_t *total;
	__device__ void operator()(uint64_t t) const { *total = t; }
};
struct OffsetDst {
	uint64_t *offsets;
	__device__ bool active() const { return true; }
	__device__ void operator()(uint64_t i, uint64_t k, uint32_t) const { offsets[i] = k; }
};

// ---- normalisation (src/Scene.cpp:62-99) -------------------------------------------------------
// Per-block bounding box of the vertices -> parts[6 * b].
__global__ void __launch_bounds__(kThreads) norm_bbox(const float *__restrict__ xyz, uint64_t nverts, float *parts) {
	float mn[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
	float mx[3] = {-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	for (uint64_t v = (uint64_t)blockIdx.x * kThreads + threadIdx.x; v < nverts; v += stride) {
#pragma unroll
		for (int c = 0; c < 3; ++c) {
			float p = xyz[3 * v + c];
			mn[c] = mn[c] < p ? mn[c] : p;
			mx[c] = p < mx[c] ? mx[c] : p;
		}
	}
	__shared__ float s[6][kThreads];
#pragma unroll
	for (int c = 0; c < 3; ++c) {
		s[c][threadIdx.x] = mn[c];
		s[3 + c][threadIdx.x] = mx[c];
	}
	__syncthreads();
	for (int w = kThreads / 2; w > 0; w >>= 1) {
		if ((int)threadIdx.x < w) {
#pragma unroll
			for (int c = 0; c < 3; ++c) {
				float a = s[c][threadIdx.x], b = s[c][threadIdx.x + w];
				s[c][threadIdx.x] = a < b ? a : b;
				float x = s[3 + c][threadIdx.x], y = s[3 + c][threadIdx.x + w];
				s[3 + c][threadIdx.x] = y < x ? x : y;
			}
		}
		__syncthreads();
	}
	if (threadIdx.x < 6)
		parts[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

// Combines the per-block boxes and computes {center.xyz, inv_extent} into out[4].
__global__ void norm_params(const float *parts, uint32_t nparts, float *out) {
	if (threadIdx.x != 0)
		return;
	float mn[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
	float mx[3] = {-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
	for (uint32_t b = 0; b < nparts; ++b)
		for (int c = 0; c < 3; ++c) {
			float a = parts[6 * b + c], x = parts[6 * b + 3 + c];
			mn[c] = mn[c] < a ? mn[c] : a;
			mx[c] = x < mx[c] ? mx[c] : x;
		}
	float ex = mx[0] - mn[0], ey = mx[1] - mn[1], ez = mx[2] - mn[2];
	float extent = gmax(ex, gmax(ey, ez)) * 0.5f;
	out[0] = (mx[0] + mn[0]) * 0.5f;
	out[1] = (mx[1] + mn[1]) * 0.5f;
	out[2] = (mx[2] + mn[2]) * 0.5f;
	out[3] = 1.0f / extent;
}

__global__ void __launch_bounds__(kThreads) norm_apply(float *xyz, uint64_t nverts, const float *params) {
	const float cx = params[0], cy = params[1], cz = params[2], inv = params[3];
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	for (uint64_t v = (uint64_t)blockIdx.x * kThreads + threadIdx.x; v < nverts; v += stride) {
		xyz[3 * v + 0] = (xyz[3 * v + 0] - cx) * inv;
		xyz[3 * v + 1] = (xyz[3 * v + 1] - cy) * inv;
		xyz[3 * v + 2] = (xyz[3 * v + 2] - cz) * inv;
	}
}

uint32_t grid_for(uint64_t items, const svo_ctx *ctx) {
	uint64_t g = (items + kThreads - 1) / kThreads;
	uint64_t cap = (uint64_t)ctx->cu_count * 64u; // 64 blocks per CU: more loads in flight than the CU holds at once
	if (g > cap)
		g = cap;
	return (uint32_t)(g ? g : 1);
}

} // namespace

int normalize_scene(svo_ctx *ctx, svo_scene *scene) {
	const uint64_t nverts = scene->T * 3u;
	if (nverts == 0)
		return SVO_OK;
	// parts: 6 floats per block + 4 result floats, carved out of the scan scratch (2 * kScanBlocks u64)
	float *parts = reinterpret_cast<float *>(ctx->d_scan_parts);
	const uint32_t blocks = 256; // 256 * 6 + 4 floats < 2 * kScanBlocks * 2 floats
	float *params = parts + 6 * blocks;
	hipError_t e;
	ctx->t_normalize.start(ctx->stream);
	if ((e = launch(ctx, "norm_bbox", norm_bbox, dim3(blocks), dim3(kThreads), 0, (const float *)scene->xyz, nverts,
	                parts)) != hipSuccess)
		return hip_check(ctx, e, "norm_bbox");
	if ((e = launch(ctx, "norm_params", norm_params, dim3(1), dim3(64), 0, (const float *)parts, blocks, params)) !=
	    hipSuccess)
		return hip_check(ctx, e, "norm_params");
	if ((e = launch(ctx, "norm_apply", norm_apply, dim3(grid_for(nverts, ctx)), dim3(kThreads), 0, scene->xyz, nverts,
	                (const float *)params)) != hipSuccess)
		return hip_check(ctx, e, "norm_apply");
	ctx->t_normalize.stop(ctx->stream);
	return SVO_OK;
}

void vox_batch_release(svo_ctx *ctx, VoxBatch *vb) {
	for (void *p : {(void *)vb->segs, (void *)vb->offsets}) // tbeg and big live in segs' allocation
		ctx->pool->free(p);
	vb->segs = nullptr;
	vb->tbeg = nullptr;
	vb->fend = nullptr;
	vb->offsets = nullptr;
	vb->big = nullptr;
}

// Count pass over a batch + scan + one read-back of every scene's F + allocation of the lists
// (Voxelizer::count_and_create_fragment_list, src/Voxelizer.cpp:134-179).
int voxelize_count_batch(svo_ctx *ctx, svo_fraglist *const *fls, uint32_t S, VoxBatch *vb) {
	if (S == 0)
		return SVO_OK;
	if (S > kMaxBatch)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "batch larger than 1024 scenes");
	vb->S = S;
	vb->h_segs.assign(S, VoxSeg{});
	vb->h_tbeg.assign(S + 1, 0);
	uint64_t T = 0;
	for (uint32_t s = 0; s < S; ++s) {
		const svo_scene *sc = fls[s]->scene;
		VoxSeg &g = vb->h_segs[s];
		g.xyz = sc->xyz;
		g.rgb = sc->rgb;
		g.T = sc->T;
		g.level = fls[s]->level;
		vb->h_tbeg[s] = T;
		T += sc->T;
	}
	vb->h_tbeg[S] = T;
	vb->Ttot = T;
	if (T > 0xFFFFFFFFull)
		return set_error(ctx, SVO_ERR_OVERFLOW, "more than 2^32-1 triangles in one voxelize call");
	Pool *pool = ctx->pool;
	hipError_t e;
	// descriptors, triangle begins (+ the fill's {total, limit}) and the big-triangle list share one
	// allocation, so one copy uploads the tables and clears the list's count
	const size_t seg_bytes = (sizeof(VoxSeg) * S + 255) & ~size_t(255);
	const size_t tbeg_bytes = (sizeof(uint64_t) * (S + 3) + 255) & ~size_t(255);
	if (pool->alloc((void **)&vb->segs, seg_bytes + tbeg_bytes + sizeof(uint32_t) * (T + 1)) != hipSuccess ||
	    pool->alloc((void **)&vb->offsets, sizeof(uint64_t) * (T ? T : 1)) != hipSuccess)
		return set_error(ctx, SVO_ERR_OUT_OF_MEMORY, "count-pass allocation failed");
	vb->tbeg = reinterpret_cast<uint64_t *>(reinterpret_cast<uint8_t *>(vb->segs) + seg_bytes);
	vb->fend = vb->tbeg + S + 1;
	vb->big = reinterpret_cast<uint32_t *>(reinterpret_cast<uint8_t *>(vb->segs) + seg_bytes + tbeg_bytes);
	{
		std::vector<uint8_t> tables(seg_bytes + tbeg_bytes + sizeof(uint32_t), 0);
		std::memcpy(tables.data(), vb->h_segs.data(), sizeof(VoxSeg) * S);
		std::memcpy(tables.data() + seg_bytes, vb->h_tbeg.data(), sizeof(uint64_t) * (S + 1));
		if (int rc = upload_table(ctx, vb->segs, tables.data(), tables.size()))
			return rc; // big[0] = 0 came with the tables
	}
	int rc = SVO_OK;
	uint64_t *total = ctx->d_scan_parts + 2 * kScanBlocks;
	ctx->t_count.start(ctx->stream);
	if (T > 0) {
		if ((e = launch(ctx, "vox_count", vox_count, dim3(grid_for(T, ctx)), dim3(kThreads), 0,
		                (const VoxSeg *)vb->segs, (const uint64_t *)vb->tbeg, S, T, vb->offsets, vb->big)) !=
		    hipSuccess)
			return hip_check(ctx, e, "vox_count");
		CountSrc src{vb->offsets, T};
		if ((e = launch(ctx, "vox_scan_reduce", scan_reduce<CountSrc>, dim3(kScanBlocks), dim3(kScanThreads), 0, src,
		                ctx->d_scan_parts)) != hipSuccess)
			return hip_check(ctx, e, "vox_scan_reduce");
		if ((e = launch(ctx, "vox_scan_down", scan_down<CountSrc, OffsetDst, TotalCommit>, dim3(kScanBlocks),
		                dim3(kScanThreads), 0, src, OffsetDst{vb->offsets}, (const uint64_t *)ctx->d_scan_parts,
		                TotalCommit{total})) != hipSuccess)
			return hip_check(ctx, e, "vox_scan_down");
	} else if ((e = launch(ctx, "vox_zero", vox_zero_counter, dim3(1), dim3(1), 0,
	                       reinterpret_cast<uint32_t *>(total))) != hipSuccess ||
	           (e = launch(ctx, "vox_zero", vox_zero_counter, dim3(1), dim3(1), 0,
	                       reinterpret_cast<uint32_t *>(total) + 1)) != hipSuccess) {
		return hip_check(ctx, e, "vox_zero");
	}
	if ((e = launch(ctx, "vox_seg_totals", vox_seg_totals, dim3(div_up(S, 64)), dim3(64), 0, vb->segs,
	                (const uint64_t *)vb->tbeg, S, T, (const uint64_t *)vb->offsets, (const uint64_t *)total)) !=
	    hipSuccess)
		return hip_check(ctx, e, "vox_seg_totals");
	ctx->t_count.stop(ctx->stream);
	if ((rc = download_table(ctx, vb->h_segs.data(), vb->segs, sizeof(VoxSeg) * S)) != SVO_OK)
		return rc;
	uint64_t Ftot = 0;
	for (uint32_t s = 0; s < S; ++s) {
		VoxSeg &g = vb->h_segs[s];
		if (g.F > 0xFFFFFFFFull)
			return set_error(ctx, SVO_ERR_OVERFLOW, "more than 2^32-1 fragments in scene " + std::to_string(s));
		svo_fraglist *fl = fls[s];
		fl->F = g.F;
		if (pool->alloc((void **)&fl->frags, sizeof(uint32_t) * 2 * (g.F ? g.F : 1)) != hipSuccess)
			return set_error(ctx, SVO_ERR_OUT_OF_MEMORY, "fragment list allocation failed");
		g.frags = reinterpret_cast<uint2 *>(fl->frags);
		Ftot += g.F;
	}
	// the descriptors with their lists, and the fill's {total, limit} = {Ftot, Ftot}
	std::vector<uint8_t> tables(seg_bytes + tbeg_bytes, 0);
	std::memcpy(tables.data(), vb->h_segs.data(), sizeof(VoxSeg) * S);
	vb->h_tbeg.push_back(Ftot);
	vb->h_tbeg.push_back(Ftot);
	std::memcpy(tables.data() + seg_bytes, vb->h_tbeg.data(), sizeof(uint64_t) * (S + 3));
	vb->h_tbeg.resize(S + 1);
	return upload_table(ctx, vb->segs, tables.data(), tables.size());
}

// Fill pass of a counted batch (CmdVoxelize, src/Voxelizer.cpp:166-179).
int voxelize_fill_batch(svo_ctx *ctx, VoxBatch *vb) {
	const uint64_t T = vb->Ttot;
	const uint32_t S = vb->S;
	uint64_t Ftot = 0;
	for (uint32_t s = 0; s < S; ++s)
		Ftot += vb->h_segs[s].F;
	hipError_t e;
	ctx->t_fill.start(ctx->stream);
	if (T > 0 && Ftot > 0) {
		if ((e = launch(ctx, "vox_fill", vox_fill, dim3(grid_for(T, ctx)), dim3(kThreads), 0, (const VoxSeg *)vb->segs,
		                (const uint64_t *)vb->tbeg, S, T, (const uint64_t *)vb->offsets, (const uint64_t *)vb->fend)) !=
		    hipSuccess)
			return hip_check(ctx, e, "vox_fill");
		if ((e = launch(ctx, "vox_fill_big", vox_fill_big, dim3((uint32_t)ctx->cu_count * 4u), dim3(kThreads), 0,
		                (const VoxSeg *)vb->segs, (const uint64_t *)vb->tbeg, S, T, (const uint64_t *)vb->offsets,
		                (const uint32_t *)vb->big, (const uint64_t *)vb->fend)) != hipSuccess)
			return hip_check(ctx, e, "vox_fill_big");
	}
	ctx->t_fill.stop(ctx->stream);
	return SVO_OK;
}

// Tables of a captured plan's voxelize pass: the lists are allocated by the plan at their
// capacities, so the descriptors (scene buffers, list pointers, levels) never change; F and
// frag_begin are rewritten by every count pass.
int voxelize_plan_init(svo_ctx *ctx, VoxBatch *vb, svo_fraglist *const *fls, uint32_t S) {
	vb->S = S;
	vb->h_segs.assign(S, VoxSeg{});
	vb->h_tbeg.assign(S + 1, 0);
	uint64_t T = 0;
	for (uint32_t s = 0; s < S; ++s) {
		VoxSeg &g = vb->h_segs[s];
		g.xyz = fls[s]->scene->xyz;
		g.rgb = fls[s]->scene->rgb;
		g.T = fls[s]->scene->T;
		g.level = fls[s]->level;
		g.frags = reinterpret_cast<uint2 *>(fls[s]->frags);
		vb->h_tbeg[s] = T;
		T += g.T;
	}
	vb->h_tbeg[S] = T;
	vb->Ttot = T;
	if (T == 0 || T > 0xFFFFFFFFull)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "a captured plan needs 1 to 2^32-1 triangles in all");
	const size_t seg_bytes = (sizeof(VoxSeg) * S + 255) & ~size_t(255);
	const size_t tbeg_bytes = (sizeof(uint64_t) * (S + 3) + 255) & ~size_t(255);
	if (ctx->pool->alloc((void **)&vb->segs, seg_bytes + tbeg_bytes + sizeof(uint32_t) * (T + 1)) != hipSuccess ||
	    ctx->pool->alloc((void **)&vb->offsets, sizeof(uint64_t) * T) != hipSuccess)
		return set_error(ctx, SVO_ERR_OUT_OF_MEMORY, "plan voxelize allocation failed");
	vb->tbeg = reinterpret_cast<uint64_t *>(reinterpret_cast<uint8_t *>(vb->segs) + seg_bytes);
	vb->fend = vb->tbeg + S + 1;
	vb->big = reinterpret_cast<uint32_t *>(reinterpret_cast<uint8_t *>(vb->segs) + seg_bytes + tbeg_bytes);
	std::vector<uint8_t> tables(seg_bytes + tbeg_bytes, 0);
	std::memcpy(tables.data(), vb->h_segs.data(), sizeof(VoxSeg) * S);
	std::memcpy(tables.data() + seg_bytes, vb->h_tbeg.data(), sizeof(uint64_t) * (S + 1));
	return upload_table(ctx, vb->segs, tables.data(), tables.size());
}

// Count pass of a plan: big-triangle list cleared by a kernel (the pass may be captured, so no
// memset), counts, scan into offsets and fend[0] = total, per-scene F and frag_begin.
hipError_t voxelize_plan_count(svo_ctx *ctx, VoxBatch *vb) {
	const uint64_t T = vb->Ttot;
	const uint32_t S = vb->S;
	hipError_t e;
	if ((e = launch(ctx, "vox_zero", vox_zero_counter, dim3(1), dim3(1), 0, vb->big)) != hipSuccess ||
	    (e = launch(ctx, "vox_count", vox_count, dim3(grid_for(T, ctx)), dim3(kThreads), 0, (const VoxSeg *)vb->segs,
	                (const uint64_t *)vb->tbeg, S, T, vb->offsets, vb->big)) != hipSuccess)
		return e;
	CountSrc src{vb->offsets, T};
	if ((e = launch(ctx, "vox_scan_reduce", scan_reduce<CountSrc>, dim3(kScanBlocks), dim3(kScanThreads), 0, src,
	                ctx->d_scan_parts)) != hipSuccess ||
	    (e = launch(ctx, "vox_scan_down", scan_down<CountSrc, OffsetDst, TotalCommit>, dim3(kScanBlocks),
	                dim3(kScanThreads), 0, src, OffsetDst{vb->offsets}, (const uint64_t *)ctx->d_scan_parts,
	                TotalCommit{vb->fend})) != hipSuccess)
		return e;
	return launch(ctx, "vox_seg_totals", vox_seg_totals, dim3(div_up(S, 64)), dim3(64), 0, vb->segs,
	              (const uint64_t *)vb->tbeg, S, T, (const uint64_t *)vb->offsets, (const uint64_t *)vb->fend);
}

// Fill pass of a plan: writes nothing when fend[1] (set between the passes) is 0.
hipError_t voxelize_plan_fill(svo_ctx *ctx, VoxBatch *vb) {
	const uint64_t T = vb->Ttot;
	const uint32_t S = vb->S;
	hipError_t e;
	if ((e = launch(ctx, "vox_fill", vox_fill, dim3(grid_for(T, ctx)), dim3(kThreads), 0, (const VoxSeg *)vb->segs,
	                (const uint64_t *)vb->tbeg, S, T, (const uint64_t *)vb->offsets, (const uint64_t *)vb->fend)) !=
	    hipSuccess)
		return e;
	return launch(ctx, "vox_fill_big", vox_fill_big, dim3((uint32_t)ctx->cu_count * 4u), dim3(kThreads), 0,
	              (const VoxSeg *)vb->segs, (const uint64_t *)vb->tbeg, S, T, (const uint64_t *)vb->offsets,
	              (const uint32_t *)vb->big, (const uint64_t *)vb->fend);
}

} // namespace svo
