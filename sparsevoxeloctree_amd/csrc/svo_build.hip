// Octree construction (OctreeBuilder::CmdBuild, src/OctreeBuilder.cpp:142-210) for a batch of
// scenes: every kernel covers all scenes of the batch in one launch (segment tables, svo_batch.hpp).
//
// The canonical breadth-first node order (child blocks in ascending parent word index) is the
// Morton order of the node prefixes at every depth: a block's 8 children sit in slot order
// x | y << 1 | z << 2 and blocks follow their parents.  So the whole internal structure follows
// from which prefixes are occupied, and the build works on occupancy bitmaps instead of one
// fragment pass per level:
//
// Reference pass              -> here
//   octree_tag_node.comp      -> bm_mark: each fragment sets the bit of its depth-(L-1) node in
//   (levels 1..L-1)              bitmap B[L-1]; bm_reduce ORs each byte of B[d+1] (8 children) into
//                                one bit of B[d], d = L-2..1
//   octree_alloc_node.comp    -> one exclusive scan over the set-bit counts of all bitmaps' 8-word
//                                chunks (stored by bm_reduce while it reads them)
//                                (rank of every occupied node among its depth = its block among
//                                the allocated children blocks); child pointer = 8 * (first block
//                                of depth + rank)
//   octree_modify_arg.comp    -> bm_levels: per-depth counts -> first block of every depth, the
//                                leaf blocks and W = (counter + 1) * 8
//   octree_init_node.comp     -> bm_write_blocks writes every block whole (pointers and zeros)
//   octree_tag_node.comp      -> build_tag_leaf: leaf block ("bin") of each fragment = rank of its
//   (level L) + leaf CAS fold    depth-(L-1) node; count per bin, scan, scatter into per-bin buckets;
//                                build_fold{,_mid,_big} sort each bucket by the 64-bit fragment and
//                                fold the 8 leaves of the block in that order (the reference's order
//                                is whatever its atomics produce), writing the leaf block whole.
// The node words are identical to the reference's pass structure run with a serial allocator.
//
// Batch layout: the scenes' fragments, bitmap words and leaf blocks each form one index space
// (begin tables fbeg, dw/dr per depth, bbeg).  Every depth's bitmap starts on an 8-word boundary;
// the rank array R holds one entry per 8-word chunk (set bits of all bitmaps before the chunk), so
// it is 1/8 of the bitmaps' size, and a node's rank is its chunk's entry plus the popcounts of the
// chunk's words before it.  Every rank is a difference of two R-based counts of the same scene,
// so modular uint32 arithmetic stays exact.
#include <algorithm>
#include <cstring>
#include <vector>
#include <string>

#include "svo_batch.hpp"
#include "svo_device.hpp"
#include "svo_runtime.hpp"
#include "svo_scan.hpp"

namespace svo {
namespace {

constexpr int kThreads = 256;
constexpr uint32_t kLdsBucket = 2048; // largest bucket sorted in LDS
constexpr uint32_t kRegBucket = 16;   // leaf blocks up to this size: one thread each (build_fold)
constexpr uint32_t kWaveBucket = 64;  // up to this size: one wave each (build_fold_mid)
constexpr uint32_t kWideBucket = 512; // up to this size: one wave each, sorted in LDS (build_fold_wide)

// Effective voxel coordinate seen by the reference's descent: the coordinate's bits, saturating
// at res-1 for coordinates >= res (shader/octree_tag_node.comp:18-30).
__device__ __forceinline__ void decode_frag(uint2 fr, uint32_t res, uint32_t &x, uint32_t &y, uint32_t &z) {
	x = fr.x & 0xfffu;
	y = (fr.x >> 12u) & 0xfffu;
	z = (fr.x >> 24u) | ((fr.y >> 28u) << 8u);
	x = umin32(x, res - 1u);
	y = umin32(y, res - 1u);
	z = umin32(z, res - 1u);
}

// Every third bit, for 12-bit inputs (36-bit output).
__device__ __forceinline__ uint64_t spread3(uint32_t v) {
	uint64_t x = v & 0xfffu;
	x = (x | (x << 16)) & 0x0000FF0000FFull;
	x = (x | (x << 8)) & 0x00F00F00F00Full;
	x = (x | (x << 4)) & 0x0C30C30C30C3ull;
	x = (x | (x << 2)) & 0x249249249249ull;
	return x;
}

// Morton code of a leaf: the child slots (x | y << 1 | z << 2) of depths 1..L, depth 1 highest.
__device__ __forceinline__ uint64_t morton(uint32_t x, uint32_t y, uint32_t z) {
	return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

__device__ __forceinline__ void cmp_swap(uint64_t &x, uint64_t &y) {
	const uint64_t lo = y < x ? y : x, hi = y < x ? x : y;
	x = lo;
	y = hi;
}

__device__ __forceinline__ uint64_t frag_key(uint2 f) { return ((uint64_t)f.x << 32) | f.y; }

// Set bits of the words of an 8-word chunk before word j (< 8), and word j (one 32-byte load).
__device__ __forceinline__ uint32_t chunk_before(const uint32_t *__restrict__ chunk, uint32_t j, uint32_t &word) {
	const uint4 *c = reinterpret_cast<const uint4 *>(chunk);
	const uint4 a = c[0], b = c[1];
	const uint32_t v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
	uint32_t r = 0u, cur = v[0];
#pragma unroll
	for (uint32_t t = 0; t < 8u; ++t) {
		r += t < j ? (uint32_t)__popc(v[t]) : 0u;
		cur = t == j ? v[t] : cur;
	}
	word = cur;
	return r;
}

// Set bits of the bitmaps before word w (a global word index) and the word itself: the rank of
// w's 8-word chunk plus the popcounts of the chunk's words before w.
__device__ __forceinline__ uint32_t bits_before(const uint32_t *__restrict__ B, const uint32_t *__restrict__ R,
                                                uint64_t w, uint32_t &word) {
	return R[w >> 3] + chunk_before(B + (w & ~7ull), (uint32_t)(w & 7u), word);
}

__global__ void __launch_bounds__(kThreads) bm_zero(uint32_t *B, uint64_t n) {
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride)
		B[i] = 0u;
}

// Fragments per thread and loop trip in the per-fragment passes: their loads are issued together,
// so a wave waits on one chain of dependent memory round trips for kU fragments.
constexpr uint32_t kU = 4;

// Where a fragment's scene keeps what the per-fragment passes read: the fragment itself, the first
// word of the scene's B[L-1] (and of its ranks), the level and the scene's first bin.
struct FragSrc {
	const uint2 *frag;
	uint64_t w0;
	uint32_t L, bin_begin;
};

// Descriptor lookups for the kU fragments base + u * kThreads + threadIdx.x of one loop trip
// (no fragment loads); frag = nullptr past the end.
__device__ __forceinline__ void frag_sources(const BuildSeg *__restrict__ segs, const uint64_t *__restrict__ fbeg,
                                             uint32_t S, uint64_t Ftot, uint64_t base, uint32_t &hint,
                                             FragSrc (&src)[kU]) {
#pragma unroll
	for (uint32_t u = 0; u < kU; ++u) {
		const uint64_t b = base + (uint64_t)u * kThreads, f = b + threadIdx.x;
		src[u] = FragSrc{nullptr, 0u, 1u, 0u};
		if (f < Ftot)
			with_seg<kThreads>(fbeg, S, hint, b, f, [&](uint32_t s) {
				const BuildSeg &g = segs[s];
				const uint32_t L = g.L;
				src[u] = FragSrc{g.frags + (f - fbeg[s]), g.bm + (L > 1u ? g.off[L - 1u] : 0u), L,
				                 (uint32_t)g.bin_begin};
			});
	}
}

// Depth-(L-1) node of a fragment (its leaf block's parent), as a Morton index.
__device__ __forceinline__ uint64_t frag_parent(uint2 fr, uint32_t L) {
	uint32_t x, y, z;
	decode_frag(fr, 1u << L, x, y, z);
	return morton(x, y, z) >> 3;
}

// Open-addressing tables in LDS that merge the kU * kThreads fragments of one loop trip by key
// before anything touches global memory: fragments come in raster order, so lanes next to each
// other alternate between leaf blocks and a wave's runs of equal keys are short, while a trip's
// 1024 fragments hold few distinct keys.
constexpr uint32_t kHash = 2048; // slots, twice the fragments of a trip
constexpr uint32_t kEmpty = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t hash_slot(uint64_t key) {
	return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64 - 11)); // 11 bits = kHash
}

// Each fragment marks its depth-(L-1) node.  The trip's bits are ORed per bitmap word in LDS, then
// each distinct word costs one load and, only when that load does not already show all of its
// bits, one atomic.
__global__ void __launch_bounds__(kThreads)
    bm_mark(const BuildSeg *__restrict__ segs, const uint64_t *__restrict__ fbeg, uint32_t S, uint32_t *B) {
	const uint64_t Ftot = fbeg[S];
	static_assert(kHash == 2048u && kHash >= 2u * kU * kThreads, "hash_slot returns 11 bits");
	__shared__ unsigned long long s_word[kHash];
	__shared__ uint32_t s_bits[kHash];
	for (uint32_t h = threadIdx.x; h < kHash; h += kThreads) {
		s_word[h] = ~0ull;
		s_bits[h] = 0u;
	}
	__syncthreads();
	const uint64_t stride = (uint64_t)gridDim.x * kThreads * kU;
	uint32_t hint = 0;
	for (uint64_t base = (uint64_t)blockIdx.x * kThreads * kU; base < Ftot; base += stride) {
		FragSrc src[kU];
		frag_sources(segs, fbeg, S, Ftot, base, hint, src);
		uint2 fr[kU];
#pragma unroll
		for (uint32_t u = 0; u < kU; ++u)
			fr[u] = src[u].frag ? *src[u].frag : make_uint2(0u, 0u);
#pragma unroll
		for (uint32_t u = 0; u < kU; ++u) {
			if (src[u].frag && src[u].L > 1u) {
				const uint64_t P = frag_parent(fr[u], src[u].L);
				const unsigned long long word = src[u].w0 + (P >> 5); // global bitmap word
				uint32_t h = hash_slot(word);
				for (;;) { // at most kU * kThreads keys in kHash slots: a free or equal slot exists
					const unsigned long long old = atomicCAS(&s_word[h], ~0ull, word);
					if (old == ~0ull || old == word)
						break;
					h = (h + 1u) & (kHash - 1u);
				}
				atomicOr(&s_bits[h], 1u << (uint32_t)(P & 31u));
			}
		}
		__syncthreads();
		uint64_t word[kHash / kThreads];
		uint32_t bits[kHash / kThreads], have[kHash / kThreads];
#pragma unroll
		for (uint32_t k = 0; k < kHash / kThreads; ++k) {
			const uint32_t h = threadIdx.x + k * kThreads;
			word[k] = s_word[h];
			bits[k] = s_bits[h];
			s_word[h] = ~0ull;
			s_bits[h] = 0u;
		}
#pragma unroll
		for (uint32_t k = 0; k < kHash / kThreads; ++k)
			have[k] = word[k] != ~0ull ? B[word[k]] : ~0u;
#pragma unroll
		for (uint32_t k = 0; k < kHash / kThreads; ++k)
			if ((have[k] & bits[k]) != bits[k]) // a stale cached word only costs a redundant atomic
				atomicOr(B + word[k], bits[k]);
		__syncthreads();
	}
}

// Bit q of B[d] = OR of byte q of B[d+1] (node q occupied iff one of its 8 children is); items are
// the words of B[d] of the scenes that have B[d+1] (begin table dr).  Word w of B[d] is made from
// chunk w of B[d+1], so the same thread also stores that chunk's set-bit count into R, which the
// rank scan then turns into ranks in place (the bitmaps are not read again for it).
__global__ void __launch_bounds__(kThreads)
    bm_reduce(const BuildSeg *__restrict__ segs, const uint64_t *__restrict__ dr, uint32_t S, uint64_t n, uint32_t d,
              uint32_t *B, uint32_t *R) {
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	uint32_t hint = 0;
	for (uint64_t base = (uint64_t)blockIdx.x * kThreads; base < n; base += stride) {
		const uint64_t i = base + threadIdx.x;
		if (i >= n)
			continue;
		with_seg<kThreads>(dr, S, hint, base, i, [&](uint32_t s) {
			const BuildSeg &g = segs[s];
			const uint64_t w = i - dr[s];
			// the children of word w's 32 nodes: the 8-word chunk 8w of B[d+1] (B[d+1] is padded to
			// whole chunks; a padding word of a small B[d] may have no chunk)
			uint32_t bits = 0u;
			if (8u * w < g.off[d + 2u] - g.off[d + 1u]) {
				const uint4 *c = reinterpret_cast<const uint4 *>(B + g.bm + g.off[d + 1u] + 8u * w);
				const uint4 a = c[0], b = c[1];
				const uint32_t v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
				uint32_t pop = 0u;
#pragma unroll
				for (uint32_t k = 0; k < 8u; ++k) {
					uint32_t x = v[k] | (v[k] >> 4u); // bit 8q of x: byte q of v[k] is nonzero
					x |= x >> 2u;
					x |= x >> 1u;
					bits |= ((x & 1u) | ((x >> 7u) & 2u) | ((x >> 14u) & 4u) | ((x >> 21u) & 8u)) << (4u * k);
					pop += (uint32_t)__popc(v[k]);
				}
				R[((g.bm + g.off[d + 1u]) >> 3) + w] = pop;
			}
			B[g.bm + g.off[d] + w] = bits;
		});
	}
}

// The small depths d = min(L-2, kTopDepth) .. 1 of one scene per block: what one bm_reduce launch
// per depth does, with B[d] (at most 8^kTopDepth / 32 = 1024 words) kept in LDS from one depth to
// the next.  Also stores the set-bit count of B[1]'s one chunk (word 0; the padding words are zero)
// for every scene with L > 1.
constexpr uint32_t kTopDepth = 5;
constexpr uint32_t kTopWords = 1u << (3u * kTopDepth - 5u);

__global__ void __launch_bounds__(kThreads) bm_reduce_top(const BuildSeg *__restrict__ segs, uint32_t *B, uint32_t *R) {
	__shared__ uint32_t s_b[2][kTopWords]; // B[d+1] of the previous depth, B[d]
	const BuildSeg &g = segs[blockIdx.x];
	const uint32_t L = g.L;
	if (L <= 1u)
		return; // no bitmaps
	const uint32_t d0 = L - 2u < kTopDepth ? L - 2u : kTopDepth;
	uint32_t cur = 0u; // s_b[cur] holds B[d+1] below depth d0
	for (uint32_t d = d0; d >= 1u; --d) {
		const uint32_t nw = (uint32_t)(g.off[d + 1u] - g.off[d]);      // words of B[d], padded
		const uint32_t nc = (uint32_t)(g.off[d + 2u] - g.off[d + 1u]); // words of B[d+1], padded
		for (uint32_t w = threadIdx.x; w < nw; w += kThreads) {
			uint32_t bits = 0u;
			if (8u * w < nc) {
				uint32_t v[8];
				if (d == d0) {
					const uint4 *c = reinterpret_cast<const uint4 *>(B + g.bm + g.off[d + 1u] + 8u * w);
					const uint4 a = c[0], b = c[1];
					v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
				} else {
#pragma unroll
					for (uint32_t k = 0; k < 8u; ++k)
						v[k] = s_b[cur][8u * w + k];
				}
				uint32_t pop = 0u;
#pragma unroll
				for (uint32_t k = 0; k < 8u; ++k) {
					uint32_t x = v[k] | (v[k] >> 4u); // bit 8q of x: byte q of v[k] is nonzero
					x |= x >> 2u;
					x |= x >> 1u;
					bits |= ((x & 1u) | ((x >> 7u) & 2u) | ((x >> 14u) & 4u) | ((x >> 21u) & 8u)) << (4u * k);
					pop += (uint32_t)__popc(v[k]);
				}
				R[((g.bm + g.off[d + 1u]) >> 3) + w] = pop;
			}
			B[g.bm + g.off[d] + w] = bits;
			s_b[cur ^ 1u][w] = bits;
		}
		__syncthreads();
		cur ^= 1u;
	}
	if (threadIdx.x == 0u)
		R[g.bm >> 3] = (uint32_t)__popc(d0 >= 1u ? s_b[cur][0] : B[g.bm]);
}

// Set bits of 8-word chunk i, as stored by rank_top and bm_reduce.
struct ChunkCounts {
	static constexpr int kItems = 4;
	const uint32_t *R;
	uint64_t n; // chunks
	__device__ uint64_t size() const { return n; }
	__device__ uint32_t operator()(uint64_t i) const { return R[i]; }
};
struct RankTotal {
	uint32_t *R;
	uint64_t n;
	__device__ void operator()(uint64_t t) const { R[n] = (uint32_t)t; }
};
struct RankDst {
	uint32_t *R;
	__device__ bool active() const { return true; }
	__device__ void operator()(uint64_t i, uint64_t k, uint32_t) const { R[i] = (uint32_t)k; }
};

// One thread per scene: per-depth node counts -> first block of every depth (fb[0] = 0 is the root
// block, fb[1] = 1), the leaf blocks and W (src/OctreeBuilder.cpp:212-214: W = (counter + 1) * 8).
__global__ void bm_levels(BuildSeg *segs, uint32_t S, const uint32_t *__restrict__ R, uint64_t cap) {
	const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= S)
		return;
	BuildSeg &g = segs[s];
	const uint32_t L = g.L;
	const uint32_t *Rs = R + (g.bm >> 3); // every depth starts a chunk
	uint64_t first = 1;
	g.fb[0] = 0u;
	g.fb[1] = 1u;
	uint64_t nlast = 1; // nodes at depth L-1 (the root when L == 1)
	for (uint32_t d = 1; d < L; ++d) {
		const uint64_t nd = (uint32_t)(Rs[g.off[d + 1u] >> 3] - Rs[g.off[d] >> 3]);
		first += nd;
		g.fb[d + 1u] = (uint32_t)(first < 0xFFFFFFFFull ? first : 0xFFFFFFFFull);
		nlast = nd;
	}
	g.W = 8u * first;
	g.overflow = g.W > cap ? 1u : 0u;
	g.nbins = nlast;
}

// The 8 words of the block of one depth-(d-1) parent: pointers to the blocks of its occupied
// children (bits of its byte in B[d]), zeros otherwise; r = rank of its first child among depth d.
__device__ __forceinline__ void store_parent_block(uint32_t *oct, uint64_t block, uint32_t byte, uint32_t fbd,
                                                   uint32_t r) {
	uint32_t wd[8];
#pragma unroll
	for (uint32_t t = 0; t < 8u; ++t) {
		const bool set = (byte >> t) & 1u;
		wd[t] = set ? (0x80000000u | ((fbd + r) << 3u)) : 0u;
		r += set ? 1u : 0u;
	}
	uint4 *dst = reinterpret_cast<uint4 *>(oct + 8u * block);
	dst[0] = make_uint4(wd[0], wd[1], wd[2], wd[3]);
	dst[1] = make_uint4(wd[4], wd[5], wd[6], wd[7]);
}

// Position of the n-th (from 0) set bit of v (n < popc(v)).
__device__ __forceinline__ uint32_t nth_bit(uint32_t v, uint32_t n) {
	uint32_t pos = 0u;
#pragma unroll
	for (uint32_t half = 16u; half >= 1u; half >>= 1) {
		const uint32_t c = (uint32_t)__popc(v & ((1u << half) - 1u));
		if (n >= c) {
			n -= c;
			pos += half;
			v >>= half;
		}
	}
	return pos;
}

// Blocks of every depth d >= 2 (the children of depth-(d-1) nodes), in one launch: every occupied
// parent's block (the parent's rank among depth d-1) is written whole.  Items are the words of
// B[d-1] of the scenes with L > d, depth after depth (begin table pw: segment k = (d - 2) S + s), so
// the scan covers the parents' bitmaps, 8x smaller than the children's.  A wave whose 64 words lie
// in one segment works cooperatively: the lanes' popcounts are prefix-summed, and lane k of each
// round writes the wave's k-th occupied parent, so every lane has work and the 64 blocks of a round
// are contiguous.  A wave across a segment boundary walks each word's set bits in its own lane.
// Block 0 also writes the root block of every scene with L > 1 (its 8 words point to the depth-1
// blocks of the occupied children, the bits of B[1]).  The same launch clears the nb leaf-block
// counters of build_tag_leaf and the three counts of the fold's work lists.
__global__ void __launch_bounds__(kThreads)
    bm_write_blocks(const BuildSeg *__restrict__ segs, const uint64_t *__restrict__ pw, uint32_t M, uint32_t S,
                    const uint32_t *__restrict__ B, const uint32_t *__restrict__ R, uint32_t *cnt,
                    const uint64_t *__restrict__ nbp, uint32_t *list) {
	const uint64_t n = pw[M], nb = *nbp;
	const uint32_t lane = threadIdx.x & 63u;
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < nb; i += stride)
		cnt[i] = 0u;
	if (blockIdx.x == 0u) {
		if (threadIdx.x < 3u)
			list[threadIdx.x] = 0u;
		for (uint32_t s = threadIdx.x; s < S; s += kThreads) {
			const BuildSeg &g = segs[s];
			if (g.L > 1u) // L == 1: the root block holds the leaves (build_fold)
				store_parent_block(g.words, 0u, B[g.bm + g.off[1]] & 0xffu, g.fb[1], 0u);
		}
	}
	uint32_t hint = 0;
	for (uint64_t base = (uint64_t)blockIdx.x * kThreads; base < n; base += stride) {
		const uint64_t i = base + threadIdx.x;
		uint32_t k = kEmpty;
		if (i < n)
			with_seg<kThreads>(pw, M, hint, base, i, [&](uint32_t t) { k = t; });
		const uint32_t k0 = __shfl(k, 0, 64);
		if (k0 == kEmpty)
			continue; // the wave is past the end (lane 0 holds its first item)
		if (__all(k == k0 || k == kEmpty)) {
			const uint32_t ku = __builtin_amdgcn_readfirstlane(k0);
			const uint32_t d = 2u + ku / S;
			const BuildSeg &g = segs[ku % S];
			const uint64_t offp = g.bm + g.off[d - 1u], offc = g.bm + g.off[d]; // both start chunks
			const uint64_t w0 = (i - lane) - pw[ku]; // lane 0's word
			const uint32_t v = k != kEmpty ? B[offp + w0 + lane] : 0u, c = (uint32_t)__popc(v);
			uint32_t incl = c;
#pragma unroll
			for (uint32_t dd = 1; dd < 64u; dd <<= 1) {
				const uint32_t t = __shfl_up(incl, dd, 64);
				incl += lane >= dd ? t : 0u;
			}
			const uint32_t excl = incl - c, T = __shfl(incl, 63, 64);
			if (T == 0u)
				continue;
			uint32_t first = 0u; // rank of the wave's first parent among depth d-1
			if (lane == 0u) {
				uint32_t dummy;
				first = bits_before(B, R, offp + w0, dummy) - R[offp >> 3];
			}
			const uint64_t blk0 = (uint64_t)g.fb[d - 1u] + __shfl(first, 0, 64);
			const uint32_t fbd = g.fb[d];
			// children of depth d before this word's 8-word chunk of B[d] (read only for occupied
			// words: an empty padding word of a small B[d-1] may have no chunk)
			const uint32_t rc = v != 0u ? R[(offc >> 3) + w0 + lane] - R[offc >> 3] : 0u;
			for (uint32_t t0 = 0; t0 < T; t0 += 64u) {
				const uint32_t j = t0 + lane; // the wave's j-th occupied parent
				uint32_t o = 0u; // owner lane: the last with excl <= j
#pragma unroll
				for (uint32_t step = 32u; step >= 1u; step >>= 1) {
					const uint32_t e = __shfl(excl, (int)(o + step), 64);
					o = e <= j ? o + step : o;
				}
				const uint32_t vo = __shfl(v, (int)o, 64), eo = __shfl(excl, (int)o, 64), ro = __shfl(rc, (int)o, 64);
				if (j < T) {
					const uint32_t b = nth_bit(vo, j - eo); // the parent's bit in the owner's word
					uint32_t cv; // the owner's chunk of B[d]: its word b / 4 holds the parent's byte
					const uint32_t before = chunk_before(B + offc + 8u * (w0 + o), b >> 2u, cv);
					const uint32_t sh = 8u * (b & 3u);
					const uint32_t r = ro + before + (uint32_t)__popc(cv & ((1u << sh) - 1u));
					store_parent_block(g.words, blk0 + j, (cv >> sh) & 0xffu, fbd, r);
				}
			}
			continue;
		}
		if (k == kEmpty)
			continue;
		const uint32_t d = 2u + k / S;
		const BuildSeg &g = segs[k % S];
		const uint64_t w = i - pw[k];
		const uint64_t offp = g.bm + g.off[d - 1u], offc = g.bm + g.off[d]; // both start chunks
		uint32_t v;
		const uint32_t rp = bits_before(B, R, offp + w, v);
		if (v == 0u)
			continue;
		const uint32_t fbd = g.fb[d];
		uint64_t block = (uint64_t)g.fb[d - 1u] + (rp - R[offp >> 3]); // first parent's block
		// the 32 parents' children bytes are the 8-word chunk offc + 8w; parents not in v have zero
		// bytes, so a running count gives every child its rank
		const uint4 *cc = reinterpret_cast<const uint4 *>(B + offc + 8u * w);
		const uint4 c0 = cc[0], c1 = cc[1];
		uint32_t r = R[(offc >> 3) + w] - R[offc >> 3];
		while (v) {
			const uint32_t b = (uint32_t)__ffs(v) - 1u, j = b >> 2u;
			v &= v - 1u;
			uint32_t cv = c0.x;
			cv = j == 1u ? c0.y : cv;
			cv = j == 2u ? c0.z : cv;
			cv = j == 3u ? c0.w : cv;
			cv = j == 4u ? c1.x : cv;
			cv = j == 5u ? c1.y : cv;
			cv = j == 6u ? c1.z : cv;
			cv = j == 7u ? c1.w : cv;
			const uint32_t byte = (cv >> (8u * (b & 3u))) & 0xffu;
			store_parent_block(g.words, block, byte, fbd, r);
			r += (uint32_t)__popc(byte);
			++block;
		}
	}
}

// Leaf block ("bin") of every fragment: the rank of its depth-(L-1) node, offset by the scene's
// first bin; node[f] = {bin, place}, where place is unique inside the bin, so the scatter needs no
// atomics.  Places come from per-bin counters in global memory whose atomics the L2 cannot perform
// (they are device-coherent, so they go to memory), and each such request stalls the L2 for the
// reads behind it.  So a loop trip first counts its kU * kThreads fragments in LDS, per 64-byte line
// of 16 counters (a line table with linear probing; the count's old value is the fragment's place
// among the trip's fragments of its bin), and then adds each line's counts with 16 neighbouring
// lanes, which the L1 merges into one request per line: a trip of 1024 fragments touches about
// 40 lines (a room scene at 2^10) to 150 (random slivers at 2^12) instead of 200+; kLines line slots
// (128 up to level 11, 256 at level 12).  A fragment that finds the table full takes its place with
// its own atomic.
constexpr uint32_t kLineBins = 16; // 4-byte counters per 64-byte line

template <uint32_t kLines>
__global__ void __launch_bounds__(kThreads)
    build_tag_leaf(const BuildSeg *__restrict__ segs, const uint64_t *__restrict__ fbeg, uint32_t S,
                   const uint32_t *__restrict__ B, const uint32_t *__restrict__ R, uint2 *__restrict__ node,
                   uint32_t *cnt) {
	const uint64_t Ftot = fbeg[S];
	__shared__ uint32_t line[kLines];
	__shared__ uint32_t n[kLines * kLineBins];
	for (uint32_t i = threadIdx.x; i < kLines; i += kThreads)
		line[i] = kEmpty;
	for (uint32_t i = threadIdx.x; i < kLines * kLineBins; i += kThreads)
		n[i] = 0u;
	__syncthreads();
	const uint64_t stride = (uint64_t)gridDim.x * kThreads * kU;
	uint32_t hint = 0;
	for (uint64_t base = (uint64_t)blockIdx.x * kThreads * kU; base < Ftot; base += stride) {
		FragSrc src[kU];
		frag_sources(segs, fbeg, S, Ftot, base, hint, src);
		uint2 fr[kU];
#pragma unroll
		for (uint32_t u = 0; u < kU; ++u)
			fr[u] = src[u].frag ? *src[u].frag : make_uint2(0u, 0u);
		// rank of the fragment's depth-(L-1) node among its scene's occupied nodes of that depth
		// (w0, the depth's first word, starts a chunk)
		uint32_t below[kU], r[kU], r0[kU], bin[kU];
#pragma unroll
		for (uint32_t u = 0; u < kU; ++u) {
			r[u] = r0[u] = below[u] = 0u;
			if (src[u].frag && src[u].L > 1u) {
				const uint64_t P = frag_parent(fr[u], src[u].L);
				uint32_t word;
				r[u] = bits_before(B, R, src[u].w0 + (P >> 5), word);
				r0[u] = R[src[u].w0 >> 3];
				below[u] = word & ((1u << (uint32_t)(P & 31u)) - 1u);
			}
		}
		uint32_t slot[kU], local[kU];
#pragma unroll
		for (uint32_t u = 0; u < kU; ++u) {
			bin[u] = src[u].frag ? src[u].bin_begin + (r[u] - r0[u]) + (uint32_t)__popc(below[u]) : kEmpty;
			slot[u] = kEmpty;
			local[u] = 0u;
			if (bin[u] != kEmpty) {
				const uint32_t key = bin[u] / kLineBins;
				static_assert(kLines == 128u || kLines == 256u, "kLines: 7 or 8 hash bits");
				uint32_t h = (key * 2654435761u) >> (kLines == 256u ? 24 : 25), probes = 0;
				for (;;) {
					const uint32_t old = atomicCAS(&line[h], kEmpty, key);
					if (old == kEmpty || old == key) {
						slot[u] = h * kLineBins + bin[u] % kLineBins;
						break;
					}
					if (++probes == kLines)
						break; // table full
					h = (h + 1u) & (kLines - 1u);
				}
				local[u] = slot[u] != kEmpty ? atomicAdd(&n[slot[u]], 1u) : atomicAdd(&cnt[bin[u]], 1u);
			}
		}
		__syncthreads();
		// one atomic per counter with fragments; lanes 16k..16k+15 hold the 16 counters of one line
		constexpr uint32_t kPer = kLines * kLineBins / kThreads, kGroup = 8; // counters per thread, per batch
		static_assert(kPer % kGroup == 0u, "whole batches");
#pragma unroll 1
		for (uint32_t k0 = 0; k0 < kPer; k0 += kGroup) {
			uint32_t key[kGroup], c[kGroup];
#pragma unroll
			for (uint32_t k = 0; k < kGroup; ++k) {
				const uint32_t p = threadIdx.x + (k0 + k) * kThreads;
				key[k] = line[p / kLineBins];
				c[k] = n[p];
			}
#pragma unroll
			for (uint32_t k = 0; k < kGroup; ++k) {
				const uint32_t p = threadIdx.x + (k0 + k) * kThreads;
				if (key[k] != kEmpty && c[k] != 0u)
					n[p] = atomicAdd(&cnt[key[k] * kLineBins + p % kLineBins], c[k]);
			}
		}
		__syncthreads();
#pragma unroll
		for (uint32_t u = 0; u < kU; ++u) {
			const uint64_t f = base + (uint64_t)u * kThreads + threadIdx.x;
			if (bin[u] != kEmpty)
				node[f] = make_uint2(bin[u], slot[u] != kEmpty ? n[slot[u]] + local[u] : local[u]);
		}
		__syncthreads();
		for (uint32_t i = threadIdx.x; i < kLines; i += kThreads)
			line[i] = kEmpty;
		for (uint32_t i = threadIdx.x; i < kLines * kLineBins; i += kThreads)
			n[i] = 0u;
		__syncthreads();
	}
}

// Scatter every fragment to its place in its bin's bucket (aux[j] = first slot of bin j).
__global__ void __launch_bounds__(kThreads)
    build_scatter(const BuildSeg *__restrict__ segs, const uint64_t *__restrict__ fbeg, uint32_t S,
                  const uint2 *__restrict__ node, const uint32_t *__restrict__ aux, uint2 *__restrict__ bucket) {
	const uint64_t Ftot = fbeg[S];
	const uint64_t stride = (uint64_t)gridDim.x * kThreads * kU;
	uint32_t hint = 0;
	for (uint64_t base = (uint64_t)blockIdx.x * kThreads * kU; base < Ftot; base += stride) {
		FragSrc src[kU];
		frag_sources(segs, fbeg, S, Ftot, base, hint, src);
		uint2 fr[kU], np[kU];
#pragma unroll
		for (uint32_t u = 0; u < kU; ++u) {
			const uint64_t f = base + (uint64_t)u * kThreads + threadIdx.x;
			fr[u] = np[u] = make_uint2(0u, 0u);
			if (src[u].frag) {
				fr[u] = *src[u].frag;
				np[u] = node[f];
			}
		}
		uint32_t first[kU];
#pragma unroll
		for (uint32_t u = 0; u < kU; ++u)
			first[u] = src[u].frag ? aux[np[u].x] : 0u;
#pragma unroll
		for (uint32_t u = 0; u < kU; ++u)
			if (src[u].frag)
				bucket[first[u] + np[u].y] = fr[u];
	}
}

// ceil(2^32 / w): floor(a / w) = umulhi(a, kDivMagic[w]) for a < 2^14 and 2 <= w <= 64 (checked
// exhaustively), the only divisions the colour fold needs.
__constant__ uint32_t kDivMagic[65] = {
    0u,         0u,         2147483648u, 1431655766u, 1073741824u, 858993460u, 715827883u, 613566757u, 536870912u,
    477218589u, 429496730u, 390451573u,  357913942u,  330382100u,  306783379u, 286331154u, 268435456u, 252645136u,
    238609295u, 226050911u, 214748365u,  204522253u,  195225787u,  186737709u, 178956971u, 171798692u, 165191050u,
    159072863u, 153391690u, 148102321u,  143165577u,  138547333u,  134217728u, 130150525u, 126322568u, 122713352u,
    119304648u, 116080198u, 113025456u,  110127367u,  107374183u,  104755300u, 102261127u, 99882961u,  97612894u,
    95443718u,  93368855u,  91382283u,   89478486u,   87652394u,   85899346u,  84215046u,  82595525u,  81037119u,
    79536432u,  78090315u,  76695845u,   75350304u,   74051161u,   72796056u,  71582789u,  70409300u,  69273667u,
    68174085u,  67108864u};

// Running average of the reference's CAS loop (octree_tag_node.comp:43-57): fold colour c into
// leaf word prev (0 = empty leaf).  The word's 6-bit count n = min(a, 63) after a fragments is the
// weight the next fragment meets: avg' = (avg * n + c) / (n + 1) per channel, count' = min(n + 1, 63).
// magic = kDivMagic staged in LDS (a per-lane index into __constant__ memory would be a global load).
__device__ __forceinline__ uint32_t fold_word(uint32_t prev, uint32_t c, const uint32_t *magic) {
	if (prev == 0u)
		return 0xC1000000u | c;
	const uint32_t n = (prev >> 24u) & 0x3fu, w = n + 1u, m = magic[w];
	const uint32_t r = __umulhi((prev & 0xffu) * n + (c & 0xffu), m);
	const uint32_t g = __umulhi(((prev >> 8u) & 0xffu) * n + ((c >> 8u) & 0xffu), m);
	const uint32_t b = __umulhi(((prev >> 16u) & 0xffu) * n + ((c >> 16u) & 0xffu), m);
	return (umin32(w, 0x3fu) << 24u) | (r & 0xffu) | ((g & 0xffu) << 8u) | ((b & 0xffu) << 16u) | 0xC0000000u;
}

// Stages the division table in LDS; every thread of the block must call it.
__device__ __forceinline__ void stage_magic(uint32_t *s_magic) {
	for (uint32_t i = threadIdx.x; i < 65u; i += blockDim.x)
		s_magic[i] = kDivMagic[i];
	__syncthreads();
}

// The leaf a fragment key belongs to inside its block (child slot at depth L).
__device__ __forceinline__ uint32_t key_slot(uint64_t key, uint32_t res) {
	uint32_t x, y, z;
	decode_frag(make_uint2((uint32_t)(key >> 32), (uint32_t)key), res, x, y, z);
	return (x & 1u) | ((y & 1u) << 1u) | ((z & 1u) << 2u);
}

// Folds one key (in ascending key order) into the 8 leaf words of a block.  The slot is selected
// without dynamic register indexing.
__device__ __forceinline__ void slot_push(uint32_t (&w)[8], uint64_t key, uint32_t res, const uint32_t *magic) {
	const uint32_t s = key_slot(key, res);
	uint32_t cw = w[0];
#pragma unroll
	for (uint32_t t = 1; t < 8u; ++t)
		cw = s == t ? w[t] : cw;
	const uint32_t nw = fold_word(cw, (uint32_t)key & 0xffffffu, magic);
#pragma unroll
	for (uint32_t t = 0; t < 8u; ++t)
		w[t] = s == t ? nw : w[t];
}

__device__ __forceinline__ void store_block(uint32_t *oct, uint64_t block, const uint32_t (&w)[8]) {
	uint4 *dst = reinterpret_cast<uint4 *>(oct + 8u * block);
	dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
	dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// Work lists of leaf blocks too large for build_fold: list[0..2] = mid, big and wide counts; mid
// blocks (17..64 fragments) at list[3 ..], big ones (> 512) at list[3 + cap ..], wide ones
// (65..512) at list[3 + 2 cap ..] (each list holds at most Ftot / 17).  Every lane of the wave calls
// it (the appends are aggregated per wave).
__device__ __forceinline__ void push_bins(uint32_t *list, uint64_t cap, uint32_t j, bool mid, bool wide, bool big) {
	const uint32_t qm = wave_append(&list[0], mid), qb = wave_append(&list[1], big), qw = wave_append(&list[2], wide);
	if (mid && qm < cap)
		list[3u + qm] = j;
	if (big && qb < cap)
		list[3u + cap + qb] = j;
	if (wide && qw < cap)
		list[3u + 2u * cap + qw] = j;
}

// Octree block of global leaf block j (scene s): the scene's first leaf block + j's rank.
struct LeafTarget {
	uint32_t *oct;
	uint64_t block;
	uint32_t res;
};
__device__ __forceinline__ LeafTarget leaf_target(const BuildSeg &g, const uint64_t *bbeg, uint32_t s, uint32_t j) {
	const uint32_t L = g.L;
	return LeafTarget{g.words, (uint64_t)(L == 1u ? 0u : g.fb[L - 1u]) + (j - bbeg[s]), 1u << L};
}

// One thread per leaf block with at most 16 fragments: sorted by the 64-bit fragment in registers
// (5-, 19- or 63-comparator networks, +inf padding), then folded in that order into the 8 leaves;
// the whole block is written, empty leaves as 0.  Each block of threads first orders its 256 leaf
// blocks by size class in LDS, so that most waves run a single network.
template <uint32_t kN>
__device__ __forceinline__ void fold_small(const uint2 *__restrict__ b, uint32_t n, uint32_t res, const uint32_t *magic,
                                           uint32_t (&w)[8]) {
	uint64_t k[kN];
#pragma unroll
	for (uint32_t a = 0; a < kN; ++a)
		k[a] = a < n ? frag_key(b[a]) : ~0ull;
	if (kN == 4) {
		cmp_swap(k[0], k[1]), cmp_swap(k[2], k[3]), cmp_swap(k[0], k[2]), cmp_swap(k[1], k[3]), cmp_swap(k[1], k[2]);
	} else if (kN == 8) {
		cmp_swap(k[0], k[1]), cmp_swap(k[2], k[3]), cmp_swap(k[4], k[5]), cmp_swap(k[6], k[7]);
		cmp_swap(k[0], k[2]), cmp_swap(k[1], k[3]), cmp_swap(k[4], k[6]), cmp_swap(k[5], k[7]);
		cmp_swap(k[1], k[2]), cmp_swap(k[5], k[6]), cmp_swap(k[0], k[4]), cmp_swap(k[3], k[7]);
		cmp_swap(k[1], k[5]), cmp_swap(k[2], k[6]);
		cmp_swap(k[1], k[4]), cmp_swap(k[3], k[6]);
		cmp_swap(k[2], k[4]), cmp_swap(k[3], k[5]);
		cmp_swap(k[3], k[4]);
	} else { // Batcher odd-even merge sort of 16
		cmp_swap(k[0], k[1]), cmp_swap(k[2], k[3]), cmp_swap(k[0], k[2]), cmp_swap(k[1], k[3]);
		cmp_swap(k[1], k[2]), cmp_swap(k[4], k[5]), cmp_swap(k[6], k[7]), cmp_swap(k[4], k[6]);
		cmp_swap(k[5], k[7]), cmp_swap(k[5], k[6]), cmp_swap(k[0], k[4]), cmp_swap(k[2], k[6]);
		cmp_swap(k[2], k[4]), cmp_swap(k[1], k[5]), cmp_swap(k[3], k[7]), cmp_swap(k[3], k[5]);
		cmp_swap(k[1], k[2]), cmp_swap(k[3], k[4]), cmp_swap(k[5], k[6]), cmp_swap(k[8], k[9]);
		cmp_swap(k[10], k[11]), cmp_swap(k[8], k[10]), cmp_swap(k[9], k[11]), cmp_swap(k[9], k[10]);
		cmp_swap(k[12], k[13]), cmp_swap(k[14], k[15]), cmp_swap(k[12], k[14]), cmp_swap(k[13], k[15]);
		cmp_swap(k[13], k[14]), cmp_swap(k[8], k[12]), cmp_swap(k[10], k[14]), cmp_swap(k[10], k[12]);
		cmp_swap(k[9], k[13]), cmp_swap(k[11], k[15]), cmp_swap(k[11], k[13]), cmp_swap(k[9], k[10]);
		cmp_swap(k[11], k[12]), cmp_swap(k[13], k[14]), cmp_swap(k[0], k[8]), cmp_swap(k[4], k[12]);
		cmp_swap(k[4], k[8]), cmp_swap(k[2], k[10]), cmp_swap(k[6], k[14]), cmp_swap(k[6], k[10]);
		cmp_swap(k[2], k[4]), cmp_swap(k[6], k[8]), cmp_swap(k[10], k[12]), cmp_swap(k[1], k[9]);
		cmp_swap(k[5], k[13]), cmp_swap(k[5], k[9]), cmp_swap(k[3], k[11]), cmp_swap(k[7], k[15]);
		cmp_swap(k[7], k[11]), cmp_swap(k[3], k[5]), cmp_swap(k[7], k[9]), cmp_swap(k[11], k[13]);
		cmp_swap(k[1], k[2]), cmp_swap(k[3], k[4]), cmp_swap(k[5], k[6]), cmp_swap(k[7], k[8]);
		cmp_swap(k[9], k[10]), cmp_swap(k[11], k[12]), cmp_swap(k[13], k[14]);
	}
#pragma unroll
	for (uint32_t a = 0; a < kN; ++a)
		if (a < n)
			slot_push(w, k[a], res, magic);
}

__global__ void __launch_bounds__(kThreads)
    build_fold(const BuildSeg *__restrict__ segs, const uint64_t *__restrict__ bbeg, uint32_t S,
               const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ aux, const uint2 *__restrict__ bucket,
               uint32_t *list, uint64_t cap) {
	const uint64_t nbins = bbeg[S];
	constexpr uint32_t kWaves = kThreads / 64;
	__shared__ uint32_t s_magic[65];
	__shared__ uint32_t s_bin[kThreads], s_n[kThreads];
	__shared__ uint32_t s_cls[kWaves][3];
	stage_magic(s_magic);
	const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
	const uint64_t below = (1ull << lane) - 1ull;
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	uint32_t hint = 0;
	for (uint64_t base = (uint64_t)blockIdx.x * kThreads; base < nbins; base += stride) {
		const uint64_t jj = base + threadIdx.x;
		uint32_t n = 0, cls = 3; // classes: <= 4, <= 8, <= 16 fragments; 3 = not folded here
		if (jj < nbins) {
			n = cnt[jj];
			cls = n <= 4u ? 0u : n <= 8u ? 1u : n <= kRegBucket ? 2u : 3u;
		}
		const bool rest = jj < nbins && cls == 3u;
		push_bins(list, cap, (uint32_t)jj, rest && n <= kWaveBucket, rest && n > kWaveBucket && n <= kWideBucket,
		          rest && n > kWideBucket);
		const uint64_t m0 = __ballot(cls == 0u), m1 = __ballot(cls == 1u), m2 = __ballot(cls == 2u);
		if (lane == 0u) {
			s_cls[wave][0] = (uint32_t)__popcll(m0);
			s_cls[wave][1] = (uint32_t)__popcll(m1);
			s_cls[wave][2] = (uint32_t)__popcll(m2);
		}
		__syncthreads();
		uint32_t pos = 0, total = 0;
#pragma unroll
		for (uint32_t c = 0; c < 3u; ++c)
#pragma unroll
			for (uint32_t v = 0; v < kWaves; ++v) {
				const uint32_t k = s_cls[v][c];
				pos += (c < cls || (c == cls && v < wave)) ? k : 0u;
				total += k;
			}
		if (cls < 3u) {
			const uint64_t mine = cls == 0u ? m0 : cls == 1u ? m1 : m2;
			pos += (uint32_t)__popcll(mine & below);
			s_bin[pos] = (uint32_t)jj;
			s_n[pos] = n;
		}
		__syncthreads();
		if (threadIdx.x < total) {
			const uint32_t j = s_bin[threadIdx.x], nj = s_n[threadIdx.x];
			LeafTarget t;
			with_seg<kThreads>(bbeg, S, hint, base, j, [&](uint32_t s) { t = leaf_target(segs[s], bbeg, s, j); });
			const uint2 *b = bucket + aux[j];
			uint32_t w[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
			if (nj <= 4u)
				fold_small<4>(b, nj, t.res, s_magic, w);
			else if (nj <= 8u)
				fold_small<8>(b, nj, t.res, s_magic, w);
			else
				fold_small<16>(b, nj, t.res, s_magic, w);
			store_block(t.oct, t.block, w);
		}
		__syncthreads();
	}
}

// Bitonic sort of m = 2^p keys (ascending) by one block; every comparator puts the minimum at the
// lower index, so keys past the end behave as +infinity padding.
template <class Get, class Put> __device__ void block_bitonic(uint32_t n, uint32_t m, Get get, Put put) {
	for (uint32_t k = 2; k <= m; k <<= 1) {
		for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
			for (uint32_t t = threadIdx.x; t < m / 2; t += blockDim.x) {
				const uint32_t blk = t / jj, off = t % jj;
				uint32_t a, b;
				if (jj == (k >> 1)) { // first step of a merge: flip
					a = blk * k + off;
					b = blk * k + k - 1u - off;
				} else {
					a = blk * 2u * jj + off;
					b = a + jj;
				}
				if (b < n) {
					const uint64_t va = get(a), vb = get(b);
					if (vb < va) {
						put(a, vb);
						put(b, va);
					}
				}
			}
			__syncthreads();
		}
	}
}

// Leaf blocks of 17..64 fragments, one wave each: lane i holds key i and its leaf slot (padding:
// slot 8, key +inf); a bitonic sort over lane shuffles orders them by (slot, key), so every leaf's
// fragments are consecutive and ascending.  Lane s < 8 then folds the run of slot s from LDS.
// Blocks blk of nblk of the list launch; s_col: 64 words of LDS per wave.
__device__ void fold_mid_blocks(uint32_t blk, uint32_t nblk, const BuildSeg *__restrict__ segs,
                                const uint64_t *__restrict__ bbeg, uint32_t S, const uint32_t *__restrict__ cnt,
                                const uint32_t *__restrict__ aux, const uint2 *__restrict__ bucket, const uint32_t *list,
                                uint64_t cap, uint32_t (*s_col)[64], const uint32_t *s_magic) {
	constexpr uint32_t kWaves = kThreads / 64;
	const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
	uint64_t count = list[0];
	if (count > cap)
		count = cap;
	for (uint64_t base = (uint64_t)blk * kWaves; base < count; base += (uint64_t)nblk * kWaves) {
		const uint64_t item = base + wave;
		if (item >= count)
			break; // wave-uniform; no block barrier follows
		const uint32_t j = list[3u + item];
		const uint32_t n = cnt[j];
		const uint32_t s = S == 1u ? 0u : seg_find(bbeg, S, j);
		const LeafTarget t = leaf_target(segs[s], bbeg, s, j);
		uint64_t key = ~0ull;
		uint32_t slot = 8u;
		if (lane < n) {
			key = frag_key(bucket[aux[j] + lane]);
			slot = key_slot(key, t.res);
		}
#pragma unroll
		for (uint32_t k = 2; k <= 64u; k <<= 1) {
#pragma unroll
			for (uint32_t d = k >> 1; d > 0u; d >>= 1) {
				const uint64_t ok = __shfl_xor(key, (int)d, 64);
				const uint32_t os = __shfl_xor(slot, (int)d, 64);
				const bool other_less = os < slot || (os == slot && ok < key);
				const bool take_min = ((lane & k) == 0u) == ((lane & d) == 0u);
				const bool take_other = take_min == other_less;
				key = take_other ? ok : key;
				slot = take_other ? os : slot;
			}
		}
		s_col[wave][lane] = (uint32_t)key & 0xffffffu;
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
		uint64_t run = 0;
#pragma unroll
		for (uint32_t u = 0; u < 8u; ++u) {
			const uint64_t mu = __ballot(slot == u);
			run = lane == u ? mu : run;
		}
		if (lane < 8u) {
			uint32_t word = 0u;
			if (run) {
				const uint32_t beg = (uint32_t)__ffsll((unsigned long long)run) - 1u, len = (uint32_t)__popcll(run);
				for (uint32_t i = beg; i < beg + len; ++i)
					word = fold_word(word, s_col[wave][i], s_magic);
			}
			t.oct[8u * t.block + lane] = word;
		}
		__builtin_amdgcn_wave_barrier();
	}
}

// Leaf blocks of 65..512 fragments, one wave each: the (slot, key) pairs, padded to a power of two
// with (8, +inf), are bitonic-sorted in the wave's LDS (no block barriers: the wave's LDS accesses
// are ordered between stages by wave barriers), so every leaf's fragments are consecutive and
// ascending; lane s < 8 then folds the run of slot s.

// Blocks blk of nblk of the list launch; s_keys / s_slots: kWideBucket entries per wave.
__device__ void fold_wide_blocks(uint32_t blk, uint32_t nblk, const BuildSeg *__restrict__ segs,
                                 const uint64_t *__restrict__ bbeg, uint32_t S, const uint32_t *__restrict__ cnt,
                                 const uint32_t *__restrict__ aux, const uint2 *__restrict__ bucket, const uint32_t *list,
                                 uint64_t cap, uint64_t *s_keys, uint8_t *s_slots, uint32_t (*s_beg)[8],
                                 uint32_t (*s_end)[8], const uint32_t *s_magic) {
	constexpr uint32_t kWaves = kThreads / 64;
	const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
	uint64_t *keys = s_keys + wave * kWideBucket;
	uint8_t *slots = s_slots + wave * kWideBucket;
	uint64_t count = list[2];
	if (count > cap)
		count = cap;
	for (uint64_t base = (uint64_t)blk * kWaves; base < count; base += (uint64_t)nblk * kWaves) {
		const uint64_t item = base + wave;
		if (item >= count)
			break; // wave-uniform; no block barrier follows
		const uint32_t j = list[3u + 2u * cap + item];
		const uint32_t n = cnt[j];
		const uint32_t s = S == 1u ? 0u : seg_find(bbeg, S, j);
		const LeafTarget t = leaf_target(segs[s], bbeg, s, j);
		const uint2 *b = bucket + aux[j];
		uint32_t m = 128u;
		while (m < n)
			m <<= 1;
		for (uint32_t i = lane; i < m; i += 64u) {
			const uint64_t key = i < n ? frag_key(b[i]) : ~0ull;
			keys[i] = key;
			slots[i] = (uint8_t)(i < n ? key_slot(key, t.res) : 8u);
		}
		if (lane < 8u)
			s_beg[wave][lane] = s_end[wave][lane] = 0u;
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
		for (uint32_t k = 2; k <= m; k <<= 1) {
			for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
				for (uint32_t q = lane; q < m / 2u; q += 64u) {
					const uint32_t blk = q / jj, off = q % jj;
					const uint32_t a = jj == (k >> 1) ? blk * k + off : blk * 2u * jj + off;
					const uint32_t c = jj == (k >> 1) ? blk * k + k - 1u - off : a + jj;
					const uint64_t ka = keys[a], kc = keys[c];
					const uint8_t sa = slots[a], sc = slots[c];
					if (sc < sa || (sc == sa && kc < ka)) {
						keys[a] = kc;
						keys[c] = ka;
						slots[a] = sc;
						slots[c] = sa;
					}
				}
				__builtin_amdgcn_wave_barrier();
				__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
			}
		}
		for (uint32_t i = lane; i < n; i += 64u) {
			const uint32_t sl = slots[i];
			if (i == 0u || slots[i - 1u] != sl)
				s_beg[wave][sl] = i;
			if (i + 1u == n || slots[i + 1u] != sl)
				s_end[wave][sl] = i + 1u;
		}
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
		if (lane < 8u) {
			uint32_t word = 0u;
			for (uint32_t i = s_beg[wave][lane]; i < s_end[wave][lane]; ++i)
				word = fold_word(word, (uint32_t)keys[i] & 0xffffffu, s_magic);
			t.oct[8u * t.block + lane] = word;
		}
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
	}
}

// Bitonic sort of n (<= m = 2^p) (slot, key) pairs held in LDS by one block, ascending; pairs past
// n behave as +infinity padding.
__device__ void block_bitonic_kv(uint32_t n, uint32_t m, uint64_t *keys, uint8_t *slots) {
	for (uint32_t k = 2; k <= m; k <<= 1) {
		for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
			for (uint32_t t = threadIdx.x; t < m / 2; t += blockDim.x) {
				const uint32_t blk = t / jj, off = t % jj;
				uint32_t a, b;
				if (jj == (k >> 1)) { // first step of a merge: flip
					a = blk * k + off;
					b = blk * k + k - 1u - off;
				} else {
					a = blk * 2u * jj + off;
					b = a + jj;
				}
				if (b < n) {
					const uint64_t ka = keys[a], kb = keys[b];
					const uint8_t sa = slots[a], sb = slots[b];
					if (sb < sa || (sb == sa && kb < ka)) {
						keys[a] = kb;
						keys[b] = ka;
						slots[a] = sb;
						slots[b] = sa;
					}
				}
			}
			__syncthreads();
		}
	}
}

// Leaf blocks of more than 512 fragments, one block each.  Up to kLdsBucket fragments: bitonic sort
// of the (slot, key) pairs in LDS, then lane s < 8 folds the run of slot s.  Beyond (a global-memory
// sort, rare): keys sorted in place in the bucket and each of the 8 lanes walks all of them,
// taking those of its slot.  Blocks blk of nblk of the list launch; s_keys / s_slots: kLdsBucket
// entries for the block.
__device__ void fold_big_blocks(uint32_t blk, uint32_t nblk, const BuildSeg *__restrict__ segs,
                                const uint64_t *__restrict__ bbeg, uint32_t S, const uint32_t *__restrict__ cnt,
                                const uint32_t *__restrict__ aux, uint2 *bucket, const uint32_t *list, uint64_t cap,
                                uint64_t *s_keys, uint8_t *s_slots, uint32_t *s_beg, uint32_t *s_end,
                                const uint32_t *s_magic) {
	uint64_t count = list[1];
	if (count > cap)
		count = cap;
	for (uint64_t item = blk; item < count; item += nblk) {
		const uint32_t j = list[3u + cap + item];
		const uint32_t n = cnt[j];
		const uint32_t s = S == 1u ? 0u : seg_find(bbeg, S, j);
		const LeafTarget t = leaf_target(segs[s], bbeg, s, j);
		uint2 *b = bucket + aux[j];
		uint32_t m = 1;
		while (m < n)
			m <<= 1;
		if (n <= kLdsBucket) {
			if (threadIdx.x < 8u)
				s_beg[threadIdx.x] = s_end[threadIdx.x] = 0u;
			for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
				const uint64_t key = frag_key(b[i]);
				s_keys[i] = key;
				s_slots[i] = (uint8_t)key_slot(key, t.res);
			}
			__syncthreads();
			block_bitonic_kv(n, m, s_keys, s_slots);
			for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
				const uint32_t sl = s_slots[i];
				if (i == 0u || s_slots[i - 1u] != sl)
					s_beg[sl] = i;
				if (i + 1u == n || s_slots[i + 1u] != sl)
					s_end[sl] = i + 1u;
			}
			__syncthreads();
			if (threadIdx.x < 8u) {
				uint32_t word = 0u;
				for (uint32_t i = s_beg[threadIdx.x]; i < s_end[threadIdx.x]; ++i)
					word = fold_word(word, (uint32_t)s_keys[i] & 0xffffffu, s_magic);
				t.oct[8u * t.block + threadIdx.x] = word;
			}
		} else {
			block_bitonic(
			    n, m, [&](uint32_t i) { return frag_key(b[i]); },
			    [&](uint32_t i, uint64_t v) { b[i] = make_uint2((uint32_t)(v >> 32), (uint32_t)v); });
			if (threadIdx.x < 8u) {
				uint32_t word = 0u;
				for (uint32_t i = 0; i < n; ++i) {
					const uint64_t key = frag_key(b[i]);
					if (key_slot(key, t.res) == threadIdx.x)
						word = fold_word(word, (uint32_t)key & 0xffffffu, s_magic);
				}
				t.oct[8u * t.block + threadIdx.x] = word;
			}
		}
		__syncthreads();
	}
}


// The fold's three work lists in one launch: blocks [0, g_mid) take the 17..64-fragment leaf blocks,
// the next g_wide the 65..512-fragment ones, the rest the larger ones.  Each role is block-uniform,
// so the big role's block barriers are safe; the LDS is shared between the roles.
__global__ void __launch_bounds__(kThreads)
    build_fold_lists(const BuildSeg *__restrict__ segs, const uint64_t *__restrict__ bbeg, uint32_t S,
                     const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ aux, uint2 *bucket,
                     const uint32_t *list, uint64_t cap, uint32_t g_mid, uint32_t g_wide) {
	constexpr uint32_t kWaves = kThreads / 64;
	static_assert(kWaves * kWideBucket == kLdsBucket, "the wide and big roles share the key arrays");
	__shared__ uint32_t s_magic[65];
	__shared__ uint64_t s_keys[kLdsBucket];
	__shared__ uint8_t s_slots[kLdsBucket];
	__shared__ uint32_t s_beg[kWaves][8], s_end[kWaves][8];
	stage_magic(s_magic);
	if (blockIdx.x < g_mid)
		fold_mid_blocks(blockIdx.x, g_mid, segs, bbeg, S, cnt, aux, bucket, list, cap,
		                reinterpret_cast<uint32_t (*)[64]>(s_keys), s_magic);
	else if (blockIdx.x < g_mid + g_wide)
		fold_wide_blocks(blockIdx.x - g_mid, g_wide, segs, bbeg, S, cnt, aux, bucket, list, cap, s_keys, s_slots, s_beg,
		                 s_end, s_magic);
	else
		fold_big_blocks(blockIdx.x - g_mid - g_wide, gridDim.x - g_mid - g_wide, segs, bbeg, S, cnt, aux, bucket, list,
		                cap, s_keys, s_slots, s_beg[0], s_end[0], s_magic);
}

// ---- scan plumbing for the bin offsets ---------------------------------------------------------
struct BinCounts {
	static constexpr int kItems = 4;
	const uint32_t *cnt;
	const uint64_t *n; // the batch's leaf blocks, in device memory (bbeg[S])
	__device__ uint64_t size() const { return *n; }
	__device__ uint32_t operator()(uint64_t i) const { return cnt[i]; }
};
struct BinOffsets {
	uint32_t *aux;
	__device__ bool active() const { return true; }
	__device__ void operator()(uint64_t i, uint64_t k, uint32_t) const { aux[i] = (uint32_t)k; }
};

// Blocks for a grid-stride loop over items (kThreads per block and trip), at most per_cu per CU:
// more blocks than the CU holds at once keep more loads in flight (64 per CU measured best for
// most passes; kernels that set up an LDS table per block want fewer, pure streams more).
uint32_t grid_for(uint64_t items, svo_ctx *ctx, uint32_t per_cu = 64u) {
	uint64_t g = (items + kThreads - 1) / kThreads;
	const uint64_t cap = (uint64_t)ctx->cu_count * per_cu;
	if (g > cap)
		g = cap;
	return (uint32_t)(g ? g : 1);
}

// Word offsets of B[1..L-1] inside one scene's bitmaps; off[L] = the scene's bitmap words.  Every
// depth starts an 8-word rank chunk (the padding words stay zero).
void levels_for(uint32_t L, uint64_t *off) {
	uint64_t o = 0;
	for (uint32_t d = 1; d <= L; ++d) {
		off[d] = o;
		if (d < L) {
			const uint64_t bits = 1ull << (3u * d);
			o += ((bits < 32u ? 1u : bits / 32u) + 7u) & ~7ull;
		}
	}
}

// Blocks for the per-fragment passes (kU fragments per thread and trip).
uint32_t grid_tiles(uint64_t F, svo_ctx *ctx, uint32_t per_cu = 64u) { return grid_for(div_up(F, kU), ctx, per_cu); }

uint64_t big_capacity(uint64_t F) { return F / (kRegBucket + 1u) + 1u; }

int structure_phase(svo_ctx *ctx, BuildBatch &bb) {
	const uint32_t S = bb.S;
	hipError_t e;
	if (bb.TW > 0) {
		if ((e = launch(ctx, "bm_zero", bm_zero, dim3(grid_for(bb.TW, ctx)), dim3(kThreads), 0, bb.bm, bb.TW)) !=
		    hipSuccess)
			return hip_check(ctx, e, "bm_zero");
		if (bb.Ftot > 0 &&
		    (e = launch(ctx, "bm_mark", bm_mark, dim3(grid_tiles(bb.Ftot, ctx)), dim3(kThreads), 0,
		                (const BuildSeg *)bb.segs, bb.fbeg(), S, bb.bm)) != hipSuccess)
			return hip_check(ctx, e, "bm_mark");
		for (uint32_t d = bb.maxL >= 3 ? bb.maxL - 2 : 0; d > kTopDepth; --d) {
			const uint64_t n = bb.h_index[(3 + kMaxLevel + d) * (uint64_t)(S + 1) + S];
			if (n > 0 && (e = launch(ctx, "bm_reduce", bm_reduce, dim3(grid_for(n, ctx)), dim3(kThreads), 0,
			                         (const BuildSeg *)bb.segs, bb.index + (3 + kMaxLevel + d) * (uint64_t)(S + 1), S, n,
			                         d, bb.bm, bb.rk)) != hipSuccess)
				return hip_check(ctx, e, "bm_reduce");
		}
		if ((e = launch(ctx, "bm_reduce_top", bm_reduce_top, dim3(S), dim3(kThreads), 0, (const BuildSeg *)bb.segs,
		                bb.bm, bb.rk)) != hipSuccess)
			return hip_check(ctx, e, "bm_reduce_top");
		// exclusive scan of the chunk counts, in place (each item is read and written by one thread)
		ChunkCounts src{bb.rk, bb.TW / 8u};
		if ((e = launch(ctx, "rank_reduce", scan_reduce<ChunkCounts>, dim3(kScanBlocks), dim3(kScanThreads), 0, src,
		                ctx->d_scan_parts)) != hipSuccess)
			return hip_check(ctx, e, "rank_reduce");
		if ((e = launch(ctx, "rank_down", scan_down<ChunkCounts, RankDst, RankTotal>, dim3(kScanBlocks),
		                dim3(kScanThreads), 0, src, RankDst{bb.rk}, (const uint64_t *)ctx->d_scan_parts,
		                RankTotal{bb.rk, bb.TW / 8u})) != hipSuccess)
			return hip_check(ctx, e, "rank_down");
	}
	if ((e = launch(ctx, "bm_levels", bm_levels, dim3(div_up(S, 64)), dim3(64), 0, bb.segs, S,
	                (const uint32_t *)bb.rk, (uint64_t)(1ull << 30))) != hipSuccess)
		return hip_check(ctx, e, "bm_levels");
	return SVO_OK;
}

int node_phase(svo_ctx *ctx, BuildBatch &bb) {
	const uint32_t S = bb.S;
	const BuildSeg *segs = bb.segs;
	hipError_t e;
	{
		// with maxL == 1 there are no parent words and no root blocks to write (the bitmaps are empty)
		const uint32_t M = bb.pw_segs();
		const uint64_t n = bb.maxL > 1u ? bb.h_index[bb.pw_at() + M] : 0u; // parent words of all depths >= 2
		if ((e = launch(ctx, "bm_write_blocks", bm_write_blocks, dim3(grid_for(std::max(n, bb.nbins), ctx)),
		                dim3(kThreads), 0, segs, bb.pw(), M, bb.maxL > 1u ? S : 0u, (const uint32_t *)bb.bm,
		                (const uint32_t *)bb.rk, bb.cnt, bb.bbeg() + S, bb.list)) != hipSuccess)
			return hip_check(ctx, e, "bm_write_blocks");
	}
	const uint32_t gF = grid_tiles(bb.Ftot, ctx);
	uint2 *node = reinterpret_cast<uint2 *>(bb.node);
	if (bb.Ftot > 0 &&
	    (e = launch(ctx, "build_tag_leaf", bb.maxL >= 12u ? build_tag_leaf<256> : build_tag_leaf<128>, dim3(gF),
	                dim3(kThreads), 0, segs, bb.fbeg(), S, (const uint32_t *)bb.bm, (const uint32_t *)bb.rk,
	                node, bb.cnt)) != hipSuccess)
		return hip_check(ctx, e, "build_tag_leaf");
	BinCounts bc{bb.cnt, bb.bbeg() + S};
	if ((e = launch(ctx, "bin_reduce", scan_reduce<BinCounts>, dim3(kScanBlocks), dim3(kScanThreads), 0, bc,
	                ctx->d_scan_parts)) != hipSuccess)
		return hip_check(ctx, e, "bin_reduce");
	if ((e = launch(ctx, "bin_offsets", scan_down<BinCounts, BinOffsets>, dim3(kScanBlocks), dim3(kScanThreads), 0, bc,
	                BinOffsets{bb.aux}, (const uint64_t *)ctx->d_scan_parts, NoCommit{})) != hipSuccess)
		return hip_check(ctx, e, "bin_offsets");
	uint2 *bucket = reinterpret_cast<uint2 *>(bb.bucket);
	if (bb.Ftot > 0 && (e = launch(ctx, "build_scatter", build_scatter, dim3(grid_tiles(bb.Ftot, ctx, 256u)),
	                               dim3(kThreads), 0, segs, bb.fbeg(), S, (const uint2 *)node,
	                               (const uint32_t *)bb.aux, bucket)) != hipSuccess)
		return hip_check(ctx, e, "build_scatter");
	const uint64_t cap = big_capacity(bb.Ftot);
	if ((e = launch(ctx, "build_fold", build_fold, dim3(grid_for(bb.nbins, ctx)), dim3(kThreads), 0, segs, bb.bbeg(),
	                S, (const uint32_t *)bb.cnt, (const uint32_t *)bb.aux, (const uint2 *)bucket, bb.list, cap)) !=
	    hipSuccess)
		return hip_check(ctx, e, "build_fold");
	const uint32_t g_mid = (uint32_t)ctx->cu_count * 8u, g_wide = (uint32_t)ctx->cu_count * 8u,
	               g_big = (uint32_t)ctx->cu_count * 2u;
	if ((e = launch(ctx, "build_fold_lists", build_fold_lists, dim3(g_mid + g_wide + g_big), dim3(kThreads), 0, segs,
	                bb.bbeg(), S, (const uint32_t *)bb.cnt, (const uint32_t *)bb.aux, bucket, (const uint32_t *)bb.list,
	                cap, g_mid, g_wide)) != hipSuccess)
		return hip_check(ctx, e, "build_fold_lists");
	return SVO_OK;
}

} // namespace

uint64_t build_bitmap_words(uint32_t L) {
	uint64_t off[16] = {};
	levels_for(L, off);
	return off[L];
}

void build_batch_release(svo_ctx *ctx, BuildBatch *bb) {
	for (void *p : {(void *)bb->segs, (void *)bb->bm, (void *)bb->rk, (void *)bb->cnt, (void *)bb->aux,
	                (void *)bb->node, (void *)bb->bucket, (void *)bb->list}) // index lives in segs' allocation
		ctx->pool->free(p);
	bb->segs = nullptr;
	bb->index = nullptr;
	bb->bm = bb->rk = bb->cnt = bb->aux = bb->node = bb->bucket = bb->list = nullptr;
}

namespace {

// Host layout of a build batch of S scenes with F[s] fragments at level L[s]: descriptors (bitmap
// offsets), fragment begins, per-depth bitmap word tables and the block writer's parent words.
void batch_layout(BuildBatch &bb, uint32_t S, const uint2 *const *frags, const uint64_t *F, const uint32_t *L) {
	bb.S = S;
	bb.h_segs.assign(S, BuildSeg{});
	const uint64_t stride = S + 1;
	bb.h_index.assign(bb.pw_at() + bb.pw_segs() + 1, 0);
	uint64_t *fbeg = bb.h_index.data();
	for (uint32_t s = 0; s < S; ++s) {
		BuildSeg &g = bb.h_segs[s];
		g.frags = frags[s];
		g.F = F[s];
		g.L = L[s];
		g.bm = bb.TW;
		levels_for(g.L, g.off);
		bb.TW += g.off[g.L];
		fbeg[s] = bb.Ftot;
		bb.Ftot += g.F;
		bb.maxL = std::max(bb.maxL, g.L);
	}
	fbeg[S] = bb.Ftot;
	// per-depth begin tables: dw[d] over scenes with B[d] (L > d), dr[d] over scenes with B[d+1] (L > d+1)
	for (uint32_t d = 1; d <= kMaxLevel; ++d) {
		uint64_t *dw = bb.h_index.data() + (2 + d) * stride, *dr = bb.h_index.data() + (3 + kMaxLevel + d) * stride;
		uint64_t aw = 0, ar = 0;
		for (uint32_t s = 0; s < S; ++s) {
			const BuildSeg &g = bb.h_segs[s];
			dw[s] = aw;
			dr[s] = ar;
			const uint64_t words = g.L > d ? g.off[d + 1] - g.off[d] : 0;
			aw += words;
			ar += g.L > d + 1 ? words : 0;
		}
		dw[S] = aw;
		dr[S] = ar;
	}
	// parent words of the block writer: B[d-1] of the scenes with L > d, for d = 2..kMaxLevel-1
	uint64_t *pw = bb.h_index.data() + bb.pw_at(), a = 0;
	for (uint32_t d = 2; d < kMaxLevel; ++d)
		for (uint32_t s = 0; s < S; ++s) {
			const BuildSeg &g = bb.h_segs[s];
			*pw++ = a;
			a += g.L > d ? g.off[d] - g.off[d - 1] : 0;
		}
	*pw = a;
}

// Descriptors and begin tables in one device allocation, so each upload is one copy.
size_t seg_table_bytes(uint32_t S) { return (sizeof(BuildSeg) * S + 255) & ~size_t(255); }

int upload_batch_tables(svo_ctx *ctx, BuildBatch &bb) {
	const size_t seg_bytes = seg_table_bytes(bb.S), idx_bytes = sizeof(uint64_t) * bb.h_index.size();
	std::vector<uint8_t> tables(seg_bytes + idx_bytes);
	std::memcpy(tables.data(), bb.h_segs.data(), sizeof(BuildSeg) * bb.S);
	std::memcpy(tables.data() + seg_bytes, bb.h_index.data(), idx_bytes);
	return upload_table(ctx, bb.segs, tables.data(), tables.size());
}

} // namespace

// The whole build of S octrees: structure phase, one read-back of every scene's W and leaf-block
// count, allocation of the octrees, node phase.  Scratch goes back to the pool at the end; the pool
// hands it only to later work on this stream, so no wait is needed for the node phase to finish.
int build_batch(svo_ctx *ctx, svo_octree *const *octs, uint32_t S) {
	if (S == 0)
		return SVO_OK;
	if (S > kMaxBatch)
		return set_error(ctx, SVO_ERR_INVALID_ARGUMENT, "batch larger than 1024 scenes");
	BuildBatch bb;
	const uint64_t stride = S + 1;
	{
		std::vector<const uint2 *> frags(S);
		std::vector<uint64_t> F(S);
		std::vector<uint32_t> L(S);
		for (uint32_t s = 0; s < S; ++s) {
			const svo_fraglist *fl = octs[s]->src;
			frags[s] = reinterpret_cast<const uint2 *>(fl->frags);
			F[s] = fl->F;
			L[s] = fl->level;
		}
		batch_layout(bb, S, frags.data(), F.data(), L.data());
	}
	if (bb.Ftot > 0xFFFFFFFEull)
		return set_error(ctx, SVO_ERR_OVERFLOW, "a build batch holds at most 2^32-2 fragments; split the batch");
	const uint64_t Fa = bb.Ftot ? bb.Ftot : 1;
	const uint64_t cap = big_capacity(bb.Ftot);
	Pool *pool = ctx->pool;
	int rc = SVO_OK;
	const size_t seg_bytes = seg_table_bytes(S), idx_bytes = sizeof(uint64_t) * bb.h_index.size();
	if (pool->alloc((void **)&bb.segs, seg_bytes + idx_bytes) != hipSuccess ||
	    pool->alloc((void **)&bb.bm, sizeof(uint32_t) * (bb.TW ? bb.TW : 1)) != hipSuccess ||
	    pool->alloc((void **)&bb.rk, sizeof(uint32_t) * (bb.TW / 8u + 1)) != hipSuccess ||
	    pool->alloc((void **)&bb.node, sizeof(uint32_t) * 2 * Fa) != hipSuccess ||
	    pool->alloc((void **)&bb.bucket, sizeof(uint32_t) * 2 * Fa) != hipSuccess ||
	    pool->alloc((void **)&bb.list, sizeof(uint32_t) * (3 + 3 * cap)) != hipSuccess)
		rc = set_error(ctx, SVO_ERR_OUT_OF_MEMORY, "octree build allocation failed");
	if (!rc) {
		bb.index = reinterpret_cast<uint64_t *>(reinterpret_cast<uint8_t *>(bb.segs) + seg_bytes);
		rc = upload_batch_tables(ctx, bb);
	}
	// structure phase: bitmaps + ranks decide every W exactly; one read-back sizes the octrees
	if (!rc) {
		ctx->t_build.start(ctx->stream);
		rc = structure_phase(ctx, bb);
		ctx->t_build.stop(ctx->stream);
	}
	if (!rc)
		rc = download_table(ctx, bb.h_segs.data(), bb.segs, sizeof(BuildSeg) * S);
	if (!rc)
		for (uint32_t s = 0; s < S; ++s)
			if (bb.h_segs[s].overflow) {
				rc = set_error(ctx, SVO_ERR_OVERFLOW,
				               "octree " + std::to_string(s) + " needs more than 2^30 words (30-bit child pointers)");
				break;
			}
	uint64_t *bbeg = bb.h_index.data() + stride;
	if (!rc) {
		for (uint32_t s = 0; s < S; ++s) {
			BuildSeg &g = bb.h_segs[s];
			svo_octree *o = octs[s];
			o->W = o->cap = g.W; // (counter + 1) * 8, src/OctreeBuilder.cpp:212-214
			if (pool->alloc((void **)&o->words, sizeof(uint32_t) * g.W) != hipSuccess) {
				rc = set_error(ctx, SVO_ERR_OUT_OF_MEMORY, "octree allocation failed");
				break;
			}
			g.words = o->words;
			g.bin_begin = bb.nbins;
			bbeg[s] = bb.nbins;
			bb.nbins += g.nbins;
		}
		bbeg[S] = bb.nbins;
	}
	if (!rc && (pool->alloc((void **)&bb.cnt, sizeof(uint32_t) * (bb.nbins ? bb.nbins : 1)) != hipSuccess ||
	            pool->alloc((void **)&bb.aux, sizeof(uint32_t) * (bb.nbins ? bb.nbins : 1)) != hipSuccess))
		rc = set_error(ctx, SVO_ERR_OUT_OF_MEMORY, "octree build allocation failed");
	if (!rc)
		rc = upload_batch_tables(ctx, bb);
	if (!rc) {
		ctx->t_build2.start(ctx->stream);
		rc = node_phase(ctx, bb);
		ctx->t_build2.stop(ctx->stream);
	}
	build_batch_release(ctx, &bb);
	return rc;
}

// ---- captured plans: voxelize + build of a fixed set of scenes with the counts kept on the device
namespace {

constexpr uint32_t kPlanThreads = kMaxBatch; // the check kernels take one scene per thread

// Between the voxelize count and fill passes: the scenes' fragment counts become the build's
// fragment begins.  A count above its list's capacity sets status bit 1 and empties the batch:
// fill limit, every fragment begin and the fragment total 0, so no later pass reads or writes a
// list past its end.  Also restores the block writer's parent-word total, which the node check
// of the previous run may have cleared.  out = {status, F[S], W[S]}.
__global__ void __launch_bounds__(kPlanThreads)
    plan_check_frags(const VoxSeg *__restrict__ vsegs, BuildSeg *bsegs, uint64_t *fbeg, uint32_t S,
                     const uint64_t *__restrict__ fcap, uint64_t *fend, uint64_t *pw_total, uint64_t pw_value,
                     uint64_t *out) {
	const uint32_t s = threadIdx.x;
	uint64_t F = 0, b = 0;
	int over = 0;
	if (s < S) {
		F = vsegs[s].F;
		b = vsegs[s].frag_begin;
		over = F > fcap[s];
		out[1u + s] = F;
		out[1u + S + s] = 0u;
		bsegs[s].F = F;
	}
	over = __syncthreads_or(over);
	if (s < S)
		fbeg[s] = over ? 0u : b;
	if (s == 0u) {
		const uint64_t total = over ? 0u : fend[0];
		fend[1] = total;
		fbeg[S] = total;
		*pw_total = pw_value;
		out[0] = over ? 1u : 0u;
	}
}

// Between the structure and node phases: leaf-block begins (an exclusive scan of the scenes' leaf
// blocks) and W against each octree's capacity.  A W above its capacity (or above 2^30 words), or
// a fragment overflow found before, sets status bit 2 and empties the node phase: leaf blocks,
// fragments and parent words 0, so it writes nothing but the root blocks.
__global__ void __launch_bounds__(kPlanThreads)
    plan_check_nodes(BuildSeg *segs, uint64_t *bbeg, uint32_t S, const uint64_t *__restrict__ wcap, uint64_t *fbeg,
                     uint64_t *pw_total, uint64_t *out) {
	__shared__ uint64_t s_wave[kPlanThreads / 64];
	const uint32_t s = threadIdx.x, lane = s & 63u, wave = s >> 6;
	uint64_t nb = 0;
	int over = 0;
	if (s < S) {
		const BuildSeg &g = segs[s];
		nb = g.nbins;
		over = g.overflow || g.W > wcap[s];
		out[1u + S + s] = g.W;
	}
	over = __syncthreads_or(over) || out[0] != 0u;
	uint64_t incl = nb;
#pragma unroll
	for (uint32_t d = 1; d < 64u; d <<= 1) {
		const uint64_t t = __shfl_up(incl, d, 64);
		incl += lane >= d ? t : 0u;
	}
	if (lane == 63u)
		s_wave[wave] = incl;
	__syncthreads();
	uint64_t below = 0, total = 0;
	for (uint32_t w = 0; w < kPlanThreads / 64; ++w) {
		below += w < wave ? s_wave[w] : 0u;
		total += s_wave[w];
	}
	const uint64_t excl = below + incl - nb;
	if (s < S) {
		segs[s].bin_begin = over ? 0u : excl;
		bbeg[s] = over ? 0u : excl;
	}
	if (s == 0u) {
		bbeg[S] = over ? 0u : total;
		if (over) {
			fbeg[S] = 0u;
			*pw_total = 0u;
			out[0] |= 2u;
		}
	}
}

// The plan's passes in stream order; nothing waits on the host, so the sequence can be captured.
int plan_enqueue(svo_ctx *ctx, BuildPlan &p) {
	BuildBatch &bb = p.bb;
	const uint32_t S = p.S;
	uint64_t *pw_total = bb.index + bb.pw_at() + bb.pw_segs();
	hipError_t e;
	if ((e = voxelize_plan_count(ctx, &p.vb)) != hipSuccess)
		return hip_check(ctx, e, "plan voxelize count");
	if ((e = launch(ctx, "plan_check_frags", plan_check_frags, dim3(1), dim3(kPlanThreads), 0,
	                (const VoxSeg *)p.vb.segs, bb.segs, bb.index, S, (const uint64_t *)p.caps, p.vb.fend, pw_total,
	                p.pw_value, p.out)) != hipSuccess)
		return hip_check(ctx, e, "plan_check_frags");
	if ((e = voxelize_plan_fill(ctx, &p.vb)) != hipSuccess)
		return hip_check(ctx, e, "plan voxelize fill");
	if (int rc = structure_phase(ctx, bb))
		return rc;
	if ((e = launch(ctx, "plan_check_nodes", plan_check_nodes, dim3(1), dim3(kPlanThreads), 0, bb.segs,
	                bb.index + (S + 1), S, (const uint64_t *)p.caps + S, bb.index, pw_total, p.out)) != hipSuccess)
		return hip_check(ctx, e, "plan_check_nodes");
	return node_phase(ctx, bb);
}

} // namespace

int build_plan_init(svo_ctx *ctx, BuildPlan *p, const svo_scene *const *scenes, const uint32_t *levels, uint32_t S,
                    const uint64_t *fcap, const uint64_t *wcap) {
	p->S = S;
	Pool *pool = ctx->pool;
	uint64_t Fcap = 0, Bcap = 0;
	std::vector<uint64_t> caps(2 * (size_t)S);
	for (uint32_t s = 0; s < S; ++s) {
		const uint32_t L = levels[s];
		caps[s] = fcap[s] ? fcap[s] : 1u;
		caps[S + s] = std::max<uint64_t>((wcap[s] + 7u) & ~7ull, 8u);
		if (caps[S + s] > (1ull << 30))
			return set_error(ctx, SVO_ERR_OVERFLOW, "an octree capacity above 2^30 words (30-bit child pointers)");
		Fcap += caps[s];
		// leaf blocks: the root at level 1, else at most one per fragment and 8^(L-1)
		Bcap += L == 1u ? 1u : std::min<uint64_t>(caps[s], 1ull << (3u * (L - 1u)));
	}
	if (Fcap > 0xFFFFFFFEull)
		return set_error(ctx, SVO_ERR_OVERFLOW, "a plan holds at most 2^32-2 fragments of capacity; split it");
	p->fls.assign(S, nullptr);
	p->octs.assign(S, nullptr);
	for (uint32_t s = 0; s < S; ++s) {
		svo_fraglist *fl = p->fls[s] = new (std::nothrow) svo_fraglist();
		svo_octree *o = p->octs[s] = new (std::nothrow) svo_octree();
		if (!fl || !o)
			return SVO_ERR_OUT_OF_MEMORY;
		fl->ctx = o->ctx = ctx;
		fl->scene = scenes[s];
		fl->level = o->level = levels[s];
		o->cap = caps[S + s];
		if (pool->alloc((void **)&fl->frags, sizeof(uint32_t) * 2 * caps[s]) != hipSuccess ||
		    pool->alloc((void **)&o->words, sizeof(uint32_t) * o->cap) != hipSuccess)
			return set_error(ctx, SVO_ERR_OUT_OF_MEMORY, "plan list or octree allocation failed");
	}
	if (int rc = voxelize_plan_init(ctx, &p->vb, p->fls.data(), S))
		return rc;
	BuildBatch &bb = p->bb;
	{
		std::vector<const uint2 *> frags(S);
		for (uint32_t s = 0; s < S; ++s)
			frags[s] = reinterpret_cast<const uint2 *>(p->fls[s]->frags);
		batch_layout(bb, S, frags.data(), caps.data(), levels);
	}
	// host totals are capacities: they size the grids and scratch; the kernels read the counts
	bb.nbins = Bcap;
	for (uint32_t s = 0; s < S; ++s)
		bb.h_segs[s].words = p->octs[s]->words;
	p->pw_value = bb.h_index[bb.pw_at() + bb.pw_segs()];
	const size_t seg_bytes = seg_table_bytes(S), idx_bytes = sizeof(uint64_t) * bb.h_index.size();
	if (pool->alloc((void **)&bb.segs, seg_bytes + idx_bytes) != hipSuccess ||
	    pool->alloc((void **)&bb.bm, sizeof(uint32_t) * (bb.TW ? bb.TW : 1)) != hipSuccess ||
	    pool->alloc((void **)&bb.rk, sizeof(uint32_t) * (bb.TW / 8u + 1)) != hipSuccess ||
	    pool->alloc((void **)&bb.node, sizeof(uint32_t) * 2 * Fcap) != hipSuccess ||
	    pool->alloc((void **)&bb.bucket, sizeof(uint32_t) * 2 * Fcap) != hipSuccess ||
	    pool->alloc((void **)&bb.list, sizeof(uint32_t) * (3 + 3 * big_capacity(Fcap))) != hipSuccess ||
	    pool->alloc((void **)&bb.cnt, sizeof(uint32_t) * Bcap) != hipSuccess ||
	    pool->alloc((void **)&bb.aux, sizeof(uint32_t) * Bcap) != hipSuccess ||
	    pool->alloc((void **)&p->caps, sizeof(uint64_t) * 2 * S) != hipSuccess ||
	    pool->alloc((void **)&p->out, sizeof(uint64_t) * (1 + 2 * (size_t)S)) != hipSuccess)
		return set_error(ctx, SVO_ERR_OUT_OF_MEMORY, "plan build allocation failed");
	bb.index = reinterpret_cast<uint64_t *>(reinterpret_cast<uint8_t *>(bb.segs) + seg_bytes);
	if (int rc = upload_batch_tables(ctx, bb))
		return rc;
	if (int rc = upload_table(ctx, p->caps, caps.data(), sizeof(uint64_t) * 2 * S))
		return rc;
	// results before the first run: status 0, no fragments, no words
	p->h_out.assign(1 + 2 * (size_t)S, 0);
	return upload_table(ctx, p->out, p->h_out.data(), sizeof(uint64_t) * p->h_out.size());
}

int build_plan_run(svo_ctx *ctx, BuildPlan *p) {
	hipError_t e = hipSuccess;
	int rc = SVO_OK;
	p->t_run.start(ctx->stream);
	if (p->use_graph) {
		if (!p->exec) {
			// profiler events stay out of the graph
			const bool prof = ctx->prof.enabled;
			ctx->prof.enabled = false;
			hipGraph_t g = nullptr;
			e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
			if (e == hipSuccess) {
				rc = plan_enqueue(ctx, *p);
				e = hipStreamEndCapture(ctx->stream, &g);
			}
			ctx->prof.enabled = prof;
			if (!rc && e == hipSuccess)
				e = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
			if (g)
				(void)hipGraphDestroy(g);
			if (rc || e != hipSuccess) {
				p->exec = nullptr;
				return rc ? rc : hip_check(ctx, e, "plan capture");
			}
		}
		e = hipGraphLaunch(p->exec, ctx->stream);
	} else {
		rc = plan_enqueue(ctx, *p);
	}
	p->t_run.stop(ctx->stream);
	if (rc)
		return rc;
	if (e != hipSuccess)
		return hip_check(ctx, e, "plan launch");
	p->pending = true;
	return SVO_OK;
}

int build_plan_finish(svo_ctx *ctx, BuildPlan *p, uint64_t *F, uint64_t *W, double *ms) {
	if (int rc = download_table(ctx, p->h_out.data(), p->out, sizeof(uint64_t) * p->h_out.size()))
		return rc;
	if (ms)
		*ms = p->pending ? p->t_run.read() : 0.0;
	p->pending = false;
	const uint32_t S = p->S;
	const uint64_t status = p->h_out[0];
	for (uint32_t s = 0; s < S; ++s) {
		const uint64_t f = p->h_out[1 + s], w = p->h_out[1 + S + s];
		if (F)
			F[s] = f;
		if (W)
			W[s] = w;
		p->fls[s]->F = status & 1u ? 0u : f;
		p->fls[s]->filled = !(status & 1u);
		p->octs[s]->W = status ? 0u : w;
	}
	if (status & 1u)
		for (uint32_t s = 0; s < S; ++s)
			if (p->h_out[1 + s] > p->bb.h_segs[s].F)
				return set_error(ctx, SVO_ERR_OVERFLOW,
				                 "scene " + std::to_string(s) + " has " + std::to_string(p->h_out[1 + s]) +
				                     " fragments, above the plan's capacity of " +
				                     std::to_string(p->bb.h_segs[s].F));
	if (status)
		for (uint32_t s = 0; s < S; ++s)
			if (p->h_out[1 + S + s] > p->octs[s]->cap)
				return set_error(ctx, SVO_ERR_OVERFLOW,
				                 "octree " + std::to_string(s) + " needs " + std::to_string(p->h_out[1 + S + s]) +
				                     " words, above the plan's capacity of " + std::to_string(p->octs[s]->cap));
	if (status)
		return set_error(ctx, SVO_ERR_OVERFLOW, "a plan count passed its capacity");
	return SVO_OK;
}

void build_plan_release(svo_ctx *ctx, BuildPlan *p) {
	if (p->exec)
		(void)hipGraphExecDestroy(p->exec);
	p->exec = nullptr;
	if (p->pending)
		(void)hipStreamSynchronize(ctx->stream);
	p->pending = false;
	p->t_run.destroy();
	vox_batch_release(ctx, &p->vb);
	build_batch_release(ctx, &p->bb);
	ctx->pool->free(p->caps);
	ctx->pool->free(p->out);
	p->caps = p->out = nullptr;
	for (svo_fraglist *fl : p->fls)
		if (fl) {
			ctx->pool->free(fl->frags);
			delete fl;
		}
	for (svo_octree *o : p->octs)
		if (o) {
			ctx->pool->free(o->words);
			delete o;
		}
	p->fls.clear();
	p->octs.clear();
}

} // namespace svo
