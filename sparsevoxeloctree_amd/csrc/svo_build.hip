// Level-by-level octree construction (OctreeBuilder::CmdBuild, src/OctreeBuilder.cpp:142-210).
//
// Reference pass            -> kernels here
//   octree_init_node.comp   -> fused into the allocation write (each new block is zeroed by the
//                              thread that allocates it); the root block by build_init.
//   octree_tag_node.comp    -> build_tag (levels 1..L-1) / build_tag_leaf (level L).  Each fragment
//                              keeps the word index of its node at the previous level (node[f]),
//                              so a level pass does one octree read instead of re-descending from
//                              the root (shader/octree_tag_node.comp:18-30 re-descends i levels).
//   octree_alloc_node.comp  -> device-wide exclusive scan of the tag bits of the level's range:
//                              the k-th tagged word (in word order) gets block counter+k+1, which
//                              is the canonical breadth-first layout; the reference's
//                              subgroup-ballot + atomicAdd order is nondeterministic.
//   octree_modify_arg.comp  -> scan_parts' commit: next range = [begin+num, +8*total), written
//                              to the other BuildState slot.
//   leaf CAS colour fold    -> count fragments per leaf, scan, scatter into per-leaf buckets, sort
//   (octree_tag_node.comp     each bucket by the 64-bit fragment and fold in that order (the
//    :43-57)                  reference's order is whatever the atomics produce).
#include "svo_device.hpp"
#include "svo_runtime.hpp"
#include "svo_scan.hpp"

namespace svo {
namespace {

constexpr int kThreads = 256;
constexpr uint32_t kSmallBucket = 32;   // leaves with more fragments go to the block-level sorter
constexpr uint32_t kLdsBucket = 4096;   // largest bucket sorted in LDS

// Effective voxel coordinate seen by the reference's descent: levels whose bit is set in the
// coordinate, saturating at res-1 for coordinates >= res (shader/octree_tag_node.comp:18-30).
__device__ __forceinline__ void decode_frag(uint2 fr, uint32_t res, uint32_t &x, uint32_t &y, uint32_t &z) {
	x = fr.x & 0xfffu;
	y = (fr.x >> 12u) & 0xfffu;
	z = (fr.x >> 24u) | ((fr.y >> 28u) << 8u);
	x = umin32(x, res - 1u);
	y = umin32(y, res - 1u);
	z = umin32(z, res - 1u);
}

__device__ __forceinline__ uint32_t child_slot(uint32_t x, uint32_t y, uint32_t z, uint32_t shift) {
	return ((x >> shift) & 1u) | (((y >> shift) & 1u) << 1u) | (((z >> shift) & 1u) << 2u);
}

// Running average of the reference's CAS loop (octree_tag_node.comp:43-57): fold colour c into a
// leaf word that already holds n fragments (the word's 6-bit count, saturated at 63).  In a
// sequence the a-th fragment (a >= 1) always meets n = min(a, 63), so unrolled callers pass a
// constant n and the divisions become multiplications.
__device__ __forceinline__ uint32_t fold_step(uint32_t prev, uint32_t c, uint32_t n) {
	const uint32_t w = n + 1u;
	uint32_t r = ((prev & 0xffu) * n + (c & 0xffu)) / w;
	uint32_t g = (((prev >> 8u) & 0xffu) * n + ((c >> 8u) & 0xffu)) / w;
	uint32_t b = (((prev >> 16u) & 0xffu) * n + ((c >> 16u) & 0xffu)) / w;
	return (umin32(w, 0x3fu) << 24u) | (r & 0xffu) | ((g & 0xffu) << 8u) | ((b & 0xffu) << 16u) | 0xC0000000u;
}

__device__ __forceinline__ uint32_t fold_push(uint32_t prev, uint32_t c, uint32_t a) {
	return a == 0u ? (0xC1000000u | c) : fold_step(prev, c, a < 63u ? a : 63u);
}

__device__ __forceinline__ void cmp_swap(uint64_t &x, uint64_t &y) {
	const uint64_t lo = y < x ? y : x, hi = y < x ? x : y;
	x = lo;
	y = hi;
}

__device__ __forceinline__ uint64_t frag_key(uint2 f) { return ((uint64_t)f.x << 32) | f.y; }

__global__ void build_init(BuildState *st, uint32_t *oct) {
	const uint32_t t = threadIdx.x;
	if (t < 8)
		oct[t] = 0u; // octree_init_node.comp for the root block
	if (t == 0) {
		// src/OctreeBuilder.cpp:27-32: {uAllocBegin = 0, uAllocNum = 8}; level 1 uses slot 1
		st->begin[1] = 0u;
		st->num[1] = 8u;
		st->counter[1] = 0u;
		st->begin[0] = st->num[0] = st->counter[0] = 0u;
		st->overflow = 0u;
		st->need_words = 0u;
	}
}

__global__ void build_init_big(uint32_t *big) { big[0] = 0u; }

// After the node buffer was grown: clear the overflow report so the level can be allocated again.
__global__ void build_clear_overflow(BuildState *st) {
	st->overflow = 0u;
	st->need_words = 0u;
}

// Levels 1..L-1: find the level-i node of every fragment and tag it non-empty.
__global__ void __launch_bounds__(kThreads)
    build_tag(const uint2 *__restrict__ frags, uint64_t F, uint32_t *__restrict__ node, uint32_t *oct,
              const BuildState *st, uint32_t L, uint32_t i) {
	if (st->overflow)
		return;
	const uint32_t res = 1u << L, shift = L - i;
	const uint32_t lane = threadIdx.x & 63u;
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	for (uint64_t base = (uint64_t)blockIdx.x * kThreads; base < F; base += stride) {
		uint64_t f = base + threadIdx.x;
		uint32_t idx = 0xFFFFFFFFu;
		if (f < F) {
			uint32_t x, y, z;
			decode_frag(frags[f], res, x, y, z);
			uint32_t blk = (i == 1u) ? 0u : (oct[node[f]] & 0x3fffffffu);
			idx = blk + child_slot(x, y, z, shift);
			node[f] = idx;
		}
		// neighbouring fragments usually share the node: only the first lane of a run stores
		uint32_t prev = __shfl_up(idx, 1u, 64);
		if (idx != 0xFFFFFFFFu && (lane == 0u || prev != idx))
			oct[idx] = 0x80000000u;
	}
}

// Level L: leaf word of every fragment; count fragments per leaf in the leaf word itself.
__global__ void __launch_bounds__(kThreads)
    build_tag_leaf(const uint2 *__restrict__ frags, uint64_t F, uint32_t *__restrict__ node, uint32_t *oct,
                   const BuildState *st, uint32_t L) {
	if (st->overflow)
		return;
	const uint32_t res = 1u << L, slot = L & 1u;
	const uint32_t begin = st->begin[slot];
	const uint32_t lane = threadIdx.x & 63u;
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	for (uint64_t base = (uint64_t)blockIdx.x * kThreads; base < F; base += stride) {
		uint64_t f = base + threadIdx.x;
		uint32_t idx = 0xFFFFFFFFu;
		if (f < F) {
			uint32_t x, y, z;
			decode_frag(frags[f], res, x, y, z);
			uint32_t blk = (L == 1u) ? 0u : (oct[node[f]] & 0x3fffffffu);
			idx = blk + child_slot(x, y, z, 0u);
			node[f] = idx - begin;
		}
		// run-length aggregate equal neighbours: one atomic per run
		uint32_t prev = __shfl_up(idx, 1u, 64);
		bool head = (lane == 0u) || prev != idx;
		uint64_t heads = __ballot(head);
		uint64_t after = (lane == 63u) ? 0ull : (heads >> (lane + 1u));
		uint32_t run = after ? (uint32_t)__ffsll((unsigned long long)after) : 64u - lane;
		if (head && idx != 0xFFFFFFFFu)
			atomicAdd(&oct[idx], run);
	}
}

// Scatter every fragment into its leaf's bucket (aux[j] holds the running end of bucket j).
__global__ void __launch_bounds__(kThreads)
    build_scatter(const uint2 *__restrict__ frags, uint64_t F, const uint32_t *__restrict__ node, uint32_t *aux,
                  uint2 *__restrict__ bucket, const BuildState *st) {
	if (st->overflow)
		return;
	const uint32_t lane = threadIdx.x & 63u;
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	for (uint64_t base = (uint64_t)blockIdx.x * kThreads; base < F; base += stride) {
		uint64_t f = base + threadIdx.x;
		uint32_t j = f < F ? node[f] : 0xFFFFFFFFu;
		uint32_t prev = __shfl_up(j, 1u, 64);
		bool head = (lane == 0u) || prev != j;
		uint64_t heads = __ballot(head);
		uint64_t after = (lane == 63u) ? 0ull : (heads >> (lane + 1u));
		uint32_t run = after ? (uint32_t)__ffsll((unsigned long long)after) : 64u - lane;
		uint32_t pos = 0;
		if (head && j != 0xFFFFFFFFu)
			pos = atomicAdd(&aux[j], run);
		// lane of this lane's run head = highest head at or below this lane
		uint64_t upto = heads & ((lane == 63u) ? ~0ull : ((2ull << lane) - 1ull));
		uint32_t head_lane = 63u - (uint32_t)__clzll((long long)upto);
		pos = __shfl(pos, (int)head_lane, 64) + (lane - head_lane);
		if (j != 0xFFFFFFFFu)
			bucket[pos] = frags[f];
	}
}

// Small buckets: insertion sort in place + fold, one thread per leaf.  Larger buckets are queued
// for build_fold_big.
__global__ void __launch_bounds__(kThreads)
    build_fold(uint32_t *oct, const uint32_t *__restrict__ aux, uint2 *bucket, uint32_t *big, uint64_t big_cap,
               const BuildState *st, uint32_t L) {
	if (st->overflow)
		return;
	const uint32_t slot = L & 1u;
	const uint32_t begin = st->begin[slot], num = st->num[slot];
	const uint32_t stride = gridDim.x * kThreads;
	for (uint32_t j = blockIdx.x * kThreads + threadIdx.x; j < num; j += stride) {
		uint32_t n = oct[begin + j];
		if (n == 0u)
			continue;
		uint32_t end = aux[j], start = end - n;
		if (n > kSmallBucket) {
			uint32_t k = atomicAdd(&big[0], 1u);
			if (k < big_cap)
				big[1 + k] = j;
			continue;
		}
		uint2 *b = bucket + start;
		uint32_t word;
		if (n <= 8u) {
			// sort in registers: 19-comparator network for 8 keys, +inf padding
			uint64_t k[8];
#pragma unroll
			for (uint32_t a = 0; a < 8u; ++a)
				k[a] = a < n ? frag_key(b[a]) : ~0ull;
			cmp_swap(k[0], k[1]), cmp_swap(k[2], k[3]), cmp_swap(k[4], k[5]), cmp_swap(k[6], k[7]);
			cmp_swap(k[0], k[2]), cmp_swap(k[1], k[3]), cmp_swap(k[4], k[6]), cmp_swap(k[5], k[7]);
			cmp_swap(k[1], k[2]), cmp_swap(k[5], k[6]), cmp_swap(k[0], k[4]), cmp_swap(k[3], k[7]);
			cmp_swap(k[1], k[5]), cmp_swap(k[2], k[6]);
			cmp_swap(k[1], k[4]), cmp_swap(k[3], k[6]);
			cmp_swap(k[2], k[4]), cmp_swap(k[3], k[5]);
			cmp_swap(k[3], k[4]);
			word = 0xC1000000u | ((uint32_t)k[0] & 0xffffffu);
#pragma unroll
			for (uint32_t a = 1; a < 8u; ++a)
				if (a < n)
					word = fold_step(word, (uint32_t)k[a] & 0xffffffu, a);
		} else {
			for (uint32_t a = 1; a < n; ++a) {
				uint2 v = b[a];
				uint64_t kv = frag_key(v);
				uint32_t c = a;
				while (c > 0u && frag_key(b[c - 1]) > kv) {
					b[c] = b[c - 1];
					--c;
				}
				b[c] = v;
			}
			word = 0u;
			for (uint32_t a = 0; a < n; ++a)
				word = fold_push(word, b[a].y & 0xffffffu, a);
		}
		oct[begin + j] = word;
	}
}

// Bitonic sort of m = 2^p keys (ascending) by one block; every comparator puts the minimum at the
// lower index, so keys past the end behave as +infinity padding.
template <class Get, class Put>
__device__ void block_bitonic(uint32_t n, uint32_t m, Get get, Put put) {
	for (uint32_t k = 2; k <= m; k <<= 1) {
		for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
			for (uint32_t t = threadIdx.x; t < m / 2; t += blockDim.x) {
				// comparator t of this stage: lower index a, partner b
				uint32_t a, b;
				if (jj == (k >> 1)) { // first step of a merge: flip
					uint32_t blk = t / jj, off = t % jj;
					a = blk * k + off;
					b = blk * k + k - 1u - off;
				} else {
					uint32_t blk = t / jj, off = t % jj;
					a = blk * 2u * jj + off;
					b = a + jj;
				}
				if (b < n) {
					uint64_t va = get(a), vb = get(b);
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

__global__ void __launch_bounds__(kThreads)
    build_fold_big(uint32_t *oct, const uint32_t *__restrict__ aux, uint2 *bucket, const uint32_t *big,
                   uint64_t big_cap, const BuildState *st, uint32_t L) {
	if (st->overflow)
		return;
	__shared__ uint64_t s_keys[kLdsBucket];
	const uint32_t slot = L & 1u;
	const uint32_t begin = st->begin[slot];
	uint64_t count = big[0];
	if (count > big_cap)
		count = big_cap;
	for (uint64_t item = blockIdx.x; item < count; item += gridDim.x) {
		uint32_t j = big[1 + item];
		uint32_t n = oct[begin + j];
		uint32_t end = aux[j], start = end - n;
		uint2 *b = bucket + start;
		uint32_t m = 1;
		while (m < n)
			m <<= 1;
		if (n <= kLdsBucket) {
			for (uint32_t t = threadIdx.x; t < n; t += blockDim.x)
				s_keys[t] = frag_key(b[t]);
			__syncthreads();
			block_bitonic(
			    n, m, [&](uint32_t i) { return s_keys[i]; }, [&](uint32_t i, uint64_t v) { s_keys[i] = v; });
			if (threadIdx.x == 0) {
				uint32_t word = 0u;
				for (uint32_t a = 0; a < n; ++a)
					word = fold_push(word, (uint32_t)s_keys[a] & 0xffffffu, a);
				oct[begin + j] = word;
			}
		} else {
			block_bitonic(
			    n, m, [&](uint32_t i) { return frag_key(b[i]); },
			    [&](uint32_t i, uint64_t v) { b[i] = make_uint2((uint32_t)(v >> 32), (uint32_t)v); });
			if (threadIdx.x == 0) {
				uint32_t word = 0u;
				for (uint32_t a = 0; a < n; ++a)
					word = fold_push(word, b[a].y & 0xffffffu, a);
				oct[begin + j] = word;
			}
		}
		__syncthreads();
	}
}

// ---- scan plumbing for the allocation and the leaf offsets ------------------------------------
struct TagBits { // octree_alloc_node.comp:12: (uOctree[idx] & 0x80000000u) > 0u
	const uint32_t *oct;
	const BuildState *st;
	uint32_t slot;
	__device__ uint64_t size() const { return st->overflow ? 0u : st->num[slot]; }
	__device__ uint32_t operator()(uint64_t i) const { return oct[st->begin[slot] + i] >> 31u; }
};

struct AllocCommit { // octree_modify_arg.comp:9-13 plus the capacity check the reference lacks
	BuildState *st;
	uint32_t slot, level;
	uint64_t cap;
	__device__ void operator()(uint64_t total) const {
		if (st->overflow)
			return;
		uint64_t nb = (uint64_t)st->begin[slot] + st->num[slot];
		uint64_t nn = total * 8u;
		uint64_t counter = (uint64_t)st->counter[slot] + total;
		if (nb + nn > cap || (counter + 1u) * 8u > (1ull << 30)) {
			st->overflow = level;
			st->need_words = nb + nn;
			return;
		}
		const uint32_t o = slot ^ 1u;
		st->begin[o] = (uint32_t)nb;
		st->num[o] = (uint32_t)nn;
		st->counter[o] = (uint32_t)counter;
	}
};

struct AllocWrite { // octree_alloc_node.comp:16-21 with the k-th tagged word -> block counter+k+1
	uint32_t *oct;
	const BuildState *st;
	uint32_t slot;
	__device__ bool active() const { return st->overflow == 0u; }
	__device__ void operator()(uint64_t i, uint64_t k, uint32_t flag) const {
		if (!flag)
			return;
		uint32_t blk = (st->counter[slot] + (uint32_t)k + 1u) << 3u;
		oct[st->begin[slot] + i] = blk | 0x80000000u;
		uint4 *z = reinterpret_cast<uint4 *>(oct + blk); // octree_init_node.comp for the new block
		z[0] = make_uint4(0u, 0u, 0u, 0u);
		z[1] = make_uint4(0u, 0u, 0u, 0u);
	}
};

struct LeafCounts {
	const uint32_t *oct;
	const BuildState *st;
	uint32_t slot;
	__device__ uint64_t size() const { return st->overflow ? 0u : st->num[slot]; }
	__device__ uint32_t operator()(uint64_t i) const { return oct[st->begin[slot] + i]; }
};
struct NoCommit {
	__device__ void operator()(uint64_t) const {}
};
struct LeafOffsets {
	uint32_t *aux;
	const BuildState *st;
	__device__ bool active() const { return st->overflow == 0u; }
	__device__ void operator()(uint64_t i, uint64_t k, uint32_t) const { aux[i] = (uint32_t)k; }
};

uint32_t grid_for(uint64_t items, svo_ctx *ctx) {
	uint64_t g = (items + kThreads - 1) / kThreads;
	uint64_t cap = (uint64_t)ctx->cu_count * 16u;
	if (g > cap)
		g = cap;
	return (uint32_t)(g ? g : 1);
}

int alloc_level(svo_ctx *ctx, svo_octree *oct, uint32_t i) {
	const uint32_t slot = i & 1u;
	hipError_t e;
	TagBits src{oct->words, ctx->d_state, slot};
	if ((e = launch(ctx, "alloc_reduce", scan_reduce<TagBits>, dim3(kScanBlocks), dim3(kScanThreads), 0, src,
	                ctx->d_scan_parts)) != hipSuccess)
		return hip_check(ctx, e, "alloc_reduce");
	AllocCommit commit{ctx->d_state, slot, i, oct->cap};
	if ((e = launch(ctx, "alloc_commit", scan_parts<AllocCommit>, dim3(1), dim3(1024), 0, ctx->d_scan_parts,
	                commit)) != hipSuccess)
		return hip_check(ctx, e, "alloc_commit");
	AllocWrite dst{oct->words, ctx->d_state, slot};
	if ((e = launch(ctx, "alloc_write", scan_down<TagBits, AllocWrite>, dim3(kScanBlocks), dim3(kScanThreads), 0,
	                src, dst, (const uint64_t *)ctx->d_scan_parts)) != hipSuccess)
		return hip_check(ctx, e, "alloc_write");
	return SVO_OK;
}

} // namespace

// Enqueue levels resume_level..L (resume_level = 1 for a fresh build; > 1 to redo the allocation of
// a level whose tags are in place after the node buffer was grown).
int build_enqueue(svo_ctx *ctx, svo_octree *oct, uint32_t resume_level) {
	const svo_fraglist *fl = oct->src;
	const uint32_t L = oct->level;
	const uint64_t F = fl->F;
	const uint2 *frags = reinterpret_cast<const uint2 *>(fl->frags);
	hipError_t e;
	const uint32_t gF = grid_for(F, ctx);
	if (resume_level <= 1) {
		if ((e = launch(ctx, "build_init", build_init, dim3(1), dim3(64), 0, ctx->d_state, oct->words)) !=
		    hipSuccess)
			return hip_check(ctx, e, "build_init");
	} else {
		if ((e = launch(ctx, "build_clear_overflow", build_clear_overflow, dim3(1), dim3(1), 0, ctx->d_state)) !=
		    hipSuccess)
			return hip_check(ctx, e, "build_clear_overflow");
	}
	for (uint32_t i = resume_level; i < L; ++i) {
		if (i != resume_level || resume_level <= 1) {
			if ((e = launch(ctx, "build_tag", build_tag, dim3(gF), dim3(kThreads), 0, frags, F, oct->node, oct->words,
			                (const BuildState *)ctx->d_state, L, i)) != hipSuccess)
				return hip_check(ctx, e, "build_tag");
		}
		int rc = alloc_level(ctx, oct, i);
		if (rc != SVO_OK)
			return rc;
	}
	// leaf level
	if ((e = launch(ctx, "build_tag_leaf", build_tag_leaf, dim3(gF), dim3(kThreads), 0, frags, F, oct->node,
	                oct->words, (const BuildState *)ctx->d_state, L)) != hipSuccess)
		return hip_check(ctx, e, "build_tag_leaf");
	const uint32_t slot = L & 1u;
	LeafCounts cnt{oct->words, ctx->d_state, slot};
	if ((e = launch(ctx, "leaf_reduce", scan_reduce<LeafCounts>, dim3(kScanBlocks), dim3(kScanThreads), 0, cnt,
	                ctx->d_scan_parts)) != hipSuccess)
		return hip_check(ctx, e, "leaf_reduce");
	if ((e = launch(ctx, "leaf_parts", scan_parts<NoCommit>, dim3(1), dim3(1024), 0, ctx->d_scan_parts,
	                NoCommit{})) != hipSuccess)
		return hip_check(ctx, e, "leaf_parts");
	LeafOffsets offs{oct->aux, ctx->d_state};
	if ((e = launch(ctx, "leaf_offsets", scan_down<LeafCounts, LeafOffsets>, dim3(kScanBlocks), dim3(kScanThreads),
	                0, cnt, offs, (const uint64_t *)ctx->d_scan_parts)) != hipSuccess)
		return hip_check(ctx, e, "leaf_offsets");
	if ((e = launch(ctx, "build_scatter", build_scatter, dim3(gF), dim3(kThreads), 0, frags, F,
	                (const uint32_t *)oct->node, oct->aux, reinterpret_cast<uint2 *>(oct->bucket),
	                (const BuildState *)ctx->d_state)) != hipSuccess)
		return hip_check(ctx, e, "build_scatter");
	// big-bucket list counter lives in big[0]
	if ((e = launch(ctx, "big_reset", build_init_big, dim3(1), dim3(1), 0, oct->big)) != hipSuccess)
		return hip_check(ctx, e, "big_reset");
	const uint64_t big_cap = F / (kSmallBucket + 1u) + 1u;
	uint32_t gW = (uint32_t)((uint64_t)ctx->cu_count * 16u);
	if ((e = launch(ctx, "build_fold", build_fold, dim3(gW), dim3(kThreads), 0, oct->words,
	                (const uint32_t *)oct->aux, reinterpret_cast<uint2 *>(oct->bucket), oct->big, big_cap,
	                (const BuildState *)ctx->d_state, L)) != hipSuccess)
		return hip_check(ctx, e, "build_fold");
	if ((e = launch(ctx, "build_fold_big", build_fold_big, dim3((uint32_t)ctx->cu_count * 2u), dim3(kThreads), 0,
	                oct->words, (const uint32_t *)oct->aux, reinterpret_cast<uint2 *>(oct->bucket),
	                (const uint32_t *)oct->big, big_cap, (const BuildState *)ctx->d_state, L)) != hipSuccess)
		return hip_check(ctx, e, "build_fold_big");
	return SVO_OK;
}

} // namespace svo
