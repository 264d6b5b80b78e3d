// Octree construction (OctreeBuilder::CmdBuild, src/OctreeBuilder.cpp:142-210).
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
//   octree_alloc_node.comp    -> one exclusive popcount scan over all bitmaps (rank of every
//                                occupied node among its depth = its block among the allocated
//                                children blocks); child pointer = 8 * (first block of depth + rank)
//   octree_modify_arg.comp    -> bm_levels: per-depth counts -> first block of every depth, the
//                                leaf range {begin, num} and W = (counter + 1) * 8
//   octree_init_node.comp     -> bm_write_blocks writes every block whole (pointers and zeros)
//   octree_tag_node.comp      -> build_tag_leaf: leaf block ("bin") of each fragment = rank of its
//   (level L) + leaf CAS fold    depth-(L-1) node; count per bin, scan, scatter into per-bin buckets;
//                                build_fold{,_mid,_big} sort each bucket by the 64-bit fragment and
//                                fold the 8 leaves of the block in that order (the reference's order
//                                is whatever its atomics produce), writing the leaf block whole.
// The node words are identical to the reference's pass structure run with a serial allocator.
#include "svo_device.hpp"
#include "svo_runtime.hpp"
#include "svo_scan.hpp"

namespace svo {
namespace {

constexpr int kThreads = 256;
constexpr uint32_t kLdsBucket = 4096;   // largest bucket sorted in LDS

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

// Bitmap layout: B[d] (d = 1..L-1, 8^d bits) at word offset off[d]; R holds the exclusive popcount
// prefix of every word of the concatenation.
struct Levels {
	uint32_t L;
	uint64_t off[14]; // off[d] for d = 1..L, off[L] = total words
};

// Rank of node P among the occupied nodes of depth d.
__device__ __forceinline__ uint32_t node_rank(const uint32_t *__restrict__ B, const uint32_t *__restrict__ R,
                                              uint64_t off, uint64_t P) {
	const uint64_t w = off + (P >> 5);
	return (R[w] - R[off]) + (uint32_t)__popc(B[w] & ((1u << (uint32_t)(P & 31u)) - 1u));
}

__global__ void __launch_bounds__(kThreads) bm_zero(uint32_t *B, uint64_t n) {
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride)
		B[i] = 0u;
}

// Each fragment marks its depth-(L-1) node.  Neighbouring fragments mostly fall into the same
// bitmap word, so a wave ORs the bits of lanes with equal words together (segmented scan over
// shuffles) and only the last lane of every run of equal words touches memory: one atomic when a
// plain load does not already show all of the run's bits.
__global__ void __launch_bounds__(kThreads)
    bm_mark(const uint2 *__restrict__ frags, uint64_t F, uint32_t L, uint32_t *Blast) {
	const uint32_t res = 1u << L, lane = threadIdx.x & 63u;
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	for (uint64_t base = (uint64_t)blockIdx.x * kThreads; base < F; base += stride) {
		const uint64_t f = base + threadIdx.x;
		uint32_t word = 0xFFFFFFFFu, bits = 0u; // node index < 2^33, so word < 2^28
		if (f < F) {
			uint32_t x, y, z;
			decode_frag(frags[f], res, x, y, z);
			const uint64_t P = morton(x, y, z) >> 3;
			word = (uint32_t)(P >> 5);
			bits = 1u << (uint32_t)(P & 31u);
		}
#pragma unroll
		for (uint32_t d = 1; d < 64u; d <<= 1) {
			const uint32_t ow = __shfl_up(word, d, 64), ob = __shfl_up(bits, d, 64);
			bits |= (lane >= d && ow == word) ? ob : 0u; // any lane with the same word may contribute
		}
		const uint32_t next = __shfl_down(word, 1u, 64);
		if (word != 0xFFFFFFFFu && (lane == 63u || next != word)) {
			uint32_t *w = &Blast[word];
			if ((*w & bits) != bits) // a stale cached word only costs a redundant atomic
				atomicOr(w, bits);
		}
	}
}

// B[d] bit q = OR of byte q of B[d+1] (node q occupied iff one of its 8 children is).
__global__ void __launch_bounds__(kThreads)
    bm_reduce(const uint32_t *__restrict__ Bc, uint64_t nwc, uint32_t *__restrict__ Bp, uint64_t nwp) {
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	for (uint64_t w = (uint64_t)blockIdx.x * kThreads + threadIdx.x; w < nwp; w += stride) {
		uint32_t bits = 0u;
#pragma unroll
		for (uint32_t k = 0; k < 8u; ++k) {
			const uint64_t cw = 8u * w + k;
			const uint32_t v = cw < nwc ? Bc[cw] : 0u;
#pragma unroll
			for (uint32_t b = 0; b < 4u; ++b)
				bits |= (((v >> (8u * b)) & 0xffu) != 0u ? 1u : 0u) << (4u * k + b);
		}
		Bp[w] = bits;
	}
}

struct PopSrc {
	const uint32_t *B;
	uint64_t n;
	__device__ uint64_t size() const { return n; }
	__device__ uint32_t operator()(uint64_t i) const { return (uint32_t)__popc(B[i]); }
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

// Per-depth node counts -> first block of every depth (fb[0] = 0 is the root block, fb[1] = 1),
// the leaf range and W (src/OctreeBuilder.cpp:212-214: W = (counter + 1) * 8).
__global__ void bm_levels(const uint32_t *__restrict__ R, Levels lv, uint32_t *fb, BuildState *st, uint64_t cap) {
	if (threadIdx.x != 0)
		return;
	const uint32_t L = lv.L;
	uint64_t first = 1;
	fb[0] = 0u;
	fb[1] = 1u;
	uint64_t nlast = 1; // nodes at depth L-1 (the root when L == 1)
	for (uint32_t d = 1; d < L; ++d) {
		const uint64_t nd = (uint64_t)R[lv.off[d + 1]] - R[lv.off[d]];
		first += nd;
		fb[d + 1] = (uint32_t)(first < 0xFFFFFFFFull ? first : 0xFFFFFFFFull);
		nlast = nd;
	}
	const uint64_t W = 8u * first;
	const uint32_t slot = L & 1u;
	st->need_words = W;
	st->overflow = W > cap ? L : 0u;
	st->begin[slot] = (uint32_t)(8u * (L == 1u ? 0u : fb[L - 1])); // first leaf word
	st->num[slot] = (uint32_t)nlast;                              // leaf blocks (one per depth-(L-1) node)
	st->counter[slot] = (uint32_t)(first - 1u);
}

// Blocks of depth d (the children of depth-(d-1) nodes), one thread per word of B[d] = 4 parents.
// Every block is written whole: pointer words for occupied children, zeros otherwise.
__global__ void __launch_bounds__(kThreads)
    bm_write_blocks(const uint32_t *__restrict__ B, const uint32_t *__restrict__ R, Levels lv, uint32_t d,
                    const uint32_t *__restrict__ fb, uint32_t *oct) {
	const uint64_t offd = lv.off[d], nw = lv.off[d + 1] - offd;
	const uint32_t fbd = fb[d], fbp = fb[d - 1];
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	for (uint64_t w = (uint64_t)blockIdx.x * kThreads + threadIdx.x; w < nw; w += stride) {
		const uint32_t v = B[offd + w];
		const uint32_t base = R[offd + w] - R[offd];
#pragma unroll
		for (uint32_t k = 0; k < 4u; ++k) {
			const uint32_t byte = (v >> (8u * k)) & 0xffu;
			const uint64_t q = 4u * w + k; // parent at depth d-1
			if (d == 1u ? q != 0u : byte == 0u)
				continue; // the root block always exists
			const uint32_t block = d == 1u ? 0u : fbp + node_rank(B, R, lv.off[d - 1], q);
			uint32_t r = base + (uint32_t)__popc(v & ((1u << (8u * k)) - 1u));
			uint32_t wd[8];
#pragma unroll
			for (uint32_t s = 0; s < 8u; ++s) {
				const bool set = (byte >> s) & 1u;
				wd[s] = set ? (0x80000000u | ((fbd + r) << 3u)) : 0u;
				r += set ? 1u : 0u;
			}
			uint4 *dst = reinterpret_cast<uint4 *>(oct + 8u * (uint64_t)block);
			dst[0] = make_uint4(wd[0], wd[1], wd[2], wd[3]);
			dst[1] = make_uint4(wd[4], wd[5], wd[6], wd[7]);
		}
	}
}

// Leaf block ("bin") of every fragment: the rank of its depth-(L-1) node.  Fragments are counted
// per bin with one atomic per run of equal neighbours; the atomic's old value gives every fragment
// its place inside the bin, so node[f] = {bin, place} and the scatter needs no atomics.
__global__ void __launch_bounds__(kThreads)
    build_tag_leaf(const uint2 *__restrict__ frags, uint64_t F, const uint32_t *__restrict__ B,
                   const uint32_t *__restrict__ R, Levels lv, uint2 *__restrict__ node, uint32_t *cnt) {
	const uint32_t L = lv.L, res = 1u << L, lane = threadIdx.x & 63u;
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	for (uint64_t base = (uint64_t)blockIdx.x * kThreads; base < F; base += stride) {
		const uint64_t f = base + threadIdx.x;
		uint32_t bin = 0xFFFFFFFFu;
		if (f < F) {
			uint32_t x, y, z;
			decode_frag(frags[f], res, x, y, z);
			bin = L == 1u ? 0u : node_rank(B, R, lv.off[L - 1], morton(x, y, z) >> 3);
		}
		const uint32_t prev = __shfl_up(bin, 1u, 64);
		const bool head = (lane == 0u) || prev != bin;
		const uint64_t heads = __ballot(head);
		const uint64_t after = (lane == 63u) ? 0ull : (heads >> (lane + 1u));
		const uint32_t run = after ? (uint32_t)__ffsll((unsigned long long)after) : 64u - lane;
		uint32_t pos = 0u;
		if (head && bin != 0xFFFFFFFFu)
			pos = atomicAdd(&cnt[bin], run);
		// lane of this lane's run head = highest head at or below this lane
		const uint64_t upto = heads & ((lane == 63u) ? ~0ull : ((2ull << lane) - 1ull));
		const uint32_t head_lane = 63u - (uint32_t)__clzll((long long)upto);
		pos = __shfl(pos, (int)head_lane, 64) + (lane - head_lane);
		if (f < F)
			node[f] = make_uint2(bin, pos);
	}
}

// Scatter every fragment to its place in its bin's bucket (aux[j] = first slot of bin j).
__global__ void __launch_bounds__(kThreads)
    build_scatter(const uint2 *__restrict__ frags, uint64_t F, const uint2 *__restrict__ node,
                  const uint32_t *__restrict__ aux, uint2 *__restrict__ bucket) {
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	for (uint64_t f = (uint64_t)blockIdx.x * kThreads + threadIdx.x; f < F; f += stride) {
		const uint2 np = node[f];
		bucket[aux[np.x] + np.y] = frags[f];
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

// Running average of the reference's CAS loop (octree_tag_node.comp:43-57): fold colour c into a
// leaf word that has folded a fragments so far.  In a sequence the a-th fragment (a >= 1) meets the
// word's 6-bit count n = min(a, 63); avg' = (avg * n + c) / (n + 1) per channel.
__device__ __forceinline__ uint32_t fold_rt(uint32_t prev, uint32_t c, uint32_t a) {
	if (a == 0u)
		return 0xC1000000u | c;
	const uint32_t n = a < 63u ? a : 63u, w = n + 1u, m = kDivMagic[w];
	const uint32_t r = __umulhi((prev & 0xffu) * n + (c & 0xffu), m);
	const uint32_t g = __umulhi(((prev >> 8u) & 0xffu) * n + ((c >> 8u) & 0xffu), m);
	const uint32_t b = __umulhi(((prev >> 16u) & 0xffu) * n + ((c >> 16u) & 0xffu), m);
	return (umin32(w, 0x3fu) << 24u) | (r & 0xffu) | ((g & 0xffu) << 8u) | ((b & 0xffu) << 16u) | 0xC0000000u;
}

// The leaf a fragment key belongs to inside its block (child slot at depth L).
__device__ __forceinline__ uint32_t key_slot(uint64_t key, uint32_t res) {
	uint32_t x, y, z;
	decode_frag(make_uint2((uint32_t)(key >> 32), (uint32_t)key), res, x, y, z);
	return (x & 1u) | ((y & 1u) << 1u) | ((z & 1u) << 2u);
}

// Folds one key (in ascending key order) into the 8 leaves of a block: w[s] words, c[s] counts.
// The slot is selected without dynamic register indexing.
__device__ __forceinline__ void slot_push(uint32_t (&w)[8], uint32_t (&c)[8], uint64_t key, uint32_t res) {
	const uint32_t s = key_slot(key, res);
	uint32_t cw = w[0], cc = c[0];
#pragma unroll
	for (uint32_t t = 1; t < 8u; ++t) {
		cw = s == t ? w[t] : cw;
		cc = s == t ? c[t] : cc;
	}
	const uint32_t nw = fold_rt(cw, (uint32_t)key & 0xffffffu, cc);
#pragma unroll
	for (uint32_t t = 0; t < 8u; ++t) {
		w[t] = s == t ? nw : w[t];
		c[t] = s == t ? cc + 1u : c[t];
	}
}

__device__ __forceinline__ void store_block(uint32_t *oct, uint64_t block, const uint32_t (&w)[8]) {
	uint4 *dst = reinterpret_cast<uint4 *>(oct + 8u * block);
	dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
	dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

constexpr uint32_t kRegBucket = 16; // bins up to this size: one thread each (build_fold)
constexpr uint32_t kWaveBucket = 64; // up to this size: one wave each (build_fold_mid)

// Work lists of bins too large for build_fold: list[0] = mid count, list[1] = big count, mid bins
// at list[2 ..], big bins at list[2 + cap ..] (each list holds at most F / 17 bins).
__device__ __forceinline__ void push_bin(uint32_t *list, uint64_t cap, uint32_t j, uint32_t n) {
	const uint32_t which = n > kWaveBucket ? 1u : 0u;
	const uint32_t q = atomicAdd(&list[which], 1u);
	if (q < cap)
		list[2u + which * cap + q] = j;
}

// One thread per leaf block: its fragments (<= 16) are sorted by the 64-bit fragment in registers
// (19- or 63-comparator networks, +inf padding), then folded in that order into the 8 leaves; the
// whole block is written, empty leaves as 0.
__global__ void __launch_bounds__(kThreads)
    build_fold(uint32_t *oct, const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ aux,
               const uint2 *__restrict__ bucket, uint32_t *list, uint64_t cap, const BuildState *st, uint32_t L) {
	const uint32_t slot = L & 1u, res = 1u << L;
	const uint64_t first = st->begin[slot] / 8u;
	const uint32_t nbins = st->num[slot];
	const uint32_t stride = gridDim.x * kThreads;
	for (uint32_t j = blockIdx.x * kThreads + threadIdx.x; j < nbins; j += stride) {
		const uint32_t n = cnt[j];
		if (n > kRegBucket) {
			push_bin(list, cap, j, n);
			continue;
		}
		const uint2 *b = bucket + aux[j];
		uint32_t w[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u}, c[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
		uint64_t k[16];
#pragma unroll
		for (uint32_t a = 0; a < 16u; ++a)
			k[a] = a < n ? frag_key(b[a]) : ~0ull;
		if (n <= 8u) {
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
		for (uint32_t a = 0; a < 16u; ++a)
			if (a < n)
				slot_push(w, c, k[a], res);
		store_block(oct, first + j, w);
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

// Bins of 17..64 fragments, one wave each: lane i holds key i (+inf padding), a bitonic sort over
// lane shuffles orders them, the sorted keys go through LDS, and lane s < 8 folds the keys of leaf
// slot s in ascending order.
__global__ void __launch_bounds__(kThreads)
    build_fold_mid(uint32_t *oct, const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ aux,
                   const uint2 *__restrict__ bucket, const uint32_t *list, uint64_t cap, const BuildState *st,
                   uint32_t L) {
	constexpr uint32_t kWaves = kThreads / 64;
	__shared__ uint64_t s_keys[kWaves][64];
	const uint32_t res = 1u << L, lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
	const uint64_t first = st->begin[L & 1u] / 8u;
	uint64_t count = list[0];
	if (count > cap)
		count = cap;
	for (uint64_t base = (uint64_t)blockIdx.x * kWaves; base < count; base += (uint64_t)gridDim.x * kWaves) {
		const uint64_t item = base + wave;
		uint32_t j = 0u, n = 0u;
		if (item < count) {
			j = list[2u + item];
			n = cnt[j];
		}
		uint64_t key = lane < n ? frag_key(bucket[aux[j] + lane]) : ~0ull;
#pragma unroll
		for (uint32_t k = 2; k <= 64u; k <<= 1) {
#pragma unroll
			for (uint32_t d = k >> 1; d > 0u; d >>= 1) {
				const uint64_t other = __shfl_xor(key, (int)d, 64);
				const bool up = (lane & k) == 0u, lower = (lane & d) == 0u;
				const bool take_min = up == lower;
				key = take_min ? (other < key ? other : key) : (other < key ? key : other);
			}
		}
		s_keys[wave][lane] = key;
		__syncthreads();
		if (item < count && lane < 8u) {
			uint32_t word = 0u, a = 0u;
			for (uint32_t i = 0; i < n; ++i) {
				const uint64_t kk = s_keys[wave][i];
				if (key_slot(kk, res) == lane)
					word = fold_rt(word, (uint32_t)kk & 0xffffffu, a++);
			}
			oct[8u * (first + j) + lane] = word;
		}
		__syncthreads();
	}
}

// Bins of more than 64 fragments, one block each: bitonic sort of the keys (in LDS up to 4096, in
// place in the bucket beyond), then 8 lanes fold the 8 leaves, each lane walking the sorted keys
// and taking those of its slot.
__global__ void __launch_bounds__(kThreads)
    build_fold_big(uint32_t *oct, const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ aux, uint2 *bucket,
                   const uint32_t *list, uint64_t cap, const BuildState *st, uint32_t L) {
	__shared__ uint64_t s_keys[kLdsBucket];
	const uint32_t res = 1u << L;
	const uint64_t first = st->begin[L & 1u] / 8u;
	uint64_t count = list[1];
	if (count > cap)
		count = cap;
	for (uint64_t item = blockIdx.x; item < count; item += gridDim.x) {
		const uint32_t j = list[2u + cap + item];
		const uint32_t n = cnt[j];
		uint2 *b = bucket + aux[j];
		uint32_t m = 1;
		while (m < n)
			m <<= 1;
		const bool lds = n <= kLdsBucket;
		if (lds) {
			for (uint32_t t = threadIdx.x; t < n; t += blockDim.x)
				s_keys[t] = frag_key(b[t]);
			__syncthreads();
			block_bitonic(
			    n, m, [&](uint32_t i) { return s_keys[i]; }, [&](uint32_t i, uint64_t v) { s_keys[i] = v; });
		} else {
			block_bitonic(
			    n, m, [&](uint32_t i) { return frag_key(b[i]); },
			    [&](uint32_t i, uint64_t v) { b[i] = make_uint2((uint32_t)(v >> 32), (uint32_t)v); });
		}
		if (threadIdx.x < 8u) {
			uint32_t word = 0u, a = 0u;
			for (uint32_t i = 0; i < n; ++i) {
				const uint64_t key = lds ? s_keys[i] : frag_key(b[i]);
				if (key_slot(key, res) == threadIdx.x)
					word = fold_rt(word, (uint32_t)key & 0xffffffu, a++);
			}
			oct[8u * (first + j) + threadIdx.x] = word;
		}
		__syncthreads();
	}
}

__global__ void build_init_lists(uint32_t *list) {
	list[0] = 0u;
	list[1] = 0u;
}

// ---- scan plumbing for the bin offsets ---------------------------------------------------------
struct BinCounts {
	const uint32_t *cnt;
	uint64_t n;
	__device__ uint64_t size() const { return n; }
	__device__ uint32_t operator()(uint64_t i) const { return cnt[i]; }
};
struct NoCommit {
	__device__ void operator()(uint64_t) const {}
};
struct BinOffsets {
	uint32_t *aux;
	__device__ bool active() const { return true; }
	__device__ void operator()(uint64_t i, uint64_t k, uint32_t) const { aux[i] = (uint32_t)k; }
};

uint32_t grid_for(uint64_t items, svo_ctx *ctx) {
	uint64_t g = (items + kThreads - 1) / kThreads;
	const uint64_t cap = (uint64_t)ctx->cu_count * 16u;
	if (g > cap)
		g = cap;
	return (uint32_t)(g ? g : 1);
}

Levels levels_for(uint32_t L) {
	Levels lv{};
	lv.L = L;
	uint64_t off = 0;
	for (uint32_t d = 1; d <= L; ++d) {
		lv.off[d] = off;
		if (d < L) {
			const uint64_t bits = 1ull << (3u * d);
			off += bits < 32u ? 1u : bits / 32u;
		}
	}
	return lv;
}

} // namespace

uint64_t build_bitmap_words(uint32_t L) { return levels_for(L).off[L]; }

uint64_t build_big_capacity(uint64_t F) { return F / (kRegBucket + 1u) + 1u; }

// Structure phase: bitmaps, ranks, level table; leaves st (slot L & 1) and W on the device.
int build_structure(svo_ctx *ctx, svo_octree *oct) {
	const svo_fraglist *fl = oct->src;
	const uint32_t L = oct->level;
	const uint64_t F = fl->F;
	const uint2 *frags = reinterpret_cast<const uint2 *>(fl->frags);
	const Levels lv = levels_for(L);
	const uint64_t TW = lv.off[L];
	hipError_t e;
	if (TW > 0) {
		if ((e = launch(ctx, "bm_zero", bm_zero, dim3(grid_for(TW, ctx)), dim3(kThreads), 0, oct->bm, TW)) != hipSuccess)
			return hip_check(ctx, e, "bm_zero");
		if (F > 0 &&
		    (e = launch(ctx, "bm_mark", bm_mark, dim3(grid_for(F, ctx)), dim3(kThreads), 0, frags, F, L,
		                oct->bm + lv.off[L - 1])) != hipSuccess)
			return hip_check(ctx, e, "bm_mark");
		for (uint32_t d = L - 1; d-- > 1;) {
			const uint64_t nwc = lv.off[d + 2] - lv.off[d + 1], nwp = lv.off[d + 1] - lv.off[d];
			if ((e = launch(ctx, "bm_reduce", bm_reduce, dim3(grid_for(nwp, ctx)), dim3(kThreads), 0,
			                (const uint32_t *)(oct->bm + lv.off[d + 1]), nwc, oct->bm + lv.off[d], nwp)) != hipSuccess)
				return hip_check(ctx, e, "bm_reduce");
		}
		PopSrc src{oct->bm, TW};
		if ((e = launch(ctx, "rank_reduce", scan_reduce<PopSrc>, dim3(kScanBlocks), dim3(kScanThreads), 0, src,
		                ctx->d_scan_parts)) != hipSuccess)
			return hip_check(ctx, e, "rank_reduce");
		if ((e = launch(ctx, "rank_parts", scan_parts<RankTotal>, dim3(1), dim3(1024), 0, ctx->d_scan_parts,
		                RankTotal{oct->rk, TW})) != hipSuccess)
			return hip_check(ctx, e, "rank_parts");
		if ((e = launch(ctx, "rank_down", scan_down<PopSrc, RankDst>, dim3(kScanBlocks), dim3(kScanThreads), 0, src,
		                RankDst{oct->rk}, (const uint64_t *)ctx->d_scan_parts)) != hipSuccess)
			return hip_check(ctx, e, "rank_down");
	}
	if ((e = launch(ctx, "bm_levels", bm_levels, dim3(1), dim3(64), 0, (const uint32_t *)oct->rk, lv, oct->lvl,
	                ctx->d_state, (uint64_t)(1ull << 30))) != hipSuccess)
		return hip_check(ctx, e, "bm_levels");
	return SVO_OK;
}

// Node phase: the internal blocks, then the leaf blocks through bins (count, scan, scatter, sort +
// fold).  oct->words holds W words, oct->cnt and oct->aux hold nbins words.
int build_nodes(svo_ctx *ctx, svo_octree *oct) {
	const svo_fraglist *fl = oct->src;
	const uint32_t L = oct->level;
	const uint64_t F = fl->F, nbins = oct->nbins;
	const uint2 *frags = reinterpret_cast<const uint2 *>(fl->frags);
	const Levels lv = levels_for(L);
	hipError_t e;
	for (uint32_t d = 1; d < L; ++d) {
		const uint64_t nw = lv.off[d + 1] - lv.off[d];
		if ((e = launch(ctx, "bm_write_blocks", bm_write_blocks, dim3(grid_for(nw, ctx)), dim3(kThreads), 0,
		                (const uint32_t *)oct->bm, (const uint32_t *)oct->rk, lv, d, (const uint32_t *)oct->lvl,
		                oct->words)) != hipSuccess)
			return hip_check(ctx, e, "bm_write_blocks");
	}
	if ((e = launch(ctx, "bin_zero", bm_zero, dim3(grid_for(nbins, ctx)), dim3(kThreads), 0, oct->cnt, nbins)) !=
	    hipSuccess)
		return hip_check(ctx, e, "bin_zero");
	const uint32_t gF = grid_for(F, ctx);
	if (F > 0 && (e = launch(ctx, "build_tag_leaf", build_tag_leaf, dim3(gF), dim3(kThreads), 0, frags, F,
	                         (const uint32_t *)oct->bm, (const uint32_t *)oct->rk, lv, reinterpret_cast<uint2 *>(oct->node),
	                         oct->cnt)) !=
	                 hipSuccess)
		return hip_check(ctx, e, "build_tag_leaf");
	BinCounts bc{oct->cnt, nbins};
	if ((e = launch(ctx, "bin_reduce", scan_reduce<BinCounts>, dim3(kScanBlocks), dim3(kScanThreads), 0, bc,
	                ctx->d_scan_parts)) != hipSuccess)
		return hip_check(ctx, e, "bin_reduce");
	if ((e = launch(ctx, "bin_parts", scan_parts<NoCommit>, dim3(1), dim3(1024), 0, ctx->d_scan_parts, NoCommit{})) !=
	    hipSuccess)
		return hip_check(ctx, e, "bin_parts");
	if ((e = launch(ctx, "bin_offsets", scan_down<BinCounts, BinOffsets>, dim3(kScanBlocks), dim3(kScanThreads), 0, bc,
	                BinOffsets{oct->aux}, (const uint64_t *)ctx->d_scan_parts)) != hipSuccess)
		return hip_check(ctx, e, "bin_offsets");
	if (F > 0 && (e = launch(ctx, "build_scatter", build_scatter, dim3(gF), dim3(kThreads), 0, frags, F,
	                         (const uint2 *)oct->node, (const uint32_t *)oct->aux,
	                         reinterpret_cast<uint2 *>(oct->bucket))) != hipSuccess)
		return hip_check(ctx, e, "build_scatter");
	if ((e = launch(ctx, "list_reset", build_init_lists, dim3(1), dim3(1), 0, oct->big)) != hipSuccess)
		return hip_check(ctx, e, "list_reset");
	const uint64_t cap = build_big_capacity(F);
	const BuildState *st = ctx->d_state;
	if ((e = launch(ctx, "build_fold", build_fold, dim3(grid_for(nbins, ctx)), dim3(kThreads), 0, oct->words,
	                (const uint32_t *)oct->cnt, (const uint32_t *)oct->aux, (const uint2 *)oct->bucket, oct->big, cap,
	                st, L)) != hipSuccess)
		return hip_check(ctx, e, "build_fold");
	if ((e = launch(ctx, "build_fold_mid", build_fold_mid, dim3((uint32_t)ctx->cu_count * 8u), dim3(kThreads), 0,
	                oct->words, (const uint32_t *)oct->cnt, (const uint32_t *)oct->aux, (const uint2 *)oct->bucket,
	                (const uint32_t *)oct->big, cap, st, L)) != hipSuccess)
		return hip_check(ctx, e, "build_fold_mid");
	if ((e = launch(ctx, "build_fold_big", build_fold_big, dim3((uint32_t)ctx->cu_count * 2u), dim3(kThreads), 0,
	                oct->words, (const uint32_t *)oct->cnt, (const uint32_t *)oct->aux,
	                reinterpret_cast<uint2 *>(oct->bucket), (const uint32_t *)oct->big, cap, st, L)) != hipSuccess)
		return hip_check(ctx, e, "build_fold_big");
	return SVO_OK;
}

} // namespace svo
