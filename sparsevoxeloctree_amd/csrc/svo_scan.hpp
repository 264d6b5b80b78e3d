// Device-wide exclusive scan (reduce-then-scan, no inter-block waiting) over a sequence whose
// length may live in device memory, so a scan can follow a kernel that decides its size without a
// host round trip.  Two launches:
//   scan_reduce<Src>          kScanBlocks blocks: block b sums its contiguous chunk -> parts[b]
//   scan_down<Src,Dst,Commit> kScanBlocks blocks: block b adds up parts[0..b) for its chunk's offset
//                             (block 0 also passes the grand total to commit), re-reads the chunk,
//                             block-local scan + chunk offset, calls dst(i, exclusive_prefix, value)
//                             for every item
// Src: __device__ uint64_t size() const; __device__ uint32_t operator()(uint64_t i) const;
// static constexpr int kItems: consecutive items per thread and tile of scan_down
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "svo_runtime.hpp"

namespace svo {

template <class Src> constexpr uint64_t scan_tile() { return (uint64_t)kScanThreads * Src::kItems; }

// Block b's chunk of the n items: whole tiles of tile items.
__device__ __forceinline__ void scan_chunk(uint64_t n, uint32_t b, uint64_t tile, uint64_t &lo, uint64_t &hi) {
	uint64_t per = (n + kScanBlocks - 1) / kScanBlocks;
	per = (per + tile - 1) / tile * tile;
	lo = b * per;
	lo = lo < n ? lo : n;
	hi = lo + per;
	hi = hi < n ? hi : n;
}

// Inclusive scan across the 64 lanes of a wave.
__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v) {
	const int lane = threadIdx.x & 63;
#pragma unroll
	for (int d = 1; d < 64; d <<= 1) {
		uint64_t o = __shfl_up(v, d, 64);
		if (lane >= d)
			v += o;
	}
	return v;
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
#pragma unroll
	for (int d = 32; d >= 1; d >>= 1)
		v += __shfl_xor(v, d, 64);
	return v;
}

// Block-wide exclusive scan of one value per thread (kScanThreads threads); returns the block total.
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t v, uint64_t &excl) {
	__shared__ uint64_t s_wave[kScanThreads / 64];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	uint64_t incl = wave_incl_scan(v);
	if (lane == 63)
		s_wave[wave] = incl;
	__syncthreads();
	uint64_t wave_off = 0, total = 0;
#pragma unroll
	for (int w = 0; w < kScanThreads / 64; ++w) {
		uint64_t t = s_wave[w];
		wave_off += (w < wave) ? t : 0;
		total += t;
	}
	__syncthreads();
	excl = wave_off + incl - v;
	return total;
}

template <class Src> __global__ void __launch_bounds__(kScanThreads) scan_reduce(Src src, uint64_t *parts) {
	uint64_t lo, hi;
	scan_chunk(src.size(), blockIdx.x, scan_tile<Src>(), lo, hi);
	uint64_t acc = 0;
	uint64_t i = lo + threadIdx.x;
	// 8 independent loads in flight per thread
	for (; i + 7u * kScanThreads < hi; i += 8u * kScanThreads) {
		uint32_t v[8];
#pragma unroll
		for (int k = 0; k < 8; ++k)
			v[k] = src(i + (uint64_t)k * kScanThreads);
#pragma unroll
		for (int k = 0; k < 8; ++k)
			acc += v[k];
	}
	for (; i < hi; i += kScanThreads)
		acc += src(i);
	acc = wave_sum(acc);
	__shared__ uint64_t s[kScanThreads / 64];
	if ((threadIdx.x & 63) == 0)
		s[threadIdx.x >> 6] = acc;
	__syncthreads();
	if (threadIdx.x == 0) {
		uint64_t t = 0;
		for (int w = 0; w < kScanThreads / 64; ++w)
			t += s[w];
		parts[blockIdx.x] = t;
	}
}

// Sum of parts[0..blockIdx.x) (the chunk totals of the blocks before this one); block 0 passes the
// sum of all parts to commit.  Every thread of the block must call it.
template <class Commit> __device__ __forceinline__ uint64_t chunk_offset(const uint64_t *parts, Commit &commit) {
	static_assert(kScanBlocks % kScanThreads == 0, "parts are read kScanThreads at a time");
	__shared__ uint64_t s_below[kScanThreads / 64], s_all[kScanThreads / 64];
	uint64_t below = 0, all = 0;
#pragma unroll
	for (int j0 = 0; j0 < kScanBlocks; j0 += kScanThreads) {
		const int j = j0 + (int)threadIdx.x;
		const uint64_t p = parts[j];
		below += j < (int)blockIdx.x ? p : 0;
		all += p;
	}
	below = wave_sum(below);
	all = wave_sum(all);
	if ((threadIdx.x & 63) == 0) {
		s_below[threadIdx.x >> 6] = below;
		s_all[threadIdx.x >> 6] = all;
	}
	__syncthreads();
	below = all = 0;
#pragma unroll
	for (int w = 0; w < kScanThreads / 64; ++w) {
		below += s_below[w];
		all += s_all[w];
	}
	if (blockIdx.x == 0 && threadIdx.x == 0)
		commit(all);
	return below;
}

struct NoCommit {
	__device__ void operator()(uint64_t) const {}
};

template <class Src, class Dst, class Commit = NoCommit>
__global__ void __launch_bounds__(kScanThreads)
    scan_down(Src src, Dst dst, const uint64_t *parts, Commit commit = Commit{}) {
	constexpr int kItems = Src::kItems;
	uint64_t base = chunk_offset(parts, commit);
	uint64_t lo, hi;
	scan_chunk(src.size(), blockIdx.x, scan_tile<Src>(), lo, hi);
	if (lo >= hi || !dst.active())
		return;
	for (uint64_t t0 = lo; t0 < hi; t0 += scan_tile<Src>()) {
		// each thread owns kItems consecutive items of the tile
		uint32_t v[kItems];
		uint64_t first = t0 + (uint64_t)threadIdx.x * kItems;
		uint64_t sum = 0;
#pragma unroll
		for (int k = 0; k < kItems; ++k) {
			uint64_t i = first + k;
			v[k] = i < hi ? src(i) : 0u;
			sum += v[k];
		}
		uint64_t excl;
		uint64_t total = block_excl_scan(sum, excl);
		uint64_t run = base + excl;
#pragma unroll
		for (int k = 0; k < kItems; ++k) {
			uint64_t i = first + k;
			if (i < hi)
				dst(i, run, v[k]);
			run += v[k];
		}
		base += total;
	}
}

} // namespace svo
