// Device-wide exclusive scan (reduce-then-scan, no inter-block waiting) over a sequence whose
// length may live in device memory, so a scan can follow a kernel that decides its size without a
// host round trip.  Three launches:
//   scan_reduce<Src>   kScanBlocks blocks: block b sums its contiguous chunk -> parts[b]
//   scan_parts<Commit> 1 block: exclusive scan of parts -> parts[kScanBlocks + b], commit(total)
//   scan_down<Src,Dst> kScanBlocks blocks: re-reads the chunk, block-local scan + chunk offset,
//                      calls dst(i, exclusive_prefix, value) for every item
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

// One block of 1024 threads (= kScanBlocks).
template <class Commit> __global__ void __launch_bounds__(1024) scan_parts(uint64_t *parts, Commit commit) {
	static_assert(kScanBlocks == 1024, "scan_parts assumes one part per thread");
	__shared__ uint64_t s_wave[16];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	uint64_t v = parts[threadIdx.x];
	uint64_t incl = wave_incl_scan(v);
	if (lane == 63)
		s_wave[wave] = incl;
	__syncthreads();
	uint64_t off = 0, total = 0;
	for (int w = 0; w < 16; ++w) {
		off += (w < wave) ? s_wave[w] : 0;
		total += s_wave[w];
	}
	parts[kScanBlocks + threadIdx.x] = off + incl - v;
	if (threadIdx.x == 0)
		commit(total);
}

template <class Src, class Dst>
__global__ void __launch_bounds__(kScanThreads) scan_down(Src src, Dst dst, const uint64_t *parts) {
	constexpr int kItems = Src::kItems;
	uint64_t lo, hi;
	scan_chunk(src.size(), blockIdx.x, scan_tile<Src>(), lo, hi);
	if (lo >= hi || !dst.active())
		return;
	uint64_t base = parts[kScanBlocks + blockIdx.x];
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
