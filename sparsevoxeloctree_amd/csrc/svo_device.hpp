// Device-side arithmetic with the GLSL semantics the hot path depends on, pinned to one definition:
// binary32 round-to-nearest, no contraction (the library is built with -ffp-contract=off),
// correctly rounded '/' and sqrt, GLSL min/max, saturating float->uint conversion.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace svo {

// GLSL min(x, y) = (y < x) ? y : x and max(x, y) = (x < y) ? y : x (written as selects, so the
// compiler cannot turn them into v_min/v_max with different signed-zero rules).
__device__ __forceinline__ float gmin(float x, float y) { return y < x ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return x < y ? y : x; }
__device__ __forceinline__ uint32_t umin32(uint32_t a, uint32_t b) { return b < a ? b : a; }
__device__ __forceinline__ uint32_t umax32(uint32_t a, uint32_t b) { return a < b ? b : a; }

// uint(float): truncation, negative and NaN -> 0, saturating above.
__device__ __forceinline__ uint32_t f2u_sat(float f) {
	if (!(f > 0.0f))
		return 0u;
	if (f >= 4294967296.0f)
		return 0xFFFFFFFFu;
	return (uint32_t)f;
}

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float bitsf(uint32_t u) { return __uint_as_float(u); }

// findMSB with findMSB(0) = 0xFFFFFFFF.
__device__ __forceinline__ uint32_t find_msb(uint32_t x) { return x ? 31u - (uint32_t)__clz(x) : 0xFFFFFFFFu; }

struct f3 {
	float x, y, z;
};

__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
	return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ float dot3(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ f3 normalize3(f3 v) {
	float s = 1.0f / __builtin_sqrtf((v.x * v.x + v.y * v.y) + v.z * v.z);
	return {v.x * s, v.y * s, v.z * s};
}

// Slot of this lane in a list appended to by the lanes whose pred holds: one atomicAdd on *counter
// per wave (a shared counter hit once per item serialises in one L2 channel).  Call it from every
// active lane; the result is meaningful where pred holds.
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool pred) {
	const uint64_t m = __ballot(pred);
	const uint32_t lane = threadIdx.x & 63u;
	uint32_t base = 0u;
	if (m) {
		const int leader = __ffsll((unsigned long long)m) - 1;
		if ((int)lane == leader)
			base = atomicAdd(counter, (uint32_t)__popcll(m));
		base = __shfl(base, leader, 64);
	}
	return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

} // namespace svo
