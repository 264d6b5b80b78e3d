// Plain device-to-device copy on the context's stream (grid-stride, 16-byte vectors when aligned).
#include "svo_runtime.hpp"

namespace svo {
namespace {

constexpr int kThreads = 256;

// One 16-byte element per thread over a grid that covers the copy (a grid-stride loop over a few
// blocks per CU reaches less of the HBM bandwidth), in launches of at most 2^31 - 1 blocks.
__global__ void __launch_bounds__(kThreads) copy_u128(uint4 *__restrict__ dst, const uint4 *__restrict__ src, uint64_t n) {
	const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
	if (i < n)
		dst[i] = src[i];
}

__global__ void __launch_bounds__(kThreads) copy_u8(uint8_t *__restrict__ dst, const uint8_t *__restrict__ src, uint64_t n) {
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride)
		dst[i] = src[i];
}

} // namespace

int copy_device(svo_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
	if (bytes == 0)
		return SVO_OK;
	const uint32_t grid = (uint32_t)ctx->cu_count * 8u;
	hipError_t e;
	if (((uintptr_t)dst | (uintptr_t)src | bytes) % 16 == 0) {
		const uint64_t n = bytes / 16, per = (uint64_t)0x7fffffff * kThreads; // elements per launch
		e = hipSuccess;
		for (uint64_t b = 0; b < n && e == hipSuccess; b += per) {
			const uint64_t m = n - b < per ? n - b : per;
			e = launch(ctx, "copy_u128", copy_u128, dim3((uint32_t)((m + kThreads - 1) / kThreads)), dim3(kThreads), 0,
			           reinterpret_cast<uint4 *>(dst) + b, reinterpret_cast<const uint4 *>(src) + b, m);
		}
	} else
		e = launch(ctx, "copy_u8", copy_u8, dim3(grid), dim3(kThreads), 0, reinterpret_cast<uint8_t *>(dst),
		           reinterpret_cast<const uint8_t *>(src), bytes);
	return hip_check(ctx, e, "copy");
}

} // namespace svo
