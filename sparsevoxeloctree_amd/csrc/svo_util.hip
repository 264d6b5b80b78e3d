// Plain device-to-device copy on the context's stream (grid-stride, 16-byte vectors when aligned).
#include "svo_runtime.hpp"

namespace svo {
namespace {

constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads) copy_u128(uint4 *__restrict__ dst, const uint4 *__restrict__ src, uint64_t n) {
	const uint64_t stride = (uint64_t)gridDim.x * kThreads;
	for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride)
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
	if (((uintptr_t)dst | (uintptr_t)src | bytes) % 16 == 0)
		e = launch(ctx, "copy_u128", copy_u128, dim3(grid), dim3(kThreads), 0, reinterpret_cast<uint4 *>(dst),
		           reinterpret_cast<const uint4 *>(src), bytes / 16);
	else
		e = launch(ctx, "copy_u8", copy_u8, dim3(grid), dim3(kThreads), 0, reinterpret_cast<uint8_t *>(dst),
		           reinterpret_cast<const uint8_t *>(src), bytes);
	return hip_check(ctx, e, "copy");
}

} // namespace svo
