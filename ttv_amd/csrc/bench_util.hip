// Diagnostic stream kernels used by the benchmark: read streams and read streams with a write
// fraction (the practical HBM ceilings the product kernels are compared with, taken as the best of
// two access patterns) and a plain copy (a known-size kernel that calibrates the FETCH_SIZE /
// WRITE_SIZE counters).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/ttvx.h"
#include "runtime.hpp"

namespace ttvx
{

// Each block streams one contiguous chunk with non-temporal 16-byte loads (the fastest read
// pattern measured on MI355X), sums it and writes one float, so the loads cannot be dropped.
__global__ __launch_bounds__(kBlock) void k_read_stream(const u32x4* __restrict__ src, int64_t n16, int64_t chunk,
                                                        float* __restrict__ out)
{
	constexpr int U = 8;
	float s = 0.f;
	int64_t const begin = static_cast<int64_t>(blockIdx.x) * chunk;
	int64_t const end = begin + chunk < n16 ? begin + chunk : n16;
	int64_t i = begin + threadIdx.x;
	for(; i + (U - 1) * kBlock < end; i += U * kBlock) {
		u32x4 v[U];
#pragma unroll
		for(int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(src + i + u * kBlock);
#pragma unroll
		for(int u = 0; u < U; ++u) s += __uint_as_float(v[u].x) + __uint_as_float(v[u].y) + __uint_as_float(v[u].z) + __uint_as_float(v[u].w);
	}
	for(; i < end; i += kBlock) {
		u32x4 v = __builtin_nontemporal_load(src + i);
		s += __uint_as_float(v.x) + __uint_as_float(v.y) + __uint_as_float(v.z) + __uint_as_float(v.w);
	}
	for(int d = 32; d > 0; d >>= 1) s += __shfl_xor(s, d, 64);
	__shared__ float ws[kBlock / 64];
	if((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
	__syncthreads();
	if(threadIdx.x == 0) out[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// Read stream plus a write stream of a fixed fraction, interleaved like a product's: block b streams
// its contiguous chunk of src in rounds of U 16-byte loads per thread and, after each round, writes
// the round's share of its contiguous chunk of dst with non-temporal 16-byte stores.
__global__ __launch_bounds__(kBlock) void k_mix_stream(const u32x4* __restrict__ src, int64_t n16, int64_t chunk,
                                                       u32x4* __restrict__ dst, int64_t m16, int64_t dchunk)
{
	constexpr int U = 8;
	float s = 0.f;
	int64_t const begin = static_cast<int64_t>(blockIdx.x) * chunk;
	int64_t const end = begin + chunk < n16 ? begin + chunk : n16;
	int64_t const dbegin = static_cast<int64_t>(blockIdx.x) * dchunk;
	int64_t const dend = dbegin + dchunk < m16 ? dbegin + dchunk : m16;
	int64_t const rounds = (end - begin + U * kBlock - 1) / (U * kBlock);
	int64_t const dper = rounds > 0 ? (dend - dbegin + rounds - 1) / rounds : 0;    // stores per round
	int64_t d = dbegin;
	for(int64_t i = begin + threadIdx.x; i - threadIdx.x < end; i += U * kBlock) {
		u32x4 v[U];
#pragma unroll
		for(int u = 0; u < U; ++u) {
			int64_t const j = i + u * kBlock < end ? i + u * kBlock : end - 1;
			v[u] = __builtin_nontemporal_load(src + j);
		}
#pragma unroll
		for(int u = 0; u < U; ++u) s += __uint_as_float(v[u].x) + __uint_as_float(v[u].y) + __uint_as_float(v[u].z) + __uint_as_float(v[u].w);
		int64_t const de = d + dper < dend ? d + dper : dend;
		u32x4 w;
		w.x = w.y = w.z = w.w = __float_as_uint(s);
		for(int64_t j = d + threadIdx.x; j < de; j += kBlock) __builtin_nontemporal_store(w, dst + j);
		d = de;
	}
}

// Compact sweep: block b reads the U*kBlock 16-byte vectors [b*U*kBlock, (b+1)*U*kBlock) of src,
// all loads issued at once, writes its share dst[b*dchunk, (b+1)*dchunk) of the write stream and
// exits; the dispatcher hands out blocks in address order, so the loads in flight form one compact
// front moving through src (the pattern of the one-unit-per-wave product kernels).  Without a
// write stream the sum goes to sink only when it equals a value it cannot take, which keeps the
// loads alive.
template<int U>
__global__ __launch_bounds__(kBlock) void k_sweep_stream(const u32x4* __restrict__ src, int64_t n16,
                                                         u32x4* __restrict__ dst, int64_t m16, int64_t dchunk,
                                                         float* __restrict__ sink)
{
	int64_t const i0 = static_cast<int64_t>(blockIdx.x) * (U * kBlock) + threadIdx.x;
	u32x4 v[U];
#pragma unroll
	for(int u = 0; u < U; ++u) {
		int64_t const j = i0 + u * kBlock;
		v[u] = __builtin_nontemporal_load(src + (j < n16 ? j : n16 - 1));
	}
	float s = 0.f;
#pragma unroll
	for(int u = 0; u < U; ++u) s += __uint_as_float(v[u].x) + __uint_as_float(v[u].y) + __uint_as_float(v[u].z) + __uint_as_float(v[u].w);
	int64_t const d0 = static_cast<int64_t>(blockIdx.x) * dchunk;
	int64_t const de = d0 + dchunk < m16 ? d0 + dchunk : m16;
	u32x4 w;
	w.x = w.y = w.z = w.w = __float_as_uint(s);
	for(int64_t j = d0 + threadIdx.x; j < de; j += kBlock) __builtin_nontemporal_store(w, dst + j);
	if(m16 == 0 && s == -1.0f) sink[0] = s;     // inputs are non-negative bit patterns of the benchmark
}

__global__ __launch_bounds__(kBlock) void k_copy_stream(const u32x4* __restrict__ src, u32x4* __restrict__ dst, int64_t n16)
{
	int64_t const stride = static_cast<int64_t>(gridDim.x) * kBlock;
	for(int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

} // namespace ttvx

extern "C" {

int ttvx_bench_read(const void* src, size_t bytes, float* partial, size_t partial_count, void* stream)
{
	using namespace ttvx;
	clear_error();
	DeviceInfo di;
	if(!current_device(di)) return TTVX_ERR_DEVICE;
	if(bytes % 16 != 0 || reinterpret_cast<uintptr_t>(src) % 16 != 0)
		return fail(TTVX_ERR_ARGUMENT, "ttvx_bench_read: 16-byte aligned buffer and size required");
	int rc = check_device_span(src, bytes, "src");
	if(rc == TTVX_OK) rc = check_device_span(partial, partial_count * 4, "partial");
	if(rc != TTVX_OK) return rc;
	unsigned const grid = static_cast<unsigned>(std::min<size_t>(partial_count, static_cast<size_t>(di.cus) * 64));
	if(grid == 0) return fail(TTVX_ERR_ARGUMENT, "ttvx_bench_read: partial_count must be positive");
	int64_t const n16 = static_cast<int64_t>(bytes / 16);
	int64_t const chunk = (n16 + grid - 1) / grid;
	hipLaunchKernelGGL(k_read_stream, dim3(grid), dim3(kBlock), 0, static_cast<hipStream_t>(stream),
	                   static_cast<const u32x4*>(src), n16, chunk, partial);
	hipError_t e = hipGetLastError();
	return e == hipSuccess ? TTVX_OK : hip_fail(e, "read stream launch");
}

int ttvx_bench_mix(const void* src, size_t src_bytes, void* dst, size_t dst_bytes, void* stream)
{
	using namespace ttvx;
	clear_error();
	DeviceInfo di;
	if(!current_device(di)) return TTVX_ERR_DEVICE;
	if(src_bytes % 16 != 0 || dst_bytes % 16 != 0 || reinterpret_cast<uintptr_t>(src) % 16 != 0 ||
	   reinterpret_cast<uintptr_t>(dst) % 16 != 0 || src_bytes == 0)
		return fail(TTVX_ERR_ARGUMENT, "ttvx_bench_mix: 16-byte aligned buffers and sizes required");
	int rc = check_device_span(src, src_bytes, "src");
	if(rc == TTVX_OK && dst_bytes > 0) rc = check_device_span(dst, dst_bytes, "dst");
	if(rc != TTVX_OK) return rc;
	// the read-stream grid (32 blocks per CU, 128 KiB chunks for 1 GiB)
	int64_t const n16 = static_cast<int64_t>(src_bytes / 16), m16 = static_cast<int64_t>(dst_bytes / 16);
	int64_t const grid = std::max<int64_t>(1, std::min<int64_t>(static_cast<int64_t>(di.cus) * 32, n16 / 256));
	int64_t const chunk = (n16 + grid - 1) / grid, dchunk = (m16 + grid - 1) / grid;
	hipLaunchKernelGGL(k_mix_stream, dim3(static_cast<unsigned>(grid)), dim3(kBlock), 0, static_cast<hipStream_t>(stream),
	                   static_cast<const u32x4*>(src), n16, chunk, static_cast<u32x4*>(dst), m16, dchunk);
	hipError_t e = hipGetLastError();
	return e == hipSuccess ? TTVX_OK : hip_fail(e, "mix stream launch");
}

int ttvx_bench_sweep(const void* src, size_t src_bytes, void* dst, size_t dst_bytes, int loads_per_thread,
                     float* sink, void* stream)
{
	using namespace ttvx;
	clear_error();
	DeviceInfo di;
	if(!current_device(di)) return TTVX_ERR_DEVICE;
	if(src_bytes % 16 != 0 || dst_bytes % 16 != 0 || reinterpret_cast<uintptr_t>(src) % 16 != 0 ||
	   reinterpret_cast<uintptr_t>(dst) % 16 != 0 || src_bytes == 0 || sink == nullptr)
		return fail(TTVX_ERR_ARGUMENT, "ttvx_bench_sweep: 16-byte aligned buffers and sizes and a sink required");
	if(loads_per_thread != 2 && loads_per_thread != 4 && loads_per_thread != 8)
		return fail(TTVX_ERR_ARGUMENT, "ttvx_bench_sweep: loads_per_thread must be 2, 4 or 8");
	int rc = check_device_span(src, src_bytes, "src");
	if(rc == TTVX_OK && dst_bytes > 0) rc = check_device_span(dst, dst_bytes, "dst");
	if(rc == TTVX_OK) rc = check_device_span(sink, 4, "sink");
	if(rc != TTVX_OK) return rc;
	int64_t const n16 = static_cast<int64_t>(src_bytes / 16), m16 = static_cast<int64_t>(dst_bytes / 16);
	int64_t const per = static_cast<int64_t>(loads_per_thread) * kBlock;
	int64_t const grid = (n16 + per - 1) / per;
	if(grid > kMaxLeanBlocks) return fail(TTVX_ERR_ARGUMENT, "ttvx_bench_sweep: source too large for one launch");
	int64_t const dchunk = (m16 + grid - 1) / grid;
	auto* s = static_cast<const u32x4*>(src);
	auto* d = static_cast<u32x4*>(dst);
	hipStream_t st = static_cast<hipStream_t>(stream);
	dim3 const g(static_cast<unsigned>(grid));
	if(loads_per_thread == 2) hipLaunchKernelGGL(k_sweep_stream<2>, g, dim3(kBlock), 0, st, s, n16, d, m16, dchunk, sink);
	else if(loads_per_thread == 4) hipLaunchKernelGGL(k_sweep_stream<4>, g, dim3(kBlock), 0, st, s, n16, d, m16, dchunk, sink);
	else hipLaunchKernelGGL(k_sweep_stream<8>, g, dim3(kBlock), 0, st, s, n16, d, m16, dchunk, sink);
	hipError_t e = hipGetLastError();
	return e == hipSuccess ? TTVX_OK : hip_fail(e, "sweep stream launch");
}

int ttvx_bench_copy(const void* src, void* dst, size_t bytes, void* stream)
{
	using namespace ttvx;
	clear_error();
	DeviceInfo di;
	if(!current_device(di)) return TTVX_ERR_DEVICE;
	if(bytes % 16 != 0 || reinterpret_cast<uintptr_t>(src) % 16 != 0 || reinterpret_cast<uintptr_t>(dst) % 16 != 0)
		return fail(TTVX_ERR_ARGUMENT, "ttvx_bench_copy: 16-byte aligned buffers and size required");
	int rc = check_device_span(src, bytes, "src");
	if(rc == TTVX_OK) rc = check_device_span(dst, bytes, "dst");
	if(rc != TTVX_OK) return rc;
	hipLaunchKernelGGL(k_copy_stream, dim3(static_cast<unsigned>(di.cus) * 8), dim3(kBlock), 0,
	                   static_cast<hipStream_t>(stream), static_cast<const u32x4*>(src), static_cast<u32x4*>(dst),
	                   static_cast<int64_t>(bytes / 16));
	hipError_t e = hipGetLastError();
	return e == hipSuccess ? TTVX_OK : hip_fail(e, "copy stream launch");
}

} // extern "C"
