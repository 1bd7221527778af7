// Quantisation kernels for reduced-precision storage of A (see quant.hpp for the formats).
#include "quant.hpp"

#include <algorithm>

#include "../../include/ttvx.h"
#include "runtime.hpp"

namespace ttvx
{

__global__ __launch_bounds__(kBlock) void k_quant_bf16(const float* __restrict__ src, uint16_t* __restrict__ dst, int64_t n)
{
	for(int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
	    i += static_cast<int64_t>(gridDim.x) * kBlock)
		dst[i] = f32_to_bf16_rne(src[i]);
}

// One block per scale group: absmax over the group's fibers, then encode.
// Element e of group g: fiber j = g*fps + e % nf, k = e / nf (consecutive threads walk
// consecutive fibers, i.e. consecutive i for a fixed k, so loads are coalesced when sAi == 1).
__global__ __launch_bounds__(kBlock) void k_quant_fp8(const float* __restrict__ A, uint8_t* __restrict__ Q,
                                                      float* __restrict__ scales, QuantArgs P)
{
	__shared__ float wmax[kBlock / 64];
	int64_t const fibers = P.O * P.I;
	for(int64_t g = blockIdx.x; g < P.groups; g += gridDim.x) {
		int64_t const j0 = g * P.fps;
		int64_t const nf = std::min(P.fps, fibers - j0);
		int64_t const total = nf * P.K;
		float m = 0.f;
		for(int64_t e = threadIdx.x; e < total; e += kBlock) {
			int64_t const k = e / nf, j = j0 + (e - k * nf);
			int64_t const o = j / P.I, i = j - o * P.I;
			m = fmaxf(m, fabsf(A[o * P.sAo + k * P.sAk + i * P.sAi]));
		}
		for(int d = 32; d > 0; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, 64));
		if((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
		__syncthreads();
		float amax = wmax[0];
		for(int w = 1; w < kBlock / 64; ++w) amax = fmaxf(amax, wmax[w]);
		__syncthreads();
		float const scale = amax > 0.f ? amax / 448.0f : 1.0f;
		float const inv   = amax > 0.f ? 448.0f / amax : 1.0f;
		if(threadIdx.x == 0) scales[g] = scale;
		for(int64_t e = threadIdx.x; e < total; e += kBlock) {
			int64_t const k = e / nf, j = j0 + (e - k * nf);
			int64_t const o = j / P.I, i = j - o * P.I;
			int64_t const off = o * P.sAo + k * P.sAk + i * P.sAi;
			Q[off] = f32_to_e4m3_rne_sat(A[off] * inv);
		}
	}
}

int quantize_bf16(std::size_t count, const float* src, uint16_t* dst, hipStream_t stream)
{
	if(count == 0) return TTVX_OK;
	int64_t const n = static_cast<int64_t>(count);
	unsigned const grid = static_cast<unsigned>(std::min<int64_t>((n + kBlock - 1) / kBlock, 8192));
	hipLaunchKernelGGL(k_quant_bf16, dim3(grid), dim3(kBlock), 0, stream, src, dst, n);
	hipError_t e = hipGetLastError();
	return e == hipSuccess ? TTVX_OK : hip_fail(e, "bf16 quantisation launch");
}

int quantize_fp8(const Fold& f, const float* a, uint8_t* aq, float* scales, int64_t fps, hipStream_t stream)
{
	QuantArgs P;
	P.O = f.O; P.K = f.K; P.I = f.I;
	P.sAo = f.sAo; P.sAk = f.sAk; P.sAi = f.sAi;
	P.fps = fps;
	P.groups = (f.O * f.I + fps - 1) / fps;
	unsigned const grid = static_cast<unsigned>(std::min<int64_t>(P.groups, 65536));
	hipLaunchKernelGGL(k_quant_fp8, dim3(grid), dim3(kBlock), 0, stream, a, aq, scales, P);
	hipError_t e = hipGetLastError();
	return e == hipSuccess ? TTVX_OK : hip_fail(e, "fp8 quantisation launch");
}

} // namespace ttvx
