// Weights of a fused multi-vector contraction (ttvs: every mode but q contracted with its vector).
// With A dense in its layout, A[o, j, i] (i: the modes stored before q, j: mode q, o: the modes
// after q) contracts as c[j] = sum_o wO[o] * sum_i A[o, j, i] * wI[i], where wI and wO are the
// Kronecker products of the vectors of the inner and of the outer modes.  The two sums are one row
// product and one column product of the library, whatever the order of A, instead of the p - 1
// products (and p - 2 materialised intermediates) of the reference's chains
// (ttvpy/src/wrapped_ttv.cpp:141-194).  This file builds wI and wO (no counterpart in the reference).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/ttvx.h"
#include "runtime.hpp"

namespace ttvx
{

constexpr int kMaxKron = 32;

template<typename T>
struct KronArgs {
	int count;
	int64_t n[kMaxKron];
	const T* v[kMaxKron];
};

// out[i] = prod_m v_m[i_m], i = i_0 + n_0 * (i_1 + n_1 * (...)): vector 0 varies fastest
template<typename T>
__global__ __launch_bounds__(kBlock) void k_kron(KronArgs<T> a, T* __restrict__ out, int64_t total)
{
	int64_t const stride = static_cast<int64_t>(gridDim.x) * kBlock;
	for(int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < total; i += stride) {
		int64_t rest = i;
		T prod = T(1);
		for(int m = 0; m < a.count; ++m) {
			int64_t const im = rest % a.n[m];
			rest /= a.n[m];
			prod *= a.v[m][im];
		}
		out[i] = prod;
	}
}

namespace
{
template<typename T>
int kron(const T* const* vecs, const size_t* lens, size_t count, T* out, void* stream, const char* name)
{
	clear_error();
	if(count == 0 || count > static_cast<size_t>(kMaxKron) || vecs == nullptr || lens == nullptr || out == nullptr)
		return fail(TTVX_ERR_ARGUMENT, std::string(name) + ": 1 to 32 vectors, their lengths and an output required");
	DeviceInfo di;
	if(!current_device(di)) return TTVX_ERR_DEVICE;
	KronArgs<T> a{};
	a.count = static_cast<int>(count);
	int64_t total = 1;
	for(size_t m = 0; m < count; ++m) {
		if(lens[m] == 0 || vecs[m] == nullptr) return fail(TTVX_ERR_ARGUMENT, std::string(name) + ": empty or null vector");
		int rc = check_device_span(vecs[m], lens[m] * sizeof(T), "vector");
		if(rc != TTVX_OK) return rc;
		a.n[m] = static_cast<int64_t>(lens[m]);
		a.v[m] = vecs[m];
		if(total > (int64_t{1} << 40) / a.n[m]) return fail(TTVX_ERR_ARGUMENT, std::string(name) + ": product too long");
		total *= a.n[m];
	}
	int rc = check_device_span(out, static_cast<size_t>(total) * sizeof(T), "out");
	if(rc != TTVX_OK) return rc;
	unsigned const grid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((total + kBlock - 1) / kBlock,
	                                                                                    static_cast<int64_t>(di.cus) * 8)));
	hipLaunchKernelGGL((k_kron<T>), dim3(grid), dim3(kBlock), 0, static_cast<hipStream_t>(stream), a, out, total);
	hipError_t const e = hipGetLastError();
	return e == hipSuccess ? TTVX_OK : hip_fail(e, name);
}
} // namespace

} // namespace ttvx

extern "C" {

int ttvx_kron_f32(const float* const* vectors, const size_t* lengths, size_t count, float* out, void* stream)
{
	return ttvx::kron<float>(vectors, lengths, count, out, stream, "ttvx_kron_f32");
}

int ttvx_kron_f64(const double* const* vectors, const size_t* lengths, size_t count, double* out, void* stream)
{
	return ttvx::kron<double>(vectors, lengths, count, out, stream, "ttvx_kron_f64");
}

} // extern "C"
