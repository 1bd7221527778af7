// extern "C" entry points of libttvx (declared in include/ttvx.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ttvx.h"
#include "plan.hpp"
#include "runtime.hpp"

namespace ttvx
{
const char* last_error_cstr();
int set_selection(int mode, int forced);
int quantize_bf16(std::size_t count, const float* src, uint16_t* dst, hipStream_t stream);
int quantize_fp8(const Fold& f, const float* a, uint8_t* aq, float* scales, int64_t fps, hipStream_t stream);
} // namespace ttvx

using namespace ttvx;

namespace
{

bool is_device_pointer(const void* p)
{
	hipPointerAttribute_t attr;
	std::memset(&attr, 0, sizeof(attr));
	hipError_t e = hipPointerGetAttributes(&attr, p);
	if(e != hipSuccess) {
		(void)hipGetLastError();
		return false;
	}
	return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

// Device staging of the synchronous entry points: one grow-only buffer per operand and device,
// held until the process ends.  Calls with host operands on one device are serialised by the
// cache's mutex (they block until C is back on the host anyway).
struct StageCache {
	std::mutex mutex;
	void* buf[4] = {nullptr, nullptr, nullptr, nullptr};     // A, b, C, product before accumulation
	std::size_t cap[4] = {0, 0, 0, 0};
};
StageCache g_stage[64];

struct Staged {
	void* dev = nullptr;
	bool staged = false;
};

// Device buffer of slot `slot` holding at least `bytes` (grown when too small).
int stage_buffer(StageCache& cache, int slot, std::size_t bytes, void*& dev)
{
	if(injected(TTVX_FAULT_ALLOC)) return hip_fail(hipErrorOutOfMemory, "hipMalloc (staging)");
	if(cache.cap[slot] < bytes) {
		if(cache.buf[slot] != nullptr) (void)hipFree(cache.buf[slot]);
		cache.buf[slot] = nullptr;
		cache.cap[slot] = 0;
		std::size_t const want = std::max<std::size_t>(bytes, 1 << 20);
		hipError_t e = hipMalloc(&cache.buf[slot], want);
		if(e != hipSuccess) return hip_fail(e, "hipMalloc (staging)");
		cache.cap[slot] = want;
	}
	dev = cache.buf[slot];
	return TTVX_OK;
}

int stage_in(StageCache& cache, int slot, const void* host_or_dev, std::size_t bytes, bool copy, Staged& s,
             const char* what)
{
	if(is_device_pointer(host_or_dev)) {
		int rc = check_device_span(host_or_dev, bytes, what);
		if(rc != TTVX_OK) return rc;
		s.dev = const_cast<void*>(host_or_dev);
		return TTVX_OK;
	}
	int const rc = stage_buffer(cache, slot, bytes, s.dev);
	if(rc != TTVX_OK) return rc;
	s.staged = true;
	if(copy) {
		hipError_t e = injected(TTVX_FAULT_COPY) ? hipErrorInvalidValue
		                                         : hipMemcpy(s.dev, host_or_dev, bytes, hipMemcpyHostToDevice);
		if(e != hipSuccess) return hip_fail(e, "hipMemcpy (host to device)");
	}
	return TTVX_OK;
}

// c[j] += t[j], j < n
template<typename T>
__global__ __launch_bounds__(kBlock) void k_accumulate(T* __restrict__ c, const T* __restrict__ t, int64_t n)
{
	for(int64_t j = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; j < n; j += static_cast<int64_t>(gridDim.x) * kBlock)
		c[j] += t[j];
}

// Synchronous product with host or device operands.  accumulate: C += A x_q b (C contiguous), the
// semantics of the reference's non-BLAS gemv_col (matrix_times_vector.h:117-126).
template<typename T>
int ttv_sync(int code, std::size_t q, std::size_t p,
             const T* a, const std::size_t* na, const std::size_t* wa, const std::size_t* pia,
             const T* b, const std::size_t* nb,
             T* c, const std::size_t* nc, const std::size_t* wc, const std::size_t* pic, bool accumulate = false)
{
	clear_error();
	Check ck = validate(q, p, a, na, wa, pia, b, nb, c, nc, wc, pic);
	if(ck.status != TTVX_OK) return fail(ck.status, ck.message);
	DeviceInfo di;
	if(!current_device(di)) return TTVX_ERR_DEVICE;
	Fold const f = fold(q, p, na, wa, pia, wc, pic);
	std::size_t const ab = static_cast<std::size_t>(f.a_span) * sizeof(T);
	std::size_t const bb = static_cast<std::size_t>(f.K) * sizeof(T);
	std::size_t const cb = static_cast<std::size_t>(f.c_span) * sizeof(T);
	StageCache& cache = g_stage[di.id & 63];
	std::lock_guard<std::mutex> lock(cache.mutex);
	Staged sa, sb, sc;
	int rc = stage_in(cache, 0, a, ab, true, sa, "A");
	if(rc == TTVX_OK) rc = stage_in(cache, 1, b, bb, true, sb, "b");
	// a strided C with gaps keeps the gap contents, an accumulated C its values: copy it in first
	if(rc == TTVX_OK) rc = stage_in(cache, 2, c, cb, accumulate || f.c_span > f.outputs, sc, "C");
	if(rc != TTVX_OK) return rc;
	if(accumulate) {
		if(f.c_span != f.outputs) return fail(TTVX_ERR_UNSUPPORTED, "ttvx: accumulation needs a contiguous C");
		// the product goes to a device temporary (slot 3), then C += product
		void* prod = nullptr;
		rc = stage_buffer(cache, 3, cb, prod);
		if(rc != TTVX_OK) return rc;
		rc = launch(f, dtype_of(code), sa.dev, sb.dev, prod, nullptr, 1, 0u, -1, nullptr);
		if(rc != TTVX_OK) return rc;
		unsigned const grid = static_cast<unsigned>(std::min<int64_t>(4096, (f.outputs + kBlock - 1) / kBlock));
		hipLaunchKernelGGL((k_accumulate<T>), dim3(std::max(grid, 1u)), dim3(kBlock), 0, nullptr, static_cast<T*>(sc.dev),
		                   static_cast<const T*>(prod), f.outputs);
		hipError_t const e = hipGetLastError();
		if(e != hipSuccess) return hip_fail(e, "accumulate launch");
	} else {
		rc = launch(f, dtype_of(code), sa.dev, sb.dev, sc.dev, nullptr, 1, 0u, -1, nullptr);
		if(rc != TTVX_OK) return rc;
	}
	hipError_t e;
	if(sc.staged) e = injected(TTVX_FAULT_COPY) ? hipErrorInvalidValue : hipMemcpy(c, sc.dev, cb, hipMemcpyDeviceToHost);
	else         e = hipStreamSynchronize(nullptr);
	if(e != hipSuccess) return hip_fail(e, "completion of the product");
	return TTVX_OK;
}

template<typename S>
int ttv_async(int code, std::size_t q, std::size_t p,
              const S* a, const std::size_t* na, const std::size_t* wa, const std::size_t* pia,
              const float* scales, std::size_t fps,
              const void* b, const std::size_t* nb,
              void* c, const std::size_t* nc, const std::size_t* wc, const std::size_t* pic,
              unsigned flags, void* stream, const Scratch* scratch = nullptr)
{
	clear_error();
	Check ck = validate(q, p, a, na, wa, pia, b, nb, c, nc, wc, pic);
	if(ck.status != TTVX_OK) return fail(ck.status, ck.message);
	DeviceInfo di;
	if(!current_device(di)) return TTVX_ERR_DEVICE;
	Fold const f = fold(q, p, na, wa, pia, wc, pic);
	DType const dt = dtype_of(code);
	std::size_t const acc = code == TTVX_F64 ? 8 : 4;
	int rc = check_device_span(a, static_cast<std::size_t>(f.a_span) * sizeof(S), "A");
	if(rc == TTVX_OK) rc = check_device_span(b, static_cast<std::size_t>(f.K) * acc, "b");
	if(rc == TTVX_OK) rc = check_device_span(c, static_cast<std::size_t>(f.c_span) * acc, "C");
	if(rc != TTVX_OK) return rc;
	if(code == TTVX_FP8E4M3) {
		if(scales == nullptr || fps == 0) return fail(TTVX_ERR_ARGUMENT, "ttvx: FP8 storage needs scales and fibers_per_scale > 0");
		std::size_t const ns = (static_cast<std::size_t>(f.outputs) + fps - 1) / fps;
		rc = check_device_span(scales, ns * sizeof(float), "scales");
		if(rc != TTVX_OK) return rc;
	}
	return launch(f, dt, a, b, c, scales, static_cast<int64_t>(fps), flags, -1, static_cast<hipStream_t>(stream), scratch);
}

// The 2-D view A(j, k) = a[j*wj + k*wk] as an order-2 tensor contracted over mode 2; its layout
// puts the mode with the smaller stride first (mode 1 on ties).
template<typename T>
int mtv_sync(int code, const T* a, const T* b, T* c, size_t m, size_t n, size_t wj, size_t wk, int accumulate)
{
	std::size_t const na[2] = {m, n};
	std::size_t const wa[2] = {wj, wk};
	std::size_t const pia[2] = {wj <= wk ? std::size_t{1} : std::size_t{2}, wj <= wk ? std::size_t{2} : std::size_t{1}};
	std::size_t const nb[1] = {n};
	std::size_t const nc[1] = {m};
	std::size_t const wc[1] = {1};
	std::size_t const pic[1] = {1};
	return ttv_sync<T>(code, 2, 2, a, na, wa, pia, b, nb, c, nc, wc, pic, accumulate != 0);
}
// Split partial sums of ttvx_ttv_f32_async_noalloc: a static array that every GPU the library runs
// on holds until the process ends.  Host code never passes `scratch` itself (its host-side name is
// not a device address): it uses the address hipGetSymbolAddress returns for the current device.
constexpr std::size_t kNoAllocScratchBytes = std::size_t{16} << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mutex;
void* g_scratch_ptr[64] = {};

// The current device's copy of `scratch` (looked up once per device, with that device current).
int device_scratch(int device, Scratch& out)
{
	std::lock_guard<std::mutex> lock(g_scratch_mutex);
	void*& ptr = g_scratch_ptr[device & 63];
	if(ptr == nullptr) {
		void* p = nullptr;
		hipError_t const e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if(e != hipSuccess || p == nullptr) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress (scratch)");
		ptr = p;
	}
	out.ptr = ptr;
	out.bytes = kNoAllocScratchBytes;
	return TTVX_OK;
}
} // namespace

extern "C" {

const char* ttvx_last_error(void) { return last_error_cstr(); }

int ttvx_version(void) { return 10000; }

int ttvx_device_count(void)
{
	int n = 0;
	if(hipGetDeviceCount(&n) != hipSuccess) {
		(void)hipGetLastError();
		return 0;
	}
	return n;
}

int ttvx_device_alloc(size_t bytes, void** out)
{
	clear_error();
	if(out == nullptr) return fail(TTVX_ERR_ARGUMENT, "ttvx_device_alloc: null output pointer");
	*out = nullptr;
	DeviceInfo di;
	if(!current_device(di)) return TTVX_ERR_DEVICE;
	if(bytes == 0) return TTVX_OK;
	hipError_t const e = injected(TTVX_FAULT_ALLOC) ? hipErrorOutOfMemory : hipMalloc(out, bytes);
	if(e != hipSuccess) {
		*out = nullptr;
		return hip_fail(e, "hipMalloc");
	}
	return TTVX_OK;
}

int ttvx_device_free(void* ptr)
{
	clear_error();
	if(ptr == nullptr) return TTVX_OK;
	hipError_t const e = hipFree(ptr);
	return e == hipSuccess ? TTVX_OK : hip_fail(e, "hipFree");
}

int ttvx_copy(void* dst, const void* src, size_t bytes)
{
	clear_error();
	if(bytes == 0) return TTVX_OK;
	if(dst == nullptr || src == nullptr) return fail(TTVX_ERR_ARGUMENT, "ttvx_copy: null pointer");
	hipError_t const e = hipMemcpy(dst, src, bytes, hipMemcpyDefault);
	return e == hipSuccess ? TTVX_OK : hip_fail(e, "hipMemcpy");
}

int ttvx_ttv_f32(int ep, int sp, int fp, size_t q, size_t p,
                 const float* a, const size_t* na, const size_t* wa, const size_t* pia,
                 const float* b, const size_t* nb,
                 float* c, const size_t* nc, const size_t* wc, const size_t* pic)
{
	(void)ep; (void)sp; (void)fp;   // policy tags are hints; the planner picks the kernel
	return ttv_sync<float>(TTVX_F32, q, p, a, na, wa, pia, b, nb, c, nc, wc, pic);
}

int ttvx_ttv_f64(int ep, int sp, int fp, size_t q, size_t p,
                 const double* a, const size_t* na, const size_t* wa, const size_t* pia,
                 const double* b, const size_t* nb,
                 double* c, const size_t* nc, const size_t* wc, const size_t* pic)
{
	(void)ep; (void)sp; (void)fp;
	return ttv_sync<double>(TTVX_F64, q, p, a, na, wa, pia, b, nb, c, nc, wc, pic);
}

int ttvx_mtv_f32(const float* a, const float* b, float* c, size_t m, size_t n, size_t wj, size_t wk, int accumulate)
{
	return mtv_sync<float>(TTVX_F32, a, b, c, m, n, wj, wk, accumulate);
}

int ttvx_mtv_f64(const double* a, const double* b, double* c, size_t m, size_t n, size_t wj, size_t wk, int accumulate)
{
	return mtv_sync<double>(TTVX_F64, a, b, c, m, n, wj, wk, accumulate);
}

int ttvx_ttv_f32_async(size_t q, size_t p,
                       const float* a, const size_t* na, const size_t* wa, const size_t* pia,
                       const float* b, const size_t* nb,
                       float* c, const size_t* nc, const size_t* wc, const size_t* pic,
                       unsigned flags, void* stream)
{
	return ttv_async<float>(TTVX_F32, q, p, a, na, wa, pia, nullptr, 0, b, nb, c, nc, wc, pic, flags, stream);
}

int ttvx_ttv_f32_async_noalloc(size_t q, size_t p,
                               const float* a, const size_t* na, const size_t* wa, const size_t* pia,
                               const float* b, const size_t* nb,
                               float* c, const size_t* nc, const size_t* wc, const size_t* pic,
                               unsigned flags, void* stream)
{
	clear_error();
	DeviceInfo di;
	if(!current_device(di)) return TTVX_ERR_DEVICE;
	Scratch sc;
	int const rc = device_scratch(di.id, sc);
	if(rc != TTVX_OK) return rc;
	return ttv_async<float>(TTVX_F32, q, p, a, na, wa, pia, nullptr, 0, b, nb, c, nc, wc, pic, flags, stream, &sc);
}

int ttvx_ttv_f64_async(size_t q, size_t p,
                       const double* a, const size_t* na, const size_t* wa, const size_t* pia,
                       const double* b, const size_t* nb,
                       double* c, const size_t* nc, const size_t* wc, const size_t* pic,
                       unsigned flags, void* stream)
{
	return ttv_async<double>(TTVX_F64, q, p, a, na, wa, pia, nullptr, 0, b, nb, c, nc, wc, pic, flags, stream);
}

int ttvx_quantize_bf16_async(size_t count, const float* src, uint16_t* dst, void* stream)
{
	clear_error();
	if(count == 0) return TTVX_OK;
	if(src == nullptr || dst == nullptr) return fail(TTVX_ERR_ARGUMENT, "ttvx: null pointer passed to ttvx_quantize_bf16_async");
	DeviceInfo di;
	if(!current_device(di)) return TTVX_ERR_DEVICE;
	int rc = check_device_span(src, count * 4, "src");
	if(rc == TTVX_OK) rc = check_device_span(dst, count * 2, "dst");
	if(rc != TTVX_OK) return rc;
	return quantize_bf16(count, src, dst, static_cast<hipStream_t>(stream));
}

size_t ttvx_fp8_num_scales(size_t q, size_t p, const size_t* na, const size_t* wa, const size_t* pia,
                           size_t fibers_per_scale)
{
	clear_error();
	Check ck = validate_a(q, p, na, wa, pia);
	if(ck.status != TTVX_OK) { set_error(ck.message); return 0; }
	if(fibers_per_scale == 0) { set_error("ttvx: fibers_per_scale must be positive"); return 0; }
	Fold const f = fold(q, p, na, wa, pia, nullptr, nullptr);
	return (static_cast<size_t>(f.outputs) + fibers_per_scale - 1) / fibers_per_scale;
}

int ttvx_quantize_fp8_async(size_t q, size_t p,
                            const float* a, const size_t* na, const size_t* wa, const size_t* pia,
                            uint8_t* aq, float* scales, size_t fibers_per_scale, void* stream)
{
	clear_error();
	Check ck = validate_a(q, p, na, wa, pia);
	if(ck.status != TTVX_OK) return fail(ck.status, ck.message);
	if(a == nullptr || aq == nullptr || scales == nullptr)
		return fail(TTVX_ERR_ARGUMENT, "ttvx: null pointer passed to ttvx_quantize_fp8_async");
	if(fibers_per_scale == 0) return fail(TTVX_ERR_ARGUMENT, "ttvx: fibers_per_scale must be positive");
	if(p < 2) return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: shape vector of C is not valid.");
	DeviceInfo di;
	if(!current_device(di)) return TTVX_ERR_DEVICE;
	Fold const f = fold(q, p, na, wa, pia, nullptr, nullptr);
	if(!f.three_d) return fail(TTVX_ERR_UNSUPPORTED, "ttvx: FP8 quantisation needs A to fold to three dimensions");
	std::size_t const ns = (static_cast<std::size_t>(f.outputs) + fibers_per_scale - 1) / fibers_per_scale;
	int rc = check_device_span(a, static_cast<std::size_t>(f.a_span) * 4, "A");
	if(rc == TTVX_OK) rc = check_device_span(aq, static_cast<std::size_t>(f.a_span), "Aq");
	if(rc == TTVX_OK) rc = check_device_span(scales, ns * 4, "scales");
	if(rc != TTVX_OK) return rc;
	return quantize_fp8(f, a, aq, scales, static_cast<int64_t>(fibers_per_scale), static_cast<hipStream_t>(stream));
}

int ttvx_ttv_bf16_async(size_t q, size_t p,
                        const uint16_t* a, const size_t* na, const size_t* wa, const size_t* pia,
                        const float* b, const size_t* nb,
                        float* c, const size_t* nc, const size_t* wc, const size_t* pic,
                        unsigned flags, void* stream)
{
	return ttv_async<uint16_t>(TTVX_BF16, q, p, a, na, wa, pia, nullptr, 0, b, nb, c, nc, wc, pic, flags, stream);
}

int ttvx_ttv_fp8_async(size_t q, size_t p,
                       const uint8_t* a, const size_t* na, const size_t* wa, const size_t* pia,
                       const float* scales, size_t fibers_per_scale,
                       const float* b, const size_t* nb,
                       float* c, const size_t* nc, const size_t* wc, const size_t* pic,
                       unsigned flags, void* stream)
{
	return ttv_async<uint8_t>(TTVX_FP8E4M3, q, p, a, na, wa, pia, scales, fibers_per_scale, b, nb, c, nc, wc, pic,
	                          flags, stream);
}

int ttvx_num_variants(void) { return V_COUNT; }

const char* ttvx_variant_name(int v) { return variant_name(v); }

int ttvx_set_selection(int mode, int forced_variant)
{
	clear_error();
	return set_selection(mode, forced_variant);
}

int ttvx_plan_info(int dtype, size_t q, size_t p, const size_t* na, const size_t* wa, const size_t* pia,
                   int64_t* out6)
{
	clear_error();
	Check ck = validate_a(q, p, na, wa, pia);
	if(ck.status != TTVX_OK) return fail(ck.status, ck.message);
	DType const dt = dtype_of(dtype);
	if(dt.code < 0) return fail(TTVX_ERR_ARGUMENT, "ttvx: unknown element type");
	if(out6 == nullptr) return fail(TTVX_ERR_ARGUMENT, "ttvx: null output array");
	Fold const f = fold(q, p, na, wa, pia, nullptr, nullptr);
	out6[0] = f.O; out6[1] = f.K; out6[2] = f.I;
	out6[3] = f.three_d ? 1 : 0;
	out6[4] = f.ref_case;
	out6[5] = select_variant(f, dt, 256, 0u);
	return TTVX_OK;
}

int ttvx_variant_valid(int dtype, size_t q, size_t p, const size_t* na, const size_t* wa, const size_t* pia, int v)
{
	clear_error();
	Check ck = validate_a(q, p, na, wa, pia);
	if(ck.status != TTVX_OK) { set_error(ck.message); return 0; }
	DType const dt = dtype_of(dtype);
	if(dt.code < 0) return 0;
	Fold const f = fold(q, p, na, wa, pia, nullptr, nullptr);
	Config cfg;
	return configure(v, f, dt, 256, 256, 0u, cfg) ? 1 : 0;
}

int ttvx_features(int dtype, size_t q, size_t p, const size_t* na, const size_t* wa, const size_t* pia,
                  float* features)
{
	clear_error();
	Check ck = validate_a(q, p, na, wa, pia);
	if(ck.status != TTVX_OK) return fail(ck.status, ck.message);
	DType const dt = dtype_of(dtype);
	if(dt.code < 0) return fail(TTVX_ERR_ARGUMENT, "ttvx: unknown element type");
	if(features == nullptr) return fail(TTVX_ERR_ARGUMENT, "ttvx: null output array");
	Fold const f = fold(q, p, na, wa, pia, nullptr, nullptr);
	features_of(f, dt, 256, features);
	return TTVX_OK;
}

int ttvx_debug_fail_next(int site, int count)
{
	clear_error();
	if(site < TTVX_FAULT_ALLOC || site > TTVX_FAULT_LAUNCH || count < 0)
		return fail(TTVX_ERR_ARGUMENT, "ttvx_debug_fail_next: site must be a TTVX_FAULT_* value and count >= 0");
	arm_fault(site, count);
	return TTVX_OK;
}

} // extern "C"
