// GPU kernels of the learned selector (layout in mlp.hpp): batched forward pass, masked-MSE
// loss and gradient, and an Adam step.  Training runs on the GPU through these kernels; the
// planner uses the host forward pass (selector.cpp) with the same parameters.
//
// Work buffer of a batch of B rows (floats): xs[B][F], z1[B][H], z2[B][H], dz1[B][H], dz2[B][H],
// dy[B][V], row_loss[B]  -> ttvx_mlp_work_floats(B).
// All reductions over the batch are in a fixed order (no atomics), so results are reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/ttvx.h"
#include "mlp.hpp"
#include "runtime.hpp"

namespace ttvx
{

namespace
{
struct WorkLayout {
	float *xs, *z1, *z2, *dz1, *dz2, *dy, *loss;
};
WorkLayout layout_of(float* w, int B)
{
	WorkLayout L;
	std::size_t const b = static_cast<std::size_t>(B);
	L.xs   = w;
	L.z1   = L.xs + b * kMlpF;
	L.z2   = L.z1 + b * kMlpH;
	L.dz1  = L.z2 + b * kMlpH;
	L.dz2  = L.dz1 + b * kMlpH;
	L.dy   = L.dz2 + b * kMlpH;
	L.loss = L.dy + b * kMlpV;
	return L;
}
} // namespace

__global__ __launch_bounds__(kBlock) void k_mlp_standardize(const float* __restrict__ P, const float* __restrict__ X, int B,
                                                            float* __restrict__ xs)
{
	int64_t const n = static_cast<int64_t>(B) * kMlpF;
	for(int64_t e = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; e < n; e += static_cast<int64_t>(gridDim.x) * kBlock) {
		int const i = static_cast<int>(e % kMlpF);
		xs[e] = (X[e] - P[kMlpOffMean + i]) * P[kMlpOffStd + i];
	}
}

// Z[b][o] = sum_i act(X[b][i]) W[o][i] + bias[o], act = relu when relu_in
__global__ __launch_bounds__(kBlock) void k_mlp_dense(const float* __restrict__ X, int in, const float* __restrict__ W,
                                                      const float* __restrict__ bias, int out, int B, int relu_in,
                                                      float* __restrict__ Z)
{
	int64_t const n = static_cast<int64_t>(B) * out;
	for(int64_t e = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; e < n; e += static_cast<int64_t>(gridDim.x) * kBlock) {
		int64_t const b = e / out;
		int const o = static_cast<int>(e - b * out);
		const float* x = X + b * in;
		const float* w = W + static_cast<int64_t>(o) * in;
		float s = bias[o];
		for(int i = 0; i < in; ++i) {
			float const xi = relu_in ? fmaxf(x[i], 0.f) : x[i];
			s = __builtin_fmaf(w[i], xi, s);
		}
		Z[e] = s;
	}
}

// dy = M*(y - T)/B, row_loss[b] = 0.5 * sum_v M*(y - T)^2
__global__ __launch_bounds__(kBlock) void k_mlp_dloss(const float* __restrict__ Y, const float* __restrict__ T,
                                                      const float* __restrict__ M, int B, float* __restrict__ dy,
                                                      float* __restrict__ row_loss)
{
	for(int64_t b = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; b < B; b += static_cast<int64_t>(gridDim.x) * kBlock) {
		float l = 0.f;
		for(int v = 0; v < kMlpV; ++v) {
			int64_t const e = b * kMlpV + v;
			float const d = M[e] * (Y[e] - T[e]);
			dy[e] = d / static_cast<float>(B);
			l += 0.5f * d * (Y[e] - T[e]);
		}
		row_loss[b] = l;
	}
}

// gW[o][i] = sum_b dZ[b][o] act(X[b][i]);  gb[o] = sum_b dZ[b][o].  One block per o; thread
// (i, g) sums rows b = g, g + G, ... then the G partial sums are added in a fixed order.
__global__ __launch_bounds__(kBlock) void k_mlp_grad_w(const float* __restrict__ dZ, int out, const float* __restrict__ X,
                                                       int in, int B, int relu_in, float* __restrict__ gW,
                                                       float* __restrict__ gb)
{
	__shared__ float part[kBlock];
	__shared__ float partb[kBlock];
	int const o = blockIdx.x;
	int const G = kBlock / 64;                 // row groups (in <= 64)
	int const i = threadIdx.x % 64;
	int const g = threadIdx.x / 64;
	float s = 0.f, sb = 0.f;
	for(int b = g; b < B; b += G) {
		float const d = dZ[static_cast<int64_t>(b) * out + o];
		if(i < in) {
			float const x = X[static_cast<int64_t>(b) * in + i];
			s = __builtin_fmaf(d, relu_in ? fmaxf(x, 0.f) : x, s);
		}
		sb += d;
	}
	part[threadIdx.x] = s;
	partb[threadIdx.x] = sb;
	__syncthreads();
	if(g == 0) {
		float t = part[i], tb = partb[i];
		for(int h = 1; h < G; ++h) { t += part[h * 64 + i]; tb += partb[h * 64 + i]; }
		if(i < in) gW[static_cast<int64_t>(o) * in + i] = t;
		if(i == 0) gb[o] = tb;
	}
}

// dX[b][i] = (sum_o dZ[b][o] W[o][i]) * (Zin[b][i] > 0)
__global__ __launch_bounds__(kBlock) void k_mlp_grad_x(const float* __restrict__ dZ, int out, const float* __restrict__ W,
                                                       int in, const float* __restrict__ Zin, int B, float* __restrict__ dX)
{
	int64_t const n = static_cast<int64_t>(B) * in;
	for(int64_t e = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; e < n; e += static_cast<int64_t>(gridDim.x) * kBlock) {
		int64_t const b = e / in;
		int const i = static_cast<int>(e - b * in);
		float s = 0.f;
		for(int o = 0; o < out; ++o) s = __builtin_fmaf(dZ[b * out + o], W[static_cast<int64_t>(o) * in + i], s);
		dX[e] = Zin[e] > 0.f ? s : 0.f;
	}
}

// loss = sum_b row_loss[b] / B, in one block with a fixed summation order
__global__ __launch_bounds__(kBlock) void k_mlp_sum_loss(const float* __restrict__ row_loss, int B, float* __restrict__ loss)
{
	__shared__ float part[kBlock];
	float s = 0.f;
	for(int b = threadIdx.x; b < B; b += kBlock) s += row_loss[b];
	part[threadIdx.x] = s;
	__syncthreads();
	if(threadIdx.x == 0) {
		float t = 0.f;
		for(int j = 0; j < kBlock; ++j) t += part[j];
		*loss = t / static_cast<float>(B);
	}
}

__global__ __launch_bounds__(kBlock) void k_mlp_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                                                     float* __restrict__ v, int64_t n, float lr, float b1, float b2,
                                                     float eps, float c1, float c2, float wd)
{
	for(int64_t e = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; e < n; e += static_cast<int64_t>(gridDim.x) * kBlock) {
		float const ge = g[e] + wd * w[e];
		float const me = b1 * m[e] + (1.f - b1) * ge;
		float const ve = b2 * v[e] + (1.f - b2) * ge * ge;
		m[e] = me;
		v[e] = ve;
		w[e] -= lr * (me / c1) / (sqrtf(ve / c2) + eps);
	}
}

// Bounded busy wait: keeps the queue busy while the host enqueues timed launches, so that event
// intervals contain kernel time only.  Always exits after `iters` sleeps.
__global__ void k_spin(int iters, int* __restrict__ sink)
{
	int x = 0;
	for(int i = 0; i < iters; ++i) {
		__builtin_amdgcn_s_sleep(127);
		x += i;
	}
	if(threadIdx.x == 0 && x == -1) *sink = x;
}

namespace
{
unsigned grid_for(int64_t n) { return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 4096))); }
int last(const char* what)
{
	hipError_t e = hipGetLastError();
	return e == hipSuccess ? TTVX_OK : hip_fail(e, what);
}
} // namespace

} // namespace ttvx

using namespace ttvx;

extern "C" {

size_t ttvx_mlp_work_floats(int batch)
{
	std::size_t const b = batch > 0 ? static_cast<std::size_t>(batch) : 0;
	return b * (kMlpF + 4 * kMlpH + kMlpV + 1);
}

int ttvx_mlp_forward(const float* params, const float* x, int batch, float* y, float* work, void* stream)
{
	clear_error();
	if(batch <= 0 || !params || !x || !y || !work) return fail(TTVX_ERR_ARGUMENT, "ttvx_mlp_forward: bad arguments");
	DeviceInfo di;
	if(!current_device(di)) return TTVX_ERR_DEVICE;
	int rc = check_device_span(params, kMlpParams * 4, "params");
	if(rc == TTVX_OK) rc = check_device_span(x, static_cast<size_t>(batch) * kMlpF * 4, "x");
	if(rc == TTVX_OK) rc = check_device_span(y, static_cast<size_t>(batch) * kMlpV * 4, "y");
	if(rc == TTVX_OK) rc = check_device_span(work, ttvx_mlp_work_floats(batch) * 4, "work");
	if(rc != TTVX_OK) return rc;
	hipStream_t s = static_cast<hipStream_t>(stream);
	WorkLayout L = layout_of(work, batch);
	int64_t const B = batch;
	hipLaunchKernelGGL(k_mlp_standardize, dim3(grid_for(B * kMlpF)), dim3(kBlock), 0, s, params, x, batch, L.xs);
	hipLaunchKernelGGL(k_mlp_dense, dim3(grid_for(B * kMlpH)), dim3(kBlock), 0, s, L.xs, kMlpF, params + kMlpOffW1,
	                   params + kMlpOffB1, kMlpH, batch, 0, L.z1);
	hipLaunchKernelGGL(k_mlp_dense, dim3(grid_for(B * kMlpH)), dim3(kBlock), 0, s, L.z1, kMlpH, params + kMlpOffW2,
	                   params + kMlpOffB2, kMlpH, batch, 1, L.z2);
	hipLaunchKernelGGL(k_mlp_dense, dim3(grid_for(B * kMlpV)), dim3(kBlock), 0, s, L.z2, kMlpH, params + kMlpOffW3,
	                   params + kMlpOffB3, kMlpV, batch, 1, y);
	return last("mlp forward");
}

// Gradient of the masked MSE w.r.t. the trainable parameters (W1..b3, kMlpTrainable floats,
// same order as the parameter vector); `work` must hold the forward pass of the same x.
int ttvx_mlp_backward(const float* params, const float* y, const float* t, const float* m, int batch, float* work,
                      float* grad, float* loss, void* stream)
{
	clear_error();
	if(batch <= 0 || !params || !y || !t || !m || !work || !grad || !loss)
		return fail(TTVX_ERR_ARGUMENT, "ttvx_mlp_backward: bad arguments");
	DeviceInfo di;
	if(!current_device(di)) return TTVX_ERR_DEVICE;
	int rc = check_device_span(params, kMlpParams * 4, "params");
	if(rc == TTVX_OK) rc = check_device_span(y, static_cast<size_t>(batch) * kMlpV * 4, "y");
	if(rc == TTVX_OK) rc = check_device_span(t, static_cast<size_t>(batch) * kMlpV * 4, "t");
	if(rc == TTVX_OK) rc = check_device_span(m, static_cast<size_t>(batch) * kMlpV * 4, "m");
	if(rc == TTVX_OK) rc = check_device_span(work, ttvx_mlp_work_floats(batch) * 4, "work");
	if(rc == TTVX_OK) rc = check_device_span(grad, kMlpTrainable * 4, "grad");
	if(rc == TTVX_OK) rc = check_device_span(loss, 4, "loss");
	if(rc != TTVX_OK) return rc;
	hipStream_t s = static_cast<hipStream_t>(stream);
	WorkLayout L = layout_of(work, batch);
	int64_t const B = batch;
	float* gW1 = grad;
	float* gb1 = gW1 + kMlpH * kMlpF;
	float* gW2 = gb1 + kMlpH;
	float* gb2 = gW2 + kMlpH * kMlpH;
	float* gW3 = gb2 + kMlpH;
	float* gb3 = gW3 + kMlpV * kMlpH;
	hipLaunchKernelGGL(k_mlp_dloss, dim3(grid_for(B)), dim3(kBlock), 0, s, y, t, m, batch, L.dy, L.loss);
	hipLaunchKernelGGL(k_mlp_sum_loss, dim3(1), dim3(kBlock), 0, s, L.loss, batch, loss);
	// layer 3: y = W3 relu(z2) + b3
	hipLaunchKernelGGL(k_mlp_grad_w, dim3(kMlpV), dim3(kBlock), 0, s, L.dy, kMlpV, L.z2, kMlpH, batch, 1, gW3, gb3);
	hipLaunchKernelGGL(k_mlp_grad_x, dim3(grid_for(B * kMlpH)), dim3(kBlock), 0, s, L.dy, kMlpV, params + kMlpOffW3, kMlpH,
	                   L.z2, batch, L.dz2);
	// layer 2: z2 = W2 relu(z1) + b2
	hipLaunchKernelGGL(k_mlp_grad_w, dim3(kMlpH), dim3(kBlock), 0, s, L.dz2, kMlpH, L.z1, kMlpH, batch, 1, gW2, gb2);
	hipLaunchKernelGGL(k_mlp_grad_x, dim3(grid_for(B * kMlpH)), dim3(kBlock), 0, s, L.dz2, kMlpH, params + kMlpOffW2, kMlpH,
	                   L.z1, batch, L.dz1);
	// layer 1: z1 = W1 xs + b1
	hipLaunchKernelGGL(k_mlp_grad_w, dim3(kMlpH), dim3(kBlock), 0, s, L.dz1, kMlpH, L.xs, kMlpF, batch, 0, gW1, gb1);
	return last("mlp backward");
}

int ttvx_mlp_adam(float* params, const float* grad, float* m1, float* m2, int step, float lr, float beta1, float beta2,
                  float eps, float weight_decay, void* stream)
{
	clear_error();
	if(!params || !grad || !m1 || !m2 || step < 1) return fail(TTVX_ERR_ARGUMENT, "ttvx_mlp_adam: bad arguments");
	DeviceInfo di;
	if(!current_device(di)) return TTVX_ERR_DEVICE;
	int rc = check_device_span(params, kMlpParams * 4, "params");
	if(rc == TTVX_OK) rc = check_device_span(grad, kMlpTrainable * 4, "grad");
	if(rc == TTVX_OK) rc = check_device_span(m1, kMlpTrainable * 4, "m1");
	if(rc == TTVX_OK) rc = check_device_span(m2, kMlpTrainable * 4, "m2");
	if(rc != TTVX_OK) return rc;
	float const c1 = 1.f - std::pow(beta1, static_cast<float>(step));
	float const c2 = 1.f - std::pow(beta2, static_cast<float>(step));
	int64_t const n = static_cast<int64_t>(kMlpTrainable);
	hipLaunchKernelGGL(k_mlp_adam, dim3(grid_for(n)), dim3(kBlock), 0, static_cast<hipStream_t>(stream),
	                   params + kMlpOffW1, grad, m1, m2, n, lr, beta1, beta2, eps, c1, c2, weight_decay);
	return last("mlp adam");
}

int ttvx_bench_spin(int iterations, int* sink, void* stream)
{
	clear_error();
	if(iterations < 0 || iterations > 100000 || sink == nullptr) return fail(TTVX_ERR_ARGUMENT, "ttvx_bench_spin: bad arguments");
	DeviceInfo di;
	if(!current_device(di)) return TTVX_ERR_DEVICE;
	int rc = check_device_span(sink, 4, "sink");
	if(rc != TTVX_OK) return rc;
	hipLaunchKernelGGL(k_spin, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream), iterations, sink);
	return last("spin launch");
}

} // extern "C"
