// Compile-time dispatch of a launch configuration to the kernel instantiations of one storage
// type.  Each launch_<type>.hip includes this file once for its type, so the four storage types
// compile in parallel.
#pragma once

#include <algorithm>

#include "runtime.hpp"

namespace ttvx
{

// Launches kernel with the configuration's grid, or, for contiguous schedules, with as many
// blocks as the GPU holds at once (the kernel's occupancy), so every block owns one range.
#ifndef TTVX_CONTIG_RESIDENT
#define TTVX_CONTIG_RESIDENT 1
#endif
template<typename Kern, typename... Xs>
void launch_sched(Kern kernel, const Config& cfg, hipStream_t s, Xs... xs)
{
	unsigned gx = static_cast<unsigned>(cfg.grid);
	if(cfg.args.contig && TTVX_CONTIG_RESIDENT) {
		int nb = 0;
		if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kBlock, cfg.shmem) != hipSuccess || nb < 1) nb = 1;
		int64_t const work = cfg.kind == KIND_DOT ? (cfg.args.groups + kBlock / 64 - 1) / (kBlock / 64) : cfg.args.items;
		int64_t const resident = static_cast<int64_t>(cfg.cus) * nb;
		gx = static_cast<unsigned>(std::max<int64_t>(1, std::min(work, resident)));
	}
	hipLaunchKernelGGL(kernel, dim3(gx, static_cast<unsigned>(cfg.ksplit)), dim3(kBlock), cfg.shmem, s, xs...);
}

// k_rows with the lanes per row fixed at compile time for full-width vectors (4..64 lanes), else read
// at run time
template<typename S, int VEC, int U, bool KH>
int launch_rows(const Config& cfg, const S* A, const typename Traits<S>::acc* B, void* C, hipStream_t s)
{
	dim3 const grid(static_cast<unsigned>(cfg.grid)), block(kBlock);
	if(cfg.args.tight) {
		hipLaunchKernelGGL((k_rows<S, VEC, U, -1, KH>), grid, block, 0, s, A, B, C, cfg.args);
		return TTVX_OK;
	}
	if constexpr(VEC == Traits<S>::vmax) {
		switch(cfg.args.LPR) {
			case 4:  hipLaunchKernelGGL((k_rows<S, VEC, U, 4, KH>), grid, block, 0, s, A, B, C, cfg.args); return TTVX_OK;
			case 8:  hipLaunchKernelGGL((k_rows<S, VEC, U, 8, KH>), grid, block, 0, s, A, B, C, cfg.args); return TTVX_OK;
			case 16: hipLaunchKernelGGL((k_rows<S, VEC, U, 16, KH>), grid, block, 0, s, A, B, C, cfg.args); return TTVX_OK;
			case 32: hipLaunchKernelGGL((k_rows<S, VEC, U, 32, KH>), grid, block, 0, s, A, B, C, cfg.args); return TTVX_OK;
			case 64: hipLaunchKernelGGL((k_rows<S, VEC, U, 64, KH>), grid, block, 0, s, A, B, C, cfg.args); return TTVX_OK;
			default: break;
		}
	}
	hipLaunchKernelGGL((k_rows<S, VEC, U, 0, KH>), grid, block, 0, s, A, B, C, cfg.args);
	return TTVX_OK;
}

template<typename S, int VEC, bool KH>
int launch_vec(const Config& cfg, const S* A, const typename Traits<S>::acc* B, void* C, hipStream_t s)
{
	dim3 const grid(static_cast<unsigned>(cfg.grid)), block(kBlock);
	if(cfg.lean == LEAN_SLAB) {
		if constexpr(!kRealigned<S>) {
			hipLaunchKernelGGL((k_slab<S, VEC, KH>), grid, block, 0, s, A, B, C, cfg.args);
			return TTVX_OK;
		}
		return fail(TTVX_ERR_UNSUPPORTED, "ttvx: realigned loads for row and column products only");
	}
	if(cfg.lean == LEAN_ROWS) {
		switch(cfg.U) {
			case 1: return launch_rows<S, VEC, 1, KH>(cfg, A, B, C, s);
			case 2: return launch_rows<S, VEC, 2, KH>(cfg, A, B, C, s);
			case 4: return launch_rows<S, VEC, 4, KH>(cfg, A, B, C, s);
			case 8: return launch_rows<S, VEC, 8, KH>(cfg, A, B, C, s);
			default: return fail(TTVX_ERR_UNSUPPORTED, "ttvx: unsupported row unroll");
		}
	}
	if(cfg.kind == KIND_DOT) {
		switch(cfg.U) {
			case 1: launch_sched(k_dot<S, VEC, 1, KH>, cfg, s, A, B, C, cfg.args); break;
			case 2: launch_sched(k_dot<S, VEC, 2, KH>, cfg, s, A, B, C, cfg.args); break;
			case 4: launch_sched(k_dot<S, VEC, 4, KH>, cfg, s, A, B, C, cfg.args); break;
			case 8: launch_sched(k_dot<S, VEC, 8, KH>, cfg, s, A, B, C, cfg.args); break;
			default: return fail(TTVX_ERR_UNSUPPORTED, "ttvx: unsupported row unroll");
		}
	} else {
		switch(cfg.U) {
			case 1: launch_sched(k_col<S, VEC, 1, KH>, cfg, s, A, B, C, cfg.args); break;
			case 2: launch_sched(k_col<S, VEC, 2, KH>, cfg, s, A, B, C, cfg.args); break;
			case 4: launch_sched(k_col<S, VEC, 4, KH>, cfg, s, A, B, C, cfg.args); break;
			default: return fail(TTVX_ERR_UNSUPPORTED, "ttvx: unsupported slab unroll");
		}
	}
	return TTVX_OK;
}

// Tries the vector widths V, V/2, ..., 1 (all instantiated).
template<typename S, bool KH, int V>
int launch_width(const Config& cfg, const S* A, const typename Traits<S>::acc* B, void* C, hipStream_t s)
{
	if(cfg.vec == V) return launch_vec<S, V, KH>(cfg, A, B, C, s);
	if constexpr(V > 1) return launch_width<S, KH, V / 2>(cfg, A, B, C, s);
	else return fail(TTVX_ERR_UNSUPPORTED, "ttvx: unsupported vector width");
}

template<typename S, bool KH>
int launch_kh(const Config& cfg, const S* A, const typename Traits<S>::acc* B, void* C, hipStream_t s)
{
	if(cfg.kind == KIND_STRIDED) {
		if constexpr(!kRealigned<S>) {
			hipLaunchKernelGGL((k_strided<S, KH>), dim3(static_cast<unsigned>(cfg.grid)), dim3(kBlock), 0, s, A, B, C, cfg.sargs);
			return TTVX_OK;
		}
		return fail(TTVX_ERR_UNSUPPORTED, "ttvx: realigned loads for row and column products only");
	}
	return launch_width<S, KH, Traits<S>::vmax>(cfg, A, B, C, s);
}

template<typename S>
int launch_typed(const Config& cfg, const void* a, const void* b, void* c, bool kahan, hipStream_t s)
{
	using Acc = typename Traits<S>::acc;
	auto const* A = static_cast<const S*>(a);
	auto const* B = static_cast<const Acc*>(b);
	int rc;
	if constexpr(std::is_same_v<Acc, double>) {
		(void)kahan;   // FP64 accumulation needs no compensation
		rc = launch_kh<S, false>(cfg, A, B, c, s);
	} else {
		rc = kahan ? launch_kh<S, true>(cfg, A, B, c, s) : launch_kh<S, false>(cfg, A, B, c, s);
	}
	if(rc != TTVX_OK) return rc;
	hipError_t const e = hipGetLastError();
	if(e != hipSuccess) return hip_fail(e, "kernel launch");
	return TTVX_OK;
}

} // namespace ttvx
