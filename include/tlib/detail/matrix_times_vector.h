// 2-D slice products of the reference (include/tlib/detail/matrix_times_vector.h:51-295), the
// building blocks its Loops-over-GEMM engine calls per slice, computed on the GPU through the
// C-ABI as mode-2 products of a 2-D tensor:
//
//   gemv_row*(a,b,c,M,N,lda):  c[i]  = sum_k a[i*lda + k] * b[k]     i < M, k < N   (:51-91, :231-262)
//   gemv_col*(a,b,c,M,N,lda):  c[j] (+)= sum_i a[j + i*lda] * b[i]    j < M, i < N   (:108-221)
//   dot*(a,b,c,M):             c[0]  = sum_k a[k] * b[k]                              (:264-295)
//
// Same semantics as the reference: gemv_row* and dot* overwrite c; the non-BLAS gemv_col and
// gemv_col_parallel add into c (matrix_times_vector.h:117-126); gemv_col_blas overwrites
// (beta = 0, :213-215).  Host or device pointers; synchronous.
#pragma once

#include <stdexcept>
#include <type_traits>
#include <vector>

#include "../../ttvx.h"

namespace tlib::ttv::detail
{

namespace mtv_impl
{
// C[j] = sum_k A(j, k) b[k] for the 2-D view A(j,k) = a[j*wj + k*wk], j < M, k < N
template<class value_t, class size_t>
inline void run(value_t const* a, value_t const* b, value_t* c, size_t M, size_t N, size_t wj, size_t wk)
{
	static_assert(std::is_same_v<value_t, float> || std::is_same_v<value_t, double>,
	              "tlib::ttv::detail gemv/dot on MI355X support float and double.");
	std::size_t const na[2] = {static_cast<std::size_t>(M), static_cast<std::size_t>(N)};
	std::size_t const wa[2] = {static_cast<std::size_t>(wj), static_cast<std::size_t>(wk)};
	// layout: the mode with the smaller stride first (ties keep mode 1 first)
	std::size_t const pia[2] = {wj <= wk ? std::size_t{1} : std::size_t{2}, wj <= wk ? std::size_t{2} : std::size_t{1}};
	std::size_t const nb[1] = {static_cast<std::size_t>(N)};
	std::size_t const nc[1] = {static_cast<std::size_t>(M)};
	std::size_t const wc[1] = {1};
	std::size_t const pic[1] = {1};
	int status;
	if constexpr(std::is_same_v<value_t, float>)
		status = ttvx_ttv_f32(TTVX_EXEC_PAR_BLAS, TTVX_SLICE_SLICE, TTVX_FUSE_NONE, 2, 2, a, na, wa, pia, b, nb, c, nc, wc, pic);
	else
		status = ttvx_ttv_f64(TTVX_EXEC_PAR_BLAS, TTVX_SLICE_SLICE, TTVX_FUSE_NONE, 2, 2, a, na, wa, pia, b, nb, c, nc, wc, pic);
	if(status != TTVX_OK) throw std::runtime_error(ttvx_last_error());
}

template<class value_t, class size_t>
inline void run_accumulate(value_t const* a, value_t const* b, value_t* c, size_t M, size_t N, size_t wj, size_t wk)
{
	std::vector<value_t> tmp(static_cast<std::size_t>(M));
	run(a, b, tmp.data(), M, N, wj, wk);
	for(std::size_t j = 0; j < static_cast<std::size_t>(M); ++j) c[j] += tmp[j];
}
} // namespace mtv_impl

template<class value_t, class size_t>
inline void gemv_row(value_t const* const a, value_t const* const b, value_t* const c, size_t const M, size_t const N, size_t const lda)
{
	mtv_impl::run(a, b, c, M, N, lda, size_t{1});
}

template<class value_t, class size_t>
inline void gemv_row_parallel(value_t const* const a, value_t const* const b, value_t* const c, size_t const M, size_t const N, size_t const lda)
{
	mtv_impl::run(a, b, c, M, N, lda, size_t{1});
}

template<class value_t, class size_t>
inline void gemv_row_blas(value_t const* const a, value_t const* const b, value_t* const c, size_t const M, size_t const N, size_t const lda)
{
	mtv_impl::run(a, b, c, M, N, lda, size_t{1});
}

template<class value_t, class size_t>
inline void gemv_col(value_t const* const a, value_t const* const b, value_t* const c, size_t const M, size_t const N, size_t const lda)
{
	mtv_impl::run_accumulate(a, b, c, M, N, size_t{1}, lda);
}

template<class value_t, class size_t>
inline void gemv_col_parallel(value_t const* const a, value_t const* const b, value_t* const c, size_t const M, size_t const N, size_t const lda)
{
	mtv_impl::run_accumulate(a, b, c, M, N, size_t{1}, lda);
}

template<class value_t, class size_t>
inline void gemv_col_blas(value_t const* const a, value_t const* const b, value_t* const c, size_t const M, size_t const N, size_t const lda)
{
	mtv_impl::run(a, b, c, M, N, size_t{1}, lda);
}

template<class value_t, class size_t>
inline void dot(value_t const* const a, value_t const* const b, value_t* const c, size_t const M)
{
	mtv_impl::run(a, b, c, size_t{1}, M, M, size_t{1});
}

template<class value_t, class size_t>
inline void dot_parallel(value_t const* const a, value_t const* const b, value_t* const c, size_t const M)
{
	mtv_impl::run(a, b, c, size_t{1}, M, M, size_t{1});
}

} // namespace tlib::ttv::detail
