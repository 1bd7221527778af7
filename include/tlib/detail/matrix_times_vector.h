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

#include "../../ttvx.h"

namespace tlib::ttv::detail
{

namespace mtv_impl
{
// c[j] (+)= sum_k A(j, k) b[k] for the 2-D view A(j,k) = a[j*wj + k*wk], j < M, k < N; the
// accumulation runs on the GPU, so c may be host or device memory in both forms
template<class value_t, class size_t>
inline void run(value_t const* a, value_t const* b, value_t* c, size_t M, size_t N, size_t wj, size_t wk,
                bool accumulate = false)
{
	static_assert(std::is_same_v<value_t, float> || std::is_same_v<value_t, double>,
	              "tlib::ttv::detail gemv/dot on MI355X support float and double.");
	auto const m = static_cast<std::size_t>(M), n = static_cast<std::size_t>(N);
	auto const sj = static_cast<std::size_t>(wj), sk = static_cast<std::size_t>(wk);
	int status;
	if constexpr(std::is_same_v<value_t, float>)
		status = ttvx_mtv_f32(a, b, c, m, n, sj, sk, accumulate ? 1 : 0);
	else
		status = ttvx_mtv_f64(a, b, c, m, n, sj, sk, accumulate ? 1 : 0);
	if(status != TTVX_OK) throw std::runtime_error(ttvx_last_error());
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
	mtv_impl::run(a, b, c, M, N, size_t{1}, lda, true);
}

template<class value_t, class size_t>
inline void gemv_col_parallel(value_t const* const a, value_t const* const b, value_t* const c, size_t const M, size_t const N, size_t const lda)
{
	mtv_impl::run(a, b, c, M, N, size_t{1}, lda, true);
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
