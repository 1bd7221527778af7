// Drop-in C++17 front end of the MI355X tensor-times-vector library.
//
// Same three interfaces as the reference:
//   * C-like low-level  ttv(ep, sp, fp, q, p, a, na, wa, pia, b, nb, c, nc, wc, pic)
//       (reference include/tlib/ttv.h:54-92)
//   * high-level        ttv(q, A, b, ep, sp, fp) -> tensor   (reference include/tlib/ttv.h:99-114)
//       plus ttv(q, A, b) with the default policy
//   * operator*         A(q) * b                              (reference include/tlib/ttv.h:122-127)
//
// The arithmetic runs on a gfx950 GPU through libttvx.so (ttvx.h).  Validation happens in the
// library with the reference's messages; failures are re-thrown as std::runtime_error.  The
// policy tags are accepted for source compatibility and passed on as hints.  C is overwritten.
// Element types: float and double (other types are rejected at compile time).
#pragma once

#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../ttvx.h"
#include "detail/cases.h"
#include "detail/index.h"
#include "detail/layout.h"
#include "detail/matrix_times_vector.h"
#include "detail/shape.h"
#include "detail/strides.h"
#include "detail/tags.h"
#include "detail/tensor.h"
#include "detail/workload_computation.h"

namespace tlib::ttv
{

namespace detail
{
inline void throw_on_error(int status)
{
	if(status != TTVX_OK) throw std::runtime_error(ttvx_last_error());
}
} // namespace detail

template<class value_t, class size_t, class execution_t, class slicing_t, class fusion_t>
inline void ttv(
	execution_t ep, slicing_t sp, fusion_t fp,
	size_t const q, size_t const p,
	value_t const* const a, size_t const* const na, size_t const* const wa, size_t const* const pia,
	value_t const* const b, size_t const* const nb,
	value_t* const c, size_t const* const nc, size_t const* const wc, size_t const* const pic)
{
	static_assert(std::is_same_v<value_t, float> || std::is_same_v<value_t, double>,
	              "tlib::ttv::ttv on MI355X supports float and double elements.");
	static_assert(sizeof(size_t) == sizeof(std::size_t) && std::is_unsigned_v<size_t>,
	              "tlib::ttv::ttv on MI355X takes extents, strides and layouts as std::size_t.");
	auto const* na_ = reinterpret_cast<std::size_t const*>(na);
	auto const* wa_ = reinterpret_cast<std::size_t const*>(wa);
	auto const* pa_ = reinterpret_cast<std::size_t const*>(pia);
	auto const* nb_ = reinterpret_cast<std::size_t const*>(nb);
	auto const* nc_ = reinterpret_cast<std::size_t const*>(nc);
	auto const* wc_ = reinterpret_cast<std::size_t const*>(wc);
	auto const* pc_ = reinterpret_cast<std::size_t const*>(pic);
	int const e = detail::policy_code(ep), s = detail::policy_code(sp), f = detail::policy_code(fp);
	if constexpr(std::is_same_v<value_t, float>)
		detail::throw_on_error(ttvx_ttv_f32(e, s, f, q, p, a, na_, wa_, pa_, b, nb_, c, nc_, wc_, pc_));
	else
		detail::throw_on_error(ttvx_ttv_f64(e, s, f, q, p, a, na_, wa_, pa_, b, nb_, c, nc_, wc_, pc_));
}

template<class value_t, class execution_t, class slicing_t, class fusion_t>
inline auto ttv(std::size_t q, tensor<value_t> const& a, tensor<value_t> const& b,
                execution_t ep, slicing_t sp, fusion_t fp)
{
	auto c_shape  = detail::generate_output_shape (a.shape(),  q);
	auto c_layout = detail::generate_output_layout(a.layout(), q);
	auto c        = tensor<value_t>(c_shape, c_layout);
	auto const wa = a.strides();
	auto const wc = c.strides();
	ttv(ep, sp, fp, q, a.order(),
	    a.data().data(), a.shape().data(), wa.data(), a.layout().data(),
	    b.data().data(), b.shape().data(),
	    c.data().data(), c.shape().data(), wc.data(), c.layout().data());
	return c;
}

// Default policy of operator* (reference include/tlib/ttv.h:126).
template<class value_t>
inline auto ttv(std::size_t q, tensor<value_t> const& a, tensor<value_t> const& b)
{
	return ttv(q, a, b, execution_policy::par_loop, slicing_policy::subtensor, fusion_policy::all);
}

} // namespace tlib::ttv

template<class value_t>
inline auto operator*(tlib::ttv::tensor_view<value_t> const& a, tlib::ttv::tensor<value_t> const& b)
{
	return tlib::ttv::ttv(a.contraction_mode(), a.get_tensor(), b,
	                      tlib::ttv::execution_policy::par_loop, tlib::ttv::slicing_policy::subtensor,
	                      tlib::ttv::fusion_policy::all);
}
