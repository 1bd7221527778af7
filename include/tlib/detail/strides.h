// Dense strides of a (shape, layout) pair and their validity rule.
//
// Behaviour follows include/tlib/detail/strides.h of the reference:
//   compute_strides  :31-57  w[pi[0]-1] = 1, w[pi[r]-1] = w[pi[r-1]-1] * n[pi[r-1]-1];
//                            all ones for vector- or scalar-shaped tensors (:44-45)
//   generate_strides :59-74
//   is_valid_strides :76-101 strides must not decrease along the layout
#pragma once

#include <algorithm>
#include <array>
#include <cstddef>
#include <iterator>
#include <stdexcept>
#include <vector>

#include "layout.h"
#include "shape.h"

namespace tlib::ttv::detail
{

template<class InputIt1, class InputIt2, class OutputIt>
inline void compute_strides(InputIt1 shape_begin, InputIt1 shape_end, InputIt2 layout_begin, OutputIt strides_begin)
{
	if(!is_valid_shape(shape_begin, shape_end))
		throw std::runtime_error("Error in tlib::detail::compute_strides(): input shape is not valid.");
	auto const n = std::distance(shape_begin, shape_end);
	if(!is_valid_layout(layout_begin, layout_begin + n))
		throw std::runtime_error("Error in tlib::detail::compute_strides(): input layout is not valid.");

	for(decltype(+n) r = 0; r < n; ++r) strides_begin[r] = 1u;

	if(is_vector(shape_begin, shape_end) || is_scalar(shape_begin, shape_end))
		return;
	if(!is_matrix(shape_begin, shape_end) && !is_tensor(shape_begin, shape_end))
		return;

	for(decltype(+n) r = 1; r < n; ++r) {
		auto const cur  = layout_begin[r] - 1;
		auto const prev = layout_begin[r - 1] - 1;
		strides_begin[cur] = strides_begin[prev] * shape_begin[prev];
	}
}

template<class size_type>
inline auto generate_strides(std::vector<size_type> const& shape, std::vector<size_type> const& layout)
{
	auto strides = std::vector<size_type>(shape.size());
	compute_strides(shape.begin(), shape.end(), layout.begin(), strides.begin());
	return strides;
}

template<class size_type, std::size_t N>
inline auto generate_strides(std::array<size_type, N> const& shape, std::array<size_type, N> const& layout)
{
	static_assert(N > 0, "Static error in tlib::detail::generate_strides(): N, i.e. length of array should be greater than zero.");
	auto strides = std::array<size_type, N>{};
	compute_strides(shape.begin(), shape.end(), layout.begin(), strides.begin());
	return strides;
}

template<class InputIt1, class InputIt2>
inline bool is_valid_strides(InputIt1 layout_begin, InputIt1 layout_end, InputIt2 stride_begin)
{
	if(!is_valid_layout(layout_begin, layout_end))
		throw std::runtime_error("Error in tlib::detail::is_valid_strides(): input layout is not valid.");
	for(auto it = layout_begin; it + 1 != layout_end && it != layout_end; ++it) {
		auto const lo = it[0] - 1;
		auto const hi = it[1] - 1;
		if(stride_begin[lo] > stride_begin[hi]) return false;
	}
	return true;
}

} // namespace tlib::ttv::detail
