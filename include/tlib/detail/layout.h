// Layout tuples pi: pi[r] is the one-based mode stored at layout position r (position 0 is
// the fastest).  First-order = {1,2,...,p}, last-order = {p,...,1}.
//
// Behaviour follows include/tlib/detail/layout.h of the reference:
//   is_valid_layout          :29-55    a permutation of 1..p
//   compute_k_order_layout   :57-76    {k,k-1,...,1,k+1,...,p}; k = 0 means last-order
//   generate_k_order_layout  :78-84
//   compute_inverse_layout   :99-109   out[pi[r]-1] = r+1
//   inverse_mode             :114-138  one-based layout position of a mode
//   compute/generate_output_layout :143-207   pi without q, entries > q decremented
#pragma once

#include <algorithm>
#include <array>
#include <cassert>
#include <cstddef>
#include <iterator>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "shape.h"

namespace tlib::ttv::detail
{

template<class InputIt>
inline bool is_valid_layout(InputIt begin, InputIt end)
{
	auto const n_signed = std::distance(begin, end);
	if(n_signed <= 0) return false;
	auto const n = static_cast<std::size_t>(n_signed);
	// every entry in 1..n and no entry twice
	std::vector<char> seen(n + 1, 0);
	for(auto it = begin; it != end; ++it) {
		auto const v = *it;
		if(v == 0u || static_cast<std::size_t>(v) > n) return false;
		if(seen[static_cast<std::size_t>(v)]) return false;
		seen[static_cast<std::size_t>(v)] = 1;
	}
	return true;
}

template<class OutputIt, class size_t>
inline void compute_k_order_layout(OutputIt begin, OutputIt end, size_t k)
{
	auto const n_signed = std::distance(begin, end);
	if(n_signed <= 0)
		throw std::runtime_error("Error in tlib::detail::compute_k_order: range provided by begin and end not correct!");
	using value_type = std::remove_reference_t<decltype(*begin)>;
	std::size_t const n = static_cast<std::size_t>(n_signed);
	if(k == 0u) k = static_cast<size_t>(n);
	std::size_t const m = static_cast<std::size_t>(k) > n ? n : static_cast<std::size_t>(k);
	// positions 0..m-1 hold m,...,1; positions m..n-1 hold m+1,...,n
	for(std::size_t r = 0; r < n; ++r, ++begin)
		*begin = static_cast<value_type>(r < m ? m - r : r + 1);
}

template<class size_t>
inline auto generate_k_order_layout(size_t p, size_t k)
{
	std::vector<size_t> v(p);
	compute_k_order_layout(v.begin(), v.end(), k);
	return v;
}

template<class OutputIt>
inline void compute_first_order_layout(OutputIt begin, OutputIt end)
{
	compute_k_order_layout(begin, end, 1u);
}

template<class OutputIt>
inline void compute_last_order_layout(OutputIt begin, OutputIt end)
{
	compute_k_order_layout(begin, end, 0u);
}

template<class InputIt, class OutputIt>
inline void compute_inverse_layout(InputIt begin, InputIt end, OutputIt output)
{
	if(!is_valid_layout(begin, end))
		throw std::runtime_error("Error in tlib::detail::compute_inverse_layout: input layout is not valid!");
	unsigned r = 1;
	for(auto it = begin; it != end; ++it, ++r)
		output[*it - 1] = r;
}

// One-based position of the one-based mode in the layout tuple.
template<class InputIt, class SizeType>
inline auto inverse_mode(InputIt layout_begin, InputIt layout_end, SizeType mode)
{
	using value_type = typename std::iterator_traits<InputIt>::value_type;
	if(!is_valid_layout(layout_begin, layout_end))
		throw std::runtime_error("Error in tlib::detail::inverse_mode(): input layout is not valid.");
	auto const p_ = std::distance(layout_begin, layout_end);
	if(p_ <= 0)
		throw std::runtime_error("Error in tlib::detail::inverse_mode(): input layout is invalid.");
	auto const p = static_cast<value_type>(p_);
	if(mode == 0u || mode > SizeType(p))
		throw std::runtime_error("Error in tlib::detail::inverse_mode(): mode should be one-based and equal to or less than layout size.");
	value_type r = 0;
	while(r < p && layout_begin[r] != value_type(mode)) ++r;
	assert(r < p);
	return r + 1;
}

template<class InputIt, class OutputIt, class ModeType>
inline void compute_output_layout(InputIt begin, InputIt end, OutputIt begin2, ModeType q)
{
	using value_type = typename std::iterator_traits<InputIt>::value_type;
	if(!is_valid_layout(begin, end))
		throw std::runtime_error("Error in tlib::detail::compute_inverse_layout: input layout is not valid!");
	auto const p_ = std::distance(begin, end);
	if(p_ < 1)
		throw std::runtime_error("Error in tlib::detail::compute_inverse_layout(): input layout is invalid.");
	auto const p = static_cast<value_type>(p_);
	if(1u > q || static_cast<value_type>(q) > p)
		throw std::runtime_error("Error in tlib::detail::compute_inverse_layout: mode must be greater zero and less than or equal to the order!");
	// drop the entry equal to q, renumber the modes above q
	for(auto it = begin; it != end; ++it) {
		auto const v = *it;
		if(v == static_cast<value_type>(q)) continue;
		*begin2++ = v > static_cast<value_type>(q) ? v - 1 : v;
	}
}

template<class SizeType, class ModeType>
inline auto generate_output_layout(std::vector<SizeType> const& input_layout, ModeType q)
{
	if(!is_valid_layout(input_layout.begin(), input_layout.end()))
		throw std::runtime_error("Error in tlib::detail::generate_output_layout(): input layout is not valid.");
	if(q == 0 || static_cast<std::size_t>(q) > input_layout.size())
		throw std::runtime_error("Error in tlib::detail::generate_output_layout(): constraction mode q should be greater than 0 and less than or equal to the tensor order.");
	auto output_layout = std::vector<SizeType>(input_layout.size() - 1);
	compute_output_layout(input_layout.begin(), input_layout.end(), output_layout.begin(), q);
	return output_layout;
}

template<class SizeType, class ModeType, std::size_t N>
inline auto generate_output_layout(std::array<SizeType, N> const& input_layout, ModeType q)
{
	static_assert(N > 0, "tlib::detail::generate_output_layout: order must be greater than zero.");
	// the reference checks this overload with the shape predicate (layout.h:192-193)
	if(!is_valid_shape(input_layout.begin(), input_layout.end()))
		throw std::runtime_error("Error in tlib::detail::generate_output_layout(): input layout is not valid.");
	if(q == 0 || static_cast<std::size_t>(q) > N)
		throw std::runtime_error("Error in tlib::detail::generate_output_layout(): constraction mode q should be greater than 0 and less than or equal to the tensor order.");
	auto output_layout = std::array<SizeType, N - 1>{};
	compute_output_layout(input_layout.begin(), input_layout.end(), output_layout.begin(), q);
	return output_layout;
}

} // namespace tlib::ttv::detail
