// Shape predicates and the output shape of C = A x_q b.
//
// Behaviour follows include/tlib/detail/shape.h of the reference:
//   is_valid_shape  :30-34   non-empty, no zero extent
//   is_scalar       :38-46   all extents 1
//   is_vector       :49-64   one of the first two extents > 1, the other 1, the rest 1
//   is_matrix       :66-77   first two extents > 1, the rest 1
//   is_tensor       :80-89   order >= 3 and some extent after the second > 1
//   compute/generate_output_shape :103-158   the input shape without entry q
#pragma once

#include <algorithm>
#include <array>
#include <cassert>
#include <cstddef>
#include <iterator>
#include <stdexcept>
#include <vector>

namespace tlib::ttv::detail
{

template<class InputIt>
inline bool is_valid_shape(InputIt begin, InputIt end)
{
	if(begin == end) return false;
	for(auto it = begin; it != end; ++it)
		if(*it == 0u) return false;
	return true;
}

template<class InputIt>
inline bool is_scalar(InputIt begin, InputIt end)
{
	if(!is_valid_shape(begin, end)) return false;
	for(auto it = begin; it != end; ++it)
		if(*it != 1u) return false;
	return true;
}

template<class InputIt>
inline bool is_vector(InputIt begin, InputIt end)
{
	if(!is_valid_shape(begin, end)) return false;
	auto const p = std::distance(begin, end);
	if(p == 1) return *begin > 1u;
	auto const n0 = begin[0], n1 = begin[1];
	bool const one_big  = (n0 > 1u) || (n1 > 1u);
	bool const one_unit = (n0 == 1u) || (n1 == 1u);
	for(auto it = begin + 2; it != end; ++it)
		if(*it != 1u) return false;
	return one_big && one_unit;
}

template<class InputIt>
inline bool is_matrix(InputIt begin, InputIt end)
{
	if(!is_valid_shape(begin, end)) return false;
	if(std::distance(begin, end) < 2) return false;
	if(!(begin[0] > 1u && begin[1] > 1u)) return false;
	for(auto it = begin + 2; it != end; ++it)
		if(*it != 1u) return false;
	return true;
}

template<class InputIt>
inline bool is_tensor(InputIt begin, InputIt end)
{
	if(!is_valid_shape(begin, end)) return false;
	if(std::distance(begin, end) < 3) return false;
	for(auto it = begin + 2; it != end; ++it)
		if(*it > 1u) return true;
	return false;
}

// Writes the extents of A without the q-th one (q is one-based) to begin_out.
template<class InputIt, class OutputIt, class SizeType>
inline void compute_output_shape(InputIt begin, InputIt end, OutputIt begin_out, SizeType q)
{
	if(!is_valid_shape(begin, end))
		throw std::runtime_error("Error in tlib::detail::compute_output_shape(): input shape is not valid.");
	auto const p_ = std::distance(begin, end);
	if(p_ < 1)
		throw std::runtime_error("Error in tlib::detail::compute_output_shape(): input range is not valid.");
	auto const p = static_cast<SizeType>(p_);
	if(q == 0u || q > p)
		throw std::runtime_error("Error in tlib::detail::compute_output_shape(): constraction mode q should be greater than 0 and less than or equal to the tensor order.");
	SizeType r = 1;
	for(auto it = begin; it != end; ++it, ++r)
		if(r != q) *begin_out++ = *it;
}

template<class SizeType, class ModeType>
inline auto generate_output_shape(std::vector<SizeType> const& input_shape, ModeType q)
{
	if(!is_valid_shape(input_shape.begin(), input_shape.end()))
		throw std::runtime_error("Error in tlib::detail::generate_output_shape(): input shape is not valid.");
	if(q == 0 || static_cast<std::size_t>(q) > input_shape.size())
		throw std::runtime_error("Error in tlib::detail::generate_output_shape(): constraction mode q should be greater than 0 and less than or equal to the tensor order.");
	auto output_shape = std::vector<SizeType>(input_shape.size() - 1);
	compute_output_shape(input_shape.begin(), input_shape.end(), output_shape.begin(), static_cast<SizeType>(q));
	return output_shape;
}

template<class SizeType, class ModeType, std::size_t N>
inline auto generate_output_shape(std::array<SizeType, N> const& input_shape, ModeType q)
{
	static_assert(N > 0, "tlib::detail::generate_output_shape: order must be greater than zero.");
	if(!is_valid_shape(input_shape.begin(), input_shape.end()))
		throw std::runtime_error("Error in tlib::detail::generate_output_shape(): input shape is not valid.");
	if(q == 0 || static_cast<std::size_t>(q) > N)
		throw std::runtime_error("Error in tlib::detail::generate_output_shape(): constraction mode q should be greater than 0 and less than or equal to the tensor order.");
	auto output_shape = std::array<SizeType, N - 1>{};
	compute_output_shape(input_shape.begin(), input_shape.end(), output_shape.begin(), static_cast<SizeType>(q));
	return output_shape;
}

} // namespace tlib::ttv::detail
