// Index-space division used by the fused-loop formulation of the product.
//
// Behaviour follows include/tlib/detail/workload_computation.h of the reference:
//   divide_layout :36-70   removes pi[0] and the contraction mode m from pi and renumbers;
//                          first = {1,2} if pi[0] < m else {2,1}
//   divide        :77-99   splits v into {v[pi[0]-1], v[m-1]} and the rest in mode order
//   divide        :106-124 output variant: {v[pi[0]-1]} and the rest
// On the GPU the fused loop is the launch grid; these functions remain for callers and tests.
#pragma once

#include <cassert>
#include <cstddef>
#include <stdexcept>
#include <utility>
#include <vector>

namespace tlib::ttv::detail
{

template<class size_type>
auto divide_layout(size_type const* const pi, unsigned const p, unsigned const m)
{
	if(p < m)
		throw std::runtime_error("Error in tlib::detail::divide_layout: contraction mode cannot be greater than the length of layout tuple.");
	if(m == 0)
		throw std::runtime_error("Error in tlib::detail::divide_layout: contraction mode cannot be zero.");
	if(p < 3)
		throw std::runtime_error("Error in tlib::detail::divide_layout: length of layout tuple must be greater than 2.");

	auto const first_mode = pi[0];
	auto const cmode      = static_cast<size_type>(m);
	auto rest = std::vector<size_type>();
	rest.reserve(p - 2);
	for(unsigned s = 0; s < p; ++s) {
		auto v = pi[s];
		if(v == first_mode || v == cmode) continue;
		// renumber: every removed mode below v shifts v down by one
		size_type shifted = v;
		if(v > first_mode) --shifted;
		if(v > cmode)      --shifted;
		rest.push_back(shifted);
	}
	auto head = first_mode < cmode ? std::vector<size_type>{1, 2} : std::vector<size_type>{2, 1};
	return std::make_pair(head, rest);
}

template<class size_type>
auto divide(size_type const* const v, size_type const* const pi, unsigned const p, unsigned const m)
{
	auto const first_mode = pi[0];
	auto rest = std::vector<size_type>();
	rest.reserve(p >= 2 ? p - 2 : 0);
	for(unsigned r = 1; r <= p; ++r)
		if(r != first_mode && r != m) rest.push_back(v[r - 1]);
	auto head = std::vector<size_type>{v[first_mode - 1], v[m - 1]};
	return std::make_pair(head, rest);
}

template<class size_type>
auto divide(size_type const* const v, size_type const* const pi, unsigned const p)
{
	auto const first_mode = pi[0];
	auto rest = std::vector<size_type>();
	rest.reserve(p >= 1 ? p - 1 : 0);
	for(unsigned r = 1; r <= p; ++r)
		if(r != first_mode) rest.push_back(v[r - 1]);
	auto head = std::vector<size_type>{v[first_mode - 1]};
	return std::make_pair(head, rest);
}

} // namespace tlib::ttv::detail
