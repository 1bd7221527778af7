// Multi-index <-> memory offset maps.
//
// Behaviour follows include/tlib/detail/index.h of the reference:
//   at       :30-48   j = sum_r i[r] * w[r]
//   at_1     :52-72   inverse of at for dense strides, peeling modes from the slowest layout position
//   at_at_1  :76-124  maps a linear index of a dense view (strides w_view) to an offset in an
//                     array with strides w_array, optionally in layout order pi
#pragma once

#include <cstddef>

namespace tlib::ttv::detail
{

template<class size_type>
constexpr auto at(unsigned p, size_type const* const i, size_type const* const w)
{
	size_type j = 0;
	for(unsigned r = 0; r < p; ++r) j += i[r] * w[r];
	return j;
}

template<class container_type>
constexpr auto at(container_type const& i, container_type const& w)
{
	return at(static_cast<unsigned>(i.size()), i.data(), w.data());
}

template<class size_type>
constexpr void at_1(unsigned const p, size_type* const i, size_type const j, size_type const* const w, size_type const* const pi)
{
	size_type rest = j;
	for(unsigned s = p; s-- > 0;) {
		auto const m = pi[s] - 1;
		i[m] = rest / w[m];
		rest -= i[m] * w[m];
	}
}

template<class container_type, class size_type>
constexpr auto at_1(size_type const j, container_type const& w, container_type const& pi)
{
	auto i = w;
	at_1(static_cast<unsigned>(i.size()), i.data(), j, w.data(), pi.data());
	return i;
}

template<class size_type>
constexpr auto at_at_1(unsigned const p, size_type const j_view, size_type const* const w_view, size_type const* const w_array)
{
	size_type rest = j_view, j = 0;
	for(unsigned r = 0; r < p; ++r) {
		auto const i = rest / w_view[r];
		rest -= i * w_view[r];
		j += i * w_array[r];
	}
	return j;
}

template<class container_type, class size_type>
constexpr auto at_at_1(size_type const j_view, container_type const& w_view, container_type const& w_array)
{
	return at_at_1(static_cast<unsigned>(w_view.size()), j_view, w_view.data(), w_array.data());
}

template<class size_type>
constexpr auto at_at_1(unsigned const p, size_type const j_view, size_type const* const w_view, size_type const* const w_array, size_type const* const pi)
{
	size_type rest = j_view, j = 0;
	for(unsigned s = p; s-- > 0;) {
		auto const m = pi[s] - 1;
		auto const i = rest / w_view[m];
		rest -= i * w_view[m];
		j += i * w_array[m];
	}
	return j;
}

template<class container_type, class size_type>
constexpr auto at_at_1(size_type const j_view, container_type const& w_view, container_type const& w_array, container_type const& pi)
{
	return at_at_1(static_cast<unsigned>(w_view.size()), j_view, w_view.data(), w_array.data(), pi.data());
}

} // namespace tlib::ttv::detail
