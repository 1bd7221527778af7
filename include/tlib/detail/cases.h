// Case classification of (p, q, pi), same numbering as the reference (include/tlib/detail/cases.h:24-36):
//   1: p == 1                      (scalar dot; unreachable through ttv() because C would be empty)
//   2..5: p == 2 with q in {1,2} and pi[0] in {1,2}
//   6: p >= 3 and q is the fastest mode      -> one dot product per output element
//   7: p >= 3 and q is the slowest mode      -> one column GEMV over the whole tensor
//   8: p >= 3 otherwise                      -> GEMV per slice / subtensor
// The planner does not branch on these numbers (it folds every case to a canonical
// [O][K][I] view); they are kept for callers and for the layout-consistency check.
#pragma once

namespace tlib::ttv::detail
{

template<unsigned case_nr, typename size_t>
inline constexpr bool is_case(unsigned p, unsigned q, size_t const* const pi)
{
	static_assert(case_nr >= 1u && case_nr <= 8u, "tlib::ttv::detail::is_case: only 8 cases from 1 to 8 are covered.");
	if constexpr(case_nr == 1u) return p == 1u;
	if constexpr(case_nr == 2u) return p == 2u && q == 1u && pi[0] == 1u;
	if constexpr(case_nr == 3u) return p == 2u && q == 2u && pi[0] == 1u;
	if constexpr(case_nr == 4u) return p == 2u && q == 1u && pi[0] == 2u;
	if constexpr(case_nr == 5u) return p == 2u && q == 2u && pi[0] == 2u;
	if constexpr(case_nr == 6u) return p >= 3u && pi[0] == q;
	if constexpr(case_nr == 7u) return p >= 3u && pi[p - 1] == q;
	if constexpr(case_nr == 8u) return p >= 3u && pi[0] != q && pi[p - 1] != q;
	return false;
}

} // namespace tlib::ttv::detail
