// Validation and folding; see plan.hpp.
#include "plan.hpp"

#include <algorithm>

#include "../../include/tlib/detail/cases.h"
#include "../../include/tlib/detail/layout.h"
#include "../../include/tlib/detail/shape.h"
#include "../../include/tlib/detail/strides.h"
#include "../../include/ttvx.h"

namespace ttvx
{

namespace
{
Check fail(int status, const char* msg)
{
	Check c;
	c.status  = status;
	c.message = msg;
	return c;
}
} // namespace

Check validate(std::size_t q, std::size_t p,
               const void* a, const std::size_t* na, const std::size_t* wa, const std::size_t* pia,
               const void* b, const std::size_t* nb,
               const void* c, const std::size_t* nc, const std::size_t* wc, const std::size_t* pic)
{
	using namespace tlib::ttv::detail;
	// include/tlib/ttv.h:64-89, same order and text
	if(p == 0)        return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: input tensor order should be greater zero.");
	if(q == 0 || q > p) return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: contraction mode should be greater zero or less than or equal to p.");
	if(a == nullptr)  return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: pointer to input tensor A should not be zero.");
	if(b == nullptr)  return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: pointer to input vector B should not be zero.");
	if(c == nullptr)  return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: pointer to output tensor C should not be zero.");
	if(na == nullptr) return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: pointer to input tensor shape vector na should not be zero.");
	if(nb == nullptr) return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: pointer to input vector shape vector nb should not be zero.");
	if(nc == nullptr) return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: pointer to output tensor shape vector nc should not be zero.");
	if(wa == nullptr) return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: pointer to input tensor stride vector wa should not be zero.");
	if(wc == nullptr) return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: pointer to output tensor stride vector wc should not be zero.");
	if(pia == nullptr) return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: pointer to input tensor permutation vector pia should not be zero.");
	if(pic == nullptr) return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: pointer to output tensor permutation vector pic should not be zero.");
	if(na[q - 1] != nb[0]) return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: contraction dimension of A and B are not equal.");
	if(!is_valid_shape(na, na + p))      return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: shape vector of A is not valid.");
	if(!is_valid_shape(nc, nc + p - 1))  return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: shape vector of C is not valid.");
	if(!is_valid_layout(pia, pia + p))     return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: layout vector of A is not valid.");
	if(!is_valid_layout(pic, pic + p - 1)) return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: layout vector of C is not valid.");
	if(!is_valid_strides(pia, pia + p, wa))     return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: stride vector of A is not valid.");
	if(!is_valid_strides(pic, pic + p - 1, wc)) return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: stride vector of C is not valid.");

	// The reference checks pia against pic when it reaches its slice/subtensor recursion
	// (case 8; tensor_times_vector.h:147-168): the layout of C must be pia without q, with the
	// modes above q renumbered.  Same test, same messages.
	if(is_case<8>(static_cast<unsigned>(p), static_cast<unsigned>(q), pia)) {
		std::size_t pos = 0;
		while(pos < p && pia[pos] != q) ++pos;
		auto expected = [q](std::size_t m) { return m > q ? m - 1 : m; };
		for(std::size_t r = 0; r < pos; ++r)
			if(pic[r] != expected(pia[r]))
				return fail(TTVX_ERR_LAYOUT, "Error in tlib::detail::compute_inverse_pia_m: beginning of layout tuples of both tensors are not correct.");
		for(std::size_t r = pos; r + 1 < p; ++r)
			if(pic[r] != expected(pia[r + 1]))
				return fail(TTVX_ERR_LAYOUT, "Error in tlib::detail::compute_inverse_pia_m: end of layout tuples of both tensors are not correct.");
	}
	return Check{};
}

Check validate_a(std::size_t q, std::size_t p, const std::size_t* na, const std::size_t* wa, const std::size_t* pia)
{
	using namespace tlib::ttv::detail;
	if(p == 0)          return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: input tensor order should be greater zero.");
	if(q == 0 || q > p) return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: contraction mode should be greater zero or less than or equal to p.");
	if(na == nullptr)   return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: pointer to input tensor shape vector na should not be zero.");
	if(wa == nullptr)   return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: pointer to input tensor stride vector wa should not be zero.");
	if(pia == nullptr)  return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: pointer to input tensor permutation vector pia should not be zero.");
	if(!is_valid_shape(na, na + p))       return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: shape vector of A is not valid.");
	if(!is_valid_layout(pia, pia + p))    return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: layout vector of A is not valid.");
	if(!is_valid_strides(pia, pia + p, wa)) return fail(TTVX_ERR_ARGUMENT, "Error in tlib::tensor_times_vector: stride vector of A is not valid.");
	return Check{};
}

Fold fold(std::size_t q, std::size_t p, const std::size_t* na, const std::size_t* wa, const std::size_t* pia,
          const std::size_t* wc, const std::size_t* pic)
{
	using namespace tlib::ttv::detail;
	Fold f;
	f.K   = static_cast<int64_t>(na[q - 1]);
	f.sAk = static_cast<int64_t>(wa[q - 1]);

	// Dense output strides derived from A when the caller gives none (planning queries).
	std::vector<std::size_t> wc_own;
	if(wc == nullptr) {
		std::vector<std::size_t> nav(na, na + p), piv(pia, pia + p);
		if(p >= 2) {
			auto nc  = generate_output_shape(nav, q);
			auto pc  = generate_output_layout(piv, q);
			wc_own   = generate_strides(nc, pc);
		}
		wc = wc_own.data();
	}
	(void)pic;

	std::size_t qpos = 0;
	while(qpos < p && pia[qpos] != q) ++qpos;
	f.p    = static_cast<int>(p);
	f.qpos = static_cast<int>(qpos);
	for(std::size_t s = 0; s < p; ++s) f.n_layout.push_back(static_cast<int64_t>(na[pia[s] - 1]));

	std::vector<FreeDim> inner, outer;
	f.a_span = 1;
	f.c_span = 1;
	f.outputs = 1;
	f.a_span += (f.K - 1) * f.sAk;
	for(std::size_t s = 0; s < p; ++s) {
		std::size_t const m = pia[s];
		if(m == q) continue;
		std::size_t const cm = m < q ? m : m - 1;   // mode of C
		FreeDim d{static_cast<int64_t>(na[m - 1]), static_cast<int64_t>(wa[m - 1]), static_cast<int64_t>(wc[cm - 1])};
		f.a_span  += (d.n - 1) * d.sa;
		f.c_span  += (d.n - 1) * d.sc;
		f.outputs *= d.n;
		if(d.n == 1) continue;
		auto& list = s < qpos ? inner : outer;
		if(!list.empty()) {
			auto& last = list.back();
			if(d.sa == last.sa * last.n && d.sc == last.sc * last.n) {
				last.n *= d.n;
				continue;
			}
		}
		list.push_back(d);
	}

	f.three_d = inner.size() <= 1 && outer.size() <= 1;
	if(f.three_d) {
		if(!inner.empty()) { f.I = inner[0].n; f.sAi = inner[0].sa; f.sCi = inner[0].sc; }
		else               { f.I = 1;          f.sAi = 1;           f.sCi = 1; }
		if(!outer.empty()) { f.O = outer[0].n; f.sAo = outer[0].sa; f.sCo = outer[0].sc; }
		else               { f.O = 1;          f.sAo = 0;           f.sCo = 0; }
	}
	f.dims = inner;
	f.dims.insert(f.dims.end(), outer.begin(), outer.end());

	auto const pu = static_cast<unsigned>(p), qu = static_cast<unsigned>(q);
	for(unsigned cnr = 1; cnr <= 8; ++cnr) {
		bool hit = false;
		switch(cnr) {
			case 1: hit = is_case<1>(pu, qu, pia); break;
			case 2: hit = is_case<2>(pu, qu, pia); break;
			case 3: hit = is_case<3>(pu, qu, pia); break;
			case 4: hit = is_case<4>(pu, qu, pia); break;
			case 5: hit = is_case<5>(pu, qu, pia); break;
			case 6: hit = is_case<6>(pu, qu, pia); break;
			case 7: hit = is_case<7>(pu, qu, pia); break;
			case 8: hit = is_case<8>(pu, qu, pia); break;
		}
		if(hit) { f.ref_case = static_cast<int>(cnr); break; }
	}
	return f;
}

} // namespace ttvx
