// Validation and folding of the host planner (ttv_amd/csrc/plan.cpp), checked by brute force.
//
// For random problems (order 1-6, extents 1-4, every layout, dense, padded and arbitrary
// non-decreasing strides of A, dense or padded strides of C) the folded description must name
// exactly the (A address, C address, contraction index) triples that enumerating every element of A
// gives, as a multiset: the 3-D view A[o*sAo + k*sAk + i*sAi] -> C[o*sCo + i*sCi] when it exists,
// else the merged free dimensions of the strided kernel.  outputs, a_span and c_span must equal the
// enumerated counts and address bounds.  Random and malformed arguments (zero extents, layouts
// that are not permutations, decreasing strides, q out of range, mismatched b and C) must give the
// status a direct reading of the reference's checks gives (include/tlib/ttv.h:64-89) and nothing
// else.  Built by tests/test_host_sanitizers.py with AddressSanitizer and UndefinedBehaviorSanitizer
// (host code only), so an out-of-bounds access or overflow in the planner fails the test.
#include "../../ttv_amd/csrc/plan.hpp"
#include "../../include/ttvx.h"

#include <algorithm>
#include <cstdio>
#include <cstdint>
#include <random>
#include <tuple>
#include <vector>

using sz = std::size_t;
using Triple = std::tuple<int64_t, int64_t, int64_t>;

static int g_fail = 0, g_checks = 0;
#define CHECK(cond, ...)                                                   \
	do {                                                                   \
		++g_checks;                                                        \
		if(!(cond)) {                                                      \
			++g_fail;                                                      \
			if(g_fail < 20) {                                              \
				std::fprintf(stderr, "FAIL line %d: %s: ", __LINE__, #cond); \
				std::fprintf(stderr, __VA_ARGS__);                         \
				std::fprintf(stderr, "\n");                                \
			}                                                              \
		}                                                                  \
	} while(0)

// strides along a layout: stride of layout position r = stride of position r-1 times its extent
// (dense), plus a padding (padded), or any non-decreasing value (arbitrary)
static std::vector<sz> make_strides(std::vector<sz> const& n, std::vector<sz> const& pi, int kind, std::mt19937& rng)
{
	std::vector<sz> w(n.size(), 1);
	if(n.empty()) return w;
	w[pi[0] - 1] = kind == 2 ? rng() % 3 : 1;
	for(sz r = 1; r < n.size(); ++r) {
		sz const prev = w[pi[r - 1] - 1];
		if(kind == 0) w[pi[r] - 1] = prev * n[pi[r - 1] - 1];
		else if(kind == 1) w[pi[r] - 1] = prev * n[pi[r - 1] - 1] + rng() % 3;
		else w[pi[r] - 1] = prev + rng() % 7;
	}
	return w;
}

static std::string str(std::vector<sz> const& v)
{
	std::string s = "[";
	for(sz x : v) s += std::to_string(x) + ",";
	return s + "]";
}

static void fold_case(std::mt19937& rng)
{
	sz const p = 1 + rng() % 6;
	std::vector<sz> na(p), pia(p);
	for(sz r = 0; r < p; ++r) na[r] = 1 + rng() % 4;
	for(sz r = 0; r < p; ++r) pia[r] = r + 1;
	std::shuffle(pia.begin(), pia.end(), rng);
	sz const q = 1 + rng() % p;
	int const akind = static_cast<int>(rng() % 3);
	std::vector<sz> wa = make_strides(na, pia, akind, rng);

	ttvx::Check ck = ttvx::validate_a(q, p, na.data(), wa.data(), pia.data());
	CHECK(ck.status == TTVX_OK, "valid problem refused: %s", ck.message.c_str());
	if(ck.status != TTVX_OK) return;

	// C: the reference's output shape and layout (A's without mode q, modes above q renumbered)
	std::vector<sz> nc, pic;
	for(sz r = 0; r < p; ++r)
		if(r + 1 != q) nc.push_back(na[r]);
	for(sz r = 0; r < p; ++r)
		if(pia[r] != q) pic.push_back(pia[r] > q ? pia[r] - 1 : pia[r]);
	bool const own_wc = p >= 2 && rng() % 2 == 0;
	std::vector<sz> wc = p >= 2 ? make_strides(nc, pic, own_wc ? 1 : 0, rng) : std::vector<sz>{1};
	ttvx::Fold const f = ttvx::fold(q, p, na.data(), wa.data(), pia.data(), own_wc ? wc.data() : nullptr, nullptr);

	// brute force over every element of A
	std::vector<Triple> want;
	int64_t amax = 0, cmax = 0, outputs = 1;
	for(sz r = 0; r < p; ++r)
		if(r + 1 != q) outputs *= static_cast<int64_t>(na[r]);
	std::vector<sz> j(p, 0);
	for(;;) {
		int64_t a = 0, c = 0;
		for(sz r = 0, m = 0; r < p; ++r) {
			a += static_cast<int64_t>(j[r] * wa[r]);
			if(r + 1 != q) c += static_cast<int64_t>(j[r] * wc[m++]);
		}
		want.emplace_back(a, c, static_cast<int64_t>(j[q - 1]));
		amax = std::max(amax, a);
		cmax = std::max(cmax, c);
		sz r = 0;
		while(r < p && ++j[r] == na[r]) j[r++] = 0;
		if(r == p) break;
	}

	// the fold's description
	std::vector<Triple> got;
	if(f.three_d) {
		for(int64_t o = 0; o < f.O; ++o)
			for(int64_t k = 0; k < f.K; ++k)
				for(int64_t i = 0; i < f.I; ++i)
					got.emplace_back(o * f.sAo + k * f.sAk + i * f.sAi, o * f.sCo + i * f.sCi, k);
	} else {
		std::vector<int64_t> d(f.dims.size(), 0);
		for(;;) {
			int64_t a = 0, c = 0;
			for(sz r = 0; r < d.size(); ++r) {
				a += d[r] * f.dims[r].sa;
				c += d[r] * f.dims[r].sc;
			}
			for(int64_t k = 0; k < f.K; ++k) got.emplace_back(a + k * f.sAk, c, k);
			sz r = 0;
			while(r < d.size() && ++d[r] == f.dims[r].n) d[r++] = 0;
			if(r == d.size()) break;
		}
	}
	std::sort(want.begin(), want.end());
	std::sort(got.begin(), got.end());
	CHECK(got == want, "fold differs: q=%zu na=%s wa=%s pia=%s wc=%s three_d=%d O=%ld K=%ld I=%ld", q,
	      str(na).c_str(), str(wa).c_str(), str(pia).c_str(), str(wc).c_str(), int(f.three_d), long(f.O), long(f.K),
	      long(f.I));
	CHECK(f.outputs == outputs, "outputs %ld != %ld", long(f.outputs), long(outputs));
	CHECK(f.a_span == amax + 1, "a_span %ld != %ld (na=%s wa=%s)", long(f.a_span), long(amax + 1), str(na).c_str(),
	      str(wa).c_str());
	CHECK(f.c_span == cmax + 1, "c_span %ld != %ld", long(f.c_span), long(cmax + 1));
	CHECK(f.K == static_cast<int64_t>(na[q - 1]), "K");
	// a dense A always folds to the 3-D view
	if(akind == 0 && !own_wc) CHECK(f.three_d, "dense problem did not fold: q=%zu na=%s pia=%s", q, str(na).c_str(), str(pia).c_str());
}

// the reference's argument checks read directly: 0 when the arguments are valid
static int expected_status(sz q, sz p, std::vector<sz> const& na, std::vector<sz> const& wa, std::vector<sz> const& pia,
                           sz nb, std::vector<sz> const& nc, std::vector<sz> const& wc, std::vector<sz> const& pic)
{
	auto perm = [](std::vector<sz> const& v, sz m) {
		if(m == 0) return false;
		std::vector<int> seen(m + 1, 0);
		for(sz r = 0; r < m; ++r) {
			if(v[r] < 1 || v[r] > m || seen[v[r]]) return false;
			seen[v[r]] = 1;
		}
		return true;
	};
	auto shape = [](std::vector<sz> const& v, sz m) {
		if(m == 0) return false;
		for(sz r = 0; r < m; ++r)
			if(v[r] == 0) return false;
		return true;
	};
	auto ordered = [](std::vector<sz> const& pi, std::vector<sz> const& w, sz m) {
		for(sz r = 0; r + 1 < m; ++r)
			if(w[pi[r] - 1] > w[pi[r + 1] - 1]) return false;
		return true;
	};
	if(p == 0 || q == 0 || q > p) return TTVX_ERR_ARGUMENT;
	if(na[q - 1] != nb) return TTVX_ERR_ARGUMENT;
	if(!shape(na, p) || !shape(nc, p - 1) || !perm(pia, p) || !perm(pic, p - 1)) return TTVX_ERR_ARGUMENT;
	if(!ordered(pia, wa, p) || !ordered(pic, wc, p - 1)) return TTVX_ERR_ARGUMENT;
	return -1;      // argument checks pass; the case-8 layout rule and the shape of C decide
}

static void validate_case(std::mt19937& rng)
{
	sz const p = rng() % 6;
	sz const q = rng() % (p + 2);
	std::vector<sz> na(std::max<sz>(p, 1)), pia(std::max<sz>(p, 1));
	for(sz r = 0; r < p; ++r) na[r] = rng() % 8 == 0 ? 0 : 1 + rng() % 4;
	for(sz r = 0; r < p; ++r) pia[r] = r + 1;
	std::shuffle(pia.begin(), pia.begin() + p, rng);
	if(p > 0 && rng() % 6 == 0) pia[rng() % p] = rng() % (p + 2);        // not a permutation
	std::vector<sz> wa(std::max<sz>(p, 1), 1);
	bool const perm_ok = [&] {
		std::vector<int> seen(p + 2, 0);
		for(sz r = 0; r < p; ++r) {
			if(pia[r] < 1 || pia[r] > p || seen[pia[r]]) return false;
			seen[pia[r]] = 1;
		}
		return true;
	}();
	if(perm_ok) wa = make_strides(std::vector<sz>(na.begin(), na.begin() + p), std::vector<sz>(pia.begin(), pia.begin() + p),
	                              static_cast<int>(rng() % 3), rng);
	for(sz r = 0; r < p; ++r)
		if(rng() % 10 == 0) wa[r] = rng() % 50;                              // possibly decreasing
	sz const m = p > 0 ? p - 1 : 0;
	std::vector<sz> nc(std::max<sz>(m, 1), 1), pic(std::max<sz>(m, 1), 1), wc(std::max<sz>(m, 1), 1);
	if(q >= 1 && q <= p) {
		for(sz r = 0, k = 0; r < p; ++r)
			if(r + 1 != q) nc[k++] = na[r];
		for(sz r = 0, k = 0; r < p; ++r)
			if(pia[r] != q && k < m) pic[k++] = pia[r] > q ? pia[r] - 1 : pia[r];
	}
	for(sz r = 0; r < m; ++r) pic[r] = std::min<sz>(std::max<sz>(pic[r], 0), m + 1);
	if(m > 0 && rng() % 8 == 0) nc[rng() % m] = rng() % 3;
	if(m > 0) {
		std::vector<int> seen(m + 2, 0);
		bool ok = true;
		for(sz r = 0; r < m; ++r) {
			if(pic[r] < 1 || pic[r] > m || seen[pic[r]]) ok = false;
			else seen[pic[r]] = 1;
		}
		if(ok) wc = make_strides(std::vector<sz>(nc.begin(), nc.begin() + m), std::vector<sz>(pic.begin(), pic.begin() + m), 1, rng);
	}
	sz const nb = (q >= 1 && q <= p) ? (rng() % 10 == 0 ? na[q - 1] + 1 : na[q - 1]) : 1;
	float dummy = 0.f;
	ttvx::Check const ck = ttvx::validate(q, p, &dummy, na.data(), wa.data(), pia.data(), &dummy, &nb, &dummy, nc.data(),
	                                      wc.data(), pic.data());
	int const want = expected_status(q, p, na, wa, pia, nb, nc, wc, pic);
	if(want == TTVX_ERR_ARGUMENT)
		CHECK(ck.status == TTVX_ERR_ARGUMENT, "accepted invalid arguments q=%zu p=%zu na=%s wa=%s pia=%s nb=%zu", q, p,
		      str(na).c_str(), str(wa).c_str(), str(pia).c_str(), nb);
	else
		CHECK(ck.status == TTVX_OK || ck.status == TTVX_ERR_ARGUMENT, "status %d", ck.status);
	CHECK((ck.status == TTVX_OK) == ck.message.empty(), "message and status disagree: %d '%s'", ck.status,
	      ck.message.c_str());
	if(q >= 1 && q <= p && p >= 1) {
		ttvx::Check const ca = ttvx::validate_a(q, p, na.data(), wa.data(), pia.data());
		if(ca.status == TTVX_OK) {
			ttvx::Fold const f = ttvx::fold(q, p, na.data(), wa.data(), pia.data(), nullptr, nullptr);
			CHECK(f.K == static_cast<int64_t>(na[q - 1]), "K of a validated problem");
		}
	}
	// null pointers: refused, never dereferenced
	CHECK(ttvx::validate_a(1, 1, nullptr, wa.data(), pia.data()).status == TTVX_ERR_ARGUMENT, "null na");
	CHECK(ttvx::validate(1, 1, &dummy, na.data(), wa.data(), pia.data(), nullptr, &nb, &dummy, nc.data(), wc.data(),
	                     pic.data()).status == TTVX_ERR_ARGUMENT, "null b");
}

int main(int argc, char** argv)
{
	int const n = argc > 1 ? std::atoi(argv[1]) : 4000;
	std::mt19937 rng(20261203u);
	for(int t = 0; t < n; ++t) fold_case(rng);
	for(int t = 0; t < 4 * n; ++t) validate_case(rng);
	std::printf("%d checks, %d failures\n", g_checks, g_fail);
	return g_fail == 0 ? 0 : 1;
}
