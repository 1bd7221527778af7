// Planning and selection entry points of libttvx (host code: ttv_amd/csrc/selector.cpp, the
// configuration and feature code of runtime.hip, the planning calls of api.hip), run without a GPU
// on a library whose host code is built with AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/test_host_sanitizers.py).
//
// Over random problems (order 1-10, extents 1-300, every layout, the four storage types) and every
// selection mode (fixed rule; learned with random weights; the tuning table with rows naming random
// variants, valid for the problem or not; a cleared table; forced), the variant ttvx_plan_info reports must be
// one ttvx_variant_valid accepts, features must be finite, and repeated queries must agree (the
// decision cache).  Malformed arguments (null pointers, wrong weight counts, q out of range,
// variants out of range) must return an error status.
#include "../../include/ttvx.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <random>
#include <string>
#include <vector>

using sz = std::size_t;

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

struct Problem {
	int dtype;
	sz q;
	std::vector<sz> na, wa, pia;
};

static Problem random_problem(std::mt19937& rng)
{
	static const int dtypes[] = {TTVX_F32, TTVX_F64, TTVX_BF16, TTVX_FP8E4M3};
	Problem pr;
	pr.dtype = dtypes[rng() % 4];
	sz const p = 1 + rng() % 10;
	pr.na.resize(p);
	for(sz r = 0; r < p; ++r) pr.na[r] = rng() % 5 == 0 ? 1 : 1 + rng() % (p <= 3 ? 300 : 12);
	pr.pia.resize(p);
	for(sz r = 0; r < p; ++r) pr.pia[r] = r + 1;
	std::shuffle(pr.pia.begin(), pr.pia.end(), rng);
	pr.q = 1 + rng() % p;
	// dense strides along the layout, sometimes padded (views of a larger tensor)
	pr.wa.assign(p, 1);
	sz w = 1;
	bool const padded = rng() % 3 == 0;
	for(sz r = 0; r < p; ++r) {
		pr.wa[pr.pia[r] - 1] = w;
		w = w * pr.na[pr.pia[r] - 1] + (padded ? rng() % 3 : 0);
	}
	return pr;
}

static void check_problem(Problem const& pr, const char* mode)
{
	sz const p = pr.na.size();
	int64_t info[6] = {0, 0, 0, 0, 0, -1};
	int rc = ttvx_plan_info(pr.dtype, pr.q, p, pr.na.data(), pr.wa.data(), pr.pia.data(), info);
	CHECK(rc == TTVX_OK, "%s: plan_info failed: %s", mode, ttvx_last_error());
	if(rc != TTVX_OK) return;
	int const v = static_cast<int>(info[5]);
	CHECK(v >= 0 && v < ttvx_num_variants(), "%s: variant %d out of range", mode, v);
	if(v >= 0 && v < ttvx_num_variants())
		CHECK(ttvx_variant_valid(pr.dtype, pr.q, p, pr.na.data(), pr.wa.data(), pr.pia.data(), v) == 1,
		      "%s: picked variant %s cannot run the problem", mode, ttvx_variant_name(v));
	CHECK(info[1] == static_cast<int64_t>(pr.na[pr.q - 1]), "%s: K", mode);
	CHECK(info[4] >= 1 && info[4] <= 8, "%s: reference case %ld", mode, long(info[4]));
	int64_t again[6];
	ttvx_plan_info(pr.dtype, pr.q, p, pr.na.data(), pr.wa.data(), pr.pia.data(), again);
	CHECK(again[5] == info[5], "%s: repeated query picked another variant", mode);
	float feat[TTVX_NUM_FEATURES];
	rc = ttvx_features(pr.dtype, pr.q, p, pr.na.data(), pr.wa.data(), pr.pia.data(), feat);
	CHECK(rc == TTVX_OK, "%s: features failed", mode);
	for(int k = 0; k < TTVX_NUM_FEATURES && rc == TTVX_OK; ++k) CHECK(std::isfinite(feat[k]), "%s: feature %d", mode, k);
	int valid = 0;
	for(int u = 0; u < ttvx_num_variants(); ++u)
		valid += ttvx_variant_valid(pr.dtype, pr.q, p, pr.na.data(), pr.wa.data(), pr.pia.data(), u);
	CHECK(valid >= 1, "%s: no variant can run the problem", mode);
	CHECK(ttvx_variant_valid(pr.dtype, pr.q, p, pr.na.data(), pr.wa.data(), pr.pia.data(), ttvx_num_variants()) == 0,
	      "%s: out-of-range variant accepted", mode);
}

int main(int argc, char** argv)
{
	int const n = argc > 1 ? std::atoi(argv[1]) : 1500;
	std::mt19937 rng(20261204u);
	std::vector<Problem> probs;
	for(int t = 0; t < n; ++t) probs.push_back(random_problem(rng));

	CHECK(ttvx_set_selection(TTVX_SELECT_HEURISTIC, 0) == TTVX_OK, "heuristic mode");
	for(auto const& pr : probs) check_problem(pr, "heuristic");

	// learned selection with random weights (and rejected weight vectors)
	std::normal_distribution<float> nd(0.f, 1.f);
	std::vector<float> w(ttvx_selector_num_params());
	for(auto& x : w) x = nd(rng);
	for(int k = TTVX_NUM_FEATURES; k < 2 * TTVX_NUM_FEATURES; ++k) w[k] = std::fabs(w[k]) + 0.1f;    // inv_std > 0
	CHECK(ttvx_selector_set_weights(w.data(), w.size() - 1) != TTVX_OK, "short weight vector accepted");
	CHECK(ttvx_selector_set_weights(nullptr, w.size()) != TTVX_OK, "null weights accepted");
	CHECK(ttvx_selector_set_weights(w.data(), w.size()) == TTVX_OK, "weights refused: %s", ttvx_last_error());
	CHECK(ttvx_set_selection(TTVX_SELECT_LEARNED, 0) == TTVX_OK, "learned mode");
	for(auto const& pr : probs) check_problem(pr, "learned");

	// tuning table: rows for half of the problems with random (possibly invalid) variants
	std::vector<int64_t> rows;
	for(size_t t = 0; t < probs.size(); t += 2) {
		auto const& pr = probs[t];
		int64_t info[6];
		if(ttvx_plan_info(pr.dtype, pr.q, pr.na.size(), pr.na.data(), pr.wa.data(), pr.pia.data(), info) != TTVX_OK) continue;
		rows.insert(rows.end(), {static_cast<int64_t>(pr.dtype), info[0], info[1], info[2],
		                         static_cast<int64_t>(rng() % ttvx_num_variants())});
	}
	CHECK(ttvx_selector_set_table(nullptr, 3) != TTVX_OK, "null table accepted");
	for(int64_t bad : {int64_t{-1}, static_cast<int64_t>(ttvx_num_variants())}) {
		int64_t const row[5] = {TTVX_F32, 4, 5, 6, bad};
		CHECK(ttvx_selector_set_table(row, 1) != TTVX_OK, "table row naming variant %ld accepted", long(bad));
	}
	CHECK(ttvx_selector_set_table(rows.data(), rows.size() / 5) == TTVX_OK, "table refused: %s", ttvx_last_error());
	CHECK(ttvx_set_selection(TTVX_SELECT_AUTO, 0) == TTVX_OK, "auto mode");
	for(auto const& pr : probs) check_problem(pr, "auto+table");
	CHECK(ttvx_selector_set_table(nullptr, 0) == TTVX_OK, "clearing the table");
	CHECK(ttvx_selector_set_weights(nullptr, 0) == TTVX_OK, "unloading the weights");
	for(auto const& pr : probs) check_problem(pr, "auto, nothing loaded");

	// forced variants: plan_info reports the forced variant where it is valid
	for(int v = 0; v < ttvx_num_variants(); ++v) {
		CHECK(ttvx_set_selection(TTVX_SELECT_FORCED, v) == TTVX_OK, "forced %d", v);
		for(size_t t = 0; t < probs.size(); t += 7) {
			auto const& pr = probs[t];
			sz const p = pr.na.size();
			int64_t info[6];
			if(ttvx_plan_info(pr.dtype, pr.q, p, pr.na.data(), pr.wa.data(), pr.pia.data(), info) != TTVX_OK) continue;
			if(ttvx_variant_valid(pr.dtype, pr.q, p, pr.na.data(), pr.wa.data(), pr.pia.data(), v))
				CHECK(info[5] == v, "forced %s not taken", ttvx_variant_name(v));
		}
	}
	CHECK(ttvx_set_selection(TTVX_SELECT_FORCED, ttvx_num_variants()) != TTVX_OK, "forced variant out of range accepted");
	CHECK(ttvx_set_selection(42, 0) != TTVX_OK, "unknown mode accepted");
	ttvx_set_selection(TTVX_SELECT_AUTO, 0);

	// malformed planning queries
	sz na[3] = {4, 5, 6}, wa[3] = {1, 4, 20}, pia[3] = {1, 2, 3};
	int64_t info[6];
	float feat[TTVX_NUM_FEATURES];
	CHECK(ttvx_plan_info(TTVX_F32, 4, 3, na, wa, pia, info) != TTVX_OK, "q > p accepted");
	CHECK(ttvx_plan_info(TTVX_F32, 1, 3, nullptr, wa, pia, info) != TTVX_OK, "null na accepted");
	CHECK(ttvx_plan_info(TTVX_F32, 1, 3, na, wa, pia, nullptr) != TTVX_OK, "null out accepted");
	CHECK(ttvx_features(TTVX_F32, 1, 3, na, wa, pia, nullptr) != TTVX_OK, "null features accepted");
	CHECK(ttvx_features(99, 1, 3, na, wa, pia, feat) != TTVX_OK, "unknown dtype accepted");
	CHECK(std::string(ttvx_variant_name(-1)).empty() && std::string(ttvx_variant_name(ttvx_num_variants())).empty(),
	      "names out of range");

	std::printf("%d checks, %d failures\n", g_checks, g_fail);
	return g_fail == 0 ? 0 : 1;
}
