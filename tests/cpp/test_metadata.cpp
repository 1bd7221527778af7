// Metadata layer of the drop-in headers (include/tlib/detail) against the golden tables of the
// reference's test suite, transcribed as data:
//   test/src/gtest_tlib_shape.cpp:17-232    shape classes and output shapes
//   test/src/gtest_tlib_layout.cpp:26-192   k-order layouts, invalid layouts, inverse, inverse_mode
//   test/src/gtest_tlib_strides.cpp:32-140  stride tables
//   test/src/gtest_tlib_workload.cpp:30-129 index-space division invariants
// Built with g++ (host code only) by tests/test_metadata.py.  With argument "dump" it reads
// queries from stdin and prints the product's answers, which the Python test compares with
// tests/golden/metadata_ref.json written from the compiled reference.
#include <tlib/detail/cases.h>
#include <tlib/detail/index.h>
#include <tlib/detail/layout.h>
#include <tlib/detail/shape.h>
#include <tlib/detail/strides.h>
#include <tlib/detail/workload_computation.h>

#include <algorithm>
#include <cstdio>
#include <iostream>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

using namespace tlib::ttv::detail;
using vec = std::vector<std::size_t>;

static int g_fail = 0, g_checks = 0;
#define CHECK(cond)                                                                      \
	do {                                                                                 \
		++g_checks;                                                                      \
		if(!(cond)) { ++g_fail; std::fprintf(stderr, "FAIL %s:%d %s\n", __FILE__, __LINE__, #cond); } \
	} while(0)

static void shape_tests()
{
	std::vector<vec> shapes = {{}, {0}, {1}, {2}, {0, 1}, {1, 0}, {1, 1}, {2, 1}, {3, 1}, {1, 2}, {1, 3}, {2, 2},
	                           {3, 3}, {0, 1, 1}, {1, 1, 0}, {1, 1, 1}, {1, 1, 2}, {1, 2, 1}, {1, 2, 2},
	                           {2, 1, 1}, {2, 1, 2}, {2, 2, 1}, {2, 2, 2}};
	auto in = [](std::vector<unsigned> const& v, unsigned i) { return std::find(v.begin(), v.end(), i) != v.end(); };
	std::vector<unsigned> scal = {2, 6, 15}, vecs = {3, 7, 8, 9, 10, 17, 19}, mats = {11, 12, 21}, tens = {16, 18, 20, 22},
	                      invalid = {0, 1, 4, 5, 13, 14};
	for(unsigned i = 0; i < shapes.size(); ++i) {
		auto const& s = shapes[i];
		CHECK(is_scalar(s.begin(), s.end()) == in(scal, i));
		CHECK(is_vector(s.begin(), s.end()) == in(vecs, i));
		CHECK(is_matrix(s.begin(), s.end()) == in(mats, i));
		CHECK(is_tensor(s.begin(), s.end()) == in(tens, i));
		CHECK(is_valid_shape(s.begin(), s.end()) == !in(invalid, i));
	}
	std::vector<vec> r1 = {{}, {}, {}, {}, {}, {}, {1}, {1}, {1}, {2}, {3}, {2}, {3}, {}, {}, {1, 1}, {1, 2}, {2, 1}, {2, 2}, {1, 1}, {1, 2}, {2, 1}, {2, 2}};
	std::vector<vec> r2 = {{}, {}, {}, {}, {}, {}, {1}, {2}, {3}, {1}, {1}, {2}, {3}, {}, {}, {1, 1}, {1, 2}, {1, 1}, {1, 2}, {2, 1}, {2, 2}, {2, 1}, {2, 2}};
	std::vector<vec> r3 = {{}, {}, {}, {}, {}, {}, {}, {}, {}, {}, {}, {}, {}, {}, {}, {1, 1}, {1, 1}, {1, 2}, {1, 2}, {2, 1}, {2, 1}, {2, 2}, {2, 2}};
	auto test_out = [&](std::vector<vec> const& refs, unsigned q) {
		for(unsigned i = 0; i < refs.size(); ++i) {
			if(!is_valid_shape(refs[i].begin(), refs[i].end()) || q > shapes[i].size()) continue;
			auto out = generate_output_shape(shapes[i], q);
			CHECK(out == refs[i]);
		}
	};
	test_out(r1, 1); test_out(r2, 2); test_out(r3, 3);
}

static void layout_tests()
{
	std::vector<std::vector<vec>> k_ref = {
		{{1}, {1, 2}, {1, 2, 3}, {1, 2, 3, 4}},
		{{1}, {2, 1}, {2, 1, 3}, {2, 1, 3, 4}},
		{{1}, {2, 1}, {3, 2, 1}, {3, 2, 1, 4}},
		{{1}, {2, 1}, {3, 2, 1}, {4, 3, 2, 1}}};
	for(std::size_t k = 1; k <= 4; ++k)
		for(std::size_t p = 1; p <= 4; ++p) {
			vec l(p);
			compute_k_order_layout(l.begin(), l.end(), k);
			CHECK(l == k_ref[k - 1][p - 1]);
			CHECK(is_valid_layout(l.begin(), l.end()));
		}
	for(std::size_t p = 1; p <= 4; ++p) {
		vec a(p), b(p);
		compute_first_order_layout(a.begin(), a.end());
		CHECK(a == k_ref[0][p - 1]);
		compute_last_order_layout(b.begin(), b.end());
		CHECK(b == k_ref[3][p - 1]);
	}
	std::vector<std::vector<unsigned>> bad = {{}, {0}, {0, 1}, {1, 0}, {0, 2}, {2, 0}, {2, 1, 0}, {3, 0, 2}, {3, 1, 4}, {1, 3, 4}, {1, 3, 5}};
	for(auto const& l : bad) CHECK(!is_valid_layout(l.begin(), l.end()));
	for(std::size_t p = 1; p <= 10; ++p)
		for(std::size_t k = 1; k <= p; ++k) {
			vec l(p), inv(p), inv2(p);
			compute_k_order_layout(l.begin(), l.end(), k);
			CHECK(is_valid_layout(l.begin(), l.end()));
			compute_inverse_layout(l.begin(), l.end(), inv.begin());
			CHECK(is_valid_layout(inv.begin(), inv.end()));
			compute_inverse_layout(inv.begin(), inv.end(), inv2.begin());
			CHECK(inv2 == l);
			for(std::size_t m = 1; m <= p; ++m) {
				auto r = inverse_mode(l.begin(), l.end(), m);
				CHECK(r >= 1 && r <= p && l[r - 1] == m);
			}
		}
}

static void stride_tests()
{
	{
		vec l1 = {1};
		auto v1 = generate_strides(vec{1}, l1);
		CHECK(v1.size() == 1 && v1[0] == 1);
		CHECK(is_valid_strides(l1.begin(), l1.end(), v1.begin()));
		vec l0 = {1, 2};
		auto v0 = generate_strides(vec{1, 1}, l0);
		CHECK(v0 == (vec{1, 1}));
	}
	for(auto const& e : std::vector<vec>{{1, 2}, {1, 4}, {1, 8}, {2, 1}, {4, 1}, {8, 1}})
		for(auto const& l : std::vector<vec>{{1, 2}, {2, 1}}) {
			CHECK(is_vector(e.begin(), e.end()));
			auto v = generate_strides(e, l);
			CHECK(v == (vec{1, 1}));
			CHECK(is_valid_strides(l.begin(), l.end(), v.begin()));
		}
	for(auto const& e : std::vector<vec>{{2, 2}, {2, 4}, {4, 2}, {4, 4}, {2, 8}, {8, 2}, {4, 8}, {8, 4}, {8, 8}}) {
		auto s1 = generate_strides(e, vec{1, 2});
		CHECK(s1 == (vec{1, e[0]}));
		auto s2 = generate_strides(e, vec{2, 1});
		CHECK(s2 == (vec{e[1], 1}));
	}
	auto table = [](std::vector<vec> const& ext, std::vector<vec> const& ref, vec const& l) {
		for(std::size_t i = 0; i < ext.size(); ++i) {
			CHECK(is_tensor(ext[i].begin(), ext[i].end()));
			auto s = generate_strides(ext[i], l);
			CHECK(s == ref[i]);
			CHECK(is_valid_strides(l.begin(), l.end(), s.begin()));
		}
	};
	table({{2, 1, 2}, {1, 2, 2}, {2, 2, 2}, {2, 4, 2}, {4, 4, 2}, {2, 4, 4}}, {{1, 2, 2}, {1, 1, 2}, {1, 2, 4}, {1, 2, 8}, {1, 4, 16}, {1, 2, 8}}, {1, 2, 3});
	table({{2, 1, 2}, {1, 2, 2}, {2, 2, 2}, {2, 4, 2}, {2, 4, 4}}, {{2, 2, 1}, {4, 2, 1}, {4, 2, 1}, {8, 2, 1}, {16, 4, 1}}, {3, 2, 1});
	table({{2, 1, 2}, {1, 2, 2}, {2, 2, 2}, {2, 4, 2}, {2, 2, 4}, {4, 2, 2}}, {{1, 1, 2}, {2, 1, 2}, {2, 1, 4}, {4, 1, 8}, {2, 1, 4}, {2, 1, 8}}, {2, 1, 3});
}

// gtest_tlib_workload.cpp:30-129: division of the index space
template<std::size_t rank>
static void division_tests(std::size_t init, std::size_t steps)
{
	std::vector<vec> shapes;
	{
		vec cur(rank, init);
		std::vector<std::size_t> idx(rank, 0);
		for(;;) {
			vec s(rank);
			for(std::size_t r = 0; r < rank; ++r) s[r] = init << idx[r];
			shapes.push_back(s);
			std::size_t r = 0;
			while(r < rank && ++idx[r] == steps) idx[r++] = 0;
			if(r == rank) break;
		}
	}
	vec perm(rank);
	std::iota(perm.begin(), perm.end(), 1);
	std::vector<vec> layouts;
	do layouts.push_back(perm); while(std::next_permutation(perm.begin(), perm.end()));
	for(auto const& n : shapes)
		for(auto const& l : layouts) {
			auto w = generate_strides(n, l);
			for(unsigned m = 1; m <= rank; ++m) {
				if(l[0] == m) continue;
				auto lay = divide_layout(l.data(), rank, m);
				auto l1 = lay.first, l2 = lay.second;
				auto ws = divide(w.data(), l.data(), rank, m);
				auto ns = divide(n.data(), l.data(), rank, m);
				auto v1 = generate_strides(ns.first, l1);
				auto v2 = generate_strides(ns.second, l2);
				auto nprod = std::accumulate(n.begin(), n.end(), std::size_t{1}, std::multiplies<>());
				for(std::size_t j = 0; j < nprod; ++j) {
					auto i = at_1(j, w, l);
					CHECK(at(i, w) == j);
					auto is = divide(i.data(), l.data(), rank, m);
					auto j1 = at(is.first, ws.first), j2 = at(is.second, ws.second);
					CHECK(j1 + j2 == j);
					auto k1 = at(is.first, v1), k2 = at(is.second, v2);
					auto ii1 = at_1(k1, v1, l1), ii2 = at_1(k2, v2, l2);
					CHECK(ii1 == is.first);
					CHECK(ii2 == is.second);
					CHECK(at_at_1(k1, v1, ws.first, l1) + at_at_1(k2, v2, ws.second, l2) == j);
				}
			}
		}
}

static void case_tests()
{
	vec p2a = {1, 2}, p2b = {2, 1}, p3 = {1, 2, 3}, p3l = {3, 2, 1};
	CHECK((is_case<2>(2, 1, p2a.data()) && is_case<3>(2, 2, p2a.data())));
	CHECK((is_case<4>(2, 1, p2b.data()) && is_case<5>(2, 2, p2b.data())));
	CHECK((is_case<6>(3, 1, p3.data()) && is_case<8>(3, 2, p3.data()) && is_case<7>(3, 3, p3.data())));
	CHECK((is_case<7>(3, 1, p3l.data()) && is_case<6>(3, 3, p3l.data())));
}

// Answers queries of the form  "<kind> <p> <ints...>"  one per line.
static int dump()
{
	std::string line;
	while(std::getline(std::cin, line)) {
		std::istringstream is(line);
		std::string kind;
		std::size_t p;
		is >> kind >> p;
		auto read = [&](std::size_t n) { vec v(n); for(auto& x : v) is >> x; return v; };
		std::ostringstream os;
		try {
			if(kind == "oshape") { auto n = read(p); std::size_t q; is >> q; for(auto x : generate_output_shape(n, q)) os << x << ' '; }
			else if(kind == "olayout") { auto l = read(p); std::size_t q; is >> q; for(auto x : generate_output_layout(l, q)) os << x << ' '; }
			else if(kind == "strides") { auto n = read(p); auto l = read(p); for(auto x : generate_strides(n, l)) os << x << ' '; }
			else if(kind == "korder") { std::size_t k; is >> k; for(auto x : generate_k_order_layout(p, k)) os << x << ' '; }
			else if(kind == "sclass") {
				auto n = read(p);
				int m = (is_valid_shape(n.begin(), n.end()) ? 1 : 0) | (is_scalar(n.begin(), n.end()) ? 2 : 0) |
				        (is_vector(n.begin(), n.end()) ? 4 : 0) | (is_matrix(n.begin(), n.end()) ? 8 : 0) |
				        (is_tensor(n.begin(), n.end()) ? 16 : 0);
				os << m;
			}
			else if(kind == "invmode") { auto l = read(p); std::size_t q; is >> q; os << inverse_mode(l.begin(), l.end(), q); }
		} catch(std::exception const& e) {
			os << "ERR";
		}
		std::cout << os.str() << '\n';
	}
	return 0;
}

int main(int argc, char** argv)
{
	if(argc > 1 && std::string(argv[1]) == "dump") return dump();
	shape_tests();
	layout_tests();
	stride_tests();
	division_tests<3>(2, 4);
	division_tests<4>(2, 3);
	case_tests();
	std::printf("%d checks, %d failures\n", g_checks, g_fail);
	return g_fail == 0 ? 0 : 1;
}
