// The drop-in C++ interfaces (include/tlib/ttv.h) on the GPU, written like the reference's
// test/src/gtest_tlib_ttv.cpp:27-425: for every policy-tag combination the reference tests,
// ranks 2..4, shapes {2,4,8}^p, every layout and every q, A is filled with the known-answer
// pattern (consecutive integers along each mode-q fiber, :27-71) and b = 1, so that
// C[j] = (nq^2 (2(j+1)-1) + nq)/2 exactly (:132).  Host std::vector buffers go through the
// library's synchronous entry points (staged to the GPU).  Also runs the three example
// programs' computations (example/interface1.cpp, interface2.cpp, interface3.cpp) and the
// argument checks.  Exit status 0 = all checks passed.
#include <tlib/device_tensor.h>
#include <tlib/ttv.h>

#include <algorithm>
#include <cstdio>
#include <functional>
#include <numeric>
#include <string>
#include <vector>

using namespace tlib::ttv;
using vec = std::vector<std::size_t>;

static long g_fail = 0, g_checks = 0;
#define CHECK(cond, ...)                                                          \
	do {                                                                          \
		++g_checks;                                                               \
		if(!(cond)) {                                                             \
			if(++g_fail < 20) { std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); std::fprintf(stderr, __VA_ARGS__); std::fprintf(stderr, "\n"); } \
		}                                                                         \
	} while(0)

// fills every mode-q fiber with consecutive integers, fibers in C's storage order
template<class T>
static void known_init(std::size_t q, std::vector<T>& a, vec const& na, vec const& wa, vec const& pia)
{
	std::size_t const p = na.size();
	std::size_t const qpos = detail::inverse_mode(pia.begin(), pia.end(), q) - 1;
	std::size_t nfib = 1;
	for(std::size_t r = 0; r < p; ++r)
		if(r + 1 != q) nfib *= na[r];
	std::size_t counter = 0;
	for(std::size_t f = 0; f < nfib; ++f) {
		std::size_t rest = f, off = 0;
		for(std::size_t s = 0; s < p; ++s) {
			if(s == qpos) continue;
			std::size_t const m = pia[s] - 1;
			off += (rest % na[m]) * wa[m];
			rest /= na[m];
		}
		for(std::size_t k = 0; k < na[q - 1]; ++k) a[off + k * wa[q - 1]] = static_cast<T>(++counter);
	}
}

static std::vector<vec> shapes_of(std::size_t rank)
{
	std::vector<vec> out;
	std::vector<std::size_t> idx(rank, 0);
	for(;;) {
		vec s(rank);
		for(std::size_t r = 0; r < rank; ++r) s[r] = std::size_t{2} << idx[r];
		out.push_back(s);
		std::size_t r = 0;
		while(r < rank && ++idx[r] == 3) idx[r++] = 0;
		if(r == rank) break;
	}
	return out;
}

template<class T, class E, class S, class F>
static void check_policy(E ep, S sp, F fp, const char* name)
{
	for(std::size_t rank = 2; rank <= 4; ++rank) {
		vec perm(rank);
		std::iota(perm.begin(), perm.end(), 1);
		std::vector<vec> layouts;
		do layouts.push_back(perm); while(std::next_permutation(perm.begin(), perm.end()));
		for(auto const& na : shapes_of(rank)) {
			std::size_t const nn = std::accumulate(na.begin(), na.end(), std::size_t{1}, std::multiplies<>());
			std::vector<T> a(nn);
			for(auto const& pia : layouts) {
				auto wa = detail::generate_strides(na, pia);
				for(std::size_t q = 1; q <= rank; ++q) {
					known_init(q, a, na, wa, pia);
					std::vector<T> b(na[q - 1], T{1});
					auto pic = detail::generate_output_layout(pia, q);
					auto nc  = detail::generate_output_shape(na, q);
					auto wc  = detail::generate_strides(nc, pic);
					vec nb   = {b.size()};
					std::size_t const nnc = nn / na[q - 1];
					std::vector<T> c(nnc, T{-1});
					ttv(ep, sp, fp, q, rank, a.data(), na.data(), wa.data(), pia.data(), b.data(), nb.data(),
					    c.data(), nc.data(), wc.data(), pic.data());
					std::size_t const nq = na[q - 1];
					bool ok = true;
					for(std::size_t j = 0; j < nnc; ++j) {
						T const expect = static_cast<T>((nq * nq * (2 * (j + 1) - 1) + nq) / 2);
						ok = ok && c[j] == expect;
					}
					CHECK(ok, "%s rank %zu q %zu", name, rank, q);
				}
			}
		}
	}
}

template<class T>
static void all_policies()
{
	namespace ep = execution_policy;
	namespace sp = slicing_policy;
	namespace fp = fusion_policy;
	check_policy<T>(ep::seq, sp::subtensor, fp::none, "seq/subtensor/none");
	check_policy<T>(ep::seq_blas, sp::subtensor, fp::none, "seq_blas/subtensor/none");
	check_policy<T>(ep::par, sp::subtensor, fp::none, "par/subtensor/none");
	check_policy<T>(ep::par_loop, sp::subtensor, fp::none, "par_loop/subtensor/none");
	check_policy<T>(ep::par_loop, sp::subtensor, fp::all, "par_loop/subtensor/all");
	check_policy<T>(ep::par_task, sp::subtensor, fp::none, "par_task/subtensor/none");
	check_policy<T>(ep::par_taskloop, sp::subtensor, fp::none, "par_taskloop/subtensor/none");
	check_policy<T>(ep::par_blas, sp::subtensor, fp::all, "par_blas/subtensor/all");
	check_policy<T>(ep::seq, sp::slice, fp::none, "seq/slice/none");
	check_policy<T>(ep::seq_blas, sp::slice, fp::none, "seq_blas/slice/none");
	check_policy<T>(ep::par, sp::slice, fp::none, "par/slice/none");
	check_policy<T>(ep::par_loop, sp::slice, fp::none, "par_loop/slice/none");
	check_policy<T>(ep::par_task, sp::slice, fp::none, "par_task/slice/none");
	check_policy<T>(ep::par_taskloop, sp::slice, fp::none, "par_taskloop/slice/none");
	check_policy<T>(ep::par_loop, sp::slice, fp::outer, "par_loop/slice/outer");
	check_policy<T>(ep::par_loop, sp::slice, fp::all, "par_loop/slice/all");
	check_policy<T>(ep::par_blas, sp::slice, fp::all, "par_blas/slice/all");
	check_policy<T>(ep::par_blas_loop, sp::slice, fp::all, "par_blas_loop/slice/all");
}

// test/src/gtest_tlib_mtv.cpp:17-180: the 2-D slice kernels on (2..2^steps)^2 shapes, row- and
// column-major fills 1..M*N, b = 1, row sums fn(i) - fn(i-1) with fn(i) = i*n*(i*n+1)/2.
template<class T, class F1, class F2>
static void check_mtv(F1 col_fn, F2 row_fn, std::size_t steps, const char* name)
{
	for(std::size_t e0 = 0; e0 < steps; ++e0)
		for(std::size_t e1 = 0; e1 < steps; ++e1) {
			std::size_t const m = std::size_t{2} << e1, n = std::size_t{2} << e0;
			std::vector<T> a(m * n), b(n, T{1}), c(m, T{0});
			auto fn = [n](std::size_t i) { return (i * n * (i * n + 1)) / 2; };
			auto expect_ok = [&]() {
				bool ok = true;
				for(std::size_t i = 1; i <= m; ++i) ok = ok && c[i - 1] == static_cast<T>(fn(i) - fn(i - 1));
				return ok;
			};
			for(std::size_t j = 0; j < n; ++j)          // column-major fill
				for(std::size_t i = 0; i < m; ++i) a[i + j * m] = static_cast<T>(i * n + j + 1);
			col_fn(a.data(), b.data(), c.data(), m, n, m);
			CHECK(expect_ok(), "%s column-major %zux%zu", name, m, n);
			std::fill(c.begin(), c.end(), T{0});
			for(std::size_t i = 0; i < m; ++i)          // row-major fill
				for(std::size_t j = 0; j < n; ++j) a[j + i * n] = static_cast<T>(j + 1 + i * n);
			row_fn(a.data(), b.data(), c.data(), m, n, n);
			CHECK(expect_ok(), "%s row-major %zux%zu", name, m, n);
		}
}

static void mtv_tests()
{
	using namespace tlib::ttv::detail;
	check_mtv<double>(gemv_col<double, std::size_t>, gemv_row<double, std::size_t>, 10, "Gemv");
	check_mtv<double>(gemv_col_parallel<double, std::size_t>, gemv_row_parallel<double, std::size_t>, 10, "GemvParallel");
	check_mtv<float>(gemv_col_blas<float, std::size_t>, gemv_row_blas<float, std::size_t>, 8, "GemvBLAS");
	// the non-BLAS column kernel adds into c; the BLAS one overwrites
	std::vector<double> a = {1, 2, 3, 4}, b = {1, 1}, c = {10, 20};
	gemv_col<double, std::size_t>(a.data(), b.data(), c.data(), 2, 2, 2);
	CHECK(c[0] == 14 && c[1] == 26, "gemv_col accumulates");
	gemv_col_blas<double, std::size_t>(a.data(), b.data(), c.data(), 2, 2, 2);
	CHECK(c[0] == 4 && c[1] == 6, "gemv_col_blas overwrites");
	double d = 0;
	dot<double, std::size_t>(a.data(), a.data(), &d, 4);
	CHECK(d == 30, "dot");

	// the same kernels on device-resident operands (the accumulating forms add on the GPU)
	{
		void *da = nullptr, *db = nullptr, *dc = nullptr;
		std::vector<double> ha = {1, 2, 3, 4}, hb = {1, 1}, hc = {10, 20}, back(2, 0.0);
		CHECK(ttvx_device_alloc(4 * sizeof(double), &da) == TTVX_OK && ttvx_device_alloc(2 * sizeof(double), &db) == TTVX_OK &&
		          ttvx_device_alloc(2 * sizeof(double), &dc) == TTVX_OK, "device alloc: %s", ttvx_last_error());
		ttvx_copy(da, ha.data(), 4 * sizeof(double));
		ttvx_copy(db, hb.data(), 2 * sizeof(double));
		ttvx_copy(dc, hc.data(), 2 * sizeof(double));
		auto* A = static_cast<double*>(da);
		auto* B = static_cast<double*>(db);
		auto* C = static_cast<double*>(dc);
		auto fetch = [&]() { ttvx_copy(back.data(), dc, 2 * sizeof(double)); };
		gemv_col<double, std::size_t>(A, B, C, 2, 2, 2);
		fetch();
		CHECK(back[0] == 14 && back[1] == 26, "device gemv_col accumulates: %g %g", back[0], back[1]);
		gemv_col_parallel<double, std::size_t>(A, B, C, 2, 2, 2);
		fetch();
		CHECK(back[0] == 18 && back[1] == 32, "device gemv_col_parallel accumulates: %g %g", back[0], back[1]);
		gemv_col_blas<double, std::size_t>(A, B, C, 2, 2, 2);
		fetch();
		CHECK(back[0] == 4 && back[1] == 6, "device gemv_col_blas overwrites: %g %g", back[0], back[1]);
		gemv_row<double, std::size_t>(A, B, C, 2, 2, 2);
		fetch();
		CHECK(back[0] == 3 && back[1] == 7, "device gemv_row: %g %g", back[0], back[1]);
		dot<double, std::size_t>(A, A, C, 4);
		fetch();
		CHECK(back[0] == 30, "device dot: %g", back[0]);
		ttvx_device_free(da);
		ttvx_device_free(db);
		ttvx_device_free(dc);
	}
	// the column-major closed forms again with every operand on the device, accumulating onto c = 1
	for(std::size_t e = 0; e < 6; ++e) {
		std::size_t const m = std::size_t{2} << e, n = std::size_t{4} << e;
		std::vector<float> ha(m * n), hb(n, 1.f), hc(m, 1.f), back(m);
		for(std::size_t j = 0; j < n; ++j)
			for(std::size_t i = 0; i < m; ++i) ha[i + j * m] = static_cast<float>(i * n + j + 1);
		void *da = nullptr, *db = nullptr, *dc = nullptr;
		ttvx_device_alloc(ha.size() * sizeof(float), &da);
		ttvx_device_alloc(n * sizeof(float), &db);
		ttvx_device_alloc(m * sizeof(float), &dc);
		ttvx_copy(da, ha.data(), ha.size() * sizeof(float));
		ttvx_copy(db, hb.data(), n * sizeof(float));
		ttvx_copy(dc, hc.data(), m * sizeof(float));
		gemv_col<float, std::size_t>(static_cast<float*>(da), static_cast<float*>(db), static_cast<float*>(dc), m, n, m);
		ttvx_copy(back.data(), dc, m * sizeof(float));
		auto fn = [n](std::size_t i) { return (i * n * (i * n + 1)) / 2; };
		bool ok = true;
		for(std::size_t i = 1; i <= m; ++i) ok = ok && back[i - 1] == static_cast<float>(fn(i) - fn(i - 1) + 1);
		CHECK(ok, "device gemv_col %zux%zu onto ones", m, n);
		ttvx_device_free(da);
		ttvx_device_free(db);
		ttvx_device_free(dc);
	}
}

// Device-resident tensors (include/tlib/device_tensor.h): the interface1 example with the data in
// HBM, a chain of three products that never leaves the GPU (every layout of an order-4 tensor,
// exact integer answers), moves, and the reference's error text.
static void device_tensors()
{
	{
		auto Ah = tensor<float>({4, 3, 2});
		auto Bh = tensor<float>({3, 1});
		std::iota(Ah.begin(), Ah.end(), 1);
		std::fill(Bh.begin(), Bh.end(), 1);
		device_tensor<float> A(Ah), B(Bh);
		auto C = A(2) * B;
		std::vector<float> expect = {15, 18, 21, 24, 51, 54, 57, 60};
		CHECK(C.to_host().data() == expect, "device interface1");
		CHECK((C.shape() == vec{4, 2}) && (C.layout() == vec{1, 2}), "device interface1 metadata");
		auto D = std::move(C);
		CHECK(C.data() == nullptr && D.to_host().data() == expect, "device tensor move");
	}
	vec const na = {8, 4, 6, 2};
	vec perm = {1, 2, 3, 4};
	do {
		auto Ah = tensor<double>(na, perm);
		std::iota(Ah.begin(), Ah.end(), 1.0);
		auto const wa = Ah.strides();
		device_tensor<double> A(Ah);
		std::vector<device_tensor<double>> ones;
		for(std::size_t m = 1; m < 4; ++m) {
			auto o = tensor<double>({na[m], 1});
			o = 1.0;
			ones.emplace_back(o);
		}
		// modes 4, 3, 2 in turn: C[i] = sum over j,k,l of A(i,j,k,l)
		auto C = ttv(2, ttv(3, ttv(4, A, ones[2]), ones[1]), ones[0]);
		auto Ch = C.to_host();
		bool ok = Ch.shape() == vec{8};
		for(std::size_t i = 0; i < na[0] && ok; ++i) {
			double want = 0;
			for(std::size_t j = 0; j < na[1]; ++j)
				for(std::size_t k = 0; k < na[2]; ++k)
					for(std::size_t l = 0; l < na[3]; ++l)
						want += 1.0 + double(i * wa[0] + j * wa[1] + k * wa[2] + l * wa[3]);
			ok = Ch.data()[i] == want;
		}
		CHECK(ok, "device chain layout %zu%zu%zu%zu", perm[0], perm[1], perm[2], perm[3]);
	} while(std::next_permutation(perm.begin(), perm.end()));
	{
		device_tensor<double> A(tensor<double>({4, 3, 2})), B(tensor<double>({5, 1}));
		std::string msg;
		try {
			auto C = A(2) * B;
			(void)C;
		} catch(std::runtime_error const& e) {
			msg = e.what();
		}
		CHECK(msg == "Error in tlib::tensor_times_vector: contraction dimension of A and B are not equal.", "device: %s", msg.c_str());
	}
}

static void examples()
{
	// interface1.cpp: auto C1 = A(mode) * B;
	{
		auto A = tensor<float>({4, 3, 2});
		auto B = tensor<float>({3, 1});
		std::iota(A.begin(), A.end(), 1);
		std::fill(B.begin(), B.end(), 1);
		auto C1 = A(2) * B;
		std::vector<float> expect = {15, 18, 21, 24, 51, 54, 57, 60};
		CHECK(C1.data() == expect, "interface1");
		CHECK((C1.shape() == vec{4, 2}) && (C1.layout() == vec{1, 2}), "interface1 metadata");
	}
	// interface2.cpp: the high-level function with two policy triples
	{
		auto A = tensor<float>({4, 3, 2});
		auto B = tensor<float>({3, 1});
		std::iota(A.begin(), A.end(), 1);
		std::fill(B.begin(), B.end(), 1);
		auto C1 = ttv(2, A, B, execution_policy::seq, slicing_policy::subtensor, fusion_policy::none);
		auto C2 = ttv(2, A, B, execution_policy::par_loop, slicing_policy::slice, fusion_policy::all);
		auto C3 = ttv(2, A, B);
		std::vector<float> expect = {15, 18, 21, 24, 51, 54, 57, 60};
		CHECK(C1.data() == expect && C2.data() == expect && C3.data() == expect, "interface2");
	}
	// interface3.cpp: the C-like low-level function on std::vector buffers
	{
		auto na = vec{4, 3, 2};
		std::size_t p = na.size(), q = 2;
		auto pia = detail::generate_k_order_layout(p, std::size_t{1});
		auto wa  = detail::generate_strides(na, pia);
		auto nb  = vec{na[q - 1]};
		auto nc  = detail::generate_output_shape(na, q);
		auto pic = detail::generate_output_layout(pia, q);
		auto wc  = detail::generate_strides(nc, pic);
		std::vector<float> A(24), B(3, 1.0f), C1(8, 0.0f), C2(8, 0.0f);
		std::iota(A.begin(), A.end(), 1.0f);
		ttv(execution_policy::seq, slicing_policy::slice, fusion_policy::none, q, p, A.data(), na.data(), wa.data(),
		    pia.data(), B.data(), nb.data(), C1.data(), nc.data(), wc.data(), pic.data());
		ttv(execution_policy::par_loop, slicing_policy::subtensor, fusion_policy::all, q, p, A.data(), na.data(),
		    wa.data(), pia.data(), B.data(), nb.data(), C2.data(), nc.data(), wc.data(), pic.data());
		std::vector<float> expect = {15, 18, 21, 24, 51, 54, 57, 60};
		CHECK(C1 == expect && C2 == expect, "interface3");
		CHECK(wa == (vec{1, 4, 12}) && wc == (vec{1, 4}), "interface3 strides");
	}
	// argument errors surface as std::runtime_error with the reference's text
	{
		auto A = tensor<double>({4, 3, 2});
		auto B = tensor<double>({5, 1});
		std::string msg;
		try {
			ttv(2, A, B);
		} catch(std::runtime_error const& e) {
			msg = e.what();
		}
		CHECK(msg == "Error in tlib::tensor_times_vector: contraction dimension of A and B are not equal.", "%s", msg.c_str());
		std::string msg2;
		try {
			auto C = A(4) * B;
			(void)C;
		} catch(std::runtime_error const& e) {
			msg2 = e.what();
		}
		CHECK(msg2 == "Error in tlib::tensor: specified contraction mode should be greater than one and equal to or less than the order.", "%s", msg2.c_str());
	}
}

// Device errors forced with the fault-injection hook: the drop-in interfaces throw
// std::runtime_error with the library's message, and the next product works.
static void injected_failures()
{
	auto A = tensor<float>({4, 3, 2});
	auto B = tensor<float>({3, 1});
	std::iota(A.begin(), A.end(), 1);
	std::fill(B.begin(), B.end(), 1);
	std::vector<float> const expect = {15, 18, 21, 24, 51, 54, 57, 60};
	for(int site : {TTVX_FAULT_ALLOC, TTVX_FAULT_COPY, TTVX_FAULT_LAUNCH}) {
		ttvx_debug_fail_next(site, 1);
		std::string msg;
		try {
			auto C = A(2) * B;
			(void)C;
		} catch(std::runtime_error const& e) {
			msg = e.what();
		}
		CHECK(msg.find("HIP error") != std::string::npos, "injected failure %d: '%s'", site, msg.c_str());
		auto C = A(2) * B;
		CHECK(C.data() == expect, "product after injected failure %d", site);
	}
}

int main()
{
	examples();
	injected_failures();
	mtv_tests();
	device_tensors();
	all_policies<double>();
	check_policy<float>(execution_policy::par_loop, slicing_policy::subtensor, fusion_policy::all, "float default");
	std::printf("%ld checks, %ld failures\n", g_checks, g_fail);
	return g_fail == 0 ? 0 : 1;
}
