// ref_shim.cpp -- C entry points around the reference's own headers (TEST INFRASTRUCTURE ONLY).
//
// Compiled by oracle/Makefile (target `ref`) against the reference sources where they lie; the
// result goes to oracle/_ref/, which is never committed.  tests/ use it to pin the oracle and the
// product's metadata layer against the real reference; bench.py may time it as a CPU baseline.
// No reference source is copied here: this file only includes <tlib/ttv.h> and forwards.
#include <tlib/ttv.h>

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{
std::string g_err;

using namespace tlib::ttv;

// The 17 policy combinations exercised by the reference's test suite
// (test/src/gtest_tlib_ttv.cpp:192-425), numbered 0..16.
template<class T>
void run_policy(int id, std::size_t q, std::size_t p, const T* a, const std::size_t* na, const std::size_t* wa,
                const std::size_t* pia, const T* b, const std::size_t* nb, T* c, const std::size_t* nc,
                const std::size_t* wc, const std::size_t* pic)
{
	namespace ep = execution_policy;
	namespace sp = slicing_policy;
	namespace fp = fusion_policy;
#define TTV_CALL(E, S, F) ttv(E, S, F, q, p, a, na, wa, pia, b, nb, c, nc, wc, pic)
	switch(id) {
		case 0:  TTV_CALL(ep::seq,          sp::subtensor, fp::none);  break;
		case 1:  TTV_CALL(ep::seq_blas,     sp::subtensor, fp::none);  break;
		case 2:  TTV_CALL(ep::par,          sp::subtensor, fp::none);  break;
		case 3:  TTV_CALL(ep::par_loop,     sp::subtensor, fp::none);  break;
		case 4:  TTV_CALL(ep::par_loop,     sp::subtensor, fp::all);   break;
		case 5:  TTV_CALL(ep::par_task,     sp::subtensor, fp::none);  break;
		case 6:  TTV_CALL(ep::par_taskloop, sp::subtensor, fp::none);  break;
		case 7:  TTV_CALL(ep::par_blas,     sp::subtensor, fp::all);   break;
		case 8:  TTV_CALL(ep::seq,          sp::slice,     fp::none);  break;
		case 9:  TTV_CALL(ep::seq_blas,     sp::slice,     fp::none);  break;
		case 10: TTV_CALL(ep::par,          sp::slice,     fp::none);  break;
		case 11: TTV_CALL(ep::par_loop,     sp::slice,     fp::none);  break;
		case 12: TTV_CALL(ep::par_task,     sp::slice,     fp::none);  break;
		case 13: TTV_CALL(ep::par_taskloop, sp::slice,     fp::none);  break;
		case 14: TTV_CALL(ep::par_loop,     sp::slice,     fp::outer); break;
		case 15: TTV_CALL(ep::par_loop,     sp::slice,     fp::all);   break;
		case 16: TTV_CALL(ep::par_blas,     sp::slice,     fp::all);   break;
		default: throw std::runtime_error("ref_shim: unknown policy id");
	}
#undef TTV_CALL
}

template<class T>
int ttv_any(int id, std::size_t q, std::size_t p, const T* a, const std::size_t* na, const std::size_t* wa,
            const std::size_t* pia, const T* b, const std::size_t* nb, T* c, const std::size_t* nc,
            const std::size_t* wc, const std::size_t* pic)
{
	try {
		run_policy<T>(id, q, p, a, na, wa, pia, b, nb, c, nc, wc, pic);
	} catch(std::exception const& e) {
		g_err = e.what();
		return 1;
	}
	return 0;
}

template<class F>
int guarded(F&& f)
{
	try {
		f();
	} catch(std::exception const& e) {
		g_err = e.what();
		return 1;
	}
	return 0;
}
} // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

// C must be zero-filled by the caller: the reference's non-BLAS slice kernels accumulate
// (matrix_times_vector.h:117-126).
int ref_ttv_f64(int id, std::size_t q, std::size_t p, const double* a, const std::size_t* na, const std::size_t* wa,
                const std::size_t* pia, const double* b, const std::size_t* nb, double* c, const std::size_t* nc,
                const std::size_t* wc, const std::size_t* pic)
{
	return ttv_any<double>(id, q, p, a, na, wa, pia, b, nb, c, nc, wc, pic);
}

int ref_ttv_f32(int id, std::size_t q, std::size_t p, const float* a, const std::size_t* na, const std::size_t* wa,
                const std::size_t* pia, const float* b, const std::size_t* nb, float* c, const std::size_t* nc,
                const std::size_t* wc, const std::size_t* pic)
{
	return ttv_any<float>(id, q, p, a, na, wa, pia, b, nb, c, nc, wc, pic);
}

// Metadata functions (for cross-checking the product's C++ headers).
int ref_output_shape(const std::size_t* na, std::size_t p, std::size_t q, std::size_t* out)
{
	return guarded([&] {
		auto v = tlib::ttv::detail::generate_output_shape(std::vector<std::size_t>(na, na + p), q);
		for(std::size_t i = 0; i < v.size(); ++i) out[i] = v[i];
	});
}

int ref_output_layout(const std::size_t* pia, std::size_t p, std::size_t q, std::size_t* out)
{
	return guarded([&] {
		auto v = tlib::ttv::detail::generate_output_layout(std::vector<std::size_t>(pia, pia + p), q);
		for(std::size_t i = 0; i < v.size(); ++i) out[i] = v[i];
	});
}

int ref_strides(const std::size_t* n, const std::size_t* pi, std::size_t p, std::size_t* out)
{
	return guarded([&] {
		auto v = tlib::ttv::detail::generate_strides(std::vector<std::size_t>(n, n + p), std::vector<std::size_t>(pi, pi + p));
		for(std::size_t i = 0; i < v.size(); ++i) out[i] = v[i];
	});
}

int ref_k_order_layout(std::size_t p, std::size_t k, std::size_t* out)
{
	return guarded([&] {
		auto v = tlib::ttv::detail::generate_k_order_layout(p, k);
		for(std::size_t i = 0; i < v.size(); ++i) out[i] = v[i];
	});
}

// Shape predicates as a bit mask: 1 valid, 2 scalar, 4 vector, 8 matrix, 16 tensor.
int ref_shape_class(const std::size_t* n, std::size_t p)
{
	using namespace tlib::ttv::detail;
	int m = 0;
	if(is_valid_shape(n, n + p)) m |= 1;
	if(is_scalar(n, n + p)) m |= 2;
	if(is_vector(n, n + p)) m |= 4;
	if(is_matrix(n, n + p)) m |= 8;
	if(is_tensor(n, n + p)) m |= 16;
	return m;
}

int ref_is_valid_layout(const std::size_t* pi, std::size_t p)
{
	return tlib::ttv::detail::is_valid_layout(pi, pi + p) ? 1 : 0;
}

int ref_is_valid_strides(const std::size_t* pi, std::size_t p, const std::size_t* w)
{
	int r = -1;
	guarded([&] { r = tlib::ttv::detail::is_valid_strides(pi, pi + p, w) ? 1 : 0; });
	return r;
}

int ref_inverse_mode(const std::size_t* pi, std::size_t p, std::size_t mode)
{
	int r = -1;
	guarded([&] { r = static_cast<int>(tlib::ttv::detail::inverse_mode(pi, pi + p, mode)); });
	return r;
}

unsigned ref_number_cores() { return tlib::ttv::detail::get_number_cores(); }

} // extern "C"
