// Device-resident counterpart of the reference's tensor<T> (include/tlib/detail/tensor.h:36-114)
// and of its high-level interfaces (include/tlib/ttv.h:99-127): the elements live in the GPU's
// HBM, so a chain of products (or repeated products with the same A) moves no data over PCIe.
//
//   tlib::ttv::device_tensor<float> A(host_tensor);        // upload once
//   tlib::ttv::device_tensor<float> b(host_vector);
//   auto C = A(2) * b;                                     // runs on the GPU, C stays there
//   tlib::ttv::tensor<float> Ch = C.to_host();             // download when needed
//
// Same shape/layout/strides rules and messages as tensor<T>; move-only (one owner per buffer).
// The products are synchronous, like the reference's; memory comes from ttvx_device_alloc.
#pragma once

#include <cstddef>
#include <functional>
#include <numeric>
#include <stdexcept>
#include <type_traits>
#include <utility>
#include <vector>

#include "../ttvx.h"
#include "detail/layout.h"
#include "detail/shape.h"
#include "detail/strides.h"
#include "detail/tags.h"
#include "detail/tensor.h"

namespace tlib::ttv
{

template<class value_t>
class device_tensor;

template<class _value_t>
struct device_tensor_view
{
public:
	using value_t  = _value_t;
	using tensor_t = device_tensor<value_t>;
	device_tensor_view() = delete;
	device_tensor_view(device_tensor_view const&) = delete;
	device_tensor_view& operator=(device_tensor_view const&) = delete;
	inline tensor_t const& get_tensor() const { return _tensor; }
	inline std::size_t contraction_mode() const { return _q; }
private:
	friend tensor_t;
	device_tensor_view(tensor_t const& t, std::size_t q) : _tensor(t), _q(q) {}
	tensor_t const& _tensor;
	std::size_t _q;
};

template<class value_t>
class device_tensor
{
	static_assert(std::is_same_v<value_t, float> || std::is_same_v<value_t, double>,
	              "tlib::ttv::device_tensor supports float and double elements.");
	using shape_t  = std::vector<std::size_t>;
	using layout_t = std::vector<std::size_t>;

	static void check(int status)
	{
		if(status != TTVX_OK) throw std::runtime_error(ttvx_last_error());
	}

public:
	device_tensor() = delete;
	device_tensor(device_tensor const&) = delete;
	device_tensor& operator=(device_tensor const&) = delete;

	// uninitialised device storage of the given shape and layout
	device_tensor(shape_t const& n, layout_t const& pi) : _n(n), _pi(pi)
	{
		if(n.size() != pi.size())
			throw std::runtime_error("Error in tlib::tensor: shape vector and layout vector must have the same length.");
		if(!detail::is_valid_shape(n.begin(), n.end()))
			throw std::runtime_error("Error in tlib::tensor: shape vector of tensor is not valid.");
		if(!detail::is_valid_layout(pi.begin(), pi.end()))
			throw std::runtime_error("Error in tlib::tensor: layout vector of tensor is not valid.");
		_size = std::accumulate(n.begin(), n.end(), std::size_t{1}, std::multiplies<>());
		void* p = nullptr;
		check(ttvx_device_alloc(_size * sizeof(value_t), &p));
		_data = static_cast<value_t*>(p);
	}

	explicit device_tensor(shape_t const& n) : device_tensor(n, detail::generate_k_order_layout(n.size(), std::size_t{1})) {}

	// upload of a host tensor (same shape and layout)
	explicit device_tensor(tensor<value_t> const& host) : device_tensor(host.shape(), host.layout())
	{
		upload(host);
	}

	device_tensor(device_tensor&& o) noexcept
		: _n(std::move(o._n)), _pi(std::move(o._pi)), _size(o._size), _data(std::exchange(o._data, nullptr))
	{
	}

	device_tensor& operator=(device_tensor&& o) noexcept
	{
		if(this != &o) {
			(void)ttvx_device_free(_data);
			_n = std::move(o._n);
			_pi = std::move(o._pi);
			_size = o._size;
			_data = std::exchange(o._data, nullptr);
		}
		return *this;
	}

	~device_tensor() { (void)ttvx_device_free(_data); }

	void upload(tensor<value_t> const& host)
	{
		if(host.shape() != _n || host.layout() != _pi)
			throw std::runtime_error("Error in tlib::device_tensor: host tensor has another shape or layout.");
		check(ttvx_copy(_data, host.data().data(), _size * sizeof(value_t)));
	}

	void download(tensor<value_t>& host) const
	{
		if(host.shape() != _n || host.layout() != _pi)
			throw std::runtime_error("Error in tlib::device_tensor: host tensor has another shape or layout.");
		check(ttvx_copy(host.data().data(), _data, _size * sizeof(value_t)));
	}

	tensor<value_t> to_host() const
	{
		tensor<value_t> host(_n, _pi);
		download(host);
		return host;
	}

	device_tensor_view<value_t> operator()(std::size_t contraction_mode) const
	{
		if(1ul > contraction_mode || contraction_mode > this->order())
			throw std::runtime_error("Error in tlib::tensor: specified contraction mode should be greater than one and equal to or less than the order.");
		return device_tensor_view<value_t>(*this, contraction_mode);
	}

	value_t const* data() const { return _data; }
	value_t* data() { return _data; }
	auto const& shape() const { return _n; }
	auto const& layout() const { return _pi; }
	auto strides() const { return detail::generate_strides(_n, _pi); }
	auto order() const { return _n.size(); }
	auto size() const { return _size; }

private:
	shape_t _n;
	layout_t _pi;
	std::size_t _size = 0;
	value_t* _data = nullptr;
};

// C = A x_q b with A, b and C in device memory (reference include/tlib/ttv.h:99-114); the policy
// tags are accepted as hints, as for the host tensors.
template<class value_t, class execution_t, class slicing_t, class fusion_t>
inline device_tensor<value_t> ttv(std::size_t q, device_tensor<value_t> const& a, device_tensor<value_t> const& b,
                                  execution_t ep, slicing_t sp, fusion_t fp)
{
	auto c = device_tensor<value_t>(detail::generate_output_shape(a.shape(), q),
	                                 detail::generate_output_layout(a.layout(), q));
	auto const wa = a.strides();
	auto const wc = c.strides();
	int const e = detail::policy_code(ep), s = detail::policy_code(sp), f = detail::policy_code(fp);
	int status;
	if constexpr(std::is_same_v<value_t, float>)
		status = ttvx_ttv_f32(e, s, f, q, a.order(), a.data(), a.shape().data(), wa.data(), a.layout().data(),
		                      b.data(), b.shape().data(), c.data(), c.shape().data(), wc.data(), c.layout().data());
	else
		status = ttvx_ttv_f64(e, s, f, q, a.order(), a.data(), a.shape().data(), wa.data(), a.layout().data(),
		                      b.data(), b.shape().data(), c.data(), c.shape().data(), wc.data(), c.layout().data());
	if(status != TTVX_OK) throw std::runtime_error(ttvx_last_error());
	return c;
}

template<class value_t>
inline device_tensor<value_t> ttv(std::size_t q, device_tensor<value_t> const& a, device_tensor<value_t> const& b)
{
	return ttv(q, a, b, execution_policy::par_loop, slicing_policy::subtensor, fusion_policy::all);
}

} // namespace tlib::ttv

// A(q) * b on device tensors (reference include/tlib/ttv.h:122-127)
template<class value_t>
inline auto operator*(tlib::ttv::device_tensor_view<value_t> const& a, tlib::ttv::device_tensor<value_t> const& b)
{
	return tlib::ttv::ttv(a.contraction_mode(), a.get_tensor(), b);
}
