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

// What A(q) returns for a device tensor.
template<class value_t>
using device_tensor_view = detail::mode_view<device_tensor<value_t>>;

template<class value_t>
class device_tensor
{
	static_assert(std::is_same_v<value_t, float> || std::is_same_v<value_t, double>,
	              "tlib::ttv::device_tensor supports float and double elements.");

	static void check(int status)
	{
		if(status != TTVX_OK) throw std::runtime_error(ttvx_last_error());
	}

public:
	using value_type = value_t;
	using shape_t    = std::vector<std::size_t>;
	using layout_t   = std::vector<std::size_t>;

	device_tensor() = delete;
	device_tensor(device_tensor const&) = delete;
	device_tensor& operator=(device_tensor const&) = delete;

	// uninitialised device storage of the given shape and layout
	device_tensor(shape_t const& n, layout_t const& pi) : m_meta(n, pi) { allocate(); }
	explicit device_tensor(shape_t const& n) : m_meta(n) { allocate(); }

	// upload of a host tensor (same shape and layout)
	explicit device_tensor(tensor<value_t> const& host) : m_meta(host.shape(), host.layout())
	{
		allocate();
		upload(host);
	}

	device_tensor(device_tensor&& o) noexcept : m_meta(std::move(o.m_meta)), m_dev(std::exchange(o.m_dev, nullptr)) {}

	device_tensor& operator=(device_tensor&& o) noexcept
	{
		if(this != &o) {
			(void)ttvx_device_free(m_dev);
			m_meta = std::move(o.m_meta);
			m_dev = std::exchange(o.m_dev, nullptr);
		}
		return *this;
	}

	~device_tensor() { (void)ttvx_device_free(m_dev); }

	void upload(tensor<value_t> const& host)
	{
		require_same_form(host);
		check(ttvx_copy(m_dev, host.data().data(), bytes()));
	}

	void download(tensor<value_t>& host) const
	{
		require_same_form(host);
		check(ttvx_copy(host.data().data(), m_dev, bytes()));
	}

	tensor<value_t> to_host() const
	{
		tensor<value_t> host(shape(), layout());
		download(host);
		return host;
	}

	device_tensor_view<value_t> operator()(std::size_t contraction_mode) const
	{
		m_meta.require_mode(contraction_mode);
		return device_tensor_view<value_t>(*this, contraction_mode);
	}

	value_t const* data() const { return m_dev; }
	value_t* data() { return m_dev; }
	shape_t const& shape() const { return m_meta.extents(); }
	layout_t const& layout() const { return m_meta.layout(); }
	std::vector<std::size_t> strides() const { return m_meta.strides(); }
	std::size_t order() const { return m_meta.order(); }
	std::size_t size() const { return m_meta.elements(); }

private:
	std::size_t bytes() const { return m_meta.elements() * sizeof(value_t); }

	void allocate()
	{
		void* p = nullptr;
		check(ttvx_device_alloc(bytes(), &p));
		m_dev = static_cast<value_t*>(p);
	}

	void require_same_form(tensor<value_t> const& host) const
	{
		if(host.shape() != shape() || host.layout() != layout())
			throw std::runtime_error("Error in tlib::device_tensor: host tensor has another shape or layout.");
	}

	detail::tensor_meta m_meta;
	value_t* m_dev = nullptr;
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
