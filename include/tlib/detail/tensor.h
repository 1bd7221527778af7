// Owning host tensor and the mode view used by A(q) * b.
//
// Same interface as the reference (include/tlib/detail/tensor.h:36-114): shape, layout,
// contiguous std::vector storage, strides derived from (shape, layout), operator()(q) giving a
// tensor_view, default first-order layout.  The product itself runs on the GPU through the
// C-ABI (ttvx.h); host buffers are staged by the library.
#pragma once

#include <functional>
#include <numeric>
#include <stdexcept>
#include <vector>

#include "layout.h"
#include "shape.h"
#include "strides.h"

namespace tlib::ttv
{

template<class value_t>
class tensor;

template<class _value_t>
struct tensor_view
{
public:
	using value_t  = _value_t;
	using tensor_t = tensor<value_t>;
	tensor_view() = delete;
	tensor_view(tensor_view const&) = delete;
	tensor_view& operator=(tensor_view const&) = delete;
	inline tensor_t const& get_tensor() const { return _tensor; }
	inline std::size_t contraction_mode() const { return _q; }
private:
	friend tensor_view tensor_t::operator()(std::size_t q) const;
	tensor_view(tensor_t const& t, std::size_t q) : _tensor(t), _q(q) {}
	tensor_t const& _tensor;
	std::size_t _q;
};

template<class value_t>
class tensor
{
	using shape_t       = std::vector<std::size_t>;
	using layout_t      = std::vector<std::size_t>;
	using strides_t     = std::vector<std::size_t>;
	using vector_t      = std::vector<value_t>;
	using tensor_view_t = tensor_view<value_t>;
public:
	tensor() = delete;

	tensor(shape_t const& n, layout_t const& pi)
		: _n(n)
		, _pi(pi)
		, _data(std::accumulate(n.begin(), n.end(), std::size_t{1}, std::multiplies<>()))
	{
		if(n.size() != pi.size())
			throw std::runtime_error("Error in tlib::tensor: shape vector and layout vector must have the same length.");
		if(!detail::is_valid_shape(n.begin(), n.end()))
			throw std::runtime_error("Error in tlib::tensor: shape vector of tensor is not valid.");
		if(!detail::is_valid_layout(pi.begin(), pi.end()))
			throw std::runtime_error("Error in tlib::tensor: layout vector of tensor is not valid.");
	}

	tensor(shape_t const& n)
		: tensor(n, detail::generate_k_order_layout(n.size(), std::size_t{1}))
	{
	}

	tensor& operator=(value_t v)
	{
		std::fill(_data.begin(), _data.end(), v);
		return *this;
	}

	tensor_view_t operator()(std::size_t contraction_mode) const
	{
		if(1ul > contraction_mode || contraction_mode > this->order())
			throw std::runtime_error("Error in tlib::tensor: specified contraction mode should be greater than one and equal to or less than the order.");
		return tensor_view_t(*this, contraction_mode);
	}

	auto begin() const { return _data.begin(); }
	auto end  () const { return _data.end  (); }
	decltype(auto) begin() { return _data.begin(); }
	decltype(auto) end  () { return _data.end  (); }

	auto const& data   () const { return _data; }
	auto      & data   ()       { return _data; }
	auto const& shape  () const { return _n; }
	auto const& layout () const { return _pi; }
	auto        strides() const { return detail::generate_strides(_n, _pi); }
	auto        order  () const { return _n.size(); }

private:
	shape_t  _n;
	layout_t _pi;
	vector_t _data;
};

} // namespace tlib::ttv
