// Host tensor of the high-level interface and the mode view that A(q) returns.
//
// Interface of the reference's tensor<T> / tensor_view<T> (include/tlib/detail/tensor.h:36-114):
// construction from a shape (first-order layout) or from a shape and a layout, the same three
// constructor messages and the same operator()(q) message, element access through data() /
// begin() / end(), and shape() / layout() / strides() / order().  The products run on the GPU
// through the C-ABI (ttvx.h); the library stages these host buffers.
//
// Internally a tensor is its validated metadata (detail::tensor_meta: extents, layout and the
// dense strides, computed once) plus a contiguous element vector.  A(q) is a
// detail::mode_view, a non-owning (tensor, q) pair that can only be made by the tensor itself
// and cannot be copied, so `A(q) * b` is the one way to use it.  device_tensor<T>
// (device_tensor.h) shares both pieces.
#pragma once

#include <algorithm>
#include <cstddef>
#include <stdexcept>
#include <utility>
#include <vector>

#include "layout.h"
#include "shape.h"
#include "strides.h"

namespace tlib::ttv
{

namespace detail
{

// Extents, layout tuple and dense strides of a tensor, checked on construction.
class tensor_meta
{
public:
	using index_vector = std::vector<std::size_t>;

	tensor_meta(index_vector extents, index_vector layout) : m_extents(std::move(extents)), m_layout(std::move(layout))
	{
		if(m_extents.size() != m_layout.size())
			throw std::runtime_error("Error in tlib::tensor: shape vector and layout vector must have the same length.");
		if(!is_valid_shape(m_extents.begin(), m_extents.end()))
			throw std::runtime_error("Error in tlib::tensor: shape vector of tensor is not valid.");
		if(!is_valid_layout(m_layout.begin(), m_layout.end()))
			throw std::runtime_error("Error in tlib::tensor: layout vector of tensor is not valid.");
		m_strides = generate_strides(m_extents, m_layout);
		m_elements = 1;
		for(std::size_t n : m_extents) m_elements *= n;
	}

	// first-order (column-major) layout 1, 2, ..., p
	explicit tensor_meta(index_vector extents)
		: tensor_meta(extents, generate_k_order_layout(extents.size(), std::size_t{1}))
	{
	}

	index_vector const& extents() const { return m_extents; }
	index_vector const& layout() const { return m_layout; }
	index_vector const& strides() const { return m_strides; }
	std::size_t order() const { return m_extents.size(); }
	std::size_t elements() const { return m_elements; }

	// the contraction mode of A(q) must name one of the tensor's modes
	void require_mode(std::size_t q) const
	{
		if(q < 1 || q > order())
			throw std::runtime_error("Error in tlib::tensor: specified contraction mode should be greater than one and equal to or less than the order.");
	}

private:
	index_vector m_extents;
	index_vector m_layout;
	index_vector m_strides;
	std::size_t m_elements = 0;
};

// (tensor, contraction mode) pair returned by A(q); only the tensor type can make one.
template<class Tensor>
class mode_view
{
public:
	using value_t  = typename Tensor::value_type;
	using tensor_t = Tensor;

	mode_view() = delete;
	mode_view(mode_view const&) = delete;
	mode_view& operator=(mode_view const&) = delete;

	tensor_t const& get_tensor() const { return *m_source; }
	std::size_t contraction_mode() const { return m_mode; }

private:
	friend Tensor;
	mode_view(tensor_t const& source, std::size_t mode) noexcept : m_source(&source), m_mode(mode) {}

	tensor_t const* m_source;
	std::size_t m_mode;
};

} // namespace detail

template<class value_t>
class tensor;

// What A(q) returns (reference tensor_view, include/tlib/detail/tensor.h:36-52).
template<class value_t>
using tensor_view = detail::mode_view<tensor<value_t>>;

// Owning host tensor with contiguous storage in its layout.
template<class value_t>
class tensor
{
public:
	using value_type = value_t;
	using shape_t    = std::vector<std::size_t>;
	using layout_t   = std::vector<std::size_t>;

	tensor() = delete;
	tensor(shape_t const& n, layout_t const& pi) : m_meta(n, pi), m_elems(m_meta.elements()) {}
	tensor(shape_t const& n) : m_meta(n), m_elems(m_meta.elements()) {}

	// every element set to v
	tensor& operator=(value_t v)
	{
		std::fill(m_elems.begin(), m_elems.end(), v);
		return *this;
	}

	tensor_view<value_t> operator()(std::size_t contraction_mode) const
	{
		m_meta.require_mode(contraction_mode);
		return tensor_view<value_t>(*this, contraction_mode);
	}

	auto begin() const { return m_elems.cbegin(); }
	auto end() const { return m_elems.cend(); }
	auto begin() { return m_elems.begin(); }
	auto end() { return m_elems.end(); }

	std::vector<value_t> const& data() const { return m_elems; }
	std::vector<value_t>& data() { return m_elems; }
	shape_t const& shape() const { return m_meta.extents(); }
	layout_t const& layout() const { return m_meta.layout(); }
	std::vector<std::size_t> strides() const { return m_meta.strides(); }
	std::size_t order() const { return m_meta.order(); }

private:
	detail::tensor_meta m_meta;
	std::vector<value_t> m_elems;
};

} // namespace tlib::ttv
