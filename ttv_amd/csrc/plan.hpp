// Host-side validation and folding of a mode-q tensor-times-vector problem.
//
// Validation reproduces the checks and messages of the reference's low-level interface
// (include/tlib/ttv.h:64-89) in the same order, plus the pia/pic consistency check that the
// reference performs for its case 8 (include/tlib/detail/tensor_times_vector.h:147-168).
//
// Folding replaces the reference's Loops-over-GEMM recursion (tensor_times_vector.h:189-398)
// and its fused-index arithmetic (index.h:76-124): free modes of A are visited in layout
// order, extent-1 modes are dropped, and adjacent modes whose A and C strides chain
// (w_next == w * n) are merged.  A dense tensor always folds to the 3-D view
//     A[o*sAo + k*sAk + i*sAi],  C[o*sCo + i*sCi],   o < O, k < K = n_q, i < I
// where I collects the free modes stored before q and O those stored after q.  Anything that
// does not fold (general strides) keeps its list of free dimensions for the strided kernel.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace ttvx
{

constexpr int kMaxDims = 32;  // free dimensions the strided kernel handles after merging

struct FreeDim {
	int64_t n, sa, sc;
};

struct Fold {
	int64_t O = 1, K = 1, I = 1;
	int64_t sAo = 0, sAk = 0, sAi = 0;
	int64_t sCo = 0, sCi = 0;
	bool    three_d = true;           // the 3-D view above exists
	std::vector<FreeDim> dims;        // free dims after merging, fastest first (strided kernel)
	int64_t a_span = 0;               // elements from A's base to its last element + 1
	int64_t c_span = 0;               // same for C
	int64_t outputs = 0;              // number of C elements
	int     ref_case = 0;             // reference case number (cases.h)
	int     p = 0;                    // order of A
	int     qpos = 0;                 // zero-based layout position of q
	std::vector<int64_t> n_layout;    // extents of A in layout order (fastest first)
};

// Status + message of a failed validation.
struct Check {
	int status = 0;         // TTVX_OK or TTVX_ERR_*
	std::string message;
};

// Validates the arguments of the low-level interface.  Pointers a/b/c are only tested for null.
Check validate(std::size_t q, std::size_t p,
               const void* a, const std::size_t* na, const std::size_t* wa, const std::size_t* pia,
               const void* b, const std::size_t* nb,
               const void* c, const std::size_t* nc, const std::size_t* wc, const std::size_t* pic);

// Folds a validated problem.  wc/pic may be null, meaning the dense output derived from A.
Fold fold(std::size_t q, std::size_t p, const std::size_t* na, const std::size_t* wa, const std::size_t* pia,
          const std::size_t* wc, const std::size_t* pic);

// Validates only the A side (q, p, na, wa, pia) for the planning queries; empty message = OK.
Check validate_a(std::size_t q, std::size_t p, const std::size_t* na, const std::size_t* wa, const std::size_t* pia);

} // namespace ttvx
