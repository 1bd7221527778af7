// Policy tags of the mode-q tensor-times-vector product.
//
// Same type and constant names as the reference (include/tlib/detail/tags.h:21-63) so that
// callers compile unchanged.  On MI355X the tags are hints only: every combination is
// accepted (the reference rejects undefined combinations at link time,
// tensor_times_vector.h:416-421) and the kernel variant is chosen by the planner.
#pragma once

namespace tlib::ttv::execution_policy
{
struct sequential_t         {};
struct sequential_blas_t    {};
struct parallel_t           {};
struct parallel_taskloop_t  {};
struct parallel_task_t      {};
struct parallel_loop_t      {};
struct parallel_blas_t      {};
struct parallel_loop_blas_t {};

inline constexpr sequential_t         seq;
inline constexpr sequential_blas_t    seq_blas;
inline constexpr parallel_t           par;
inline constexpr parallel_loop_t      par_loop;
inline constexpr parallel_taskloop_t  par_taskloop;
inline constexpr parallel_task_t      par_task;
inline constexpr parallel_blas_t      par_blas;
inline constexpr parallel_loop_blas_t par_blas_loop;
} // namespace tlib::ttv::execution_policy

namespace tlib::ttv::slicing_policy
{
struct slice_t      {};
struct subtensor_t  {};

inline constexpr slice_t     slice;
inline constexpr subtensor_t subtensor;
} // namespace tlib::ttv::slicing_policy

namespace tlib::ttv::fusion_policy
{
struct none_t   {};
struct outer_t  {};
struct all_t    {};

inline constexpr none_t    none;
inline constexpr outer_t   outer;
inline constexpr all_t     all;
} // namespace tlib::ttv::fusion_policy

namespace tlib::ttv::detail
{
// Integer codes carried across the C-ABI (ttvx.h TTVX_EXEC_*, TTVX_SLICE_*, TTVX_FUSE_*).
constexpr int policy_code(execution_policy::sequential_t)         { return 0; }
constexpr int policy_code(execution_policy::sequential_blas_t)    { return 1; }
constexpr int policy_code(execution_policy::parallel_t)           { return 2; }
constexpr int policy_code(execution_policy::parallel_taskloop_t)  { return 3; }
constexpr int policy_code(execution_policy::parallel_task_t)      { return 4; }
constexpr int policy_code(execution_policy::parallel_loop_t)      { return 5; }
constexpr int policy_code(execution_policy::parallel_blas_t)      { return 6; }
constexpr int policy_code(execution_policy::parallel_loop_blas_t) { return 7; }
constexpr int policy_code(slicing_policy::slice_t)                { return 0; }
constexpr int policy_code(slicing_policy::subtensor_t)            { return 1; }
constexpr int policy_code(fusion_policy::none_t)                  { return 0; }
constexpr int policy_code(fusion_policy::outer_t)                 { return 1; }
constexpr int policy_code(fusion_policy::all_t)                   { return 2; }
} // namespace tlib::ttv::detail
