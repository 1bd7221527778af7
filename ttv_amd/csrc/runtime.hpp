// Shared host-side runtime of libttvx: error state, device properties, variant catalogue and
// launch configuration.  Used by the plain, quantised, batched and selector entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/ttvx.h"
#include "kernels.hpp"
#include "plan.hpp"

namespace ttvx
{

// ---- error state --------------------------------------------------------------------------
void set_error(const std::string& msg);
void clear_error();
int  fail(int status, const std::string& msg);          // set_error + return status
int  hip_fail(hipError_t e, const char* where);         // TTVX_ERR_DEVICE with HIP's text

// ---- fault injection (test hook, ttvx_debug_fail_next) ---------------------------------------
// True when a failure is armed at `site` (TTVX_FAULT_*); consumes one armed failure.
bool injected(int site);
// Arms `count` failures at `site` (replacing what was armed there).
void arm_fault(int site, int count);

// ---- device --------------------------------------------------------------------------------
struct DeviceInfo {
	int id = -1;
	int cus = 0;
	bool gfx950 = false;
	std::string arch;
};
// Properties of the current device; returns false (and sets the error) without a usable GPU.
bool current_device(DeviceInfo& info);

// Checks that [p, p + bytes) lies inside one device allocation (guards against host pointers
// and out-of-bounds spans reaching a kernel).
int check_device_span(const void* p, std::size_t bytes, const char* what);

// ---- variants --------------------------------------------------------------------------------
enum Kind { KIND_DOT = 0, KIND_COL = 1, KIND_STRIDED = 2 };

enum Variant {
	V_DOT_U1 = 0, V_DOT_U2, V_DOT_U4, V_DOT_U8, V_DOT_SPLIT,
	V_COL_U1, V_COL_U2, V_COL_U4, V_COL_NARROW, V_COL_SPLIT,
	V_COL_K_LANES, V_STRIDED, V_COL_WIDE,
	V_COL_U1_V8, V_COL_U2_V8,         // col_u1 / col_u2 with 8-byte vectors (1- and 2-byte storage)
	V_COUNT
};
const char* variant_name(int v);

// Element type code (TTVX_F32 ...) -> storage properties
struct DType {
	int code;
	int size;     // bytes per stored element
	int vmax;     // elements per 16-byte vector
};
DType dtype_of(int code);

// Kernels of the one-wave-per-group regime (kernels.hpp): k_rows for rows of one vector per lane,
// k_slab for col_narrow; LEAN_NONE for the looping k_dot / k_col / k_strided.
enum Lean { LEAN_NONE = 0, LEAN_ROWS = 1, LEAN_SLAB = 2 };

// Most blocks of kBlock threads in one launch (fewer than 2^32 work-items per grid dimension).
constexpr int64_t kMaxLeanBlocks = (int64_t{1} << 32) / kBlock - 1;

struct Config {
	Kind kind = KIND_STRIDED;
	Lean lean = LEAN_NONE;
	int vec = 1;        // elements per load
	int U = 1;
	int variant = V_STRIDED;
	Args args{};
	StridedArgs sargs{};
	int ksplit = 1;
	int grid = 1;          // work-bound grid; contiguous schedules shrink it to the resident blocks
	int cus = 1;
	std::size_t shmem = 0;
};

// Builds the launch configuration of variant v; false when v cannot run this problem.
bool configure(int v, const Fold& f, const DType& dt, uintptr_t a_addr, int cus, unsigned flags, Config& cfg);

// Output-store form of a column configuration for C at c_addr: widest aligned store (args.cvec) and
// wave-transposed stores (args.tstore, with their LDS staging in shmem).
void set_output_stores(Config& cfg, const Fold& f, const DType& dt, uintptr_t c_addr);

// Variant chosen by the fixed rule.
int heuristic_variant(const Fold& f, const DType& dt, uintptr_t a_addr, unsigned flags);

// Variant chosen by the process-wide selection mode (heuristic / learned / forced).
int select_variant(const Fold& f, const DType& dt, uintptr_t a_addr, unsigned flags);

// Selector features of a folded problem (TTVX_NUM_FEATURES floats).
void features_of(const Fold& f, const DType& dt, uintptr_t a_addr, float* out);

// Learned selector: argmin of the MLP's predicted relative time over valid variants, or -1
// when no weights are loaded.
int learned_variant(const Fold& f, const DType& dt, uintptr_t a_addr, unsigned flags);

// Loads/clears the learned selector's weights (layout in mlp.hpp); count 0 clears.
int set_selector_weights(const float* params, std::size_t count);
// Loads the shipped selector weights once per process unless weights were set (or cleared)
// explicitly; failure leaves the fixed rule in charge.
void load_default_selector_weights();
int set_selector_table(const int64_t* rows, std::size_t count);

// Typed launchers, one per storage type (launch_<type>.hip).
int launch_f32 (const Config& cfg, const void* a, const void* b, void* c, bool kahan, hipStream_t s);
int launch_f64 (const Config& cfg, const void* a, const void* b, void* c, bool kahan, hipStream_t s);
int launch_bf16(const Config& cfg, const void* a, const void* b, void* c, bool kahan, hipStream_t s);
int launch_fp8 (const Config& cfg, const void* a, const void* b, void* c, bool kahan, hipStream_t s);
// BF16 / FP8 kernels with realigned loads, for tensors whose vectors do not all start on 4-byte boundaries
int launch_bf16_sub(const Config& cfg, const void* a, const void* b, void* c, bool kahan, hipStream_t s);
int launch_fp8_sub (const Config& cfg, const void* a, const void* b, void* c, bool kahan, hipStream_t s);

// Device memory a caller provides for the split partial sums instead of the library's buffers.
struct Scratch {
	void* ptr = nullptr;
	std::size_t bytes = 0;
};

// Grow-only buffer of at least `bytes` for the split partials of launches on (device, stream).
int partial_buffer(int device, hipStream_t stream, std::size_t bytes, void** out);

// Enqueues C = A x_q b for a folded problem with device pointers.  Split partial sums go to
// `scratch` when given (an error when they do not fit), else to the stream's partials buffer.
int launch(const Fold& f, const DType& dt, const void* a, const void* b, void* c,
           const float* scales, int64_t fps, unsigned flags, int variant, hipStream_t stream,
           const Scratch* scratch = nullptr);

} // namespace ttvx
