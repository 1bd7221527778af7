// Host runtime of libttvx: error state, device checks, variant catalogue, launch configuration,
// fixed selection rule and the launch itself.
#include "runtime.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ttvx.h"
#include "mlp.hpp"

// Waves per CU that the split row kernel aims for (its splits x row groups).  Measured on MI355X,
// rows of 2^27-2^29 bytes, O = 2..100: 32 beats 8 and 16 by up to 30 % for 30 and more rows.
#ifndef TTVX_SPLIT_WPC
#define TTVX_SPLIT_WPC 32
#endif

namespace ttvx
{

// ---- error state --------------------------------------------------------------------------
namespace
{
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }
int fail(int status, const std::string& msg)
{
	set_error(msg);
	return status;
}
int hip_fail(hipError_t e, const char* where)
{
	std::string m = "ttvx: HIP error in ";
	m += where;
	m += ": ";
	m += hipGetErrorString(e);
	return fail(TTVX_ERR_DEVICE, m);
}
const char* last_error_cstr() { return g_last_error.c_str(); }

// ---- fault injection -------------------------------------------------------------------------
namespace
{
std::atomic<int> g_faults[3];
} // namespace

void arm_fault(int site, int count)
{
	if(site >= 0 && site < 3) g_faults[site].store(count);
}

bool injected(int site)
{
	if(site < 0 || site >= 3) return false;
	int v = g_faults[site].load();
	while(v > 0)
		if(g_faults[site].compare_exchange_weak(v, v - 1)) return true;
	return false;
}

// ---- device --------------------------------------------------------------------------------
namespace
{
std::mutex g_dev_mutex;
std::vector<DeviceInfo> g_devices;   // cached per device id
} // namespace

bool current_device(DeviceInfo& info)
{
	int id = 0;
	hipError_t e = hipGetDevice(&id);
	if(e != hipSuccess) {
		(void)hipGetLastError();
		set_error("ttvx: no HIP device available");
		return false;
	}
	std::lock_guard<std::mutex> lock(g_dev_mutex);
	if(static_cast<int>(g_devices.size()) <= id) g_devices.resize(id + 1);
	DeviceInfo& d = g_devices[id];
	if(d.id < 0) {
		hipDeviceProp_t prop;
		e = hipGetDeviceProperties(&prop, id);
		if(e != hipSuccess) {
			(void)hipGetLastError();
			set_error("ttvx: hipGetDeviceProperties failed");
			return false;
		}
		d.id     = id;
		d.cus    = prop.multiProcessorCount;
		d.arch   = prop.gcnArchName;
		d.gfx950 = d.arch.rfind("gfx950", 0) == 0;
	}
	if(!d.gfx950) {
		set_error("ttvx: device " + std::to_string(id) + " is " + d.arch + "; this library is built for gfx950 (MI355X)");
		return false;
	}
	info = d;
	return true;
}

int check_device_span(const void* p, std::size_t bytes, const char* what)
{
	hipDeviceptr_t base = nullptr;
	std::size_t size = 0;
	hipError_t e = hipMemGetAddressRange(&base, &size, const_cast<void*>(p));
	if(e != hipSuccess || base == nullptr) {
		(void)hipGetLastError();
		return fail(TTVX_ERR_ARGUMENT, std::string("ttvx: ") + what + " is not device memory (stream-ordered entry points take device pointers)");
	}
	auto const lo = reinterpret_cast<uintptr_t>(base), at = reinterpret_cast<uintptr_t>(p);
	if(at < lo || at + bytes > lo + size)
		return fail(TTVX_ERR_ARGUMENT, std::string("ttvx: ") + what + " spans past the end of its device allocation");
	return TTVX_OK;
}

// ---- variants --------------------------------------------------------------------------------
namespace
{
static_assert(kMlpV == V_COUNT, "the selector network has one output per kernel variant");
const char* const kVariantNames[V_COUNT] = {
	"dot_u1", "dot_u2", "dot_u4", "dot_u8", "dot_split",
	"col_u1", "col_u2", "col_u4", "col_narrow", "col_split",
	"col_k_lanes", "strided", "col_wide",
	"col_u1_v8", "col_u2_v8",
};

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
int64_t pow2_ceil(int64_t x)
{
	int64_t p = 1;
	while(p < x) p <<= 1;
	return p;
}
bool is_pow2(int64_t x) { return x > 0 && (x & (x - 1)) == 0; }
int ilog2(int64_t x)
{
	int r = 0;
	while((int64_t{1} << (r + 1)) <= x) ++r;
	return r;
}
} // namespace

const char* variant_name(int v) { return (v >= 0 && v < V_COUNT) ? kVariantNames[v] : ""; }

DType dtype_of(int code)
{
	switch(code) {
		case TTVX_F32:     return {TTVX_F32, 4, 4};
		case TTVX_F64:     return {TTVX_F64, 8, 2};
		case TTVX_BF16:    return {TTVX_BF16, 2, 8};
		case TTVX_FP8E4M3: return {TTVX_FP8E4M3, 1, 16};
		default:           return {-1, 0, 0};
	}
}

namespace
{
// Vector width (elements) of a row of `inner` elements: the widest power of two that fits, at most
// `bytes` bytes (16, or 8 for the col_*_v8 variants).  Vectors need only the element's alignment,
// and a row whose length the width does not divide ends in a vector clamped to the row's end
// (kernels.hpp: col_clamp, row_clamp).
int pick_vec(const DType& dt, int64_t inner, int bytes = 16)
{
	int v = std::min(dt.vmax, bytes / dt.size);
	while(v > 1 && v > inner) v >>= 1;
	return v;
}

// Widest power-of-two vector whose loads are all naturally aligned (the selector's alignment feature).
int aligned_vec(const DType& dt, uintptr_t a_addr, int64_t inner, std::initializer_list<int64_t> strides)
{
	int v = dt.vmax;
	while(v > 1) {
		bool ok = inner % v == 0 && a_addr % static_cast<uintptr_t>(v * dt.size) == 0;
		for(int64_t s : strides) ok = ok && s % v == 0;
		if(ok) break;
		v >>= 1;
	}
	return v;
}

std::size_t acc_size(const DType& dt) { return dt.code == TTVX_F64 ? 8 : 4; }

// Build-time tuning constants (the defaults are the values measured best on MI355X; the
// benchmark sweeps rebuild with -D overrides to compare):
//   TTVX_CONTIG_DEFAULT       work assignment of k_dot / k_col: 0 = grid-stride over up to
//                             TTVX_GRID_PER_CU blocks per CU (faster for rows and narrow column
//                             tiles), 1 = one contiguous range per block
//   TTVX_GRID_PER_CU          block cap per CU of the col_u4, col_wide, col_split, col_k_lanes and
//                             strided launches
//   TTVX_COL_WIDE_U           slab units per lane of the col_wide variant
//   TTVX_DOT_GROUPS_PER_WAVE  row groups per wave the looping row launches aim at (twice that for
//                             1- and 2-byte types, whose groups hold more rows)
//   TTVX_NARROW_LANES         lanes per row chunk of col_narrow
#ifndef TTVX_CONTIG_DEFAULT
#define TTVX_CONTIG_DEFAULT 0
#endif
#ifndef TTVX_GRID_PER_CU
#define TTVX_GRID_PER_CU 16
#endif
constexpr int kGridPerCU = TTVX_GRID_PER_CU;
// column kernels: a slab row narrower than this many bytes gets more k rows per wave so that each
// wave-wide load covers whole 128-byte lines (16^7 with I = 16 FP32: 4.4 -> 5.1 TB/s)
#ifndef TTVX_MIN_ROW_BYTES
#define TTVX_MIN_ROW_BYTES 128
#endif
#ifndef TTVX_COL_WIDE_U
#define TTVX_COL_WIDE_U 4
#endif
#ifndef TTVX_DOT_GROUPS_PER_WAVE
#define TTVX_DOT_GROUPS_PER_WAVE 2
#endif
constexpr int kDotGroupsPerWave = TTVX_DOT_GROUPS_PER_WAVE;
constexpr int kDotGridPerCU = 64;
// largest grid of k_rows and k_slab, which take one work unit per wave or block and so need a
// block for every unit: a launch holds fewer than 2^32 work-items
// col_narrow's row chunk in lanes (16 lanes x 16-byte vectors = 256 bytes: 6.61 -> 6.81 TB/s on
// FP32 slabs of 128 x 128 against whole 512-byte rows)
#ifndef TTVX_NARROW_LANES
#define TTVX_NARROW_LANES 16
#endif
constexpr int64_t kNarrowLanes = TTVX_NARROW_LANES;
} // namespace

bool configure(int v, const Fold& f, const DType& dt, uintptr_t a_addr, int cus, unsigned flags, Config& cfg)
{
	cfg = Config{};
	cfg.variant = v;
	cfg.cus = std::max(cus, 1);
	Args& P = cfg.args;
	std::memset(&P, 0, sizeof(P));
	P.fps = 1;
	bool const fp64_final = (flags & TTVX_FLAG_FP64_FINAL) != 0 || f.K > (int64_t{1} << 20);
	int64_t const gridcap = static_cast<int64_t>(std::max(cus, 1)) * kGridPerCU;

	if(v == V_STRIDED) {
		if(static_cast<int>(f.dims.size()) > kMaxFree) return false;
		cfg.kind = KIND_STRIDED;
		StridedArgs& G = cfg.sargs;
		std::memset(&G, 0, sizeof(G));
		G.nd = static_cast<int>(f.dims.size());
		for(int d = 0; d < G.nd; ++d) { G.n[d] = f.dims[d].n; G.sa[d] = f.dims[d].sa; G.sc[d] = f.dims[d].sc; }
		G.K = f.K; G.sAk = f.sAk; G.outputs = f.outputs; G.fps = 1;
		cfg.grid = static_cast<int>(std::max<int64_t>(1, std::min(ceil_div(f.outputs, kBlock), gridcap)));
		return true;
	}
	if(!f.three_d) return false;

	P.O = f.O; P.K = f.K; P.I = f.I;
	P.sAo = f.sAo; P.sAk = f.sAk; P.sCo = f.sCo;

	bool const dot_shape = f.I == 1 && (f.sAk == 1 || f.K == 1);
	bool const col_shape = f.I > 1 && f.sAi == 1 && f.sCi == 1;

	if(v >= V_DOT_U1 && v <= V_DOT_SPLIT) {
		if(!dot_shape) return false;
		cfg.kind = KIND_DOT;
		cfg.vec  = pick_vec(dt, f.K);
		int64_t const vsteps = ceil_div(f.K, cfg.vec);
		P.LPR = static_cast<int>(std::min<int64_t>(64, pow2_ceil(vsteps)));
		P.lprShift = ilog2(P.LPR);
		int64_t const kstride = static_cast<int64_t>(P.LPR) * cfg.vec;
		static const int us[] = {1, 2, 4, 8, 2};
		cfg.U = us[v - V_DOT_U1];
		int64_t const rpw = 64 / P.LPR;
		P.groups = ceil_div(f.O, rpw * cfg.U);
		int64_t const waves_wanted = static_cast<int64_t>(std::max(cus, 1)) * 4 * 4;
		int64_t ksplit = 1;
		if(v == V_DOT_SPLIT) {
			if(P.groups >= waves_wanted || vsteps < 4 * P.LPR) return false;  // nothing to gain
			// up to 4096 splits: a few long rows (O = 2, K = 2^27) still fill the GPU with waves
			// (0.7 -> 6.5 TB/s in FP32); their partials are summed by one block per output
			// (k_finalize_wide)
			ksplit = std::min<int64_t>(4096, ceil_div(static_cast<int64_t>(std::max(cus, 1)) * TTVX_SPLIT_WPC, P.groups));
		}
		int64_t kchunk = ceil_div(ceil_div(f.K, ksplit), kstride) * kstride;
		ksplit = ceil_div(f.K, kchunk);
		P.kchunk = kchunk;
		P.kiter  = static_cast<int>(std::min<int64_t>(1 << 30, ceil_div(std::min(kchunk, f.K), kstride)));
		P.kpart  = f.K % cfg.vec != 0 ? 1 : 0;
		cfg.ksplit = static_cast<int>(ksplit);
		if(P.kiter == 1 && ksplit == 1 && ceil_div(P.groups, kBlock / 64) <= kMaxLeanBlocks) {
			// rows of one vector per lane: k_rows, one group per wave and no cap on the grid (short-lived
			// blocks handed out in address order sweep A as one compact front).  When a power-of-two
			// row group leaves many lanes idle (9 vectors per row: 9 of 16), rows take exactly their
			// vectors' lanes ("tight": 7 rows of 9 lanes per wave) if that keeps >= 15 % more lanes busy
			double const util_pow2 = static_cast<double>(vsteps) / P.LPR;
			double const util_tight = static_cast<double>((64 / vsteps) * vsteps) / 64.0;
			if(vsteps > 1 && !is_pow2(vsteps) && util_tight > 1.15 * util_pow2) {
				P.tight = 1;
				P.LPR = static_cast<int>(vsteps);
				P.lprShift = 0;
				P.groups = ceil_div(f.O, (64 / vsteps) * cfg.U);
			}
			cfg.lean = LEAN_ROWS;
			cfg.grid = static_cast<int>(std::max<int64_t>(1, ceil_div(P.groups, kBlock / 64)));
		} else {
			// about kDotGroupsPerWave row groups per wave (twice that for the narrow types, whose groups
			// hold more rows): enough waves to fill the GPU, and each wave still has a next group to load
			// while it sums the current one
			int64_t const dot_cap = static_cast<int64_t>(std::max(cus, 1)) * kDotGridPerCU;
			// (the split variant: one group per wave, the splits were counted for that)
			int64_t const per_wave = v == V_DOT_SPLIT ? 1 : (dt.size >= 4 ? kDotGroupsPerWave : 2 * kDotGroupsPerWave);
			cfg.grid = static_cast<int>(std::max<int64_t>(1, std::min(ceil_div(P.groups, (kBlock / 64) * per_wave), dot_cap)));
		}
	} else if((v >= V_COL_U1 && v <= V_COL_SPLIT) || v == V_COL_WIDE || v == V_COL_U1_V8 || v == V_COL_U2_V8) {
		if(!col_shape) return false;
		cfg.kind = KIND_COL;
		// the _v8 variants: col_u1 / col_u2 with 8-byte lane vectors (twice the lanes per row, half the
		// accumulators per lane; for 1- and 2-byte storage, where 16 bytes are 8-16 columns per lane)
		bool const v8 = v == V_COL_U1_V8 || v == V_COL_U2_V8;
		int const vb = v8 ? (v == V_COL_U1_V8 ? V_COL_U1 : V_COL_U2) : v;     // the tiling it follows
		cfg.vec  = pick_vec(dt, f.I, v8 ? 8 : 16);
		if(v8 && (dt.size > 2 || cfg.vec == pick_vec(dt, f.I))) return false;  // nothing narrower to take
		int64_t const iv = ceil_div(f.I, cfg.vec);          // vectors per row (the last one clamped)
		int64_t const itmax = kBlock;
		P.IT = static_cast<int>(std::min<int64_t>(iv, itmax));
		// k groups per block: "wide" spreads the whole block over k (a flat walk through each
		// slab); the other column variants keep >= 16 k steps per lane and fill the block with slabs
		int64_t const ks_cap = kBlock / P.IT;
		int64_t ks_deep = std::max<int64_t>(1, std::min<int64_t>(ks_cap, f.K / 16));
#if TTVX_MIN_ROW_BYTES
		{
			// rows narrower than a cache line: enough k rows per wave instruction to fill whole lines
			int64_t const row = static_cast<int64_t>(P.IT) * cfg.vec * dt.size;
			if(row < TTVX_MIN_ROW_BYTES)
				ks_deep = std::max<int64_t>(ks_deep, std::min<int64_t>(ks_cap, std::min<int64_t>(f.K, TTVX_MIN_ROW_BYTES / row)));
		}
#endif
		if(is_pow2(P.IT)) ks_deep = int64_t{1} << ilog2(ks_deep);     // power-of-two IT*KS <= 64 sums inside one wave
		P.KS = static_cast<int>(vb == V_COL_WIDE ? std::max<int64_t>(1, std::min<int64_t>(ks_cap, f.K)) : ks_deep);
		P.OS = std::max(1, kBlock / (P.IT * P.KS));
		static const int us[] = {1, 2, 4, 1, 1};
		cfg.U = vb == V_COL_WIDE ? TTVX_COL_WIDE_U : us[vb - V_COL_U1];
		// wide rows (a full-width block still leaves U vectors per lane): the U units of a lane
		// step along i, so a block streams U*IT*16 bytes of every row it touches
		P.along_i = (cfg.U > 1 && iv >= static_cast<int64_t>(P.IT) * cfg.U && P.IT == kBlock) ? 1 : 0;
		if(P.along_i) {
			P.iTiles = ceil_div(iv, static_cast<int64_t>(P.IT) * cfg.U);
			P.items  = ceil_div(f.O, static_cast<int64_t>(P.OS)) * P.iTiles;
		} else {
			P.iTiles = ceil_div(iv, P.IT);
			P.items  = ceil_div(f.O, static_cast<int64_t>(P.OS) * cfg.U) * P.iTiles;
		}
		if(vb == V_COL_NARROW) {
			// k_slab: rows cut into chunks of IT <= kNarrowLanes vectors (a power of two), the block's
			// other lanes along k (as many as K uses), OS slabs per block, one item per block
			P.IT = static_cast<int>(std::min<int64_t>(int64_t{1} << ilog2(iv), kNarrowLanes));
			P.KS = static_cast<int>(std::min<int64_t>(kBlock / P.IT, pow2_ceil(f.K)));
			P.OS = kBlock / (P.IT * P.KS);
			P.itShift = ilog2(P.IT);
			P.spanShift = ilog2(static_cast<int64_t>(P.IT) * P.KS);
			cfg.U = 1;
			P.along_i = 0;
			P.iTiles = ceil_div(iv, P.IT);
			P.items  = ceil_div(f.O, static_cast<int64_t>(P.OS)) * P.iTiles;
			P.red    = 2;
			P.kchunk = f.K;
			cfg.ksplit = 1;
			cfg.lean = LEAN_SLAB;
			cfg.grid = static_cast<int>(std::max<int64_t>(1, P.items));
			if(P.items > kMaxLeanBlocks) return false;               // one block per item
			if(fp64_final && dt.code != TTVX_F64) return false;      // no split partials in k_slab
			return true;
		}
		int64_t ksplit = 1;
		if(vb == V_COL_SPLIT) {
			// only when the output tiles cannot fill the GPU and every split keeps long k runs; up to
			// 4096 splits (8 blocks per CU), so that a few narrow columns (O = 1, I = 2..64, K = 2^22
			// and more) still fill the GPU
			int64_t const want = static_cast<int64_t>(std::max(cus, 1)) * 8;
			if(P.items >= std::max(cus, 1) || f.K / P.KS < 64) return false;
			ksplit = std::min<int64_t>(4096, std::min<int64_t>(ceil_div(want, P.items), f.K / P.KS / 32));
			if(ksplit < 2) return false;
		}
		int64_t const kchunk = ceil_div(f.K, ksplit);
		P.kchunk = kchunk;
		cfg.ksplit = static_cast<int>(ceil_div(f.K, kchunk));
		// one- and two-unit tiles: one item per block on an uncapped grid (short-lived blocks handed
		// out in address order; MI355X, FP32: +1-2 % on 128^4 with I = 128, +4-7 % on 16^7 with
		// I = 16..256 against blocks looping over items); wider tiles keep the capped grid
		int64_t const cap = (vb == V_COL_U1 || vb == V_COL_U2) ? kMaxLeanBlocks : gridcap;
		cfg.grid   = static_cast<int>(std::max<int64_t>(1, std::min(P.items, cap)));
	} else if(v == V_COL_K_LANES) {
		if(!dot_shape) return false;
		// the column kernel on a row problem: lanes along k (scalar loads), summed across 64 lanes
		cfg.kind = KIND_COL;
		cfg.vec  = 1;
		P.I = 1; P.IT = 1;
		P.KS = static_cast<int>(std::min<int64_t>(64, pow2_ceil(f.K)));
		P.OS = kBlock / P.KS;
		P.sAk = 1;
		cfg.U = 1;
		P.iTiles = 1;
		P.items  = ceil_div(f.O, P.OS);
		P.kchunk = f.K;
		cfg.ksplit = 1;
		cfg.grid = static_cast<int>(std::max<int64_t>(1, std::min(P.items, gridcap)));
	} else {
		return false;
	}

	P.contig = TTVX_CONTIG_DEFAULT;
	if(cfg.kind == KIND_COL) {
		if(P.KS == 1) P.red = 0;
		else if(is_pow2(P.IT) && is_pow2(P.KS) && P.IT * P.KS <= 64) P.red = 1;
		else P.red = 2;
		cfg.shmem = P.red == 2 ? static_cast<std::size_t>(cfg.U) * cfg.vec * kBlock * acc_size(dt) : 0;
		if(cfg.shmem > 64 * 1024) return false;
		P.one_step = (P.red != 2 && P.kchunk <= P.KS && P.K <= P.KS) ? 1 : 0;
	}
	if(fp64_final && dt.code != TTVX_F64) P.part64 = 1;
	return true;
}

int heuristic_variant(const Fold& f, const DType& dt, uintptr_t a_addr, unsigned flags)
{
	Config cfg;
	int const cus = 256;
	if(!f.three_d) return V_STRIDED;
	// FP64 below ~192 MiB: the launch is short and its ramp and tail weigh more, so the variants
	// that spread the work over the most waves win (measured on 256^3: U=1 tiles 1.5x faster than
	// U=4); for 4-byte and narrower types the same rule loses on average, and is not applied
	bool const small = dt.size == 8 && f.O * f.K * f.I * dt.size < (int64_t{192} << 20);
	if(f.I == 1 && (f.sAk == 1 || f.K == 1)) {
		// rows of K elements: enough rows in flight per lane, split long rows when rows are few
		if(configure(V_DOT_SPLIT, f, dt, a_addr, cus, flags, cfg)) return V_DOT_SPLIT;
		configure(V_DOT_U1, f, dt, a_addr, cus, flags, cfg);
		int64_t const kiter = cfg.args.kiter;
		if(small && kiter >= 2) return V_DOT_U1;
		if(dt.size == 1) return kiter <= 2 ? V_DOT_U4 : V_DOT_U2;   // 16 decoded values per vector
		if(kiter <= 1) return V_DOT_U8;
		if(kiter <= 2) return V_DOT_U4;
		if(kiter <= 4) return V_DOT_U2;
		return V_DOT_U2;
	}
	if(f.I > 1 && f.sAi == 1 && f.sCi == 1) {
		if(configure(V_COL_SPLIT, f, dt, a_addr, cus, flags, cfg)) return V_COL_SPLIT;
		configure(V_COL_U1, f, dt, a_addr, cus, flags, cfg);
		if(small) return cfg.args.items >= cus ? V_COL_U1 : V_COL_NARROW;
		int64_t const iv = ceil_div(f.I, cfg.vec);
		// fp8: a 16-byte vector already holds 16 columns per lane
		if(dt.size == 1) return iv >= 4 * kBlock ? V_COL_U2 : V_COL_NARROW;
		if(iv >= 4 * kBlock) return V_COL_U4;            // wide rows: 4 vectors per lane along i
		int64_t const per_thread = f.K / std::max(1, cfg.args.KS);
		if(per_thread >= 8) return V_COL_U2;
		return V_COL_U4;
	}
	return V_STRIDED;
}

namespace
{
std::atomic<int> g_sel_mode{TTVX_SELECT_AUTO};
std::atomic<int> g_forced{0};
} // namespace

int set_selection(int mode, int forced)
{
	if(mode != TTVX_SELECT_HEURISTIC && mode != TTVX_SELECT_LEARNED && mode != TTVX_SELECT_FORCED &&
	   mode != TTVX_SELECT_AUTO)
		return fail(TTVX_ERR_ARGUMENT, "ttvx: unknown selection mode");
	if(mode == TTVX_SELECT_FORCED && (forced < 0 || forced >= V_COUNT))
		return fail(TTVX_ERR_ARGUMENT, "ttvx: forced variant out of range");
	g_forced.store(forced);
	g_sel_mode.store(mode);
	return TTVX_OK;
}

int select_variant(const Fold& f, const DType& dt, uintptr_t a_addr, unsigned flags)
{
	int const mode = g_sel_mode.load();
	if(mode == TTVX_SELECT_FORCED) {
		Config cfg;
		int const v = g_forced.load();
		if(configure(v, f, dt, a_addr, 256, flags, cfg)) return v;
	} else if(mode == TTVX_SELECT_LEARNED || mode == TTVX_SELECT_AUTO) {
		if(mode == TTVX_SELECT_AUTO) load_default_selector_weights();
		int const v = learned_variant(f, dt, a_addr, flags);
		if(v >= 0) return v;
	}
	return heuristic_variant(f, dt, a_addr, flags);
}

void features_of(const Fold& f, const DType& dt, uintptr_t a_addr, float* out)
{
	for(int i = 0; i < TTVX_NUM_FEATURES; ++i) out[i] = 0.f;
	auto lg = [](double x) { return x > 0 ? static_cast<float>(std::log2(x)) : 0.f; };
	int64_t N = 1;
	for(auto n : f.n_layout) N *= n;
	out[0] = static_cast<float>(f.p);
	out[1] = f.p > 1 ? static_cast<float>(f.qpos) / static_cast<float>(f.p - 1) : 0.f;
	out[2] = lg(static_cast<double>(f.I));
	out[3] = lg(static_cast<double>(f.K));
	out[4] = lg(static_cast<double>(f.O));
	out[5] = lg(static_cast<double>(N));
	out[6] = dt.code == TTVX_F32 ? 1.f : 0.f;
	out[7] = dt.code == TTVX_BF16 ? 1.f : 0.f;
	out[8] = dt.code == TTVX_FP8E4M3 ? 1.f : 0.f;
	out[9] = dt.code == TTVX_F64 ? 1.f : 0.f;
	out[10] = f.I == 1 ? 1.f : 0.f;
	out[11] = static_cast<float>(aligned_vec(dt, a_addr, f.I > 1 ? f.I : f.K, {})) / static_cast<float>(dt.vmax);
	out[12] = f.three_d ? 1.f : 0.f;
	out[13] = lg(static_cast<double>(N) * dt.size);
	out[23] = static_cast<float>(f.qpos);
	if(!f.three_d) return;
	// tiling of the candidate kernels: parallelism, tile utilisation, work per lane
	Config c1, c2, c4;
	int const cus = 256;
	if(configure(V_COL_U1, f, dt, a_addr, cus, 0u, c1)) {
		int64_t const iv = ceil_div(f.I, c1.vec);
		out[14] = lg(static_cast<double>(c1.args.items));
		out[17] = static_cast<float>(static_cast<double>(iv) / static_cast<double>(c1.args.iTiles * c1.args.IT));
		if(configure(V_COL_U2, f, dt, a_addr, cus, 0u, c2))
			out[16] = static_cast<float>(static_cast<double>(f.O * iv) /
			                             static_cast<double>(c2.args.items * c2.args.OS * c2.U * c2.args.IT));
		if(configure(V_COL_U4, f, dt, a_addr, cus, 0u, c4))
			out[15] = static_cast<float>(static_cast<double>(f.O * iv) /
			                             static_cast<double>(c4.args.items * c4.args.OS * c4.U * c4.args.IT));
		out[18] = lg(static_cast<double>(iv));
		out[19] = lg(static_cast<double>(f.K) / c1.args.KS);
		out[20] = static_cast<float>(c1.args.red);
		out[21] = lg(static_cast<double>(f.O) * f.I);
		out[22] = lg(static_cast<double>(f.K) * f.I * dt.size);
	} else if(configure(V_DOT_U1, f, dt, a_addr, cus, 0u, c1)) {
		int64_t const vsteps = (f.K + c1.vec - 1) / c1.vec;
		out[14] = lg(static_cast<double>(c1.args.groups));
		out[15] = out[16] = out[17] = static_cast<float>(static_cast<double>(vsteps) /
		                                                 static_cast<double>(c1.args.LPR * c1.args.kiter));
		out[18] = lg(static_cast<double>(vsteps));
		out[19] = static_cast<float>(c1.args.kiter);
		out[20] = lg(static_cast<double>(c1.args.LPR));
		out[21] = lg(static_cast<double>(f.O));
		out[22] = lg(static_cast<double>(f.K) * dt.size);
	}
}

// ---- launch ----------------------------------------------------------------------------------
void set_output_stores(Config& cfg, const Fold& f, const DType& dt, uintptr_t c_addr)
{
	if(cfg.kind == KIND_COL) {
		// largest power-of-two store (<= 16 bytes) that every lane's VEC outputs start aligned to
		int64_t const ab = static_cast<int64_t>(acc_size(dt));
		int chunk = 16;
		while(chunk > static_cast<int>(ab) &&
		      (c_addr % static_cast<uintptr_t>(chunk) != 0 || (f.O > 1 && (f.sCo * ab) % chunk != 0) ||
		       (cfg.vec * ab) % chunk != 0))
			chunk >>= 1;
		cfg.args.cvec = chunk > static_cast<int>(ab) ? chunk : 0;
		// wave-transposed stores: no reduction, a whole wave along i, lane vectors wider than 16
		// bytes in an instantiation that carries them, 16-byte aligned C, no split partials
		cfg.args.tstore = (cfg.args.red == 0 && cfg.args.IT % 64 == 0 && cfg.args.cvec >= 16 && f.I % cfg.vec == 0 &&
		                   tstore_compiled(static_cast<int>(cfg.vec * ab), cfg.U) && cfg.ksplit == 1 &&
		                   !cfg.args.part64) ? 1 : 0;
		// their LDS staging: one run of 64*VEC outputs per wave
		if(cfg.args.tstore)
			cfg.shmem = std::max<std::size_t>(cfg.shmem, static_cast<std::size_t>(kBlock) * cfg.vec * static_cast<std::size_t>(ab));
	}

}

namespace
{
// Split partial sums: one grow-only device buffer per (device, stream), reused in stream order by
// every launch on that stream and held until the process ends (no allocation once it has grown to
// the largest partials the stream needed).
struct PartialBuffers {
	std::mutex mutex;
	std::map<hipStream_t, std::pair<void*, std::size_t>> by_stream;
};
PartialBuffers g_partials[64];
} // namespace

int partial_buffer(int device, hipStream_t stream, std::size_t bytes, void** out)
{
	if(injected(TTVX_FAULT_ALLOC)) return hip_fail(hipErrorOutOfMemory, "hipMalloc (split partials)");
	PartialBuffers& pb = g_partials[device & 63];
	std::lock_guard<std::mutex> lock(pb.mutex);
	auto& slot = pb.by_stream[stream];
	if(slot.second < bytes) {
		if(slot.first != nullptr) {
			// earlier launches on this stream may still read the old buffer
			hipError_t e = hipStreamSynchronize(stream);
			if(e != hipSuccess) return hip_fail(e, "hipStreamSynchronize (growing the split partials)");
			(void)hipFree(slot.first);
			slot = {nullptr, 0};
		}
		std::size_t const want = std::max<std::size_t>(bytes, std::size_t{1} << 20);
		hipError_t e = hipMalloc(&slot.first, want);
		if(e != hipSuccess) {
			slot = {nullptr, 0};
			return hip_fail(e, "hipMalloc (split partials)");
		}
		slot.second = want;
	}
	*out = slot.first;
	return TTVX_OK;
}

int launch(const Fold& f, const DType& dt, const void* a, const void* b, void* c,
           const float* scales, int64_t fps, unsigned flags, int variant, hipStream_t stream, const Scratch* scratch)
{
	DeviceInfo di;
	if(!current_device(di)) return TTVX_ERR_DEVICE;
	uintptr_t const a_addr = reinterpret_cast<uintptr_t>(a);
	if(variant < 0) variant = select_variant(f, dt, a_addr, flags);
	Config cfg;
	if(!configure(variant, f, dt, a_addr, di.cus, flags, cfg)) {
		variant = heuristic_variant(f, dt, a_addr, flags);
		if(!configure(variant, f, dt, a_addr, di.cus, flags, cfg))
			return fail(TTVX_ERR_UNSUPPORTED, "ttvx: no kernel variant can run this problem");
	}
	bool const kahan = (flags & TTVX_FLAG_KAHAN) != 0;
	set_output_stores(cfg, f, dt, reinterpret_cast<uintptr_t>(c));

	if(cfg.kind == KIND_STRIDED) {
		cfg.sargs.scales = scales;
		cfg.sargs.fps    = fps > 0 ? fps : 1;
	}
	bool const use_partial = cfg.kind != KIND_STRIDED && (cfg.ksplit > 1 || cfg.args.part64);
	void* partial = nullptr;
	if(use_partial) {
		std::size_t const esz = cfg.args.part64 ? 8 : acc_size(dt);
		std::size_t const bytes = static_cast<std::size_t>(cfg.ksplit) * static_cast<std::size_t>(f.O * f.I) * esz;
		if(scratch != nullptr) {
			if(bytes > scratch->bytes)
				return fail(TTVX_ERR_UNSUPPORTED, "ttvx: the split partial sums need " + std::to_string(bytes) +
				                                      " bytes, more than the " + std::to_string(scratch->bytes) +
				                                      "-byte static scratch buffer holds");
			partial = scratch->ptr;
		} else {
			int const rc = partial_buffer(di.id, stream, bytes, &partial);
			if(rc != TTVX_OK) return rc;
		}
		cfg.args.partial = partial;
	} else if(cfg.kind != KIND_STRIDED) {
		cfg.args.scales = scales;
		cfg.args.fps    = fps > 0 ? fps : 1;
	}

	// 2- and 1-byte products whose vectors do not all start on a 4-byte boundary run the
	// realigned-load kernels (loads from the 4-byte boundary, funnel-shifted):
	// * rows: a row vector starts at o*sAo + k, k a multiple of the width or K - width for the
	//   clamped last one (MI355X, BF16 rows of 129 elements: 3.65 -> 4.29 TB/s, FP8: 2.68 -> 2.95);
	// * columns whose i-groups span at least a wave (the dword after a lane's vector comes from the
	//   next lane): a slab row starts at o*sAo + k*sAk (MI355X, K = 16, I = 1025..65537: BF16
	//   +2-8 %, FP8 +2-8 %).  Narrower i-groups keep the element-aligned loads: more of their lanes
	//   sit at a group end and load the extra dword themselves (I = 129: -4 % BF16, -18 % FP8).
	uint64_t const a_low = reinterpret_cast<uintptr_t>(a) | static_cast<uint64_t>(f.sAo * dt.size);
	bool const sub4 = dt.size < 4 && cfg.vec * dt.size >= 4 &&
	                  ((cfg.kind == KIND_DOT && ((a_low | static_cast<uint64_t>(f.K * dt.size)) & 3u) != 0) ||
	                   (cfg.kind == KIND_COL && cfg.lean != LEAN_SLAB && !cfg.args.one_step && cfg.args.IT >= 64 &&
	                    ((a_low | static_cast<uint64_t>(f.sAk * dt.size)) & 3u) != 0));
	if(injected(TTVX_FAULT_LAUNCH)) return hip_fail(hipErrorLaunchFailure, "kernel launch");
	int rc = TTVX_OK;
	switch(dt.code) {
		case TTVX_F32:     rc = launch_f32 (cfg, a, b, c, kahan, stream); break;
		case TTVX_F64:     rc = launch_f64 (cfg, a, b, c, kahan, stream); break;
		case TTVX_BF16:    rc = sub4 ? launch_bf16_sub(cfg, a, b, c, kahan, stream) : launch_bf16(cfg, a, b, c, kahan, stream); break;
		case TTVX_FP8E4M3: rc = sub4 ? launch_fp8_sub(cfg, a, b, c, kahan, stream) : launch_fp8(cfg, a, b, c, kahan, stream); break;
		default:           rc = fail(TTVX_ERR_UNSUPPORTED, "ttvx: unknown element type");
	}
	if(rc == TTVX_OK && use_partial) {
		int64_t const n = f.O * f.I;
		unsigned const grid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kBlock), 4096)));
		if(cfg.ksplit > 64) {
			// many splits of few outputs: one block sums each output's partials
			unsigned const gw = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(n, 4096)));
			if(dt.code == TTVX_F64)
				hipLaunchKernelGGL((k_finalize_wide<double>), dim3(gw), dim3(kBlock), 0, stream, partial, cfg.args.part64,
				                   cfg.ksplit, f.O, f.I, f.sCo, c, scales, fps > 0 ? fps : 1);
			else
				hipLaunchKernelGGL((k_finalize_wide<float>), dim3(gw), dim3(kBlock), 0, stream, partial, cfg.args.part64,
				                   cfg.ksplit, f.O, f.I, f.sCo, c, scales, fps > 0 ? fps : 1);
		} else if(dt.code == TTVX_F64)
			hipLaunchKernelGGL((k_finalize<double>), dim3(grid), dim3(kBlock), 0, stream, partial, cfg.args.part64,
			                   cfg.ksplit, f.O, f.I, f.sCo, c, scales, fps > 0 ? fps : 1);
		else
			hipLaunchKernelGGL((k_finalize<float>), dim3(grid), dim3(kBlock), 0, stream, partial, cfg.args.part64,
			                   cfg.ksplit, f.O, f.I, f.sCo, c, scales, fps > 0 ? fps : 1);
		hipError_t e = hipGetLastError();
		if(e != hipSuccess) rc = hip_fail(e, "finalize launch");
	}
	return rc;
}

} // namespace ttvx
