// Batched mode-q products: many independent TTVs of different order, shape, mode and layout in
// one persistent launch (no counterpart in the reference, which runs one product per call).
//
// At creation every problem is validated and folded (plan.hpp), and the library's selection (the
// tuning table, the learned selector, the fixed rule: the same choice a single call makes) picks its
// kernel variant; the batch keeps that variant's tile configuration (column items of 1, 2 or 4
// slab units, 256-byte column chunks, row groups of 1-8 row sets, all vector widths), with the
// contraction never split.  A problem's work units (column items or row groups) are cut into batch
// tiles of about kTileBytes of A each, so a tile costs one counter increment and one table search
// however small its units are.  Problems are ordered largest first and their tiles numbered
// consecutively.  The persistent kernel runs as many blocks as the GPU holds at once; each takes
// tile numbers from a device counter (dynamic load balance), finds the owning problem by binary
// search over the tile offsets and runs the tile body of the problem's variant, the same device
// function the single-problem kernel runs, so a problem's result does not depend on whether it ran
// batched or alone.  A fill kernel clears the counter before every launch; both kernels can be
// captured once into a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/ttvx.h"
#include "kernels.hpp"
#include "plan.hpp"
#include "runtime.hpp"

namespace ttvx
{

// tile body of a problem
enum BatchBody { BB_COL = 0, BB_SLAB = 1, BB_ROWS = 2, BB_DOT = 3 };

struct BatchDesc {
	Args args;
	const float* a;
	const float* b;
	float* c;
	int64_t units;     // column items, or row groups
	int64_t per_tile;  // units per batch tile (a multiple of the block's waves for row groups)
	int32_t body;      // BatchBody
	int32_t U;         // slab units (BB_COL) or row sets (BB_ROWS, BB_DOT) per unit
	int32_t vec;
	int32_t variant;   // the selected variant (ttvx_batch_variant)
};

constexpr int64_t kTileBytes = int64_t{1} << 20;
constexpr int kBatchRed = kBlock * 4 * 4;      // LDS floats: column items of U*VEC <= 16 partials per lane

__global__ __launch_bounds__(kBlock) void k_batch_reset(unsigned int* counter) { *counter = 0u; }

template<int VEC, int U>
__device__ __forceinline__ void batch_col(const BatchDesc& d, int64_t ub, int64_t ue, float* red)
{
	if(d.args.one_step) {
		col_items_pipelined<float, VEC, U>(d.a, d.b, d.c, d.args, ub, ue, 1, 0, d.args.K, red);
		return;
	}
	for(int64_t item = ub; item < ue; ++item)
		col_item<float, VEC, U, false>(d.a, d.b, d.c, d.args, item, 0, d.args.K, red);
}

template<int VEC, int U>
__device__ __forceinline__ void batch_rows(const BatchDesc& d, int64_t gb, int64_t ge)
{
	for(int64_t g = gb + (threadIdx.x >> 6); g < ge; g += kBlock / 64)
		rows_group<float, VEC, U, 0, false>(d.a, d.b, d.c, d.args, g);
}

template<int VEC, int U>
__device__ __forceinline__ void batch_dot(const BatchDesc& d, int64_t gb, int64_t ge)
{
	if(d.args.kiter <= dot_breg(VEC)) {
		float bb[dot_breg(VEC)][VEC];
		dot_load_b<float, VEC>(d.b, d.args, 0, d.args.K, bb);
		for(int64_t g = gb + (threadIdx.x >> 6); g < ge; g += kBlock / 64)
			dot_group_short<float, VEC, U, false>(d.a, d.c, d.args, g, 0, d.args.K, bb);
	} else {
		for(int64_t g = gb + (threadIdx.x >> 6); g < ge; g += kBlock / 64)
			dot_group_long<float, VEC, U, false>(d.a, d.b, d.c, d.args, g, 0, d.args.K);
	}
}

template<int VEC>
__device__ __forceinline__ void batch_tile(const BatchDesc& d, int64_t ub, int64_t ue, float* red)
{
	switch(d.body) {
		case BB_COL:
			if(d.U == 1) batch_col<VEC, 1>(d, ub, ue, red);
			else if(d.U == 2) batch_col<VEC, 2>(d, ub, ue, red);
			else batch_col<VEC, 4>(d, ub, ue, red);
			break;
		case BB_SLAB:
			for(int64_t item = ub; item < ue; ++item)
				slab_item<float, VEC, false>(d.a, d.b, d.c, d.args, static_cast<uint32_t>(item),
				                             reinterpret_cast<float (*)[64 * VEC]>(red));
			break;
		case BB_ROWS:
			if(d.U == 1) batch_rows<VEC, 1>(d, ub, ue);
			else if(d.U == 2) batch_rows<VEC, 2>(d, ub, ue);
			else if(d.U == 4) batch_rows<VEC, 4>(d, ub, ue);
			else batch_rows<VEC, 8>(d, ub, ue);
			break;
		default:
			if(d.U == 1) batch_dot<VEC, 1>(d, ub, ue);
			else if(d.U == 2) batch_dot<VEC, 2>(d, ub, ue);
			else if(d.U == 4) batch_dot<VEC, 4>(d, ub, ue);
			else batch_dot<VEC, 8>(d, ub, ue);
			break;
	}
}

__global__ __launch_bounds__(kBlock) void k_batch(const BatchDesc* __restrict__ descs, const int64_t* __restrict__ tile0,
                                                  int nprob, int64_t total_tiles, unsigned int* __restrict__ counter)
{
	__shared__ float red[kBatchRed];
	__shared__ int64_t s_tile;
	for(;;) {
		if(threadIdx.x == 0) s_tile = static_cast<int64_t>(atomicAdd(counter, 1u));
		__syncthreads();
		int64_t const t = s_tile;
		__syncthreads();
		if(t >= total_tiles) break;
		// owning problem: last table entry with tile0 <= t
		int lo = 0, hi = nprob - 1;
		while(lo < hi) {
			int const mid = (lo + hi + 1) >> 1;
			if(tile0[mid] <= t) lo = mid;
			else hi = mid - 1;
		}
		BatchDesc const& d = descs[lo];
		int64_t const ub = (t - tile0[lo]) * d.per_tile;
		int64_t const ue = ub + d.per_tile < d.units ? ub + d.per_tile : d.units;
		switch(d.vec) {
			case 4:  batch_tile<4>(d, ub, ue, red); break;
			case 2:  batch_tile<2>(d, ub, ue, red); break;
			default: batch_tile<1>(d, ub, ue, red); break;
		}
	}
}

struct Batch {
	int device = -1;
	int nprob = 0;
	int64_t total_tiles = 0;
	int64_t bytes = 0;
	int grid = 0;
	void* dmem = nullptr;              // descriptors followed by the tile counter
	BatchDesc* descs = nullptr;
	int64_t* tile0 = nullptr;
	unsigned int* counter = nullptr;
	hipStream_t capture_stream = nullptr;
	hipGraphExec_t graph_exec = nullptr;
	hipGraph_t graph = nullptr;
	std::vector<int> order;            // descriptor slot -> caller's problem index
	std::vector<int> variant;          // caller's problem index -> variant of its tiles
};

namespace
{
int enqueue(Batch& bt, hipStream_t s)
{
	hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(1), 0, s, bt.counter);
	hipLaunchKernelGGL(k_batch, dim3(static_cast<unsigned>(bt.grid)), dim3(kBlock), 0, s, bt.descs, bt.tile0,
	                   bt.nprob, bt.total_tiles, bt.counter);
	hipError_t e = hipGetLastError();
	return e == hipSuccess ? TTVX_OK : hip_fail(e, "batch launch");
}
} // namespace

} // namespace ttvx

using namespace ttvx;

extern "C" {

int ttvx_batch_create(size_t count, const ttvx_batch_problem* problems, void** out)
{
	clear_error();
	if(out == nullptr) return fail(TTVX_ERR_ARGUMENT, "ttvx_batch_create: null output handle");
	*out = nullptr;
	if(count == 0 || problems == nullptr) return fail(TTVX_ERR_ARGUMENT, "ttvx_batch_create: empty batch");
	DeviceInfo di;
	if(!current_device(di)) return TTVX_ERR_DEVICE;
	DType const dt = dtype_of(TTVX_F32);

	struct Item {
		BatchDesc d;
		int64_t tiles;
		int64_t cost;
		int index;
	};
	std::vector<Item> items;
	items.reserve(count);
	int64_t bytes = 0;
	for(size_t n = 0; n < count; ++n) {
		ttvx_batch_problem const& pr = problems[n];
		std::string const where = "ttvx_batch_create: problem " + std::to_string(n) + ": ";
		Check ck = validate_a(pr.q, pr.p, pr.na, pr.wa, pr.pia);
		if(ck.status != TTVX_OK) return fail(ck.status, where + ck.message);
		if(pr.p < 2) return fail(TTVX_ERR_ARGUMENT, where + "Error in tlib::tensor_times_vector: shape vector of C is not valid.");
		if(pr.a == nullptr || pr.b == nullptr || pr.c == nullptr) return fail(TTVX_ERR_ARGUMENT, where + "null data pointer");
		Fold const f = fold(pr.q, pr.p, pr.na, pr.wa, pr.pia, pr.wc, nullptr);
		int rc = check_device_span(pr.a, static_cast<size_t>(f.a_span) * 4, "A");
		if(rc == TTVX_OK) rc = check_device_span(pr.b, static_cast<size_t>(f.K) * 4, "b");
		if(rc == TTVX_OK) rc = check_device_span(pr.c, static_cast<size_t>(f.c_span) * 4, "C");
		if(rc != TTVX_OK) return fail(rc, where + ttvx_last_error());
		if(f.K > (int64_t{1} << 20))
			return fail(TTVX_ERR_UNSUPPORTED, where + "the batched kernel does not split contractions longer than 2^20; "
			                                          "run such products alone (they use an FP64 final reduction)");
		uintptr_t const aaddr = reinterpret_cast<uintptr_t>(pr.a);
		// the variant a single call would take, with the contraction never split
		int v = select_variant(f, dt, aaddr, 0u);
		if(v == V_COL_SPLIT) v = V_COL_U1;
		if(v == V_DOT_SPLIT) v = V_DOT_U2;
		Config cfg;
		if(v == V_STRIDED || !configure(v, f, dt, aaddr, di.cus, 0u, cfg) || cfg.ksplit != 1 || cfg.args.part64)
			return fail(TTVX_ERR_UNSUPPORTED, where + "the batched kernel needs A and C to fold to three dense-inner dimensions");
		Item it;
		std::memset(&it, 0, sizeof(it));
		int64_t unit_bytes;
		if(cfg.lean == LEAN_SLAB) {
			it.d.body  = BB_SLAB;
			it.d.units = cfg.args.items;
			unit_bytes = static_cast<int64_t>(cfg.args.OS) * cfg.args.IT * cfg.vec * f.K * 4;
		} else if(cfg.kind == KIND_COL) {
			it.d.body  = BB_COL;
			it.d.units = cfg.args.items;
			unit_bytes = static_cast<int64_t>(cfg.args.OS) * cfg.U * cfg.args.IT * cfg.vec * f.K * 4;
		} else {
			it.d.body  = cfg.lean == LEAN_ROWS ? BB_ROWS : BB_DOT;
			it.d.units = cfg.args.groups;
			unit_bytes = static_cast<int64_t>(64 / cfg.args.LPR) * cfg.U * f.K * 4;
		}
		if(cfg.vec > 4 || cfg.shmem > sizeof(float) * kBatchRed || (cfg.lean == LEAN_SLAB && 4 * 64 * cfg.vec > kBatchRed))
			return fail(TTVX_ERR_UNSUPPORTED, where + "tile does not fit the batched kernel");
		it.d.per_tile = std::max<int64_t>(1, kTileBytes / std::max<int64_t>(unit_bytes, 1));
		if(it.d.body == BB_ROWS || it.d.body == BB_DOT)
			it.d.per_tile = (it.d.per_tile + kBlock / 64 - 1) / (kBlock / 64) * (kBlock / 64);
		it.tiles = (it.d.units + it.d.per_tile - 1) / it.d.per_tile;
		it.cost  = f.outputs * f.K;
		set_output_stores(cfg, f, dt, reinterpret_cast<uintptr_t>(pr.c));
		it.d.args = cfg.args;
		it.d.args.kchunk  = f.K;
		it.d.args.partial = nullptr;
		it.d.args.part64  = 0;
		it.d.args.scales  = nullptr;
		it.d.args.fps     = 1;
		it.d.args.tstore  = 0;
		it.d.U = cfg.U;
		it.d.variant = v;
		it.d.vec = cfg.vec;
		it.d.a = pr.a;
		it.d.b = pr.b;
		it.d.c = pr.c;
		it.index = static_cast<int>(n);
		items.push_back(it);
		bytes += f.outputs * f.K * 4 + f.outputs * 4 + f.K * 4;
	}
	std::stable_sort(items.begin(), items.end(), [](Item const& x, Item const& y) { return x.cost > y.cost; });

	auto* bt = new Batch();
	bt->device = di.id;
	bt->nprob = static_cast<int>(items.size());
	bt->bytes = bytes;
	bt->variant.assign(items.size(), -1);
	std::vector<BatchDesc> host(items.size());
	std::vector<int64_t> tile0(items.size());
	int64_t tile = 0;
	for(size_t n = 0; n < items.size(); ++n) {
		host[n]  = items[n].d;
		tile0[n] = tile;
		tile += items[n].tiles;
		bt->order.push_back(items[n].index);
		bt->variant[items[n].index] = items[n].d.variant;
	}
	bt->total_tiles = tile;
	// every block draws once past the end: the counter must not wrap
	if(tile >= (int64_t{1} << 32) - kBlock * 8192) {
		delete bt;
		return fail(TTVX_ERR_UNSUPPORTED, "ttvx_batch_create: too many tiles for one launch");
	}
	int resident = 0;
	if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, k_batch, kBlock, 0) != hipSuccess || resident < 1) resident = 1;
	bt->grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(tile, static_cast<int64_t>(di.cus) * resident)));
	size_t const dbytes  = host.size() * sizeof(BatchDesc);
	size_t const t0_off  = (dbytes + 255) / 256 * 256;
	size_t const tbytes  = tile0.size() * sizeof(int64_t);
	size_t const cnt_off = t0_off + (tbytes + 255) / 256 * 256;     // counter on its own 256-byte line
	hipError_t e = hipMalloc(&bt->dmem, cnt_off + 256);
	if(e != hipSuccess) {
		delete bt;
		return hip_fail(e, "hipMalloc (batch descriptors)");
	}
	bt->descs   = static_cast<BatchDesc*>(bt->dmem);
	bt->tile0   = reinterpret_cast<int64_t*>(static_cast<char*>(bt->dmem) + t0_off);
	bt->counter = reinterpret_cast<unsigned int*>(static_cast<char*>(bt->dmem) + cnt_off);
	e = hipMemcpy(bt->descs, host.data(), dbytes, hipMemcpyHostToDevice);
	if(e == hipSuccess) e = hipMemcpy(bt->tile0, tile0.data(), tbytes, hipMemcpyHostToDevice);
	if(e != hipSuccess) {
		(void)hipFree(bt->dmem);
		delete bt;
		return hip_fail(e, "hipMemcpy (batch descriptors)");
	}
	*out = bt;
	return TTVX_OK;
}

int ttvx_batch_launch(void* handle, int use_graph, void* stream)
{
	clear_error();
	auto* bt = static_cast<Batch*>(handle);
	if(bt == nullptr) return fail(TTVX_ERR_ARGUMENT, "ttvx_batch_launch: null handle");
	hipStream_t s = static_cast<hipStream_t>(stream);
	if(!use_graph) return enqueue(*bt, s);
	if(bt->graph_exec == nullptr) {
		// capture once on a private stream (the legacy default stream cannot be captured); a failed
		// earlier capture leaves the stream for this attempt and its graph is dropped
		hipError_t e = hipSuccess;
		if(bt->capture_stream == nullptr) {
			e = hipStreamCreateWithFlags(&bt->capture_stream, hipStreamNonBlocking);
			if(e != hipSuccess) {
				bt->capture_stream = nullptr;
				return hip_fail(e, "hipStreamCreateWithFlags");
			}
		}
		if(bt->graph != nullptr) {
			(void)hipGraphDestroy(bt->graph);
			bt->graph = nullptr;
		}
		e = hipStreamBeginCapture(bt->capture_stream, hipStreamCaptureModeThreadLocal);
		if(e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
		int rc = enqueue(*bt, bt->capture_stream);
		e = hipStreamEndCapture(bt->capture_stream, &bt->graph);
		if(rc != TTVX_OK) return rc;
		if(e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
		e = hipGraphInstantiate(&bt->graph_exec, bt->graph, nullptr, nullptr, 0);
		if(e != hipSuccess) {
			bt->graph_exec = nullptr;
			return hip_fail(e, "hipGraphInstantiate");
		}
	}
	hipError_t e = hipGraphLaunch(bt->graph_exec, s);
	return e == hipSuccess ? TTVX_OK : hip_fail(e, "hipGraphLaunch");
}

int ttvx_batch_info(void* handle, int64_t* out4)
{
	clear_error();
	auto* bt = static_cast<Batch*>(handle);
	if(bt == nullptr || out4 == nullptr) return fail(TTVX_ERR_ARGUMENT, "ttvx_batch_info: null argument");
	out4[0] = bt->total_tiles;
	out4[1] = bt->bytes;
	out4[2] = bt->grid;
	out4[3] = bt->nprob;
	return TTVX_OK;
}

int ttvx_batch_variant(void* handle, size_t index)
{
	clear_error();
	auto* bt = static_cast<Batch*>(handle);
	if(bt == nullptr || index >= bt->variant.size()) {
		fail(TTVX_ERR_ARGUMENT, "ttvx_batch_variant: null handle or problem index out of range");
		return -1;
	}
	return bt->variant[index];
}

int ttvx_batch_destroy(void* handle)
{
	clear_error();
	auto* bt = static_cast<Batch*>(handle);
	if(bt == nullptr) return TTVX_OK;
	if(bt->graph_exec) (void)hipGraphExecDestroy(bt->graph_exec);
	if(bt->graph) (void)hipGraphDestroy(bt->graph);
	if(bt->capture_stream) (void)hipStreamDestroy(bt->capture_stream);
	if(bt->dmem) (void)hipFree(bt->dmem);
	delete bt;
	return TTVX_OK;
}

} // extern "C"
