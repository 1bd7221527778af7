// Batched mode-q products: many independent TTVs of different order, shape, mode and layout in
// one persistent launch (no counterpart in the reference, which runs one product per call).
//
// At creation every problem is validated and folded (plan.hpp), gets the launch configuration of
// its fixed-rule variant restricted to the batch instantiations (column tiles with U = 1, row
// groups with U = 2; no split of n_q), and is cut into block tiles.  Problems are ordered by tile
// cost, largest first, and their tiles numbered consecutively.  The persistent kernel's blocks
// take tile numbers from a device counter (dynamic load balance), find the owning problem by
// binary search over the tile offsets and run the same tile bodies as the single-problem kernels,
// so a problem's result does not depend on whether it ran batched or alone.  A fill kernel
// clears the counter before every launch; both kernels can be captured once into a hipGraph
// and replayed.
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

struct BatchDesc {
	Args args;
	const float* a;
	const float* b;
	float* c;
	int64_t tile0;     // first tile number of this problem
	int32_t kind;      // KIND_DOT or KIND_COL
	int32_t vec;
};

constexpr int kBatchDotU = 2;

__global__ __launch_bounds__(kBlock) void k_batch_reset(unsigned int* counter) { *counter = 0u; }

template<int VEC>
__device__ __forceinline__ void batch_dot_tile(const BatchDesc& d, int64_t item)
{
	int64_t const g = item * (kBlock / 64) + (threadIdx.x >> 6);
	if(g >= d.args.groups) return;
	if(d.args.kiter <= dot_breg(VEC)) {
		float bb[dot_breg(VEC)][VEC];
		dot_load_b<float, VEC>(d.b, d.args, 0, d.args.K, bb);
		dot_group_short<float, VEC, kBatchDotU, false>(d.a, d.c, d.args, g, 0, d.args.K, bb);
	} else {
		dot_group_long<float, VEC, kBatchDotU, false>(d.a, d.b, d.c, d.args, g, 0, d.args.K);
	}
}

__global__ __launch_bounds__(kBlock) void k_batch(const BatchDesc* __restrict__ descs, int nprob, int64_t total_tiles,
                                                  unsigned int* __restrict__ counter)
{
	__shared__ float red[kBlock * 4];
	__shared__ int64_t s_tile;
	for(;;) {
		if(threadIdx.x == 0) s_tile = static_cast<int64_t>(atomicAdd(counter, 1u));
		__syncthreads();
		int64_t const t = s_tile;
		__syncthreads();
		if(t >= total_tiles) break;
		// owning problem: last descriptor with tile0 <= t
		int lo = 0, hi = nprob - 1;
		while(lo < hi) {
			int const mid = (lo + hi + 1) >> 1;
			if(descs[mid].tile0 <= t) lo = mid;
			else hi = mid - 1;
		}
		BatchDesc const d = descs[lo];
		int64_t const item = t - d.tile0;
		if(d.kind == KIND_COL) {
			switch(d.vec) {
				case 4:  col_item<float, 4, 1, false>(d.a, d.b, d.c, d.args, item, 0, d.args.K, red); break;
				case 2:  col_item<float, 2, 1, false>(d.a, d.b, d.c, d.args, item, 0, d.args.K, red); break;
				default: col_item<float, 1, 1, false>(d.a, d.b, d.c, d.args, item, 0, d.args.K, red); break;
			}
		} else {
			switch(d.vec) {
				case 4:  batch_dot_tile<4>(d, item); break;
				case 2:  batch_dot_tile<2>(d, item); break;
				default: batch_dot_tile<1>(d, item); break;
			}
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
	unsigned int* counter = nullptr;
	hipStream_t capture_stream = nullptr;
	hipGraphExec_t graph_exec = nullptr;
	hipGraph_t graph = nullptr;
	std::vector<int> order;            // descriptor slot -> caller's problem index
};

namespace
{
int enqueue(Batch& bt, hipStream_t s)
{
	hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(1), 0, s, bt.counter);
	hipLaunchKernelGGL(k_batch, dim3(static_cast<unsigned>(bt.grid)), dim3(kBlock), 0, s, bt.descs, bt.nprob,
	                   bt.total_tiles, bt.counter);
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
		int64_t tile_cost;
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
		uintptr_t const aaddr = reinterpret_cast<uintptr_t>(pr.a);
		Config cfg;
		Item it;
		std::memset(&it, 0, sizeof(it));
		if(configure(V_COL_U1, f, dt, aaddr, di.cus, 0u, cfg)) {
			it.d.kind = KIND_COL;
			it.tiles  = cfg.args.items;
			it.tile_cost = static_cast<int64_t>(cfg.args.OS) * cfg.args.IT * cfg.vec * f.K;
		} else if(configure(V_DOT_U2, f, dt, aaddr, di.cus, 0u, cfg)) {
			it.d.kind = KIND_DOT;
			it.tiles  = (cfg.args.groups + kBlock / 64 - 1) / (kBlock / 64);
			it.tile_cost = static_cast<int64_t>(kBlock / 64) * (64 / cfg.args.LPR) * kBatchDotU * f.K;
		} else {
			return fail(TTVX_ERR_UNSUPPORTED, where + "the batched kernel needs A and C to fold to three dense-inner dimensions");
		}
		if(cfg.kind == KIND_COL && cfg.args.red == 2 && cfg.vec > 4)
			return fail(TTVX_ERR_UNSUPPORTED, where + "tile does not fit the batched kernel");
		it.d.args = cfg.args;
		it.d.args.kchunk  = f.K;
		it.d.args.partial = nullptr;
		it.d.args.part64  = 0;
		it.d.args.scales  = nullptr;
		it.d.args.fps     = 1;
		it.d.vec = cfg.vec;
		it.d.a = pr.a;
		it.d.b = pr.b;
		it.d.c = pr.c;
		it.index = static_cast<int>(n);
		items.push_back(it);
		bytes += f.outputs * f.K * 4 + f.outputs * 4 + f.K * 4;
	}
	std::stable_sort(items.begin(), items.end(), [](Item const& x, Item const& y) { return x.tile_cost > y.tile_cost; });

	auto* bt = new Batch();
	bt->device = di.id;
	bt->nprob = static_cast<int>(items.size());
	bt->bytes = bytes;
	std::vector<BatchDesc> host(items.size());
	int64_t tile = 0;
	for(size_t n = 0; n < items.size(); ++n) {
		host[n] = items[n].d;
		host[n].tile0 = tile;
		tile += items[n].tiles;
		bt->order.push_back(items[n].index);
	}
	bt->total_tiles = tile;
	if(tile >= (int64_t{1} << 32) - kBlock * 8192) {
		delete bt;
		return fail(TTVX_ERR_UNSUPPORTED, "ttvx_batch_create: too many tiles for one launch");
	}
	bt->grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(tile, static_cast<int64_t>(di.cus) * 8)));
	size_t const dbytes = host.size() * sizeof(BatchDesc);
	size_t const counter_off = (dbytes + 255) / 256 * 256;     // counter on its own 256-byte line
	hipError_t e = hipMalloc(&bt->dmem, counter_off + 256);
	if(e != hipSuccess) {
		delete bt;
		return hip_fail(e, "hipMalloc (batch descriptors)");
	}
	bt->descs = static_cast<BatchDesc*>(bt->dmem);
	bt->counter = reinterpret_cast<unsigned int*>(static_cast<char*>(bt->dmem) + counter_off);
	e = hipMemcpy(bt->descs, host.data(), dbytes, hipMemcpyHostToDevice);
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
		// capture once on a private stream (the legacy default stream cannot be captured)
		hipError_t e = hipStreamCreateWithFlags(&bt->capture_stream, hipStreamNonBlocking);
		if(e != hipSuccess) return hip_fail(e, "hipStreamCreateWithFlags");
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
