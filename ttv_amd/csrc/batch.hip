// Batched mode-q products: many independent TTVs of different order, shape, mode and layout in a
// few launches (no counterpart in the reference, which runs one product per call).
//
// At creation every problem is validated and folded (plan.hpp), and the library's selection (the
// tuning table, the learned selector, the fixed rule: the same choice a single call makes) picks its
// kernel variant, with the contraction never split.  The problem keeps that variant's configuration
// and its block decomposition: one block per four row groups (k_rows), per 256-byte column chunk
// item (k_slab), per run of column items or row groups (k_col, k_dot) as the single launch's grid
// divides them.  Problems whose blocks run the same tile body (body, vector width, slab units or row
// sets, lanes per row) form a group, and a group is ONE launch whose grid is the concatenation of
// its problems' grids: block b reads its problem from a per-block owner table, then runs the same
// device function the single-problem kernel runs, so a problem's result does not depend on whether
// it ran batched or alone.  Each group kernel has the register budget of its own body (a single
// kernel holding every body would run at the occupancy of the largest), blocks stay short-lived
// (the dispatcher balances the load), and a batch of hundreds of problems costs a handful of
// launches instead of one launch, one ramp and one drain per problem.  The launches of a batch can
// be captured once into a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/ttvx.h"
#include "kernels.hpp"
#include "plan.hpp"
#include "runtime.hpp"

namespace ttvx
{

// tile body of a group
enum BatchBody { BB_COL = 0, BB_SLAB = 1, BB_ROWS = 2, BB_DOT = 3 };

struct BatchDesc {
	Args args;
	const float* a;
	const float* b;
	float* c;
	int64_t block0;    // first block of the problem in its group's grid
	int64_t per;       // column items or row groups per block (BB_COL, BB_DOT)
};

template<int VEC, int U>
__device__ __forceinline__ void group_col(const BatchDesc& d, const Args& P, int64_t lb, float* red)
{
	int64_t const ib = lb * d.per;
	int64_t const ie = ib + d.per < P.items ? ib + d.per : P.items;
	if(P.one_step) {
		col_items_pipelined<float, VEC, U>(d.a, d.b, d.c, P, ib, ie, 1, 0, P.K, red);
		return;
	}
	for(int64_t item = ib; item < ie; ++item) col_item<float, VEC, U, false>(d.a, d.b, d.c, P, item, 0, P.K, red);
}

template<int VEC, int U>
__device__ __forceinline__ void group_dot(const BatchDesc& d, const Args& P, int64_t lb)
{
	int64_t const gb = lb * d.per;
	int64_t const ge = gb + d.per < P.groups ? gb + d.per : P.groups;
	int64_t const g0 = gb + (threadIdx.x >> 6);
	if(P.kiter <= dot_breg(VEC)) {
		float bb[dot_breg(VEC)][VEC];
		dot_load_b<float, VEC>(d.b, P, 0, P.K, bb);
		for(int64_t g = g0; g < ge; g += kBlock / 64) dot_group_short<float, VEC, U, false>(d.a, d.c, P, g, 0, P.K, bb);
	} else {
		for(int64_t g = g0; g < ge; g += kBlock / 64) dot_group_long<float, VEC, U, false>(d.a, d.b, d.c, P, g, 0, P.K);
	}
}

// One launch of a group: block b runs local block b - block0 of problem owner[b].
template<int BODY, int VEC, int U, int LPRC>
__global__ __launch_bounds__(kBlock) void k_batch_group(const BatchDesc* __restrict__ descs,
                                                        const uint32_t* __restrict__ owner)
{
	BatchDesc const& d = descs[owner[blockIdx.x]];
	Args const P = d.args;
	int64_t const lb = static_cast<int64_t>(blockIdx.x) - d.block0;
	if constexpr(BODY == BB_ROWS) {
		int64_t const g = lb * (kBlock / 64) + (threadIdx.x >> 6);
		if(g < P.groups) rows_group<float, VEC, U, LPRC, false>(d.a, d.b, d.c, P, g);
	} else if constexpr(BODY == BB_SLAB) {
		__shared__ float red[kBlock / 64][64 * VEC];
		slab_item<float, VEC, false>(d.a, d.b, d.c, P, static_cast<uint32_t>(lb), red);
	} else if constexpr(BODY == BB_COL) {
		extern __shared__ unsigned char smem_raw[];
		group_col<VEC, U>(d, P, lb, reinterpret_cast<float*>(smem_raw));
	} else {
		group_dot<VEC, U>(d, P, lb);
	}
}

namespace
{
using GroupKernel = void (*)(const BatchDesc*, const uint32_t*);

// group key: body, vector width, slab units / row sets, lanes per row fixed at compile time
using Key = std::tuple<int, int, int, int>;

template<int VEC, int U>
GroupKernel rows_kernel(int lprc)
{
	// compile-time lanes per row for full-width vectors, as the single-problem k_rows; -1: tight rows
	if(lprc < 0) return k_batch_group<BB_ROWS, VEC, U, -1>;
	if constexpr(VEC == 4) {
		switch(lprc) {
			case 4:  return k_batch_group<BB_ROWS, VEC, U, 4>;
			case 8:  return k_batch_group<BB_ROWS, VEC, U, 8>;
			case 16: return k_batch_group<BB_ROWS, VEC, U, 16>;
			case 32: return k_batch_group<BB_ROWS, VEC, U, 32>;
			case 64: return k_batch_group<BB_ROWS, VEC, U, 64>;
			default: return k_batch_group<BB_ROWS, VEC, U, 0>;
		}
	} else {
		return k_batch_group<BB_ROWS, VEC, U, 0>;
	}
}

template<int BODY, int VEC>
GroupKernel kernel_u(int U, int lprc)
{
	if constexpr(BODY == BB_SLAB) {
		return U == 1 ? k_batch_group<BB_SLAB, VEC, 1, 0> : nullptr;
	} else if constexpr(BODY == BB_COL) {
		switch(U) {
			case 1: return k_batch_group<BB_COL, VEC, 1, 0>;
			case 2: return k_batch_group<BB_COL, VEC, 2, 0>;
			case 4: return k_batch_group<BB_COL, VEC, 4, 0>;
			default: return nullptr;
		}
	} else if constexpr(BODY == BB_DOT) {
		switch(U) {
			case 1: return k_batch_group<BB_DOT, VEC, 1, 0>;
			case 2: return k_batch_group<BB_DOT, VEC, 2, 0>;
			case 4: return k_batch_group<BB_DOT, VEC, 4, 0>;
			case 8: return k_batch_group<BB_DOT, VEC, 8, 0>;
			default: return nullptr;
		}
	} else {
		switch(U) {
			case 1: return rows_kernel<VEC, 1>(lprc);
			case 2: return rows_kernel<VEC, 2>(lprc);
			case 4: return rows_kernel<VEC, 4>(lprc);
			case 8: return rows_kernel<VEC, 8>(lprc);
			default: return nullptr;
		}
	}
}

template<int BODY>
GroupKernel kernel_vec(int vec, int U, int lprc)
{
	switch(vec) {
		case 1: return kernel_u<BODY, 1>(U, lprc);
		case 2: return kernel_u<BODY, 2>(U, lprc);
		case 4: return kernel_u<BODY, 4>(U, lprc);
		default: return nullptr;
	}
}

GroupKernel group_kernel(Key const& k)
{
	int const body = std::get<0>(k), vec = std::get<1>(k), U = std::get<2>(k), lprc = std::get<3>(k);
	switch(body) {
		case BB_COL:  return kernel_vec<BB_COL>(vec, U, lprc);
		case BB_SLAB: return kernel_vec<BB_SLAB>(vec, U, lprc);
		case BB_ROWS: return kernel_vec<BB_ROWS>(vec, U, lprc);
		default:      return kernel_vec<BB_DOT>(vec, U, lprc);
	}
}
} // namespace

struct BatchLaunch {
	GroupKernel kernel = nullptr;
	int64_t blocks = 0;
	std::size_t shmem = 0;
	std::size_t desc_off = 0;    // byte offsets into the batch's device memory
	std::size_t owner_off = 0;
};

struct Batch {
	int device = -1;
	int nprob = 0;
	int64_t total_blocks = 0;
	int64_t bytes = 0;
	void* dmem = nullptr;              // descriptors and owner tables of every launch
	std::vector<BatchLaunch> launches;
	hipStream_t capture_stream = nullptr;
	hipGraphExec_t graph_exec = nullptr;
	hipGraph_t graph = nullptr;
	std::vector<int> variant;          // caller's problem index -> variant of its tiles
};

namespace
{
int enqueue(Batch& bt, hipStream_t s)
{
	auto* base = static_cast<char*>(bt.dmem);
	for(BatchLaunch const& L : bt.launches) {
		hipLaunchKernelGGL(L.kernel, dim3(static_cast<unsigned>(L.blocks)), dim3(kBlock), L.shmem, s,
		                   reinterpret_cast<const BatchDesc*>(base + L.desc_off),
		                   reinterpret_cast<const uint32_t*>(base + L.owner_off));
		hipError_t e = hipGetLastError();
		if(e != hipSuccess) return hip_fail(e, "batch launch");
	}
	return TTVX_OK;
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
	if(count > (size_t{1} << 31)) return fail(TTVX_ERR_ARGUMENT, "ttvx_batch_create: too many problems");
	DeviceInfo di;
	if(!current_device(di)) return TTVX_ERR_DEVICE;
	DType const dt = dtype_of(TTVX_F32);

	struct Item {
		BatchDesc d;
		Key key;
		int64_t blocks;
		std::size_t shmem;
		int64_t cost;
	};
	std::vector<Item> items;
	items.reserve(count);
	std::vector<int> variant(count, -1);
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
		// the problem's own variant, or the one a single call would take; the contraction is never split
		if(pr.variant >= ttvx_num_variants())
			return fail(TTVX_ERR_ARGUMENT, where + "no kernel variant " + std::to_string(pr.variant));
		int v = pr.variant >= 0 ? pr.variant : select_variant(f, dt, aaddr, 0u);
		if(v == V_COL_SPLIT) v = V_COL_U1;
		if(v == V_DOT_SPLIT) v = V_DOT_U2;
		Config cfg;
		if(v == V_STRIDED || !configure(v, f, dt, aaddr, di.cus, 0u, cfg) || cfg.ksplit != 1 || cfg.args.part64)
			return fail(TTVX_ERR_UNSUPPORTED, where + (pr.variant >= 0 ? std::string("variant ") + ttvx_variant_name(pr.variant) +
			                                                            " cannot run this problem in a batch"
			                                                          : std::string("the batched kernel needs A and C to fold "
			                                                                        "to three dense-inner dimensions")));
		Item it{};
		int body, lprc = 0;
		if(cfg.lean == LEAN_SLAB) {
			body = BB_SLAB;
			it.blocks = cfg.args.items;
		} else if(cfg.kind == KIND_COL) {
			body = BB_COL;
			it.d.per = std::max<int64_t>(1, (cfg.args.items + cfg.grid - 1) / cfg.grid);
			it.blocks = (cfg.args.items + it.d.per - 1) / it.d.per;
		} else if(cfg.lean == LEAN_ROWS) {
			body = BB_ROWS;
			it.blocks = (cfg.args.groups + kBlock / 64 - 1) / (kBlock / 64);
			if(cfg.args.tight) lprc = -1;
			else if(cfg.vec == 4 && cfg.args.LPR >= 4) lprc = cfg.args.LPR;
		} else {
			body = BB_DOT;
			int64_t const waves = kBlock / 64;
			int64_t const per = std::max<int64_t>(1, (cfg.args.groups + cfg.grid - 1) / cfg.grid);
			it.d.per = (per + waves - 1) / waves * waves;
			it.blocks = (cfg.args.groups + it.d.per - 1) / it.d.per;
		}
		it.key = Key(body, cfg.vec, cfg.U, lprc);
		if(group_kernel(it.key) == nullptr || cfg.shmem > 64 * 1024 || it.blocks > kMaxLeanBlocks)
			return fail(TTVX_ERR_UNSUPPORTED, where + "tile does not fit the batched kernels");
		it.shmem = body == BB_COL ? cfg.shmem : 0;
		set_output_stores(cfg, f, dt, reinterpret_cast<uintptr_t>(pr.c));
		it.d.args = cfg.args;
		it.d.args.kchunk  = f.K;
		it.d.args.partial = nullptr;
		it.d.args.part64  = 0;
		it.d.args.scales  = nullptr;
		it.d.args.fps     = 1;
		it.d.args.tstore  = 0;
		it.d.a = pr.a;
		it.d.b = pr.b;
		it.d.c = pr.c;
		it.cost = f.outputs * f.K;
		variant[n] = v;
		items.push_back(it);
		bytes += f.outputs * f.K * 4 + f.outputs * 4 + f.K * 4;
	}

	// groups: problems of one tile body, largest first inside a group (the small ones fill the
	// drain), groups launched largest first; a group whose grid would pass the launch limit
	// continues in another launch
	std::map<Key, std::vector<int>> groups;
	for(size_t n = 0; n < items.size(); ++n) groups[items[n].key].push_back(static_cast<int>(n));
	struct Plan {
		Key key;
		std::vector<int> members;
		int64_t blocks = 0;
		int64_t cost = 0;
		std::size_t shmem = 0;
	};
	std::vector<Plan> plans;
	for(auto& g : groups) {
		std::stable_sort(g.second.begin(), g.second.end(),
		                 [&](int x, int y) { return items[x].cost > items[y].cost; });
		Plan cur;
		cur.key = g.first;
		for(int m : g.second) {
			if(cur.blocks + items[m].blocks > kMaxLeanBlocks) {
				plans.push_back(cur);
				cur = Plan{};
				cur.key = g.first;
			}
			items[m].d.block0 = cur.blocks;
			cur.members.push_back(m);
			cur.blocks += items[m].blocks;
			cur.cost += items[m].cost;
			cur.shmem = std::max(cur.shmem, items[m].shmem);
		}
		plans.push_back(cur);
	}
	std::stable_sort(plans.begin(), plans.end(), [](Plan const& x, Plan const& y) { return x.cost > y.cost; });

	// device memory: per launch, its descriptors then its owner table (each 256-byte aligned)
	auto* bt = new Batch();
	bt->device = di.id;
	bt->nprob = static_cast<int>(items.size());
	bt->bytes = bytes;
	bt->variant = variant;
	std::size_t off = 0;
	for(Plan const& p : plans) {
		BatchLaunch L;
		L.kernel = group_kernel(p.key);
		L.blocks = p.blocks;
		L.shmem = p.shmem;
		L.desc_off = off;
		off += (p.members.size() * sizeof(BatchDesc) + 255) / 256 * 256;
		L.owner_off = off;
		off += (static_cast<std::size_t>(p.blocks) * sizeof(uint32_t) + 255) / 256 * 256;
		bt->launches.push_back(L);
		bt->total_blocks += p.blocks;
	}
	std::vector<unsigned char> host(off);
	for(size_t l = 0; l < plans.size(); ++l) {
		auto* descs = reinterpret_cast<BatchDesc*>(host.data() + bt->launches[l].desc_off);
		auto* owner = reinterpret_cast<uint32_t*>(host.data() + bt->launches[l].owner_off);
		for(size_t j = 0; j < plans[l].members.size(); ++j) {
			Item const& it = items[plans[l].members[j]];
			std::memcpy(&descs[j], &it.d, sizeof(BatchDesc));
			std::fill(owner + it.d.block0, owner + it.d.block0 + it.blocks, static_cast<uint32_t>(j));
		}
	}
	hipError_t e = injected(TTVX_FAULT_ALLOC) ? hipErrorOutOfMemory : hipMalloc(&bt->dmem, std::max<std::size_t>(off, 256));
	if(e != hipSuccess) {
		bt->dmem = nullptr;
		delete bt;
		return hip_fail(e, "hipMalloc (batch descriptors)");
	}
	e = hipMemcpy(bt->dmem, host.data(), off, hipMemcpyHostToDevice);
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
	int dev = -1;
	hipError_t e = hipGetDevice(&dev);
	if(e != hipSuccess) return hip_fail(e, "hipGetDevice");
	if(dev != bt->device) return fail(TTVX_ERR_ARGUMENT, "ttvx_batch_launch: the batch was created on another device");
	hipStream_t s = static_cast<hipStream_t>(stream);
	if(!use_graph) return enqueue(*bt, s);
	if(bt->graph_exec == nullptr) {
		// capture once on a private stream (the legacy default stream cannot be captured); a failed
		// earlier capture leaves the stream for this attempt and its graph is dropped
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
	e = hipGraphLaunch(bt->graph_exec, s);
	return e == hipSuccess ? TTVX_OK : hip_fail(e, "hipGraphLaunch");
}

int ttvx_batch_info(void* handle, int64_t* out4)
{
	clear_error();
	auto* bt = static_cast<Batch*>(handle);
	if(bt == nullptr || out4 == nullptr) return fail(TTVX_ERR_ARGUMENT, "ttvx_batch_info: null argument");
	out4[0] = bt->total_blocks;
	out4[1] = bt->bytes;
	out4[2] = static_cast<int64_t>(bt->launches.size());
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
