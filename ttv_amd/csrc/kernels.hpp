// gfx950 kernels of the mode-q tensor-times-vector product.
//
// Every problem arrives folded to  C[o,i] = sum_k A[o*sAo + k*sAk + i] * b[k]  (plan.hpp).
// The launch grid is the fused loop nest of the reference's Loops-over-GEMM formulation
// (tensor_times_vector.h:1207-1256 and :705-772 fuse the free modes into one OpenMP loop and
// call one BLAS gemv per slice): here one work item is a tile of the [O][I] output, and the
// contraction over k runs inside the tile.  There is no transposed copy of A.
//
//   k_dot  (I == 1, q is the fastest mode; reference gemv_row / dot, matrix_times_vector.h:51-91,
//           231-295): LPR lanes share a row of K contiguous elements, each lane loads 16-byte
//           vectors, rows are summed across lanes by recursive halving.  U rows per lane in
//           flight; rows of one vector per lane are double-buffered across row groups.
//   k_col  (I > 1; reference gemv_col / gemv_col_blas, matrix_times_vector.h:108-221): IT lanes
//           along i (16-byte vectors, coalesced), KS lane groups along k, OS output slabs per
//           item; partial sums over k meet by recursive halving (one wave) or in LDS (several
//           waves).  Outputs leave with non-temporal stores where one wave instruction writes
//           whole 128-byte lines (BF16/FP8 lane vectors are first transposed through LDS).
//   k_strided  general strides that do not fold to three dimensions: one thread per output.
//   k_finalize sums the split-n_q partials (FP32 or FP64) and applies FP8 scales.
//
// Storage types: float, double, bfloat16 (uint16_t) and OCP FP8-e4m3 (uint8_t).  Accumulation
// is FP64 for double storage and FP32 otherwise, with optional Kahan compensation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace ttvx
{

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef float    f32x2 __attribute__((ext_vector_type(2)));

// k unroll of the column tiles (independent loads in flight per lane and slab unit) for 4/8-byte
// and for 1/2-byte elements
#ifndef TTVX_KL_WIDE
#define TTVX_KL_WIDE 8
#endif
#ifndef TTVX_KL_NARROW
#define TTVX_KL_NARROW 8
#endif
constexpr int kBlock = 256;   // threads per block for every kernel below (4 waves of 64)

template<typename S> struct Traits;
template<> struct Traits<float>    { using acc = float;  static constexpr int vmax = 4;  };
template<> struct Traits<double>   { using acc = double; static constexpr int vmax = 2;  };
template<> struct Traits<uint16_t> { using acc = float;  static constexpr int vmax = 8;  };
template<> struct Traits<uint8_t>  { using acc = float;  static constexpr int vmax = 16; };

// BF16 and FP8 storage whose vectors may start off a 4-byte boundary (odd row lengths, odd base
// offsets): the same bytes, read by realigned loads (below).  Base<S> is the storage format a type
// decodes as.
struct bf16_sub { uint16_t bits; };
struct fp8_sub  { uint8_t bits; };
template<> struct Traits<bf16_sub> { using acc = float; static constexpr int vmax = 8;  };
template<> struct Traits<fp8_sub>  { using acc = float; static constexpr int vmax = 16; };
template<typename S> struct Base { using type = S; };
template<> struct Base<bf16_sub> { using type = uint16_t; };
template<> struct Base<fp8_sub>  { using type = uint8_t; };
template<typename S> constexpr bool kRealigned = !std::is_same_v<typename Base<S>::type, S>;

// ---- loads ---------------------------------------------------------------------------------
// Loads VEC consecutive elements (VEC*sizeof(S) bytes, naturally aligned) and converts them to
// the accumulation type.

// A is streamed exactly once, so its loads are non-temporal (no reuse to keep in the caches);
// measured on MI355X this lifts a pure 1 GiB read stream from ~6.0 to ~7.0 TB/s.
// A is device memory: the loads go through the global (not the flat) path, so they only count
// against vmcnt.
#define TTVX_GLOBAL(T, p) ((const __attribute__((address_space(1))) T*)(p))
#define TTVX_GLOBAL_W(T, p) ((__attribute__((address_space(1))) T*)(p))
// Vectors of A need only the element's alignment: global memory on gfx950 takes vector loads at
// any byte alignment (rows whose length is not a multiple of the vector width keep 16-byte loads),
// so the load types below promise no more than one byte of alignment.
typedef uint32_t u32x4_b __attribute__((ext_vector_type(4), aligned(1)));
typedef uint32_t u32x2_b __attribute__((ext_vector_type(2), aligned(1)));
typedef uint32_t u32_b __attribute__((aligned(1)));
template<int BYTES> struct RawLoad;
template<> struct RawLoad<16> {
	static __device__ __forceinline__ void get(const void* p, uint32_t (&w)[4]) {
		u32x4_b v = __builtin_nontemporal_load(TTVX_GLOBAL(u32x4_b, p));
		w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
	}
};
template<> struct RawLoad<8> {
	static __device__ __forceinline__ void get(const void* p, uint32_t (&w)[2]) {
		u32x2_b v = __builtin_nontemporal_load(TTVX_GLOBAL(u32x2_b, p));
		w[0] = v.x; w[1] = v.y;
	}
};
template<> struct RawLoad<4> {
	static __device__ __forceinline__ void get(const void* p, uint32_t (&w)[1]) {
		w[0] = __builtin_nontemporal_load(TTVX_GLOBAL(u32_b, p));
	}
};
typedef uint16_t u16_b __attribute__((aligned(1)));
template<> struct RawLoad<2> {
	static __device__ __forceinline__ void get(const void* p, uint32_t (&w)[1]) {
		w[0] = __builtin_nontemporal_load(TTVX_GLOBAL(u16_b, p));
	}
};
template<> struct RawLoad<1> {
	static __device__ __forceinline__ void get(const void* p, uint32_t (&w)[1]) {
		w[0] = __builtin_nontemporal_load(TTVX_GLOBAL(uint8_t, p));
	}
};

// Vector loads of 1-byte-aligned BF16/FP8 data cost up to twice as much as 4-byte-aligned ones
// (MI355X: BF16 rows of 129 elements 3.7 TB/s against 6.5 TB/s for 136).  A realigned load reads the
// BYTES/4 dwords from the 4-byte boundary at or below p plus the dword after them, and funnel-shifts
// each pair (v_alignbyte) by p's offset.  The extra dword lies in the aligned dword that holds the
// vector's last byte (or, for an aligned p, re-reads the vector's own last dword), so no byte outside
// the dwords the vector touches is read.
typedef uint32_t u32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));
typedef uint32_t u32x2_a4 __attribute__((ext_vector_type(2), aligned(4)));
template<int BYTES>
__device__ __forceinline__ void realigned_load(const void* p, uint32_t (&w)[BYTES / 4])
{
	constexpr int W = BYTES / 4;
	uintptr_t const a = reinterpret_cast<uintptr_t>(p);
	uint32_t const m = static_cast<uint32_t>(a & 3u);
	uintptr_t const base = a - m;
	uint32_t d[W + 1];
	if constexpr(W == 4) {
		u32x4_a4 v = __builtin_nontemporal_load(TTVX_GLOBAL(u32x4_a4, base));
		d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
	} else if constexpr(W == 2) {
		u32x2_a4 v = __builtin_nontemporal_load(TTVX_GLOBAL(u32x2_a4, base));
		d[0] = v.x; d[1] = v.y;
	} else {
		d[0] = __builtin_nontemporal_load(TTVX_GLOBAL(uint32_t, base));
	}
	d[W] = __builtin_nontemporal_load(TTVX_GLOBAL(uint32_t, base + (m != 0u ? BYTES : BYTES - 4)));
#pragma unroll
	for(int i = 0; i < W; ++i) w[i] = __builtin_amdgcn_alignbyte(d[i + 1], d[i], m);
}

template<typename S, int VEC> struct Vec {
	static constexpr int bytes = VEC * static_cast<int>(sizeof(S));
	static constexpr int words = bytes >= 4 ? bytes / 4 : 1;
	uint32_t w[words];
	__device__ __forceinline__ void load(const S* p)
	{
		if constexpr(kRealigned<S> && bytes >= 4) realigned_load<bytes>(p, w);
		else RawLoad<bytes>::get(p, w);
	}
};

// Column loads of realigned storage without a second load per vector.  The lanes of an i-group hold
// consecutive vectors of one slab row, so the dword after a lane's own W dwords (read from the
// 4-byte boundary at or below its vector) is the first dword of the next lane: it comes over with one
// ds_bpermute.  Lanes whose next lane holds no adjacent vector (the last lane of an i-group or of the
// wave, lanes at the clamped end of a row) load that dword themselves, as realigned_load does.
typedef uint32_t u32x2_a4x __attribute__((ext_vector_type(2), aligned(4)));
template<typename S, int VEC> struct XVec {
	static constexpr int W = Vec<S, VEC>::words;
	uint32_t d[W];
	uint32_t ext;
	uint32_t m;
	__device__ __forceinline__ void load(const S* p, bool edge)
	{
		uintptr_t const a = reinterpret_cast<uintptr_t>(p);
		m = static_cast<uint32_t>(a & 3u);
		uintptr_t const base = a - m;
		if constexpr(W == 4) {
			u32x4_a4 v = __builtin_nontemporal_load(TTVX_GLOBAL(u32x4_a4, base));
			d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
		} else if constexpr(W == 2) {
			u32x2_a4x v = __builtin_nontemporal_load(TTVX_GLOBAL(u32x2_a4x, base));
			d[0] = v.x; d[1] = v.y;
		} else {
			d[0] = __builtin_nontemporal_load(TTVX_GLOBAL(uint32_t, base));
		}
		ext = 0u;
		if(edge) ext = __builtin_nontemporal_load(TTVX_GLOBAL(uint32_t, base + (m != 0u ? 4 * W : 4 * W - 4)));
	}
	// every lane of the i-group executes this together (the exchange reads the next lane's d[0])
	__device__ __forceinline__ void finish(bool edge, int lane, Vec<S, VEC>& v) const
	{
		uint32_t const nxt = static_cast<uint32_t>(
		    __builtin_amdgcn_ds_bpermute(((lane + 1) & 63) << 2, static_cast<int>(d[0])));
		uint32_t const next = edge ? ext : nxt;
#pragma unroll
		for(int i = 0; i + 1 < W; ++i) v.w[i] = __builtin_amdgcn_alignbyte(d[i + 1], d[i], m);
		v.w[W - 1] = __builtin_amdgcn_alignbyte(next, d[W - 1], m);
	}
};

// decode element e of a loaded vector
template<typename S, int VEC>
__device__ __forceinline__ void decode(const Vec<S, VEC>& v, typename Traits<S>::acc (&x)[VEC]);

template<int VEC>
__device__ __forceinline__ void decode(const Vec<float, VEC>& v, float (&x)[VEC])
{
#pragma unroll
	for(int e = 0; e < VEC; ++e) x[e] = __uint_as_float(v.w[e]);
}

template<int VEC>
__device__ __forceinline__ void decode(const Vec<double, VEC>& v, double (&x)[VEC])
{
#pragma unroll
	for(int e = 0; e < VEC; ++e)
		x[e] = __hiloint2double(static_cast<int>(v.w[2 * e + 1]), static_cast<int>(v.w[2 * e]));
}

template<int VEC>
__device__ __forceinline__ void decode(const Vec<uint16_t, VEC>& v, float (&x)[VEC])
{
	if constexpr(VEC == 1) {
		x[0] = __uint_as_float(v.w[0] << 16);
	} else {
#pragma unroll
		for(int e = 0; e < VEC / 2; ++e) {
			x[2 * e]     = __uint_as_float(v.w[e] << 16);
			x[2 * e + 1] = __uint_as_float(v.w[e] & 0xffff0000u);
		}
	}
}

// OCP e4m3 -> float with the gfx950 packed converter (two codes per instruction).
template<int VEC>
__device__ __forceinline__ void decode(const Vec<uint8_t, VEC>& v, float (&x)[VEC])
{
	if constexpr(VEC == 1) {
		f32x2 f = __builtin_amdgcn_cvt_pk_f32_fp8(static_cast<int>(v.w[0]), false);
		x[0] = f.x;
	} else if constexpr(VEC == 2) {
		f32x2 f = __builtin_amdgcn_cvt_pk_f32_fp8(static_cast<int>(v.w[0]), false);
		x[0] = f.x; x[1] = f.y;
	} else {
#pragma unroll
		for(int e = 0; e < VEC / 4; ++e) {
			f32x2 lo = __builtin_amdgcn_cvt_pk_f32_fp8(static_cast<int>(v.w[e]), false);
			f32x2 hi = __builtin_amdgcn_cvt_pk_f32_fp8(static_cast<int>(v.w[e]), true);
			x[4 * e] = lo.x; x[4 * e + 1] = lo.y; x[4 * e + 2] = hi.x; x[4 * e + 3] = hi.y;
		}
	}
}

// realigned storage decodes as its base format (same words)
template<typename S, int VEC>
__device__ __forceinline__ void decode_as_base(const Vec<S, VEC>& v, float (&x)[VEC])
{
	Vec<typename Base<S>::type, VEC> u;
#pragma unroll
	for(int i = 0; i < Vec<S, VEC>::words; ++i) u.w[i] = v.w[i];
	decode(u, x);
}
template<int VEC>
__device__ __forceinline__ void decode(const Vec<bf16_sub, VEC>& v, float (&x)[VEC]) { decode_as_base(v, x); }
template<int VEC>
__device__ __forceinline__ void decode(const Vec<fp8_sub, VEC>& v, float (&x)[VEC]) { decode_as_base(v, x); }

// ---- accumulation ----------------------------------------------------------------------------
template<typename Acc, bool KAHAN> struct Sum;
template<typename Acc> struct Sum<Acc, false> {
	Acc s = Acc(0);
	__device__ __forceinline__ void add(Acc x) { s += x; }
	// one fused multiply-add in the accumulation type (the float form must not promote to double)
	__device__ __forceinline__ void fma(Acc a, Acc b)
	{
		if constexpr(std::is_same_v<Acc, float>) s = __builtin_fmaf(a, b, s);
		else s = __builtin_fma(a, b, s);
	}
	__device__ __forceinline__ Acc value() const { return s; }
};
template<typename Acc> struct Sum<Acc, true> {
	Acc s = Acc(0), c = Acc(0);     // running sum and its compensation (Kahan 1965)
	__device__ __forceinline__ void add(Acc x) {
		Acc y = x - c;
		Acc t = s + y;
		c = (t - s) - y;
		s = t;
	}
	__device__ __forceinline__ void fma(Acc a, Acc b) { add(a * b); }
	__device__ __forceinline__ Acc value() const { return s - c; }
};

// ---- launch arguments ------------------------------------------------------------------------
struct Args {
	int64_t O, K, I;          // folded extents
	int64_t sAo, sAk;         // A strides (elements); the i stride is 1 for k_dot/k_col
	int64_t sCo;              // C stride of o (the i stride is 1)
	// k_col tiling
	int IT, KS, OS;           // lanes along i (vectors), lane groups along k, slabs per item
	int red;                  // cross-lane reduction: 0 none, 1 shuffles in one wave, 2 LDS
	int along_i;              // the U units of a thread lie along i (wide rows) instead of along o
	int64_t iTiles;           // tiles along i (IT*VEC columns, or U*IT*VEC when along_i)
	int64_t items;            // ceil(O / (OS*U)) * iTiles, or ceil(O / OS) * iTiles when along_i
	// k_dot tiling
	int LPR;                  // lanes per row (a power of two <= 64; any 2..64 for tight k_rows)
	int lprShift;             // log2(LPR) (power-of-two LPR)
	int64_t groups;           // ceil(O / (rows per wave * U))
	int kiter;                // vector steps per lane per row (for b caching)
	// split of the contraction over gridDim.y
	int64_t kchunk;           // contraction elements per split (multiple of the lane k step)
	void*   partial;          // [gridDim.y][O*I] partial sums, or null when gridDim.y == 1
	int     part64;           // partial sums stored as double
	// FP8 scales (null for other types)
	const float* scales;
	int64_t fps;              // fibers per scale
	// work assignment: 0 = grid-stride over items/groups, 1 = each block one contiguous range
	int contig;
	// column items with one k step per lane and no LDS reduction: double-buffered item loop
	int one_step;
	// bytes per vector store of C allowed by its alignment (0: element stores only)
	int cvec;
	// column outputs without reduction whose lane vectors exceed 16 bytes (BF16, FP8), a whole wave
	// along i and 16-byte aligned C: stored wave-transposed (col_store_transposed)
	int tstore;
	// k_slab: log2(IT) and log2(IT*KS)
	int itShift, spanShift;
	// rows (I == 1) whose length the vector width does not divide: their last vector is clamped
	int kpart;
	// k_rows with LPR = ceil(K / VEC) lanes per row, not a power of two (rows_group, LPRC < 0)
	int tight;
};

// Range of work units [begin, end) with step of block blockIdx.x (lane offset added by callers)
__device__ __forceinline__ void block_range(int contig, int64_t n, int64_t& begin, int64_t& end, int64_t& step)
{
	if(contig) {
		int64_t const per = (n + gridDim.x - 1) / gridDim.x;
		begin = static_cast<int64_t>(blockIdx.x) * per;
		end   = begin + per < n ? begin + per : n;
		step  = 1;
	} else {
		begin = blockIdx.x;
		end   = n;
		step  = gridDim.x;
	}
}

// SC: the storage type carries per-group scales (FP8); for the other types the scale code is not
// generated at all.
template<typename S> constexpr bool kScaled = std::is_same_v<typename Base<S>::type, uint8_t>;

// NT: the final store of C is non-temporal.  Chosen per store site at compile time: it pays where one
// wave instruction writes whole 128-byte lines (+4..12 % on outputs of 1/16..1/4 of A) and costs
// where lanes write pieces of lines that other instructions complete (up to 2x slower), because
// non-temporal stores are not combined in L2.  Split partials are re-read and stay cached.
template<typename Acc, bool SC, bool NT = false>
__device__ __forceinline__ void store_out(void* __restrict__ C, const Args& P, int64_t o, int64_t i, Acc v)
{
	int64_t const j = o * P.I + i;
	if(P.partial != nullptr) {
		int64_t const off = static_cast<int64_t>(blockIdx.y) * P.O * P.I + j;
		if(P.part64) static_cast<double*>(P.partial)[off] = static_cast<double>(v);
		else         static_cast<Acc*>(P.partial)[off]    = v;
		return;
	}
	if constexpr(SC)
		if(P.scales != nullptr) v *= static_cast<Acc>(P.scales[j / P.fps]);
	Acc* const dst = static_cast<Acc*>(C) + (o * P.sCo + i);
	if constexpr(NT) __builtin_nontemporal_store(v, TTVX_GLOBAL_W(Acc, dst));
	else             *dst = v;
}

// ---- cross-lane sums ------------------------------------------------------------------------
// A lane holds partial sums v[0..N) of N outputs; the lanes l ^ (m*stride), m < width (stride and
// width powers of two), hold partial sums of the same N outputs.  Recursive halving: at every
// exchange step a lane keeps half of its outputs and adds its partner's half of them, so N outputs
// take N-1 (+ log2(width/N)) shuffles instead of N*log2(width).  On return the lane holds the
// full sums of outputs base .. base+n-1 in v[0..n), n = max(1, N/width); when width > N, the
// width/N lanes that differ only in the low bits hold the same sums, and the one with those bits
// zero is the lane that stores them (the function returns whether this lane stores).
template<typename Acc, int N>
__device__ __forceinline__ bool halving_sum(Acc (&v)[N], int lane, int stride, int width, int& base, int& n)
{
	base = 0;
	n = N;
	int d = width >> 1;
#pragma unroll
	for(int h = N / 2; h >= 1; h >>= 1) {
		if(d >= 1) {
			bool const up = (lane & (d * stride)) != 0;
#pragma unroll
			for(int j = 0; j < h; ++j) {
				Acc const send = up ? v[j] : v[j + h];
				Acc const keep = up ? v[j + h] : v[j];
				v[j] = keep + __shfl_xor(send, d * stride, 64);
			}
			if(up) base += h;
			n = h;
			d >>= 1;
		}
	}
	bool const store = d < 1 || ((lane / stride) & ((d << 1) - 1)) == 0;
	for(; d >= 1; d >>= 1) v[0] += __shfl_xor(v[0], d * stride, 64);
	return store;
}

// x from lane (lane ^ D): D < 32 through ds_swizzle in bit-mask mode (the lane index is an
// immediate, no address arithmetic), D = 32 through __shfl_xor
template<int D, typename Acc>
__device__ __forceinline__ Acc xor_lane(Acc x)
{
	if constexpr(D < 32) {
		constexpr int kPattern = 0x1f | (D << 10);     // and-mask 31, or-mask 0, xor-mask D
		if constexpr(sizeof(Acc) == 4) {
			return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(x), kPattern));
		} else {
			int const lo = __builtin_amdgcn_ds_swizzle(__double2loint(x), kPattern);
			int const hi = __builtin_amdgcn_ds_swizzle(__double2hiint(x), kPattern);
			return __hiloint2double(hi, lo);
		}
	} else {
		return __shfl_xor(x, D, 64);
	}
}

// halving_sum (stride 1) for a width W known at compile time: constant lane offsets
template<typename Acc, int N, int W, int D = W / 2, int H = N / 2>
__device__ __forceinline__ void halving_steps(Acc (&v)[N], int lane, int& base, int& n)
{
	if constexpr(H >= 1 && D >= 1) {
		bool const up = (lane & D) != 0;
#pragma unroll
		for(int j = 0; j < H; ++j) {
			Acc const send = up ? v[j] : v[j + H];
			Acc const keep = up ? v[j + H] : v[j];
			v[j] = keep + xor_lane<D>(send);
		}
		if(up) base += H;
		n = H;
		halving_steps<Acc, N, W, D / 2, H / 2>(v, lane, base, n);
	} else if constexpr(D >= 1) {
		v[0] += xor_lane<D>(v[0]);
		halving_steps<Acc, N, W, D / 2, 0>(v, lane, base, n);
	}
}

template<typename Acc, int N, int W>
__device__ __forceinline__ bool halving_sum_ct(Acc (&v)[N], int lane, int& base, int& n)
{
	base = 0;
	n = N;
	halving_steps<Acc, N, W>(v, lane, base, n);
	// the lanes that differ only in the bits below the last halving step hold the same sums
	constexpr int kRest = W / N > 1 ? W / N : 1;       // lanes that end with identical sums
	return (lane & (kRest - 1)) == 0;
}

// FP8 scales of the VEC consecutive outputs (o, i .. i+VEC-1), one group-index division per vector
template<typename Acc, int VEC>
__device__ __forceinline__ void scale_vec(const Args& P, int64_t o, int64_t i, Acc (&w)[VEC])
{
	if(P.scales == nullptr) return;
	int64_t const j0 = o * P.I + i;
	int64_t const g0 = j0 / P.fps;
	if(P.fps >= VEC) {
		// at most one group boundary inside the vector
		int64_t const rem = j0 - g0 * P.fps;
		Acc const s0 = static_cast<Acc>(P.scales[g0]);
		Acc const s1 = rem + VEC > P.fps ? static_cast<Acc>(P.scales[g0 + 1]) : s0;
#pragma unroll
		for(int e = 0; e < VEC; ++e) w[e] *= rem + e < P.fps ? s0 : s1;
	} else {
#pragma unroll
		for(int e = 0; e < VEC; ++e) w[e] *= static_cast<Acc>(P.scales[(j0 + e) / P.fps]);
	}
}

// VEC consecutive outputs of one lane whose vector starts at column iu and was loaded from column
// ic = col_clamp(iu) (ic < iu for the last, partial vector of a row whose length VEC does not
// divide): vector stores when the vector is whole, C's alignment allows (P.cvec) and no split
// partials are pending; FP8 scales with one group-index division per vector.
// When a lane's vector is exactly one 16-byte store, consecutive lanes write consecutive 16-byte
// pieces and the store is non-temporal; wider vectors (BF16, FP8) take several interleaved stores.
template<typename Acc, int VEC, bool SC>
__device__ __forceinline__ void store_vec(void* __restrict__ C, const Args& P, int64_t o, int64_t iu, int64_t ic,
                                          const Acc (&v)[VEC])
{
	constexpr int BYTES = VEC * static_cast<int>(sizeof(Acc));
	int64_t const i = iu;
	if(ic != iu || P.partial != nullptr || i + VEC > P.I) {
		// element stores; v[e] is column ic + e, and the columns before iu belong to the lane before
#pragma unroll
		for(int e = 0; e < VEC; ++e)
			if(ic + e >= iu && ic + e < P.I) store_out<Acc, SC>(C, P, o, ic + e, v[e]);
		return;
	}
	Acc w[VEC];
#pragma unroll
	for(int e = 0; e < VEC; ++e) w[e] = v[e];
	if constexpr(SC) scale_vec<Acc, VEC>(P, o, i, w);
	Acc* dst = static_cast<Acc*>(C) + o * P.sCo + i;
	if(P.cvec == 0 || (BYTES >= 16 ? P.cvec < 16 : P.cvec < BYTES)) {
		// C rows off the vector grid (a row length that 4 outputs do not divide): element-aligned
		// vector stores (gfx950 takes them at the rate of aligned ones), the scales looked up once
		// per vector as above
		if constexpr(BYTES >= 16) {
#pragma unroll
			for(int c = 0; c < BYTES / 16; ++c) {
				u32x4_a4 q;
				uint32_t const* src = reinterpret_cast<const uint32_t*>(w) + 4 * c;
				q.x = src[0]; q.y = src[1]; q.z = src[2]; q.w = src[3];
				*TTVX_GLOBAL_W(u32x4_a4, dst + c * (16 / static_cast<int>(sizeof(Acc)))) = q;
			}
		} else {
#pragma unroll
			for(int e = 0; e < VEC; ++e) dst[e] = w[e];
		}
		return;
	}
	if constexpr(BYTES == 16) {
		u32x4 q;
		uint32_t const* src = reinterpret_cast<const uint32_t*>(w);
		q.x = src[0]; q.y = src[1]; q.z = src[2]; q.w = src[3];
		__builtin_nontemporal_store(q, TTVX_GLOBAL_W(u32x4, dst));
	} else if constexpr(BYTES > 16) {
#pragma unroll
		for(int c = 0; c < BYTES / 16; ++c) {
			u32x4 q;
			uint32_t const* src = reinterpret_cast<const uint32_t*>(w) + 4 * c;
			q.x = src[0]; q.y = src[1]; q.z = src[2]; q.w = src[3];
			reinterpret_cast<u32x4*>(dst)[c] = q;
		}
	} else if constexpr(BYTES == 8) {
		u32x2 q;
		uint32_t const* src = reinterpret_cast<const uint32_t*>(w);
		q.x = src[0]; q.y = src[1];
		*reinterpret_cast<u32x2*>(dst) = q;
	} else {
#pragma unroll
		for(int e = 0; e < VEC; ++e) dst[e] = w[e];
	}
}

// item -> (slab group, i tile): no division for a single i tile, 32-bit division when it fits
__device__ __forceinline__ void item_split(const Args& P, int64_t item, int64_t& og, int64_t& itile)
{
	if(P.iTiles == 1) {
		og = item;
		itile = 0;
	} else if(P.items <= 0xffffffffll) {
		uint32_t const it32 = static_cast<uint32_t>(item), n32 = static_cast<uint32_t>(P.iTiles);
		uint32_t const q = it32 / n32;
		og = q;
		itile = it32 - q * n32;
	} else {
		og = item / P.iTiles;
		itile = item - og * P.iTiles;
	}
}

// ---- column tiles -------------------------------------------------------------------------
// Block tile: OS*U output slabs x IT*VEC columns (or OS slabs x U*IT*VEC columns when the U
// units lie along i), all of the block's k range.
// Thread t: it = t % IT, ks = (t / IT) % KS, os = t / (IT*KS).  Every thread of the block calls
// col_item for the same item (the LDS reduction synchronises the block).
// Unit u of a thread: slab o_u and first column i_u of its VEC-vector.
// First column a lane's VEC-vector starting at column iu is loaded from: iu itself, or I - VEC for
// the partial last vector of a row (and for lanes past the row, which store nothing); I >= VEC
__device__ __forceinline__ int64_t col_clamp(const Args& P, int64_t iu, int VEC)
{
	return iu + VEC <= P.I ? iu : P.I - VEC;
}

__device__ __forceinline__ void col_unit(const Args& P, int64_t og, int64_t itile, int os, int it, int u, int U, int VEC,
                                         int64_t& o, int64_t& i)
{
	if(P.along_i) {
		o = og * P.OS + os;
		i = ((itile * U + u) * P.IT + it) * VEC;
	} else {
		o = (og * U + u) * P.OS + os;
		i = (itile * P.IT + it) * VEC;
	}
}
template<typename Acc, int VEC, int U, bool SC>
__device__ __forceinline__ void col_store(void* __restrict__ C, const Args& P, int64_t og, int64_t itile, int os,
                                          int it, int ks, int t, bool active, Acc (&v)[U][VEC], Acc* red);

// Instantiations that carry the wave-transposed store (its staging raises the register count): BF16
// lane vectors (32 bytes) with 2 or 4 slab units, FP8 lane vectors (64 bytes) with 4 units.  The
// one-unit kernels and the two-unit FP8 kernel would lose a wave per SIMD, which costs more than
// the stores gain.
__host__ __device__ constexpr bool tstore_compiled(int lane_bytes, int units)
{
	return (lane_bytes == 32 && units >= 2) || (lane_bytes == 64 && units == 4);
}

// Column outputs of one wave whose 64 lanes lie along i (IT a multiple of 64, no reduction) and
// whose lane vectors span G = VEC*sizeof(Acc)/16 > 1 16-byte pieces (BF16: 2, FP8: 4).  Stored
// directly, store instruction c would write piece c of every lane, a quarter or half of each
// 128-byte line, and lines are only completed by later instructions.  Instead the wave stages its
// run of 64*VEC outputs in its own LDS slice (64*VEC values) and store c writes pieces
// c*64 .. c*64+63 of the run, one per lane and contiguous, so each instruction covers whole lines
// and can be non-temporal.  Only the wave's own lanes touch the slice, so wavefront-scope fences
// order the accesses; every lane of the wave executes this (uniform control flow).
template<typename Acc, int VEC, int U, bool SC>
__device__ __forceinline__ void col_store_transposed(void* __restrict__ C, const Args& P, int64_t og, int64_t itile,
                                                     int os, int it, bool active, const Acc (&v)[U][VEC], Acc* lds)
{
	static_assert(sizeof(Acc) == 4, "4-byte outputs");
	constexpr int G = VEC * static_cast<int>(sizeof(Acc)) / 16;
	int const lane = static_cast<int>(threadIdx.x) & 63;
	Acc* const slice = lds + (static_cast<int>(threadIdx.x) >> 6) * (64 * VEC);
#pragma unroll
	for(int u = 0; u < U; ++u) {
		int64_t oo, ii;
		col_unit(P, og, itile, os, it, u, U, VEC, oo, ii);
		bool const row_ok = active && oo < P.O;
		Acc w[VEC];
#pragma unroll
		for(int e = 0; e < VEC; ++e) w[e] = v[u][e];
		if constexpr(SC)
			if(row_ok && ii < P.I) scale_vec<Acc, VEC>(P, oo, ii, w);
		// the previous unit's reads of the slice are done before it is written again
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
#pragma unroll
		for(int g = 0; g < G; ++g) {
			u32x4 q;
			q.x = __float_as_uint(w[4 * g]);     q.y = __float_as_uint(w[4 * g + 1]);
			q.z = __float_as_uint(w[4 * g + 2]); q.w = __float_as_uint(w[4 * g + 3]);
			reinterpret_cast<u32x4*>(slice)[lane * G + g] = q;
		}
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		int64_t const i0 = ii - static_cast<int64_t>(lane) * VEC;      // first column of the wave's run
#pragma unroll
		for(int c = 0; c < G; ++c) {
			int const piece = c * 64 + lane;
			u32x4 const q = reinterpret_cast<const u32x4*>(slice)[piece];
			// the VEC columns of the piece's source lane are inside the row (VEC divides I)
			if(row_ok && i0 + static_cast<int64_t>(piece / G) * VEC < P.I) {
				Acc* const dst = static_cast<Acc*>(C) + oo * P.sCo + i0 + static_cast<int64_t>(piece) * 4;
				__builtin_nontemporal_store(q, TTVX_GLOBAL_W(u32x4, dst));
			}
		}
	}
}

template<typename S, int VEC, int U, bool KAHAN>
__device__ __forceinline__ void col_item(const S* __restrict__ A, const typename Traits<S>::acc* __restrict__ b,
                                         void* __restrict__ C, const Args& P, int64_t item, int64_t kb, int64_t ke,
                                         typename Traits<S>::acc* red)
{
	using Acc = typename Traits<S>::acc;
	// k unroll: U*KU independent 16-byte loads in flight per lane (TTVX_KL_WIDE for 4- and 8-byte
	// elements, TTVX_KL_NARROW for the 1- and 2-byte types)
	constexpr int KL = sizeof(S) >= 4 ? TTVX_KL_WIDE : TTVX_KL_NARROW;
	constexpr int KU = KL / U > 1 ? KL / U : 1;
	int const t  = static_cast<int>(threadIdx.x);
	int const it = t % P.IT;
	int const ks = (t / P.IT) % P.KS;
	int const os = t / (P.IT * P.KS);
	bool const active = os < P.OS;
	int64_t const kstep = static_cast<int64_t>(P.KS) * P.sAk;

	int64_t og, itile;
	item_split(P, item, og, itile);
	int64_t o[U], iu[U];
	const S* pa[U];
#pragma unroll
	for(int u = 0; u < U; ++u) {
		col_unit(P, og, itile, os, it, u, U, VEC, o[u], iu[u]);
		// clamped: tail lanes re-read live data and store nothing
		int64_t const oc = o[u] < P.O ? o[u] : P.O - 1;
		int64_t const ic = col_clamp(P, iu[u], VEC);
		pa[u] = A + oc * P.sAo + (kb + ks) * P.sAk + ic;
	}
	Sum<Acc, KAHAN> acc[U][VEC];
	// realigned storage: lanes whose next lane does not hold the following vector
	bool edge[U];
#pragma unroll
	for(int u = 0; u < U; ++u) edge[u] = it == P.IT - 1 || (t & 63) == 63 || iu[u] + 2 * VEC > P.I;

	if(active) {
		int64_t k = kb + ks;
		for(; k + (KU - 1) * P.KS < ke; k += KU * P.KS) {
			Acc bk[KU];
			Vec<S, VEC> raw[KU][U];
			XVec<S, VEC> xr[KU][U];
#pragma unroll
			for(int j = 0; j < KU; ++j) bk[j] = b[k + j * P.KS];
#pragma unroll
			for(int j = 0; j < KU; ++j)
#pragma unroll
				for(int u = 0; u < U; ++u) {
					if constexpr(kRealigned<S>) xr[j][u].load(pa[u] + j * kstep, edge[u]);
					else raw[j][u].load(pa[u] + j * kstep);
				}
			// every load of the step is issued before any of them is used: without this fence the
			// scheduler may trade loads in flight for registers and serialise them
			__builtin_amdgcn_sched_barrier(0);
#pragma unroll
			for(int j = 0; j < KU; ++j)
#pragma unroll
				for(int u = 0; u < U; ++u) {
					Acc x[VEC];
					if constexpr(kRealigned<S>) xr[j][u].finish(edge[u], t & 63, raw[j][u]);
					decode(raw[j][u], x);
#pragma unroll
					for(int e = 0; e < VEC; ++e) acc[u][e].fma(x[e], bk[j]);
				}
#pragma unroll
			for(int u = 0; u < U; ++u) pa[u] += KU * kstep;
		}
		for(; k < ke; k += P.KS) {
			Acc const bk = b[k];
#pragma unroll
			for(int u = 0; u < U; ++u) {
				Vec<S, VEC> raw;
				if constexpr(kRealigned<S>) {
					XVec<S, VEC> xr;
					xr.load(pa[u], edge[u]);
					xr.finish(edge[u], t & 63, raw);
				} else {
					raw.load(pa[u]);
				}
				Acc x[VEC];
				decode(raw, x);
#pragma unroll
				for(int e = 0; e < VEC; ++e) acc[u][e].fma(x[e], bk);
				pa[u] += kstep;
			}
		}
	}

	Acc v[U][VEC];
#pragma unroll
	for(int u = 0; u < U; ++u)
#pragma unroll
		for(int e = 0; e < VEC; ++e) v[u][e] = acc[u][e].value();
	col_store<Acc, VEC, U, kScaled<S>>(C, P, og, itile, os, it, ks, t, active, v, red);
}

// Sums the lanes' partial sums of one column item over its KS lane groups and stores the item's
// outputs (shuffles inside a wave, or LDS across waves).
template<typename Acc, int VEC, int U, bool SC>
__device__ __forceinline__ void col_store(void* __restrict__ C, const Args& P, int64_t og, int64_t itile, int os,
                                          int it, int ks, int t, bool active, Acc (&v)[U][VEC], Acc* red)
{
	if(P.red == 1) {
		// lanes it + IT*ks of one wave hold partial sums of the same U*VEC outputs
		Acc w[U * VEC];
#pragma unroll
		for(int u = 0; u < U; ++u)
#pragma unroll
			for(int e = 0; e < VEC; ++e) w[u * VEC + e] = v[u][e];
		int base, n;
		bool const st = halving_sum<Acc, U * VEC>(w, t & 63, P.IT, P.KS, base, n);
		if(active && st) {
			if constexpr(VEC * sizeof(Acc) == 16) {
				// the lane holds whole 16-byte lane vectors (KS <= U): one vector store each, so
				// the lanes of one instruction write runs of slab rows (non-temporal, whole lines)
				if(n >= VEC) {
#pragma unroll
					for(int g = 0; g < U; ++g)
						if(g * VEC < n) {
							int64_t oo, ii;
							col_unit(P, og, itile, os, it, base / VEC + g, U, VEC, oo, ii);
							Acc x[VEC];
#pragma unroll
							for(int e = 0; e < VEC; ++e) x[e] = w[g * VEC + e];
							if(oo < P.O && ii < P.I) store_vec<Acc, VEC, SC>(C, P, oo, ii, col_clamp(P, ii, VEC), x);
						}
					return;
				}
			}
#pragma unroll
			for(int j = 0; j < U * VEC; ++j)
				if(j < n) {
					int const f = base + j;
					int64_t oo, iu;
					col_unit(P, og, itile, os, it, f / VEC, U, VEC, oo, iu);
					int64_t const ii = col_clamp(P, iu, VEC) + f % VEC;
					if(oo < P.O && ii >= iu && ii < P.I) store_out<Acc, SC>(C, P, oo, ii, w[j]);
				}
		}
		return;
	} else if(P.red == 2) {
		// partial sums meet in LDS; every thread then finishes some of the item's outputs
		// (output r = ((os*U + u)*IT + it)*VEC + e, partials added in ascending ks)
#pragma unroll
		for(int u = 0; u < U; ++u)
#pragma unroll
			for(int e = 0; e < VEC; ++e) red[(u * VEC + e) * kBlock + t] = v[u][e];
		__syncthreads();
		int const per = P.OS * U * P.IT * VEC;
		for(int r = t; r < per; r += kBlock) {
			int const e  = r % VEC;
			int const r1 = r / VEC;
			int const ir = r1 % P.IT;
			int const r2 = r1 / P.IT;
			int const ur = r2 % U;
			int const orr = r2 / U;
			Acc s = Acc(0);
			const Acc* src = red + (ur * VEC + e) * kBlock + orr * P.IT * P.KS + ir;
			for(int q = 0; q < P.KS; ++q) s += src[q * P.IT];
			int64_t oo, iu;
			col_unit(P, og, itile, orr, ir, ur, U, VEC, oo, iu);
			int64_t const ii = col_clamp(P, iu, VEC) + e;
			if(oo < P.O && ii >= iu && ii < P.I) store_out<Acc, SC, true>(C, P, oo, ii, s);
		}
		__syncthreads();
		return;
	}

	if constexpr(tstore_compiled(VEC * static_cast<int>(sizeof(Acc)), U)) {
		if(P.tstore) {
			col_store_transposed<Acc, VEC, U, SC>(C, P, og, itile, os, it, active, v, red);
			return;
		}
	}
	if(active && ks == 0) {
#pragma unroll
		for(int u = 0; u < U; ++u) {
			int64_t oo, ii;
			col_unit(P, og, itile, os, it, u, U, VEC, oo, ii);
			if(oo < P.O && ii < P.I) store_vec<Acc, VEC, SC>(C, P, oo, ii, col_clamp(P, ii, VEC), v[u]);
		}
	}
}

// Items whose contraction is one k step per lane (K <= KS, shuffle or no reduction): the next
// item's vectors are loaded before the current item is summed and stored, so loads stay in
// flight through the epilogue (short contractions over narrow slabs have little else to overlap).
template<typename S, int VEC, int U>
__device__ __forceinline__ void col_items_pipelined(const S* __restrict__ A, const typename Traits<S>::acc* __restrict__ b,
                                                    void* __restrict__ C, const Args& P, int64_t ib, int64_t ie,
                                                    int64_t is, int64_t kb, int64_t ke, typename Traits<S>::acc* red)
{
	using Acc = typename Traits<S>::acc;
	int const t  = static_cast<int>(threadIdx.x);
	int const it = t % P.IT;
	int const ks = (t / P.IT) % P.KS;
	int const os = t / (P.IT * P.KS);
	bool const active = os < P.OS;
	int64_t const k = kb + ks;
	bool const live = active && k < ke;          // lanes past the end of the contraction add zero
	Acc const bk = live ? b[k] : Acc(0);
	auto load = [&](int64_t item, Vec<S, VEC> (&raw)[U]) {
		int64_t og, itile;
		item_split(P, item, og, itile);
#pragma unroll
		for(int u = 0; u < U; ++u) {
			int64_t o, i;
			col_unit(P, og, itile, os, it, u, U, VEC, o, i);
			int64_t const oc = o < P.O ? o : P.O - 1;
			int64_t const ic = col_clamp(P, i, VEC);
			if(live) raw[u].load(A + oc * P.sAo + k * P.sAk + ic);
			else
#pragma unroll
				for(int w = 0; w < Vec<S, VEC>::words; ++w) raw[u].w[w] = 0u;
		}
	};
	auto finish = [&](int64_t item, const Vec<S, VEC> (&raw)[U]) {
		int64_t og, itile;
		item_split(P, item, og, itile);
		Acc v[U][VEC];
#pragma unroll
		for(int u = 0; u < U; ++u) {
			Acc x[VEC];
			decode(raw[u], x);
#pragma unroll
			for(int e = 0; e < VEC; ++e) v[u][e] = x[e] * bk;
		}
		col_store<Acc, VEC, U, kScaled<S>>(C, P, og, itile, os, it, ks, t, active, v, red);
	};
	Vec<S, VEC> ra[U], rb[U];
	if(ib < ie) load(ib, ra);
	for(int64_t item = ib; item < ie; item += 2 * is) {
		bool const has1 = item + is < ie;
		if(has1) load(item + is, rb);
		finish(item, ra);
		if(item + 2 * is < ie) load(item + 2 * is, ra);
		if(has1) finish(item + is, rb);
	}
}

template<typename S, int VEC, int U, bool KAHAN>
__global__ __launch_bounds__(kBlock) void k_col(const S* __restrict__ A, const typename Traits<S>::acc* __restrict__ b,
                                                void* __restrict__ C, Args P)
{
	using Acc = typename Traits<S>::acc;
	extern __shared__ unsigned char smem_raw[];
	Acc* red = reinterpret_cast<Acc*>(smem_raw);
	int64_t const kb = static_cast<int64_t>(blockIdx.y) * P.kchunk;
	int64_t const ke = kb + P.kchunk < P.K ? kb + P.kchunk : P.K;
	int64_t ib, ie, is;
	block_range(P.contig, P.items, ib, ie, is);
	if(P.one_step) {
		col_items_pipelined<S, VEC, U>(A, b, C, P, ib, ie, is, kb, ke, red);
		return;
	}
	for(int64_t item = ib; item < ie; item += is)
		col_item<S, VEC, U, KAHAN>(A, b, C, P, item, kb, ke, red);
}

// ---- row tiles (q fastest) ----------------------------------------------------------------
// A wave handles a row group g: U*RPW rows (RPW = 64/LPR).  Lane l: row r = l / LPR inside each
// of the U row sets, k offset (l % LPR)*VEC.  Rows are summed over their LPR lanes by recursive
// halving, after which the U*RPW sums of a group are stored by one instruction (consecutive rows).
// vector steps of b a lane keeps in registers for short rows: 16 values of b whatever the type
__host__ __device__ constexpr int dot_breg(int vec) { return vec >= 16 ? 1 : (vec >= 8 ? 2 : 4); }

// Start of the VEC elements loaded for a lane's vector at k (k < ke): k itself, or ke - VEC for the
// partial last vector of a row whose length VEC does not divide (VEC <= the row length)
__device__ __forceinline__ int64_t row_clamp(int64_t k, int64_t ke, int VEC)
{
	return k + VEC <= ke ? k : ke - VEC;
}

// Zeroes the decoded elements before position d: in a clamped vector they belong to the lane before
// (a select, so that no value of another lane reaches the sum, finite or not)
template<typename Acc, int VEC>
__device__ __forceinline__ void mask_head(Acc (&x)[VEC], int64_t d)
{
#pragma unroll
	for(int e = 0; e < VEC; ++e) x[e] = e < d ? Acc(0) : x[e];
}

template<typename S, int VEC>
__device__ __forceinline__ void dot_load_b(const typename Traits<S>::acc* __restrict__ b, const Args& P, int64_t kb,
                                           int64_t ke, typename Traits<S>::acc (&bb)[dot_breg(VEC)][VEC])
{
	using Acc = typename Traits<S>::acc;
	int const kl = static_cast<int>(threadIdx.x) & (P.LPR - 1);
	int64_t const k0 = kb + static_cast<int64_t>(kl) * VEC;
	int64_t const kstride = static_cast<int64_t>(P.LPR) * VEC;
#pragma unroll
	for(int s = 0; s < dot_breg(VEC); ++s) {
		int64_t const ks = k0 + s * kstride;
		int64_t const kc = row_clamp(ks, ke, VEC);        // where the step's vector is loaded from
#pragma unroll
		for(int e = 0; e < VEC; ++e) bb[s][e] = (s < P.kiter && ks < ke) ? b[kc + e] : Acc(0);
	}
}

// sums the lanes' partial row sums of group g and stores the U*RPW rows (non-temporal for 4- and
// 8-byte storage, whose short rows leave whole runs of rows to one store instruction)
template<typename Acc, int U, bool SC, bool NT>
__device__ __forceinline__ void dot_store(void* __restrict__ C, const Args& P, int64_t g, Acc (&v)[U])
{
	int const lane = static_cast<int>(threadIdx.x) & 63;
	int const r    = lane >> P.lprShift;
	int const rpw  = 64 >> P.lprShift;
	int base, n;
	if(halving_sum<Acc, U>(v, lane, 1, P.LPR, base, n)) {
#pragma unroll
		for(int j = 0; j < U; ++j)
			if(j < n) {
				int64_t const row = (g * U + base + j) * rpw + r;
				if(row < P.O) store_out<Acc, SC, NT>(C, P, row, 0, v[j]);
			}
	}
}

// pointer to the start of this lane's piece of row set u of group g (rows past the end clamped:
// those lanes re-read a live row and store nothing)
template<typename S, int U>
__device__ __forceinline__ const S* dot_row_ptr(const S* __restrict__ A, const Args& P, int64_t g, int u, int r, int64_t k0)
{
	int64_t const row = (g * U + u) * (64 >> P.lprShift) + r;
	int64_t const rc  = row < P.O ? row : P.O - 1;
	return A + rc * P.sAo + k0;
}

// rows of at most LPR*VEC elements (one vector per lane and row): double-buffered, the next
// group's vectors are in flight while the current group is summed and stored
template<typename S, int VEC, int U, bool KAHAN>
__device__ __forceinline__ void dot_rows_pipelined(const S* __restrict__ A, void* __restrict__ C, const Args& P,
                                                   int64_t g0, int64_t ge, int64_t gs, int64_t kb, int64_t ke,
                                                   const typename Traits<S>::acc (&bb)[VEC])
{
	using Acc = typename Traits<S>::acc;
	int const lane = static_cast<int>(threadIdx.x) & 63;
	int const r    = lane >> P.lprShift;
	int64_t const k0 = kb + static_cast<int64_t>(lane & (P.LPR - 1)) * VEC;
	bool const live = k0 < ke;      // lanes past the end of the row contribute zero
	int64_t const kc = live ? row_clamp(k0, ke, VEC) : k0;
	auto load = [&](int64_t g, Vec<S, VEC> (&raw)[U]) {
#pragma unroll
		for(int u = 0; u < U; ++u) {
			if(live) raw[u].load(dot_row_ptr<S, U>(A, P, g, u, r, kc));
			else
#pragma unroll
				for(int w = 0; w < Vec<S, VEC>::words; ++w) raw[u].w[w] = 0u;
		}
	};
	auto finish = [&](int64_t g, const Vec<S, VEC> (&raw)[U]) {
		Acc v[U];
#pragma unroll
		for(int u = 0; u < U; ++u) {
			Sum<Acc, KAHAN> acc;
			Acc x[VEC];
			decode(raw[u], x);
			if(P.kpart) mask_head<Acc, VEC>(x, k0 - kc);
#pragma unroll
			for(int e = 0; e < VEC; ++e) acc.fma(x[e], bb[e]);
			v[u] = acc.value();
		}
		dot_store<Acc, U, kScaled<S>, (sizeof(S) >= 4)>(C, P, g, v);
	};
	Vec<S, VEC> ra[U], rb[U];
	if(g0 < ge) load(g0, ra);
	for(int64_t g = g0; g < ge; g += 2 * gs) {
		bool const has1 = g + gs < ge;
		if(has1) load(g + gs, rb);
		finish(g, ra);
		if(g + 2 * gs < ge) load(g + 2 * gs, ra);
		if(has1) finish(g + gs, rb);
	}
}

// short rows: every k step of the lane uses b from registers (P.kiter <= dot_breg(VEC))
template<typename S, int VEC, int U, bool KAHAN>
__device__ __forceinline__ void dot_group_short(const S* __restrict__ A, void* __restrict__ C, const Args& P, int64_t g,
                                                int64_t kb, int64_t ke,
                                                const typename Traits<S>::acc (&bb)[dot_breg(VEC)][VEC])
{
	using Acc = typename Traits<S>::acc;
	int const lane = static_cast<int>(threadIdx.x) & 63;
	int const r    = lane >> P.lprShift;
	int const kl   = lane & (P.LPR - 1);
	int64_t const kstride = static_cast<int64_t>(P.LPR) * VEC;
	int64_t const k0 = kb + static_cast<int64_t>(kl) * VEC;
	const S* pa[U];
#pragma unroll
	for(int u = 0; u < U; ++u) pa[u] = dot_row_ptr<S, U>(A, P, g, u, r, k0);
	Sum<Acc, KAHAN> acc[U];
#pragma unroll
	for(int s = 0; s < dot_breg(VEC); ++s) {
		int64_t const ks = k0 + s * kstride;
		if(s < P.kiter && ks < ke) {
			int64_t const kc = row_clamp(ks, ke, VEC);
			Vec<S, VEC> raw[U];
#pragma unroll
			for(int u = 0; u < U; ++u) raw[u].load(pa[u] + (kc - k0));
#pragma unroll
			for(int u = 0; u < U; ++u) {
				Acc x[VEC];
				decode(raw[u], x);
				if(P.kpart) mask_head<Acc, VEC>(x, ks - kc);
#pragma unroll
				for(int e = 0; e < VEC; ++e) acc[u].fma(x[e], bb[s][e]);
			}
		}
	}
	Acc v[U];
#pragma unroll
	for(int u = 0; u < U; ++u) v[u] = acc[u].value();
	dot_store<Acc, U, kScaled<S>, (sizeof(S) >= 4)>(C, P, g, v);
}

// long rows: b is re-read through the cache, loads of U rows and two k steps overlap
template<typename S, int VEC, int U, bool KAHAN>
__device__ __forceinline__ void dot_group_long(const S* __restrict__ A, const typename Traits<S>::acc* __restrict__ b,
                                               void* __restrict__ C, const Args& P, int64_t g, int64_t kb, int64_t ke)
{
	using Acc = typename Traits<S>::acc;
	int const lane = static_cast<int>(threadIdx.x) & 63;
	int const r    = lane >> P.lprShift;
	int const kl   = lane & (P.LPR - 1);
	int64_t const kstride = static_cast<int64_t>(P.LPR) * VEC;
	int64_t const k0 = kb + static_cast<int64_t>(kl) * VEC;
	const S* pa[U];
#pragma unroll
	for(int u = 0; u < U; ++u) pa[u] = dot_row_ptr<S, U>(A, P, g, u, r, 0);
	Sum<Acc, KAHAN> acc[U];
	int64_t k = k0;
	for(; k + kstride < ke; k += 2 * kstride) {
		// the first step of the pair is whole; the second may be the row's partial last vector
		int64_t const kk[2] = {k, row_clamp(k + kstride, ke, VEC)};
		Vec<S, VEC> raw[2][U];
		Acc bv[2][VEC];
#pragma unroll
		for(int h = 0; h < 2; ++h)
#pragma unroll
			for(int u = 0; u < U; ++u) raw[h][u].load(pa[u] + kk[h]);
#pragma unroll
		for(int h = 0; h < 2; ++h)
#pragma unroll
			for(int e = 0; e < VEC; ++e) bv[h][e] = b[kk[h] + e];
#pragma unroll
		for(int h = 0; h < 2; ++h)
#pragma unroll
			for(int u = 0; u < U; ++u) {
				Acc x[VEC];
				decode(raw[h][u], x);
				if(h == 1 && P.kpart) mask_head<Acc, VEC>(x, k + kstride - kk[1]);
#pragma unroll
				for(int e = 0; e < VEC; ++e) acc[u].fma(x[e], bv[h][e]);
			}
	}
	if(k < ke) {
		int64_t const kc = row_clamp(k, ke, VEC);
		Acc bv[VEC];
#pragma unroll
		for(int e = 0; e < VEC; ++e) bv[e] = b[kc + e];
#pragma unroll
		for(int u = 0; u < U; ++u) {
			Vec<S, VEC> raw;
			raw.load(pa[u] + kc);
			Acc x[VEC];
			decode(raw, x);
			if(P.kpart) mask_head<Acc, VEC>(x, k - kc);
#pragma unroll
			for(int e = 0; e < VEC; ++e) acc[u].fma(x[e], bv[e]);
		}
	}
	Acc v[U];
#pragma unroll
	for(int u = 0; u < U; ++u) v[u] = acc[u].value();
	dot_store<Acc, U, kScaled<S>, (sizeof(S) >= 4)>(C, P, g, v);
}

template<typename S, int VEC, int U, bool KAHAN>
__global__ __launch_bounds__(kBlock) void k_dot(const S* __restrict__ A, const typename Traits<S>::acc* __restrict__ b,
                                                void* __restrict__ C, Args P)
{
	using Acc = typename Traits<S>::acc;
	int64_t const kb = static_cast<int64_t>(blockIdx.y) * P.kchunk;
	int64_t const ke = kb + P.kchunk < P.K ? kb + P.kchunk : P.K;
	// groups: contiguous per block (the block's waves interleaved) or grid-stride over all waves
	int64_t g0, ge, gs;
	if(P.contig) {
		block_range(1, P.groups, g0, ge, gs);
		g0 += threadIdx.x >> 6;
		gs = kBlock / 64;
	} else {
		g0 = static_cast<int64_t>(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
		ge = P.groups;
		gs = static_cast<int64_t>(gridDim.x) * (kBlock / 64);
	}
	if(P.kiter <= dot_breg(VEC)) {
		Acc bb[dot_breg(VEC)][VEC];
		dot_load_b<S, VEC>(b, P, kb, ke, bb);
		if(P.kiter == 1) {
			dot_rows_pipelined<S, VEC, U, KAHAN>(A, C, P, g0, ge, gs, kb, ke, bb[0]);
		} else {
			for(int64_t g = g0; g < ge; g += gs) dot_group_short<S, VEC, U, KAHAN>(A, C, P, g, kb, ke, bb);
		}
	} else {
		for(int64_t g = g0; g < ge; g += gs) dot_group_long<S, VEC, U, KAHAN>(A, b, C, P, g, kb, ke);
	}
}

// ---- one-wave-per-group kernels ----------------------------------------------------------------
// Streaming A as one compact front: every wave (k_rows) or block (k_slab) reads one small work unit
// (2-32 KiB), finishes it and exits, and the dispatcher hands out the next units in address order.
// Per unit there is no loop and little set-up, so the kernels stay lean (the per-unit cost is
// what limits this regime).  Measured on MI355X, FP32 128^4: rows of 128 elements 6.71 -> 7.02 TB/s
// against row groups looped over a capped grid; 128 x 128 slabs in 256-byte column chunks 6.61 ->
// 6.81 TB/s against whole-row items looped over a capped grid.

// b[k0 .. k0+VEC) into registers, as 16-byte loads whatever b's alignment (global memory takes
// element-aligned vector loads) when a lane's values fill whole 16-byte vectors
typedef uint32_t u32x4_e __attribute__((ext_vector_type(4), aligned(4)));
template<typename Acc, int VEC>
__device__ __forceinline__ void load_b_vec(const Acc* __restrict__ b, int64_t k0, Acc (&bb)[VEC])
{
	constexpr int BYTES = VEC * static_cast<int>(sizeof(Acc));
	if constexpr(BYTES % 16 == 0) {
#pragma unroll
		for(int c = 0; c < BYTES / 16; ++c) {
			u32x4_e const q = *reinterpret_cast<const u32x4_e*>(reinterpret_cast<const unsigned char*>(b + k0) + 16 * c);
			reinterpret_cast<uint32_t*>(bb)[4 * c]     = q.x;
			reinterpret_cast<uint32_t*>(bb)[4 * c + 1] = q.y;
			reinterpret_cast<uint32_t*>(bb)[4 * c + 2] = q.z;
			reinterpret_cast<uint32_t*>(bb)[4 * c + 3] = q.w;
		}
	} else {
#pragma unroll
		for(int e = 0; e < VEC; ++e) bb[e] = b[k0 + e];
	}
}

// Sums each segment of `len` consecutive lanes (segment s: lanes s*len .. s*len+len-1) into the
// segment's first lane, for U values per lane: shifts by 1, 2, 4, ... with the lanes past their
// segment's end adding nothing.
template<typename Acc, int U>
__device__ __forceinline__ void segment_sum(Acc (&v)[U], int kl, int len)
{
	for(int d = 1; d < len; d <<= 1) {
		bool const take = kl + d < len;
#pragma unroll
		for(int u = 0; u < U; ++u) {
			Acc const t = __shfl_down(v[u], d, 64);
			v[u] += take ? t : Acc(0);
		}
	}
}

// Rows of K <= 64*VEC contiguous elements (one vector per lane and row, P.kiter == 1): wave
// blockIdx.x*4 + w sums the U*RPW rows of its group and stores their sums.  Every load is issued
// before the first use and none sits behind a branch: lanes past the end of a row re-read the row's
// first vector and add nothing.  LPRC > 0 fixes the lanes per row (a power of two) at compile time
// (constant cross-lane offsets); LPRC == 0 reads the power of two from P.LPR; LPRC < 0 ("tight")
// takes any P.LPR = ceil(K / VEC) lanes per row, RPW = 64 / LPR rows per wave, the lanes past RPW*LPR
// idle, and sums rows by segmented shifts (9 lanes per row: 63 of 64 lanes busy instead of 9 of 16).
template<typename S, int VEC, int U, int LPRC, bool KAHAN>
__device__ __forceinline__ void rows_group(const S* __restrict__ A, const typename Traits<S>::acc* __restrict__ b,
                                           void* __restrict__ C, const Args& P, int64_t g)
{
	using Acc = typename Traits<S>::acc;
	int const lane = static_cast<int>(threadIdx.x) & 63;
	if constexpr(LPRC < 0) {
		int const lpr = P.LPR;
		int const rpw = 64 / lpr;
		int const r   = lane / lpr;
		int const kl  = lane - r * lpr;
		int64_t const k0 = static_cast<int64_t>(kl) * VEC;
		bool const live = r < rpw && k0 < P.K;
		int64_t const kc = live ? row_clamp(k0, P.K, VEC) : 0;
		Acc bb[VEC];
		load_b_vec<Acc, VEC>(b, kc, bb);
		Vec<S, VEC> raw[U];
		int64_t const row0 = g * U * rpw + (r < rpw ? r : 0);
#pragma unroll
		for(int u = 0; u < U; ++u) {
			int64_t const row = row0 + static_cast<int64_t>(u) * rpw;
			raw[u].load(A + (row < P.O ? row : P.O - 1) * P.sAo + kc);
		}
		Acc v[U];
#pragma unroll
		for(int u = 0; u < U; ++u) {
			Sum<Acc, KAHAN> acc;
			Acc x[VEC];
			decode(raw[u], x);
			if(P.kpart) mask_head<Acc, VEC>(x, k0 - kc);
#pragma unroll
			for(int e = 0; e < VEC; ++e) acc.fma(x[e], bb[e]);
			v[u] = live ? acc.value() : Acc(0);
		}
		segment_sum<Acc, U>(v, kl, lpr);
		if(kl == 0 && r < rpw) {
#pragma unroll
			for(int u = 0; u < U; ++u) {
				int64_t const row = row0 + static_cast<int64_t>(u) * rpw;
				if(row < P.O) store_out<Acc, kScaled<S>, (sizeof(S) >= 4)>(C, P, row, 0, v[u]);
			}
		}
		return;
	}
	int const lpr   = LPRC > 0 ? LPRC : P.LPR;
	int const shift = LPRC > 0 ? __builtin_ctz(LPRC) : P.lprShift;
	int const rpw   = 64 >> shift;
	int const r    = lane >> shift;
	int64_t const k0 = static_cast<int64_t>(lane & (lpr - 1)) * VEC;
	bool const live = k0 < P.K;
	int64_t const kc = live ? row_clamp(k0, P.K, VEC) : 0;
	Acc bb[VEC];
	load_b_vec<Acc, VEC>(b, kc, bb);
	// rows of the group: (g*U + u)*rpw + r; a group past the last row re-reads the last row
	Vec<S, VEC> raw[U];
	int64_t const row0 = g * U * rpw + r;
#pragma unroll
	for(int u = 0; u < U; ++u) {
		int64_t const row = row0 + static_cast<int64_t>(u) * rpw;
		raw[u].load(A + (row < P.O ? row : P.O - 1) * P.sAo + kc);
	}
	Acc v[U];
#pragma unroll
	for(int u = 0; u < U; ++u) {
		Sum<Acc, KAHAN> acc;
		Acc x[VEC];
		decode(raw[u], x);
		if(P.kpart) mask_head<Acc, VEC>(x, k0 - kc);
#pragma unroll
		for(int e = 0; e < VEC; ++e) acc.fma(x[e], bb[e]);
		v[u] = live ? acc.value() : Acc(0);
	}
	int base, n;
	bool store;
	if constexpr(LPRC > 0) store = halving_sum_ct<Acc, U, LPRC>(v, lane, base, n);
	else                   store = halving_sum<Acc, U>(v, lane, 1, lpr, base, n);
	if(store) {
#pragma unroll
		for(int j = 0; j < U; ++j)
			if(j < n) {
				int64_t const row = row0 + static_cast<int64_t>(base + j) * rpw;
				if(row < P.O) store_out<Acc, kScaled<S>, (sizeof(S) >= 4)>(C, P, row, 0, v[j]);
			}
	}
}

template<typename S, int VEC, int U, int LPRC, bool KAHAN>
__global__ __launch_bounds__(kBlock) void k_rows(const S* __restrict__ A, const typename Traits<S>::acc* __restrict__ b,
                                                 void* __restrict__ C, Args P)
{
	int64_t const g = static_cast<int64_t>(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
	if(g < P.groups) rows_group<S, VEC, U, LPRC, KAHAN>(A, b, C, P, g);
}

// Column items of one block each (col_narrow): item -> (OS slabs, chunk of IT*VEC columns); thread t
// -> it = t % IT, ks = (t / IT) % KS (lane groups along k), os = t / (IT*KS), all of K; IT, KS, OS
// powers of two with IT*KS*OS = 256.  Partial sums meet by shuffles inside a wave and, when one slab's
// lanes span several waves, in LDS; the slab's first IT lanes store its IT*VEC outputs.
template<typename S, int VEC, bool KAHAN>
__device__ __forceinline__ void slab_item(const S* __restrict__ A, const typename Traits<S>::acc* __restrict__ b,
                                          void* __restrict__ C, const Args& P, uint32_t item,
                                          typename Traits<S>::acc (*red)[64 * VEC])
{
	using Acc = typename Traits<S>::acc;
	constexpr int KL = sizeof(S) >= 4 ? TTVX_KL_WIDE : TTVX_KL_NARROW;    // loads in flight per lane
	int const t    = static_cast<int>(threadIdx.x);
	int const span = P.IT * P.KS;                // lanes of one slab
	int const it   = t & (P.IT - 1);
	int const ks   = (t >> P.itShift) & (P.KS - 1);
	int const os   = t >> P.spanShift;
	// item -> (slab group, column chunk): one 32-bit division at most
	uint32_t const q32 = P.iTiles == 1 ? item : item / static_cast<uint32_t>(P.iTiles);
	int64_t const og = q32, itile = item - q32 * static_cast<uint32_t>(P.iTiles);
	int64_t const o  = og * P.OS + os;
	int64_t const oc = o < P.O ? o : P.O - 1;    // tail slabs and columns re-read live data, store nothing
	int64_t const i  = (itile * P.IT + it) * VEC;
	int64_t const ic = col_clamp(P, i, VEC);
	const S* pa = A + oc * P.sAo + static_cast<int64_t>(ks) * P.sAk + ic;
	int64_t const kstep = static_cast<int64_t>(P.KS) * P.sAk;
	Sum<Acc, KAHAN> acc[VEC];
	// rounds of KL loads issued together; a lane's steps past its last row re-read the round's first
	// row (and are not added), so short contractions keep all their loads in flight too
	int64_t const nk = ks < P.K ? (P.K - 1 - ks) / P.KS + 1 : 0;
	for(int64_t j0 = 0; j0 < nk; j0 += KL) {
		Vec<S, VEC> raw[KL];
		Acc bk[KL];
#pragma unroll
		for(int j = 0; j < KL; ++j) raw[j].load(pa + (j0 + j < nk ? j : 0) * kstep);
#pragma unroll
		for(int j = 0; j < KL; ++j) bk[j] = b[ks + (j0 + j < nk ? j0 + j : j0) * P.KS];
		__builtin_amdgcn_sched_barrier(0);
#pragma unroll
		for(int j = 0; j < KL; ++j) {
			if(j0 + j < nk) {          // re-read rows are not added (a non-finite value must not leak)
				Acc x[VEC];
				decode(raw[j], x);
#pragma unroll
				for(int e = 0; e < VEC; ++e) acc[e].fma(x[e], bk[j]);
			}
		}
		pa += KL * kstep;
	}
	Acc v[VEC];
#pragma unroll
	for(int e = 0; e < VEC; ++e) v[e] = acc[e].value();
	// lanes it + IT*m of one wave hold partial sums of the same columns of the same slab
	int const wspan = span < 64 ? span : 64;
	for(int d = P.IT; d < wspan; d <<= 1)
#pragma unroll
		for(int e = 0; e < VEC; ++e) v[e] += __shfl_xor(v[e], d, 64);
	bool const live = o < P.O && i < P.I;
	if(span <= 64) {
		if(ks == 0 && live) store_vec<Acc, VEC, kScaled<S>>(C, P, o, i, ic, v);
		return;
	}
	// a slab's lanes span span/64 waves: their first IT lanes meet in LDS
	int const lane = t & 63, w = t >> 6;
	if(lane < P.IT)
#pragma unroll
		for(int e = 0; e < VEC; ++e) red[w][lane * VEC + e] = v[e];
	__syncthreads();
	if((t & (span - 1)) < P.IT && live) {
		int const w0 = w, gw = span / 64;
#pragma unroll
		for(int e = 0; e < VEC; ++e) {
			Acc s = red[w0][it * VEC + e];
			for(int ww = 1; ww < gw; ++ww) s += red[w0 + ww][it * VEC + e];
			v[e] = s;
		}
		store_vec<Acc, VEC, kScaled<S>>(C, P, o, i, ic, v);
	}
	__syncthreads();       // the next item of this block may write red
}

template<typename S, int VEC, bool KAHAN>
__global__ __launch_bounds__(kBlock) void k_slab(const S* __restrict__ A, const typename Traits<S>::acc* __restrict__ b,
                                                 void* __restrict__ C, Args P)
{
	__shared__ typename Traits<S>::acc red[kBlock / 64][64 * VEC];
	slab_item<S, VEC, KAHAN>(A, b, C, P, blockIdx.x, red);
}

// ---- k_strided ------------------------------------------------------------------------------
constexpr int kMaxFree = 32;
struct StridedArgs {
	int     nd;
	int64_t n[kMaxFree], sa[kMaxFree], sc[kMaxFree];   // free dims, fastest first
	int64_t K, sAk, outputs;
	const float* scales;
	int64_t fps;
};

template<typename S, bool KAHAN>
__global__ __launch_bounds__(kBlock) void k_strided(const S* __restrict__ A, const typename Traits<S>::acc* __restrict__ b,
                                                    void* __restrict__ C, StridedArgs G)
{
	using Acc = typename Traits<S>::acc;
	for(int64_t j = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; j < G.outputs;
	    j += static_cast<int64_t>(gridDim.x) * kBlock) {
		int64_t rest = j, offa = 0, offc = 0;
		for(int d = 0; d < G.nd; ++d) {
			int64_t const idx = rest % G.n[d];
			rest /= G.n[d];
			offa += idx * G.sa[d];
			offc += idx * G.sc[d];
		}
		Sum<Acc, KAHAN> acc;
		const S* pa = A + offa;
		for(int64_t k = 0; k < G.K; ++k) {
			Vec<S, 1> raw;
			raw.load(pa + k * G.sAk);
			Acc x[1];
			decode(raw, x);
			acc.fma(x[0], b[k]);
		}
		Acc v = acc.value();
		if(G.scales != nullptr) v *= static_cast<Acc>(G.scales[j / G.fps]);
		static_cast<Acc*>(C)[offc] = v;
	}
}

// ---- k_finalize -----------------------------------------------------------------------------
template<typename Acc>
__global__ __launch_bounds__(kBlock) void k_finalize(const void* __restrict__ partial, int part64, int splits,
                                                     int64_t O, int64_t I, int64_t sCo, void* __restrict__ C,
                                                     const float* __restrict__ scales, int64_t fps)
{
	int64_t const n = O * I;
	for(int64_t j = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; j < n;
	    j += static_cast<int64_t>(gridDim.x) * kBlock) {
		Acc v;
		if(part64) {
			double s = 0.0;
			for(int q = 0; q < splits; ++q) s += static_cast<const double*>(partial)[q * n + j];
			v = static_cast<Acc>(s);
		} else {
			Acc s = Acc(0);
			for(int q = 0; q < splits; ++q) s += static_cast<const Acc*>(partial)[q * n + j];
			v = s;
		}
		if(scales != nullptr) v *= static_cast<Acc>(scales[j / fps]);
		int64_t const o = j / I, i = j - o * I;
		static_cast<Acc*>(C)[o * sCo + i] = v;
	}
}

// The same sums for many splits of few outputs: block j sums output j's partials (thread t takes
// splits t, t + kBlock, ...; then the waves and the block, in a fixed order), in FP64.
template<typename Acc>
__global__ __launch_bounds__(kBlock) void k_finalize_wide(const void* __restrict__ partial, int part64, int splits,
                                                          int64_t O, int64_t I, int64_t sCo, void* __restrict__ C,
                                                          const float* __restrict__ scales, int64_t fps)
{
	__shared__ double red[kBlock / 64];
	int64_t const n = O * I;
	int const t = static_cast<int>(threadIdx.x);
	for(int64_t j = blockIdx.x; j < n; j += gridDim.x) {
		double s = 0.0;
		for(int q = t; q < splits; q += kBlock)
			s += part64 ? static_cast<const double*>(partial)[q * n + j]
			            : static_cast<double>(static_cast<const Acc*>(partial)[q * n + j]);
#pragma unroll
		for(int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
		if((t & 63) == 0) red[t >> 6] = s;
		__syncthreads();
		if(t == 0) {
			double sum = 0.0;
#pragma unroll
			for(int w = 0; w < kBlock / 64; ++w) sum += red[w];
			Acc v = static_cast<Acc>(sum);
			if(scales != nullptr) v *= static_cast<Acc>(scales[j / fps]);
			int64_t const o = j / I, i = j - o * I;
			static_cast<Acc*>(C)[o * sCo + i] = v;
		}
		__syncthreads();
	}
}

} // namespace ttvx
