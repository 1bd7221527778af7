/*
 * ttvx.h -- C-ABI of the MI355X (gfx950) mode-q tensor-times-vector library libttvx.so.
 *
 *   C(i_1..i_{q-1}, i_{q+1}..i_p) = sum_{i_q} A(i_1..i_p) * b(i_q)
 *
 * Conventions are those of the reference's low-level interface (include/tlib/ttv.h:54-92):
 * q and layout tuples are one-based, extents/strides are indexed by mode, pi[r] is the mode
 * at layout position r (r = 0 fastest), nb[0] must equal na[q-1], b is contiguous.
 * Strides are counted in elements.  C is overwritten (beta = 0), never accumulated into.
 *
 * Every entry point returns TTVX_OK (0) or a nonzero status; the text of the last error of
 * the calling thread is returned by ttvx_last_error().  Argument errors carry exactly the
 * std::runtime_error messages of the reference (include/tlib/ttv.h:64-89 and
 * include/tlib/detail/tensor_times_vector.h:160,164), so C++ and Python wrappers can
 * re-throw them verbatim.
 *
 * Pointers: the synchronous entry points accept host or device memory for a, b and c (host
 * buffers are staged through device memory and the call blocks until C is written back); the
 * *_async entry points take device pointers only and enqueue on `stream` (a hipStream_t; NULL
 * selects the default stream).
 *
 * No torch types and no C++ types cross this boundary.
 */
#ifndef TTVX_H
#define TTVX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------- */
#define TTVX_OK               0
#define TTVX_ERR_ARGUMENT     1  /* argument validation failed (reference message)               */
#define TTVX_ERR_LAYOUT       2  /* pia/pic inconsistent (tensor_times_vector.h:147-168 message)  */
#define TTVX_ERR_DEVICE       3  /* HIP runtime error or no usable gfx950 device                 */
#define TTVX_ERR_UNSUPPORTED  4  /* combination not provided by this build                       */

/* ---- element types -------------------------------------------------------------------- */
#define TTVX_F32      0  /* float  A, b, C                                      */
#define TTVX_F64      1  /* double A, b, C                                      */
#define TTVX_BF16     2  /* bfloat16 A (uint16_t storage), float b and C         */
#define TTVX_FP8E4M3  3  /* OCP FP8 e4m3 A (uint8_t storage) + float scales; float b and C */

/* ---- policy codes (include/tlib/detail/tags.h:21-63); hints only, every combination accepted */
#define TTVX_EXEC_SEQ           0
#define TTVX_EXEC_SEQ_BLAS      1
#define TTVX_EXEC_PAR           2
#define TTVX_EXEC_PAR_TASKLOOP  3
#define TTVX_EXEC_PAR_TASK      4
#define TTVX_EXEC_PAR_LOOP      5
#define TTVX_EXEC_PAR_BLAS      6
#define TTVX_EXEC_PAR_BLAS_LOOP 7
#define TTVX_SLICE_SLICE        0
#define TTVX_SLICE_SUBTENSOR    1
#define TTVX_FUSE_NONE          0
#define TTVX_FUSE_OUTER         1
#define TTVX_FUSE_ALL           2

/* ---- accumulation flags ------------------------------------------------------------- */
#define TTVX_FLAG_KAHAN      1u  /* compensated (Kahan) summation along n_q                       */
#define TTVX_FLAG_FP64_FINAL 2u  /* FP64 cross-lane / cross-block reduction (automatic when n_q > 2^20) */

/* Text of the calling thread's last error ("" when the last call succeeded). */
const char* ttvx_last_error(void);

/* Library version as major*10000 + minor*100 + patch. */
int ttvx_version(void);

/* Number of visible HIP devices; 0 without a GPU.  Never fails. */
int ttvx_device_count(void);

/* ---- device memory for resident tensors ------------------------------------------------ */
/* Back the device-resident tensor of include/tlib/device_tensor.h (the reference's tensor<T>,
   include/tlib/detail/tensor.h:56-114, keeps std::vector storage; on the GPU the data stays in HBM
   between products).  Allocation on the current device (gfx950 required); copies are synchronous
   and take host<->device or device<->device pointers.  ttvx_device_free(NULL) is a no-op. */
int ttvx_device_alloc(size_t bytes, void** out);
int ttvx_device_free(void* ptr);
int ttvx_copy(void* dst, const void* src, size_t bytes);

/* ---- plain drop-in ------------------------------------------------------------------- */
/* Replaces tlib::ttv::ttv<float,std::size_t,ep,sp,fp>(ep,sp,fp,q,p,a,na,wa,pia,b,nb,c,nc,wc,pic)
 * (reference include/tlib/ttv.h:54-92, dispatching to include/tlib/detail/tensor_times_vector.h).
 * ep/sp/fp: TTVX_EXEC_* / TTVX_SLICE_* / TTVX_FUSE_* codes.  Synchronous. */
int ttvx_ttv_f32(int ep, int sp, int fp, size_t q, size_t p,
                 const float* a, const size_t* na, const size_t* wa, const size_t* pia,
                 const float* b, const size_t* nb,
                 float* c, const size_t* nc, const size_t* wc, const size_t* pic);

/* Same for double (tlib::ttv::ttv<double,...>, include/tlib/ttv.h:54-92). */
int ttvx_ttv_f64(int ep, int sp, int fp, size_t q, size_t p,
                 const double* a, const size_t* na, const size_t* wa, const size_t* pia,
                 const double* b, const size_t* nb,
                 double* c, const size_t* nc, const size_t* wc, const size_t* pic);

/* Stream-ordered forms of the two above: device pointers only; flags = TTVX_FLAG_*. */
int ttvx_ttv_f32_async(size_t q, size_t p,
                       const float* a, const size_t* na, const size_t* wa, const size_t* pia,
                       const float* b, const size_t* nb,
                       float* c, const size_t* nc, const size_t* wc, const size_t* pic,
                       unsigned flags, void* stream);
int ttvx_ttv_f64_async(size_t q, size_t p,
                       const double* a, const size_t* na, const size_t* wa, const size_t* pia,
                       const double* b, const size_t* nb,
                       double* c, const size_t* nc, const size_t* wc, const size_t* pic,
                       unsigned flags, void* stream);

/* 2-D slice product (matrix times vector) on the strided view A(j, k) = a[j*wj + k*wk], j < m, k < n:
 *   c[j] = sum_k A(j, k) * b[k]            accumulate == 0  (gemv_row*, gemv_col_blas, dot*)
 *   c[j] = c[j] + sum_k A(j, k) * b[k]     accumulate != 0  (gemv_col, gemv_col_parallel)
 * Replaces the slice kernels of include/tlib/detail/matrix_times_vector.h:51-295 (the non-BLAS
 * gemv_col forms add into c, :117-126).  c is contiguous; host or device pointers; synchronous. */
int ttvx_mtv_f32(const float* a, const float* b, float* c, size_t m, size_t n, size_t wj, size_t wk, int accumulate);
int ttvx_mtv_f64(const double* a, const double* b, double* c, size_t m, size_t n, size_t wj, size_t wk, int accumulate);

/* Same as ttvx_ttv_f32_async, but allocates no device memory (no counterpart in the reference: for
 * callers that cannot allocate, such as code under HIP graph capture or allocation-free serving
 * loops).  The split-n_q / FP64-final partial sums go to a static 16 MiB device array of the
 * library (one per GPU, held until the process ends) instead of the library's cached buffers;
 * problems whose partials need more return TTVX_ERR_UNSUPPORTED.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (other streams, other threads) race. */
int ttvx_ttv_f32_async_noalloc(size_t q, size_t p,
                               const float* a, const size_t* na, const size_t* wa, const size_t* pia,
                               const float* b, const size_t* nb,
                               float* c, const size_t* nc, const size_t* wc, const size_t* pic,
                               unsigned flags, void* stream);

/* ---- reduced-precision storage of A (no counterpart in the reference) ------------------ */
/* A keeps its extents, strides and layout; only the element type changes.  Accumulation is
 * FP32 (optionally Kahan), b and C are float.
 *
 * FP8 scales: the mode-q fibers of A are numbered in the order of C's dense memory offsets
 * j (j = o*I + i in the folded [O][K][I] view).  Scale g multiplies the fibers
 * g*fibers_per_scale <= j < (g+1)*fibers_per_scale; it is absmax(group)/448 (1 for an all-zero
 * group) and the stored code of x is RNE-e4m3(x * (448/absmax)), saturated to +-448. */

/* Round-to-nearest-even float -> bfloat16 over `count` contiguous elements. */
int ttvx_quantize_bf16_async(size_t count, const float* src, uint16_t* dst, void* stream);

/* Number of FP8 scales for (q, p, na, wa, pia) and the given group size, 0 on error. */
size_t ttvx_fp8_num_scales(size_t q, size_t p, const size_t* na, const size_t* wa, const size_t* pia,
                           size_t fibers_per_scale);

/* Quantises FP32 A into e4m3 codes `aq` (same strides) and per-group `scales`. */
int ttvx_quantize_fp8_async(size_t q, size_t p,
                            const float* a, const size_t* na, const size_t* wa, const size_t* pia,
                            uint8_t* aq, float* scales, size_t fibers_per_scale, void* stream);

/* C = A x_q b with bfloat16 A. */
int ttvx_ttv_bf16_async(size_t q, size_t p,
                        const uint16_t* a, const size_t* na, const size_t* wa, const size_t* pia,
                        const float* b, const size_t* nb,
                        float* c, const size_t* nc, const size_t* wc, const size_t* pic,
                        unsigned flags, void* stream);

/* C = A x_q b with e4m3 A and its scales. */
int ttvx_ttv_fp8_async(size_t q, size_t p,
                       const uint8_t* a, const size_t* na, const size_t* wa, const size_t* pia,
                       const float* scales, size_t fibers_per_scale,
                       const float* b, const size_t* nb,
                       float* c, const size_t* nc, const size_t* wc, const size_t* pic,
                       unsigned flags, void* stream);

/* ---- planning, variants and selection (no counterpart in the reference: the reference's
 *      policy tags, tags.h:21-63, are chosen by the caller) ---------------------------- */
#define TTVX_SELECT_HEURISTIC 0  /* fixed rule                                     */
#define TTVX_SELECT_LEARNED   1  /* MLP selector (falls back to the rule without weights) */
#define TTVX_SELECT_FORCED    2  /* the variant given to ttvx_set_selection      */
#define TTVX_SELECT_AUTO      3  /* default: the tuning table, then the MLP selector, then the fixed
                                    rule.  Unless set explicitly, the shipped table and weights are
                                    loaded on first use ($TTVX_SELECTOR_TABLE / $TTVX_SELECTOR_WEIGHTS,
                                    else data/selector_table_gfx950.bin and data/selector_gfx950.bin
                                    next to the library's lib/ directory) */

#define TTVX_NUM_FEATURES 24

/* Number of kernel variants and the name of variant v ("" if out of range). */
int ttvx_num_variants(void);
const char* ttvx_variant_name(int v);

/* Selection mode for subsequent calls of this process (all threads). */
int ttvx_set_selection(int mode, int forced_variant);

/* Folded view of (q, p, na, wa, pia) for element type dtype:
 *   out[0]=O, out[1]=K, out[2]=I, out[3]=1 if the folded 3-D view exists else 0,
 *   out[4]=reference case number 1..8 (cases.h:24-36), out[5]=variant the current selection picks. */
int ttvx_plan_info(int dtype, size_t q, size_t p, const size_t* na, const size_t* wa, const size_t* pia,
                   int64_t* out6);

/* 1 if variant v can run (q, p, na, wa, pia) in dtype, else 0. */
int ttvx_variant_valid(int dtype, size_t q, size_t p, const size_t* na, const size_t* wa, const size_t* pia,
                       int v);

/* Selector input features of a problem (TTVX_NUM_FEATURES floats). */
int ttvx_features(int dtype, size_t q, size_t p, const size_t* na, const size_t* wa, const size_t* pia,
                  float* features);

/* Number of floats in the selector MLP's parameter vector, and loading it (count 0 unloads).
 * Layout: mean[24], inv_std[24], W1[64][24], b1[64], W2[64][64], b2[64], W3[15][64], b3[15];
 * the network predicts log2(t_variant / t_fastest) for every variant. */
size_t ttvx_selector_num_params(void);
int ttvx_selector_set_weights(const float* params, size_t count);

/* Tuning table of the exhaustive autotuner: `count` rows of 5 int64 (dtype, O, K, I, variant), the
 * measured fastest variant of folded problems (O, K, I as in ttvx_plan_info); the learned and auto
 * modes take a row's variant for a 16-byte aligned A before asking the MLP.  count 0 clears it. */
int ttvx_selector_set_table(const int64_t* rows, size_t count);

/* ---- selector network on the GPU (no counterpart in the reference) --------------------- */
/* Device-pointer kernels that train and evaluate the selector MLP (parameter layout above).
 * work: ttvx_mlp_work_floats(batch) floats holding the activations of the last forward pass. */
size_t ttvx_mlp_work_floats(int batch);
/* y[batch][15] = MLP(x[batch][24]). */
int ttvx_mlp_forward(const float* params, const float* x, int batch, float* y, float* work, void* stream);
/* Masked mean-squared error 0.5*sum(m*(y-t)^2)/batch into *loss and its gradient w.r.t. the
 * trainable parameters (everything after mean/inv_std, in parameter order) into grad. */
int ttvx_mlp_backward(const float* params, const float* y, const float* t, const float* m, int batch, float* work,
                      float* grad, float* loss, void* stream);
/* One Adam step (bias-corrected, step >= 1, decoupled weight decay) on the trainable parameters;
 * m1/m2 are the moment buffers. */
int ttvx_mlp_adam(float* params, const float* grad, float* m1, float* m2, int step, float lr, float beta1, float beta2,
                  float eps, float weight_decay, void* stream);

/* ---- batched products (no counterpart in the reference: it runs one product per call) ---- */
/* One FP32 problem of a batch; C's extents and layout are derived from A as in the reference
 * (shape.h:103-158, layout.h:143-207); wc = NULL means C's dense strides in that layout.
 * variant < 0 takes the library's selection; variant >= 0 fixes the kernel variant of this problem
 * (the per-product counterpart of the reference's policy tags, tags.h:21-63). */
typedef struct ttvx_batch_problem {
	size_t q, p;
	const float* a;
	const size_t* na;
	const size_t* wa;
	const size_t* pia;
	const float* b;
	float* c;
	const size_t* wc;
	int variant;
} ttvx_batch_problem;

/* Validates and plans `count` problems (device pointers, fixed for the batch's lifetime) and
 * returns an opaque handle.  Each problem keeps the tile configuration of its variant (its own, or
 * the one a single call selects, split variants in their unsplit form); contractions longer than
 * 2^20, problems that fold to the strided kernel and variants that cannot run their problem are
 * rejected (TTVX_ERR_UNSUPPORTED). */
int ttvx_batch_create(size_t count, const ttvx_batch_problem* problems, void** batch);
/* Computes every problem of the batch on `stream`: one launch per group of problems that share a
 * tile body (a handful per batch); with use_graph != 0 the launches are captured into a hipGraph
 * on first use and replayed afterwards.  Must be called on the device the batch was created on. */
int ttvx_batch_launch(void* batch, int use_graph, void* stream);
/* out4 = {blocks over all launches, algorithmic bytes, launches, problems}. */
int ttvx_batch_info(void* batch, int64_t* out4);
/* Kernel variant (0 .. ttvx_num_variants()-1) whose tiles run problem `index` of the batch: the
 * problem's own variant, or the one a single call selects, split variants replaced by their
 * unsplit forms (col_split -> col_u1, dot_split -> dot_u2).  A single call forced to it gives
 * bit-identical outputs.  -1 on error. */
int ttvx_batch_variant(void* batch, size_t index);
int ttvx_batch_destroy(void* batch);

/* ---- fused multi-vector contraction (ttvs) ------------------------------------------------ */
/* out[i] = prod_m vectors[m][i_m] with i = i_0 + n_0*(i_1 + n_1*(...)) (vector 0 fastest), on the
 * device, stream-ordered; 1..32 device vectors.  The weights of the fused form of ttvpy's ttvs
 * (ttvpy/src/wrapped_ttv.cpp:141-194): with A dense, every mode but q contracts as one row product
 * with the Kronecker product of the vectors of the modes stored before q and one column product
 * with that of the modes stored after q (no counterpart in the reference, which chains p - 1
 * products). */
int ttvx_kron_f32(const float* const* vectors, const size_t* lengths, size_t count, float* out, void* stream);
int ttvx_kron_f64(const double* const* vectors, const size_t* lengths, size_t count, double* out, void* stream);

/* ---- fault injection (no counterpart in the reference; a test hook) --------------------- */
#define TTVX_FAULT_ALLOC  0  /* device allocations: staging, split partials, batches, ttvx_device_alloc */
#define TTVX_FAULT_COPY   1  /* host <-> device copies of the synchronous entry points */
#define TTVX_FAULT_LAUNCH 2  /* product launches (the kernel is not enqueued) */
/* The next `count` operations at `site` fail as if HIP had reported out of memory, an invalid copy
 * or a launch failure, without touching the device (count 0 disarms): the library's error paths
 * (status, message, state left for the next call) can be exercised on a healthy GPU.  Process-wide. */
int ttvx_debug_fail_next(int site, int count);

/* ---- diagnostics (no counterpart in the reference; used by the benchmark) -------------- */
/* Pure read stream over `bytes` (16-byte aligned), each block one contiguous chunk, writing one
 * float per block to `partial` (blocks = min(partial_count, 64 per CU)). */
int ttvx_bench_read(const void* src, size_t bytes, float* partial, size_t partial_count, void* stream);
/* Reads src_bytes and writes dst_bytes (16-byte aligned) in one launch, each block its contiguous
 * share of both (a product whose output is dst_bytes/src_bytes of its input). */
int ttvx_bench_mix(const void* src, size_t src_bytes, void* dst, size_t dst_bytes, void* stream);
/* The same traffic as a compact sweep: one short-lived block per loads_per_thread (2, 4 or 8) x 256
 * 16-byte vectors of src, in address order, each writing its share of dst (dst_bytes may be 0;
 * sink is one device float).  The benchmark's ceilings are the faster of the two patterns. */
int ttvx_bench_sweep(const void* src, size_t src_bytes, void* dst, size_t dst_bytes, int loads_per_thread,
                     float* sink, void* stream);
/* Plain copy of `bytes` (16-byte aligned): a known-size kernel for counter calibration. */
int ttvx_bench_copy(const void* src, void* dst, size_t bytes, void* stream);
/* Bounded busy wait of one wave (iterations <= 100000 sleeps of ~8k cycles): keeps the queue
 * busy while the host enqueues timed launches. */
int ttvx_bench_spin(int iterations, int* sink, void* stream);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* TTVX_H */
