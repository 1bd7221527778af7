/*
 * ttv_oracle.c -- CPU oracle of the mode-q tensor-times-vector product.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libttvx.so, include/, ttv_amd/) links,
 * loads or calls this file.  It is used by tests/ as the checker, by __graft_entry__.smoke()
 * as the checker, and by bench.py as the timed CPU baseline ("port" of the reference's CPU
 * path).  Every function names the reference file:line it restates.
 *
 * Contents
 *   oracle_ttv_naive_f64 / _f32in   textbook definition, FP64 accumulation, any strides
 *                                    (README.md:13-18; include/tlib/ttv.h:28-53)
 *   oracle_ttv_log_f32 / _f64       the reference's default CPU path: mtv(par_blas) for cases
 *                                    1-7 (matrix_times_vector.h:356-373) and
 *                                    par_loop/subtensor/all for case 8
 *                                    (tensor_times_vector.h:1207-1256), BLAS gemv per slice
 *   oracle_blas_load                 binds cblas_[sd]gemv from an OpenBLAS shared library
 *   oracle_number_cores              physical cores (tensor_times_vector.h:56-88)
 *   oracle_ttv_init_known            known-answer initialisation (test/src/gtest_tlib_ttv.cpp:27-71)
 *
 * Pinning (how each function is checked, see tests/test_oracle.py):
 *   naive      known-answer closed form of gtest_tlib_ttv.cpp:132, numpy.einsum, the
 *              ttvpy test vectors (ttvpy/tests/test.py), brute force on tiny shapes, and the
 *              compiled reference (oracle/_ref) on seeded fixtures
 *   log        the same known answers, the compiled reference on fixtures, and naive
 *   init_known the closed form itself (C[j] = nq^2 j + nq(nq+1)/2 for b = 1)
 *   cores      sockets x cores per socket of /sys topology, capped by online CPUs
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <omp.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ---- naive definition ----------------------------------------------------------------------
 * C(i_1..i_{q-1},i_{q+1}..i_p) = sum_{i_q} A(i) b(i_q); C modes are A's modes without q.  wc is
 * indexed by C mode.  Accumulates in double in ascending i_q order.  Any strides. */
static void naive_core(size_t q, size_t p, const void* a, int a_is_f32, const size_t* na, const size_t* wa,
                       const void* b, int b_is_f32, double* c, const size_t* wc)
{
	size_t nout = 1;
	for(size_t r = 0; r < p; ++r)
		if(r != q - 1) nout *= na[r];
	size_t const nq = na[q - 1], wq = wa[q - 1];
#pragma omp parallel for schedule(static)
	for(long long jj = 0; jj < (long long)nout; ++jj) {
		/* mixed-radix decode of jj over the free modes in mode order (mode 1 fastest) */
		size_t rest = (size_t)jj, offa = 0, offc = 0, cm = 0;
		for(size_t r = 0; r < p; ++r) {
			if(r == q - 1) continue;
			size_t const idx = rest % na[r];
			rest /= na[r];
			offa += idx * wa[r];
			offc += idx * wc[cm];
			++cm;
		}
		double s = 0.0;
		for(size_t k = 0; k < nq; ++k) {
			double const av = a_is_f32 ? (double)((const float*)a)[offa + k * wq] : ((const double*)a)[offa + k * wq];
			double const bv = b_is_f32 ? (double)((const float*)b)[k] : ((const double*)b)[k];
			s += av * bv;
		}
		c[offc] = s;
	}
}

void oracle_ttv_naive_f64(size_t q, size_t p, const double* a, const size_t* na, const size_t* wa,
                          const double* b, double* c, const size_t* wc)
{
	naive_core(q, p, a, 0, na, wa, b, 0, c, wc);
}

void oracle_ttv_naive_f32in(size_t q, size_t p, const float* a, const size_t* na, const size_t* wa,
                            const float* b, double* c, const size_t* wc)
{
	naive_core(q, p, a, 1, na, wa, b, 1, c, wc);
}

/* ---- BLAS binding --------------------------------------------------------------------------- */
typedef void (*sgemv_t)(int, int, int, int, float, const float*, int, const float*, int, float, float*, int);
typedef void (*dgemv_t)(int, int, int, int, double, const double*, int, const double*, int, double, double*, int);
typedef void (*setthr_t)(int);
enum { ROW_MAJOR = 101, COL_MAJOR = 102, NO_TRANS = 111 };   /* CBLAS enumerators */

static sgemv_t  g_sgemv = NULL;
static dgemv_t  g_dgemv = NULL;
static setthr_t g_setthr = NULL;

/* Binds the gemv and thread-count symbols; tries plain and "scipy_"-prefixed names.
 * Returns 1 when BLAS is bound, 0 otherwise (the non-BLAS loops are used then). */
int oracle_blas_load(const char* path)
{
	void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
	if(!h) return 0;
	const char* pre[2] = {"", "scipy_"};
	for(int i = 0; i < 2 && !g_sgemv; ++i) {
		char n1[64], n2[64], n3[64];
		snprintf(n1, sizeof n1, "%scblas_sgemv", pre[i]);
		snprintf(n2, sizeof n2, "%scblas_dgemv", pre[i]);
		snprintf(n3, sizeof n3, "%sopenblas_set_num_threads", pre[i]);
		g_sgemv  = (sgemv_t)dlsym(h, n1);
		g_dgemv  = (dgemv_t)dlsym(h, n2);
		g_setthr = (setthr_t)dlsym(h, n3);
	}
	if(!g_sgemv || !g_dgemv) { g_sgemv = NULL; g_dgemv = NULL; g_setthr = NULL; return 0; }
	return 1;
}

int oracle_blas_loaded(void) { return g_sgemv != NULL; }

/* ---- core count (tensor_times_vector.h:56-88: sockets x cores per socket, capped by the
 *      number of logical CPUs) --------------------------------------------------------------- */
unsigned oracle_number_cores(void)
{
	long online = sysconf(_SC_NPROCESSORS_ONLN);
	if(online < 1) online = 1;
	/* distinct (package, core) pairs of the online CPUs */
	int pairs[4096][2];
	int npairs = 0;
	for(long cpu = 0; cpu < 4096 && npairs < 4096; ++cpu) {
		char path[128];
		int pkg = -1, core = -1;
		FILE* f;
		snprintf(path, sizeof path, "/sys/devices/system/cpu/cpu%ld/topology/physical_package_id", cpu);
		if(!(f = fopen(path, "r"))) { if(cpu >= online * 4) break; continue; }
		if(fscanf(f, "%d", &pkg) != 1) pkg = -1;
		fclose(f);
		snprintf(path, sizeof path, "/sys/devices/system/cpu/cpu%ld/topology/core_id", cpu);
		if(!(f = fopen(path, "r"))) continue;
		if(fscanf(f, "%d", &core) != 1) core = -1;
		fclose(f);
		int seen = 0;
		for(int i = 0; i < npairs; ++i)
			if(pairs[i][0] == pkg && pairs[i][1] == core) { seen = 1; break; }
		if(!seen) { pairs[npairs][0] = pkg; pairs[npairs][1] = core; ++npairs; }
	}
	unsigned n = (unsigned)npairs;
	if(n == 0 || n > (unsigned)online) n = (unsigned)online;
	return n;
}

/* ---- slice kernels (matrix_times_vector.h) ------------------------------------------------- */
/* gemv_col: c[j] = sum_i a[j + i*lda] b[i], j < M, i < N (column-major slice, :108-127 with
 * the BLAS overwrite semantics of gemv_col_blas :189-221). */
static void gemv_col_f32(const float* a, const float* b, float* c, size_t M, size_t N, size_t lda)
{
	if(g_sgemv) { g_sgemv(COL_MAJOR, NO_TRANS, (int)M, (int)N, 1.0f, a, (int)lda, b, 1, 0.0f, c, 1); return; }
	for(size_t j = 0; j < M; ++j) c[j] = 0.f;
	for(size_t i = 0; i < N; ++i) {
		float const bb = b[i];
		const float* a0 = a + i * lda;
		for(size_t j = 0; j < M; ++j) c[j] += a0[j] * bb;
	}
}
static void gemv_col_f64(const double* a, const double* b, double* c, size_t M, size_t N, size_t lda)
{
	if(g_dgemv) { g_dgemv(COL_MAJOR, NO_TRANS, (int)M, (int)N, 1.0, a, (int)lda, b, 1, 0.0, c, 1); return; }
	for(size_t j = 0; j < M; ++j) c[j] = 0.0;
	for(size_t i = 0; i < N; ++i) {
		double const bb = b[i];
		const double* a0 = a + i * lda;
		for(size_t j = 0; j < M; ++j) c[j] += a0[j] * bb;
	}
}
/* gemv_row: c[i] = sum_k a[i*lda + k] b[k], i < M, k < N (row-major slice, :51-69, :231-262) */
static void gemv_row_f32(const float* a, const float* b, float* c, size_t M, size_t N, size_t lda)
{
	if(g_sgemv) { g_sgemv(ROW_MAJOR, NO_TRANS, (int)M, (int)N, 1.0f, a, (int)lda, b, 1, 0.0f, c, 1); return; }
#pragma omp parallel for schedule(static)
	for(long long i = 0; i < (long long)M; ++i) {
		float s = 0.f;
		const float* r = a + (size_t)i * lda;
		for(size_t k = 0; k < N; ++k) s += r[k] * b[k];
		c[i] = s;
	}
}
static void gemv_row_f64(const double* a, const double* b, double* c, size_t M, size_t N, size_t lda)
{
	if(g_dgemv) { g_dgemv(ROW_MAJOR, NO_TRANS, (int)M, (int)N, 1.0, a, (int)lda, b, 1, 0.0, c, 1); return; }
#pragma omp parallel for schedule(static)
	for(long long i = 0; i < (long long)M; ++i) {
		double s = 0.0;
		const double* r = a + (size_t)i * lda;
		for(size_t k = 0; k < N; ++k) s += r[k] * b[k];
		c[i] = s;
	}
}

/* ---- the reference's default CPU path ------------------------------------------------------
 * Returns the reference case number (cases.h:24-36).  `threads` = OpenMP threads for case 8 and
 * BLAS threads for cases 2-7 (0 = oracle_number_cores()).  Dense strides are assumed, as in the
 * reference (mtv ignores wa/wc, matrix_times_vector.h:309-311; the subtensor path asserts
 * n == wa_m, tensor_times_vector.h:1038). */
#define LOG_BODY(T, ROW, COL)                                                                          \
	unsigned const cores = threads > 0 ? (unsigned)threads : oracle_number_cores();                  \
	size_t const m = q;                                                                               \
	size_t N = 1;                                                                                     \
	for(size_t r = 0; r < p; ++r) N *= na[r];                                                         \
	size_t const n = N / na[m - 1];                                                                   \
	int cnr;                                                                                          \
	if(p == 1) cnr = 1;                                                                               \
	else if(p == 2) cnr = (q == 1 ? (pia[0] == 1 ? 2 : 4) : (pia[0] == 1 ? 3 : 5));                   \
	else if(pia[0] == q) cnr = 6;                                                                     \
	else if(pia[p - 1] == q) cnr = 7;                                                                 \
	else cnr = 8;                                                                                     \
	if(cnr != 8) {                                                                                    \
		/* mtv(par_blas), matrix_times_vector.h:356-373, after set_blas_threads_max */                \
		if(g_setthr) g_setthr((int)cores);                                                            \
		omp_set_num_threads((int)cores);                                                              \
		switch(cnr) {                                                                                 \
			case 1: { T s = 0; for(size_t k = 0; k < na[0]; ++k) s += a[k] * b[k]; c[0] = s; } break; \
			case 2: ROW(a, b, c, na[1], na[0], na[0]); break;                                         \
			case 3: COL(a, b, c, na[0], na[1], na[0]); break;                                         \
			case 4: COL(a, b, c, na[1], na[0], na[1]); break;                                         \
			case 5: ROW(a, b, c, na[0], na[1], na[1]); break;                                         \
			case 6: ROW(a, b, c, n, na[m - 1], na[m - 1]); break;                                     \
			case 7: COL(a, b, c, n, na[m - 1], n); break;                                             \
		}                                                                                             \
		return cnr;                                                                                   \
	}                                                                                                 \
	/* case 8: par_loop/subtensor/all, tensor_times_vector.h:1207-1256: one single-threaded gemv    \
	 * of the (w_q x n_q, lda w_q) subtensor per index of the fused outer modes */                   \
	size_t inv = 0;                                                                                   \
	while(pia[inv] != m) ++inv;                                                                       \
	size_t const inv_pia_m = inv + 1;                                                                 \
	size_t num = 1;                                                                                   \
	for(size_t i = inv_pia_m; i < p; ++i) num *= na[pia[i] - 1];                                      \
	size_t const wa_m = wa[m - 1], na_m = na[m - 1];                                                  \
	size_t const wa_m1 = wa[pia[inv_pia_m] - 1];                                                      \
	size_t const wc_m1 = wc[pic[inv_pia_m - 1] - 1];                                                  \
	if(g_setthr) g_setthr(1);                                                                         \
	_Pragma("omp parallel for schedule(static) num_threads(cores) proc_bind(spread)")                  \
	for(long long i = 0; i < (long long)num; ++i)                                                     \
		COL(a + (size_t)i * wa_m1, b, c + (size_t)i * wc_m1, wa_m, na_m, wa_m);                       \
	return 8;

int oracle_ttv_log_f32(size_t q, size_t p, const float* a, const size_t* na, const size_t* wa, const size_t* pia,
                       const float* b, float* c, const size_t* wc, const size_t* pic, int threads)
{
	LOG_BODY(float, gemv_row_f32, gemv_col_f32)
}

int oracle_ttv_log_f64(size_t q, size_t p, const double* a, const size_t* na, const size_t* wa, const size_t* pia,
                       const double* b, double* c, const size_t* wc, const size_t* pic, int threads)
{
	LOG_BODY(double, gemv_row_f64, gemv_col_f64)
}

/* ---- known-answer initialisation (test/src/gtest_tlib_ttv.cpp:27-71) ------------------------
 * Every mode-q fiber is filled with consecutive integers 1,2,...; fibers are visited with the
 * layout positions other than q's nested from the slowest (outermost loop) to the fastest, so
 * the fiber with C memory offset j holds j*n_q + 1 ... j*n_q + n_q. */
void oracle_ttv_init_known(size_t q, size_t p, double* a, const size_t* na, const size_t* wa, const size_t* pia)
{
	size_t qpos = 0;
	while(pia[qpos] != q) ++qpos;
	size_t nfib = 1;
	for(size_t r = 0; r < p; ++r)
		if(r != q - 1) nfib *= na[r];
	size_t const nq = na[q - 1], wq = wa[q - 1];
	for(size_t f = 0; f < nfib; ++f) {
		/* decode f over layout positions except qpos, fastest first */
		size_t rest = f, off = 0;
		for(size_t s = 0; s < p; ++s) {
			if(s == qpos) continue;
			size_t const md = pia[s] - 1;
			off += (rest % na[md]) * wa[md];
			rest /= na[md];
		}
		for(size_t k = 0; k < nq; ++k) a[off + k * wq] = (double)(f * nq + k + 1);
	}
}
