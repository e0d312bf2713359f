/*
 * tn_hip.h - C ABI of the MI355X (gfx950) tensor-network kernel library.
 *
 * This is the backend seam of Uni10: the free functions that the tensor
 * layer calls with raw element pointers (reference:
 * uni10/src/uni10/tools/uni10_tools.h:56-135 and
 * uni10/src/uni10/numeric/lapack/uni10_lapack.h:54-83), restated as a plain C
 * ABI over device memory.  Every entry point:
 *   - takes plain pointers and sizes (no C++ or torch types),
 *   - returns 0 on success or a negative TN_E* code; the message of the last
 *     failure on the calling thread is available from tn_last_error(),
 *   - never throws and never calls exit()/abort() (the reference's CUDA
 *     helpers exit() on error, helper_uni10.h:86-105),
 *   - enqueues on the given stream (NULL = the library's default stream) and
 *     returns without synchronising unless it says otherwise.
 * All element buffers are FP64, row-major, resident in device memory.
 */
#ifndef TN_HIP_H
#define TN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define TN_OK 0
#define TN_EINVAL -1     /* bad argument (shape, permutation, mode, NULL)   */
#define TN_ENOMEM -2     /* device allocation failed or budget exceeded      */
#define TN_EHIP -3       /* a HIP runtime call failed                        */
#define TN_ENODEV -4     /* no usable gfx950 device                          */
#define TN_ENOCONV -5    /* iterative solver did not converge                */
#define TN_EUNSUPP -6    /* shape outside what the kernel supports           */

/* Message of the last failure on this thread ("" if none). */
const char* tn_last_error(void);

/* ---- device, streams, memory --------------------------------------------
 * Replaces guni10_create/guni10_destroy and the elem* family
 * (uni10_tools_gpu.cu:7-163: cudaMallocManaged, cudaMemcpy, cudaMemset).
 * Device memory comes from a caching pool with an optional byte budget. */
int tn_init(int device);                 /* select the device; idempotent     */
int tn_device_info(int* cu_count, size_t* hbm_bytes, char* arch, int arch_len);
int tn_set_budget(size_t bytes);         /* 0 = unlimited                      */
int tn_mem_stats(size_t* in_use, size_t* cached, size_t* peak, size_t* budget);
int tn_malloc(void** dptr, size_t bytes);            /* elemAlloc             */
int tn_free(void* dptr);                             /* elemFree              */
int tn_pool_trim(void);                              /* release cached blocks */
/* The stream the calling thread's tn_malloc/tn_free are ordered on (NULL =
 * the library's default stream).  A freed block is reused at once by an
 * allocation on the stream it was freed on; an allocation on another stream
 * first makes that stream wait for the freeing stream's work.  So a caller
 * frees a buffer after enqueuing its last use on its pool stream, or after
 * synchronising the stream that used it.  A stream given here outlives the
 * blocks freed on it, or is destroyed through tn_stream_destroy (which
 * drains it first).  No counterpart in the reference, whose frees are
 * synchronous (cudaFree). */
int tn_pool_set_stream(void* stream);
int tn_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int tn_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
/* Pinned host memory and an asynchronous device-to-host copy into it: the
 * copy is ordered on `stream` and complete once an event recorded after it
 * is (tn_event_query), so a caller polls results without blocking. */
int tn_host_alloc(void** p, size_t bytes);
int tn_host_free(void* p);
int tn_memcpy_d2h_async(void* dst_pinned, const void* src, size_t bytes, void* stream);
int tn_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
int tn_zero(void* dst, size_t bytes, void* stream);  /* elemBzero (fill kernel) */
int tn_stream_create(void** stream);
int tn_stream_destroy(void* stream);
int tn_stream_sync(void* stream);
/* *done = 1 when every piece of work enqueued on the stream has finished,
 * 0 while some is still running (never blocks) */
int tn_stream_query(void* stream, int* done);
int tn_device_sync(void);
/* Events for timing on a given stream (milliseconds between two records). */
int tn_event_create(void** ev);
int tn_event_destroy(void* ev);
int tn_event_record(void* ev, void* stream);
int tn_event_elapsed_ms(void* ev_start, void* ev_stop, float* ms);
int tn_event_query(void* ev, int* done);   /* 1 once the recorded work is done; never blocks */

/* HIP graphs: capture every call made on `stream` between begin and end into
 * an executable graph (no host round trips inside: calls that copy from or
 * to host memory synchronise and are not capturable).  Replaying it issues
 * the same kernels with no host gaps.  No memset nodes are recorded: the
 * library clears memory with fill kernels.  Device memory allocated from the
 * pool during a capture belongs to the graph (its replays write it) and is
 * returned to the pool only when the graph is destroyed, whenever its holder
 * frees it; memory freed during a capture is likewise held until then.  One
 * capture at a time per process; tn_graph_destroy synchronises the device. */
int tn_graph_begin(void* stream);
int tn_graph_end(void* stream, void** graph_exec);
int tn_graph_launch(void* graph_exec, void* stream);
int tn_graph_destroy(void* graph_exec);

/* ---- permute / reshapeElem ----------------------------------------------
 * Dense (non-symmetric) permute.  Output axis j takes source axis perm[j]
 * (rsp_outin of UniTensorReal.cpp:545-558); dst has dims[perm[j]].
 * Bit-exact copy.  Replaces reshapeElem/_reshapeElem
 * (uni10_tools_gpu.cu:319-339) and the host odometer of
 * UniTensorReal.cpp:601-646. */
int tn_permute_dense_f64(const double* src, double* dst, int rank,
                         const int64_t* dims, const int* perm, void* stream);

/* reshapeElem with the reference's own argument convention
 * (uni10_tools_gpu.cu:332-339, table built at UniTensorReal.cpp:586-595):
 * offset[0..r) are the source row-major strides, offset[r+i] is the
 * destination stride of source axis i. */
int tn_reshape_elem_f64(const double* old_elem, int bond_num, size_t elem_num,
                        const size_t* offset, double* new_elem, void* stream);

/* Symmetric (block-sparse) permute: one move per source sub-block (Qidx),
 * the per-Qidx copy of UniTensorReal.cpp:683-756, in one launch.
 * Host arrays, nmoves entries:
 *   meta[6*m+0..5] = {src corner offset, src block cols (Bin_cDim),
 *                     src sub-block cols (sBin_cDim), dst corner offset,
 *                     dst block cols (Bot_cDim), dst sub-block cols (sBot_cDim)}
 *   sub_dims[rank*m+b] = source sub-block extent of bond b (sBin_sBdims)
 *   bot_acc[rank*m+b]  = destination linear stride of source bond b (sBot_acc)
 *   sign[m]            = +1/-1 (fermionic swap sign; NULL = all +1)
 * A plan uploads these once; executing it is one kernel launch. */
typedef struct tn_blockperm_plan_s* tn_blockperm_plan;
int tn_blockperm_plan_create(tn_blockperm_plan* plan, int rank, int64_t nmoves,
                             const int64_t* meta, const int64_t* sub_dims,
                             const int64_t* bot_acc, const double* sign,
                             int64_t dst_elems);
int tn_blockperm_exec(tn_blockperm_plan plan, const double* src, double* dst,
                      void* stream);
int tn_blockperm_plan_destroy(tn_blockperm_plan plan);

/* In-place scaling of rectangular sub-blocks of an element buffer: the
 * swap-gate sign loop of addGate (UniTensorReal.cpp:868-921), which
 * multiplies every Qidx sub-block whose fermion-parity product is odd by -1.
 * rects[4*i..4*i+3] = {element offset of the corner, rows, cols, row pitch}
 * (host array, uploaded once); exec multiplies every listed element by a.
 * Rectangles must not overlap. */
typedef struct tn_rectscale_plan_s* tn_rectscale_plan;
int tn_rectscale_plan_create(tn_rectscale_plan* plan, int64_t nrects, const int64_t* rects);
int tn_rectscale_exec(tn_rectscale_plan plan, double* base, double a, void* stream);
int tn_rectscale_plan_destroy(tn_rectscale_plan plan);

/* Row-major transpose AT(N x M) = A(M x N)^T; replaces setTranspose
 * (uni10_lapack_cpu.cpp:350-355; GPU uni10_lapack_gpu.cu:568-598). */
int tn_transpose_f64(const double* A, int64_t M, int64_t N, double* AT, void* stream);

/* ---- grouped sector GEMM -------------------------------------------------
 * The per-Qnum loop of contract (UniTensorTools.cpp:191-203), each step a
 * matrixMul (uni10_lapack_cpu.cpp:41-44; GPU uni10_lapack_gpu.cu:42-152):
 * row-major C_g(M x N) = A_g(M x K) B_g(K x N) for every group member g, in
 * one launch.  Offsets are in elements from the A/B/C base pointers.
 * Modes: arithmetic of the products; storage stays FP64. */
#define TN_GEMM_F64 0      /* FP64 MFMA, FP64 accumulate (plain drop-in)        */
#define TN_GEMM_F32 1      /* FP32 MFMA on FP64->FP32 rounded inputs, FP32 acc. */
#define TN_GEMM_BF16X3 2   /* split-BF16 (hi+lo), 3 products, FP32 accumulate    */
#define TN_GEMM_BF16X6 3   /* split-BF16 (hi+mid+lo), 6 products, FP32 accumulate*/
/* FP64-grade split-BF16: each A row / B column scaled by a power of two and
 * cut into six 7-bit integer slices (exact in BF16); the 21 slice products of
 * total weight >= 2^-49 on BF16 MFMA, whose FP32 sums stay exact integers
 * and are moved to FP64 every 128 k; error about K 2^-49 relative to
 * max|A row| max|B column|.  Needs finite inputs. */
#define TN_GEMM_BF16_EXACT 4
typedef struct {
  int64_t M, N, K;
  int64_t a_off, b_off, c_off;
  int64_t lda, ldb, ldc;
} tn_gemm_desc;
typedef struct tn_gemm_plan_s* tn_gemm_plan;
int tn_gemm_plan_create(tn_gemm_plan* plan, int mode, int64_t ngroups,
                        const tn_gemm_desc* descs);
/* Multi-stride operand addressing, so a contraction reads its inputs (and
 * writes its output) in any bond order without permuting them.  An index map
 * sends a row or column index i of an operand to an element offset: with
 * n >= 1 levels, i is split into digits over ext[0..n) (last level fastest)
 * and offset(i) = sum_l digit_l * stride[l]; the product of ext[] must equal
 * the index range (M, N or K).  n = 0 means the plain layout: row i ->
 * i * ld, column i -> i.  Indices must stay below 2^31.
 *   A(m,k) = A[a_off + a_row(m) + a_col(k)]
 *   B(k,n) = B[b_off + b_row(k) + b_col(n)]
 *   C(m,n) = C[c_off + c_row(m) + c_col(n)]                                  */
#define TN_MAX_MAP_LEVELS 4
typedef struct {
  int32_t n;
  int32_t pad;
  int64_t ext[TN_MAX_MAP_LEVELS];
  int64_t stride[TN_MAX_MAP_LEVELS];
} tn_index_map;
/* Optional row-scaled second output (all zero: none).  With rs_div > 0 the
 * launch also writes D(m,n) = C(m,n) * S[rs_off + m / rs_div] to
 * C[d_off + c_row(m) + c_col(n)]; S and D are addressed from C's base
 * pointer.  The stored C value times the scale is rounded once, exactly as
 * the separate diagRowMul pass it replaces (uni10_lapack_cpu.cpp:46-50,
 * uni10_lapack_gpu.cu:312-339), which it saves: one read of C and one
 * launch (the batched TEBD engine's theta = lambda C). */
typedef struct {
  int64_t M, N, K;
  int64_t a_off, b_off, c_off;
  int64_t lda, ldb, ldc;
  tn_index_map a_row, a_col, b_row, b_col, c_row, c_col;
  int64_t d_off, rs_off, rs_div;
} tn_gemm_desc_ex;
int tn_gemm_plan_create_ex(tn_gemm_plan* plan, int mode, int64_t ngroups,
                           const tn_gemm_desc_ex* descs);
int tn_gemm_plan_set_tile(tn_gemm_plan plan, int tile_id); /* autotuner hook  */
/* Autotuner: runs the plan with every tile shape on the given operands
 * (reps timed launches each, HIP events on `stream`), keeps the fastest and
 * reports its time.  Every tile shape accumulates K in the same order, so
 * the choice never changes the results.  C is overwritten. */
int tn_gemm_plan_autotune(tn_gemm_plan plan, const double* A, const double* B, double* C,
                          int reps, void* stream, float* best_ms);
int tn_gemm_plan_info(tn_gemm_plan plan, int64_t* tiles, double* flops, int* tile_id);
/* Device-side activity mask (no counterpart in the reference; for batched
 * engines whose slots are not all busy): with mask != NULL, group g is
 * skipped -- its output keeps its contents -- while mask[group_idx[g]] == 0
 * (group_idx[g] < 0: never).  mask is a device int array the caller updates
 * between launches (captured graphs read it on every replay); NULL removes
 * the mask.  The same for SVD plans (per problem: outputs, warm-start data
 * and convergence counts untouched) and vector-op plans (per item). */
int tn_gemm_plan_set_mask(tn_gemm_plan plan, const int* mask, const int32_t* group_idx);
int tn_gemm_exec(tn_gemm_plan plan, const double* A, const double* B, double* C,
                 void* stream);
int tn_gemm_plan_destroy(tn_gemm_plan plan);
/* One matrixMul (a group of one), same argument order as the reference. */
int tn_matmul_f64(const double* A, const double* B, int64_t M, int64_t N, int64_t K,
                  double* C, int mode, void* stream);
/* Complex FP64, interleaved (re, im) pairs.  tn_permute_dense_c128: the
 * complex dense permute (UniTensorComplex.cpp) as the real permute with the
 * pair axis kept last, bit-exact.  tn_matmul_c128: row-major C = A B with
 * A (M x K), B (K x N) complex (matrixMul on std::complex<double>,
 * uni10_lapack_cpu.cpp), as one FP64 MFMA product of A (M x 2K real) with
 * the 2K x 2N real expansion of B; synchronises the stream. */
int tn_permute_dense_c128(const double* src, double* dst, int rank, const int64_t* dims,
                          const int* perm, void* stream);
int tn_matmul_c128(const double* A, const double* B, int64_t M, int64_t N, int64_t K, double* C,
                   void* stream);
/* Complex SVD (matrixSVD on std::complex<double>, uni10_lapack_cpu.cpp:606-633,
 * zgesvd): A (M x N, interleaved, row-major) = U diag(S) VT with U (M x k) and
 * VT (k x N) complex row-major, k = min(M, N), S descending; U may be NULL.
 * One-sided complex Jacobi (each column pair rotated by a real Jacobi angle
 * after a phase), one workgroup; waits for the stream. */
int tn_svd_c128(const double* A, int64_t M, int64_t N, double* U, double* S, double* VT, void* stream);
/* Hermitian eigen solver (eigSyDecompose on std::complex<double>,
 * uni10_lapack_cpu.cpp:865-890, zheev): eigenvalues ascending; row k of
 * vec_rows (N x N complex) is the conjugate of eigenvector k, so that
 * H = vec_rows^H diag(eig) vec_rows (zheev's vectors read back row-major).
 * Computed as the SVD of H + 2 ||H||_F I (positive definite, so its right
 * singular vectors are H's eigenvectors); waits for the stream. */
int tn_eigh_c128(const double* A, int64_t N, double* eig, double* vec_rows, void* stream);
/* X (n x 2 x 2 real, per complex element b: [re, im, -im, re]): the factor
 * that turns a complex product into a real one (C = A X over the pair
 * axis), for block-sparse complex contractions. */
int tn_expand_c128(const double* B, int64_t n, double* X, void* stream);

/* ---- vectors / diagonal scaling -----------------------------------------
 * vectorNorm (uni10_lapack_cpu.cpp:397-414), vectorScal (:75-89),
 * diagRowMul / diagColMul (:46-57).  tn_norm2_f64 writes ||x||_2 to a device
 * double with a fixed reduction order (deterministic, run to run). */
int tn_norm2_f64(const double* x, int64_t n, double* out_dev, void* stream);
/* vectorSum (uni10_lapack_cpu.cpp:387-395): x[0] + x[inc] + ... (n terms)
 * to a device double, in a fixed order; vectorAdd (:59-74, daxpy):
 * y += a x; the strided copy dst[i*dst_inc] = src[i*src_inc] (i < n) of
 * setDiag / getDiag (uni10_tools_cpu.cpp:170-181). */
int tn_sum_f64(const double* x, int64_t n, int64_t inc, double* out_dev, void* stream);
/* Seeded fill of elemRand / randomize (uni10_tools.h:67,101): dst[i] =
 * SplitMix64 uniform[-1, 1) of (seed, i + 1), generated on the device. */
int tn_fill_uniform_f64(double* dst, int64_t n, uint64_t seed, void* stream);
int tn_axpy_f64(double a, const double* x, double* y, int64_t n, void* stream);
int tn_copy_strided_f64(double* dst, int64_t dst_inc, const double* src, int64_t src_inc, int64_t n,
                        void* stream);
int tn_scale_f64(double* x, int64_t n, double a, void* stream);
int tn_scale_inv_dev_f64(double* x, int64_t n, const double* denom_dev, void* stream);
int tn_diag_row_mul_f64(double* mat, const double* diag, int64_t M, int64_t N, void* stream);
int tn_diag_col_mul_f64(double* mat, const double* diag, int64_t M, int64_t N, void* stream);

/* ---- dense FP64 decompositions -------------------------------------------
 * matrixSVD (uni10_lapack_cpu.cpp:289-313): thin SVD of the row-major
 * M x N matrix A: U (M x k), S (k, descending), VT (k x N), k = min(M, N).
 * One-sided Jacobi; singular vectors are unique up to a sign per pair.
 * U may be NULL (only S and VT are formed; for M >= N and min(M, N) <= 64
 * this takes the faster QR-preconditioned sweeps, as does the block path
 * for M >= N, M <= 4096). */
int tn_svd_f64(const double* A, int64_t M, int64_t N, double* U, double* S,
               double* VT, void* stream);
/* tn_svd_f64 without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * kernel and results (bit for bit) as tn_svd_f64, limited to the
 * one-workgroup size min(M, N) <= 64, max(M, N) <= 128 (larger shapes fail
 * with TN_EUNSUPP).  The problem descriptor and convergence counters live in
 * a 256-byte static device array per device, held until the process ends and
 * shared by every call on that device: calls that overlap on one device race.
 * Synchronises the stream (it reads the convergence counters). */
int tn_svd_f64_noalloc(const double* A, int64_t M, int64_t N, double* U, double* S, double* VT,
                       void* stream);
/* Batched: problem i reads A + a_off[i] (m[i] x n[i]) and writes
 * U + u_off[i], S + s_off[i], VT + vt_off[i]. */
int tn_svd_f64_batched(int64_t batch, const int64_t* m, const int64_t* n,
                       const double* A, const int64_t* a_off,
                       double* U, const int64_t* u_off,
                       double* S, const int64_t* s_off,
                       double* VT, const int64_t* vt_off, void* stream);
/* Jacobi sweeps used by the last SVD call (batch: the largest) on this
 * thread; a convergence diagnostic. */
int tn_svd_last_sweeps(int* sweeps);
/* Solver choice (tuning hook, no counterpart in the reference): a problem
 * with min(M, N) <= 64 and max(M, N) <= 128 is solved by one wavefront-
 * parallel workgroup; a larger one with min(M, N) < n_min that fits one
 * workgroup's LDS by one workgroup in LDS; the rest (min(M, N) <= 8192) by
 * block Jacobi (32-column panels, FP64 MFMA Gram and update products, many
 * workgroups per matrix).  n_min = 0 restores the default, 65 (block Jacobi
 * for everything beyond the small kernel).  Applies to plans created
 * afterwards. */
int tn_svd_set_block_min(int64_t n_min);
/* Sweep cap (test and tuning hook, no counterpart in the reference): plans
 * and calls created afterwards report a problem as not converged once it
 * has run n sweeps without converging (n = 0 restores the defaults, 60 for
 * the one-workgroup kernels and 40 for block Jacobi).  The kernels stop
 * normally at the cap. */
int tn_svd_set_max_sweeps(int n);

/* Warm start for a nearby matrix (M >= N): V0T is an N x N matrix with
 * orthonormal rows, e.g. the VT of a previous SVD of a matrix close to A;
 * the Jacobi sweeps start from W = A V0 instead of A, which cuts the sweeps
 * when A changes little between calls (successive TEBD updates).  The result
 * is the SVD of A, as from tn_svd_f64 (U may be NULL likewise).  No
 * counterpart in the reference. */
int tn_svd_f64_warm(const double* A, int64_t M, int64_t N, const double* V0T, double* U,
                    double* S, double* VT, void* stream);

/* A reusable batch of SVD problems (shapes and offsets fixed at creation;
 * v0_off[i] >= 0 gives problem i a warm start from V0T + v0_off[i], NULL or
 * -1 = cold; u_off[i] = -1: U of problem i is not formed, and U may be NULL
 * at execution when no problem forms it).  tn_svd_plan_exec only launches kernels: it neither allocates
 * nor synchronises, so it can be captured into a graph (tn_graph_begin).
 * tn_svd_plan_check waits for the stream and reports the problems that did
 * not converge and the most sweeps any needed, counted over every execution
 * since the previous check (or the plan's creation), then clears the counts.
 * The batched matrixSVD of many chains (uni10_lapack_cpu.cpp:289-313 per
 * chain in the reference). */
typedef struct tn_svd_plan_s* tn_svd_plan;
int tn_svd_plan_create(tn_svd_plan* plan, int64_t batch, const int64_t* m, const int64_t* n,
                       const int64_t* a_off, const int64_t* v0_off, const int64_t* u_off,
                       const int64_t* s_off, const int64_t* vt_off);
int tn_svd_plan_exec(tn_svd_plan plan, const double* A, const double* V0T, double* U, double* S,
                     double* VT, void* stream);
int tn_svd_plan_check(tn_svd_plan plan, int* nonconverged, int* sweeps, void* stream);
int tn_svd_plan_set_mask(tn_svd_plan plan, const int* mask, const int32_t* prob_idx);
int tn_svd_plan_destroy(tn_svd_plan plan);

/* eigSyDecompose (uni10_lapack_cpu.cpp:131-152): symmetric N x N;
 * eigenvalues ascending; eigenvectors returned as ROWS of EigVec. */
int tn_eigh_f64(const double* A, int64_t N, double* eig, double* eigvec_rows,
                void* stream);
/* Many eigSyDecompose calls (uni10_lapack_cpu.cpp:131-152) in one launch,
 * one workgroup per matrix: problem i is the n[i] x n[i] symmetric matrix at
 * A + a_off[i], eigenvalues ascending to eig + e_off[i], eigenvector rows to
 * vec_rows + v_off[i] (n[i] <= 90; n[i] = 0 is skipped).  Waits for the
 * stream; TN_ENOCONV names how many did not converge. */
int tn_eigh_f64_batched(int64_t batch, const int64_t* n, const double* A, const int64_t* a_off, double* eig,
                        const int64_t* e_off, double* vec_rows, const int64_t* v_off, void* stream);

/* lanczosEV (uni10_lapack_cpu.cpp:416-509; the reference's GPU entry
 * uni10_lapack_gpu.cu:681-687 only throws): lowest eigenpair of the n x n
 * row-major matrix A by the plain Lanczos recurrence, every vector operation
 * on the device.  psi (n doubles, device) is the start vector on entry and
 * the ground state sum_k z_k v_k on return (sign: <start|psi> >= 0).
 * *iters: the maximum step count on entry (must exceed 2), the steps taken
 * on return.  Stops when the lowest Ritz value changes by at most err_tol
 * relative to max(|E0|, 1), when the Krylov space closes (beta <= 1e-15),
 * or at the maximum; *converged = 0 only in the last case.  A space that
 * closes after one step reports *E0 = 0 and leaves psi as given, as the
 * reference does.  Synchronises the stream. */
int tn_lanczos_f64(const double* A, int64_t n, double* psi, int64_t* iters, double err_tol,
                   double* E0, int* converged, void* stream);
/* The complex form (lanczosEV on std::complex<double>,
 * uni10_lapack_cpu.cpp:905-1000): A (n x n) and psi (n) complex, interleaved
 * (re, im) pairs, 16-byte aligned; alpha = Re(v^H w); same arguments,
 * stopping rule and sign convention (<start|psi> real and >= 0). */
int tn_lanczos_c128(const double* A, int64_t n, double* psi, int64_t* iters, double err_tol,
                    double* E0, int* converged, void* stream);

/* ---- TEBD truncation ------------------------------------------------------
 * The truncate-and-renormalise of iTEBD.cpp:157-161 plus the discarded
 * weight: given S (k, descending) on the device, keep dim = min(chi, k)
 * values, write lambda = S[:dim] / ||S[:dim]||, norm_out = ||S[:dim]|| and
 * trunc_err_out = sum_{i>=dim} S_i^2 (all device doubles). */
int tn_truncate_f64(const double* S, int64_t k, int64_t chi, double* lambda_out,
                    double* norm_out, double* trunc_err_out, void* stream);

/* ---- batched vector operations ---------------------------------------------
 * Many independent short vector operations (one item per chain and step of
 * a batched TEBD update) in one launch; all offsets are in doubles from one
 * arena base pointer.  Items of one plan must not read what another item of
 * the same plan writes.  Executing a plan neither allocates nor
 * synchronises (graph-capturable).  Batched forms of diagRowMul
 * (uni10_lapack_cpu.cpp:46-57), the truncation of iTEBD.cpp:157-161, the VT
 * row copy of iTEBD.cpp:162-166 and vectorScal (uni10_lapack_cpu.cpp:75-89).
 *   ROWSCALE   dst[i] = src[i] * base[aux + i / rep]          (i < n)
 *   COPY       dst[i] = src[i]                                (i < n)
 *   SCALE_INV  dst[i] = src[i] / base[aux]                    (i < n)
 *   TRUNCATE   src = S (n values, descending), keep rep of them:
 *              dst[i] = S[i] / ||S[:rep]|| (i < rep),
 *              base[aux] = ||S[:rep]||, base[aux+1] = sum_{i>=rep} S_i^2 */
#define TN_VECOP_ROWSCALE 0
#define TN_VECOP_COPY 1
#define TN_VECOP_SCALE_INV 2
#define TN_VECOP_TRUNCATE 3
typedef struct {
  int32_t op, pad;
  int64_t n, src, dst, aux, rep;
} tn_vecop;
typedef struct tn_vecops_plan_s* tn_vecops_plan;
int tn_vecops_plan_create(tn_vecops_plan* plan, int64_t nitems, const tn_vecop* items);
int tn_vecops_plan_exec(tn_vecops_plan plan, double* base, void* stream);
int tn_vecops_plan_set_mask(tn_vecops_plan plan, const int* mask, const int32_t* item_idx);
int tn_vecops_plan_destroy(tn_vecops_plan plan);

#ifdef __cplusplus
}
#endif
#endif /* TN_HIP_H */
