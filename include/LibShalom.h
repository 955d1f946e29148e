/*
 * LibShalom.h -- MI355X (gfx950) implementation of LibShalom's small and
 * irregular-shaped GEMM API.
 *
 * Drop-in for the reference header NN_LIB/LibShalom.h with these changes:
 *   - every symbol is extern "C" (the reference's are C++-mangled because
 *     its header is C++, NN_LIB/LibShalom.h:5-7);
 *   - A, B and C may be device pointers (hipMalloc / torch tensors / managed
 *     memory) on the current HIP device, or host memory as in the
 *     reference (the reference-API calls stage it; see below);
 *   - transa is honoured (the reference ignores it, NN_SMM.c:8,
 *     SMM_thread.c:81) and dgemm supports Trans (a no-op in the reference,
 *     NN_SMM.c:33-37);
 *   - only the public entry points are exported; the reference's internal
 *     kernels and pack routines (SGEMM_NN, SGEMM_NT, Small_*, NPACK, TPACK,
 *     Dete_grad_*) and the GEMM_K/N/M blocking macros have no counterpart
 *     because the packing and partition are done inside the GPU kernels.
 *
 * Storage is row-major with the reference's implicit leading dimensions:
 * op(A) is M x K (A stored M x K, lda = K; or Trans: K x M, lda = M),
 * op(B) is K x N (B stored K x N, ldb = N; or Trans: N x K, ldb = K),
 * C is M x N (ldc = N).  C = op(A) * op(B); C is overwritten (beta = 0).
 */
#ifndef LIBSHALOM_AMD_H
#define LIBSHALOM_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* NN_LIB/LibShalom.h:10-12 */
#define NoTrans 0
#define Trans 1

/* A hipStream_t.  NULL is the null (default) stream. */
typedef void *LibShalom_stream_t;

/* Status codes (the reference has no error reporting; its functions return
 * void and silently skip unsupported cases). */
#define LIBSHALOM_SUCCESS 0
#define LIBSHALOM_INVALID_VALUE 1   /* negative size, bad trans / precision flag */
#define LIBSHALOM_INVALID_POINTER 2 /* NULL pointer for a non-empty operand */
#define LIBSHALOM_HIP_ERROR 3       /* a HIP runtime call failed */
#define LIBSHALOM_ALLOC_FAILED 4    /* workspace allocation failed */
#define LIBSHALOM_NOT_SUPPORTED 5   /* size beyond the supported range (dims < 2^31) */
#define LIBSHALOM_NO_DEVICE 6       /* no usable gfx950 device */

/* Arithmetic of the float GEMMs.  FP32: FP32 MFMA.  BF16X3: each FP32
 * operand split as hi + lo (bf16), three BF16 MFMA products hi*hi + hi*lo +
 * lo*hi, FP32 accumulation.  BF16: one BF16 product.  BF16X6: three-way
 * split (exact for every FP32 value), the eight products of the pieces but
 * lo2*lo2, FP32 accumulation: within |C - C64| <= 2*K*2^-24*(|A||B|), as
 * FP32.  In the split modes every result element that comes out Inf or NaN
 * (an Inf or NaN operand, or a finite one beyond the largest bf16) is
 * recomputed in FP32 by the same launch, so special values give the
 * classes (finite, +-Inf, NaN) of the FP32 mode.  FP64 GEMMs always use
 * FP64 MFMA. */
#define LIBSHALOM_FP32 0
#define LIBSHALOM_BF16X3 1
#define LIBSHALOM_BF16 2
#define LIBSHALOM_BF16X6 3
#define LIBSHALOM_FP64 4 /* id of the DGEMM path in LibShalom_query_plan */

/* ------------------------------------------------------------------------
 * Reference API (synchronous: C is complete when the call returns).  Errors
 * leave C untouched and are reported by LibShalom_get_last_error().
 * Operands may be device pointers or, as in the reference, host memory: host
 * operands are staged through a device buffer (inputs copied in, C copied
 * back, through pinned bounce buffers with the copies overlapped, one
 * synchronisation per call), so the reference's harnesses run unchanged.
 * Calls from several host threads are safe.  The other entry points take
 * device-accessible pointers only.
 * ------------------------------------------------------------------------ */

/* Replaces LibShalom_sgemm (NN_LIB/LibShalom.h:22, NN_SMM.c:5-27). */
void LibShalom_sgemm(int transa, int transb, float *C, float *A, float *B, long M, long N, long K);

/* Replaces LibShalom_sgemm_mp (NN_LIB/LibShalom.h:27, SMM_thread.c:75-114):
 * the irregular-shape path (tall-skinny / short-wide); the OpenMP Tm x Tn
 * thread grid becomes a workgroup schedule with split-K, or, for a short
 * side of at most 64 and a long K, workgroups of 16 long-side indices that
 * split K between their waves (one launch, no scratch).  LibShalom_sgemm
 * takes the same plans: on the GPU every call spreads over the whole
 * device, so the reference's one-core / thread-grid distinction falls away. */
void LibShalom_sgemm_mp(int transa, int transb, float *C, float *A, float *B, long M, long N, long K);

/* Replaces LibShalom_dgemm (NN_LIB/LibShalom.h:39, NN_SMM.c:30-38). */
void LibShalom_dgemm(int transa, int transb, double *C, double *A, double *B, long M, long N, long K);

/* Replace LibShalom_set_thread_nums / _set_thread_ways (NN_LIB/LibShalom.h:41-42,
 * SMM_thread.c:116-137).  No-ops: the GPU schedule does not use a CPU thread
 * count.  The values are recorded and can be read back. */
void LibShalom_set_thread_nums(int num);
void LibShalom_set_thread_ways(int tm, int tn);
void LibShalom_get_thread_config(int *num, int *tm, int *tn);

/* Precision used by the reference-API float calls above (default FP32).
 * Returns LIBSHALOM_INVALID_VALUE for an unknown precision. */
int LibShalom_set_precision(int precision);
int LibShalom_get_precision(void);

/* Status of the last LibShalom call made by this host thread. */
int LibShalom_get_last_error(void);
const char *LibShalom_status_string(int status);

/* ------------------------------------------------------------------------
 * Extended API (no reference counterpart): explicit precision and stream,
 * asynchronous on `stream`, returns a status.
 * ------------------------------------------------------------------------ */
int LibShalom_sgemm_ex(int transa, int transb, float *C, const float *A, const float *B,
                       long M, long N, long K, int precision, LibShalom_stream_t stream);
int LibShalom_sgemm_mp_ex(int transa, int transb, float *C, const float *A, const float *B,
                          long M, long N, long K, int precision, LibShalom_stream_t stream);
int LibShalom_dgemm_ex(int transa, int transb, double *C, const double *A, const double *B,
                       long M, long N, long K, LibShalom_stream_t stream);

/* BLAS-style general GEMM (no reference counterpart; the reference's API has
 * implicit leading dimensions and beta = 0, NN_LIB/LibShalom.h:22-42):
 *   C = alpha * op(A) op(B) + beta * C, row-major,
 * A is M x K (lda >= K) or, with transa, K x M (lda >= M); B is K x N
 * (ldb >= N) or, with transb, N x K (ldb >= K); C is M x N (ldc >= N).  C is
 * not read when beta is 0 (NaN or Inf in C do not propagate then), and A
 * and B are not read when alpha is 0 (they may then be NULL).  Leading
 * dimensions below the minimum return LIBSHALOM_INVALID_VALUE; at or above
 * 2^21 elements, LIBSHALOM_NOT_SUPPORTED.  Asynchronous on `stream`.
 *
 * Streams and capture (every asynchronous entry point): device buffers the
 * library keeps per stream (split-K partial tiles and arrival counters, a
 * convolution's flipped filter) grow on demand and an outgrown buffer is
 * kept, not freed, until the process ends, so graphs that recorded it stay
 * valid.  A call under stream capture that would need a buffer to grow
 * returns LIBSHALOM_NOT_SUPPORTED; one eager call of that size on the
 * stream beforehand sizes it. */
int LibShalom_sgemm_general(int transa, int transb, long M, long N, long K, float alpha, const float *A,
                            long lda, const float *B, long ldb, float beta, float *C, long ldc,
                            int precision, LibShalom_stream_t stream);
int LibShalom_dgemm_general(int transa, int transb, long M, long N, long K, double alpha, const double *A,
                            long lda, const double *B, long ldb, double beta, double *C, long ldc,
                            LibShalom_stream_t stream);

/* LibShalom_sgemm_general, allocating nothing (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Same arguments, results and errors.  Its split-K scratch (arrival counters
 * and partial tiles) is a static device array of 64 KiB + 32 MiB per GPU,
 * held until the process ends; the planner picks fewer K slices (or none)
 * where more would not fit, and a plan needing more returns
 * LIBSHALOM_NOT_SUPPORTED.  The first call on a device takes the array's
 * address and zeroes its counters (a fill kernel and a wait on the null
 * stream: make that first call outside a stream capture).  Calls that
 * overlap on one GPU share the array and race: serialize them. */
int LibShalom_sgemm_general_noalloc(int transa, int transb, long M, long N, long K, float alpha,
                                    const float *A, long lda, const float *B, long ldb, float beta, float *C,
                                    long ldc, int precision, LibShalom_stream_t stream);

/* Batched GEMMs of uneven sizes (no reference counterpart; the reference
 * makes one call per GEMM, e.g. benchmark/CP2K_kernels/LibShalom_dgemm.c:102-106).
 * A, B, C, M, N, K are HOST arrays of `count` entries (device pointers and
 * sizes); one transa / transb for the whole batch.
 *
 * A plan buckets the GEMMs by tile shape, uploads the work descriptors once
 * and captures the launches in a hipGraph; LibShalom_plan_execute replays it
 * on `stream`.  The pointers are baked into the plan. */
typedef struct LibShalom_plan_s *LibShalom_plan_t;
int LibShalom_sgemm_batched_plan(LibShalom_plan_t *plan, int transa, int transb, float *const *C,
                                 const float *const *A, const float *const *B, const long *M,
                                 const long *N, const long *K, long count, int precision);
int LibShalom_dgemm_batched_plan(LibShalom_plan_t *plan, int transa, int transb, double *const *C,
                                 const double *const *A, const double *const *B, const long *M,
                                 const long *N, const long *K, long count);
int LibShalom_plan_execute(LibShalom_plan_t plan, LibShalom_stream_t stream);
int LibShalom_plan_destroy(LibShalom_plan_t plan);
/* flops = sum 2*M*N*K; bytes = sum s*(M*K + K*N + M*N) (algorithmic). */
int LibShalom_plan_info(LibShalom_plan_t plan, double *flops, double *bytes, long *work_items);

/* One-shot batched calls: plan, execute on `stream`, wait, destroy. */
int LibShalom_sgemm_batched(int transa, int transb, float *const *C, const float *const *A,
                            const float *const *B, const long *M, const long *N, const long *K,
                            long count, int precision, LibShalom_stream_t stream);
int LibShalom_dgemm_batched(int transa, int transb, double *const *C, const double *const *A,
                            const double *const *B, const long *M, const long *N, const long *K,
                            long count, LibShalom_stream_t stream);

/* Strided batch of GEMMs in one launch: C_i = op(A_i) op(B_i) with
 * X_i = X + i*strideX.  strideA or strideB 0: the operand is shared by the
 * batch (e.g. the weights of an im2col convolution); strideC 0: C is the sum
 * over the batch.  Long reductions are cut into slices whose partial tiles
 * are summed inside the launch in a fixed order.  FP32-input modes only.  No
 * reference counterpart (the reference has no batched entry point). */
int LibShalom_sgemm_strided_batched(int transa, int transb, float *C, const float *A,
                                    const float *B, long M, long N, long K, long batch,
                                    long strideA, long strideB, long strideC, int precision,
                                    LibShalom_stream_t stream);

/* Backward of C = op(A) op(B) given dC (M x N): writes dA with the storage
 * shape of A and dB with the storage shape of B,
 *   op(A)' = dC op(B)^T,   op(B)' = op(A)^T dC,
 * both products in ONE kernel launch (no reference counterpart: the
 * reference has no differentiation).  Asynchronous on `stream`. */
int LibShalom_sgemm_backward(int transa, int transb, const float *dC, const float *A,
                             const float *B, float *dA, float *dB, long M, long N, long K,
                             int precision, LibShalom_stream_t stream);

/* The same for a strided batch: image i uses A + i*strideA, B + i*strideB,
 * dC + i*strideC and writes dA + i*strideA, dB + i*strideB.  A stride of 0
 * for A (or B) means the operand is shared by the batch -- the weights of an
 * im2col convolution -- and its gradient is the sum over the images, written
 * once.  One launch; split reductions are summed inside it in a fixed order,
 * so results do not depend on scheduling.  No reference counterpart. */
int LibShalom_sgemm_backward_strided(int transa, int transb, const float *dC, const float *A,
                                     const float *B, float *dA, float *dB, long M, long N, long K,
                                     long batch, long strideA, long strideB, long strideC,
                                     int precision, LibShalom_stream_t stream);

/* 2-D convolution forward as an implicit GEMM (no reference counterpart: the
 * reference's CNN benchmark hands LibShalom_sgemm_mp the GEMMs of im2col
 * convolutions, M = Cout, N = H*W, K = Cin*9 --
 * benchmark/CNN_kernels/LibShalom_sgemm_mp.c:86-116 -- with the im2col
 * matrix already stored).  Y[b] (Cout x P x Q) = W (Cout x C*R*S) times the
 * im2col matrix of X[b] (C*R*S x P*Q), which is never stored: the operand
 * loads gather it from X.  X is NCHW (batch x C x H x W), W is OIHW
 * (Cout x C x R x S), Y is batch x Cout x P x Q with
 * P = (H + 2 pad_h - dil_h (R - 1) - 1) / stride_h + 1 (Q likewise).  The
 * precision flags of the float GEMMs apply.  One launch; long reductions are
 * cut into slices of whole channels and summed inside it in a fixed order.
 * Asynchronous on `stream`.  LIBSHALOM_NOT_SUPPORTED when one image
 * (C*H*W floats) reaches 2 GiB, P*Q*Q or (C*R*S + R*S)*R*S reaches 2^32,
 * H + 2 pad_h (W + 2 pad_w) reaches 2^29, or C*R*S or P*Q reaches 2^21. */
int LibShalom_sconv2d_forward(float *Y, const float *W, const float *X, long batch, long C, long H, long Wd,
                              long Cout, long R, long S, long stride_h, long stride_w, long pad_h, long pad_w,
                              long dil_h, long dil_w, int precision, LibShalom_stream_t stream);
/* Backward of LibShalom_sconv2d_forward, both gradients in ONE launch, the
 * im2col matrix again never stored: dX (batch x C x H x W) from dY
 * (batch x Cout x P x Q) and W, and dW (Cout x C x R x S) = the sum over the
 * batch, from dY and X.  Either output may be NULL to skip it.  dX is the
 * convolution of dY with the flipped, transposed filter (written by the
 * library into its per-stream workspace), fractionally strided for a
 * strided convolution, with strides up to 8 (LIBSHALOM_NOT_SUPPORTED
 * otherwise; dW takes any geometry).  Split reductions
 * are summed inside the launch in a fixed order.  Limits as for the forward,
 * with H + 2 pad_h (W + 2 pad_w) below 2^28, one dY image below 2 GiB, and
 * Cout*R*S and H*W also below 2^21. */
int LibShalom_sconv2d_backward(float *dX, float *dW, const float *dY, const float *W, const float *X, long batch,
                               long C, long H, long Wd, long Cout, long R, long S, long stride_h, long stride_w,
                               long pad_h, long pad_w, long dil_h, long dil_w, int precision,
                               LibShalom_stream_t stream);
/* P and Q of that convolution (LIBSHALOM_INVALID_VALUE for a bad geometry;
 * sizes <= 0 mean the dilated filter does not fit the padded input). */
int LibShalom_conv2d_output_size(long H, long Wd, long R, long S, long stride_h, long stride_w, long pad_h,
                                 long pad_w, long dil_h, long dil_w, long *P, long *Q);

/* Autotuner for batched plans (no reference counterpart).  It runs the given
 * batch of shapes on the device (synthetic operands in a scratch buffer; the
 * caller's memory is not touched), groups the GEMMs by max(M, N) and K
 * ranges, and for each group in order of work tries every other candidate
 * square tile (32, 64, 128; FP64: 32, 64), keeping it when the whole batch
 * runs faster.  The chosen tile is recorded per (precision, transa, transb,
 * M/N/K class); plans built afterwards use it.  Synchronous on `stream`;
 * `classes_tuned` receives the number of classes recorded.
 * LIBSHALOM_TUNE_FILE names a text file the table is read from at first use
 * and appended to.  query_tuned_tile reports the tile a plan would use
 * (recorded, or the fixed rule); autotune_reset forgets the table. */
int LibShalom_autotune_batched(int precision, int transa, int transb, const long *M, const long *N,
                               const long *K, long count, int *classes_tuned,
                               LibShalom_stream_t stream);
int LibShalom_autotune_reset(void);
int LibShalom_query_tuned_tile(int precision, int transa, int transb, long M, long N, long K,
                               int *tile_m, int *tile_n);

/* Planner introspection (host only, no device needed): the workgroup tile
 * (tile_m x tile_n of C), K-split count and K slice length the library would
 * use for one GEMM on a device with num_cu compute units (mp != 0: the
 * LibShalom_sgemm_mp path), and the tile a batched plan gives one GEMM
 * (FP64 plans run the GEMMs of the 32 x 32 bucket with M, N <= 16 on
 * one-wave 16 x 16 tiles; both tile queries report that).  A GEMM with a
 * short side (N <= 64 and M >= 1024, or M <= 64 and N >= 1024; in the
 * split-BF16 modes a short side of at most 32) and a moderate K may run on
 * the short-side kernel instead, 16 long-side indices per workgroup over
 * all of K: reported as a 16 x 16/32/64 (or 16/32/64 x 16) tile and no
 * split.
 * Replaces the reference's thread-grid planner Dete_grad_N_threads_nums /
 * Dete_grad_M_threads_nums (SMM_thread.c:14-72). */
int LibShalom_query_plan(int precision, int mp, long M, long N, long K, int num_cu, int *tile_m,
                         int *tile_n, int *ksplit, int *kper);
int LibShalom_query_batched_tile(int precision, long M, long N, long K, int *tile_m, int *tile_n);

#ifdef __cplusplus
}
#endif
#endif /* LIBSHALOM_AMD_H */
