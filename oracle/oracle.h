/*
 * oracle.h -- CPU restatement of LibShalom's small/irregular GEMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it, and only as the checker (or, for
 * the CPU baseline, as the timed CPU loop).
 *
 * What each part restates (reference paths are relative to the LibShalom
 * source tree):
 *   - oracle_drand48 / oracle_random_matrix: the harness input recipe
 *     `a[i*n+j] = 2.0 * (float)drand48() - 1.0` with glibc's unseeded
 *     drand48 state X = 0 (benchmark/small_SGEMM/LibShalom_sgemm.c:26-34; no
 *     srand48 call anywhere in the reference, which runs on Linux).  Pinned against glibc's
 *     drand48 by tests/golden/drand48_first.json.
 *   - oracle_hello_fill: Hello.c:20-29 (B filled as B[i*K+j], i<K, j<N).
 *   - oracle_gemm_ref: C = op(A) * op(B), beta = 0 overwrite, row-major,
 *     lda = K (NoTrans) / M (Trans), ldb = N (NoTrans) / K (Trans),
 *     ldc = N -- the semantics of LibShalom_sgemm / _sgemm_mp / _dgemm
 *     (NN_SMM.c:5-38, SMM_thread.c:75-114; overwrite as in
 *     test_temp_L1.c:8-70; NT layout B[N][K] as in test_SMM_NT.c:2228-2250).
 *     Accumulates in long double and also returns (|A||B|)_ij for the
 *     elementwise tolerance |C - C_ref| <= 2*K*u*(|A||B|)_ij.
 *     Pinned by closed forms (all-ones KAT, evaluate/test.c:38-46),
 *     transpose identities and exact small-integer products in
 *     tests/test_oracle.py.
 *   - oracle_sconv2d_ref: the convolution whose im2col GEMM the reference's
 *     CNN benchmark times (benchmark/CNN_kernels/LibShalom_sgemm_mp.c:86-116,
 *     M = Cout, N = H*W, K = Cin*9); the reference stores the im2col matrix,
 *     the implicit-GEMM path gathers it.  Pinned against torch's float64
 *     conv2d, against torch's unfold (im2col) followed by oracle_sgemm_ref,
 *     and by closed forms (a one-tap filter shifts the input, all-ones
 *     operands count the taps inside the image) in tests/test_oracle.py.
 *   - oracle_sconv2d_bwd_ref: its gradients; pinned against torch's float64
 *     autograd of conv2d and by the adjoint identities
 *     <conv(X, W), dY> = <X, dX> = <W, dW> in tests/test_oracle.py.
 *   - cpu_sgemm / cpu_dgemm (+ _batched): the reference's C row-major
 *     verification loop as a triple loop (i-k-j for NN, i-j-k dot for NT),
 *     OpenMP over rows or over batch entries.  Used only as the reported
 *     CPU baseline; the NEON assembly cannot run on x86-64.
 *
 * Not restated: the AArch64 micro-kernels' summation order (bit-level
 * parity with the ARMv8 build is unpinnable here: no AArch64 toolchain or
 * qemu), and the OpenMP thread-grid planner Dete_grad_* (SMM_thread.c:14-72),
 * which changes work placement but never results.
 */
#ifndef LIBSHALOM_ORACLE_H
#define LIBSHALOM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* drand48 LCG state: X_{n+1} = (a X_n + c) mod 2^48 (POSIX a, c).  The
 * default state is glibc's unseeded one, X = 0. */
void   oracle_seed_default(uint64_t *state);
double oracle_drand48(uint64_t *state);
void   oracle_drand48_fill(uint64_t *state, double *out, long n);
void   oracle_random_matrix(uint64_t *state, long m, long n, float *a);
void   oracle_random_matrix_d(uint64_t *state, long m, long n, double *a);
void   oracle_hello_fill(uint64_t *state, long M, long N, long K, float *A, float *B);

/* Reference GEMM: C_ref = op(A) op(B) in long double, absAB = |op(A)||op(B)|. */
void oracle_sgemm_ref(int transa, int transb, const float *A, const float *B,
                      double *Cref, double *absAB, long M, long N, long K);
void oracle_dgemm_ref(int transa, int transb, const double *A, const double *B,
                      double *Cref, double *absAB, long M, long N, long K);

/* Many GEMMs stored back to back in flat arrays (A of GEMM g at offA[g], ...);
 * Kelem receives K[g] for every element of C_g so one check covers the batch. */
void oracle_sgemm_ref_flat(int transa, int transb, long count, const float *A, const float *B,
                           const long *offA, const long *offB, const long *offC, const long *M,
                           const long *N, const long *K, double *Cref, double *absAB, double *Kelem);
void oracle_dgemm_ref_flat(int transa, int transb, long count, const double *A, const double *B,
                           const long *offA, const long *offB, const long *offC, const long *M,
                           const long *N, const long *K, double *Cref, double *absAB, double *Kelem);

/* 2-D convolution (NCHW input, OIHW filter, stride / padding / dilation per
 * axis), the GEMM of the filter and each image's im2col matrix; long double
 * sums, absWX = sum |W||X| over the taps inside the image. */
void oracle_sconv2d_ref(const float *X, const float *Wt, double *Yref, double *absWX, long batch, long C, long H,
                        long W, long Cout, long R, long S, long sh, long sw, long ph, long pw, long dh, long dw,
                        long P, long Q);
/* Its backward for the loss sum(Y * dY): dX (like X) and dW (like Wt), each
 * forward product scattered to the operand it came from; long double sums
 * and the sums of |.| products for the bound. */
void oracle_sconv2d_bwd_ref(const float *X, const float *Wt, const float *dY, double *dXref, double *absdX,
                            double *dWref, double *absdW, long batch, long C, long H, long W, long Cout, long R,
                            long S, long sh, long sw, long ph, long pw, long dh, long dw, long P, long Q);
/* Elementwise check against the reference.
 * stats[0] = max |C - Cref| / (2*K*u*absAB)   (<= 1 means within tolerance)
 * stats[1] = max |C - Cref| / absAB            (componentwise relative error)
 * stats[2] = mean |C - Cref| / absAB
 * stats[3] = max |C - Cref|
 * Elements with absAB == 0 must match exactly (they count as a violation
 * otherwise).  NaN in C counts as a violation.  Returns the violation count. */
long oracle_check_f32(const float *C, const double *Cref, const double *absAB,
                      long n, long K, double u, double *stats);
long oracle_check_f64(const double *C, const double *Cref, const double *absAB,
                      long n, long K, double u, double *stats);
/* Same with a per-element K (from oracle_*_ref_flat) and an extra per-product
 * term: bound = (2*K*u + extra) * absAB. */
long oracle_check_kv(const void *C, int is_f64, const double *Cref, const double *absAB,
                     const double *Kelem, long n, double u, double extra, double *stats);

/* CPU baseline (reported, not targeted). */
void cpu_sgemm(int transa, int transb, float *C, const float *A, const float *B,
               long M, long N, long K);
void cpu_dgemm(int transa, int transb, double *C, const double *A, const double *B,
               long M, long N, long K);
void cpu_sgemm_batched(int transa, int transb, long count, float *const *C,
                       const float *const *A, const float *const *B,
                       const long *M, const long *N, const long *K);
void cpu_dgemm_batched(int transa, int transb, long count, double *const *C,
                       const double *const *A, const double *const *B,
                       const long *M, const long *N, const long *K);
int  cpu_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
