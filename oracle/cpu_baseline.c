/*
 * cpu_baseline.c -- the reported CPU baseline (TEST INFRASTRUCTURE ONLY).
 *
 * LibShalom's micro-kernels are AArch64 NEON assembly and cannot run on an
 * x86-64 host, so the baseline is the reference harnesses' C row-major
 * verification loop restated as a triple loop (the loop the zip prototype
 * SMM_m8n12_new_single.c and NN_LIB/benchmark.c:130-173 check against):
 * i-k-j order for op(B) = B (row of C updated by a*row of B), i-j-k dot
 * order for op(B) = B^T.  OpenMP over rows for one GEMM, over entries for a
 * batch.  Built with -O3 -fopenmp -march=x86-64-v3.
 */
#include "oracle.h"
#include <omp.h>
#include <string.h>

#define DEFINE_GEMM(NAME, T)                                                         \
    static void NAME##_rows(int transa, int transb, T *C, const T *A, const T *B,    \
                            long M, long N, long K, long i0, long i1)                \
    {                                                                                \
        for (long i = i0; i < i1; ++i) {                                             \
            T *c = C + i * N;                                                        \
            if (!transb) {                                                           \
                for (long j = 0; j < N; ++j) c[j] = (T)0;                            \
                for (long k = 0; k < K; ++k) {                                       \
                    T a = transa ? A[k * M + i] : A[i * K + k];                      \
                    const T *b = B + k * N;                                          \
                    for (long j = 0; j < N; ++j) c[j] += a * b[j];                   \
                }                                                                    \
            } else {                                                                 \
                for (long j = 0; j < N; ++j) {                                       \
                    const T *b = B + j * K;                                          \
                    T s = (T)0;                                                      \
                    if (!transa) {                                                   \
                        const T *a = A + i * K;                                      \
                        for (long k = 0; k < K; ++k) s += a[k] * b[k];               \
                    } else {                                                         \
                        for (long k = 0; k < K; ++k) s += A[k * M + i] * b[k];       \
                    }                                                                \
                    c[j] = s;                                                        \
                }                                                                    \
            }                                                                        \
        }                                                                            \
    }

DEFINE_GEMM(sg, float)
DEFINE_GEMM(dg, double)

void cpu_sgemm(int transa, int transb, float *C, const float *A, const float *B,
               long M, long N, long K)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < M; ++i) sg_rows(transa, transb, C, A, B, M, N, K, i, i + 1);
}

void cpu_dgemm(int transa, int transb, double *C, const double *A, const double *B,
               long M, long N, long K)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < M; ++i) dg_rows(transa, transb, C, A, B, M, N, K, i, i + 1);
}

void cpu_sgemm_batched(int transa, int transb, long count, float *const *C,
                       const float *const *A, const float *const *B,
                       const long *M, const long *N, const long *K)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (long g = 0; g < count; ++g)
        sg_rows(transa, transb, C[g], A[g], B[g], M[g], N[g], K[g], 0, M[g]);
}

void cpu_dgemm_batched(int transa, int transb, long count, double *const *C,
                       const double *const *A, const double *const *B,
                       const long *M, const long *N, const long *K)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (long g = 0; g < count; ++g)
        dg_rows(transa, transb, C[g], A[g], B[g], M[g], N[g], K[g], 0, M[g]);
}

int cpu_num_threads(void) { return omp_get_max_threads(); }
