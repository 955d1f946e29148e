/*
 * oracle.c -- CPU restatement used as the checker for LibShalom's GEMM path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h for what each function restates
 * and how it is pinned).  Built with -ffp-contract=off so the reference sums
 * do not depend on the compiler's FMA contraction choices.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define LCG_A 0x5DEECE66DULL
#define LCG_C 0xBULL
#define LCG_MASK ((1ULL << 48) - 1)

/* Unseeded drand48 state on the reference's platform (Linux / glibc): glibc
 * zero-initialises the 48-bit X and only fills in a and c on first use, so
 * the first value is 11 / 2^48.  (0x1234ABCD330E is the BSD libc default;
 * LibShalom's harnesses never call srand48 and run on Linux --
 * benchmark/small_SGEMM/LibShalom_sgemm.c:26-34, NN_LIB/run.sh.) */
void oracle_seed_default(uint64_t *state) { *state = 0ULL; }

/* X_{n+1} = (a X_n + c) mod 2^48; drand48 returns X_{n+1} / 2^48. */
double oracle_drand48(uint64_t *state)
{
    uint64_t x = (*state * LCG_A + LCG_C) & LCG_MASK;
    *state = x;
    return ldexp((double)x, -48);
}

void oracle_drand48_fill(uint64_t *state, double *out, long n)
{
    for (long i = 0; i < n; ++i) out[i] = oracle_drand48(state);
}

/* `a[i*n+j] = 2.0 * (float)drand48() - 1.0;` with a float destination:
 * the double is rounded to float first, the affine map is done in double,
 * then rounded to float again (benchmark/small_SGEMM/LibShalom_sgemm.c:33). */
static inline float recipe_f32(uint64_t *state)
{
    float f = (float)oracle_drand48(state);
    return (float)(2.0 * (double)f - 1.0);
}

void oracle_random_matrix(uint64_t *state, long m, long n, float *a)
{
    for (long i = 0; i < m; ++i)
        for (long j = 0; j < n; ++j) a[i * n + j] = recipe_f32(state);
}

/* FP64 inputs: the same recipe with a double destination (the CP2K DGEMM
 * harness leaves its inputs uninitialised, CP2K_kernels/LibShalom_dgemm.c:97-99,
 * so the seeded drand48 stream stands in for it). */
void oracle_random_matrix_d(uint64_t *state, long m, long n, double *a)
{
    for (long i = 0; i < m; ++i)
        for (long j = 0; j < n; ++j) a[i * n + j] = 2.0 * oracle_drand48(state) - 1.0;
}

/* Hello.c:20-29: A row-major M x K, then B filled as B[i*K + j] for i < K,
 * j < N (identical to row-major K x N only when N == K). */
void oracle_hello_fill(uint64_t *state, long M, long N, long K, float *A, float *B)
{
    for (long i = 0; i < M; ++i)
        for (long j = 0; j < K; ++j) A[i * K + j] = recipe_f32(state);
    for (long i = 0; i < K; ++i)
        for (long j = 0; j < N; ++j) B[i * K + j] = recipe_f32(state);
}

/* op(A)[i][k] and op(B)[k][j] for row-major storage with the reference's
 * implicit leading dimensions. */
#define OPA(i, k) (transa ? A[(k) * M + (i)] : A[(i) * K + (k)])
#define OPB(k, j) (transb ? B[(j) * K + (k)] : B[(k) * N + (j)])

void oracle_sgemm_ref(int transa, int transb, const float *A, const float *B,
                      double *Cref, double *absAB, long M, long N, long K)
{
    for (long i = 0; i < M; ++i)
        for (long j = 0; j < N; ++j) {
            long double s = 0.0L, sa = 0.0L;
            for (long k = 0; k < K; ++k) {
                long double a = OPA(i, k), b = OPB(k, j);
                s += a * b;
                sa += fabsl(a) * fabsl(b);
            }
            Cref[i * N + j] = (double)s;
            absAB[i * N + j] = (double)sa;
        }
}

void oracle_dgemm_ref(int transa, int transb, const double *A, const double *B,
                      double *Cref, double *absAB, long M, long N, long K)
{
    for (long i = 0; i < M; ++i)
        for (long j = 0; j < N; ++j) {
            long double s = 0.0L, sa = 0.0L;
            for (long k = 0; k < K; ++k) {
                long double a = OPA(i, k), b = OPB(k, j);
                s += a * b;
                sa += fabsl(a) * fabsl(b);
            }
            Cref[i * N + j] = (double)s;
            absAB[i * N + j] = (double)sa;
        }
}

#undef OPA
#undef OPB

void oracle_sgemm_ref_flat(int transa, int transb, long count, const float *A, const float *B,
                           const long *offA, const long *offB, const long *offC, const long *M,
                           const long *N, const long *K, double *Cref, double *absAB, double *Kelem)
{
    for (long g = 0; g < count; ++g) {
        oracle_sgemm_ref(transa, transb, A + offA[g], B + offB[g], Cref + offC[g], absAB + offC[g],
                         M[g], N[g], K[g]);
        for (long i = 0; i < M[g] * N[g]; ++i) Kelem[offC[g] + i] = (double)K[g];
    }
}

void oracle_dgemm_ref_flat(int transa, int transb, long count, const double *A, const double *B,
                           const long *offA, const long *offB, const long *offC, const long *M,
                           const long *N, const long *K, double *Cref, double *absAB, double *Kelem)
{
    for (long g = 0; g < count; ++g) {
        oracle_dgemm_ref(transa, transb, A + offA[g], B + offB[g], Cref + offC[g], absAB + offC[g],
                         M[g], N[g], K[g]);
        for (long i = 0; i < M[g] * N[g]; ++i) Kelem[offC[g] + i] = (double)K[g];
    }
}

/* 2-D convolution, NCHW input X (batch x C x H x W), OIHW filter Wt
 * (Cout x C x R x S), output batch x Cout x P x Q:
 *   Y[b][o][p][q] = sum_{c,r,s} Wt[o][c][r][s] * X[b][c][p*sh - ph + r*dh][q*sw - pw + s*dw]
 * with taps in the padding contributing 0.  This is the product of the
 * filter (Cout x C*R*S) and the im2col matrix of each image that the
 * reference's CNN benchmark hands to LibShalom_sgemm_mp
 * (benchmark/CNN_kernels/LibShalom_sgemm_mp.c:86-116: M = Cout, N = H*W,
 * K = Cin*9 for 3 x 3, stride 1, padding 1).  Long double sums; absWX gets
 * the sum of |Wt||X| for the elementwise bound, K = C*R*S. */
void oracle_sconv2d_ref(const float *X, const float *Wt, double *Yref, double *absWX, long batch, long C, long H,
                        long W, long Cout, long R, long S, long sh, long sw, long ph, long pw, long dh, long dw,
                        long P, long Q)
{
    for (long b = 0; b < batch; ++b)
        for (long o = 0; o < Cout; ++o)
            for (long p = 0; p < P; ++p)
                for (long q = 0; q < Q; ++q) {
                    long double acc = 0.0L, aa = 0.0L;
                    for (long c = 0; c < C; ++c)
                        for (long r = 0; r < R; ++r) {
                            const long ih = p * sh - ph + r * dh;
                            if (ih < 0 || ih >= H) continue;
                            for (long s = 0; s < S; ++s) {
                                const long iw = q * sw - pw + s * dw;
                                if (iw < 0 || iw >= W) continue;
                                const long double w = Wt[((o * C + c) * R + r) * S + s];
                                const long double x = X[((b * C + c) * H + ih) * W + iw];
                                acc += w * x;
                                aa += fabsl(w) * fabsl(x);
                            }
                        }
                    const long i = ((b * Cout + o) * P + p) * Q + q;
                    Yref[i] = (double)acc;
                    absWX[i] = (double)aa;
                }
}

/* Backward of oracle_sconv2d_ref for the loss sum(Y * dY):
 *   dX[b][c][ih][iw] = sum over (o, r, s, p, q) that read X[b][c][ih][iw] of Wt[o][c][r][s] * dY[b][o][p][q]
 *   dW[o][c][r][s]   = sum over (b, p, q) of dY[b][o][p][q] * X[b][c][ih][iw]
 * (ih = p*sh - ph + r*dh, iw = q*sw - pw + s*dw inside the image): every
 * product of the forward sum, scattered to the operand it came from.  Long
 * double sums; absdX / absdW get the sums of |.| products. */
void oracle_sconv2d_bwd_ref(const float *X, const float *Wt, const float *dY, double *dXref, double *absdX,
                            double *dWref, double *absdW, long batch, long C, long H, long W, long Cout, long R,
                            long S, long sh, long sw, long ph, long pw, long dh, long dw, long P, long Q)
{
    const long nx = batch * C * H * W, nw = Cout * C * R * S;
    long double *gx = (long double *)calloc((size_t)(2 * nx + 2 * nw), sizeof(long double));
    if (!gx) return;
    long double *ax = gx + nx, *gw = ax + nx, *aw = gw + nw;
    for (long b = 0; b < batch; ++b)
        for (long o = 0; o < Cout; ++o)
            for (long p = 0; p < P; ++p)
                for (long q = 0; q < Q; ++q) {
                    const long double g = dY[((b * Cout + o) * P + p) * Q + q];
                    for (long c = 0; c < C; ++c)
                        for (long r = 0; r < R; ++r) {
                            const long ih = p * sh - ph + r * dh;
                            if (ih < 0 || ih >= H) continue;
                            for (long s = 0; s < S; ++s) {
                                const long iw = q * sw - pw + s * dw;
                                if (iw < 0 || iw >= W) continue;
                                const long xi = ((b * C + c) * H + ih) * W + iw, wi = ((o * C + c) * R + r) * S + s;
                                const long double x = X[xi], w = Wt[wi];
                                gx[xi] += w * g;
                                ax[xi] += fabsl(w) * fabsl(g);
                                gw[wi] += g * x;
                                aw[wi] += fabsl(g) * fabsl(x);
                            }
                        }
                }
    for (long i = 0; i < nx; ++i) {
        dXref[i] = (double)gx[i];
        absdX[i] = (double)ax[i];
    }
    for (long i = 0; i < nw; ++i) {
        dWref[i] = (double)gw[i];
        absdW[i] = (double)aw[i];
    }
    free(gx);
}

static long check_core2(const void *Cv, int is_f64, const double *Cref, const double *absAB,
                        const double *Kelem, long n, long K, double u, double extra, double *stats)
{
    const float *Cf = (const float *)Cv;
    const double *Cd = (const double *)Cv;
    double max_ratio = 0.0, max_rel = 0.0, sum_rel = 0.0, max_abs = 0.0;
    long bad = 0, nrel = 0;
    for (long i = 0; i < n; ++i) {
        double kk = Kelem ? Kelem[i] : (double)K;
        if (kk < 1.0) kk = 1.0;
        double c = is_f64 ? Cd[i] : (double)Cf[i];
        double err = fabs(c - Cref[i]);
        if (isnan(c) || isnan(err)) { ++bad; max_ratio = INFINITY; continue; }
        if (err > max_abs) max_abs = err;
        double bound = (2.0 * kk * u + extra) * absAB[i];
        if (absAB[i] == 0.0) {
            if (err != 0.0) { ++bad; max_ratio = INFINITY; }
            continue;
        }
        double ratio = err / bound;
        double rel = err / absAB[i];
        if (ratio > 1.0) ++bad;
        if (ratio > max_ratio) max_ratio = ratio;
        if (rel > max_rel) max_rel = rel;
        sum_rel += rel;
        ++nrel;
    }
    stats[0] = max_ratio;
    stats[1] = max_rel;
    stats[2] = nrel ? sum_rel / (double)nrel : 0.0;
    stats[3] = max_abs;
    return bad;
}

long oracle_check_f32(const float *C, const double *Cref, const double *absAB,
                      long n, long K, double u, double *stats)
{
    return check_core2(C, 0, Cref, absAB, 0, n, K, u, 0.0, stats);
}

long oracle_check_f64(const double *C, const double *Cref, const double *absAB,
                      long n, long K, double u, double *stats)
{
    return check_core2(C, 1, Cref, absAB, 0, n, K, u, 0.0, stats);
}

long oracle_check_kv(const void *C, int is_f64, const double *Cref, const double *absAB,
                     const double *Kelem, long n, double u, double extra, double *stats)
{
    return check_core2(C, is_f64, Cref, absAB, Kelem, n, 0, u, extra, stats);
}
