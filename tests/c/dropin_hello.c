/*
 * dropin_hello.c -- a plain C program written against the reference's ABI
 * (the LibShalom_sgemm / LibShalom_sgemm_mp signatures, NoTrans / Trans, host
 * buffers from malloc), built with a C compiler against include/LibShalom.h
 * and linked with -lsmm, the library name the reference's Makefile produces
 * (NN_LIB/Makefile:7-8).  It is how a C caller of the reference switches
 * over: no source change, a different -I / -L.
 *
 *   dropin_hello <out.bin> [hello | mp]
 *
 * hello: C = A * B^T at M = N = K = 80 (LibShalom_sgemm, NoTrans, Trans);
 * mp:    C = A * B at 4096 x 16 x 1024 (LibShalom_sgemm_mp, NoTrans, NoTrans).
 * The operands are filled from a fixed integer recurrence; A, B and C are
 * written to out.bin (float32, in that order) for the caller to check, and
 * the call's status (LibShalom_get_last_error) is the exit code.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "LibShalom.h"

/* values in [-1, 1) from a 32-bit LCG (Numerical Recipes constants) */
static unsigned lcg_state = 12345u;
static float next_value(void)
{
    lcg_state = lcg_state * 1664525u + 1013904223u;
    return (float)(lcg_state >> 8) / 8388608.0f - 1.0f;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s out.bin [hello|mp]\n", argv[0]);
        return 64;
    }
    const int mp = argc > 2 && strcmp(argv[2], "mp") == 0;
    const long M = mp ? 4096 : 80, N = mp ? 16 : 80, K = mp ? 1024 : 80;
    float *A = (float *)malloc(sizeof(float) * M * K);
    float *B = (float *)malloc(sizeof(float) * K * N);
    float *C = (float *)malloc(sizeof(float) * M * N);
    if (!A || !B || !C) return 65;
    for (long i = 0; i < M * K; ++i) A[i] = next_value();
    for (long i = 0; i < K * N; ++i) B[i] = next_value();
    for (long i = 0; i < M * N; ++i) C[i] = -12345.0f; /* every element must be written */

    if (mp)
        LibShalom_sgemm_mp(NoTrans, NoTrans, C, A, B, M, N, K);
    else
        LibShalom_sgemm(NoTrans, Trans, C, A, B, M, N, K); /* B stored N x K */
    const int status = LibShalom_get_last_error();
    printf("%s %ldx%ldx%ld: %s\n", mp ? "LibShalom_sgemm_mp" : "LibShalom_sgemm", M, N, K,
           LibShalom_status_string(status));

    FILE *f = fopen(argv[1], "wb");
    if (!f) return 66;
    fwrite(A, sizeof(float), (size_t)(M * K), f);
    fwrite(B, sizeof(float), (size_t)(K * N), f);
    fwrite(C, sizeof(float), (size_t)(M * N), f);
    fclose(f);
    free(A);
    free(B);
    free(C);
    return status;
}
