// host_asan.cpp -- drives the host layer of the library (argument checks,
// the planner and tile queries, the tune-file parser, the thread config and
// status bookkeeping, the batched-plan builder's input checks) with no GPU
// present, from four threads at once.  It is linked with the library's host
// code built under AddressSanitizer and UndefinedBehaviorSanitizer (tests/c/
// Makefile, target host_asan), so an out-of-bounds access, a use after free,
// a leak in these paths or undefined arithmetic ends it with an error.
// Exit status 0: every call returned the status it should.
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "LibShalom.h"

static int failures = 0;
#define EXPECT(call, want)                                                                       \
    do {                                                                                         \
        const int got_ = (call);                                                                 \
        if (got_ != (want)) {                                                                    \
            std::fprintf(stderr, "%s:%d: %s returned %d, expected %d\n", __FILE__, __LINE__, #call, \
                         got_, (want));                                                          \
            __atomic_add_fetch(&failures, 1, __ATOMIC_RELAXED);                                  \
        }                                                                                        \
    } while (0)

static void host_paths(int t)
{
    float *const fake = reinterpret_cast<float *>(0x1000);
    double *const dfake = reinterpret_cast<double *>(0x1000);
    // argument checks, before any device call
    EXPECT(LibShalom_sgemm_ex(2, 0, fake, fake, fake, 4, 4, 4, LIBSHALOM_FP32, nullptr), LIBSHALOM_INVALID_VALUE);
    EXPECT(LibShalom_sgemm_ex(0, 0, fake, fake, fake, -1, 4, 4, LIBSHALOM_FP32, nullptr), LIBSHALOM_INVALID_VALUE);
    EXPECT(LibShalom_sgemm_ex(0, 0, nullptr, fake, fake, 4, 4, 4, LIBSHALOM_FP32, nullptr), LIBSHALOM_INVALID_POINTER);
    EXPECT(LibShalom_sgemm_ex(0, 0, fake, fake, fake, 4, 4, 4, 9, nullptr), LIBSHALOM_INVALID_VALUE);
    EXPECT(LibShalom_sgemm_mp_ex(0, 0, fake, fake, fake, 1L << 31, 4, 4, LIBSHALOM_BF16X3, nullptr),
           LIBSHALOM_NOT_SUPPORTED);
    EXPECT(LibShalom_dgemm_ex(0, 1, dfake, dfake, dfake, 4, 4, -4, nullptr), LIBSHALOM_INVALID_VALUE);
    EXPECT(LibShalom_sgemm_general(0, 0, 4, 4, 4, 1.f, fake, 3, fake, 4, 0.f, fake, 4, LIBSHALOM_FP32, nullptr),
           LIBSHALOM_INVALID_VALUE);
    EXPECT(LibShalom_sgemm_general(0, 0, 4, 4, 4, 0.f, nullptr, 4, nullptr, 4, 1.f, fake, 4, LIBSHALOM_FP32,
                                   nullptr),
           LIBSHALOM_NO_DEVICE);  // alpha = 0: A and B are not needed
    EXPECT(LibShalom_dgemm_general(1, 1, 4, 4, 4, 1.0, dfake, 4, dfake, 4, 0.0, nullptr, 4, nullptr),
           LIBSHALOM_INVALID_POINTER);
    EXPECT(LibShalom_sgemm_ex(0, 0, fake, fake, fake, 4, 4, 4, LIBSHALOM_FP32, nullptr), LIBSHALOM_NO_DEVICE);
    EXPECT(LibShalom_sgemm_ex(0, 0, fake, fake, fake, 0, 4, 4, LIBSHALOM_FP32, nullptr), LIBSHALOM_NO_DEVICE);
    // the reference API reports through the per-thread last status
    LibShalom_sgemm(NoTrans, Trans, nullptr, nullptr, nullptr, 8, 8, 8);
    EXPECT(LibShalom_get_last_error(), LIBSHALOM_INVALID_POINTER);
    LibShalom_dgemm(NoTrans, 3, dfake, dfake, dfake, 8, 8, 8);
    EXPECT(LibShalom_get_last_error(), LIBSHALOM_INVALID_VALUE);
    for (int s = 0; s <= 7; ++s)
        if (!LibShalom_status_string(s)) __atomic_add_fetch(&failures, 1, __ATOMIC_RELAXED);
    // planner queries over a grid of shapes, every precision and mp flag
    const long dims[] = {1, 2, 7, 16, 17, 33, 80, 127, 128, 129, 1000, 4096, 65536, 1L << 20};
    for (int prec = 0; prec <= 4; ++prec)
        for (long M : dims)
            for (long N : dims)
                for (long K : {1L, 5L, 64L, 1024L, 1L << 16}) {
                    int tm, tn, ks, kp;
                    EXPECT(LibShalom_query_plan(prec, (int)((M + N + t) & 1), M, N, K, 256, &tm, &tn, &ks, &kp),
                           LIBSHALOM_SUCCESS);
                    if (ks < 1 || kp < 1 || (long)ks * kp < K || tm < 16 || tn < 16)
                        __atomic_add_fetch(&failures, 1, __ATOMIC_RELAXED);
                    EXPECT(LibShalom_query_batched_tile(prec, M, N, K, &tm, &tn), LIBSHALOM_SUCCESS);
                    EXPECT(LibShalom_query_tuned_tile(prec == 4 ? 0 : prec, (int)(M & 1), (int)(N & 1), M, N, K,
                                                      &tm, &tn),
                           LIBSHALOM_SUCCESS);
                }
    int tm, tn, ks, kp;
    EXPECT(LibShalom_query_plan(7, 0, 4, 4, 4, 256, &tm, &tn, &ks, &kp), LIBSHALOM_INVALID_VALUE);
    EXPECT(LibShalom_query_plan(0, 0, -4, 4, 4, 256, &tm, &tn, &ks, &kp), LIBSHALOM_INVALID_VALUE);
    // convolution geometry
    long P, Q;
    EXPECT(LibShalom_conv2d_output_size(56, 56, 3, 3, 1, 1, 1, 1, 1, 1, &P, &Q), LIBSHALOM_SUCCESS);
    if (P != 56 || Q != 56) __atomic_add_fetch(&failures, 1, __ATOMIC_RELAXED);
    EXPECT(LibShalom_conv2d_output_size(2, 8, 5, 3, 1, 1, 0, 1, 1, 1, &P, &Q), LIBSHALOM_SUCCESS);
    if (P != 0 || Q != 8) __atomic_add_fetch(&failures, 1, __ATOMIC_RELAXED);  // the filter does not fit
    EXPECT(LibShalom_sconv2d_forward(fake, fake, fake, 2, 3, 8, 8, 4, 3, 3, 1, 1, 1, 1, 1, 1, 7, nullptr),
           LIBSHALOM_INVALID_VALUE);
    // batched plans: the input checks run before the device is touched
    LibShalom_plan_t plan = nullptr;
    std::vector<float *> Cs(3, fake);
    std::vector<const float *> As(3, fake), Bs(3, fake);
    std::vector<long> Ms = {4, -1, 4}, Ns = {4, 4, 4}, Ks = {4, 4, 4};
    EXPECT(LibShalom_sgemm_batched_plan(&plan, 0, 0, Cs.data(), As.data(), Bs.data(), Ms.data(), Ns.data(),
                                        Ks.data(), 3, LIBSHALOM_FP32),
           LIBSHALOM_INVALID_VALUE);
    EXPECT(LibShalom_sgemm_batched_plan(&plan, 0, 0, Cs.data(), As.data(), Bs.data(), Ms.data(), Ns.data(),
                                        Ks.data(), 3, 11),
           LIBSHALOM_INVALID_VALUE);
    // thread configuration (recorded, no effect) and precision
    LibShalom_set_thread_ways(2 + t, 3);
    int num, a, b;
    LibShalom_get_thread_config(&num, &a, &b);
    EXPECT(LibShalom_set_precision(9), LIBSHALOM_INVALID_VALUE);
}

int main()
{
    std::vector<std::thread> th;
    for (int t = 0; t < 4; ++t) th.emplace_back(host_paths, t);
    for (auto &x : th) x.join();
    EXPECT(LibShalom_set_precision(LIBSHALOM_BF16X6), LIBSHALOM_SUCCESS);
    EXPECT(LibShalom_get_precision(), LIBSHALOM_BF16X6);
    EXPECT(LibShalom_set_precision(LIBSHALOM_FP32), LIBSHALOM_SUCCESS);
    std::printf("host_asan: %d failures\n", failures);
    return failures ? 1 : 0;
}
