// api.cpp -- the extern "C" boundary of libshalom.so: argument checking, the
// tile / split-K planner that replaces the reference's OpenMP thread-grid
// planner, per-stream workspaces, and batched plans captured in hipGraphs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/LibShalom.h"
#include "backward.hpp"
#include "kernels.hpp"
#include "planner.hpp"
#include "skinny.hpp"

using namespace ls;

namespace ls {
LaunchTable &launch_table()
{
    static LaunchTable *t = [] {
        LaunchTable *p = new LaunchTable();
        std::memset(p, 0, sizeof(LaunchTable));
        register_fp32_ta0(*p);
        register_fp32_ta1(*p);
        register_bf16x3_ta0(*p);
        register_bf16x3_ta1(*p);
        register_bf16_ta0(*p);
        register_bf16_ta1(*p);
        register_bf16x6_ta0(*p);
        register_bf16x6_ta1(*p);
        register_fp64_ta0(*p);
        register_fp64_ta1(*p);
        return p;
    }();
    return *t;
}
BwdTable &bwd_table()
{
    static BwdTable *t = [] {
        BwdTable *p = new BwdTable();
        std::memset(p, 0, sizeof(BwdTable));
        register_backward_fp32(*p);
        register_backward_bf16x3(*p);
        register_backward_bf16(*p);
        register_backward_bf16x6(*p);
        return p;
    }();
    return *t;
}
__global__ void k_xcc_probe(int *out, int n)
{
    if (threadIdx.x == 0 && (int)blockIdx.x < n) out[blockIdx.x] = xcc_id();
}
}  // namespace ls

namespace {

thread_local int g_last_status = LIBSHALOM_SUCCESS;
constexpr long kMaxLd = 1L << 21;  // leading dimensions (elements) the kernels' 32-bit offsets allow
std::atomic<int> g_precision{LIBSHALOM_FP32};
std::atomic<int> g_thread_nums{64}, g_tm{0}, g_tn{0};

int set_status(int s)
{
    g_last_status = s;
    return s;
}

// ---------------------------------------------------------------------------
// Devices.
// ---------------------------------------------------------------------------
struct DeviceInfo {
    bool checked = false;
    bool ok = false;
    int num_cu = 256;
    // workgroup L of a 1-D launch runs on XCD L % 8 (checked once by
    // k_xcc_probe): split-K slices of a tile can then share one L2
    bool xcd_rr = false;
};
std::mutex g_dev_mu;
std::map<int, DeviceInfo> g_devices;

constexpr int kProbeBlocks = 64;
__device__ int g_xcc_probe[kProbeBlocks];

// Whether the device deals the workgroups of a 1-D launch to its 8 XCDs
// round-robin (from a start that depends on the queue and its history): a
// launch of kProbeBlocks workgroups records each one's XCD.
// Any failure (or a stream capture in progress) answers no.
bool probe_xcd_round_robin()
{
    int *p = nullptr;
    if (hipGetSymbolAddress((void **)&p, HIP_SYMBOL(g_xcc_probe)) != hipSuccess || !p) {
        (void)hipGetLastError();
        return false;
    }
    hipLaunchKernelGGL(k_xcc_probe, dim3(kProbeBlocks), dim3(64), 0, nullptr, p, kProbeBlocks);
    int h[kProbeBlocks];
    if (hipGetLastError() != hipSuccess || hipMemcpy(h, p, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    // round-robin from some start: (XCD - workgroup) mod 8 the same for all
    for (int i = 0; i < kProbeBlocks; ++i)
        if (h[i] < 0 || h[i] > 7 || (h[i] - i + 64) % 8 != (h[0] + 64) % 8) return false;
    return true;
}

int current_device(int *dev, DeviceInfo *info)
{
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) {
        (void)hipGetLastError();
        return LIBSHALOM_NO_DEVICE;
    }
    std::lock_guard<std::mutex> lk(g_dev_mu);
    DeviceInfo &di = g_devices[d];
    if (!di.checked) {
        di.checked = true;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) == hipSuccess) {
            di.ok = std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
            di.num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
            if (di.ok) di.xcd_rr = probe_xcd_round_robin();
        } else {
            (void)hipGetLastError();
        }
    }
    if (!di.ok) return LIBSHALOM_NO_DEVICE;
    *dev = d;
    *info = di;
    return LIBSHALOM_SUCCESS;
}

// A pointer the GPU can dereference: device, managed or pinned-host memory.
bool device_accessible(const void *p)
{
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
           attr.type == hipMemoryTypeUnified || attr.devicePointer != nullptr;
}

// ---------------------------------------------------------------------------
// Per (device, stream) device buffers: the workspace (a convolution's
// flipped filter) and the fixup space of the in-launch reductions.
//
// A buffer that has to grow is replaced, and the old one is kept, not
// freed, until the process ends: a kernel another host thread is about to
// enqueue with it (its caller took the pointer before the growth) and a
// captured graph that recorded it both stay valid.  Growth at least doubles,
// so the kept buffers add up to less than the live one.  A stream in graph
// capture cannot grow its buffers (allocation and the zeroing of a new
// fixup space would happen outside the graph): such a call returns
// LIBSHALOM_NOT_SUPPORTED; an eager call of the same size on that stream
// beforehand makes it possible.
// ---------------------------------------------------------------------------
struct GrowBuf {
    char *ptr = nullptr;
    size_t size = 0;
    std::vector<char *> retired;
};

bool stream_capturing(hipStream_t s)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return cs != hipStreamCaptureStatusNone;
}

// Makes b hold at least `bytes` (at least `floor`), retiring the old buffer.
// *grown tells whether a new buffer was allocated.
int grow_buf(GrowBuf &b, hipStream_t s, size_t bytes, size_t floor, bool *grown)
{
    *grown = false;
    if (b.size >= bytes) return LIBSHALOM_SUCCESS;
    if (stream_capturing(s)) return LIBSHALOM_NOT_SUPPORTED;
    const size_t sz = std::max({bytes, floor, 2 * b.size});
    char *p = nullptr;
    if (hipMalloc((void **)&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        return LIBSHALOM_ALLOC_FAILED;
    }
    if (b.ptr) b.retired.push_back(b.ptr);
    b.ptr = p;
    b.size = sz;
    *grown = true;
    return LIBSHALOM_SUCCESS;
}

std::mutex g_ws_mu;
std::map<std::pair<int, hipStream_t>, GrowBuf> g_ws;

int workspace(int dev, hipStream_t s, size_t bytes, void **out)
{
    if (const char *f = std::getenv("LIBSHALOM_FAIL_WORKSPACE"))
        if (f[0] == '1') return LIBSHALOM_ALLOC_FAILED;  // failure injection for tests
    std::lock_guard<std::mutex> lk(g_ws_mu);
    GrowBuf &w = g_ws[{dev, s}];
    bool grown;
    const int st = grow_buf(w, s, bytes, (size_t)1 << 20, &grown);
    if (st) return st;
    *out = w.ptr;
    return LIBSHALOM_SUCCESS;
}

// ---------------------------------------------------------------------------
// In-launch reductions (split-K of single GEMMs, sliced backward / strided
// problems).
// ---------------------------------------------------------------------------
// The fixup space of a (device, stream): arrival counters (zero between
// launches; every launch leaves them zero) and the partial-tile slots.
std::mutex g_fx_mu;
std::map<std::pair<int, hipStream_t>, GrowBuf> g_fx;
constexpr size_t kCounterBytes = 1 << 20;  // counters of 128 K tiles (two ints each)
// Partial tiles a single GEMM's split-K may use (the planner picks fewer K
// slices when more would be needed).
constexpr size_t kSplitKSlotLimit = (size_t)256 << 20;

__global__ void k_zero_ints(int *p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

int fixup_space(int dev, hipStream_t s, size_t slot_bytes, int **counters, float **slots)
{
    if (const char *f = std::getenv("LIBSHALOM_FAIL_WORKSPACE"))
        if (f[0] == '1') return LIBSHALOM_ALLOC_FAILED;  // failure injection for tests
    std::lock_guard<std::mutex> lk(g_fx_mu);
    GrowBuf &w = g_fx[{dev, s}];
    bool grown;
    int st = grow_buf(w, s, kCounterBytes + slot_bytes, (size_t)8 << 20, &grown);
    if (st) return st;
    if (grown) {
        // zeroed on the stream, ahead of any kernel that gets this pointer
        // (it is handed out under the lock, after this launch)
        hipLaunchKernelGGL(k_zero_ints, dim3(256), dim3(256), 0, s, (int *)w.ptr, (int)(kCounterBytes / sizeof(int)));
        if (hipGetLastError() != hipSuccess) return LIBSHALOM_HIP_ERROR;
    }
    *counters = (int *)w.ptr;
    *slots = (float *)(w.ptr + kCounterBytes);
    return LIBSHALOM_SUCCESS;
}

// ---------------------------------------------------------------------------
// Argument checks shared by every entry point.
// ---------------------------------------------------------------------------
int check_dims(int ta, int tb, long M, long N, long K)
{
    if ((ta != 0 && ta != 1) || (tb != 0 && tb != 1)) return LIBSHALOM_INVALID_VALUE;
    if (M < 0 || N < 0 || K < 0) return LIBSHALOM_INVALID_VALUE;
    const long lim = (1L << 31) - 1;
    if (M > lim || N > lim || K > lim) return LIBSHALOM_NOT_SUPPORTED;
    // the kernels address a tile's rows with 32-bit byte offsets
    const long ld_lim = kMaxLd;
    if ((ta ? M : K) >= ld_lim || (tb ? K : N) >= ld_lim || N >= ld_lim) return LIBSHALOM_NOT_SUPPORTED;
    return LIBSHALOM_SUCCESS;
}

int check_ptrs(const void *C, const void *A, const void *B, long M, long N, long K, bool deep)
{
    if (M == 0 || N == 0) return LIBSHALOM_SUCCESS;
    if (!C) return LIBSHALOM_INVALID_POINTER;
    if (K > 0 && (!A || !B)) return LIBSHALOM_INVALID_POINTER;
    if (deep) {
        if (!device_accessible(C)) return LIBSHALOM_INVALID_POINTER;
        if (K > 0 && (!device_accessible(A) || !device_accessible(B))) return LIBSHALOM_INVALID_POINTER;
    }
    return LIBSHALOM_SUCCESS;
}

GemmDesc make_desc(int ta, int tb, const void *A, const void *B, void *C, long M, long N, long K, int esize)
{
    GemmDesc d;
    d.A = A;
    d.B = B;
    d.C = C;
    d.M = (int)M;
    d.N = (int)N;
    d.K = (int)K;
    d.lda = (int)(ta ? M : K);
    d.ldb = (int)(tb ? K : N);
    d.ldc = (int)N;
    (void)esize;
    d.flags = 0;
    return d;
}

// Leading dimensions of a row-major GEMM: A is M x K (lda >= K) or, with
// transa, K x M (lda >= M); B is K x N (ldb >= N) or, with transb, N x K
// (ldb >= K); C is M x N (ldc >= N).  Every one is at least 1.
int check_lds(int ta, int tb, long M, long N, long K, long lda, long ldb, long ldc)
{
    const long a_min = std::max(1L, ta ? M : K), b_min = std::max(1L, tb ? K : N), c_min = std::max(1L, N);
    if (lda < a_min || ldb < b_min || ldc < c_min) return LIBSHALOM_INVALID_VALUE;
    if (lda >= kMaxLd || ldb >= kMaxLd || ldc >= kMaxLd) return LIBSHALOM_NOT_SUPPORTED;
    return LIBSHALOM_SUCCESS;
}

// LIBSHALOM_SINGLE_PLAN="tile:ksplit" (tile = index into kTileShapes) forces
// the plan of single GEMMs where it is valid -- a tuning and testing knob,
// read once.
void forced_plan(int prec, long M, long N, long K, SinglePlan *p)
{
    static const std::pair<int, int> forced = [] {
        int t = -1, ks = 0;
        if (const char *e = std::getenv("LIBSHALOM_SINGLE_PLAN"))
            if (std::sscanf(e, "%d:%d", &t, &ks) != 2) t = -1;
        return std::make_pair(t, ks);
    }();
    const int t = forced.first, ks = forced.second;
    if (t < 0 || t >= NTILE || ks < 1 || ks > 64 || (prec == P_FP64 && t == T128x128)) return;
    const int kd = chunk_depth(prec), esize = prec == P_FP64 ? 8 : 4;
    const long kper = ks == 1 ? std::max<long>(K, 1) : cdiv(cdiv(K, ks), kd) * kd;
    const int nks = ks == 1 ? 1 : (int)cdiv(K, kper);
    const TileShape ts = kTileShapes[t];
    if (nks > 1 && ((size_t)cdiv(M, ts.bm) * cdiv(N, ts.bn) > kCounterBytes / (2 * sizeof(int)) ||
                    splitk_slot_bytes(t, M, N, nks, esize) > kSplitKSlotLimit))
        return;
    *p = SinglePlan{t, nks, (int)kper};
}

// One GEMM on `stream`, C = alpha * op(A) op(B) + beta * C with explicit
// leading dimensions: plan tiles / split-K and launch.
// ---------------------------------------------------------------------------
// Scratch of the no-allocation entry point (LibShalom_sgemm_general_noalloc):
// split-K arrival counters and partial-tile slots in a file-scope device
// array, one copy per GPU, held until the process ends.  Host code uses only
// the device address hipGetSymbolAddress returns for the current device
// (noalloc_scratch), never the array's name.  The counters are zeroed once,
// when that address is taken, and every launch leaves them zero again.
// Calls that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocCounterBytes = (size_t)64 << 10;  // counters of 8 K tiles
constexpr size_t kNoAllocSlotBytes = (size_t)32 << 20;     // partial tiles
__device__ __align__(256) char scratch[kNoAllocCounterBytes + kNoAllocSlotBytes];
std::mutex g_noalloc_mu;
std::map<int, char *> g_noalloc_ptr;

int noalloc_scratch(int dev, char **out)
{
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    auto it = g_noalloc_ptr.find(dev);
    if (it != g_noalloc_ptr.end()) {
        *out = it->second;
        return LIBSHALOM_SUCCESS;
    }
    char *ptr = nullptr;
    if (hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return LIBSHALOM_HIP_ERROR;
    }
    // a fill kernel, then wait: the counters must be zero before any use
    hipLaunchKernelGGL(k_zero_ints, dim3(64), dim3(256), 0, nullptr, (int *)ptr,
                       (int)(kNoAllocCounterBytes / sizeof(int)));
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(nullptr) != hipSuccess) {
        (void)hipGetLastError();
        return LIBSHALOM_HIP_ERROR;
    }
    g_noalloc_ptr[dev] = ptr;
    *out = ptr;
    return LIBSHALOM_SUCCESS;
}

// The short-side kernel (skinny.hpp) for a GEMM whose C is tall (N <=
// 64) or wide (M <= 64) with a long side L of at least 1024, when it beats
// the tile kernels' split-K plans (measured on MI355X, sgemm_mp shapes, A/B
// against the tile plans): K >= 512 (below that both are launch-bound); the
// K x short-side work a CU's workgroups carry in series (rounds of L / 16
// workgroups over the CUs, times K times the 16-column blocks) at most 4096,
// and at most 2048 when the workgroups leave CUs idle (1024 x 64 x 4096:
// 35 us against 19; 2048 x 32 x 2048: 14.5 against 12.0); a wide C in one
// round (16 x 16384 x 1024 ran slower).  Wins: 4096 x 16 x 1024 11.6 -> 6.3
// us, 2048 x 16 x 1024 9.6 -> 6.1 us, 16 x 4096 x 1024 11.3 -> 7.0 us,
// 4096 x 16 x 4096 21.8 -> 17.1 us; in BF16X3 4096 x 16 x 1024 11.7 -> 6.5 us,
// 4096 x 32 x 1024 23.3 -> 16.0 us.  And the operands' byte offsets (past
// their extents by at most a group of 16 indices) must stay in 31 bits.
// Returns the kernel's column-block count index (0, 1, 2 for 16, 32, 64
// short-side columns), or -1 for the tile kernels.  LIBSHALOM_SKINNY=0
// turns it off (measurement).
int skinny_choice(int prec, int ta, int tb, long M, long N, long K, long lda, long ldb, int num_cu)
{
    static const bool off = [] {
        const char *e = std::getenv("LIBSHALOM_SKINNY");
        return e && e[0] == '0';
    }();
    if (off || prec == P_FP64) return -1;
    const bool wide = M <= 64 && N >= 1024, tall = N <= 64 && M >= 1024;
    if (!wide && !tall) return -1;
    const long L = wide ? N : M, S = wide ? M : N;
    // X: op(A) (tall) or op(B)^T (wide); Y: the other, with their index counts
    const bool xkc = wide ? tb != 0 : ta == 0, ykc = wide ? ta == 0 : tb != 0;
    const long xld = wide ? ldb : lda, yld = wide ? lda : ldb;
    auto fits = [&](bool kc, long count, long ld) {
        const long span = kc ? (count + 16) * ld + K + 64 : (K + 64) * ld + count + 64;
        return span * 4 < (1L << 31);
    };
    if (!fits(xkc, L, xld) || !fits(ykc, S + 48, yld)) return -1;
    const int nc = S <= 16 ? 0 : S <= 32 ? 1 : 2;
    const long nwg = (L + 15) / 16, rounds = (nwg + num_cu - 1) / num_cu, chain = K << nc;
    if (K < 512 || rounds * chain > 4096 || (nwg < num_cu && chain > 2048) || (wide && rounds > 1)) return -1;
    // the split modes' tile plans keep up on short sides above 32 (their
    // MFMA work per byte is a third of FP32's): 4096 x 64 x 1024 NT and
    // 64 x 4096 x 1024 NT ran 10-15% slower on the short-side kernel
    if (prec != P_FP32 && nc == 2) return -1;
    return nc;
}

// One GEMM on `stream`: plan tiles / split-K and launch.  With noalloc the
// split-K counters and slots come from the static scratch array and the plan
// is limited to what it holds; otherwise from the per-stream fixup space.
int run_general(int prec, int ta, int tb, void *C, const void *A, const void *B, long M, long N, long K, long lda,
                long ldb, long ldc, double alpha, double beta, bool mp, hipStream_t stream, bool noalloc = false)
{
    int dev;
    DeviceInfo info;
    int st = current_device(&dev, &info);
    if (st) return st;
    if (M == 0 || N == 0) return LIBSHALOM_SUCCESS;
    const bool f64 = prec == P_FP64;
    const int esize = f64 ? 8 : 4;
    GemmDesc d = make_desc(ta, tb, A, B, C, M, N, K, esize);
    d.lda = (int)lda;
    d.ldb = (int)ldb;
    d.ldc = (int)ldc;
    const int nc = skinny_choice(prec, ta, tb, M, N, K, lda, ldb, info.num_cu);
    if (nc >= 0) {
        // eight waves (K in eighths) for a tall NN GEMM whose workgroups
        // carry K x short-side work of 4096 in one round: 4096 x 64 x 1024
        // 12.7 -> 10.3 us, 4096 x 16 x 4096 17.0 -> 14.7 us; the other
        // cases measured the same or slower with them
        const bool wide = M <= 64;
        const long nwg = ((wide ? N : M) + 15) / 16;
        const int nw8 = prec == P_FP32 && !wide && !ta && !tb && nwg <= info.num_cu && (K << nc) >= 4096;
        SkinnyLaunch sk = skinny_table().fn[prec][ta][tb][wide][nc][nw8];
        if (!sk) return LIBSHALOM_NOT_SUPPORTED;
        const SkinnyArgs a{(const float *)d.A, (const float *)d.B, (float *)C, (int)M, (int)N, (int)K,
                           (int)lda, (int)ldb, (int)ldc, (float)alpha, (float)beta};
        return sk(a, stream) == hipSuccess ? LIBSHALOM_SUCCESS : LIBSHALOM_HIP_ERROR;
    }
    // split-K needs the XCD placement of its slices (k_gemm_tiles); without
    // it the agent-scope fences make a split slower than none
    const size_t slot_limit = noalloc ? kNoAllocSlotBytes : kSplitKSlotLimit;
    const size_t counter_limit = (noalloc ? kNoAllocCounterBytes : kCounterBytes) / (2 * sizeof(int));
    SinglePlan p = plan_single(prec, M, N, K, mp, info.num_cu, slot_limit, info.xcd_rr ? counter_limit : 0);
    if (info.xcd_rr && !noalloc) forced_plan(prec, M, N, K, &p);
    TilesLaunch fn = launch_table().tiles[prec][ta][tb][p.tile];
    if (!fn) return LIBSHALOM_NOT_SUPPORTED;
    int *counters = nullptr;
    float *slots = nullptr;
    if (p.ksplit > 1) {
        const size_t need = splitk_slot_bytes(p.tile, M, N, p.ksplit, esize);
        if (noalloc) {
            char *ptr;
            if (need > kNoAllocSlotBytes) return LIBSHALOM_NOT_SUPPORTED;
            st = noalloc_scratch(dev, &ptr);
            if (st) return st;
            counters = (int *)ptr;
            slots = (float *)(ptr + kNoAllocCounterBytes);
        } else {
            st = fixup_space(dev, stream, need, &counters, &slots);
            if (st) return st;
        }
    }
    if (fn(d, p.ksplit, p.kper, slots, counters, alpha, beta, info.xcd_rr, stream) != hipSuccess)
        return LIBSHALOM_HIP_ERROR;
    return LIBSHALOM_SUCCESS;
}

// One GEMM with the reference's semantics: implicit leading dimensions,
// C overwritten (alpha 1, beta 0).
int run_single(int prec, int ta, int tb, void *C, const void *A, const void *B, long M, long N, long K,
               bool mp, hipStream_t stream)
{
    return run_general(prec, ta, tb, C, A, B, M, N, K, ta ? M : K, tb ? K : N, N, 1.0, 0.0, mp, stream);
}

int sync_null()
{
    if (hipStreamSynchronize(nullptr) != hipSuccess) {
        (void)hipGetLastError();
        return LIBSHALOM_HIP_ERROR;
    }
    return LIBSHALOM_SUCCESS;
}

int valid_precision(int p)
{
    return p == LIBSHALOM_FP32 || p == LIBSHALOM_BF16X3 || p == LIBSHALOM_BF16 || p == LIBSHALOM_BF16X6;
}

// LIBSHALOM_* precision flag -> kernel precision id (they coincide).
int kprec(int p) { return p; }

long ceil_div(long a, long b) { return (a + b - 1) / b; }

// One gradient as a problem: output R x Cc (leading dimension ldo), left
// operand R x red, right operand red x Cc, per-image strides; sO == 0 with a
// batch means one output summed over the images.
void bwd_side(BwdProblem &P, const void *L, long long sL, int lda, const void *Rt, long long sR, int ldb, void *O,
              long long sO, int ldo, long R, long Cc, long red, int nimg)
{
    std::memset(&P, 0, sizeof(P));
    P.L = L;
    P.Rt = Rt;
    P.O = O;
    P.sL = sL;
    P.sR = sR;
    P.sO = nimg > 1 ? sO : 0;
    P.lda = lda;
    P.ldb = ldb;
    P.ldo = ldo;
    P.R = (int)R;
    P.Cc = (int)Cc;
    P.red = (int)red;
    P.nimg = nimg;
    P.kalign = 32;
}

// Cut the reductions of both problems into slices of about `len` (a
// multiple of the K chunk) and count units.
void bwd_slice(BwdProblem &P, int T, long len)
{
    const bool shared = P.sO == 0;
    P.tiles_c = (int)ceil_div(P.Cc, T);
    P.ntiles = (int)ceil_div(P.R, T) * P.tiles_c;
    const long outs = shared ? 1 : P.nimg;
    const long total = shared ? (long)P.nimg * P.red : P.red;
    if (P.ntiles == 0) {
        P.units = P.grid = 0;
        P.nslices = 1;
        return;
    }
    if (total <= len || P.red == 0) {
        P.nslices = 1;
        P.spi = shared ? 0 : 1;
        P.gimg = shared ? P.nimg : 1;
        P.slen = P.red;
    } else if (!shared || P.red >= len) {
        long spi = ceil_div(P.red, len);
        const long slen = ceil_div(ceil_div(P.red, spi), P.kalign) * P.kalign;
        spi = ceil_div(P.red, slen);
        P.spi = (int)spi;
        P.slen = (int)slen;
        P.gimg = 1;
        P.nslices = (int)(shared ? P.nimg * spi : spi);
    } else {
        const long g = std::max(1L, len / std::max(1, P.red));
        P.spi = 0;
        P.gimg = (int)g;
        P.slen = P.red;
        P.nslices = (int)ceil_div(P.nimg, g);
    }
    P.units = (int)(outs * P.ntiles * P.nslices);
    // with slices, the tile groups padded to a multiple of 8 (unit_of)
    P.grid = P.nslices > 1 ? (int)(((outs * P.ntiles + 7) / 8) * 8 * P.nslices) : P.units;
}

// Tile size and reduction slices for problems that share one launch.
// 128 x 128 tiles when every output dimension fills them and they alone give
// at least big_quarters / 4 of the compute units a unit of work (reduction
// slices fill the rest), else 64 x 64.  (Measured on the conv layers with
// three quarters: the 128x784 layer, 224 big tiles for 256 CUs, runs
// 1.1-1.35x faster on them.  The 256x196 layer, 128 big tiles, runs its
// FP32 forward 3% slower on them (with two slices per tile) and its
// implicit backward 7% slower, but in BF16X3 its stored forward 1.16x, its
// implicit forward 1.27x and its implicit backward 1.2x faster, so the
// split-BF16 modes plan with one half.)  A reduction is cut into
// slices of about W / (2 * CUs) for 64 x 64 tiles, W / CUs for 128 x 128
// tiles (W = all tiles' reduction work) so that no unit is much longer than
// a compute unit's share: a 128 x 128 workgroup fills its CU alone, and one
// unit per CU there measured neutral to 1.4x faster than two on the conv
// layers' backward launches; slices cost partial-tile traffic, which is
// kept under a quarter of the operand traffic -- the larger of the
// algorithmic bytes and what the units load (a shared operand is read by
// every image's tiles: on the 512 x 49 x 4608 layer, batch 32, the units
// load 13x the algorithmic bytes, and two slices per tile took its BF16X3
// forward from 0.094 to 0.064 ms).  occ64: resident 64 x 64 workgroups per
// CU the makespan model assumes (below).  Returns the tile size (0: nothing
// to launch, < 0: -status).
int plan_problems(BwdProblem **probs, int n, bool allow_big, int dev, int num_cu, hipStream_t s,
                  int big_quarters = 3, int occ64 = 2)
{
    auto base_units = [&](int T) {
        long u = 0;
        for (int i = 0; i < n; ++i) {
            const BwdProblem &P = *probs[i];
            u += (P.sO == 0 ? 1L : P.nimg) * ceil_div(P.R, T) * ceil_div(P.Cc, T);
        }
        return u;
    };
    const int T = (allow_big && 4 * base_units(128) >= (long)big_quarters * num_cu) ? 128 : 64;
    double W = 0, ops = 0;
    for (int i = 0; i < n; ++i) {
        const BwdProblem &P = *probs[i];
        const double outs = P.sO == 0 ? 1 : P.nimg;
        const double total = (P.sO == 0 ? (double)P.nimg : 1.0) * std::max(P.red, 1);
        W += outs * ceil_div(P.R, T) * ceil_div(P.Cc, T) * total;
        ops += 4.0 * ((P.sL == 0 ? 1.0 : P.nimg) * P.R * P.red + (P.sR == 0 ? 1.0 : P.nimg) * P.red * P.Cc +
                      outs * P.R * P.Cc);
    }
    long len0 = std::max(256L, (long)std::ceil(W / ((T == 128 ? 1.0 : 2.0) * num_cu)));
    len0 = ceil_div(len0, 32) * 32;
    size_t slot_bytes = 0, counters = 0;
    long units = 0;
    // Slices every problem at slice length len, doubled until the partial
    // tiles fit their limits.
    auto slice_all = [&](long len) {
        for (int it = 0; it < 40; ++it) {
            slot_bytes = counters = 0;
            units = 0;
            double loads = 0;  // operand bytes the units load: each its L and R panels of the slice
            for (int i = 0; i < n; ++i) {
                BwdProblem &P = *probs[i];
                bwd_slice(P, T, len);
                units += P.units;
                loads += (double)P.units * 2.0 * T * (P.spi > 0 ? (double)P.slen : (double)P.gimg * P.red) * 4.0;
                if (P.nslices > 1) {
                    slot_bytes += (size_t)P.units * T * T * sizeof(float);
                    counters += (size_t)(P.sO == 0 ? 1 : P.nimg) * P.ntiles;
                }
            }
            const bool traffic_ok = 2.0 * slot_bytes <= 0.25 * std::max(ops, loads);
            if (traffic_ok && slot_bytes <= ((size_t)256 << 20) && counters * 2 * sizeof(int) <= kCounterBytes) break;
            len *= 2;
        }
    };
    // Of 1, 2 and 4 times that length (and a quarter and a half of it when
    // more than two 64 x 64 workgroups are resident), the one with the least
    // estimated makespan: rounds of units over the resident workgroups (one
    // per CU for 128 x 128 tiles, occ64 for 64 x 64) times the longest
    // unit's reduction steps; ties go to the longer slices (fewer partial
    // tiles) unless those leave more than a quarter of the resident
    // workgroups without a unit where the shorter ones did not (the 14x14
    // conv layer's implicit backward, BF16X3: 170 units of 128 x 128 in one
    // round against 340 in two, 0.175 -> 0.155 ms; the 28x28 layer's
    // forward keeps its 224 unsliced units against 448 halves, 0.038 against
    // 0.058 ms).  (The 7x7 conv layer's backward, 300 units of 128 x 128 for
    // 256 CUs at the base length, runs in one round of 248 at twice it:
    // 0.21 -> 0.15 ms.  Single-problem launches -- the strided and the
    // implicit-convolution forward -- plan with four resident workgroups:
    // measured on the conv layers at batch 32, the 14x14 and 7x7 layers'
    // forwards ran 8-18% faster (BF16X3 implicit 0.093 -> 0.076 ms on 7x7),
    // the others unchanged; the two-problem backward launches ran up to 1.5x
    // slower with it and keep two.)
    const long resident = (long)num_cu * (T == 128 ? 1 : occ64);
    long best = len0;
    double best_cost = 0;
    bool best_fills = false;
    const int mlo = occ64 > 2 ? 1 : 4;  // quarters of len0: from a quarter with more resident units
    for (int m = mlo; m <= 16; m *= 2) {
        const long lenm = std::max(256L, ceil_div(len0 * m / 4, 32) * 32);
        slice_all(lenm);
        double longest = 0;
        for (int i = 0; i < n; ++i) {
            const BwdProblem &P = *probs[i];
            if (P.units > 0) longest = std::max(longest, P.spi > 0 ? (double)P.slen : (double)P.gimg * P.red);
        }
        const double cost = (double)ceil_div(units, resident) * longest;
        const bool fills = 4 * units >= 3 * resident;
        if (m == mlo || cost < best_cost || (cost == best_cost && (fills || !best_fills))) {
            best = lenm;
            best_cost = cost;
            best_fills = fills;
        }
    }
    slice_all(best);
    if (units > INT32_MAX) return -LIBSHALOM_NOT_SUPPORTED;
    if (units == 0) return 0;
    if (counters == 0) return T;
    int *cnt;
    float *slots;
    int st = fixup_space(dev, s, slot_bytes, &cnt, &slots);
    if (st) return -st;
    for (int i = 0; i < n; ++i) {
        BwdProblem &P = *probs[i];
        if (P.nslices <= 1) continue;
        P.counters = cnt;
        P.slots = slots;
        cnt += 2 * (size_t)(P.sO == 0 ? 1 : P.nimg) * P.ntiles;  // two ints per tile (splitk_arrive)
        slots += (size_t)P.units * T * T;
    }
    return T;
}

int backward_run(int prec, int ta, int tb, const float *dC, const float *A, const float *B, float *dA, float *dB,
                 long M, long N, long K, int nimg, long sA, long sB, long sC, hipStream_t s)
{
    int dev;
    DeviceInfo info;
    int st = current_device(&dev, &info);
    if (st) return st;
    BwdProblem X, Y;
    // dA (stored like A): dC op(B)^T, or op(B) dC^T when A is stored K x M
    if (!ta)
        bwd_side(X, dC, sC, (int)N, B, sB, (int)(tb ? K : N), dA, sA, (int)K, M, K, N, nimg);
    else
        bwd_side(X, B, sB, (int)(tb ? K : N), dC, sC, (int)N, dA, sA, (int)M, K, M, N, nimg);
    // dB (stored like B): op(A)^T dC, or dC^T op(A) when B is stored N x K
    if (!tb)
        bwd_side(Y, A, sA, (int)(ta ? M : K), dC, sC, (int)N, dB, sB, (int)N, K, N, M, nimg);
    else
        bwd_side(Y, dC, sC, (int)N, A, sA, (int)(ta ? M : K), dB, sB, (int)K, N, K, M, nimg);
    BwdProblem *probs[2] = {&X, &Y};
    const int T = plan_problems(probs, 2, std::min(std::min(X.R, X.Cc), std::min(Y.R, Y.Cc)) >= 128, dev,
                                info.num_cu, s);
    if (T <= 0) return -T;
    BwdLaunch fn = bwd_table().fn[prec][ta][tb][T == 128];
    if (!fn) return LIBSHALOM_NOT_SUPPORTED;
    X.xcd = Y.xcd = info.xcd_rr;
    return fn(X, Y, s) == hipSuccess ? LIBSHALOM_SUCCESS : LIBSHALOM_HIP_ERROR;
}

// C_i = op(A_i) op(B_i) over a strided batch; sC == 0 sums over the batch.
int strided_run(int prec, int ta, int tb, float *C, const float *A, const float *B, long M, long N, long K, int nimg,
                long sA, long sB, long sC, hipStream_t s)
{
    int dev;
    DeviceInfo info;
    int st = current_device(&dev, &info);
    if (st) return st;
    BwdProblem P;
    bwd_side(P, A, sA, (int)(ta ? M : K), B, sB, (int)(tb ? K : N), C, sC, (int)N, M, N, K, nimg);
    BwdProblem *probs[1] = {&P};
    // one problem: its 64 x 64 workgroups are light enough for four per CU
    const int T = plan_problems(probs, 1, std::min(M, N) >= 128, dev, info.num_cu, s, prec == P_FP32 ? 3 : 2, 4);
    if (T <= 0) return -T;
    StridedLaunch fn = bwd_table().strided[prec][ta][tb][T == 128];
    if (!fn) return LIBSHALOM_NOT_SUPPORTED;
    P.xcd = info.xcd_rr;
    return fn(P, s) == hipSuccess ? LIBSHALOM_SUCCESS : LIBSHALOM_HIP_ERROR;
}

// Output size of a convolution along one axis: (in + 2 pad - dil (f - 1) - 1) / stride + 1
// (floor division; <= 0 when the dilated filter does not fit).
long conv_out_size(long in, long f, long stride, long pad, long dil)
{
    const long span = in + 2 * pad - dil * (f - 1) - 1;
    return span < 0 ? 0 : span / stride + 1;
}

// ceil(2^32 / d) for the device's division by d (0 for d == 1).
unsigned div_magic(long d) { return d == 1 ? 0u : (unsigned)((((unsigned long long)1 << 32) + d - 1) / d); }

ConvGeom conv_geom(long C, long H, long W, long R, long S, long Q, long sh, long sw, long ph, long pw, long dh,
                   long dw, long th = 1, long tw = 1)
{
    ConvGeom g;
    g.th = (int)th;
    g.tw = (int)tw;
    g.mTh = div_magic(th);
    g.mTw = div_magic(tw);
    g.C = (int)C;
    g.H = (int)H;
    g.W = (int)W;
    g.R = (int)R;
    g.S = (int)S;
    g.RS = (int)(R * S);
    g.Q = (int)Q;
    g.sh = (int)sh;
    g.sw = (int)sw;
    g.ph = (int)ph;
    g.pw = (int)pw;
    g.dh = (int)dh;
    g.dw = (int)dw;
    g.mRS = div_magic(R * S);
    g.mS = div_magic(S);
    g.mQ = div_magic(Q);
    g.fold = 0;
    g.PQ = 0;
    g.img = 0;
    g.mPQ = 0;
    return g;
}

// Fold the batch into the pixel dimension of a convolution GEMM when the
// per-image pixel count leaves part of the last 64-wide tile empty and the
// folded problem keeps 32-bit offsets and exact divisions: the input batch
// (in_img elements per image) and the output (out_img per image) below 2 GiB,
// and batch * pixels * pixels below 2^32.
bool conv_fold(long batch, long pixels, long in_img, long out_img)
{
    return batch > 1 && pixels % 64 != 0 && 4.0 * batch * in_img < 2147483648.0 &&
           4.0 * batch * out_img < 2147483648.0 && (double)batch * pixels * pixels < 4294967296.0;
}

void fold_into(BwdProblem &P, ConvGeom &g, long batch, long pixels, long in_img, long out_img)
{
    g.fold = (int)batch;
    g.PQ = (int)pixels;
    g.img = (int)in_img;
    g.mPQ = div_magic(pixels);
    P.fold_total = (int)(batch * out_img);
    P.fold_mpq = g.mPQ;
}

// Wf[c][o][R-1-r][S-1-s] = W[o][c][r][s]: the filter of the input gradient
// (any stride), as a C x Cout*R*S matrix.
__global__ void k_conv_flip(const float *__restrict__ Wt, float *__restrict__ Wf, int Cout, int C, int RS)
{
    const long n = (long)Cout * C * RS;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long o = i / ((long)C * RS), rem = i - o * C * RS;
        const long c = rem / RS, t = rem - c * RS;
        Wf[(c * Cout + o) * RS + (RS - 1 - t)] = Wt[i];
    }
}

// Convolution backward as implicit GEMMs in one launch (either gradient may
// be skipped with a null pointer).  dX: the convolution of dY (Cout
// channels, P x Q) with the flipped, transposed filter, padding
// dil*(f-1) - pad (negative padding crops), which gives back H x W; for
// stride > 1 it is fractionally strided (a tap reads dY row
// (h - pad' + r*dil) / stride only when that division is exact).  dW: the
// sum over images of dY_i times the transposed im2col matrix of X_i.
int conv_bwd_run(int prec, float *dX, float *dW, const float *dY, const float *Wt, const float *X, long batch,
                 long C, long H, long W, long Cout, long R, long S, long sh, long sw, long ph, long pw, long dh,
                 long dw, long P, long Q, hipStream_t s)
{
    int dev;
    DeviceInfo info;
    int st = current_device(&dev, &info);
    if (st) return st;
    const long RS = R * S, K = C * RS, PQ = P * Q, HW = H * W;
    const int kalign = (int)(32 / std::gcd(32L, RS) * RS);
    BwdProblem X_, Y_;
    std::memset(&X_, 0, sizeof(X_));
    std::memset(&Y_, 0, sizeof(Y_));
    ConvGeom gx = conv_geom(Cout, P, Q, R, S, W, 1, 1, dh * (R - 1) - ph, dw * (S - 1) - pw, dh, dw, sh, sw);
    const ConvGeom gy = conv_geom(C, H, W, R, S, Q, sh, sw, ph, pw, dh, dw);
    if (dX) {
        void *wf;
        st = workspace(dev, s, (size_t)Cout * K * sizeof(float), &wf);
        if (st) return st;
        const long n = Cout * K;
        hipLaunchKernelGGL(k_conv_flip, dim3((unsigned)std::min(1024L, ceil_div(n, 256))), dim3(256), 0, s, Wt,
                           (float *)wf, (int)Cout, (int)C, (int)RS);
        if (hipGetLastError() != hipSuccess) return LIBSHALOM_HIP_ERROR;
        // the input gradient's pixels are the images' H x W, batch-folded when
        // the last tile of an image would be partly empty (on 3x3 layers,
        // batch 32: 14x14 pixels 8-21% faster, 7x7 BF16X3 18%, others within 2%)
        if (conv_fold(batch, HW, Cout * PQ, C * HW)) {
            bwd_side(X_, wf, 0, (int)(Cout * RS), dY, 0, 0, dX, 0, (int)HW, C, batch * HW, Cout * RS, 1);
            fold_into(X_, gx, batch, HW, Cout * PQ, C * HW);
        } else {
            bwd_side(X_, wf, 0, (int)(Cout * RS), dY, Cout * PQ, 0, dX, C * HW, (int)HW, C, HW, Cout * RS, (int)batch);
        }
        X_.kalign = kalign;
    }
    if (dW) bwd_side(Y_, dY, Cout * PQ, (int)PQ, X, C * HW, 0, dW, 0, (int)K, Cout, K, PQ, (int)batch);
    BwdProblem *probs[2];
    int n = 0;
    if (dX) probs[n++] = &X_;
    if (dW) probs[n++] = &Y_;
    long mind = LONG_MAX;
    for (int i = 0; i < n; ++i) mind = std::min(mind, (long)std::min(probs[i]->R, probs[i]->Cc));
    const int T = plan_problems(probs, n, mind >= 128, dev, info.num_cu, s, prec == P_FP32 ? 3 : 2);
    if (T <= 0) return -T;
    // dW's gather in 16-byte runs when a thread's four pixels (slice starts
    // are multiples of 32) always share an output row and read consecutive
    // input columns
    const bool runs = sw == 1 && Q % 4 == 0;
    ConvBwdLaunch fn = bwd_table().conv_bwd[prec][T == 128][runs];
    if (!fn) return LIBSHALOM_NOT_SUPPORTED;
    X_.xcd = Y_.xcd = info.xcd_rr;
    return fn(X_, Y_, gx, gy, s) == hipSuccess ? LIBSHALOM_SUCCESS : LIBSHALOM_HIP_ERROR;
}

// Convolution forward as an implicit GEMM over a batch of NCHW images: the
// same sliced tile launch as a strided batch whose B operand is each
// image's im2col matrix, gathered by the operand loads.  Slices of the
// reduction hold whole channels.
int conv_fwd_run(int prec, float *Y, const float *Wt, const float *X, long batch, long C, long H, long W, long Cout,
                 long R, long S, long sh, long sw, long ph, long pw, long dh, long dw, long P, long Q, hipStream_t s)
{
    int dev;
    DeviceInfo info;
    int st = current_device(&dev, &info);
    if (st) return st;
    const long RS = R * S, K = C * RS, PQ = P * Q;
    ConvGeom g = conv_geom(C, H, W, R, S, Q, sh, sw, ph, pw, dh, dw);
    BwdProblem Pb;
    // Not batch-folded: on 3x3 layers of 7x7 to 28x28 pixels, batch 32, the
    // folded forward measured within 2% of this one either way.
    bwd_side(Pb, Wt, 0, (int)K, X, C * H * W, 0, Y, Cout * PQ, (int)PQ, Cout, PQ, K, (int)batch);
    Pb.kalign = (int)(32 / std::gcd(32L, RS) * RS);
    BwdProblem *probs[1] = {&Pb};
    const int T = plan_problems(probs, 1, std::min(Cout, (long)Pb.Cc) >= 128, dev, info.num_cu, s,
                                prec == P_FP32 ? 3 : 2, 4);
    if (T <= 0) return -T;
    ConvLaunch fn = bwd_table().conv_fwd[prec][T == 128];
    if (!fn) return LIBSHALOM_NOT_SUPPORTED;
    Pb.xcd = info.xcd_rr;
    return fn(Pb, g, s) == hipSuccess ? LIBSHALOM_SUCCESS : LIBSHALOM_HIP_ERROR;
}

// ---------------------------------------------------------------------------
// The reference-API calls (synchronous, null stream).  The reference takes
// host pointers; here device pointers are used as they are, and operands in
// host memory that the GPU cannot address (pageable) are staged through a
// per-device device buffer: inputs copied in, C copied back when the
// product is done.  The copies go through kBufs pinned bounce buffers of
// kChunk bytes in rotation: the CPU fills (or empties) one while the DMA
// engine moves others, and one synchronisation ends the call.  (A pageable
// hipMemcpy per operand synchronises each time, and its device-to-host
// direction runs at a third of the pinned rate.)  Small operands share one
// chunk and travel with one copy; events are recorded only where a buffer
// comes round again within the call.  The staging buffers are held for the
// whole call.
// ---------------------------------------------------------------------------
constexpr size_t kChunk = (size_t)4 << 20;  // (4 x 4 MiB measured best of 1-8 MiB x 2-8 for 16 MiB)
constexpr int kBufs = 4;

struct Staging {
    char *dev = nullptr;  // device copy of the host operands
    size_t dev_size = 0;
    char *pin[kBufs] = {};          // pinned bounce buffers
    hipEvent_t done[kBufs] = {};    // the last DMA recorded on pin[i]
};
std::mutex g_stage_mu;
std::map<int, Staging> g_stage;

int hip_ok(hipError_t e)
{
    if (e == hipSuccess) return LIBSHALOM_SUCCESS;
    (void)hipGetLastError();
    return LIBSHALOM_HIP_ERROR;
}

int staging_ready(Staging &sg, size_t dev_bytes)
{
    for (int i = 0; i < kBufs; ++i) {
        if (sg.pin[i]) continue;
        if (hipHostMalloc((void **)&sg.pin[i], kChunk, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            sg.pin[i] = nullptr;
            return LIBSHALOM_ALLOC_FAILED;
        }
        if (int st = hip_ok(hipEventCreateWithFlags(&sg.done[i], hipEventDisableTiming))) return st;
    }
    if (sg.dev_size < dev_bytes) {
        if (sg.dev) (void)hipFree(sg.dev);  // the calls are synchronous: nothing uses it now
        sg.dev = nullptr;
        sg.dev_size = 0;
        if (hipMalloc((void **)&sg.dev, dev_bytes) != hipSuccess) {
            (void)hipGetLastError();
            sg.dev = nullptr;
            return LIBSHALOM_ALLOC_FAILED;
        }
        sg.dev_size = dev_bytes;
    }
    return LIBSHALOM_SUCCESS;
}

// Host -> device: the pieces (host source, bytes) land one after another
// from dst on.  A chunk may span pieces.  Every earlier call has finished
// (the calls end with a synchronisation), so a buffer needs a wait only
// when this transfer comes round to it again.
int stage_in(Staging &sg, const std::pair<const void *, size_t> *pieces, int n, char *dst, hipStream_t s)
{
    size_t total = 0;
    for (int i = 0; i < n; ++i) total += pieces[i].second;
    const long chunks = (long)((total + kChunk - 1) / kChunk);
    int p = 0;
    size_t in_piece = 0;  // bytes of piece p already taken
    for (long c = 0; c < chunks; ++c) {
        const int b = (int)(c % kBufs);
        if (c >= kBufs)
            if (int st = hip_ok(hipEventSynchronize(sg.done[b]))) return st;  // chunk c - kBufs has left
        size_t fill = 0;
        while (p < n && fill < kChunk) {
            const size_t take = std::min(pieces[p].second - in_piece, kChunk - fill);
            std::memcpy(sg.pin[b] + fill, (const char *)pieces[p].first + in_piece, take);
            fill += take;
            in_piece += take;
            if (in_piece == pieces[p].second) {
                ++p;
                in_piece = 0;
            }
        }
        if (int st = hip_ok(hipMemcpyAsync(dst, sg.pin[b], fill, hipMemcpyHostToDevice, s))) return st;
        if (c + kBufs < chunks)
            if (int st = hip_ok(hipEventRecord(sg.done[b], s))) return st;
        dst += fill;
    }
    return LIBSHALOM_SUCCESS;
}

// Device -> host after the work queued on s: up to kBufs chunks in flight
// while the CPU copies out the oldest; the last one is waited for with a
// stream synchronisation (cheaper than an event here).
int stage_out(Staging &sg, void *dst, const char *src, size_t bytes, hipStream_t s)
{
    const long chunks = (long)((bytes + kChunk - 1) / kChunk);
    long issued = 0;
    for (long r = 0; r < chunks; ++r) {
        for (; issued < chunks && issued < r + kBufs; ++issued) {  // chunk issued - kBufs is out
            const int b = (int)(issued % kBufs);
            const size_t off = (size_t)issued * kChunk, len = std::min(kChunk, bytes - off);
            if (int st = hip_ok(hipMemcpyAsync(sg.pin[b], src + off, len, hipMemcpyDeviceToHost, s))) return st;
            if (issued + 1 < chunks)
                if (int st = hip_ok(hipEventRecord(sg.done[b], s))) return st;
        }
        const int b = (int)(r % kBufs);
        if (int st = hip_ok(r + 1 < chunks ? hipEventSynchronize(sg.done[b]) : hipStreamSynchronize(s))) return st;
        const size_t off = (size_t)r * kChunk;
        std::memcpy((char *)dst + off, sg.pin[b], std::min(kChunk, bytes - off));
    }
    return LIBSHALOM_SUCCESS;
}

template <typename T>
int reference_call(int prec, int ta, int tb, T *C, T *A, T *B, long M, long N, long K, bool mp)
{
    int st = check_dims(ta, tb, M, N, K);
    if (!st) st = check_ptrs(C, A, B, M, N, K, false);
    if (st || M == 0 || N == 0) return st;
    int dev;
    DeviceInfo info;
    st = current_device(&dev, &info);
    if (st) return st;
    const bool hC = !device_accessible(C);
    const bool hA = K > 0 && !device_accessible(A);
    const bool hB = K > 0 && !device_accessible(B);
    if (!(hC || hA || hB)) {
        st = run_single(prec, ta, tb, C, A, B, M, N, K, mp, nullptr);
        return st ? st : sync_null();
    }
    const size_t nA = hA ? (size_t)M * K * sizeof(T) : 0;
    const size_t nB = hB ? (size_t)K * N * sizeof(T) : 0;
    const size_t nC = hC ? (size_t)M * N * sizeof(T) : 0;
    const size_t offC = (nA + nB + 255) & ~(size_t)255;  // A and B adjacent: one copy when small
    std::lock_guard<std::mutex> lk(g_stage_mu);
    Staging &sg = g_stage[dev];
    st = staging_ready(sg, offC + nC);
    if (st) return st;
    // A's bytes are a multiple of 4 (of 8 for FP64): B stays aligned after it
    T *dA = hA ? (T *)sg.dev : A;
    T *dB = hB ? (T *)(sg.dev + nA) : B;
    T *dC = hC ? (T *)(sg.dev + offC) : C;
    std::pair<const void *, size_t> in[2];
    int n_in = 0;
    if (hA) in[n_in++] = {A, nA};
    if (hB) in[n_in++] = {B, nB};
    st = stage_in(sg, in, n_in, sg.dev, nullptr);
    if (!st) st = run_single(prec, ta, tb, dC, dA, dB, M, N, K, mp, nullptr);
    if (!st && hC) st = stage_out(sg, C, (const char *)dC, nC, nullptr);
    // also after an error: the next call finds the bounce buffers idle
    const int st_sync = sync_null();
    return st ? st : st_sync;
}

int sgemm_sync(int ta, int tb, float *C, float *A, float *B, long M, long N, long K, bool mp)
{
    return set_status(reference_call<float>(kprec(g_precision.load()), ta, tb, C, A, B, M, N, K, mp));
}

}  // namespace

// ===========================================================================
// Batched plans.
// ===========================================================================
struct LibShalom_plan_s {
    int device = 0;
    int prec = 0;
    void *dev_buf = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double flops = 0, bytes = 0;
    long items = 0;
};

namespace {

// How a batched plan executes.
enum BatchMode : int {
    BM_PERSIST = 0,  // one persistent launch per tile bucket, jobs streamed across
    BM_SINGLE = 1,   // every GEMM in the largest square-tile bucket: one persistent launch
};

int batch_mode_from_env()
{
    const char *m = std::getenv("LIBSHALOM_BATCHED_MODE");
    if (m && !std::strcmp(m, "single")) return BM_SINGLE;
    return BM_PERSIST;
}

// A tile of a GEMM as a GEMM of its own: pointers offset to the tile, M and N
// cut to the tile.
GemmDesc tile_job(int ta, int tb, const GemmDesc &g, int m0, int n0, int bm, int bn, int esize)
{
    GemmDesc j = g;
    const char *A = (const char *)g.A, *B = (const char *)g.B;
    char *C = (char *)g.C;
    if (g.K > 0) {
        j.A = A + (size_t)esize * (ta ? (size_t)m0 : (size_t)m0 * g.lda);
        j.B = B + (size_t)esize * (tb ? (size_t)n0 * g.ldb : (size_t)n0);
    }
    j.C = C + (size_t)esize * ((size_t)m0 * g.ldc + n0);
    j.M = std::min(bm, g.M - m0);
    j.N = std::min(bn, g.N - n0);
    return j;
}

// Relative cost of a job for ordering and for apportioning workgroups: the
// kernels are bound by moving operands, so count elements moved.
double job_cost(const GemmDesc &d) { return (double)(d.M + d.N) * std::max(d.K, 1) + (double)d.M * d.N; }

// Order a bucket's jobs for a persistent launch of `grid` workgroups, where
// workgroup b runs jobs b, b + grid, b + 2*grid, ...  A job's time is close to
// its number of K chunks (each chunk is one step of the tile pipeline, the
// same work whatever the tile's fill) plus about two thirds of a step for
// writing its C tile.  Jobs are sorted by that estimate, longest first, and
// dealt out in boustrophedon order (rounds of `grid` jobs, alternate rounds
// reversed), so every workgroup's total is within about one job of the
// others; plain round-robin over the sorted list gives workgroup 0 the
// largest job of every round.
void order_jobs(std::vector<GemmDesc> &jobs, int grid, int bk)
{
    auto est = [bk](const GemmDesc &d) { return (double)std::max(1, (d.K + bk - 1) / bk) + 0.65; };
    std::stable_sort(jobs.begin(), jobs.end(), [&](const GemmDesc &x, const GemmDesc &y) {
        const double ex = est(x), ey = est(y);
        return ex != ey ? ex > ey : job_cost(x) > job_cost(y);
    });
    if (grid <= 1) return;
    const size_t G = (size_t)grid;
    for (size_t r0 = G; r0 + G <= jobs.size(); r0 += 2 * G) std::reverse(jobs.begin() + r0, jobs.begin() + r0 + G);
}

constexpr int kTile16 = -16;  // bucket tile id of the FP64 16 x 16 persistent kernel

struct Bucket {
    int tile;
    std::vector<GemmDesc> jobs;
    int grid = 0;
    size_t off = 0;  // byte offset of the bucket's jobs in the plan buffer
};

// ---------------------------------------------------------------------------
// Autotuner for the batched tile buckets.  The table maps (precision,
// transa, transb, shape class) to a tile; shape classes are ranges of M, N
// and K.  Tuning works on a whole batch: the GEMMs are grouped by
// (max(M, N) range, K range), and, group by group in order of work, each
// other candidate square tile (32, 64, 128; FP64: 32, 64) is tried for the
// group while the rest keep their current tiles; a candidate is kept when
// the batch -- every bucket launched as a plan would launch it, on synthetic
// operands in a scratch buffer -- runs at least 1% faster.  Timing the whole
// batch keeps what the buckets share (the chip, L2, the tail of each launch)
// in the measurement.  Plans built afterwards use the table; classes never
// tuned keep the fixed rule.  LIBSHALOM_TUNE_FILE names a text file the
// table is read from at first use and appended to after each tuning run.
// ---------------------------------------------------------------------------
constexpr int kDimEdges[] = {16, 32, 48, 64, 96, 128};  // M and N classes: <= edge, then > 128
constexpr int kKEdges[] = {16, 32, 64, 128};            // K classes
int dim_class(long v, const int *edges, int n)
{
    for (int i = 0; i < n; ++i)
        if (v <= edges[i]) return i;
    return n;
}
int shape_class(long M, long N, long K)
{
    const int nd = 7, nk = 5;
    return (dim_class(M, kDimEdges, 6) * nd + dim_class(N, kDimEdges, 6)) * nk + dim_class(K, kKEdges, 4);
}
int tune_group(long M, long N, long K) { return dim_class(std::max(M, N), kDimEdges, 6) * 5 + dim_class(K, kKEdges, 4); }
long tune_key(int prec, int ta, int tb, int cls) { return (((long)prec * 2 + ta) * 2 + tb) * 1000 + cls; }

std::mutex g_tune_mu;
std::map<long, int> g_tune;  // key -> TileId
bool g_tune_loaded = false;

void tune_load_locked()
{
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const char *path = std::getenv("LIBSHALOM_TUNE_FILE");
    if (!path || !path[0]) return;
    FILE *f = std::fopen(path, "r");
    if (!f) return;
    // lines "prec ta tb class tile"; a line naming a tile without a
    // persistent launcher for its mode and transposes (or out of range) is
    // ignored, so that a plan never meets a missing kernel
    const LaunchTable &lt = launch_table();
    int prec, ta, tb, cls, tile;
    while (std::fscanf(f, "%d %d %d %d %d", &prec, &ta, &tb, &cls, &tile) == 5)
        if (prec >= 0 && prec < NPREC && (ta | tb) >= 0 && ta < 2 && tb < 2 && tile >= 0 && tile < NTILE &&
            cls >= 0 && cls < 1000 && lt.persist[prec][ta][tb][tile] != nullptr)
            g_tune[tune_key(prec, ta, tb, cls)] = tile;
    std::fclose(f);
}

int tuned_tile(int prec, int ta, int tb, long M, long N, long K)
{
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_load_locked();
        auto it = g_tune.find(tune_key(prec, ta, tb, shape_class(M, N, K)));
        if (it != g_tune.end()) return it->second;
    }
    return batched_tile(prec, M, N, K);
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

// Time the batch `descs` with tile assign[i] for GEMM i, buckets launched as
// a plan launches them (ms per execution, best of `reps`).
int time_assignment(int prec, int ta, int tb, const std::vector<GemmDesc> &descs, const std::vector<int> &assign,
                    int esize, int num_cu, hipStream_t s, DevBuf &jobs_dev, size_t &jobs_cap, hipEvent_t e0,
                    hipEvent_t e1, float *ms_out)
{
    std::vector<std::vector<GemmDesc>> b(NTILE);
    for (size_t i = 0; i < descs.size(); ++i) {
        const GemmDesc &d = descs[i];
        const TileShape ts = kTileShapes[assign[i]];
        for (int m0 = 0; m0 < d.M; m0 += ts.bm)
            for (int n0 = 0; n0 < d.N; n0 += ts.bn)
                b[assign[i]].push_back(tile_job(ta, tb, d, m0, n0, ts.bm, ts.bn, esize));
    }
    size_t total = 0;
    std::vector<size_t> off(NTILE, 0);
    for (int t = 0; t < NTILE; ++t) {
        PersistOccupancy occ = launch_table().persist_occ[prec][ta][tb][t];
        const long slots = (long)std::max(1, occ ? occ() : 1) * num_cu;
        order_jobs(b[t], (int)std::min<long>((long)b[t].size(), slots), esize == 8 ? 16 : 32);
        off[t] = total;
        total += ((b[t].size() * sizeof(GemmDesc)) + 255) & ~(size_t)255;
    }
    if (hipStreamSynchronize(s) != hipSuccess) return LIBSHALOM_HIP_ERROR;
    if (total > jobs_cap) {
        if (jobs_dev.p) (void)hipFree(jobs_dev.p);
        jobs_dev.p = nullptr;
        if (hipMalloc(&jobs_dev.p, total) != hipSuccess) {
            (void)hipGetLastError();
            jobs_cap = 0;
            return LIBSHALOM_ALLOC_FAILED;
        }
        jobs_cap = total;
    }
    std::vector<char> host(total);
    for (int t = 0; t < NTILE; ++t)
        if (!b[t].empty()) std::memcpy(host.data() + off[t], b[t].data(), b[t].size() * sizeof(GemmDesc));
    if (total && hipMemcpy(jobs_dev.p, host.data(), total, hipMemcpyHostToDevice) != hipSuccess)
        return LIBSHALOM_HIP_ERROR;
    auto run = [&]() {
        const int order[NTILE] = {T128x128, T128x64, T128x32, T32x128, T64x64, T32x32};
        for (int oi = 0; oi < NTILE; ++oi) {
            const int t = order[oi];
            if (b[t].empty()) continue;
            PersistLaunch fn = launch_table().persist[prec][ta][tb][t];
            PersistOccupancy occ = launch_table().persist_occ[prec][ta][tb][t];
            if (!fn) return false;
            const long nj = (long)b[t].size();
            const int grid = (int)std::min<long>(nj, (long)std::max(1, occ ? occ() : 1) * num_cu);
            if (fn((const GemmDesc *)((char *)jobs_dev.p + off[t]), (int)nj, grid, s) != hipSuccess) return false;
        }
        return true;
    };
    if (!run()) return LIBSHALOM_HIP_ERROR;  // warm-up
    float best = 1e30f;
    for (int r = 0; r < 3; ++r) {
        if (hipEventRecord(e0, s) != hipSuccess) return LIBSHALOM_HIP_ERROR;
        if (!run()) return LIBSHALOM_HIP_ERROR;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return LIBSHALOM_HIP_ERROR;
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return LIBSHALOM_HIP_ERROR;
        best = std::min(best, ms);
    }
    *ms_out = best;
    return LIBSHALOM_SUCCESS;
}

int autotune(int prec, int ta, int tb, const long *M, const long *N, const long *K, long count, hipStream_t s,
             int *classes_tuned)
{
    int dev;
    DeviceInfo info;
    int st = current_device(&dev, &info);
    if (st) return st;
    const bool f64 = prec == P_FP64;
    const int esize = f64 ? 8 : 4;
    // the batch (at most 64 K GEMMs and 4 GiB of operands: an even sample of larger ones)
    size_t bytes_all = 0;
    long nonempty = 0;
    for (long g = 0; g < count; ++g) {
        st = check_dims(ta, tb, M[g], N[g], K[g]);
        if (st) return st;
        if (M[g] == 0 || N[g] == 0) continue;
        bytes_all += (size_t)esize * (M[g] * K[g] + K[g] * N[g] + M[g] * N[g]) + 64;
        ++nonempty;
    }
    if (nonempty == 0) return LIBSHALOM_SUCCESS;
    const double keep = std::min(1.0, std::min(65536.0 / nonempty, (double)((size_t)4 << 30) / bytes_all));
    std::vector<long> sel;
    double acc = 0;
    for (long g = 0; g < count; ++g) {
        if (M[g] == 0 || N[g] == 0) continue;
        acc += keep;
        if (acc >= 1.0) {
            acc -= 1.0;
            sel.push_back(g);
        }
    }
    size_t need = 0;
    for (long g : sel) need += (size_t)esize * (M[g] * K[g] + K[g] * N[g] + M[g] * N[g]) + 64;
    DevBuf scratch, jobs_dev;
    if (hipMalloc(&scratch.p, need) != hipSuccess) {
        (void)hipGetLastError();
        return LIBSHALOM_ALLOC_FAILED;
    }
    if (hipMemsetAsync(scratch.p, 0, need, s) != hipSuccess) return LIBSHALOM_HIP_ERROR;
    std::vector<GemmDesc> descs;
    std::vector<int> group, assign;
    char *base = (char *)scratch.p;
    for (long g : sel) {
        const void *A = base;
        base += (size_t)esize * M[g] * K[g];
        const void *B = base;
        base += (size_t)esize * K[g] * N[g];
        void *C = base;
        base += (size_t)esize * M[g] * N[g] + 64;
        descs.push_back(make_desc(ta, tb, A, B, C, M[g], N[g], K[g], esize));
        group.push_back(tune_group(M[g], N[g], K[g]));
        assign.push_back(tuned_tile(prec, ta, tb, M[g], N[g], K[g]));
    }
    // groups by work
    std::map<int, double> gwork;
    for (size_t i = 0; i < descs.size(); ++i) gwork[group[i]] += job_cost(descs[i]);
    std::vector<std::pair<double, int>> order;
    for (auto &kv : gwork) order.push_back({-kv.second, kv.first});
    std::sort(order.begin(), order.end());

    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return LIBSHALOM_HIP_ERROR;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return LIBSHALOM_HIP_ERROR;
    }
    size_t jobs_cap = 0;
    float best = 0;
    st = time_assignment(prec, ta, tb, descs, assign, esize, info.num_cu, s, jobs_dev, jobs_cap, e0, e1, &best);
    const int cands[3] = {T32x32, T64x64, T128x128};
    const int ncand = f64 ? 2 : 3;
    for (size_t oi = 0; oi < order.size() && !st; ++oi) {
        const int g = order[oi].second;
        for (int ci = 0; ci < ncand && !st; ++ci) {
            std::vector<int> trial = assign;
            bool changed = false;
            for (size_t i = 0; i < descs.size(); ++i)
                if (group[i] == g && trial[i] != cands[ci]) {
                    trial[i] = cands[ci];
                    changed = true;
                }
            if (!changed) continue;
            float ms = 0;
            st = time_assignment(prec, ta, tb, descs, trial, esize, info.num_cu, s, jobs_dev, jobs_cap, e0, e1, &ms);
            if (!st && ms < 0.99f * best) {
                best = ms;
                assign.swap(trial);
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) return st;
    if (hipStreamSynchronize(s) != hipSuccess) return LIBSHALOM_HIP_ERROR;
    // record the tile of every shape class seen (all GEMMs of a group share it)
    std::map<long, int> found;
    for (size_t i = 0; i < descs.size(); ++i)
        found[tune_key(prec, ta, tb, shape_class(descs[i].M, descs[i].N, descs[i].K))] = assign[i];
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_load_locked();
        for (auto &f : found) g_tune[f.first] = f.second;
        const char *path = std::getenv("LIBSHALOM_TUNE_FILE");
        if (path && path[0] && !found.empty()) {
            if (FILE *fp = std::fopen(path, "a")) {
                for (auto &f : found) {
                    const long key = f.first;
                    const int cls = (int)(key % 1000), rest = (int)(key / 1000);
                    std::fprintf(fp, "%d %d %d %d %d\n", rest / 4, (rest / 2) % 2, rest % 2, cls, f.second);
                }
                std::fclose(fp);
            }
        }
    }
    if (classes_tuned) *classes_tuned = (int)found.size();
    return LIBSHALOM_SUCCESS;
}

void tune_reset()
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.clear();
    g_tune_loaded = true;  // a reset table stays empty (no reload from the file)
}

int build_plan(LibShalom_plan_t *out, int prec, int ta, int tb, void *const *C, const void *const *A,
               const void *const *B, const long *M, const long *N, const long *K, long count)
{
    if (!out) return LIBSHALOM_INVALID_POINTER;
    *out = nullptr;
    if (count < 0) return LIBSHALOM_INVALID_VALUE;
    if (count > 0 && (!C || !A || !B || !M || !N || !K)) return LIBSHALOM_INVALID_POINTER;
    if ((ta != 0 && ta != 1) || (tb != 0 && tb != 1)) return LIBSHALOM_INVALID_VALUE;
    // every GEMM's arguments, before the device is touched
    for (long g = 0; g < count; ++g) {
        int st = check_dims(ta, tb, M[g], N[g], K[g]);
        if (!st) st = check_ptrs(C[g], A[g], B[g], M[g], N[g], K[g], false);
        if (st) return st;
    }
    int dev;
    DeviceInfo info;
    int st = current_device(&dev, &info);
    if (st) return st;
    const bool f64 = prec == P_FP64;
    const int esize = f64 ? 8 : 4;
    const int mode = batch_mode_from_env();
    const int single_tile = f64 ? T64x64 : T128x128;

    std::vector<Bucket> buckets(NTILE);
    for (int t = 0; t < NTILE; ++t) buckets[t].tile = t;
    double flops = 0, bytes = 0;
    for (long g = 0; g < count; ++g) {
        if (M[g] == 0 || N[g] == 0) continue;
        flops += 2.0 * M[g] * N[g] * K[g];
        bytes += (double)esize * (M[g] * K[g] + K[g] * N[g] + M[g] * N[g]);
        const int tile = mode == BM_SINGLE ? single_tile : tuned_tile(prec, ta, tb, M[g], N[g], K[g]);
        const GemmDesc d = make_desc(ta, tb, A[g], B[g], C[g], M[g], N[g], K[g], esize);
        Bucket &b = buckets[tile];
        const TileShape ts = kTileShapes[tile];
        for (int m0 = 0; m0 < d.M; m0 += ts.bm)
            for (int n0 = 0; n0 < d.N; n0 += ts.bn) b.jobs.push_back(tile_job(ta, tb, d, m0, n0, ts.bm, ts.bn, esize));
    }
    // FP64: the GEMMs of the 32 x 32 bucket with M, N <= 16 go to one-wave
    // 16 x 16 tiles (bucket index NTILE, tile id kTile16).
    buckets.emplace_back();
    buckets[NTILE].tile = kTile16;
    if (f64 && mode != BM_SINGLE && launch_table().persist16[ta][tb]) {
        std::vector<GemmDesc> keep;
        for (const GemmDesc &j : buckets[T32x32].jobs)
            (j.M <= 16 && j.N <= 16 ? buckets[NTILE].jobs : keep).push_back(j);
        buckets[T32x32].jobs.swap(keep);
    }
    // Per bucket: as many workgroups as can be resident, jobs dealt so that
    // their totals balance.
    long total_jobs = 0;
    size_t total = 0;
    for (auto &b : buckets) {
        const long nj = (long)b.jobs.size();
        if (nj == 0) continue;
        if (nj > (long)INT32_MAX / 2) return LIBSHALOM_NOT_SUPPORTED;
        total_jobs += nj;
        PersistOccupancy occ = b.tile == kTile16 ? launch_table().persist16_occ[ta][tb]
                                                 : launch_table().persist_occ[prec][ta][tb][b.tile];
        const long slots = (long)std::max(1, occ ? occ() : 1) * info.num_cu;
        b.grid = (int)std::min(nj, slots);
        order_jobs(b.jobs, b.grid, f64 ? 16 : 32);
        b.off = total;
        total += ((size_t)nj * sizeof(GemmDesc) + 255) & ~(size_t)255;
    }

    LibShalom_plan_s *plan = new (std::nothrow) LibShalom_plan_s();
    if (!plan) return LIBSHALOM_ALLOC_FAILED;
    plan->device = dev;
    plan->prec = prec;
    plan->flops = flops;
    plan->bytes = bytes;
    plan->items = total_jobs;

    hipStream_t cs = nullptr;
    auto fail = [&](int code) {
        if (cs) (void)hipStreamDestroy(cs);
        LibShalom_plan_destroy(plan);
        (void)hipGetLastError();
        return code;
    };
    if (total > 0) {
        if (hipMalloc(&plan->dev_buf, total) != hipSuccess) {
            plan->dev_buf = nullptr;
            return fail(LIBSHALOM_ALLOC_FAILED);
        }
        std::vector<char> host(total);
        for (auto &b : buckets)
            if (!b.jobs.empty()) std::memcpy(host.data() + b.off, b.jobs.data(), b.jobs.size() * sizeof(GemmDesc));
        if (hipMemcpy(plan->dev_buf, host.data(), total, hipMemcpyHostToDevice) != hipSuccess)
            return fail(LIBSHALOM_HIP_ERROR);
    }
    // Capture the bucket launches into one graph: the largest tile's bucket on
    // the capture stream, every other bucket on a branch of its own forked
    // from the start and joined at the end, so that the small buckets' kernels
    // start on compute units the large one's last workgroups leave idle
    // (measured: 2.8% less time per step for the 50k BF16X3 batch, 4% for
    // BF16X6).  Streams keep the default priority.
    constexpr int NB = NTILE + 1;  // the tile buckets and the FP64 16 x 16 one
    const int order[NB] = {T128x128, T128x64, T128x32, T32x128, T64x64, T32x32, NTILE};
    hipStream_t fs[NB] = {};
    hipEvent_t ev[NB + 1] = {};
    auto release = [&]() {
        for (int i = 0; i < NB; ++i)
            if (fs[i]) (void)hipStreamDestroy(fs[i]);
        for (int i = 0; i <= NB; ++i)
            if (ev[i]) (void)hipEventDestroy(ev[i]);
    };
    bool launch_ok = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) == hipSuccess &&
                     hipEventCreateWithFlags(&ev[NB], hipEventDisableTiming) == hipSuccess;
    int lead = -1;  // the bucket on the capture stream
    for (int oi = 0; oi < NB && launch_ok; ++oi) {
        if (buckets[order[oi]].grid == 0) continue;
        if (lead < 0) {
            lead = oi;
            continue;
        }
        launch_ok = hipStreamCreateWithFlags(&fs[oi], hipStreamNonBlocking) == hipSuccess &&
                    hipEventCreateWithFlags(&ev[oi], hipEventDisableTiming) == hipSuccess;
    }
    if (!launch_ok || hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        release();
        return fail(LIBSHALOM_HIP_ERROR);
    }
    // the fork point (only when there is a bucket to fork)
    bool forked = false;
    for (int oi = 0; oi < NB; ++oi) forked |= fs[oi] != nullptr;
    if (forked) launch_ok = hipEventRecord(ev[NB], cs) == hipSuccess;
    for (int oi = 0; oi < NB && launch_ok; ++oi) {
        const Bucket &b = buckets[order[oi]];
        if (b.grid == 0) continue;
        const GemmDesc *jobs = (const GemmDesc *)((char *)plan->dev_buf + b.off);
        PersistLaunch fn = b.tile == kTile16 ? launch_table().persist16[ta][tb]
                                             : launch_table().persist[prec][ta][tb][b.tile];
        const hipStream_t bs = oi == lead ? cs : fs[oi];
        if (oi != lead) launch_ok = hipStreamWaitEvent(bs, ev[NB], 0) == hipSuccess;
        if (!launch_ok || !fn || fn(jobs, (int)b.jobs.size(), b.grid, bs) != hipSuccess) launch_ok = false;
        if (launch_ok && oi != lead)
            launch_ok = hipEventRecord(ev[oi], bs) == hipSuccess && hipStreamWaitEvent(cs, ev[oi], 0) == hipSuccess;
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(cs, &g);
    release();
    plan->graph = g;
    if (!launch_ok || e != hipSuccess) return fail(LIBSHALOM_HIP_ERROR);
    if (hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0) != hipSuccess) {
        plan->exec = nullptr;
        return fail(LIBSHALOM_HIP_ERROR);
    }
    (void)hipStreamDestroy(cs);
    *out = plan;
    return LIBSHALOM_SUCCESS;
}

}  // namespace

// ===========================================================================
// extern "C" entry points.
// ===========================================================================
extern "C" {

void LibShalom_sgemm(int transa, int transb, float *C, float *A, float *B, long M, long N, long K)
{
    sgemm_sync(transa, transb, C, A, B, M, N, K, false);
}

void LibShalom_sgemm_mp(int transa, int transb, float *C, float *A, float *B, long M, long N, long K)
{
    sgemm_sync(transa, transb, C, A, B, M, N, K, true);
}

void LibShalom_dgemm(int transa, int transb, double *C, double *A, double *B, long M, long N, long K)
{
    set_status(reference_call<double>(P_FP64, transa, transb, C, A, B, M, N, K, false));
}

void LibShalom_set_thread_nums(int num)
{
    g_thread_nums.store(num);
    set_status(LIBSHALOM_SUCCESS);
}

void LibShalom_set_thread_ways(int tm, int tn)
{
    g_tm.store(tm);
    g_tn.store(tn);
    g_thread_nums.store(tm * tn);
    set_status(LIBSHALOM_SUCCESS);
}

void LibShalom_get_thread_config(int *num, int *tm, int *tn)
{
    if (num) *num = g_thread_nums.load();
    if (tm) *tm = g_tm.load();
    if (tn) *tn = g_tn.load();
}

int LibShalom_set_precision(int precision)
{
    if (!valid_precision(precision)) return set_status(LIBSHALOM_INVALID_VALUE);
    g_precision.store(precision);
    return set_status(LIBSHALOM_SUCCESS);
}

int LibShalom_get_precision(void) { return g_precision.load(); }

int LibShalom_get_last_error(void) { return g_last_status; }

const char *LibShalom_status_string(int status)
{
    switch (status) {
    case LIBSHALOM_SUCCESS: return "success";
    case LIBSHALOM_INVALID_VALUE: return "invalid value";
    case LIBSHALOM_INVALID_POINTER: return "invalid pointer";
    case LIBSHALOM_HIP_ERROR: return "HIP runtime error";
    case LIBSHALOM_ALLOC_FAILED: return "workspace allocation failed";
    case LIBSHALOM_NOT_SUPPORTED: return "not supported";
    case LIBSHALOM_NO_DEVICE: return "no gfx950 device";
    default: return "unknown status";
    }
}

int LibShalom_sgemm_ex(int transa, int transb, float *C, const float *A, const float *B, long M, long N, long K,
                       int precision, LibShalom_stream_t stream)
{
    int st = valid_precision(precision) ? LIBSHALOM_SUCCESS : LIBSHALOM_INVALID_VALUE;
    if (!st) st = check_dims(transa, transb, M, N, K);
    if (!st) st = check_ptrs(C, A, B, M, N, K, false);
    if (!st) st = run_single(kprec(precision), transa, transb, C, A, B, M, N, K, false, (hipStream_t)stream);
    return set_status(st);
}

int LibShalom_sgemm_mp_ex(int transa, int transb, float *C, const float *A, const float *B, long M, long N,
                          long K, int precision, LibShalom_stream_t stream)
{
    int st = valid_precision(precision) ? LIBSHALOM_SUCCESS : LIBSHALOM_INVALID_VALUE;
    if (!st) st = check_dims(transa, transb, M, N, K);
    if (!st) st = check_ptrs(C, A, B, M, N, K, false);
    if (!st) st = run_single(kprec(precision), transa, transb, C, A, B, M, N, K, true, (hipStream_t)stream);
    return set_status(st);
}

int LibShalom_dgemm_ex(int transa, int transb, double *C, const double *A, const double *B, long M, long N, long K,
                       LibShalom_stream_t stream)
{
    int st = check_dims(transa, transb, M, N, K);
    if (!st) st = check_ptrs(C, A, B, M, N, K, false);
    if (!st) st = run_single(P_FP64, transa, transb, C, A, B, M, N, K, false, (hipStream_t)stream);
    return set_status(st);
}

int LibShalom_sgemm_general(int transa, int transb, long M, long N, long K, float alpha, const float *A, long lda,
                            const float *B, long ldb, float beta, float *C, long ldc, int precision,
                            LibShalom_stream_t stream)
{
    int st = valid_precision(precision) ? LIBSHALOM_SUCCESS : LIBSHALOM_INVALID_VALUE;
    if (!st) st = check_dims(transa, transb, M, N, K);
    if (!st) st = check_lds(transa, transb, M, N, K, lda, ldb, ldc);
    // alpha == 0: C = beta * C without reading A or B (as BLAS), which may then be null
    const long Kr = alpha == 0 ? 0 : K;
    if (!st) st = check_ptrs(C, A, B, M, N, Kr, false);
    if (!st)
        st = run_general(kprec(precision), transa, transb, C, A, B, M, N, Kr, lda, ldb, ldc, alpha, beta, false,
                         (hipStream_t)stream);
    return set_status(st);
}

int LibShalom_sgemm_general_noalloc(int transa, int transb, long M, long N, long K, float alpha, const float *A,
                                    long lda, const float *B, long ldb, float beta, float *C, long ldc, int precision,
                                    LibShalom_stream_t stream)
{
    int st = valid_precision(precision) ? LIBSHALOM_SUCCESS : LIBSHALOM_INVALID_VALUE;
    if (!st) st = check_dims(transa, transb, M, N, K);
    if (!st) st = check_lds(transa, transb, M, N, K, lda, ldb, ldc);
    // alpha == 0: C = beta * C without reading A or B (as BLAS), which may then be null
    const long Kr = alpha == 0 ? 0 : K;
    if (!st) st = check_ptrs(C, A, B, M, N, Kr, false);
    if (!st)
        st = run_general(kprec(precision), transa, transb, C, A, B, M, N, Kr, lda, ldb, ldc, alpha, beta, false,
                         (hipStream_t)stream, true);
    return set_status(st);
}

int LibShalom_dgemm_general(int transa, int transb, long M, long N, long K, double alpha, const double *A, long lda,
                            const double *B, long ldb, double beta, double *C, long ldc, LibShalom_stream_t stream)
{
    int st = check_dims(transa, transb, M, N, K);
    if (!st) st = check_lds(transa, transb, M, N, K, lda, ldb, ldc);
    // alpha == 0: C = beta * C without reading A or B (as BLAS), which may then be null
    const long Kr = alpha == 0 ? 0 : K;
    if (!st) st = check_ptrs(C, A, B, M, N, Kr, false);
    if (!st)
        st = run_general(P_FP64, transa, transb, C, A, B, M, N, Kr, lda, ldb, ldc, alpha, beta, false,
                         (hipStream_t)stream);
    return set_status(st);
}

int LibShalom_sgemm_batched_plan(LibShalom_plan_t *plan, int transa, int transb, float *const *C,
                                 const float *const *A, const float *const *B, const long *M, const long *N,
                                 const long *K, long count, int precision)
{
    if (!valid_precision(precision)) return set_status(LIBSHALOM_INVALID_VALUE);
    return set_status(build_plan(plan, kprec(precision), transa, transb, (void *const *)C, (const void *const *)A,
                                 (const void *const *)B, M, N, K, count));
}

int LibShalom_dgemm_batched_plan(LibShalom_plan_t *plan, int transa, int transb, double *const *C,
                                 const double *const *A, const double *const *B, const long *M, const long *N,
                                 const long *K, long count)
{
    return set_status(build_plan(plan, P_FP64, transa, transb, (void *const *)C, (const void *const *)A,
                                 (const void *const *)B, M, N, K, count));
}

int LibShalom_plan_execute(LibShalom_plan_t plan, LibShalom_stream_t stream)
{
    if (!plan || !plan->exec) return set_status(LIBSHALOM_INVALID_POINTER);
    if (hipGraphLaunch(plan->exec, (hipStream_t)stream) != hipSuccess) {
        (void)hipGetLastError();
        return set_status(LIBSHALOM_HIP_ERROR);
    }
    return set_status(LIBSHALOM_SUCCESS);
}

int LibShalom_plan_destroy(LibShalom_plan_t plan)
{
    if (!plan) return set_status(LIBSHALOM_SUCCESS);
    if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
    if (plan->graph) (void)hipGraphDestroy(plan->graph);
    if (plan->dev_buf) (void)hipFree(plan->dev_buf);
    delete plan;
    return set_status(LIBSHALOM_SUCCESS);
}

int LibShalom_plan_info(LibShalom_plan_t plan, double *flops, double *bytes, long *work_items)
{
    if (!plan) return set_status(LIBSHALOM_INVALID_POINTER);
    if (flops) *flops = plan->flops;
    if (bytes) *bytes = plan->bytes;
    if (work_items) *work_items = plan->items;
    return set_status(LIBSHALOM_SUCCESS);
}

static int batched_oneshot(int prec, int ta, int tb, void *const *C, const void *const *A, const void *const *B,
                           const long *M, const long *N, const long *K, long count, hipStream_t s)
{
    LibShalom_plan_t plan = nullptr;
    int st = build_plan(&plan, prec, ta, tb, C, A, B, M, N, K, count);
    if (st) return st;
    if (hipGraphLaunch(plan->exec, s) != hipSuccess) st = LIBSHALOM_HIP_ERROR;
    if (hipStreamSynchronize(s) != hipSuccess) st = LIBSHALOM_HIP_ERROR;
    (void)hipGetLastError();
    LibShalom_plan_destroy(plan);
    return st;
}

int LibShalom_sgemm_batched(int transa, int transb, float *const *C, const float *const *A, const float *const *B,
                            const long *M, const long *N, const long *K, long count, int precision,
                            LibShalom_stream_t stream)
{
    if (!valid_precision(precision)) return set_status(LIBSHALOM_INVALID_VALUE);
    return set_status(batched_oneshot(kprec(precision), transa, transb, (void *const *)C, (const void *const *)A,
                                      (const void *const *)B, M, N, K, count, (hipStream_t)stream));
}

int LibShalom_dgemm_batched(int transa, int transb, double *const *C, const double *const *A,
                            const double *const *B, const long *M, const long *N, const long *K, long count,
                            LibShalom_stream_t stream)
{
    return set_status(batched_oneshot(P_FP64, transa, transb, (void *const *)C, (const void *const *)A,
                                      (const void *const *)B, M, N, K, count, (hipStream_t)stream));
}

int LibShalom_sgemm_backward(int transa, int transb, const float *dC, const float *A, const float *B, float *dA,
                             float *dB, long M, long N, long K, int precision, LibShalom_stream_t stream)
{
    return LibShalom_sgemm_backward_strided(transa, transb, dC, A, B, dA, dB, M, N, K, 1, 0, 0, 0, precision,
                                            stream);
}

int LibShalom_sgemm_backward_strided(int transa, int transb, const float *dC, const float *A, const float *B,
                                     float *dA, float *dB, long M, long N, long K, long batch, long strideA,
                                     long strideB, long strideC, int precision, LibShalom_stream_t stream)
{
    int st = valid_precision(precision) ? LIBSHALOM_SUCCESS : LIBSHALOM_INVALID_VALUE;
    if (!st) st = check_dims(transa, transb, M, N, K);
    if (!st && (batch < 0 || batch > INT32_MAX || strideA < 0 || strideB < 0 || strideC < 0))
        st = LIBSHALOM_INVALID_VALUE;
    if (st) return set_status(st);
    if (batch == 0) return set_status(LIBSHALOM_SUCCESS);
    if ((M * N > 0 && !dC) || (M * K > 0 && (!A || !dA)) || (K * N > 0 && (!B || !dB)))
        return set_status(LIBSHALOM_INVALID_POINTER);
    return set_status(backward_run(kprec(precision), transa, transb, dC, A, B, dA, dB, M, N, K, (int)batch,
                                   strideA, strideB, strideC, (hipStream_t)stream));
}

int LibShalom_sgemm_strided_batched(int transa, int transb, float *C, const float *A, const float *B, long M,
                                    long N, long K, long batch, long strideA, long strideB, long strideC,
                                    int precision, LibShalom_stream_t stream)
{
    int st = valid_precision(precision) ? LIBSHALOM_SUCCESS : LIBSHALOM_INVALID_VALUE;
    if (!st) st = check_dims(transa, transb, M, N, K);
    if (!st && (batch < 0 || batch > INT32_MAX || strideA < 0 || strideB < 0 || strideC < 0))
        st = LIBSHALOM_INVALID_VALUE;
    if (st) return set_status(st);
    if (batch == 0) return set_status(LIBSHALOM_SUCCESS);
    if ((M * N > 0 && !C) || (M * N > 0 && K > 0 && (!A || !B))) return set_status(LIBSHALOM_INVALID_POINTER);
    return set_status(strided_run(kprec(precision), transa, transb, C, A, B, M, N, K, (int)batch, strideA, strideB,
                                  strideC, (hipStream_t)stream));
}

int LibShalom_sconv2d_forward(float *Y, const float *Wt, const float *X, long batch, long C, long H, long W,
                              long Cout, long R, long S, long stride_h, long stride_w, long pad_h, long pad_w,
                              long dil_h, long dil_w, int precision, LibShalom_stream_t stream)
{
    if (!valid_precision(precision)) return set_status(LIBSHALOM_INVALID_VALUE);
    if (batch < 0 || C <= 0 || H <= 0 || W <= 0 || Cout <= 0 || R <= 0 || S <= 0 || stride_h <= 0 ||
        stride_w <= 0 || pad_h < 0 || pad_w < 0 || dil_h <= 0 || dil_w <= 0)
        return set_status(LIBSHALOM_INVALID_VALUE);
    const long P = conv_out_size(H, R, stride_h, pad_h, dil_h), Q = conv_out_size(W, S, stride_w, pad_w, dil_w);
    if (P <= 0 || Q <= 0) return set_status(LIBSHALOM_INVALID_VALUE);
    // 32-bit offsets within an image and exact multiply-high divisions
    const double K = (double)C * R * S, img = (double)C * H * W, PQ = (double)P * Q;
    if (batch > INT32_MAX || Cout > INT32_MAX || 4.0 * img >= 2147483648.0 || PQ >= 2147483648.0 ||
        (K + R * S) * R * S >= 4294967296.0 || PQ * Q >= 4294967296.0 || H + 2.0 * pad_h >= 536870912.0 ||
        W + 2.0 * pad_w >= 536870912.0 ||
        // leading dimensions of the filter (C*R*S) and of each output image (P*Q)
        K >= (double)kMaxLd || PQ >= (double)kMaxLd)
        return set_status(LIBSHALOM_NOT_SUPPORTED);
    if (batch == 0) return set_status(LIBSHALOM_SUCCESS);
    if (!Y || !Wt || !X) return set_status(LIBSHALOM_INVALID_POINTER);
    return set_status(conv_fwd_run(kprec(precision), Y, Wt, X, batch, C, H, W, Cout, R, S, stride_h, stride_w, pad_h,
                                   pad_w, dil_h, dil_w, P, Q, (hipStream_t)stream));
}

int LibShalom_sconv2d_backward(float *dX, float *dWt, const float *dY, const float *Wt, const float *X, long batch,
                               long C, long H, long W, long Cout, long R, long S, long stride_h, long stride_w,
                               long pad_h, long pad_w, long dil_h, long dil_w, int precision,
                               LibShalom_stream_t stream)
{
    if (!valid_precision(precision)) return set_status(LIBSHALOM_INVALID_VALUE);
    if (batch < 0 || C <= 0 || H <= 0 || W <= 0 || Cout <= 0 || R <= 0 || S <= 0 || stride_h <= 0 ||
        stride_w <= 0 || pad_h < 0 || pad_w < 0 || dil_h <= 0 || dil_w <= 0)
        return set_status(LIBSHALOM_INVALID_VALUE);
    const long P = conv_out_size(H, R, stride_h, pad_h, dil_h), Q = conv_out_size(W, S, stride_w, pad_w, dil_w);
    if (P <= 0 || Q <= 0) return set_status(LIBSHALOM_INVALID_VALUE);
    const double K = (double)C * R * S, img = (double)C * H * W, PQ = (double)P * Q;
    if (batch > INT32_MAX || Cout > INT32_MAX || 4.0 * img >= 2147483648.0 || PQ >= 2147483648.0 ||
        (K + R * S) * R * S >= 4294967296.0 || PQ * Q >= 4294967296.0 || H + 2.0 * pad_h >= 268435456.0 ||
        W + 2.0 * pad_w >= 268435456.0 || 4.0 * Cout * PQ >= 2147483648.0 || K >= 2147483648.0 ||
        // the input gradient's own gather: dY as the image, H x W as its output
        ((double)Cout * R * S + R * S) * R * S >= 4294967296.0 || (double)H * W * W >= 4294967296.0 ||
        // leading dimensions: dW and the flipped filter (C*R*S, Cout*R*S), dY and dX images (P*Q, H*W)
        K >= (double)kMaxLd || (double)Cout * R * S >= (double)kMaxLd || PQ >= (double)kMaxLd ||
        (double)H * W >= (double)kMaxLd)
        return set_status(LIBSHALOM_NOT_SUPPORTED);
    // the input gradient of a strided convolution: strides up to 8 (exact
    // multiply-high division of r*dil by the stride)
    if (dX && (stride_h > 8 || stride_w > 8))
        return set_status(LIBSHALOM_NOT_SUPPORTED);
    if (batch == 0 || (!dX && !dWt)) return set_status(LIBSHALOM_SUCCESS);
    if (!dY || (dX && !Wt) || (dWt && !X)) return set_status(LIBSHALOM_INVALID_POINTER);
    return set_status(conv_bwd_run(kprec(precision), dX, dWt, dY, Wt, X, batch, C, H, W, Cout, R, S, stride_h,
                                   stride_w, pad_h, pad_w, dil_h, dil_w, P, Q, (hipStream_t)stream));
}

int LibShalom_conv2d_output_size(long H, long W, long R, long S, long stride_h, long stride_w, long pad_h,
                                 long pad_w, long dil_h, long dil_w, long *P, long *Q)
{
    if (H <= 0 || W <= 0 || R <= 0 || S <= 0 || stride_h <= 0 || stride_w <= 0 || pad_h < 0 || pad_w < 0 ||
        dil_h <= 0 || dil_w <= 0)
        return set_status(LIBSHALOM_INVALID_VALUE);
    if (P) *P = conv_out_size(H, R, stride_h, pad_h, dil_h);
    if (Q) *Q = conv_out_size(W, S, stride_w, pad_w, dil_w);
    return set_status(LIBSHALOM_SUCCESS);
}

int LibShalom_autotune_batched(int precision, int transa, int transb, const long *M, const long *N, const long *K,
                               long count, int *classes_tuned, LibShalom_stream_t stream)
{
    if (!valid_precision(precision) && precision != LIBSHALOM_FP64) return set_status(LIBSHALOM_INVALID_VALUE);
    if ((transa != 0 && transa != 1) || (transb != 0 && transb != 1) || count < 0)
        return set_status(LIBSHALOM_INVALID_VALUE);
    if (count > 0 && (!M || !N || !K)) return set_status(LIBSHALOM_INVALID_POINTER);
    if (classes_tuned) *classes_tuned = 0;
    return set_status(autotune(kprec(precision), transa, transb, M, N, K, count, (hipStream_t)stream,
                               classes_tuned));
}

int LibShalom_autotune_reset(void)
{
    tune_reset();
    return set_status(LIBSHALOM_SUCCESS);
}

// The tile a batched plan runs a GEMM on, given its bucket tile: FP64 plans
// run the GEMMs of the 32 x 32 bucket with M, N <= 16 on 16 x 16 tiles.
static void report_plan_tile(int precision, int t, long M, long N, int *tile_m, int *tile_n)
{
    const bool t16 = precision == LIBSHALOM_FP64 && t == T32x32 && M <= 16 && N <= 16;
    if (tile_m) *tile_m = t16 ? 16 : kTileShapes[t].bm;
    if (tile_n) *tile_n = t16 ? 16 : kTileShapes[t].bn;
}

int LibShalom_query_tuned_tile(int precision, int transa, int transb, long M, long N, long K, int *tile_m,
                               int *tile_n)
{
    if (!valid_precision(precision) && precision != LIBSHALOM_FP64) return set_status(LIBSHALOM_INVALID_VALUE);
    if ((transa != 0 && transa != 1) || (transb != 0 && transb != 1) || M < 0 || N < 0 || K < 0)
        return set_status(LIBSHALOM_INVALID_VALUE);
    const int t = tuned_tile(kprec(precision), transa, transb, M, N, K);
    report_plan_tile(precision, t, M, N, tile_m, tile_n);
    return set_status(LIBSHALOM_SUCCESS);
}

int LibShalom_query_plan(int precision, int mp, long M, long N, long K, int num_cu, int *tile_m, int *tile_n,
                         int *ksplit, int *kper)
{
    if (!valid_precision(precision) && precision != LIBSHALOM_FP64) return set_status(LIBSHALOM_INVALID_VALUE);
    if (M < 0 || N < 0 || K < 0 || num_cu <= 0) return set_status(LIBSHALOM_INVALID_VALUE);
    // a short-sided FP32 GEMM (row-major, no transposes) runs on the
    // short-side kernel: 16 long-side indices per workgroup, K whole
    const int nc = skinny_choice(kprec(precision), 0, 0, M, N, K, K, N, num_cu);
    if (nc >= 0) {
        const int w = 16 << nc;
        if (tile_m) *tile_m = M <= 64 ? w : 16;
        if (tile_n) *tile_n = M <= 64 ? 16 : w;
        if (ksplit) *ksplit = 1;
        if (kper) *kper = (int)std::max<long>(K, 1);
        return set_status(LIBSHALOM_SUCCESS);
    }
    const SinglePlan p = plan_single(kprec(precision), M, N, K, mp != 0, num_cu);
    if (tile_m) *tile_m = kTileShapes[p.tile].bm;
    if (tile_n) *tile_n = kTileShapes[p.tile].bn;
    if (ksplit) *ksplit = p.ksplit;
    if (kper) *kper = p.kper;
    return set_status(LIBSHALOM_SUCCESS);
}

int LibShalom_query_batched_tile(int precision, long M, long N, long K, int *tile_m, int *tile_n)
{
    if (!valid_precision(precision) && precision != LIBSHALOM_FP64) return set_status(LIBSHALOM_INVALID_VALUE);
    if (M < 0 || N < 0 || K < 0) return set_status(LIBSHALOM_INVALID_VALUE);
    const int t = batched_tile(kprec(precision), M, N, K);
    report_plan_tile(precision, t, M, N, tile_m, tile_n);
    return set_status(LIBSHALOM_SUCCESS);
}

}  // extern "C"
