// CPU emulation of the HIP surface used by libpdpd_hip (bounds checking with
// AddressSanitizer). Work-items run as ucontext fibers, one block at a time;
// __syncthreads and wave intrinsics are barriers among the live fibers.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>
#include <ucontext.h>
#include <algorithm>

#define __global__
#define __device__
#define __host__
#define __shared__ static
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __HIP_MEMORY_SCOPE_AGENT 0
#define HIP_SYMBOL(x) (&x)
#define __align__(n) alignas(n)

struct float2 { float x, y; };
struct float3 { float x, y, z; };
struct float4 { float x, y, z, w; };
inline float2 make_float2(float a, float b) { return {a, b}; }
inline float3 make_float3(float a, float b, float c) { return {a, b, c}; }
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }
struct dim3_ { uint32_t x, y, z; };

inline uint32_t __float_as_uint(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
inline float __uint_as_float(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
inline uint32_t __umulhi(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }
inline int __clzll(long long x) { return x ? __builtin_clzll((unsigned long long)x) : 64; }
inline int __ffsll(long long x) { return x ? __builtin_ctzll((unsigned long long)x) + 1 : 0; }
inline int __ffsll(unsigned long long x) { return x ? __builtin_ctzll(x) + 1 : 0; }
inline unsigned __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline unsigned __popc(unsigned x) { return __builtin_popcount(x); }
template <class A, class B> inline auto min(A a, B b) { return a < b ? a : b; }
template <class A, class B> inline auto max(A a, B b) { return a > b ? a : b; }

namespace emu {
struct Fiber {
    ucontext_t ctx;
    std::vector<char> stack;
    dim3_ tid, bid;
    bool done = false;
    int wait = 0;            // 0 running, 1 block barrier, 2 wave op
    uint64_t val = 0;        // wave op operand
    uint64_t res = 0;        // wave op result
    int op = 0, arg = 0;
};
struct State {
    std::vector<Fiber> fibers;
    int cur = -1;
    ucontext_t sched;
    dim3_ bdim, gdim;
    std::function<void()> body;
    bool capturing = false;
    std::vector<std::function<void()>> captured;
};
inline State &st() { static State s; return s; }
inline Fiber &me() { return st().fibers[st().cur]; }
inline void yield_() { State &s = st(); swapcontext(&s.fibers[s.cur].ctx, &s.sched); }
inline void trampoline() { st().body(); me().done = true; yield_(); }

// one wave-level exchange: every live lane of the wave posts, the last computes
inline uint64_t wave_op(int op, uint64_t v, int arg)
{
    Fiber &f = me();
    f.wait = 2; f.op = op; f.val = v; f.arg = arg;
    yield_();
    return me().res;
}
inline void resolve_waves()
{
    State &s = st();
    const int n = (int)s.fibers.size();
    for (int w0 = 0; w0 < n; w0 += 64) {
        int live = 0, waiting = 0, op = -1;
        for (int l = w0; l < std::min(n, w0 + 64); l++) {
            if (s.fibers[l].done) continue;
            live++;
            if (s.fibers[l].wait == 2) { waiting++; op = s.fibers[l].op; }
        }
        if (!live || waiting != live) continue;
        // all live lanes of the wave are at a wave op: check they agree, compute
        uint64_t ballot = 0;
        for (int l = w0; l < std::min(n, w0 + 64); l++) {
            Fiber &f = s.fibers[l];
            if (f.done) continue;
            if (f.op != op) { fprintf(stderr, "EMU: wave op mismatch\n"); abort(); }
            if (op == 0 && f.val) ballot |= 1ull << (l - w0);
        }
        for (int l = w0; l < std::min(n, w0 + 64); l++) {
            Fiber &f = s.fibers[l];
            if (f.done) continue;
            const int lane = l - w0;
            int src = lane;
            if (op == 0) { f.res = ballot; f.wait = 0; continue; }
            if (op == 4) { f.res = 0; f.wait = 0; continue; }
            if (op == 5) {                                                 // readfirstlane
                int first = w0;
                while (s.fibers[first].done) first++;
                f.res = s.fibers[first].val;
                f.wait = 0;
                continue;
            }
            if (op == 1) src = lane - f.arg < 0 ? lane : lane - f.arg;     // shfl_up
            if (op == 2) src = f.arg & 63;                                 // shfl
            if (op == 3) src = lane ^ f.arg;                               // shfl_xor
            const int sl = w0 + src;
            if (sl >= n || s.fibers[sl].done) { fprintf(stderr, "EMU: shuffle from inactive lane %d\n", src); abort(); }
            f.res = s.fibers[sl].val;
            f.wait = 0;
        }
        // second pass sets wait=0 for shuffles (already done above)
    }
}
inline void run_block(dim3_ bid)
{
    State &s = st();
    const uint32_t nt = s.bdim.x;
    s.fibers.assign(nt, Fiber());
    for (uint32_t t = 0; t < nt; t++) {
        Fiber &f = s.fibers[t];
        f.stack.resize(1 << 18);
        f.tid = {t, 0, 0};
        f.bid = bid;
        getcontext(&f.ctx);
        f.ctx.uc_stack.ss_sp = f.stack.data();
        f.ctx.uc_stack.ss_size = f.stack.size();
        f.ctx.uc_link = nullptr;
        makecontext(&f.ctx, (void (*)())trampoline, 0);
    }
    for (;;) {
        bool progress = false, all_done = true;
        for (uint32_t t = 0; t < nt; t++) {
            Fiber &f = s.fibers[t];
            if (f.done) continue;
            all_done = false;
            if (f.wait) continue;
            s.cur = (int)t;
            swapcontext(&s.sched, &f.ctx);
            progress = true;
        }
        if (all_done) break;
        // block barrier: all live fibers waiting at it -> release
        int live = 0, atbar = 0;
        for (auto &f : s.fibers) if (!f.done) { live++; if (f.wait == 1) atbar++; }
        if (live && atbar == live) { for (auto &f : s.fibers) if (!f.done) f.wait = 0; progress = true; }
        resolve_waves();
        bool any_run = false;
        for (auto &f : s.fibers) if (!f.done && !f.wait) any_run = true;
        if (!any_run && !progress) {
            fprintf(stderr, "EMU: deadlock in block %u (live %d, at barrier %d)\n", bid.x, live, atbar);
            abort();
        }
    }
    s.cur = -1;
}
inline void launch(uint32_t grid, uint32_t block, std::function<void()> body)
{
    State &s = st();
    if (s.capturing) { s.captured.push_back([grid, block, body]() { launch(grid, block, body); }); return; }
    s.gdim = {grid, 1, 1};
    s.bdim = {block, 1, 1};
    s.body = body;
    for (uint32_t b = 0; b < grid; b++) run_block({b, 0, 0});
}
}  // namespace emu

#define threadIdx (emu::me().tid)
#define blockIdx (emu::me().bid)
#define blockDim (emu::st().bdim)
#define gridDim (emu::st().gdim)
inline void __syncthreads() { emu::me().wait = 1; emu::yield_(); }
inline unsigned long long __ballot(int p) { return emu::wave_op(0, p ? 1 : 0, 0); }
inline unsigned long long __builtin_amdgcn_ballot_w64(bool p) { return __ballot(p); }
inline uint32_t __shfl_up(uint32_t v, unsigned d, int = 64) { return (uint32_t)emu::wave_op(1, v, (int)d); }
inline uint32_t __shfl(uint32_t v, int src, int = 64) { return (uint32_t)emu::wave_op(2, v, src); }
inline uint32_t __shfl_xor(uint32_t v, int m, int = 64) { return (uint32_t)emu::wave_op(3, v, m); }
inline uint32_t __builtin_amdgcn_readfirstlane(uint32_t v) { return (uint32_t)emu::wave_op(5, v, 0); }
inline uint32_t __ockl_wfred_max_u32(uint32_t v)
{
    for (int m = 32; m >= 1; m >>= 1) {
        const uint32_t o = __shfl_xor(v, m);
        v = o > v ? o : v;
    }
    return v;
}
inline void __builtin_amdgcn_s_sleep(int) { emu::yield_(); }
inline void __builtin_amdgcn_wave_barrier() { (void)emu::wave_op(4, 0, 0); }
inline void __builtin_amdgcn_fence(int, const char *) {}   // fibers run one at a time
template <class T, class U> inline T atomicAdd(T *p, U v) { T o = *p; *p = o + (T)v; return o; }
template <class T, class U> inline T atomicOr(T *p, U v) { T o = *p; *p = o | (T)v; return o; }
template <class T> inline T __hip_atomic_load(T *p, int, int) { return *p; }
template <class T, class U> inline void __hip_atomic_store(T *p, U v, int, int) { *p = (T)v; }

// ---- runtime ----
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorNotReady = 600, hipErrorNotSupported = 801 };
typedef void *hipStream_t;
typedef void *hipEvent_t;
typedef void *hipGraph_t;
typedef void *hipGraphExec_t;
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyHostToHost };
enum { hipStreamNonBlocking = 1, hipStreamCaptureModeThreadLocal = 1 };
typedef int hipStreamCaptureMode;
struct hipDeviceProp_t { int multiProcessorCount; char name[256]; };
inline const char *hipGetErrorString(hipError_t e) { return e ? "emulated error" : "no error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int) { p->multiProcessorCount = 256; strcpy(p->name, "emu"); return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipMalloc(void **p, size_t n) { *p = malloc(n); if (*p) memset(*p, 0xA5, n); return *p ? hipSuccess : 2; }
inline hipError_t hipFree(void *p) { free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t) { memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = (void *)1; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamQuery(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = (void *)1; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 0.001f; return hipSuccess; }
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { *e = (void *)1; return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) { *p = const_cast<void *>(sym); return hipSuccess; }
inline hipError_t hipHostMalloc(void **p, size_t n, unsigned) { *p = malloc(n); return *p ? hipSuccess : 2; }
inline hipError_t hipHostFree(void *p) { free(p); return hipSuccess; }
#define hipEventDisableTiming 2u
#define hipHostMallocDefault 0u
struct EmuGraph { std::vector<std::function<void()>> ops; };
inline hipError_t hipStreamBeginCapture(hipStream_t, int) { emu::st().capturing = true; emu::st().captured.clear(); return hipSuccess; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) {
    emu::st().capturing = false;
    auto *G = new EmuGraph{emu::st().captured};
    emu::st().captured.clear();
    *g = G;
    return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, char *, size_t) {
    *e = new EmuGraph(*(EmuGraph *)g); return hipSuccess; }
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete (EmuGraph *)g; return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete (EmuGraph *)e; return hipSuccess; }
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) { for (auto &op : ((EmuGraph *)e)->ops) op(); return hipSuccess; }
