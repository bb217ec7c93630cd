// CPU emulation of the HIP runtime and of workgroup execution (fibers); the
// execution model is described at the top of hip/hip_runtime.h.
#include "hip/hip_runtime.h"

#include <sys/mman.h>
#include <ucontext.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <thread>
#include <vector>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/asan_interface.h>
#include <sanitizer/common_interface_defs.h>
#define EMU_ASAN 1
#endif

struct emu_stream {
    bool capturing = false;
    emu_graph *graph = nullptr;
};
struct emu_event {
    long long t = 0;
};
struct emu_graph {
    std::vector<std::function<void()>> ops;
};

namespace emu {

enum FState { RUNNABLE, WAIT_BAR, WAIT_WAVE, DONE };

struct Fiber {
    ucontext_t ctx;
    void *stack = nullptr;
    size_t ssize = 0;
    dim3 tid;
    FState st = DONE;
    int pred = 0, op = 0, arg = 0;
    unsigned long long val = 0, res = 0;
    void *fake = nullptr;
};

struct Block {
    dim3 bid, bdim, gdim;
    int nt = 0;
    Fiber *f = nullptr;
    ucontext_t main;
    const std::function<void()> *body = nullptr;
    const void *main_bottom = nullptr;
    size_t main_size = 0;
};

thread_local Fiber *t_fiber = nullptr;
thread_local Block *t_block = nullptr;

static dim3 g_host_dim;  // threadIdx & co. outside a kernel: a fault of the caller

static Fiber &cur_fiber()
{
    if (!t_fiber) {
        std::fprintf(stderr, "emu: thread index read outside a kernel\n");
        std::abort();
    }
    return *t_fiber;
}
dim3 &thread_idx() { return cur_fiber().tid; }
dim3 &block_idx() { return t_block ? t_block->bid : g_host_dim; }
dim3 &block_dim() { return t_block ? t_block->bdim : g_host_dim; }
dim3 &grid_dim() { return t_block ? t_block->gdim : g_host_dim; }

long long now_ticks()
{
    using namespace std::chrono;
    return (long long)(duration_cast<nanoseconds>(steady_clock::now().time_since_epoch()).count() / 10);
}

// ---------------------------------------------------------------- fibers
constexpr size_t STACK = 256 * 1024;

static void yield_to_scheduler()
{
    Fiber &F = *t_fiber;
    Block &B = *t_block;
#ifdef EMU_ASAN
    __sanitizer_start_switch_fiber(&F.fake, B.main_bottom, B.main_size);
#endif
    swapcontext(&F.ctx, &B.main);
#ifdef EMU_ASAN
    __sanitizer_finish_switch_fiber(F.fake, &B.main_bottom, &B.main_size);
#endif
}

static void fiber_entry()
{
    Block &B = *t_block;
#ifdef EMU_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &B.main_bottom, &B.main_size);
#endif
    (*B.body)();
    t_fiber->st = DONE;
#ifdef EMU_ASAN
    __sanitizer_start_switch_fiber(nullptr, B.main_bottom, B.main_size);
#endif
    setcontext(&B.main);
}

static void run_fiber(Block &B, Fiber &F)
{
    t_fiber = &F;
#ifdef EMU_ASAN
    void *fake = nullptr;
    __sanitizer_start_switch_fiber(&fake, F.stack, F.ssize);
#endif
    swapcontext(&B.main, &F.ctx);
#ifdef EMU_ASAN
    __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#endif
    t_fiber = nullptr;
}

static const char *state_name(FState s)
{
    switch (s) {
    case RUNNABLE: return "runnable";
    case WAIT_BAR: return "barrier";
    case WAIT_WAVE: return "wave-op";
    default: return "done";
    }
}

[[noreturn]] static void deadlock(Block &B, const char *why)
{
    std::fprintf(stderr, "emu: workgroup (%u,%u,%u) of a %u-thread launch: %s\n", B.bid.x, B.bid.y, B.bid.z,
                 (unsigned)B.nt, why);
    for (int w = 0; w * WAVE < B.nt; w++) {
        int n[4] = {0, 0, 0, 0};
        for (int l = w * WAVE; l < std::min(B.nt, (w + 1) * WAVE); l++) n[B.f[l].st]++;
        std::fprintf(stderr, "  wave %d: %d %s, %d %s, %d %s, %d %s\n", w, n[0], state_name(RUNNABLE), n[1],
                     state_name(WAIT_BAR), n[2], state_name(WAIT_WAVE), n[3], state_name(DONE));
    }
    std::abort();
}

int sync_block(int pred, int op)
{
    Fiber &F = cur_fiber();
    F.pred = pred;
    F.op = op;
    F.st = WAIT_BAR;
    yield_to_scheduler();
    return (int)F.res;
}

unsigned long long wave_exchange(unsigned long long v, int op, int arg)
{
    Fiber &F = cur_fiber();
    F.val = v;
    F.op = op;
    F.arg = arg;
    F.st = WAIT_WAVE;
    yield_to_scheduler();
    return F.res;
}

// Release a wave whose live lanes all wait at a wave operation.
static bool resolve_wave(Block &B, int w)
{
    const int l0 = w * WAVE, l1 = std::min(B.nt, l0 + WAVE);
    int live = 0, waiting = 0, op = -1, arg = 0;
    for (int l = l0; l < l1; l++) {
        const Fiber &F = B.f[l];
        if (F.st == DONE) continue;
        live++;
        if (F.st != WAIT_WAVE) continue;
        if (waiting == 0) {
            op = F.op;
            arg = F.arg;
        } else if (F.op != op || (op != 2 && F.arg != arg)) {
            deadlock(B, "lanes of one wave wait at different wave operations");
        }
        waiting++;
    }
    if (waiting == 0) return false;
    if (waiting != live) deadlock(B, "a wave operation is reached by only some live lanes of its wave");
    unsigned long long ballot = 0;
    for (int l = l0; l < l1; l++)
        if (B.f[l].st == WAIT_WAVE && B.f[l].val) ballot |= 1ull << (l - l0);
    for (int l = l0; l < l1; l++) {
        Fiber &F = B.f[l];
        if (F.st != WAIT_WAVE) continue;
        const int lane = l - l0;
        if (op == 2) {
            F.res = ballot;
        } else {
            const int src = op == 0 ? (lane ^ arg) : lane - arg;
            if (op == 1 && src < 0) {
                F.res = F.val;  // shfl_up below the offset keeps its own value
            } else {
                // the source lane must take part: a lane that has ended or
                // branched around the operation holds no defined value
                const int gl = l0 + src;
                if (src < 0 || src >= WAVE || gl >= l1 || B.f[gl].st != WAIT_WAVE)
                    deadlock(B, "a shuffle reads a lane that does not take part in it");
                F.res = B.f[gl].val;
            }
        }
    }
    for (int l = l0; l < l1; l++)
        if (B.f[l].st == WAIT_WAVE) B.f[l].st = RUNNABLE;
    return true;
}

// Release the workgroup barrier when every live thread waits at it.
static bool resolve_barrier(Block &B)
{
    int live = 0, waiting = 0, op = -1;
    int all = 1, any = 0, count = 0;
    for (int l = 0; l < B.nt; l++) {
        const Fiber &F = B.f[l];
        if (F.st == DONE) continue;
        live++;
        if (F.st != WAIT_BAR) continue;
        if (waiting == 0) op = F.op;
        else if (F.op != op) deadlock(B, "threads wait at barriers of different kinds");
        waiting++;
        all = all && F.pred;
        any = any || F.pred;
        count += F.pred != 0;
    }
    if (waiting == 0 || waiting != live) return false;
    const int r = op == 1 ? all : op == 2 ? any : op == 3 ? count : 0;
    for (int l = 0; l < B.nt; l++)
        if (B.f[l].st == WAIT_BAR) {
            B.f[l].res = (unsigned long long)r;
            B.f[l].st = RUNNABLE;
        }
    return true;
}

// per OS thread: fiber stacks, kept across workgroups
thread_local std::vector<Fiber> t_fibers;

static void run_block(dim3 grid, dim3 block, unsigned long long lin, const std::function<void()> &body)
{
    Block B;
    B.gdim = grid;
    B.bdim = block;
    B.bid = dim3((unsigned)(lin % grid.x), (unsigned)(lin / grid.x % grid.y), (unsigned)(lin / grid.x / grid.y));
    B.nt = (int)(block.x * block.y * block.z);
    B.body = &body;
    if ((int)t_fibers.size() < B.nt) t_fibers.resize(B.nt);
    B.f = t_fibers.data();
    for (int l = 0; l < B.nt; l++) {
        Fiber &F = B.f[l];
        if (!F.stack) {
            F.stack = mmap(nullptr, STACK, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
            if (F.stack == MAP_FAILED) {
                std::fprintf(stderr, "emu: fiber stack allocation failed\n");
                std::abort();
            }
            F.ssize = STACK;
        }
#ifdef EMU_ASAN
        ASAN_UNPOISON_MEMORY_REGION(F.stack, F.ssize);
#endif
        getcontext(&F.ctx);
        F.ctx.uc_stack.ss_sp = F.stack;
        F.ctx.uc_stack.ss_size = F.ssize;
        F.ctx.uc_link = nullptr;
        makecontext(&F.ctx, fiber_entry, 0);
        F.tid = dim3((unsigned)(l % block.x), (unsigned)(l / block.x % block.y), (unsigned)(l / block.x / block.y));
        F.st = RUNNABLE;
        F.res = 0;
    }
    t_block = &B;
    for (;;) {
        bool live = false;
        for (int l = 0; l < B.nt; l++) {
            if (B.f[l].st == RUNNABLE) run_fiber(B, B.f[l]);
            live = live || B.f[l].st != DONE;
        }
        if (!live) break;
        bool released = false;
        for (int w = 0; w * WAVE < B.nt; w++) released = resolve_wave(B, w) || released;
        if (!released) released = resolve_barrier(B);
        if (!released) deadlock(B, "every live thread waits and no barrier or wave operation can complete");
    }
    t_block = nullptr;
}

// ---------------------------------------------------------------- workgroup pool
struct Pool {
    std::mutex launch;  // one grid at a time, like one in-order queue
    std::mutex m;
    std::condition_variable go, done;
    unsigned long long gen = 0;
    std::atomic<unsigned long long> next{0};
    unsigned long long nblocks = 0;
    dim3 grid, block;
    const std::function<void()> *body = nullptr;
    int busy = 0;
    std::vector<std::thread> th;

    Pool()
    {
        for (int i = 0; i < EMU_RESIDENT; i++) th.emplace_back([this]() { worker(); });
        for (auto &t : th) t.detach();
    }
    void worker()
    {
        unsigned long long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(m);
                go.wait(lk, [&]() { return gen != seen; });
                seen = gen;
            }
            for (;;) {
                const unsigned long long b = next.fetch_add(1);
                if (b >= nblocks) break;
                run_block(grid, block, b, *body);
            }
            std::lock_guard<std::mutex> lk(m);
            if (--busy == 0) done.notify_all();
        }
    }
    void run(dim3 g, dim3 b, const std::function<void()> &fn)
    {
        std::lock_guard<std::mutex> one(launch);
        {
            std::lock_guard<std::mutex> lk(m);
            grid = g;
            block = b;
            body = &fn;
            nblocks = (unsigned long long)g.x * g.y * g.z;
            next = 0;
            busy = EMU_RESIDENT;
            gen++;
        }
        go.notify_all();
        std::unique_lock<std::mutex> lk(m);
        done.wait(lk, [&]() { return busy == 0; });
    }
};

static Pool &pool()
{
    static Pool *p = new Pool();  // lives until the process ends
    return *p;
}

static thread_local hipError_t t_last = hipSuccess;

void run_grid(dim3 grid, dim3 block, const std::function<void()> &body)
{
    const unsigned long long nt = (unsigned long long)block.x * block.y * block.z;
    if (grid.x == 0 || grid.y == 0 || grid.z == 0 || nt == 0 || nt > 1024) {
        std::fprintf(stderr, "emu: invalid launch: grid (%u,%u,%u), block (%u,%u,%u)\n", grid.x, grid.y, grid.z,
                     block.x, block.y, block.z);
        t_last = hipErrorInvalidValue;
        return;
    }
    pool().run(grid, block, body);
}

hipError_t submit(hipStream_t s, std::function<void()> op)
{
    if (s && s->capturing) {
        s->graph->ops.push_back(std::move(op));
        return hipSuccess;
    }
    op();
    hipError_t e = t_last;
    t_last = hipSuccess;
    return e;
}

}  // namespace emu

void emu_yield() { std::this_thread::yield(); }

// ---------------------------------------------------------------- runtime API
using emu::submit;

static int g_device = 0;

// Byte pattern of fresh device allocations: 0xff (NaN doubles) by default,
// so that a read before the first write shows; EMU_FILL=<byte> sets it.
static int fill_byte()
{
    static const int b = []() {
        const char *e = std::getenv("EMU_FILL");
        return e ? (int)std::strtol(e, nullptr, 0) & 0xff : 0xff;
    }();
    return b;
}

hipError_t hipMalloc(void **p, size_t n)
{
    if (!p) return hipErrorInvalidValue;
    *p = nullptr;
    if (n == 0) return hipSuccess;
    if (posix_memalign(p, 256, n) != 0) return hipErrorOutOfMemory;
    std::memset(*p, fill_byte(), n);
    return hipSuccess;
}
hipError_t hipFree(void *p)
{
    std::free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned)
{
    if (!p) return hipErrorInvalidValue;
    *p = nullptr;
    if (n == 0) return hipSuccess;
    if (posix_memalign(p, 256, n) != 0) return hipErrorOutOfMemory;
    std::memset(*p, fill_byte(), n);
    return hipSuccess;
}
hipError_t hipHostFree(void *p)
{
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind)
{
    if (n) std::memcpy(dst, src, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind, hipStream_t s)
{
    return submit(s, [=]() {
        if (n) std::memcpy(dst, src, n);
    });
}
hipError_t hipMemset(void *p, int v, size_t n)
{
    if (n) std::memset(p, v, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void *p, int v, size_t n, hipStream_t s)
{
    return submit(s, [=]() {
        if (n) std::memset(p, v, n);
    });
}
hipError_t hipSetDevice(int d)
{
    if (d != 0) return hipErrorInvalidDevice;
    g_device = d;
    return hipSuccess;
}
hipError_t hipGetDevice(int *d)
{
    *d = g_device;
    return hipSuccess;
}
hipError_t hipGetDeviceCount(int *n)
{
    *n = 1;
    return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int d)
{
    if (d != 0) return hipErrorInvalidDevice;
    std::memset(p, 0, sizeof(*p));
    std::snprintf(p->name, sizeof(p->name), "CPU emulation");
    std::snprintf(p->gcnArchName, sizeof(p->gcnArchName), "emu");
    p->multiProcessorCount = emu::EMU_CUS;
    p->totalGlobalMem = (size_t)8 << 30;
    p->warpSize = 64;
    p->maxThreadsPerBlock = 1024;
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int d)
{
    if (d != 0) return hipErrorInvalidDevice;
    *v = a == hipDeviceAttributeWallClockRate ? 100000 : 0;  // kHz of wall_clock64
    return hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetLastError()
{
    hipError_t e = emu::t_last;
    emu::t_last = hipSuccess;
    return e;
}
const char *hipGetErrorString(hipError_t e)
{
    switch (e) {
    case hipSuccess: return "no error";
    case hipErrorInvalidValue: return "invalid argument";
    case hipErrorOutOfMemory: return "out of memory";
    case hipErrorNotReady: return "not ready";
    case hipErrorInvalidDevice: return "invalid device ordinal";
    default: return "unknown error";
    }
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned)
{
    *s = new emu_stream();
    return hipSuccess;
}
hipError_t hipStreamCreate(hipStream_t *s) { return hipStreamCreateWithFlags(s, 0); }
hipError_t hipStreamDestroy(hipStream_t s)
{
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s)
{
    if (s && s->capturing) return hipErrorInvalidValue;
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t *e)
{
    *e = new emu_event();
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e)
{
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s)
{
    if (!e) return hipErrorInvalidValue;
    return submit(s, [e]() { e->t = emu::now_ticks(); });
}
hipError_t hipEventSynchronize(hipEvent_t e) { return e ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipEventQuery(hipEvent_t e) { return e ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b)
{
    if (!a || !b) return hipErrorInvalidValue;
    *ms = (float)((double)(b->t - a->t) * 1e-5);
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode)
{
    if (!s || s->capturing) return hipErrorInvalidValue;
    s->capturing = true;
    s->graph = new emu_graph();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g)
{
    if (!s || !s->capturing) return hipErrorInvalidValue;
    s->capturing = false;
    *g = s->graph;
    s->graph = nullptr;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *x, hipGraph_t g, hipGraphNode_t *, char *, size_t)
{
    if (!g) return hipErrorInvalidValue;
    *x = new emu_graph(*g);
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s)
{
    if (!x) return hipErrorInvalidValue;
    hipError_t r = hipSuccess;
    for (auto &op : x->ops) {
        hipError_t e = submit(s, op);
        if (e != hipSuccess && r == hipSuccess) r = e;
    }
    return r;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t x)
{
    delete x;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g)
{
    delete g;
    return hipSuccess;
}
hipError_t hipGetSymbolAddress(void **p, const void *symbol)
{
    *p = const_cast<void *>(symbol);
    return hipSuccess;
}
