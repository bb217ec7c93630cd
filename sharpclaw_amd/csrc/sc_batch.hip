// Batched solves, host side: the sc_batch_* C ABI over the persistent
// kernel of sc_batch_kernel.h, its launchers and the workgroup-size tuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <queue>
#include <vector>

#include "../../include/sharpclaw_amd.h"
#include "sc_batch_host.h"
#include "sc_internal.h"

namespace sc {

const int kBatchBlocks[NBATCH_BLOCKS] = {64, 128, 256, 512, 1024};

// What one barrier of a split solve's parts costs per stage (ms, measured
// on a 1600-cell grid in 7 parts: 4.6 us per stage against 1.9 us for one
// chunk), the tuner's estimate; and how long a part waits at a barrier
// before the solve is given up (INFO_GANG_TIMEOUT).
constexpr double GANG_BARRIER_MS = 0.003;
constexpr double GANG_WAIT_MS = 10000.0;
// most workgroups held by split solves in one run (bounds the queue)
constexpr int MAX_SPLIT_WG = 1024;

// ------------------------------------------------------------ launchers

// resident workgroups per CU of one instantiation (same on every device here)
template <int P, int CHAR, class T, int NT>
static int occ_of(void)
{
    static int occ = -1;
    if (occ < 0) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<P, CHAR, T, NT>, NT, 0) !=
                hipSuccess || nb < 1)
            nb = 1;
        occ = nb;
    }
    return occ;
}

template <int P, int CHAR, class T, int NT>
static hipError_t gob(const BatchArgs &A, const PhysConst<T> &c, int ncu, int max_wg, hipStream_t st)
{
    int nwg = std::min(A.nsolves, ncu * occ_of<P, CHAR, T, NT>());
    if (max_wg > 0) nwg = std::min(nwg, max_wg);
    if (nwg < 1) return hipSuccess;
    hipLaunchKernelGGL((batch_kernel<P, CHAR, T, NT>), dim3(nwg), dim3(NT), 0, st, A, c);
    return hipGetLastError();
}

template <int P, int CHAR, class T>
static int occ_nt(int nt)
{
    if constexpr (CHAR == RK_WENO || CHAR == RK_WENO_CHAR) {
        if (nt == 512) return occ_of<P, CHAR, T, 512>();
        if (nt == 1024) return occ_of<P, CHAR, T, 1024>();
    }
    switch (nt) {
    case 64: return occ_of<P, CHAR, T, 64>();
    case 128: return occ_of<P, CHAR, T, 128>();
    case 256: return occ_of<P, CHAR, T, 256>();
    }
    return 0;
}

template <int P, int CHAR>
static int occ_t(const PhysSel &s, int nt)
{
    return s.fp32 ? occ_nt<P, CHAR, float>(nt) : occ_nt<P, CHAR, double>(nt);
}

template <int P>
static int occ_p(const PhysSel &s, int nt)
{
    switch (batch_rk(s)) {
    case RK_WENO: return occ_t<P, RK_WENO>(s, nt);
    case RK_WENO_WAVE: return occ_t<P, RK_WENO_WAVE>(s, nt);
    case RK_WENO_CHAR: return occ_t<P, RK_WENO_CHAR>(s, nt);
    case RK_WENO_TRANS: return occ_t<P, RK_WENO_TRANS>(s, nt);
    case RK_TVD_WAVE: return occ_t<P, RK_TVD_WAVE>(s, nt);
    case RK_TVD_CHAR: return occ_t<P, RK_TVD_CHAR>(s, nt);
    }
    return 0;
}

// resident workgroups per CU; 0 for options or a size the kernel is not
// built for
int batch_occupancy(const PhysSel &s, int nt)
{
    switch (s.phys) {
    case ADV1: return occ_p<ADV1>(s, nt);
    case AC1: return occ_p<AC1>(s, nt);
    case EU1: return occ_p<EU1>(s, nt);
    case SW1: return occ_p<SW1>(s, nt);
    }
    return 0;
}

template <int P, int CHAR, class T>
static hipError_t gob_nt(int nt, const BatchArgs &A, const PhysConst<T> &c, int ncu, int max_wg,
                         hipStream_t st)
{
    if constexpr (CHAR == RK_WENO || CHAR == RK_WENO_CHAR) {
        if (nt == 512) return gob<P, CHAR, T, 512>(A, c, ncu, max_wg, st);
        if (nt == 1024) return gob<P, CHAR, T, 1024>(A, c, ncu, max_wg, st);
    }
    switch (nt) {
    case 64: return gob<P, CHAR, T, 64>(A, c, ncu, max_wg, st);
    case 128: return gob<P, CHAR, T, 128>(A, c, ncu, max_wg, st);
    case 256: return gob<P, CHAR, T, 256>(A, c, ncu, max_wg, st);
    }
    return hipErrorInvalidValue;
}

template <int P, int CHAR>
static hipError_t gob_t(const PhysSel &s, int nt, const BatchArgs &A, const PhysConst<double> &cd,
                        const PhysConst<float> &cf, int ncu, int max_wg, hipStream_t st)
{
    return s.fp32 ? gob_nt<P, CHAR, float>(nt, A, cf, ncu, max_wg, st)
                  : gob_nt<P, CHAR, double>(nt, A, cd, ncu, max_wg, st);
}

template <int P>
static hipError_t gob_p(const PhysSel &s, int nt, const BatchArgs &A, const PhysConst<double> &cd,
                        const PhysConst<float> &cf, int ncu, int max_wg, hipStream_t st)
{
    switch (batch_rk(s)) {
    case RK_WENO: return gob_t<P, RK_WENO>(s, nt, A, cd, cf, ncu, max_wg, st);
    case RK_WENO_WAVE: return gob_t<P, RK_WENO_WAVE>(s, nt, A, cd, cf, ncu, max_wg, st);
    case RK_WENO_CHAR: return gob_t<P, RK_WENO_CHAR>(s, nt, A, cd, cf, ncu, max_wg, st);
    case RK_WENO_TRANS: return gob_t<P, RK_WENO_TRANS>(s, nt, A, cd, cf, ncu, max_wg, st);
    case RK_TVD_WAVE: return gob_t<P, RK_TVD_WAVE>(s, nt, A, cd, cf, ncu, max_wg, st);
    case RK_TVD_CHAR: return gob_t<P, RK_TVD_CHAR>(s, nt, A, cd, cf, ncu, max_wg, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_batch(const PhysSel &s, int nt, const BatchArgs &A, const PhysConst<double> &cd,
                        const PhysConst<float> &cf, int ncu, int max_wg, hipStream_t st)
{
    switch (s.phys) {
    case ADV1: return gob_p<ADV1>(s, nt, A, cd, cf, ncu, max_wg, st);
    case AC1: return gob_p<AC1>(s, nt, A, cd, cf, ncu, max_wg, st);
    case EU1: return gob_p<EU1>(s, nt, A, cd, cf, ncu, max_wg, st);
    case SW1: return gob_p<SW1>(s, nt, A, cd, cf, ncu, max_wg, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace sc

using namespace sc;

// ------------------------------------------------------------------ C ABI

struct sc_batch {
    sc_params p;
    PhysSel sel;
    PhysConst<double> cd;
    PhysConst<float> cf;
    int dev, ncu, nsolves, block;
    int ends;                 // work-queue ends (BatchArgs::ends)
    hipStream_t st;
    std::vector<BatchSolve> solves;
    std::vector<SolveState> state;
    std::vector<int> order;   // solves, largest first
    std::vector<int> parts;   // workgroups per solve of the next run (1: not split)
    std::vector<int> queue;   // queue entries of the next run (queue_entry)
    std::vector<int> cur;     // buffer holding each solve's current q
    BatchSolve *d_solves;
    SolveState *d_state;
    int *d_order;
    int queue_cap;               // entries d_order holds
    unsigned long long *d_sync;  // split solves' barrier words (BatchArgs::sync)
    unsigned long long *d_next;  // work-queue counters
    double *d_base;
    size_t total;             // doubles allocated
    double tick_ms;           // wall_clock64 period in ms
    double predicted[NBATCH_BLOCKS];  // tuner's makespan estimates (ms)
    hipEvent_t ev0, ev1;
};

static void batch_free(sc_batch *b)
{
    if (!b) return;
    if (b->st) (void)hipStreamSynchronize(b->st);
    if (b->d_solves) (void)hipFree(b->d_solves);
    if (b->d_state) (void)hipFree(b->d_state);
    if (b->d_order) (void)hipFree(b->d_order);
    if (b->d_sync) (void)hipFree(b->d_sync);
    if (b->d_next) (void)hipFree(b->d_next);
    if (b->d_base) (void)hipFree(b->d_base);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
}

extern "C" int sc_batch_create(const sc_params *params, int nsolves, const int *mx, const double *xlower,
                               const double *xupper, sc_batch **out)
{
    if (!out || !mx || !xlower || !xupper || nsolves < 1) return sc_fail(SC_ERR_ARG, "bad argument");
    *out = nullptr;
    int rc = sc_check_params(params, false);
    if (rc) return rc;
    if (params->ndim != 1) return sc_fail(SC_ERR_UNSUPPORTED, "batched solves are 1D grids");
    if (params->maux != 0 || params->src_term != 0)
        return sc_fail(SC_ERR_UNSUPPORTED, "batched solves take no aux arrays and no source term");
    if (params->time_integrator < 1 || params->time_integrator > 4)
        return sc_fail(SC_ERR_ARG, "time_integrator should be between 1 and 4");
    for (int k = 0; k < 2; k++) {
        int m = params->mthbc[k];
        if (m < 1 || m > 3) return sc_fail(SC_ERR_UNSUPPORTED, "mthbc must be 1, 2 or 3");
        if (m == 3 && params->meqn < 2) return sc_fail(SC_ERR_ARG, "a wall needs meqn >= 2");
    }
    if ((params->mthbc[0] == 2) != (params->mthbc[1] == 2))
        return sc_fail(SC_ERR_ARG, "periodic boundary conditions require both sides to be 2");
    for (int i = 0; i < nsolves; i++)
        if (mx[i] < 3 || !(xupper[i] > xlower[i]))
            return sc_fail(SC_ERR_ARG, "solve %d: mx must be >= 3 and xupper > xlower", i);
    if (nsolves >= ENTRY_MAX_SOLVES) return sc_fail(SC_ERR_ARG, "at most %d solves per batch", ENTRY_MAX_SOLVES - 1);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return sc_fail(SC_ERR_HIP, "no HIP device available");
    if (params->device < 0 || params->device >= ndev) return sc_fail(SC_ERR_ARG, "device out of range");
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(params->device);
    sc_batch *b = new sc_batch();
    b->p = *params;
    b->sel.phys = sc_phys_id(params);
    b->sel.chr = params->char_decomp;
    b->sel.lim = params->lim_type;
    b->sel.fp32 = params->precision == SC_FP32;
    b->cd = sc_make_const<double>(params);
    b->cf = sc_make_const<float>(params);
    b->dev = params->device;
    b->nsolves = nsolves;
    b->block = 0;  // tuned at the first run
    // every 8th workgroup takes the smallest solves first (BatchArgs::ends);
    // measured on 4096 dam breaks against all-largest-first: the same
    // makespan, p50 latency 4.7x lower, p99 5% higher.  SC_BATCH_ENDS
    // overrides it (1: largest first only).
    b->ends = 8;
    if (const char *e = getenv("SC_BATCH_ENDS")) b->ends = std::max(1, atoi(e));
    hipDeviceProp_t prop;
    (void)hipGetDeviceProperties(&prop, b->dev);
    b->ncu = prop.multiProcessorCount;
    int khz = 0;
    (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->dev);
    b->tick_ms = khz > 0 ? 1.0 / khz : 1e-5;
    b->solves.resize(nsolves);
    long long off = 0;
    for (int i = 0; i < nsolves; i++) {
        memset(&b->solves[i], 0, sizeof(BatchSolve));
        b->solves[i].off = off;
        b->solves[i].mx = mx[i];
        b->solves[i].dx = (xupper[i] - xlower[i]) / (double)mx[i];
        off += 4LL * (mx[i] + 6) * params->meqn;
    }
    b->total = (size_t)off;
    b->order.resize(nsolves);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) { return mx[x] > mx[y]; });
    b->cur.assign(nsolves, 0);
    b->parts.assign(nsolves, 1);
    b->state.resize(nsolves);
    b->queue_cap = nsolves + MAX_SPLIT_WG;
    bool ok = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(&b->d_solves, nsolves * sizeof(BatchSolve)) == hipSuccess;
    ok = ok && hipMalloc(&b->d_state, nsolves * sizeof(SolveState)) == hipSuccess;
    ok = ok && hipMalloc(&b->d_order, b->queue_cap * sizeof(int)) == hipSuccess;
    ok = ok && hipMalloc(&b->d_sync, 4 * (size_t)nsolves * sizeof(unsigned long long)) == hipSuccess;
    ok = ok && hipMalloc(&b->d_next, sizeof(unsigned long long)) == hipSuccess;
    ok = ok && hipMalloc(&b->d_base, b->total * sizeof(double)) == hipSuccess;
    ok = ok && hipEventCreate(&b->ev0) == hipSuccess && hipEventCreate(&b->ev1) == hipSuccess;
    ok = ok && hipMemset(b->d_base, 0, b->total * sizeof(double)) == hipSuccess;
    (void)hipSetDevice(prev);
    if (!ok) {
        batch_free(b);
        return sc_fail(SC_ERR_NOMEM, "batch allocation of %zu doubles failed", (size_t)off);
    }
    *out = b;
    return SC_OK;
}

extern "C" int sc_batch_destroy(sc_batch *b)
{
    batch_free(b);
    return SC_OK;
}

extern "C" int sc_batch_set_q(sc_batch *b, int i, const double *q)
{
    if (!b || !q || i < 0 || i >= b->nsolves) return sc_fail(SC_ERR_ARG, "bad argument");
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(b->dev);
    size_t len = (size_t)(b->solves[i].mx + 6) * b->p.meqn;
    hipError_t e = hipMemcpy(b->d_base + b->solves[i].off, q, len * sizeof(double), hipMemcpyHostToDevice);
    b->cur[i] = 0;
    (void)hipSetDevice(prev);
    if (e != hipSuccess) return sc_fail(SC_ERR_HIP, "sc_batch_set_q: %s", hipGetErrorString(e));
    return SC_OK;
}

extern "C" int sc_batch_get_q(sc_batch *b, int i, double *q)
{
    if (!b || !q || i < 0 || i >= b->nsolves) return sc_fail(SC_ERR_ARG, "bad argument");
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(b->dev);
    size_t len = (size_t)(b->solves[i].mx + 6) * b->p.meqn;
    hipError_t e = hipMemcpy(q, b->d_base + b->solves[i].off + b->cur[i] * len, len * sizeof(double),
                             hipMemcpyDeviceToHost);
    (void)hipSetDevice(prev);
    if (e != hipSuccess) return sc_fail(SC_ERR_HIP, "sc_batch_get_q: %s", hipGetErrorString(e));
    // the device never stores ghost cells (stages synthesise them)
    fill_ghosts_1d(q, b->solves[i].mx, b->p.meqn, b->p.mthbc[0], b->p.mthbc[1]);
    return SC_OK;
}

// The queue of a run: the split solves first (largest first), their parts
// as consecutive entries, then the others largest first.
static void build_queue(sc_batch *b, const std::vector<int> &parts)
{
    b->queue.clear();
    for (int pass = 0; pass < 2; pass++)
        for (int i : b->order) {
            const int k = parts[i];
            if ((k > 1) != (pass == 0)) continue;
            for (int p = 0; p < k; p++) b->queue.push_back(queue_entry(i, p, k));
        }
}

// One launch over all solves from fresh loop states (q in b->cur), each
// solve split into parts[i] workgroups.
static int run_once(sc_batch *b, int nt, double tstart, double dt0, int info0, int step_limit,
                    const std::vector<int> &parts, float *ms)
{
    for (int i = 0; i < b->nsolves; i++) b->state[i] = initial_state(tstart, dt0, b->cur[i], info0);
    build_queue(b, parts);
    if ((int)b->queue.size() > b->queue_cap) return sc_fail(SC_ERR_ARG, "batch: too many parts");
    BatchArgs A;
    memset(&A, 0, sizeof A);
    A.solves = b->d_solves;
    A.state = b->d_state;
    A.order = b->d_order;
    A.nsolves = (int)b->queue.size();
    A.next = b->d_next;
    A.ends = b->ends;
    A.base = b->d_base;
    A.meqn = b->p.meqn;
    A.bc[0] = b->p.mthbc[0];
    A.bc[1] = b->p.mthbc[1];
    A.ti = b->p.time_integrator;
    A.step_limit = step_limit;
    A.budget_chunks = 0;
    A.sync = b->d_sync;
    A.spin_ticks = (long long)(GANG_WAIT_MS / b->tick_ms);
    hipError_t e = hipMemcpyAsync(b->d_state, b->state.data(), b->nsolves * sizeof(SolveState),
                                  hipMemcpyHostToDevice, b->st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_order, b->queue.data(), b->queue.size() * sizeof(int), hipMemcpyHostToDevice, b->st);
    if (e == hipSuccess) e = hipMemsetAsync(b->d_sync, 0, 4 * (size_t)b->nsolves * sizeof(unsigned long long), b->st);
    if (e == hipSuccess) e = hipMemsetAsync(b->d_next, 0, sizeof(unsigned long long), b->st);
    if (e == hipSuccess) e = hipEventRecord(b->ev0, b->st);
    if (e == hipSuccess) e = launch_batch(b->sel, nt, A, b->cd, b->cf, b->ncu, 0, b->st);
    if (e == hipSuccess) e = hipEventRecord(b->ev1, b->st);
    if (e == hipSuccess) e = hipEventSynchronize(b->ev1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, b->ev0, b->ev1);
    if (e == hipSuccess)
        e = hipMemcpy(b->state.data(), b->d_state, b->nsolves * sizeof(SolveState), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return sc_fail(SC_ERR_HIP, "batch run: %s", hipGetErrorString(e));
    for (int i = 0; i < b->nsolves; i++)
        if (b->state[i].info == INFO_GANG_TIMEOUT)
            return sc_fail(SC_ERR_HIP, "batch run: the parts of solve %d did not meet at a barrier within %g ms", i,
                           GANG_WAIT_MS);
    return SC_OK;
}

// Makespan of a whole run predicted for workgroup size nt from a trial of
// one time-loop iteration of every solve (trial_ms, st: the states after
// it), and the split of the largest solves over several workgroups that
// the prediction prefers (parts).  The trial gives the time c of one
// chunk-stage on one busy workgroup; the trial's CFL gives each solve's
// step count, (tend - tstart) / dt_next (maxn for fixed dt).  A solve of ch
// chunks split over k workgroups takes iters x stages x (ceil(ch / k) c +
// barrier) on each of them.  When there are fewer solves than resident
// workgroups, starting from no splits, the solve with the longest chain
// gets one more part while its chain is longer than the balanced load
// (total work / slots) and the split still shortens it (at most max_parts
// parts, at most half the slots and MAX_SPLIT_WG workgroups in split
// solves).  The parts are then placed largest first on the least loaded of
// the resident workgroups (the device work queue does the same), and the
// largest load is the makespan.
static double predict_ms(const sc_batch *b, int nt, float trial_ms, double tstart, int stages, int max_parts,
                         std::vector<int> &parts)
{
    const int n = b->nsolves;
    std::vector<double> its(n), ch(n);
    double trial = 0.0;
    for (int i = 0; i < n; i++) {
        const BatchSolve &S = b->solves[i];
        const SolveState &s = b->state[i];
        ch[i] = (double)((S.mx + nt - 2) / (nt - 1));
        double iters = S.maxn;
        if (S.ctl.dt_variable == 1 && s.dt > 0.0)
            iters = std::min((double)S.maxn, std::max(1.0, std::ceil((S.tend - tstart) / s.dt)));
        if (s.info != 0) iters = 0.0;
        trial += s.it > 0 ? ch[i] * stages : 0.0;
        its[i] = iters * stages;
    }
    parts.assign(n, 1);
    // resident workgroups (a launch has at most one per queue entry)
    const int slots = std::max(1, b->ncu * batch_occupancy(b->sel, nt));
    if (trial <= 0.0) return trial_ms;
    // the trial's own makespan in chunk-stages: its balanced load, or its
    // largest solve when there are fewer solves than slots
    double one = 0.0;
    for (int i = 0; i < n; i++)
        if (b->state[i].it > 0) one = std::max(one, ch[i] * stages);
    const double c = trial_ms / std::max(trial / std::min(n, slots), one);
    auto chain = [&](int i, int k) {
        return its[i] * (std::ceil(ch[i] / k) * c + (k > 1 ? GANG_BARRIER_MS : 0.0));
    };
    double total = 0.0;
    for (int i = 0; i < n; i++) total += chain(i, 1);
    // largest chain first: a max-heap of (chain, solve)
    std::priority_queue<std::pair<double, int>> top;
    for (int i = 0; i < n; i++) top.push({chain(i, 1), i});
    int extra = 0;
    // with a solve for every slot no workgroup idles at the start, and the
    // barriers' cache maintenance (L2 write-back and invalidation at device
    // scope) slows the workgroups around them: measured on 4096 dam breaks,
    // splitting the 26 largest solves in two made the run 38% longer
    while (max_parts > 1 && n < slots && !top.empty()) {
        const int i = top.top().second;
        const int k = parts[i];
        const double now = chain(i, k);
        if (now <= total / slots) break;
        const int kn = k + 1;
        const int in_split = extra + (k == 1 ? kn : 1);  // workgroups held by split solves
        if (kn > std::min((int)ch[i], max_parts) || 2 * in_split > slots || in_split > MAX_SPLIT_WG) break;
        const double next = chain(i, kn);
        if (next >= now) break;
        top.pop();
        total += kn * next - k * now;
        parts[i] = kn;
        extra = in_split;
        top.push({next, i});
    }
    std::vector<double> work;
    for (int i = 0; i < n; i++)
        for (int p = 0; p < parts[i]; p++) work.push_back(chain(i, parts[i]));
    std::sort(work.begin(), work.end(), std::greater<double>());
    std::priority_queue<double, std::vector<double>, std::greater<double>> load;
    for (int k = 0; k < std::min(slots, (int)work.size()); k++) load.push(0.0);
    double worst = 0.0;
    for (double w : work) {
        double l = load.top() + w;
        load.pop();
        load.push(l);
        worst = std::max(worst, l);
    }
    return worst;
}

// Most workgroups per split solve: SC_BATCH_PARTS=k (1: no splits).
static int max_parts_of()
{
    int m = MAX_PARTS;
    if (const char *e = getenv("SC_BATCH_PARTS")) m = std::min(MAX_PARTS, std::max(1, atoi(e)));
    return m;
}

// Autotuner of the workgroup size: each candidate runs one time-loop
// iteration of every solve from the current states, and the one with the
// smallest predicted makespan (predict_ms) is kept with its split of the
// largest solves.  A trial writes only the buffers other than the one
// holding q (and the SSP(10,4) register), and its loop states are
// discarded, so the solves are untouched.
static int tune_block(sc_batch *b, double tstart, double dt0, int info0)
{
    if (b->block) return SC_OK;
    int forced = 0;
    if (const char *e = getenv("SC_BATCH_BLOCK")) {
        int v = atoi(e);
        for (int k = 0; k < NBATCH_BLOCKS; k++)
            if (kBatchBlocks[k] == v) forced = v;
    }
    static const int nstage[] = {0, 2, 4, 3, 10};
    const std::vector<int> single(b->nsolves, 1);
    const int maxp = max_parts_of();
    int best = 256;
    double bestms = 1e300;
    std::vector<int> plan, best_plan(b->nsolves, 1);
    for (int k = 0; k < NBATCH_BLOCKS; k++) {
        if (forced && kBatchBlocks[k] != forced) continue;
        if (kBatchBlocks[k] > batch_max_block(b->sel)) {
            b->predicted[k] = -1.0;  // not built for these options
            continue;
        }
        float ms = 0.f;
        int rc = run_once(b, kBatchBlocks[k], tstart, dt0, info0, 1, single, &ms);
        if (rc) return rc;
        double pred = predict_ms(b, kBatchBlocks[k], ms, tstart, nstage[b->p.time_integrator], maxp, plan);
        b->predicted[k] = pred;
        if (pred < bestms) {
            bestms = pred;
            best = kBatchBlocks[k];
            best_plan = plan;
        }
    }
    b->block = best;
    b->parts = best_plan;
    return SC_OK;
}

extern "C" int sc_batch_run(sc_batch *b, const sc_run_ctl *ctl, double tstart, double tend,
                            double dt_initial, sc_batch_result *results, double *gpu_ms)
{
    if (!b || !ctl || !results) return sc_fail(SC_ERR_ARG, "bad argument");
    int maxn, info0;
    loop_limits(*ctl, tstart, tend, dt_initial, &maxn, &info0);
    for (int i = 0; i < b->nsolves; i++) {
        b->solves[i].maxn = maxn;
        b->solves[i].tstart = tstart;
        b->solves[i].tend = tend;
        b->solves[i].ctl = *ctl;
    }
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(b->dev);
    int rc = SC_OK;
    if (hipMemcpy(b->d_solves, b->solves.data(), b->nsolves * sizeof(BatchSolve), hipMemcpyHostToDevice) !=
        hipSuccess)
        rc = sc_fail(SC_ERR_HIP, "batch descriptors upload failed");
    if (!rc) rc = tune_block(b, tstart, dt_initial, info0);
    float ms = 0.f;
    if (!rc) rc = run_once(b, b->block, tstart, dt_initial, info0, 0, b->parts, &ms);
    (void)hipSetDevice(prev);
    if (rc) return rc;
    long long t0 = 0;
    bool first = true;
    for (int i = 0; i < b->nsolves; i++)
        if (b->state[i].t_begin && (first || b->state[i].t_begin < t0)) {
            t0 = b->state[i].t_begin;
            first = false;
        }
    for (int i = 0; i < b->nsolves; i++) {
        b->cur[i] = b->state[i].cur;
        state_to_result(b->state[i], b->tick_ms, t0, &results[i]);
    }
    if (gpu_ms) *gpu_ms = ms;
    return SC_OK;
}

extern "C" int sc_batch_block(sc_batch *b, int *block)
{
    if (!b || !block) return sc_fail(SC_ERR_ARG, "bad argument");
    *block = b->block;
    return SC_OK;
}

extern "C" int sc_batch_tuning(sc_batch *b, int max, int *blocks, double *predicted_ms, int *n)
{
    if (!b || !n || (max > 0 && (!blocks || !predicted_ms))) return sc_fail(SC_ERR_ARG, "bad argument");
    int k = 0;
    for (; k < NBATCH_BLOCKS && k < max; k++) {
        blocks[k] = kBatchBlocks[k];
        predicted_ms[k] = b->predicted[k];
    }
    *n = k;
    return SC_OK;
}

extern "C" int sc_batch_parts(sc_batch *b, int *parts)
{
    if (!b || !parts) return sc_fail(SC_ERR_ARG, "bad argument");
    for (int i = 0; i < b->nsolves; i++) parts[i] = b->parts[i];
    return SC_OK;
}
