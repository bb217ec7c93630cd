// Request server: continuous batching of independent solves on one device.
//
// Requests (a grid, its options, an interval and the initial q) are queued
// by sc_server_submit.  A scheduler thread admits them in arrival order
// while their device footprint fits the free space of one device arena and
// an active slot is free, and advances every admitted solve concurrently:
//
//   1D solves  in rounds: per round, one launch of the persistent batch
//              kernel for each group of solves with identical options (its
//              own stream, so groups overlap), every solve advancing a
//              bounded number of time-loop iterations from its saved loop
//              state (sc_batch_kernel.h).  Requests admitted meanwhile join
//              at the next round; finished ones leave at the end of the
//              round that finished them.
//   2D solves  in rounds too: per round, each group of 2D grids with equal
//              options advances K time steps with no host involvement --
//              every stage launch covers the tiles of all the group's
//              grids, and a control kernel takes each grid's time-loop
//              decisions between steps (sc_batch2d.h).
//
// Each group runs its rounds back to back on its own stream, independently
// of the other groups (the scheduler polls the groups' round events), so a
// slow group does not hold up the others' arrivals and completions.
//
// A request's arithmetic is the same whatever else is in flight (the loop
// state is exact across rounds, and no reduction mixes solves), so its
// result alone and in a mix is the same.  Device memory is one hipMalloc at
// creation; admission, release and placement are host bookkeeping (first
// fit over 256-byte units with coalescing), so the hot path never calls the
// allocator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/sharpclaw_amd.h"
#include "sc_batch2d.h"
#include "sc_batch_host.h"
#include "sc_internal.h"

using namespace sc;

namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator over [0, size) in 256-byte units.
class Arena {
  public:
    void init(size_t bytes)
    {
        size_ = bytes / kUnit * kUnit;
        free_.clear();
        if (size_) free_[0] = size_;
        used_ = peak_ = 0;
    }
    static size_t round(size_t bytes) { return (bytes + kUnit - 1) / kUnit * kUnit; }
    // offset in bytes, or -1 when no free run is large enough
    long long alloc(size_t bytes)
    {
        bytes = round(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > bytes) free_[off + bytes] = len - bytes;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return (long long)off;
        }
        return -1;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = round(bytes);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t size() const { return size_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }

  private:
    static constexpr size_t kUnit = 256;
    size_t size_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;  // offset -> length
};

struct Req {
    long long id = 0;
    sc_params p;
    sc_run_ctl ctl;
    double tstart = 0, tend = 0, dt0 = 0;
    std::vector<double> q;  // initial data, then the result
    size_t bytes = 0;       // device footprint
    long long off = -1;     // arena offset (bytes)
    int status = 0;         // 0 queued, 1 active, 2 done
    int rc = SC_OK;
    int rounds = 0;
    BatchSolve bs;          // 1D: descriptor
    Grid2D g2;              // 2D: descriptor
    SolveState ss;          // loop state
    sc_run_stats run;
    double t_submit = 0, t_admit = 0, t_done = 0;
};

// 1D requests whose options select the same kernel and constants.
struct Group {
    sc_params p;
    PhysSel sel;
    PhysConst<double> cd;
    PhysConst<float> cf;
    int nt = 256;
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    std::vector<Req *> active;    // members of the round in flight / next round
    std::vector<Req *> incoming;  // admitted, join at the next launch
    bool inflight = false;
    int cap = 0;
    BatchSolve *d_solves = nullptr;
    SolveState *d_state = nullptr;
    int *d_order = nullptr;
    unsigned long long *d_next = nullptr;  // work-queue counters of a round
    BatchSolve *hs = nullptr;   // pinned staging of the descriptors
    SolveState *hst = nullptr;
    int *ho = nullptr;
};

// Workgroup size of a 1D solve by its size: large solves take wide
// workgroups (fewer chunks per stage, so fewer dependent steps on their
// critical path), small ones narrow (fewer idle lanes).
int block_class(int mx)
{
    return mx >= 1024 ? 1024 : mx >= 384 ? 512 : 256;
}

// 2D requests whose options select the same kernels and constants.
struct Group2D {
    sc_params p;
    Batch2DRunner run;
    std::vector<Req *> active;
    std::vector<Req *> incoming;
    bool inflight = false;
};

bool same_group(const sc_params &a, const sc_params &b)
{
    return block_class(a.nx[0]) == block_class(b.nx[0]) && a.physics == b.physics && a.meqn == b.meqn && a.char_decomp == b.char_decomp &&
           a.precision == b.precision && a.time_integrator == b.time_integrator &&
           a.mthbc[0] == b.mthbc[0] && a.mthbc[1] == b.mthbc[1] && a.gamma == b.gamma &&
           a.grav == b.grav && a.u == b.u && a.rho == b.rho && a.bulk == b.bulk && a.efix == b.efix &&
           a.tfluct_solver == b.tfluct_solver && a.lim_type == b.lim_type &&
           (a.lim_type != 1 || std::memcmp(a.mthlim, b.mthlim, sizeof a.mthlim) == 0);
}

}  // namespace

struct sc_server {
    int dev = 0, ncu = 1, max_active = 1, budget_chunks = 512;
    double tick_ms = 1e-5;
    Clock::time_point t0;
    double *arena_base = nullptr;
    Arena arena;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> queue;
    std::unordered_map<long long, std::unique_ptr<Req>> reqs;
    std::deque<long long> completed;
    std::vector<std::unique_ptr<Group>> groups;      // scheduler thread only
    std::vector<std::unique_ptr<Group2D>> groups2d;  // scheduler thread only
    int n_active = 0, peak_active = 0;
    long long next_id = 1, submitted = 0, n_completed = 0, rounds = 0;
    bool stop = false;
    std::string fatal;
    std::thread sched;

    double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
};

namespace {

size_t grid_doubles(const sc_params &p)
{
    size_t n = (size_t)(p.nx[0] + 2 * p.mbc) * p.meqn;
    if (p.ndim == 2) n *= (size_t)(p.nx[1] + 2 * p.mbc);
    return n;
}

// Called with mu held: the request leaves the active set.
void finish_locked(sc_server *S, Req *r, int rc)
{
    if (r->off >= 0) S->arena.release((size_t)r->off, r->bytes);
    r->off = -1;
    r->rc = rc;
    r->status = 2;
    r->t_done = S->now_ms();
    S->n_active--;
    S->n_completed++;
    S->completed.push_back(r->id);
    S->cv_done.notify_all();
    S->cv_work.notify_all();
}

bool same_group2d(const sc_params &a, const sc_params &b)
{
    return a.physics == b.physics && a.meqn == b.meqn && a.char_decomp == b.char_decomp &&
           a.precision == b.precision && a.time_integrator == b.time_integrator &&
           a.mthbc[0] == b.mthbc[0] && a.mthbc[1] == b.mthbc[1] && a.mthbc[2] == b.mthbc[2] &&
           a.mthbc[3] == b.mthbc[3] && a.gamma == b.gamma && a.grav == b.grav && a.efix == b.efix;
}

Group2D *group2d_for(sc_server *S, const sc_params &p)
{
    for (auto &g : S->groups2d)
        if (same_group2d(g->p, p)) return g.get();
    auto g = std::make_unique<Group2D>();
    g->p = p;
    if (g->run.init(p)) return nullptr;
    S->groups2d.push_back(std::move(g));
    return S->groups2d.back().get();
}

Group *group_for(sc_server *S, const sc_params &p)
{
    for (auto &g : S->groups)
        if (same_group(g->p, p)) return g.get();
    auto g = std::make_unique<Group>();
    g->p = p;
    g->sel.phys = sc_phys_id(&p);
    g->sel.chr = p.char_decomp;
    g->sel.lim = p.lim_type;
    g->sel.fp32 = p.precision == SC_FP32;
    g->cd = sc_make_const<double>(&p);
    g->cf = sc_make_const<float>(&p);
    g->nt = std::min(block_class(p.nx[0]), batch_max_block(g->sel));
    if (const char *e = getenv("SC_SERVE_BLOCK")) {
        int v = atoi(e);
        for (int k = 0; k < NBATCH_BLOCKS; k++)
            if (kBatchBlocks[k] == v && v <= batch_max_block(g->sel)) g->nt = v;
    }
    if (hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&g->ev, hipEventDisableTiming) != hipSuccess ||
        hipMalloc(&g->d_next, sizeof(unsigned long long)) != hipSuccess)
        return nullptr;
    S->groups.push_back(std::move(g));
    return S->groups.back().get();
}

void free_group_buffers(Group *g)
{
    if (g->d_solves) (void)hipFree(g->d_solves);
    if (g->d_state) (void)hipFree(g->d_state);
    if (g->d_order) (void)hipFree(g->d_order);
    if (g->hs) (void)hipHostFree(g->hs);
    if (g->hst) (void)hipHostFree(g->hst);
    if (g->ho) (void)hipHostFree(g->ho);
    g->d_solves = nullptr;
    g->d_state = nullptr;
    g->d_order = nullptr;
    g->hs = nullptr;
    g->hst = nullptr;
    g->ho = nullptr;
    g->cap = 0;
}

bool ensure_cap(Group *g, int n)
{
    if (n <= g->cap) return true;
    int cap = std::max(n, 2 * g->cap);
    (void)hipStreamSynchronize(g->st);
    free_group_buffers(g);
    if (hipMalloc(&g->d_solves, cap * sizeof(BatchSolve)) != hipSuccess ||
        hipMalloc(&g->d_state, cap * sizeof(SolveState)) != hipSuccess ||
        hipMalloc(&g->d_order, cap * sizeof(int)) != hipSuccess ||
        hipHostMalloc(&g->hs, cap * sizeof(BatchSolve), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&g->hst, cap * sizeof(SolveState), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&g->ho, cap * sizeof(int), hipHostMallocDefault) != hipSuccess)
        return false;
    g->cap = cap;
    return true;
}

// Scheduler-side start of an admitted request (no lock held).
int start_request(sc_server *S, Req *r)
{
    if (r->p.ndim == 2) {
        Group2D *g = group2d_for(S, r->p);
        if (!g) return sc_fail(SC_ERR_HIP, "server: stream/event creation failed");
        int tx, ty;
        tile2d_shape(g->run.tile, &tx, &ty);
        Grid2D &G = r->g2;
        memset(&G, 0, sizeof G);
        G.off = r->off / (long long)sizeof(double);
        G.nx = r->p.nx[0];
        G.ny = r->p.nx[1];
        G.LX = G.nx + 6;
        G.LY = G.ny + 6;
        G.ntx = (G.nx + tx - 1) / tx;
        G.nty = (G.ny + ty - 1) / ty;
        G.dx = r->p.dx[0];
        G.dy = r->p.dx[1];
        G.tstart = r->tstart;
        G.tend = r->tend;
        G.ctl = r->ctl;
        int info0 = 0;
        loop_limits(r->ctl, r->tstart, r->tend, r->dt0, &G.maxn, &info0);
        r->ss = initial_state(r->tstart, r->dt0, 0, info0);
        hipError_t e = hipMemcpyAsync(S->arena_base + G.off, r->q.data(), r->q.size() * sizeof(double),
                                      hipMemcpyHostToDevice, g->run.st);
        if (e != hipSuccess) return sc_fail(SC_ERR_HIP, "server upload: %s", hipGetErrorString(e));
        g->incoming.push_back(r);
        return SC_OK;
    }
    Group *g = group_for(S, r->p);
    if (!g) return sc_fail(SC_ERR_HIP, "server: stream/event creation failed");
    const int n = r->p.nx[0];
    memset(&r->bs, 0, sizeof r->bs);
    r->bs.off = r->off / (long long)sizeof(double);
    r->bs.mx = n;
    r->bs.dx = r->p.dx[0];
    r->bs.tstart = r->tstart;
    r->bs.tend = r->tend;
    r->bs.ctl = r->ctl;
    int info0 = 0;
    loop_limits(r->ctl, r->tstart, r->tend, r->dt0, &r->bs.maxn, &info0);
    r->ss = initial_state(r->tstart, r->dt0, 0, info0);
    size_t len = (size_t)(n + 6) * r->p.meqn;
    hipError_t e = hipMemcpyAsync(S->arena_base + r->bs.off, r->q.data(), len * sizeof(double),
                                  hipMemcpyHostToDevice, g->st);
    if (e != hipSuccess) return sc_fail(SC_ERR_HIP, "server upload: %s", hipGetErrorString(e));
    g->incoming.push_back(r);
    return SC_OK;
}

void record_run(Req *r)
{
    sc_batch_result br;
    state_to_result(r->ss, 1.0, 0, &br);
    memset(&r->run, 0, sizeof r->run);
    r->run.t = br.t;
    r->run.cfl_maxused = br.cfl_maxused;
    r->run.cfl_last = br.cfl_last;
    r->run.dt_minused = br.dt_minused;
    r->run.dt_maxused = br.dt_maxused;
    r->run.dt_last = br.dt_last;
    r->run.steps_taken = br.steps_taken;
    r->run.info = br.info;
    r->run.rejected = br.rejected;
    r->run.stage_evals = br.stage_evals;
}

int launch_round2d(sc_server *S, Group2D *g)
{
    g->active.insert(g->active.end(), g->incoming.begin(), g->incoming.end());
    g->incoming.clear();
    const int n = (int)g->active.size();
    int rc = g->run.reserve(n);
    if (rc) return rc;
    long long tiles = 0;
    for (int i = 0; i < n; i++) {
        g->run.h_grids[i] = g->active[i]->g2;
        g->run.h_state[i] = g->active[i]->ss;
        tiles += (long long)g->active[i]->g2.ntx * g->active[i]->g2.nty;
    }
    const int K = round_iterations(stages_of_integrator(g->p.time_integrator), tiles);
    return g->run.launch_round(n, K, S->arena_base);
}

int collect_round2d(sc_server *S, Group2D *g)
{
    int rc = g->run.wait();
    if (rc) return rc;
    std::vector<Req *> keep, done;
    for (size_t i = 0; i < g->active.size(); i++) {
        Req *r = g->active[i];
        r->ss = g->run.h_state[i];
        r->rounds++;
        (r->ss.done ? done : keep).push_back(r);
    }
    for (Req *r : done) {
        const Grid2D &G = r->g2;
        const size_t n = (size_t)G.LX * G.LY * r->p.meqn;
        double *q = S->arena_base + G.off + (long long)r->ss.cur * n;
        hipError_t e = launch_bc(2, r->p.meqn, q, G.nx, G.ny, G.LX, G.LY, r->p.mthbc, g->run.st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(r->q.data(), q, n * sizeof(double), hipMemcpyDeviceToHost, g->run.st);
        if (e == hipSuccess) e = hipStreamSynchronize(g->run.st);
        if (e != hipSuccess) return sc_fail(SC_ERR_HIP, "server download: %s", hipGetErrorString(e));
        record_run(r);
    }
    g->active.swap(keep);
    if (!done.empty()) {
        std::lock_guard<std::mutex> lk(S->mu);
        for (Req *r : done) finish_locked(S, r, SC_OK);
    }
    return SC_OK;
}

int launch_round(sc_server *S, Group *g)
{
    g->active.insert(g->active.end(), g->incoming.begin(), g->incoming.end());
    g->incoming.clear();
    std::stable_sort(g->active.begin(), g->active.end(),
                     [](const Req *a, const Req *b) { return a->bs.mx > b->bs.mx; });
    const int n = (int)g->active.size();
    if (!ensure_cap(g, n)) return sc_fail(SC_ERR_NOMEM, "server: descriptor allocation failed");
    for (int i = 0; i < n; i++) {
        g->hs[i] = g->active[i]->bs;
        g->hst[i] = g->active[i]->ss;
        g->ho[i] = i;
    }
    BatchArgs A;
    memset(&A, 0, sizeof A);
    A.solves = g->d_solves;
    A.state = g->d_state;
    A.order = g->d_order;
    A.nsolves = n;
    A.next = g->d_next;
    A.ends = 1;
    A.base = S->arena_base;
    A.meqn = g->p.meqn;
    A.bc[0] = g->p.mthbc[0];
    A.bc[1] = g->p.mthbc[1];
    A.ti = g->p.time_integrator;
    // rounds of similar length whatever the workgroup size
    A.budget_chunks = std::max(1, S->budget_chunks * 256 / g->nt);
    hipError_t e = hipMemcpyAsync(g->d_solves, g->hs, n * sizeof(BatchSolve), hipMemcpyHostToDevice, g->st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(g->d_state, g->hst, n * sizeof(SolveState), hipMemcpyHostToDevice, g->st);
    if (e == hipSuccess) e = hipMemcpyAsync(g->d_order, g->ho, n * sizeof(int), hipMemcpyHostToDevice, g->st);
    if (e == hipSuccess) e = hipMemsetAsync(g->d_next, 0, sizeof(unsigned long long), g->st);
    if (e == hipSuccess) e = launch_batch(g->sel, g->nt, A, g->cd, g->cf, S->ncu, 0, g->st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(g->hst, g->d_state, n * sizeof(SolveState), hipMemcpyDeviceToHost, g->st);
    if (e == hipSuccess) e = hipEventRecord(g->ev, g->st);
    if (e != hipSuccess) return sc_fail(SC_ERR_HIP, "server round: %s", hipGetErrorString(e));
    return SC_OK;
}

int collect_round(sc_server *S, Group *g)
{
    hipError_t e = hipEventSynchronize(g->ev);
    if (e != hipSuccess) return sc_fail(SC_ERR_HIP, "server round: %s", hipGetErrorString(e));
    std::vector<Req *> keep, done;
    for (size_t i = 0; i < g->active.size(); i++) {
        Req *r = g->active[i];
        r->ss = g->hst[i];
        r->rounds++;
        (r->ss.done ? done : keep).push_back(r);
    }
    for (Req *r : done) {
        const int n = r->bs.mx;
        size_t len = (size_t)(n + 6) * r->p.meqn;
        e = hipMemcpyAsync(r->q.data(), S->arena_base + r->bs.off + (long long)r->ss.cur * len,
                           len * sizeof(double), hipMemcpyDeviceToHost, g->st);
        if (e == hipSuccess) e = hipStreamSynchronize(g->st);
        if (e != hipSuccess) return sc_fail(SC_ERR_HIP, "server download: %s", hipGetErrorString(e));
        fill_ghosts_1d(r->q.data(), n, r->p.meqn, r->p.mthbc[0], r->p.mthbc[1]);
        record_run(r);
    }
    g->active.swap(keep);
    if (!done.empty()) {
        std::lock_guard<std::mutex> lk(S->mu);
        for (Req *r : done) finish_locked(S, r, SC_OK);
    }
    return SC_OK;
}

void fail_all(sc_server *S, const char *msg)
{
    std::lock_guard<std::mutex> lk(S->mu);
    S->fatal = msg;
    for (auto &g : S->groups) {
        for (Req *r : g->active) finish_locked(S, r, SC_ERR_HIP);
        for (Req *r : g->incoming) finish_locked(S, r, SC_ERR_HIP);
        g->active.clear();
        g->incoming.clear();
    }
    for (auto &g : S->groups2d) {
        for (Req *r : g->active) finish_locked(S, r, SC_ERR_HIP);
        for (Req *r : g->incoming) finish_locked(S, r, SC_ERR_HIP);
        g->active.clear();
        g->incoming.clear();
    }
    while (!S->queue.empty()) {
        Req *r = S->queue.front();
        S->queue.pop_front();
        S->n_active++;  // balanced by finish_locked
        finish_locked(S, r, SC_ERR_HIP);
    }
    S->stop = true;
}

// ready: the group's round in flight has finished (or failed: < 0)
int round_done(hipEvent_t ev)
{
    hipError_t e = hipEventQuery(ev);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    sc_fail(SC_ERR_HIP, "server round: %s", hipGetErrorString(e));
    return -1;
}

// Each group runs its rounds back to back on its own stream, independently
// of the other groups: a group is relaunched as soon as its own round has
// been collected, with the requests admitted meanwhile.
void sched_loop(sc_server *S)
{
    if (hipSetDevice(S->dev) != hipSuccess) {
        fail_all(S, "hipSetDevice in the scheduler thread failed");
        return;
    }
    for (;;) {
        std::vector<Req *> admitted;
        bool busy = false;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->stop) return;
            // admission in arrival order: footprint and a free slot
            while (!S->queue.empty() && S->n_active < S->max_active) {
                Req *r = S->queue.front();
                long long off = S->arena.alloc(r->bytes);
                if (off < 0) break;
                S->queue.pop_front();
                r->off = off;
                r->status = 1;
                r->t_admit = S->now_ms();
                S->n_active++;
                S->peak_active = std::max(S->peak_active, S->n_active);
                admitted.push_back(r);
            }
            busy = !admitted.empty();
            for (auto &g : S->groups) busy = busy || g->inflight || !g->active.empty() || !g->incoming.empty();
            for (auto &g : S->groups2d) busy = busy || g->inflight || !g->active.empty() || !g->incoming.empty();
            if (!busy) {
                S->cv_work.wait_for(lk, std::chrono::milliseconds(20));
                continue;
            }
        }
        for (Req *r : admitted) {
            int rc = start_request(S, r);
            if (rc) {
                std::lock_guard<std::mutex> lk(S->mu);
                finish_locked(S, r, rc);
            }
        }
        bool progressed = false;
        long long finished_rounds = 0;
        for (auto &g : S->groups2d) {
            if (g->inflight || (g->active.empty() && g->incoming.empty())) continue;
            if (launch_round2d(S, g.get())) {
                fail_all(S, "server round launch failed");
                return;
            }
            g->inflight = true;
            progressed = true;
        }
        for (auto &g : S->groups) {
            if (g->inflight || (g->active.empty() && g->incoming.empty())) continue;
            if (launch_round(S, g.get())) {
                fail_all(S, "server round launch failed");
                return;
            }
            g->inflight = true;
            progressed = true;
        }
        for (auto &g : S->groups) {
            if (!g->inflight) continue;
            int d = round_done(g->ev);
            if (d < 0 || (d > 0 && collect_round(S, g.get()))) {
                fail_all(S, "server round failed");
                return;
            }
            if (d > 0) {
                g->inflight = false;
                progressed = true;
                finished_rounds++;
            }
        }
        for (auto &g : S->groups2d) {
            if (!g->inflight) continue;
            int d = round_done(g->run.ev);
            if (d < 0 || (d > 0 && collect_round2d(S, g.get()))) {
                fail_all(S, "server round failed");
                return;
            }
            if (d > 0) {
                g->inflight = false;
                progressed = true;
                finished_rounds++;
            }
        }
        std::unique_lock<std::mutex> lk(S->mu);
        S->rounds += finished_rounds;
        if (!progressed) S->cv_work.wait_for(lk, std::chrono::microseconds(50));
    }
}

}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" int sc_server_create(int device, size_t hbm_bytes, int max_active, sc_server **out)
{
    if (!out || max_active < 1 || hbm_bytes < 4096) return sc_fail(SC_ERR_ARG, "bad argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return sc_fail(SC_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) return sc_fail(SC_ERR_ARG, "device out of range");
    int prev = 0;
    (void)hipGetDevice(&prev);
    hipError_t se = hipSetDevice(device);
    if (se != hipSuccess) return sc_fail(SC_ERR_HIP, "hipSetDevice(%d): %s", device, hipGetErrorString(se));
    auto S = std::make_unique<sc_server>();
    S->dev = device;
    S->max_active = max_active;
    S->t0 = Clock::now();
    hipDeviceProp_t prop;
    (void)hipGetDeviceProperties(&prop, device);
    S->ncu = prop.multiProcessorCount;
    int khz = 0;
    (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device);
    S->tick_ms = khz > 0 ? 1.0 / khz : 1e-5;
    if (const char *e = getenv("SC_SERVE_BUDGET")) {
        int v = atoi(e);
        if (v > 0) S->budget_chunks = v;
    }
    hipError_t e = hipMalloc(&S->arena_base, hbm_bytes);
    (void)hipSetDevice(prev);
    if (e != hipSuccess) return sc_fail(SC_ERR_NOMEM, "server arena of %zu bytes: %s", hbm_bytes, hipGetErrorString(e));
    S->arena.init(hbm_bytes);
    sc_server *s = S.release();
    s->sched = std::thread(sched_loop, s);
    *out = s;
    return SC_OK;
}

extern "C" int sc_server_submit(sc_server *S, const sc_params *p, const sc_run_ctl *ctl, double tstart,
                                double tend, double dt_initial, const double *q, long long *id)
{
    if (!S || !p || !ctl || !q || !id) return sc_fail(SC_ERR_ARG, "bad argument");
    int rc = sc_check_params(p, true);
    if (rc) return rc;
    if (p->device != S->dev) return sc_fail(SC_ERR_ARG, "request device %d, server device %d", p->device, S->dev);
    if (p->maux != 0 || p->src_term != 0)
        return sc_fail(SC_ERR_UNSUPPORTED, "served solves take no aux arrays and no source term");
    if (p->ndim == 2 && (p->lim_type != 2 || (p->char_decomp != 0 && p->char_decomp != 2)))
        return sc_fail(SC_ERR_UNSUPPORTED, "served 2D solves reconstruct with weno5 or weno5_char (lim_type 2, "
                                           "char_decomp 0 or 2); the others run on single 2D grids");
    if (p->time_integrator < 1 || p->time_integrator > 4)
        return sc_fail(SC_ERR_ARG, "time_integrator should be between 1 and 4");
    for (int k = 0; k < 2 * p->ndim; k++)
        if (p->mthbc[k] < 1 || p->mthbc[k] > 3)
            return sc_fail(SC_ERR_UNSUPPORTED, "served solves need mthbc 1, 2 or 3");
    auto r = std::make_unique<Req>();
    r->p = *p;
    if (r->p.ndim == 1) {
        r->p.nx[1] = 1;
        r->p.dx[1] = 1.0;
    }
    r->ctl = *ctl;
    r->tstart = tstart;
    r->tend = tend;
    r->dt0 = dt_initial;
    const size_t n = grid_doubles(r->p);
    r->q.assign(q, q + n);
    r->bytes = 4 * n * sizeof(double);  // 3 q buffers + the SSP(10,4) register
    memset(&r->run, 0, sizeof r->run);
    std::lock_guard<std::mutex> lk(S->mu);
    if (!S->fatal.empty()) return sc_fail(SC_ERR_HIP, "server stopped: %s", S->fatal.c_str());
    if (Arena::round(r->bytes) > S->arena.size())
        return sc_fail(SC_ERR_NOMEM, "request needs %zu device bytes, server holds %zu", r->bytes, S->arena.size());
    r->id = S->next_id++;
    r->t_submit = S->now_ms();
    *id = r->id;
    S->queue.push_back(r.get());
    S->reqs[r->id] = std::move(r);
    S->submitted++;
    S->cv_work.notify_all();
    return SC_OK;
}

extern "C" int sc_server_poll(sc_server *S, long long *ids, int max_ids, int *n)
{
    if (!S || !n || (max_ids > 0 && !ids)) return sc_fail(SC_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    int k = 0;
    while (k < max_ids && !S->completed.empty()) {
        ids[k++] = S->completed.front();
        S->completed.pop_front();
    }
    *n = k;
    if (!S->fatal.empty()) return sc_fail(SC_ERR_HIP, "server stopped: %s", S->fatal.c_str());
    return SC_OK;
}

extern "C" int sc_server_wait(sc_server *S, double timeout_ms, int *n_ready)
{
    if (!S || !n_ready) return sc_fail(SC_ERR_ARG, "bad argument");
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms),
                        [&] { return !S->completed.empty() || !S->fatal.empty(); });
    *n_ready = (int)S->completed.size();
    return SC_OK;
}

extern "C" int sc_server_result(sc_server *S, long long id, double *q, sc_request_stats *st)
{
    if (!S || !q || !st) return sc_fail(SC_ERR_ARG, "bad argument");
    std::unique_ptr<Req> r;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->reqs.find(id);
        if (it == S->reqs.end()) return sc_fail(SC_ERR_ARG, "unknown request %lld", id);
        if (it->second->status != 2) return sc_fail(SC_ERR_STATE, "request %lld not finished", id);
        r = std::move(it->second);
        S->reqs.erase(it);
    }
    memcpy(q, r->q.data(), r->q.size() * sizeof(double));
    memset(st, 0, sizeof *st);
    st->run = r->run;
    st->submit_ms = r->t_submit;
    st->admit_ms = r->t_admit;
    st->done_ms = r->t_done;
    st->footprint = (long long)r->bytes;
    st->rounds = r->rounds;
    if (r->rc) return sc_fail(r->rc, "request %lld failed", id);
    return SC_OK;
}

extern "C" int sc_server_drain(sc_server *S, double timeout_ms)
{
    if (!S) return sc_fail(SC_ERR_ARG, "bad argument");
    std::unique_lock<std::mutex> lk(S->mu);
    bool ok = S->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), [&] {
        return (S->queue.empty() && S->n_active == 0) || !S->fatal.empty();
    });
    if (!S->fatal.empty()) return sc_fail(SC_ERR_HIP, "server stopped: %s", S->fatal.c_str());
    if (!ok) return sc_fail(SC_ERR_STATE, "drain timed out with %d active, %zu queued", S->n_active, S->queue.size());
    return SC_OK;
}

extern "C" int sc_server_get_stats(sc_server *S, sc_server_stats *st)
{
    if (!S || !st) return sc_fail(SC_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    memset(st, 0, sizeof *st);
    st->submitted = S->submitted;
    st->completed = S->n_completed;
    st->rounds = S->rounds;
    st->peak_bytes = (long long)S->arena.peak();
    st->capacity_bytes = (long long)S->arena.size();
    st->used_bytes = (long long)S->arena.used();
    st->active = S->n_active;
    st->queued = (int)S->queue.size();
    st->peak_active = S->peak_active;
    st->clock_ms = S->now_ms();
    return SC_OK;
}

// Outstanding requests are abandoned: the scheduler ends after its current
// round.
extern "C" int sc_server_destroy(sc_server *S)
{
    if (!S) return SC_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv_work.notify_all();
    }
    if (S->sched.joinable()) S->sched.join();
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(S->dev);
    for (auto &g : S->groups2d) g->run.destroy();
    for (auto &g : S->groups) {
        if (g->st) (void)hipStreamSynchronize(g->st);
        free_group_buffers(g.get());
        if (g->d_next) (void)hipFree(g->d_next);
        if (g->ev) (void)hipEventDestroy(g->ev);
        if (g->st) (void)hipStreamDestroy(g->st);
    }
    if (S->arena_base) (void)hipFree(S->arena_base);
    (void)hipSetDevice(prev);
    delete S;
    return SC_OK;
}
