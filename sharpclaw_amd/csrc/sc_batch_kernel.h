// Batched solves, device side: the persistent kernel that advances many
// independent 1D grids of uneven size in one launch, each by a bounded
// number of time-loop iterations from its saved loop state.
//
// Each workgroup takes solves from a device work queue, largest first
// (the host passes the order), and runs a solve's whole time loop on the
// device: the sharpclaw() loop of src/sharpclaw.f90:35-235 (dt clamp to
// tend, qold save, SSP stages, CFL-driven dt update, accept or
// reject-and-retake), every stage a sweep of stage1d_chunk over the grid.
// A large solve can be split over k workgroups ("parts", consecutive queue
// entries): part p sweeps its own contiguous range of the stage's chunks,
// the k parts meet at a device-memory barrier after every stage, and the
// CFL of the last stage is the maximum over the parts.  Every part runs
// the same (uniform) control logic on the same numbers, so a split solve
// computes exactly what one workgroup computes.  A launch of one split
// solve can instead give every part a whole step between barriers: part p
// owns a contiguous range of cells and computes stage s of a step on that
// range widened by 3 cells per stage still to come (the stencil reach of a
// cell's update), in private stage buffers, so only the step's last stage
// writes the shared q and only the step's CFL maximum needs a barrier; the
// overlap cells are computed twice, by the same arithmetic.
// State stays in global memory (three rotating q buffers per solve plus the
// SSP(10,4) register), where the workgroup's L2 keeps it hot between stages;
// the block barrier that ends each chunk makes a stage's writes visible to
// the next stage.  Every thread evaluates the (uniform) control logic
// itself from the block-reduced CFL, so no broadcast is needed.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/sharpclaw_amd.h"
#include "sc_stage.h"

namespace sc {

// One solve of a launch: its grid, its slab of device memory and the
// options of its time loop.
struct BatchSolve {
    long long off;      // first double of this solve's buffers (4 x len)
    int mx;
    int maxn;           // iteration limit of the whole loop (max_steps, or the
                        // fixed-dt step count, as sharpclaw.f90 computes it)
    double dx;
    double tstart, tend;
    sc_run_ctl ctl;
};

// The loop state of one solve, kept in device memory between launches so a
// solve can be advanced a bounded number of iterations per launch and
// resumed later with identical arithmetic.
struct SolveState {
    double t, told, dt, cfl, cflmax, dtmin, dtmax;
    long long stages, t_begin, t_end;
    int cur, old, retake, steps, rejected, info, it, done;
    int go;             // 2D rounds: the current iteration's stages run
    int nbar;           // barriers a split solve's parts have passed
};

// A queue entry: solve index (bits 0-19), part (bits 20-25) and part
// count - 1 (bits 26-31); a plain solve index is a solve of one part.
constexpr int PART_SHIFT = 20, PARTS_SHIFT = 26, MAX_PARTS = 32;
constexpr int SOLVE_MASK = (1 << PART_SHIFT) - 1;
constexpr int ENTRY_MAX_SOLVES = 1 << PART_SHIFT;
__host__ __device__ inline int queue_entry(int solve, int part, int parts)
{
    return solve | (part << PART_SHIFT) | ((parts - 1) << PARTS_SHIFT);
}

// info of a split solve whose parts did not all reach a barrier within
// BatchArgs::spin_ticks (the solve is abandoned; the others continue)
constexpr int INFO_GANG_TIMEOUT = 14;

struct BatchArgs {
    const BatchSolve *solves;
    SolveState *state;
    const int *order;   // queue entries of this launch (queue_entry), largest first
    int nsolves;        // entries of order
    // work-queue counters: low 32 bits = entries taken from the front of
    // order (largest first), high 32 bits = entries taken from the back
    unsigned long long *next;
    int ends;           // 1: every workgroup takes from the front; k > 1:
                        // workgroups b with b % k == k - 1 take from the back
                        // (smallest first)
    double *base;
    int meqn;
    int bc[2];
    int ti;             // time integrator 1..4
    int step_limit;     // > 0: at most this many iterations per solve
    int budget_chunks;  // > 0: at most max(1, budget_chunks / (chunks x stages))
                        // iterations, chunks = ceil(mx / (NT - 1))
    // split solves: per solve four words (arrivals at its barrier, the CFL
    // maximum of even / odd iterations, unused), zero at a solve's start
    unsigned long long *sync;
    long long spin_ticks;  // longest wait at a barrier (wall_clock64 ticks)
    // non-null: the launch is the parts of one split solve (nsolves parts of
    // solve 0, boundaries without periodic wrap), which meet once per step;
    // part p's three private stage buffers (two stage values, the SSP(10,4)
    // register) are the len-double buffers 3p..3p+2 from priv
    double *priv;
};

struct Plan {
    int n;
    int op[10];
};

__device__ __forceinline__ Plan stage_plan(int ti)
{
    Plan P;
    switch (ti) {
    case 1: P = {2, {OP_ADD1, OP_SSP22_2}}; break;
    case 2: P = {4, {OP_ADD_THIRD, OP_ADD_THIRD, OP_ADD_THIRD, OP_SSP42_4}}; break;
    case 3: P = {3, {OP_ADD1, OP_SSP33_2, OP_SSP33_3}}; break;
    default:
        P = {10, {OP_ADD_SIXTH, OP_ADD_SIXTH, OP_ADD_SIXTH, OP_ADD_SIXTH, OP_SSP104_5, OP_ADD_SIXTH,
                  OP_ADD_SIXTH, OP_ADD_SIXTH, OP_ADD_SIXTH, OP_SSP104_10}};
    }
    return P;
}

__device__ __forceinline__ int other_buf(int a, int b)
{
    for (int k = 0; k < 3; k++)
        if (k != a && k != b) return k;
    return 0;
}

// Index into A.order of the next solve for this workgroup, A.nsolves when
// the queue is empty.  One 64-bit atomic claims an entry from either end:
// with f front and b back entries taken before it, the claim is valid while
// f + b < nsolves, so the two ends never take the same entry and exactly
// nsolves claims succeed.  With ends = k > 1 the largest solves still start
// at once on most workgroups while every k-th one finishes the small solves
// first, so completion latencies are low for most solves.
__device__ __forceinline__ int fetch_solve(const BatchArgs &A)
{
    const bool back = A.ends > 1 && (int)(blockIdx.x % A.ends) == A.ends - 1;
    const unsigned long long old = atomicAdd(A.next, back ? (1ull << 32) : 1ull);
    const unsigned f = (unsigned)(old & 0xffffffffull), b = (unsigned)(old >> 32);
    if ((unsigned long long)f + b >= (unsigned long long)A.nsolves) return A.nsolves;
    return back ? A.nsolves - 1 - (int)b : (int)f;
}

// Barrier of the parts of one split solve: every part adds one arrival to
// the solve's counter and waits until it reaches target (= parts x barriers
// passed so far).  The fences make each workgroup's global writes of the
// stage visible to the other parts before they read them (release before
// the arrival, acquire after the wait).  The wait is bounded by spin_ticks
// of the wall clock; false when it ran out.
__device__ __forceinline__ bool gang_sync(const BatchArgs &A, unsigned long long *ctr, unsigned long long target,
                                          int *flag)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        atomicAdd(ctr, 1ull);
        const long long t0 = (long long)wall_clock64();
        int ok = 1;
        while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if ((long long)wall_clock64() - t0 > A.spin_ticks) {
                ok = 0;
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *flag = ok;
    }
    __syncthreads();
    const bool ok = *flag != 0;
    __syncthreads();
    return ok;
}

// At least 4 waves per SIMD: the float64 variants otherwise take a few more
// than 128 VGPRs and drop to 3 (25% fewer resident workgroups).
template <int P, int CHAR, class T, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void batch_kernel(BatchArgs A,
                                                                                           PhysConst<T> c)
{
    constexpr int M = Shape<P>::M;
    constexpr int NW = NT / 64;
    // the reconstructions with neighbour data (weno5_wave, weno5_trans,
    // tvd2_wave) hold it in sw, and the interface exchange reuses it when
    // it is large enough (as stage1d_kernel)
    constexpr int NB = nbr_words<P, 1, CHAR, T>();
    constexpr bool AL = sx_in_sw<P, 1, CHAR, T, NT>();
    __shared__ T sq[M][NT + 5];
    __shared__ __align__(8) T sx0[AL ? 1 : 2 * M][NT];
    __shared__ T sw[NB > 0 ? NB : 1][NB > 0 ? NT + 4 : 1];
    T(*sx)[NT] = AL ? reinterpret_cast<T(*)[NT]>(&sw[0][0]) : sx0;
    __shared__ double red[NW];
    __shared__ int sb, sflag;
    const int t = threadIdx.x;
    const Plan plan = stage_plan(A.ti);

    // Work queue.  The next solve is fetched in one place, at the end of the
    // loop body in the same lane-0 block as the result store, with a barrier
    // before the shared index is read.  (Two blocks guarded by the same lane
    // test on either side of the back-edge would be merged by jump threading
    // into a nested loop that the other lanes of the wave iterate without
    // lane 0 -- a livelock under exec masking.)
    if (t == 0) sb = fetch_solve(A);
    __syncthreads();
    // the queue index is uniform over the workgroup: read as such, the
    // solve's descriptor and loop state load into scalar registers
    int k = __builtin_amdgcn_readfirstlane(sb);
    __syncthreads();
    while (k < A.nsolves) {
        const int ent = A.order[k];
        const int b = ent & SOLVE_MASK;
        const int part = (ent >> PART_SHIFT) & (MAX_PARTS * 2 - 1);
        const int parts = ((unsigned)ent >> PARTS_SHIFT) + 1;
        const BatchSolve S = A.solves[b];
        SolveState s = A.state[b];
        const int n = S.mx;
        // this part's chunks of every stage: [ch0, ch1) of the nch chunks
        const int nch = (n + NT - 2) / (NT - 1);
        const int ch0 = (int)((long long)nch * part / parts), ch1 = (int)((long long)nch * (part + 1) / parts);
        unsigned long long *sync = parts > 1 ? A.sync + 4 * (size_t)b : nullptr;
        const size_t len = (size_t)(n + 6) * A.meqn;
        // the solve's buffer j (0-2: q buffers, 3: the SSP(10,4) register),
        // formed from the argument pointer each time: a pointer kept in a
        // dynamically indexed array is a generic pointer to the compiler,
        // whose accesses it would issue as flat instructions
        auto buf = [&](int j) { return A.base + S.off + (size_t)j * len; };
        if (s.t_begin == 0) s.t_begin = (long long)wall_clock64();
        // a whole step between barriers: this part's cells [lo, hi)
        const bool whole = sync && A.priv;
        const int lo = 1 + (int)((long long)n * part / parts), hi = 1 + (int)((long long)n * (part + 1) / parts);
        auto pbuf = [&](int j) { return A.priv + (size_t)(part * 3 + j) * len; };

        StageArgs a;
        a.nx = n;
        a.ny = 1;
        a.LX = n + 6;
        a.LY = 1;
        a.bc[0] = A.bc[0];
        a.bc[1] = A.bc[1];
        a.bc[2] = a.bc[3] = 1;
        a.smax = nullptr;
        a.ntx = a.nty = 1;
        a.qbin = buf(3);
        a.qbout = buf(3);
        a.capa = nullptr;  // batched solves have no aux and no source
        a.src = 0;
        a.dt = 0.0;
        a.xlo = 0.0;
        a.dxc = S.dx;
        a.tflag = nullptr;
        a.dflag = nullptr;

        int lim = 0x7fffffff;
        if (A.step_limit > 0) lim = A.step_limit;
        else if (A.budget_chunks > 0) lim = max(1, A.budget_chunks / (((n + NT - 2) / (NT - 1)) * plan.n));

        // src/sharpclaw.f90:35-235 for this solve, from its saved state
        const sc_run_ctl ctl = S.ctl;
        bool fin = s.done != 0;
        for (int used = 0; !fin && used < lim; used++) {
            if (s.info != 0 || s.it >= S.maxn) {
                fin = true;
                break;
            }
            s.it++;
            if (!s.retake) {
                s.told = s.t;
                if (s.told + s.dt > S.tend && S.tstart < S.tend) s.dt = S.tend - s.told;
                s.old = s.cur;  // qold = q
            }
            s.retake = 0;
            s.t = s.told + s.dt;
            a.dtdx = s.dt / S.dx;
            a.qold = buf(s.old);
            int in = s.cur;
            T sm = T(0);
            bool lost = false;  // a barrier of a split solve timed out
            for (int st = 0; st < plan.n; st++) {
                int outb = other_buf(in, s.old);
                a.op = plan.op[st];
                const bool last = st == plan.n - 1;
                if (whole) {
                    // the cells the remaining stages of this part read: the
                    // stage reads q of the step (shared) or the previous
                    // stage's private values, valid 3 cells further out
                    const int h = 3 * (plan.n - 1 - st);
                    const int rlo = max(1, lo - h), rhi = min(n + 1, hi + h);
                    a.qin = st == 0 ? buf(in) : pbuf((st - 1) & 1);
                    a.qout = last ? buf(outb) : pbuf(st & 1);
                    a.qbin = a.qbout = pbuf(2);
                    for (int c0 = rlo; c0 < rhi; c0 += NT - 1) {
                        T s1 = stage1d_chunk<P, 1, CHAR, T, NT>(a, c, c0, sq, sx, &sw[0][0], rhi - 1);
                        if (last) sm = fmax(sm, s1);
                    }
                    in = outb;
                    continue;
                }
                a.qin = buf(in);
                a.qout = buf(outb);
                for (int ch = ch0; ch < ch1; ch++) {
                    T s1 = stage1d_chunk<P, 1, CHAR, T, NT>(a, c, 1 + ch * (NT - 1), sq, sx, &sw[0][0]);
                    if (last) sm = fmax(sm, s1);
                }
                in = outb;
                if (sync && !last) {
                    s.nbar++;
                    lost = lost || !gang_sync(A, sync, (unsigned long long)parts * (unsigned)s.nbar, &sflag);
                }
            }
            s.stages += plan.n;
            // block maximum of |s| of the last stage -> CFL (flux1.f90:126-133)
            double d = (double)wave_max(sm);
            if ((t & 63) == 0) red[t >> 6] = d;
            __syncthreads();
            double m = red[0];
#pragma unroll
            for (int w = 1; w < NW; w++) m = fmax(m, red[w]);
            __syncthreads();
            if (sync) {
                // maximum over the parts: slot 1 + (it & 1) collects this
                // iteration's, and part 0 clears the other one, which every
                // part read before the first barrier of this iteration and
                // none writes before the first barrier of the next
                unsigned long long *slot = sync + 1 + (s.it & 1);
                if (t == 0 && m > 0.0) atomicMax(slot, (unsigned long long)__double_as_longlong(m));
                s.nbar++;
                lost = lost || !gang_sync(A, sync, (unsigned long long)parts * (unsigned)s.nbar, &sflag);
                if (t == 0) {
                    red[0] = __longlong_as_double(
                        (long long)__hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                    if (part == 0) __hip_atomic_store(sync + 1 + ((s.it + 1) & 1), 0ull, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT);
                }
                __syncthreads();
                m = red[0];
                __syncthreads();
            }
            if (lost) {
                s.info = INFO_GANG_TIMEOUT;
                fin = true;
                break;
            }
            s.cfl = (s.dt / S.dx) * m;
            if (ctl.dt_variable == 1) {
                if (s.cfl > 0.0) {
                    s.dt = fmin(ctl.dt_max, s.dt * ctl.cfl_desired / s.cfl);
                    s.dtmin = fmin(s.dt, s.dtmin);
                    s.dtmax = fmax(s.dt, s.dtmax);
                } else {
                    s.dt = ctl.dt_max;
                }
            }
            if (s.cfl <= ctl.cfl_max) {
                s.cflmax = fmax(s.cfl, s.cflmax);
                s.cur = in;
            } else {
                s.t = s.told;
                s.cur = s.old;  // q = qold
                s.rejected++;
                if (ctl.dt_variable == 1) {
                    s.retake = 1;
                } else {
                    s.cflmax = fmax(s.cfl, s.cflmax);
                    fin = true;
                    break;
                }
            }
            s.steps++;
            if (s.t >= S.tend) fin = true;
        }
        if (!fin && (s.info != 0 || s.it >= S.maxn)) fin = true;
        if (fin && !s.done) {
            if (s.info == 0 && ctl.dt_variable == 1 && s.t < S.tend && s.steps == S.maxn) s.info = 11;
            if (s.info == 0 && ctl.dt_variable == 0 && s.cflmax > ctl.cfl_max) s.info = 12;
            s.done = 1;
            s.t_end = (long long)wall_clock64();
        }
        if (t == 0) {
            if (part == 0) A.state[b] = s;
            sb = fetch_solve(A);
        }
        __syncthreads();
        k = __builtin_amdgcn_readfirstlane(sb);
        __syncthreads();
    }
}

}  // namespace sc
