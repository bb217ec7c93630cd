// Batched solves, device side: the persistent kernel that advances many
// independent 1D grids of uneven size in one launch.
//
// Each workgroup takes solves from a device work queue, largest first
// (the host passes the order), and runs a solve's whole time loop on the
// device: the sharpclaw() loop of src/sharpclaw.f90:35-235 (dt clamp to
// tend, qold save, SSP stages, CFL-driven dt update, accept or
// reject-and-retake), every stage a sweep of stage1d_chunk over the grid.
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

struct BatchSolve {
    long long off;   // first double of this solve's buffers (4 x len)
    int mx;
    double dx;
};

struct BatchArgs {
    const BatchSolve *solves;
    sc_batch_result *out;
    const int *order;   // solve indices, largest first
    const int *start;   // buffer holding each solve's q at launch
    int nsolves;
    int *next;          // work-queue counter
    double *base;
    int meqn;
    int bc[2];
    int ti;             // time integrator 1..4
    sc_run_ctl ctl;
    int maxn;           // step limit (sharpclaw.f90: mx steps for fixed dt)
    int info0;          // nonzero: option error found by the host, no steps
    double tstart, tend, dt0;
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

template <int P, int CHAR, class T, int NT>
__global__ __launch_bounds__(NT) void batch_kernel(BatchArgs A, PhysConst<T> c)
{
    constexpr int M = Shape<P>::M;
    constexpr int NW = NT / 64;
    __shared__ T sq[M][NT + 5];
    __shared__ T sx[2 * M][NT];
    __shared__ double red[NW];
    __shared__ int sb;
    const int t = threadIdx.x;
    const Plan plan = stage_plan(A.ti);

    // Work queue.  The next solve is fetched in one place, at the end of the
    // loop body in the same lane-0 block as the result store, with a barrier
    // before the shared index is read.  (Two blocks guarded by the same lane
    // test on either side of the back-edge would be merged by jump threading
    // into a nested loop that the other lanes of the wave iterate without
    // lane 0 -- a livelock under exec masking.)
    if (t == 0) sb = atomicAdd(A.next, 1);
    __syncthreads();
    int k = sb;
    __syncthreads();
    while (k < A.nsolves) {
        const int b = A.order[k];
        const BatchSolve S = A.solves[b];
        const int n = S.mx;
        const size_t len = (size_t)(n + 6) * A.meqn;
        double *buf[4];
#pragma unroll
        for (int j = 0; j < 4; j++) buf[j] = A.base + S.off + j * len;
        long long t_begin = (long long)wall_clock64();

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
        a.qbin = buf[3];
        a.qbout = buf[3];

        // src/sharpclaw.f90:35-235 for this solve
        const sc_run_ctl ctl = A.ctl;
        double tcur = A.tstart, dt = A.dt0, told = A.tstart, cflmax = 0.0, cfl = 0.0;
        double dtmin = dt, dtmax = dt;
        int cur = A.start[b], old = cur, steps = 0, rejected = 0, info = 0;
        long long stages = 0;
        bool retake = false;
        info = A.info0;
        for (int it = 1; info == 0 && it <= A.maxn; it++) {
            if (!retake) {
                told = tcur;
                if (told + dt > A.tend && A.tstart < A.tend) dt = A.tend - told;
                old = cur;  // qold = q
            }
            retake = false;
            tcur = told + dt;
            a.dtdx = dt / S.dx;
            a.qold = buf[old];
            int in = cur;
            T sm = T(0);
            for (int st = 0; st < plan.n; st++) {
                int outb = other_buf(in, old);
                a.qin = buf[in];
                a.qout = buf[outb];
                a.op = plan.op[st];
                const bool last = st == plan.n - 1;
                for (int c0 = 1; c0 <= n; c0 += NT - 1) {
                    T s1 = stage1d_chunk<P, 1, CHAR, T, NT>(a, c, c0, sq, sx);
                    if (last) sm = fmax(sm, s1);
                }
                in = outb;
            }
            stages += plan.n;
            // block maximum of |s| of the last stage -> CFL (flux1.f90:126-133)
            double d = (double)wave_max(sm);
            if ((t & 63) == 0) red[t >> 6] = d;
            __syncthreads();
            double m = red[0];
#pragma unroll
            for (int w = 1; w < NW; w++) m = fmax(m, red[w]);
            __syncthreads();
            cfl = (dt / S.dx) * m;
            if (ctl.dt_variable == 1) {
                if (cfl > 0.0) {
                    dt = fmin(ctl.dt_max, dt * ctl.cfl_desired / cfl);
                    dtmin = fmin(dt, dtmin);
                    dtmax = fmax(dt, dtmax);
                } else {
                    dt = ctl.dt_max;
                }
            }
            if (cfl <= ctl.cfl_max) {
                cflmax = fmax(cfl, cflmax);
                cur = in;
            } else {
                tcur = told;
                cur = old;  // q = qold
                rejected++;
                if (ctl.dt_variable == 1) {
                    retake = true;
                } else {
                    cflmax = fmax(cfl, cflmax);
                    break;
                }
            }
            steps++;
            if (tcur >= A.tend) break;
        }
        if (info == 0 && ctl.dt_variable == 1 && tcur < A.tend && steps == A.maxn) info = 11;
        if (info == 0 && ctl.dt_variable == 0 && cflmax > ctl.cfl_max) info = 12;
        if (t == 0) {
            sc_batch_result r;
            r.t = tcur;
            r.cfl_maxused = cflmax;
            r.cfl_last = cfl;
            r.dt_minused = dtmin;
            r.dt_maxused = dtmax;
            r.dt_last = dt;
            r.steps_taken = steps;
            r.rejected = rejected;
            r.info = info;
            r.buffer = cur;
            r.stage_evals = stages;
            r.t_begin = t_begin;
            r.t_end = (long long)wall_clock64();
            A.out[b] = r;
            sb = atomicAdd(A.next, 1);
        }
        __syncthreads();
        k = sb;
        __syncthreads();
    }
}

}  // namespace sc
