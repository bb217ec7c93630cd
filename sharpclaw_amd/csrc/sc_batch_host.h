// Host-side helpers of the persistent batch kernel, shared by the batch
// API (sc_batch.hip) and the request server (sc_serve.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/sharpclaw_amd.h"
#include "sc_batch_kernel.h"

namespace sc {

// Workgroup sizes the batch kernel is built for.
constexpr int NBATCH_BLOCKS = 5;
extern const int kBatchBlocks[NBATCH_BLOCKS];

// Reconstruction code (RK_*) of the batch kernel for these options
// (flux1.f90:65-117): lim_type 2 with char_decomp 0-3, or lim_type 1 with
// char_decomp 1 (tvd2_wave) or 2 (tvd2_char); -1 for any other.
inline int batch_rk(const PhysSel &s)
{
    if (s.lim == 2) return s.chr >= 0 && s.chr <= 3 ? s.chr : -1;
    if (s.lim == 1) return s.chr == 1 ? RK_TVD_WAVE : s.chr == 2 ? RK_TVD_CHAR : -1;
    return -1;
}

// Largest workgroup the batch kernel is built for with these options: the
// reconstructions with neighbour data keep it in LDS per interface, and are
// built up to 256 threads.
inline int batch_max_block(const PhysSel &s)
{
    const int rk = batch_rk(s);
    return rk == RK_WENO || rk == RK_WENO_CHAR ? 1024 : 256;
}

// Launch the persistent kernel over A.order[0..A.nsolves) with nt threads
// per workgroup and at most max_wg workgroups (0: fill the device).
hipError_t launch_batch(const PhysSel &s, int nt, const BatchArgs &A, const PhysConst<double> &cd,
                        const PhysConst<float> &cf, int ncu, int max_wg, hipStream_t st);

// Resident workgroups per CU of the kernel for these options and size.
int batch_occupancy(const PhysSel &s, int nt);

// Iteration limit and pre-loop option errors of the time loop for one
// interval (sharpclaw.f90: fixed dt takes (tend - tstart)/dt steps, info 2
// when that is not whole, info 3 when cfl_desired > cfl_max).
inline void loop_limits(const sc_run_ctl &ctl, double tstart, double tend, double dt0, int *maxn, int *info0)
{
    *maxn = ctl.max_steps;
    *info0 = 0;
    if (ctl.dt_variable == 0) {
        if (tend < tstart) {
            *maxn = 1;
        } else {
            *maxn = (int)((tend - tstart + 1e-10) / dt0);
            if (std::fabs(*maxn * dt0 - (tend - tstart)) > 1e-5 * (tend - tstart)) *info0 = 2;
        }
    }
    if (ctl.dt_variable == 1 && ctl.cfl_desired > ctl.cfl_max) *info0 = 3;
}

// Fresh loop state of a solve whose q is in buffer cur.
inline SolveState initial_state(double tstart, double dt0, int cur, int info0)
{
    SolveState s;
    memset(&s, 0, sizeof s);
    s.t = s.told = tstart;
    s.dt = s.dtmin = s.dtmax = dt0;
    s.cur = s.old = cur;
    s.info = info0;
    return s;
}

// The public record of a finished (or paused) solve.
inline void state_to_result(const SolveState &s, double tick_ms, long long t0, sc_batch_result *r)
{
    r->t = s.t;
    r->cfl_maxused = s.cflmax;
    r->cfl_last = s.cfl;
    r->dt_minused = s.dtmin;
    r->dt_maxused = s.dtmax;
    r->dt_last = s.dt;
    r->steps_taken = s.steps;
    r->rejected = s.rejected;
    r->info = s.info;
    r->buffer = s.cur;
    r->stage_evals = s.stages;
    r->t_begin = s.t_begin;
    r->t_end = s.t_end;
    r->done_ms = s.done ? (double)(s.t_end - t0) * tick_ms : -1.0;
}

// Ghost cells of a 1D grid filled on the host as bc1.f90:40-107 does.
inline void fill_ghosts_1d(double *q, int n, int meqn, int bl, int br)
{
    const int L = n + 6;
    for (int m = 0; m < meqn; m++)
        for (int g = 1; g <= 3; g++) {
            int jl = 1 - g, jr = n + g;
            bool negl, negr;
            int sl = bc_src(jl, n, bl, br, negl);
            int sr = bc_src(jr, n, bl, br, negr);
            double vl = q[(sl + 2) + (size_t)m * L], vr = q[(sr + 2) + (size_t)m * L];
            q[(jl + 2) + (size_t)m * L] = (m == 1 && negl) ? -vl : vl;
            q[(jr + 2) + (size_t)m * L] = (m == 1 && negr) ? -vr : vr;
        }
}

}  // namespace sc
