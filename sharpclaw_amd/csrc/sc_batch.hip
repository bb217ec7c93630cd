// Batched solves: many independent 1D grids of uneven size in one
// persistent launch.
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
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/sharpclaw_amd.h"
#include "sc_batch_kernel.h"
#include "sc_internal.h"

namespace sc {

// ------------------------------------------------------------ launchers

template <int P, int CHAR, class T, int NT>
static hipError_t gob(const BatchArgs &A, const PhysConst<T> &c, int nwg, hipStream_t st)
{
    hipLaunchKernelGGL((batch_kernel<P, CHAR, T, NT>), dim3(nwg), dim3(NT), 0, st, A, c);
    return hipGetLastError();
}

template <int P, int CHAR, class T, int NT>
static int occ(void)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<P, CHAR, T, NT>, NT, 0) !=
        hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

template <int P, int CHAR, class T>
static hipError_t gob_nt(int nt, const BatchArgs &A, const PhysConst<T> &c, int ncu, hipStream_t st)
{
    switch (nt) {
    case 64: return gob<P, CHAR, T, 64>(A, c, std::min(A.nsolves, ncu * occ<P, CHAR, T, 64>()), st);
    case 128: return gob<P, CHAR, T, 128>(A, c, std::min(A.nsolves, ncu * occ<P, CHAR, T, 128>()), st);
    default: return gob<P, CHAR, T, 256>(A, c, std::min(A.nsolves, ncu * occ<P, CHAR, T, 256>()), st);
    }
}

template <int P>
static hipError_t gob_p(const PhysSel &s, int nt, const BatchArgs &A, const PhysConst<double> &cd,
                        const PhysConst<float> &cf, int ncu, hipStream_t st)
{
    if (s.chr == 2)
        return s.fp32 ? gob_nt<P, 2, float>(nt, A, cf, ncu, st) : gob_nt<P, 2, double>(nt, A, cd, ncu, st);
    return s.fp32 ? gob_nt<P, 0, float>(nt, A, cf, ncu, st) : gob_nt<P, 0, double>(nt, A, cd, ncu, st);
}

static hipError_t launch_batch(const PhysSel &s, int nt, const BatchArgs &A, const PhysConst<double> &cd,
                               const PhysConst<float> &cf, int ncu, hipStream_t st)
{
    switch (s.phys) {
    case ADV1: return gob_p<ADV1>(s, nt, A, cd, cf, ncu, st);
    case AC1: return gob_p<AC1>(s, nt, A, cd, cf, ncu, st);
    case EU1: return gob_p<EU1>(s, nt, A, cd, cf, ncu, st);
    case SW1: return gob_p<SW1>(s, nt, A, cd, cf, ncu, st);
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
    hipStream_t st;
    std::vector<BatchSolve> solves;
    std::vector<int> order;
    std::vector<int> cur;     // buffer holding each solve's current q
    BatchSolve *d_solves;
    int *d_order, *d_next, *d_start;
    sc_batch_result *d_out;
    double *d_base;
    size_t total;             // doubles allocated
    double tick_ms;           // wall_clock64 period in ms
    hipEvent_t ev0, ev1;
};

static void batch_free(sc_batch *b)
{
    if (!b) return;
    if (b->st) (void)hipStreamSynchronize(b->st);
    if (b->d_solves) (void)hipFree(b->d_solves);
    if (b->d_order) (void)hipFree(b->d_order);
    if (b->d_next) (void)hipFree(b->d_next);
    if (b->d_start) (void)hipFree(b->d_start);
    if (b->d_out) (void)hipFree(b->d_out);
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
    b->sel.fp32 = params->precision == SC_FP32;
    b->cd = sc_make_const<double>(params);
    b->cf = sc_make_const<float>(params);
    b->dev = params->device;
    b->nsolves = nsolves;
    b->block = 0;  // tuned at the first run
    hipDeviceProp_t prop;
    (void)hipGetDeviceProperties(&prop, b->dev);
    b->ncu = prop.multiProcessorCount;
    int khz = 0;
    (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->dev);
    b->tick_ms = khz > 0 ? 1.0 / khz : 1e-5;
    b->solves.resize(nsolves);
    long long off = 0;
    for (int i = 0; i < nsolves; i++) {
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
    bool ok = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(&b->d_solves, nsolves * sizeof(BatchSolve)) == hipSuccess;
    ok = ok && hipMalloc(&b->d_order, nsolves * sizeof(int)) == hipSuccess;
    ok = ok && hipMalloc(&b->d_next, sizeof(int)) == hipSuccess;
    ok = ok && hipMalloc(&b->d_start, nsolves * sizeof(int)) == hipSuccess;
    ok = ok && hipMalloc(&b->d_out, nsolves * sizeof(sc_batch_result)) == hipSuccess;
    ok = ok && hipMalloc(&b->d_base, b->total * sizeof(double)) == hipSuccess;
    ok = ok && hipEventCreate(&b->ev0) == hipSuccess && hipEventCreate(&b->ev1) == hipSuccess;
    ok = ok && hipMemcpy(b->d_solves, b->solves.data(), nsolves * sizeof(BatchSolve), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(b->d_order, b->order.data(), nsolves * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
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
    // the device never stores ghost cells (stages synthesise them); fill
    // them as bc1.f90:40-107 would
    const int n = b->solves[i].mx, L = n + 6;
    for (int m = 0; m < b->p.meqn; m++)
        for (int g = 1; g <= 3; g++) {
            int jl = 1 - g, jr = n + g;
            bool negl, negr;
            int sl = bc_src(jl, n, b->p.mthbc[0], b->p.mthbc[1], negl);
            int sr = bc_src(jr, n, b->p.mthbc[0], b->p.mthbc[1], negr);
            double vl = q[(sl + 2) + (size_t)m * L], vr = q[(sr + 2) + (size_t)m * L];
            q[(jl + 2) + (size_t)m * L] = (m == 1 && negl) ? -vl : vl;
            q[(jr + 2) + (size_t)m * L] = (m == 1 && negr) ? -vr : vr;
        }
    return SC_OK;
}

static int run_once(sc_batch *b, int nt, const sc_run_ctl *ctl, double tstart, double tend, double dt0,
                    int maxn_override, float *ms)
{
    BatchArgs A;
    memset(&A, 0, sizeof A);
    A.solves = b->d_solves;
    A.out = b->d_out;
    A.order = b->d_order;
    A.start = b->d_start;
    A.nsolves = b->nsolves;
    A.next = b->d_next;
    A.base = b->d_base;
    A.meqn = b->p.meqn;
    A.bc[0] = b->p.mthbc[0];
    A.bc[1] = b->p.mthbc[1];
    A.ti = b->p.time_integrator;
    A.ctl = *ctl;
    // step limit and option checks of the time loop (as sc_advance)
    A.maxn = ctl->max_steps;
    if (ctl->dt_variable == 0) {
        if (tend < tstart) {
            A.maxn = 1;
        } else {
            A.maxn = (int)((tend - tstart + 1e-10) / dt0);
            if (std::fabs(A.maxn * dt0 - (tend - tstart)) > 1e-5 * (tend - tstart)) A.info0 = 2;
        }
    }
    if (ctl->dt_variable == 1 && ctl->cfl_desired > ctl->cfl_max) A.info0 = 3;
    if (maxn_override > 0) {
        A.maxn = maxn_override;
        A.info0 = 0;
    }
    A.tstart = tstart;
    A.tend = tend;
    A.dt0 = dt0;
    hipError_t e = hipMemsetAsync(b->d_next, 0, sizeof(int), b->st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_start, b->cur.data(), b->nsolves * sizeof(int), hipMemcpyHostToDevice, b->st);
    if (e == hipSuccess) e = hipEventRecord(b->ev0, b->st);
    if (e == hipSuccess) e = launch_batch(b->sel, nt, A, b->cd, b->cf, b->ncu, b->st);
    if (e == hipSuccess) e = hipEventRecord(b->ev1, b->st);
    if (e == hipSuccess) e = hipEventSynchronize(b->ev1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, b->ev0, b->ev1);
    if (e != hipSuccess) return sc_fail(SC_ERR_HIP, "batch run: %s", hipGetErrorString(e));
    return SC_OK;
}

// Autotuner of the workgroup size: each candidate (64, 128, 256 threads,
// i.e. 63, 127, 255 cells per chunk) runs one time step of every solve from
// the current states -- a step writes only the two buffers other than the
// one holding q (and the SSP(10,4) register), and with rejection disabled
// no step is retaken, so the states are untouched afterwards -- and the
// fastest is kept.
static int tune_block(sc_batch *b, const sc_run_ctl *ctl, double tstart, double tend, double dt0)
{
    if (b->block) return SC_OK;
    if (const char *e = getenv("SC_BATCH_BLOCK")) {
        int v = atoi(e);
        if (v == 64 || v == 128 || v == 256) {
            b->block = v;
            return SC_OK;
        }
    }
    sc_run_ctl one = *ctl;
    one.cfl_max = 1e300;  // no reject, so no retake from qold
    one.cfl_desired = 1.0;
    int best = 256;
    float bestms = 1e30f;
    for (int nt : {64, 128, 256}) {
        float ms = 0.f;
        int rc = run_once(b, nt, &one, tstart, tend, dt0, 1, &ms);
        if (rc) return rc;
        if (ms < bestms) {
            bestms = ms;
            best = nt;
        }
    }
    b->block = best;
    return SC_OK;
}

extern "C" int sc_batch_run(sc_batch *b, const sc_run_ctl *ctl, double tstart, double tend,
                            double dt_initial, sc_batch_result *results, double *gpu_ms)
{
    if (!b || !ctl || !results) return sc_fail(SC_ERR_ARG, "bad argument");
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(b->dev);
    int rc = tune_block(b, ctl, tstart, tend, dt_initial);
    float ms = 0.f;
    if (!rc) rc = run_once(b, b->block, ctl, tstart, tend, dt_initial, 0, &ms);
    if (!rc) {
        hipError_t e = hipMemcpy(results, b->d_out, b->nsolves * sizeof(sc_batch_result), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = sc_fail(SC_ERR_HIP, "batch results: %s", hipGetErrorString(e));
    }
    (void)hipSetDevice(prev);
    if (rc) return rc;
    long long t0 = results[0].t_begin;
    for (int i = 0; i < b->nsolves; i++) t0 = std::min(t0, results[i].t_begin);
    for (int i = 0; i < b->nsolves; i++) {
        b->cur[i] = results[i].buffer;
        results[i].done_ms = (double)(results[i].t_end - t0) * b->tick_ms;
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
