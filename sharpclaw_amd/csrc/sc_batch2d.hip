// Batched 2D grids: the tile-parallel stage kernel over many grids and the
// per-grid control kernel of the device-resident time loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "sc_batch2d.h"
#include "sc_internal.h"
#include "sc_stage2d.h"

namespace sc {

__host__ __device__ int stages_of_integrator(int ti)
{
    return ti == 1 ? 2 : ti == 2 ? 4 : ti == 3 ? 3 : ti == 4 ? 10 : 0;
}

// Stage A.stage of the current iteration on every tile of every grid whose
// iteration is open (state.go).  Block b belongs to the last grid whose
// tile0 <= b.  The stage's input, output and q^n buffers follow from the
// grid's cur/old buffers as in a host-driven step: stage 0 reads cur, each
// stage writes the buffer that is neither its input nor q^n.
template <int P, int CHAR, class T, int TX, int TY>
__global__ __launch_bounds__(256) void stage2d_batch_kernel(Batch2DArgs A, PhysConst<T> c)
{
    const int blk = blockIdx.x;
    int lo = 0, hi = A.ngrids - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (A.grids[mid].tile0 <= blk) lo = mid;
        else hi = mid - 1;
    }
    const Grid2D &G = A.grids[lo];
    const SolveState &s = A.state[lo];
    if (!s.go) return;
    const int ntiles = G.ntx * G.nty;
    int local = blk - G.tile0;
    if (local >= ntiles) return;
    // XCD-aware order within the grid (its first tile is a multiple of 8,
    // so block b runs on XCD b mod 8): each XCD gets a contiguous band of
    // the grid's tiles and keeps the halo rows it re-reads in its L2
    if ((ntiles & 7) == 0) local = (local & 7) * (ntiles >> 3) + (local >> 3);
    const Plan plan = stage_plan(A.ti);
    int in = s.cur, out = other_buf(in, s.old);
    for (int k = 0; k < A.stage; k++) {
        in = out;
        out = other_buf(in, s.old);
    }
    const size_t n = (size_t)G.LX * G.LY * (size_t)Shape<P>::M;
    double *buf = A.base + G.off;
    StageArgs a;
    a.qin = buf + in * n;
    a.qold = buf + s.old * n;
    a.qout = buf + out * n;
    a.qbin = buf + 3 * n;
    a.qbout = buf + 3 * n;
    a.smax = A.stage == plan.n - 1 ? A.smax + (size_t)CFL_WORDS * lo : nullptr;
    a.nx = G.nx;
    a.ny = G.ny;
    a.LX = G.LX;
    a.LY = G.LY;
    for (int k = 0; k < 4; k++) a.bc[k] = A.bc[k];
    a.dtdx = s.dt / G.dx;
    a.dtdy = s.dt / G.dy;
    a.op = plan.op[A.stage];
    a.ntx = G.ntx;
    a.nty = G.nty;
    a.capa = nullptr;
    a.src = 0;
    a.dt = s.dt;
    a.xlo = 0.0;
    a.dxc = G.dx;
    a.aux = nullptr;
    a.vscale = 1.0;
    a.full = 0;
    a.order = TILE_BANDS;
    a.tflag = nullptr;
    a.tlist = nullptr;
    a.tnf = nullptr;
    a.dflag = nullptr;
    stage2d_tile<P, CHAR, T, TX, TY>(a, c, local);
}

// One thread per grid.  Post: close the iteration whose stages just ran;
// pre: open the next (loop2d_post / loop2d_pre, sc_batch2d.h) and clear the
// grid's CFL words.
__global__ void ctl2d_kernel(Batch2DArgs A, int do_post, int do_pre)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.ngrids) return;
    const Grid2D &G = A.grids[g];
    SolveState s = A.state[g];
    const int nst = stages_of_integrator(A.ti);
    if (do_post && s.go) {
        double sx = __longlong_as_double((long long)cfl_word(A.smax + (size_t)CFL_WORDS * g, 0));
        double sy = __longlong_as_double((long long)cfl_word(A.smax + (size_t)CFL_WORDS * g, 1));
        loop2d_post(G, s, nst, sx, sy);
    }
    if (do_pre && loop2d_pre(G, s)) {
        for (int k = 0; k < CFL_SUB; k++) {
            A.smax[(size_t)CFL_WORDS * g + k * CFL_STRIDE] = 0ull;
            A.smax[(size_t)CFL_WORDS * g + k * CFL_STRIDE + 1] = 0ull;
        }
    }
    A.state[g] = s;
}

// ------------------------------------------------------------ launchers

template <int P, int CHAR, class T, int TX, int TY>
static hipError_t go2b(const Batch2DArgs &A, int ntiles, const PhysConst<T> &c, hipStream_t st)
{
    hipLaunchKernelGGL((stage2d_batch_kernel<P, CHAR, T, TX, TY>), dim3(ntiles), dim3(256), 0, st, A, c);
    return hipGetLastError();
}

template <int P, int CHAR, class T>
static hipError_t go2b_t(int tile, const Batch2DArgs &A, int ntiles, const PhysConst<T> &c, hipStream_t st)
{
    switch (tile) {
    case 1: return go2b<P, CHAR, T, 15, 16>(A, ntiles, c, st);
    case 2: return go2b<P, CHAR, T, 32, 7>(A, ntiles, c, st);
    case 3: return go2b<P, CHAR, T, 14, 17>(A, ntiles, c, st);
    default: return go2b<P, CHAR, T, 16, 15>(A, ntiles, c, st);
    }
}

template <int P>
static hipError_t go2b_p(const PhysSel &s, int tile, const Batch2DArgs &A, int ntiles, const PhysConst<double> &cd,
                         const PhysConst<float> &cf, hipStream_t st)
{
    if (s.chr == 2)
        return s.fp32 ? go2b_t<P, 2, float>(tile, A, ntiles, cf, st) : go2b_t<P, 2, double>(tile, A, ntiles, cd, st);
    return s.fp32 ? go2b_t<P, 0, float>(tile, A, ntiles, cf, st) : go2b_t<P, 0, double>(tile, A, ntiles, cd, st);
}

hipError_t launch_stage2d_batch(const PhysSel &s, int tile, const Batch2DArgs &A, int ntiles,
                                const PhysConst<double> &cd, const PhysConst<float> &cf, hipStream_t st)
{
    if (ntiles < 1) return hipSuccess;
    switch (s.phys) {
    case EU2: return go2b_p<EU2>(s, tile, A, ntiles, cd, cf, st);
    case SW2: return go2b_p<SW2>(s, tile, A, ntiles, cd, cf, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_ctl2d(const Batch2DArgs &A, int do_post, int do_pre, hipStream_t st)
{
    int nb = (A.ngrids + 63) / 64;
    if (nb < 1) return hipSuccess;
    hipLaunchKernelGGL(ctl2d_kernel, dim3(nb), dim3(64), 0, st, A, do_post, do_pre);
    return hipGetLastError();
}

}  // namespace sc

// ------------------------------------------------------------ host driver

namespace sc {

int round_iterations(int stages, long long tiles)
{
    long long k = 131072 / std::max(1LL, (long long)stages * tiles);
    return (int)std::min(64LL, std::max(1LL, k));
}

int Batch2DRunner::init(const sc_params &p, hipStream_t stream)
{
    sel.phys = p.physics == SC_PHYS_EULER ? EU2 : SW2;
    sel.chr = p.char_decomp;
    sel.fp32 = p.precision == SC_FP32;
    cd = sc_make_const<double>(&p);
    cf = sc_make_const<float>(&p);
    ti = p.time_integrator;
    for (int k = 0; k < 4; k++) bc[k] = p.mthbc[k];
    owns_st = stream == nullptr;
    st = stream;
    if ((owns_st && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) ||
        hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
        return sc_fail(SC_ERR_HIP, "2D batch: stream/event creation failed");
    return SC_OK;
}

void Batch2DRunner::set_tile(int t)
{
    if (t == tile) return;
    if (g_exec) (void)hipGraphExecDestroy(g_exec);
    g_exec = nullptr;
    tile = t;
}

int Batch2DRunner::reserve(int n)
{
    if (n <= cap) return SC_OK;
    int c = std::max(n, 2 * cap);
    if (st) (void)hipStreamSynchronize(st);
    if (g_exec) (void)hipGraphExecDestroy(g_exec);
    g_exec = nullptr;
    if (d_grids) (void)hipFree(d_grids);
    if (d_state) (void)hipFree(d_state);
    if (d_smax) (void)hipFree(d_smax);
    if (h_grids) (void)hipHostFree(h_grids);
    if (h_state) (void)hipHostFree(h_state);
    d_grids = nullptr;
    d_state = nullptr;
    d_smax = nullptr;
    h_grids = nullptr;
    h_state = nullptr;
    cap = 0;
    if (hipMalloc(&d_grids, c * sizeof(Grid2D)) != hipSuccess ||
        hipMalloc(&d_state, c * sizeof(SolveState)) != hipSuccess ||
        hipMalloc(&d_smax, (size_t)CFL_WORDS * c * sizeof(unsigned long long)) != hipSuccess ||
        hipHostMalloc(&h_grids, c * sizeof(Grid2D), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&h_state, c * sizeof(SolveState), hipHostMallocDefault) != hipSuccess)
        return sc_fail(SC_ERR_NOMEM, "2D batch: descriptor allocation failed");
    cap = c;
    return SC_OK;
}

// First tiles at multiples of 8 (the blocks in between exit at once).
long long Batch2DRunner::assign_tiles(int n)
{
    long long tiles = 0;
    for (int i = 0; i < n; i++) {
        tiles = (tiles + 7) & ~7LL;
        h_grids[i].tile0 = (int)tiles;
        tiles += (long long)h_grids[i].ntx * h_grids[i].nty;
    }
    return tiles;
}

// ctl(pre), then K x [stages, ctl(post, pre unless last)], then the states
// back to the pinned staging.
hipError_t Batch2DRunner::enqueue_steps(int n, int K, long long tiles, double *base)
{
    Batch2DArgs A;
    memset(&A, 0, sizeof A);
    A.grids = d_grids;
    A.state = d_state;
    A.smax = d_smax;
    A.ngrids = n;
    A.base = base;
    for (int k = 0; k < 4; k++) A.bc[k] = bc[k];
    A.ti = ti;
    const int nst = stages_of_integrator(ti);
    hipError_t e = launch_ctl2d(A, 0, 1, st);
    for (int it = 0; it < K && e == hipSuccess; it++) {
        for (int k = 0; k < nst && e == hipSuccess; k++) {
            A.stage = k;
            e = launch_stage2d_batch(sel, tile, A, (int)tiles, cd, cf, st);
        }
        if (e == hipSuccess) e = launch_ctl2d(A, 1, it + 1 < K, st);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h_state, d_state, n * sizeof(SolveState), hipMemcpyDeviceToHost, st);
    return e;
}

int Batch2DRunner::launch_round(int n, int K, double *base)
{
    int rc = reserve(n);
    if (rc) return rc;
    long long tiles = assign_tiles(n);
    if (tiles > 0x7fffffffLL) return sc_fail(SC_ERR_ARG, "2D batch: too many tiles");
    hipError_t e = hipMemcpyAsync(d_grids, h_grids, n * sizeof(Grid2D), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_state, h_state, n * sizeof(SolveState), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = enqueue_steps(n, K, tiles, base);
    if (e == hipSuccess) e = hipEventRecord(ev, st);
    if (e != hipSuccess) return sc_fail(SC_ERR_HIP, "2D batch round: %s", hipGetErrorString(e));
    return SC_OK;
}

int Batch2DRunner::continue_round(int n, int K, double *base, bool graph)
{
    long long tiles = 0;
    for (int i = 0; i < n; i++) tiles = ((tiles + 7) & ~7LL) + (long long)h_grids[i].ntx * h_grids[i].nty;
    hipError_t e = hipSuccess;
    if (!graph) {
        e = enqueue_steps(n, K, tiles, base);
    } else {
        if (g_exec && (g_n != n || g_K != K || g_tiles != tiles || g_base != base)) {
            (void)hipGraphExecDestroy(g_exec);
            g_exec = nullptr;
        }
        if (!g_exec) {
            hipGraph_t gr = nullptr;
            e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                hipError_t e2 = enqueue_steps(n, K, tiles, base);
                e = hipStreamEndCapture(st, &gr);
                if (e == hipSuccess) e = e2;
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&g_exec, gr, nullptr, nullptr, 0);
            if (gr) (void)hipGraphDestroy(gr);
            if (e != hipSuccess) {
                g_exec = nullptr;
                return sc_fail(SC_ERR_HIP, "2D round capture: %s", hipGetErrorString(e));
            }
            g_n = n;
            g_K = K;
            g_tiles = tiles;
            g_base = base;
        }
        e = hipGraphLaunch(g_exec, st);
    }
    if (e == hipSuccess) e = hipEventRecord(ev, st);
    if (e != hipSuccess) return sc_fail(SC_ERR_HIP, "2D batch round: %s", hipGetErrorString(e));
    return SC_OK;
}

int Batch2DRunner::wait()
{
    hipError_t e = hipEventSynchronize(ev);
    if (e != hipSuccess) return sc_fail(SC_ERR_HIP, "2D batch round: %s", hipGetErrorString(e));
    return SC_OK;
}

void Batch2DRunner::destroy()
{
    if (st) (void)hipStreamSynchronize(st);
    if (g_exec) (void)hipGraphExecDestroy(g_exec);
    g_exec = nullptr;
    if (d_grids) (void)hipFree(d_grids);
    if (d_state) (void)hipFree(d_state);
    if (d_smax) (void)hipFree(d_smax);
    if (h_grids) (void)hipHostFree(h_grids);
    if (h_state) (void)hipHostFree(h_state);
    if (ev) (void)hipEventDestroy(ev);
    if (st && owns_st) (void)hipStreamDestroy(st);
    d_grids = nullptr;
    d_state = nullptr;
    d_smax = nullptr;
    h_grids = nullptr;
    h_state = nullptr;
    ev = nullptr;
    st = nullptr;
    cap = 0;
}

}  // namespace sc
