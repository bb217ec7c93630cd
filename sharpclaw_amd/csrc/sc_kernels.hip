// gfx950 kernels of the semi-discrete step.
//
// Fused stage kernels: one pass reads the stage input q (plus q^n and the
// SSP(10,4) register where the stage needs them), synthesises ghost cells
// for the boundary conditions in LDS, reconstructs both sides of every
// interface (WENO5, component-wise or characteristic), solves the interface
// Riemann problems and the in-cell problems in registers, forms
// dq = dt*L(q), combines it into the next RK stage and writes that stage.
// The maximum wave speed is reduced across each wavefront with shuffles,
// across the block in LDS, and merged into a device word with one atomic
// per block.
#include <algorithm>

#include <hip/hip_ext.h>

#include "sc_batch2d.h"
#include "sc_kernels.h"
#include "sc_stage.h"
#include "sc_stage2d.h"

namespace sc {

// ------------------------------------------------------ 1D fused stage

// One block of NT threads solves NT interfaces i = c0..c0+NT-1 and updates
// the NT-1 cells c0..c0+NT-2 (stage1d_chunk).  CHAR: reconstruction code
// (RK_*, sc_stage.h).
template <int P, int IXY, int CHAR, class T, int NT>
__global__ __launch_bounds__(NT) void stage1d_kernel(StageArgs a, PhysConst<T> c)
{
    constexpr int M = Shape<P>::M;
    constexpr int NB = nbr_words<P, IXY, CHAR, T>();
    constexpr bool AL = sx_in_sw<P, IXY, CHAR, T, NT>();
    __shared__ T sq[M][NT + 5];
    __shared__ __align__(8) T sx[AL ? 1 : 2 * M][NT];
    __shared__ T sw[NB > 0 ? NB : 1][NT + 4];
    __shared__ double red[NT / 64];
    T(*sxp)[NT] = AL ? reinterpret_cast<T(*)[NT]>(&sw[0][0]) : sx;
    T smax = stage1d_chunk<P, IXY, CHAR, T, NT>(a, c, 1 + blockIdx.x * (NT - 1), sq, sxp, &sw[0][0]);
    if (a.smax) merge_smax<T, NT / 64>(a.smax, smax, red);
}

// ------------------------------------------------- 2D stage by slices

// x sweeps (IXY = 1) over rows 0..ny+1 or y sweeps (IXY = 2) over columns
// 0..nx+1 of a 2D stage (sweep2d_chunk): one block per (chunk, group of NLN
// slices), NT interfaces per slice.  The x sweep runs one row per block
// (its loads are contiguous along the row); the y sweep runs NLN = 4
// neighbouring columns per block, so each staged row of the chunk is a
// contiguous 32-byte load and each update a contiguous store.  Workgroups
// are dealt to the 8 XCDs round-robin; the block index is remapped so that
// each XCD takes a contiguous band of (chunk, group) pairs, group fastest:
// neighbouring groups, which share cache lines, meet in one XCD's L2.
template <int P, int IXY, int CHAR, class T, int NT, int NLN>
__global__ __launch_bounds__(NT * NLN) void sweep2d_kernel(StageArgs a, PhysConst<T> c, double *dqx)
{
    constexpr int M = Shape<P>::M;
    constexpr int NB = nbr_words<P, IXY, CHAR, T>();
    constexpr bool AL = sx_in_sw<P, IXY, CHAR, T, NT>();
    __shared__ T sq[M * (NT + 5) * NLN];
    __shared__ T sx[AL ? 1 : 2 * M * NT * NLN];
    __shared__ T sw[NB > 0 ? NB * (NT + 4) * NLN : 1];
    __shared__ double red[NT * NLN / 64];
    const int nb = gridDim.x, q8 = nb >> 3, r8 = nb & 7;
    const int xcd = blockIdx.x & 7, k = blockIdx.x >> 3;
    const int b = xcd < r8 ? xcd * (q8 + 1) + k : r8 * (q8 + 1) + (xcd - r8) * q8 + k;
    const int ngrp = ((IXY == 1 ? a.ny : a.nx) + 2 + NLN - 1) / NLN;
    const int c0 = 1 + (b / ngrp) * (NT - 1), sl0 = (b % ngrp) * NLN;
    T smax = sweep2d_chunk<P, IXY, CHAR, T, NT, NLN>(a, c, dqx, c0, sl0, sq, AL ? sw : sx, sw);
    if (a.smax) merge_smax<T, NT * NLN / 64>(a.smax + (IXY - 1), smax, red);
}

// ------------------------------------------------------ 2D fused stage

// The tile a workgroup takes.  Workgroups are dealt to the 8 XCDs
// round-robin (workgroup w runs on XCD w mod 8, its (w div 8)-th there) and
// each XCD has its own L2.  TILE_BANDS: each XCD a contiguous band of tiles
// (tile rows), so halo rows shared by neighbouring tiles stay in one L2.
// TILE_ADAPTIVE: the list tile_partition builds after each step holds the
// tiles that were not constant at the last step first, then the constant
// ones, each part in strip order (nf = the count of the former, *a.tnf).
// Each XCD takes an equal share of the non-constant part, a contiguous run of
// it, with its first workgroups (a non-constant tile takes several times as
// long as a constant one, so these start first and spread evenly over the
// XCDs), then a contiguous run of the constant part: within an XCD
// consecutive workgroups take neighbouring tiles, whose shared halo rows stay
// in its L2.  Piecewise-constant data such as a four-quadrant Riemann problem
// has its non-constant tiles in a few rows of tiles, which TILE_BANDS would
// give to one or two XCDs, and in tile order they would start late and set
// the launch's tail.
__device__ __forceinline__ int tile_of_workgroup(const StageArgs &a, int w, int nf)
{
    const int ntiles = a.ntx * a.nty;
    if (w >= ntiles) return -1;
    if (a.order == TILE_ADAPTIVE) {
        const int x = w & 7, k = w >> 3;
        // XCD y runs workgroups y, y+8, ...: ceil((ntiles - y)/8) of them, the
        // first ceil((nf - y)/8) on non-constant tiles
        const int nfq = nf >> 3, nfr = nf & 7;
        const int nfx = nfq + (x < nfr ? 1 : 0);
        if (k < nfx) return a.tlist[x * nfq + min(x, nfr) + k];
        int start = 0;
        for (int y = 0; y < x; y++) start += ((ntiles - y + 7) >> 3) - (nfq + (y < nfr ? 1 : 0));
        return a.tlist[nf + start + (k - nfx)];
    }
    if ((ntiles & 7) == 0) return (w & 7) * (ntiles >> 3) + (w >> 3);
    return w;
}

// The CFL set's two words (the max over their CFL_SUB copies) as doubles
// into out[0..1] (pinned host memory: the host reads them after the stream
// has passed the kernel).  One wave: lane k reads copy k mod 32 of word k / 32.
__device__ __forceinline__ void cfl_set_to(const unsigned long long *set, double *out)
{
    const int k = threadIdx.x;
    long long v = (long long)set[(k & 31) * CFL_STRIDE + (k >> 5)];
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) {
        const long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    if ((k & 31) == 0) out[k >> 5] = __longlong_as_double(v);
}

__global__ __launch_bounds__(64) void cfl_to_host_kernel(const unsigned long long *set, double *out)
{
    cfl_set_to(set, out);
}

// The tile at position p of the strip order: strips of STRIP tile columns
// (the last one narrower), each in row order, left to right; grids of at
// most STRIP tiles per row are in row order.  In the tile list the constant
// tiles follow this order, so that an XCD's contiguous run of them goes down
// a strip: the tile above is at most a strip width later, and the halo rows
// it shares with the one below are still in the XCD's L2 (64 tiles of 16 x
// 15 cells stage ~1.8 MB; at 4096^2 a row of tiles in between overflows the
// 4 MB L2).
constexpr int STRIP = 64;

__device__ __forceinline__ int strip_tile(int p, int ntx, int nty)
{
    const int full = ntx / STRIP, per = STRIP * nty;
    int s = p / per, w = STRIP;
    if (s >= full) {
        s = full;
        w = ntx - full * STRIP;
    }
    const int r = p - s * per;
    return (r / w) * ntx + s * STRIP + r % w;
}

// The tiles of consecutive positions of the strip order, from position p
// on, without a division per step
struct StripWalk {
    int ntx, nty, x0, w, ty, tx;  // strip start and width, current row and column in the strip
    __device__ StripWalk(int p, int ntx_, int nty_) : ntx(ntx_), nty(nty_)
    {
        const int full = ntx / STRIP, per = STRIP * nty;
        int s = p / per;
        w = STRIP;
        if (s >= full) {
            s = full;
            w = ntx - full * STRIP;
        }
        x0 = s * STRIP;
        const int r = p - s * per;
        ty = w > 0 ? r / w : 0;
        tx = w > 0 ? r % w : 0;
    }
    __device__ int tile() const { return ty * ntx + x0 + tx; }
    __device__ void next()
    {
        if (++tx < w) return;
        tx = 0;
        if (++ty < nty) return;
        ty = 0;
        x0 += w;
        w = min(STRIP, ntx - x0);
    }
};

// Stable partition of the tiles by their flags in strip order (one block of
// 1024 threads): each thread counts the flags of a run of positions and a
// two-level scan (wave prefix sums, then the 16 wave totals) places the
// runs.  With set != NULL the first wave also hands the step's CFL set to the
// host (cfl_set_to).
// The partition itself, by a block of 1024 threads.
__device__ __forceinline__ int tile_partition_block(const unsigned char *flag, int ntx, int nty, int *list, int *nf)
{
    __shared__ int wsum[16];
    const int n = ntx * nty;
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const int per = (n + 1023) / 1024;
    const int lo = min(n, t * per), hi = min(n, lo + per);
    int cnt = 0;
    {
        StripWalk sw(lo, ntx, nty);
        for (int i = lo; i < hi; i++, sw.next()) cnt += flag[sw.tile()] != 0;
    }
    int v = cnt;  // inclusive prefix sum over the wave
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(v, off, 64);
        if (lane >= off) v += o;
    }
    if (lane == 63) wsum[wid] = v;
    __syncthreads();
    if (wid == 0) {
        int s = lane < 16 ? wsum[lane] : 0;
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) {
            const int o = __shfl_up(s, off, 64);
            if (lane >= off) s += o;
        }
        if (lane < 16) wsum[lane] = s;
    }
    __syncthreads();
    const int total = wsum[15];
    int pf = v - cnt + (wid > 0 ? wsum[wid - 1] : 0);  // flagged tiles before this run
    int pc = total + (lo - pf);                          // unflagged ones before it, after all flagged
    StripWalk sw(lo, ntx, nty);
    for (int i = lo; i < hi; i++, sw.next()) {
        const int tile = sw.tile();
        if (flag[tile])
            list[pf++] = tile;
        else
            list[pc++] = tile;
    }
    if (t == 0) *nf = total;
    return total;
}

__global__ __launch_bounds__(1024) void tile_partition_kernel(const unsigned char *flag, int ntx, int nty, int *list,
                                                              int *nf, const unsigned long long *set, double *out)
{
    if (set && threadIdx.x < 64) cfl_set_to(set, out);
    tile_partition_block(flag, ntx, nty, list, nf);
}

// The same partition for grids of many tiles, one position per thread over
// many blocks: tile_count_kernel writes each block's count of flagged
// positions, tile_scatter_kernel places every block's positions after the
// counts of the blocks before it.
__global__ __launch_bounds__(1024) void tile_count_kernel(const unsigned char *flag, int ntx, int nty, int *bcount)
{
    __shared__ int wsum[16];
    const int n = ntx * nty;
    const int p = blockIdx.x * 1024 + threadIdx.x, lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const bool f = p < n && flag[strip_tile(p, ntx, nty)] != 0;
    const unsigned long long b = __ballot(f);
    if (lane == 0) wsum[wid] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int k = 0; k < 16; k++) s += wsum[k];
        bcount[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(1024) void tile_scatter_kernel(const unsigned char *flag, int ntx, int nty,
                                                            const int *bcount, int nblk, int *list, int *nf,
                                                            const unsigned long long *set, double *out)
{
    __shared__ int wsum[16], red[2];
    const int n = ntx * nty;
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    if (set && blockIdx.x == 0 && wid == 0) cfl_set_to(set, out);
    if (t == 0) {
        int before = 0, total = 0;
        for (int k = 0; k < nblk; k++) {
            const int c = bcount[k];
            before += k < (int)blockIdx.x ? c : 0;
            total += c;
        }
        red[0] = before;
        red[1] = total;
        if (blockIdx.x == 0) *nf = total;
    }
    const int p = blockIdx.x * 1024 + t;
    const int tile = p < n ? strip_tile(p, ntx, nty) : 0;
    const bool f = p < n && flag[tile] != 0;
    const unsigned long long b = __ballot(f);
    if (lane == 0) wsum[wid] = __popcll(b);
    __syncthreads();
    int fw = 0;  // flagged positions of this block in the waves before this one
    for (int k = 0; k < wid; k++) fw += wsum[k];
    const int fl = fw + __popcll(b & ((1ull << lane) - 1ull));  // flagged before p in the block
    const int pf = red[0] + fl;
    const int pc = red[1] + (p - pf);  // unflagged positions before p, after all flagged
    if (p < n) list[f ? pf : pc] = tile;
}

__global__ void tile_list_reset_kernel(unsigned char *flag, int ntx, int nty, int *list, int *nf)
{
    const int n = ntx * nty;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        flag[i] = 0;
        list[i] = strip_tile(i, ntx, nty);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *nf = 0;
}

hipError_t launch_tile_partition(const unsigned char *flag, int ntx, int nty, int *list, int *nf,
                                 const unsigned long long *set, double *out, int *bcount, hipStream_t st)
{
    const int n = ntx * nty;
    if (n <= 16384) {
        hipLaunchKernelGGL(tile_partition_kernel, dim3(1), dim3(1024), 0, st, flag, ntx, nty, list, nf, set, out);
        return hipGetLastError();
    }
    const int nblk = (n + 1023) / 1024;
    hipLaunchKernelGGL(tile_count_kernel, dim3(nblk), dim3(1024), 0, st, flag, ntx, nty, bcount);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(tile_scatter_kernel, dim3(nblk), dim3(1024), 0, st, flag, ntx, nty, (const int *)bcount, nblk,
                       list, nf, set, out);
    return hipGetLastError();
}

hipError_t launch_cfl_to_host(const unsigned long long *set, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(cfl_to_host_kernel, dim3(1), dim3(64), 0, st, set, out);
    return hipGetLastError();
}

hipError_t launch_tile_list_reset(unsigned char *flag, int ntx, int nty, int *list, int *nf, hipStream_t st)
{
    const int nb = std::min(1024, (ntx * nty + 255) / 256);
    hipLaunchKernelGGL(tile_list_reset_kernel, dim3(std::max(nb, 1)), dim3(256), 0, st, flag, ntx, nty, list, nf);
    return hipGetLastError();
}

template <int P, int CHAR, class T, int TX, int TY, bool CAPA = false>
__global__ __launch_bounds__(256) void stage2d_kernel(StageArgs a, PhysConst<T> c)
{
    const int b = tile_of_workgroup(a, blockIdx.x, a.order == TILE_ADAPTIVE ? *a.tnf : 0);
    stage2d_tile<P, CHAR, T, TX, TY, CAPA>(a, c, b);
}

// Stage k of the resident loop's current iteration.  The grid's fixed
// arguments (shape, boundaries, combination, tile list, CFL set) come by
// value as for stage2d_kernel, with the three q buffers in qin, qold and
// qout; the iteration's buffer roles, dt and tile count come from the
// control block (one 64-byte record, ResStage), which the control launch
// wrote after the last iteration: a workgroup waits for one scalar load, as
// stage2d_kernel does for its tile count.  (The buffers are chosen among the
// argument pointers rather than read from memory: the compiler knows
// argument pointers to be global memory and issues global loads and stores
// through them, where a pointer read from memory would take flat ones.)
// Nothing runs once the loop is done.
template <int P, int CHAR, class T, int TX, int TY>
__global__ __launch_bounds__(256) void stage2d_res_kernel(StageArgs a, const Res2D *R, int k, PhysConst<T> c)
{
    const ResStage r = R->st[k];
    if (!r.go) return;
    const double *b0 = a.qin, *b1 = a.qold;
    double *b2 = a.qout;
    auto pick = [&](int i) { return i == 0 ? b0 : i == 1 ? b1 : (const double *)b2; };
    a.qin = pick(r.in);
    a.qold = pick(r.old);
    a.qout = r.out == 0 ? const_cast<double *>(b0) : r.out == 1 ? const_cast<double *>(b1) : b2;
    a.dtdx = r.dtdx;
    a.dtdy = r.dtdy;
    a.dt = r.dt;
    const int b = tile_of_workgroup(a, blockIdx.x, r.nf);
    stage2d_tile<P, CHAR, T, TX, TY>(a, c, b);
}

// The control launch of the resident loop (one block of 1024 threads).
// do_post, after an iteration's stages: its CFL from the last stage's set,
// the loop's decisions (loop2d_post), and the tile list from the last
// stage's flags (tile_partition_block).  do_pre: open the next iteration
// (loop2d_pre) and write what its stage launches need (ResStage: buffers
// rotated and dt formed as the host-driven step does), or mark them as
// nothing to do; clear the words of the iteration's CFL set.
__global__ __launch_bounds__(1024) void res2d_ctl_kernel(Res2D *R, const unsigned char *tflag, int *tlist, int *tnf,
                                                         unsigned long long *set, int ntx, int nty, int do_post,
                                                         int do_pre)
{
    __shared__ double sxy[2];
    __shared__ int opened;
    const int t = threadIdx.x;
    const bool ran = do_post && R->s.go;
    // thread 0 reads the loop state and the grid's limits now, so that their
    // latency passes under the partition
    SolveState s;
    Grid2D G;
    if (t == 0) {
        s = R->s;
        G = R->G;
    }
    if (ran && t < 64) cfl_set_to(set, sxy);
    // the tile list of the next iteration (kept when no iteration ran)
    const int nf = ran ? tile_partition_block(tflag, ntx, nty, tlist, tnf) : *tnf;
    __syncthreads();
    if (t == 0) {
        if (ran) loop2d_post(G, s, R->nst, sxy[0], sxy[1]);
        const bool op = do_pre && loop2d_pre(G, s);
        int in = s.cur;
        for (int k = 0; k < R->nst; k++) {
            const int out = other_buf(in, s.old);
            ResStage r;
            r.in = in;
            r.old = s.old;
            r.out = out;
            r.dtdx = s.dt / G.dx;
            r.dtdy = s.dt / G.dy;
            r.dt = s.dt;
            r.go = op ? 1 : 0;
            r.nf = nf;
            R->st[k] = r;
            in = out;
        }
        R->s = s;
        opened = op;
    }
    __syncthreads();
    // the x and y words of every copy (the words the stages merge into)
    if (opened && t < 2 * CFL_SUB) set[(t >> 1) * CFL_STRIDE + (t & 1)] = 0ull;
}

// ------------------------------------------------------------ ghost fill

__global__ void bc1_kernel(double *q, int meqn, int n, int LX, int bl, int br)
{
    int t = threadIdx.x;  // 0..5 -> ghost cells -2..0, n+1..n+3
    if (t >= 6) return;
    int j = t < 3 ? t - 2 : n + t - 2;
    bool neg;
    int js = bc_src(j, n, bl, br, neg);
    for (int m = 0; m < meqn; m++) {
        double v = q[(size_t)(js + 2) + (size_t)m * LX];
        q[(size_t)(j + 2) + (size_t)m * LX] = (m == 1 && neg) ? -v : v;
    }
}

// x ghosts for every row (ghost rows included), as bc2.f90:31-116
__global__ void bc2x_kernel(double *q, int meqn, int nx, int ny, int LX, int LY, int bl, int br)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 6 * LY) return;
    int jj = idx / 6, t = idx % 6;
    int j = jj - 2;
    int i = t < 3 ? t - 2 : nx + t - 2;
    bool neg;
    int is = bc_src(i, nx, bl, br, neg);
    size_t plane = (size_t)LX * LY;
    for (int m = 0; m < meqn; m++) {
        double v = q[(size_t)(is + 2) + (size_t)LX * (j + 2) + m * plane];
        q[(size_t)(i + 2) + (size_t)LX * (j + 2) + m * plane] = (m == 1 && neg) ? -v : v;
    }
}

// then y ghosts for every column (ghost columns included), bc2.f90:118-206
__global__ void bc2y_kernel(double *q, int meqn, int nx, int ny, int LX, int LY, int bl, int br)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 6 * LX) return;
    int ii = idx % LX, t = idx / LX;
    int i = ii - 2;
    int j = t < 3 ? t - 2 : ny + t - 2;
    bool neg;
    int js = bc_src(j, ny, bl, br, neg);
    size_t plane = (size_t)LX * LY;
    for (int m = 0; m < meqn; m++) {
        double v = q[(size_t)(i + 2) + (size_t)LX * (js + 2) + m * plane];
        q[(size_t)(i + 2) + (size_t)LX * (j + 2) + m * plane] = (m == 2 && neg) ? -v : v;
    }
}

// ------------------------------------------------- single-routine kernels

// rp on slices in the reference layout: i = 2-mbc..mx+mbc
template <int P, int IXY, class T>
__global__ void rp_kernel(int mx, int mbc, const double *ql, const double *qr, double *wave,
                          double *sp, double *amdq, double *apdq, PhysConst<T> c)
{
    constexpr int M = Shape<P>::M, W = Shape<P>::W;
    const int L = mx + 2 * mbc;
    int i = 2 - mbc + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i > mx + mbc) return;
    T lf[M], rt[M], wv[M * W], s[W], am[M], ap[M];
    int o = i + mbc - 1;
#pragma unroll
    for (int m = 0; m < M; m++) {
        lf[m] = T(qr[o - 1 + m * L]);
        rt[m] = T(ql[o + m * L]);
    }
    riemann<P, IXY, T>(c, lf, rt, wv, s, am, ap);
#pragma unroll
    for (int w = 0; w < W; w++) {
        sp[o + w * L] = s[w];
#pragma unroll
        for (int m = 0; m < M; m++) wave[o + L * (m + M * w)] = wv[m + M * w];
    }
#pragma unroll
    for (int m = 0; m < M; m++) {
        amdq[o + m * L] = am[m];
        apdq[o + m * L] = ap[m];
    }
}

template <int P, int IXY, class T>
__global__ void evec_kernel(int mx, int mbc, const double *q, double *evl, double *evr,
                            PhysConst<T> c)
{
    constexpr int M = Shape<P>::M;
    const int L = mx + 2 * mbc;
    int i = 2 - mbc + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i > mx + mbc) return;
    int o = i + mbc - 1;
    T qa[M], qb[M], l[M * M], r[M * M];
#pragma unroll
    for (int m = 0; m < M; m++) {
        qa[m] = T(q[o - 1 + m * L]);
        qb[m] = T(q[o + m * L]);
    }
    eigvec<P, IXY, T>(c, qa, qb, l, r);
#pragma unroll
    for (int p = 0; p < M; p++)
#pragma unroll
        for (int m = 0; m < M; m++) {
            evl[o + L * (p + M * m)] = l[p * M + m];  // evl(i,p,m)
            evr[o + L * (m + M * p)] = r[m * M + p];  // evr(i,m,p)
        }
}

// weno5 / weno5_char on one slice in assumed-shape layout (1..mx2); the
// interfaces I = 4..mx2-2 (both sides defined in the reference)
template <int M, int CHAR, class T>
__global__ void weno_kernel(int mx2, const double *q, const double *evl, const double *evr,
                            double *ql, double *qr)
{
    int I = 4 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (I > mx2 - 2) return;
    T v[6][M];
#pragma unroll
    for (int k = 0; k < 6; k++)
#pragma unroll
        for (int m = 0; m < M; m++) v[k][m] = T(q[(I - 4 + k) + m * mx2]);
    T d[5][M];
#pragma unroll
    for (int k = 0; k < 5; k++)
#pragma unroll
        for (int m = 0; m < M; m++) d[k][m] = v[k + 1][m] - v[k][m];
    T lf[M], rt[M];
#pragma unroll
    for (int m = 0; m < M; m++) {
        T central = (-v[1][m] + T(7) * (v[2][m] + v[3][m]) - v[4][m]) * (T(1) / T(12));
        lf[m] = central;
        rt[m] = central;
    }
    if constexpr (CHAR == 0) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            lf[m] = weno_corr<T>(T(1), d[0][m], d[1][m], d[2][m], d[3][m]) + lf[m];
            rt[m] = weno_corr<T>(T(-1), d[4][m], d[3][m], d[2][m], d[1][m]) + rt[m];
        }
    } else {
        const int mi = I - 1;
#pragma unroll
        for (int p = 0; p < M; p++) {
            T hh[5];
#pragma unroll
            for (int k = 0; k < 5; k++) {
                T acc = T(0);
#pragma unroll
                for (int m = 0; m < M; m++) acc = acc + T(evl[mi + mx2 * (p + M * m)]) * d[k][m];
                hh[k] = acc;
            }
            T u1 = weno_corr<T>(T(1), hh[0], hh[1], hh[2], hh[3]);
            T u2 = weno_corr<T>(T(-1), hh[4], hh[3], hh[2], hh[1]);
#pragma unroll
            for (int m = 0; m < M; m++) {
                T e = T(evr[mi + mx2 * (m + M * p)]);
                lf[m] = lf[m] + e * u1;
                rt[m] = rt[m] + e * u2;
            }
        }
    }
#pragma unroll
    for (int m = 0; m < M; m++) {
        qr[(I - 2) + m * mx2] = lf[m];  // qr(I-1)
        ql[(I - 1) + m * mx2] = rt[m];  // ql(I)
    }
}

// ------------------------------------------------------------ reductions

__global__ void sums_partial(int ndim, const double *q, int nx, int ny, int LX, int LY,
                             double *partial, int nblk)
{
    const int m = blockIdx.y;
    const long long ncell = (long long)nx * (ndim == 2 ? ny : 1);
    double acc = 0.0;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < ncell;
         k += (long long)gridDim.x * blockDim.x) {
        int i = (int)(k % nx) + 1;
        int j = ndim == 2 ? (int)(k / nx) + 1 : 1;
        size_t g = (size_t)(i + 2) + (ndim == 2 ? (size_t)LX * (j + 2) + (size_t)m * LX * LY
                                                  : (size_t)m * LX);
        acc += q[g];
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)m * nblk + blockIdx.x] = red[0];
}

__global__ void sums_final(const double *partial, int nblk, double *out)
{
    const int m = blockIdx.x;
    if (threadIdx.x == 0) {
        double acc = 0.0;
        for (int b = 0; b < nblk; b++) acc += partial[(size_t)m * nblk + b];
        out[m] = acc;
    }
}

// out1.f90:63-68, out2.f90:60-67: interior values with |q| < 1e-99 are
// set to +0 in q itself (NaN compares false and is kept)
__global__ void flush_small_kernel(int ndim, double *q, int nx, int ny, int LX, int LY, int meqn)
{
    const long long ncell = (long long)nx * (ndim == 2 ? ny : 1);
    const long long total = ncell * meqn;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < total;
         k += (long long)gridDim.x * blockDim.x) {
        int m = (int)(k / ncell);
        long long c = k - (long long)m * ncell;
        int i = (int)(c % nx) + 1;
        int j = ndim == 2 ? (int)(c / nx) + 1 : 1;
        size_t g = (size_t)(i + 2) + (ndim == 2 ? (size_t)LX * (j + 2) + (size_t)m * LX * LY
                                                  : (size_t)m * LX);
        if (fabs(q[g]) < 1.0e-99) q[g] = 0.0;
    }
}

__global__ void copy_kernel(double2 *__restrict__ dst, const double2 *__restrict__ src, size_t n2)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n2;
         k += (size_t)gridDim.x * blockDim.x)
        dst[k] = src[k];
}

// -------------------------------------------------------------- launchers

// Events that the next stage launch of this host thread reports its device
// start and end times to (set_launch_events; null: a plain launch).  A
// launch with events goes through hipExtLaunchKernelGGL, which records them
// at the kernel's own start and completion, so their elapsed time is the
// kernel's duration without the host's launch latency.
static thread_local hipEvent_t t_ev_start = nullptr, t_ev_stop = nullptr;

void set_launch_events(hipEvent_t start, hipEvent_t stop)
{
    t_ev_start = start;
    t_ev_stop = stop;
}

template <typename K, typename... Args>
static hipError_t launch_k(K kernel, dim3 grid, dim3 block, hipStream_t st, hipEvent_t start, hipEvent_t stop,
                           Args... args)
{
    if (start || stop)
        hipExtLaunchKernelGGL(kernel, grid, block, 0, st, start, stop, 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
    return hipGetLastError();
}

constexpr int NT1 = 256;

template <int P, int IXY, int CHAR, class T>
static hipError_t go1(const StageArgs &a, const PhysConst<T> &c, hipStream_t st)
{
    int nblk = (a.nx + NT1 - 2) / (NT1 - 1);
    return launch_k(stage1d_kernel<P, IXY, CHAR, T, NT1>, dim3(nblk), dim3(NT1), st, t_ev_start, t_ev_stop, a, c);
}

template <int P, int IXY, int CHAR>
static hipError_t go1p(const PhysSel &s, const StageArgs &a, const PhysConst<double> &cd,
                       const PhysConst<float> &cf, hipStream_t st)
{
    return s.fp32 ? go1<P, IXY, CHAR, float>(a, cf, st) : go1<P, IXY, CHAR, double>(a, cd, st);
}

template <int P, int IXY>
static hipError_t go1c(const PhysSel &s, const StageArgs &a, const PhysConst<double> &cd,
                       const PhysConst<float> &cf, hipStream_t st)
{
    if constexpr (P == ADV1 || P == AC1 || P == EU1 || P == SW1) {
        // the other reconstructions of flux1.f90:65-117 (1D grids)
        if (s.lim == 1) {
            if (s.chr == 1) return go1p<P, IXY, RK_TVD_WAVE>(s, a, cd, cf, st);
            if (s.chr == 2) return go1p<P, IXY, RK_TVD_CHAR>(s, a, cd, cf, st);
            return hipErrorInvalidValue;
        }
        if (s.chr == 1) return go1p<P, IXY, RK_WENO_WAVE>(s, a, cd, cf, st);
        if (s.chr == 3) return go1p<P, IXY, RK_WENO_TRANS>(s, a, cd, cf, st);
    }
    if (s.lim != 2 || (s.chr != 0 && s.chr != 2)) return hipErrorInvalidValue;
    return s.chr == 2 ? go1p<P, IXY, 2>(s, a, cd, cf, st) : go1p<P, IXY, 0>(s, a, cd, cf, st);
}

hipError_t launch_stage1d(const PhysSel &s, int ixy, const StageArgs &a, const PhysConst<double> &cd,
                          const PhysConst<float> &cf, hipStream_t st)
{
    switch (s.phys) {
    case ADV1: return go1c<ADV1, 1>(s, a, cd, cf, st);
    case AC1: return go1c<AC1, 1>(s, a, cd, cf, st);
    case EU1: return go1c<EU1, 1>(s, a, cd, cf, st);
    case SW1: return go1c<SW1, 1>(s, a, cd, cf, st);
    case EU2: return ixy == 1 ? go1c<EU2, 1>(s, a, cd, cf, st) : go1c<EU2, 2>(s, a, cd, cf, st);
    case SW2: return ixy == 1 ? go1c<SW2, 1>(s, a, cd, cf, st) : go1c<SW2, 2>(s, a, cd, cf, st);
    }
    return hipErrorInvalidValue;
}

// slice sweeps: x rows of 128 interfaces (a 1024-cell row is 9 chunks), y
// groups of 4 columns x 64 interfaces (1024^2 Euler stage with weno5_trans:
// 139 us, against 156 us with 8 columns x 32; the others alike)
constexpr int NTX = 128, NTY = 64, NLY = 4;

template <int P, int CHAR, class T>
static hipError_t gos(const StageArgs &a, const PhysConst<T> &c, double *dqx, hipStream_t st)
{
    // blocks: chunks x slice groups (sweep2d_kernel maps the index to both)
    const long long nbx = (long long)((a.nx + NTX - 2) / (NTX - 1)) * (a.ny + 2);
    const long long nby = (long long)((a.ny + NTY - 2) / (NTY - 1)) * ((a.nx + 2 + NLY - 1) / NLY);
    if (nbx > 0x7fffffffLL || nby > 0x7fffffffLL) return hipErrorInvalidValue;
    hipError_t e = launch_k(sweep2d_kernel<P, 1, CHAR, T, NTX, 1>, dim3((unsigned)nbx), dim3(NTX), st, t_ev_start,
                            (hipEvent_t) nullptr, a, c, dqx);
    if (e != hipSuccess) return e;
    return launch_k(sweep2d_kernel<P, 2, CHAR, T, NTY, NLY>, dim3((unsigned)nby), dim3(NTY * NLY), st,
                    (hipEvent_t) nullptr, t_ev_stop, a, c, dqx);
}

template <int P, int CHAR>
static hipError_t gosp(const PhysSel &s, const StageArgs &a, double *dqx, const PhysConst<double> &cd,
                       const PhysConst<float> &cf, hipStream_t st)
{
    return s.fp32 ? gos<P, CHAR, float>(a, cf, dqx, st) : gos<P, CHAR, double>(a, cd, dqx, st);
}

template <int P>
static hipError_t gosc(const PhysSel &s, const StageArgs &a, double *dqx, const PhysConst<double> &cd,
                       const PhysConst<float> &cf, hipStream_t st)
{
    if (s.lim == 1) {
        if (s.chr == 1) return gosp<P, RK_TVD_WAVE>(s, a, dqx, cd, cf, st);
        if (s.chr == 2) return gosp<P, RK_TVD_CHAR>(s, a, dqx, cd, cf, st);
        return hipErrorInvalidValue;
    }
    if (s.lim != 2) return hipErrorInvalidValue;
    if (s.chr == 1) return gosp<P, RK_WENO_WAVE>(s, a, dqx, cd, cf, st);
    if (s.chr == 3) return gosp<P, RK_WENO_TRANS>(s, a, dqx, cd, cf, st);
    return hipErrorInvalidValue;  // weno5 / weno5_char: the tile kernel
}

bool sweep2d_path(const PhysSel &s) { return s.lim != 2 || s.chr == 1 || s.chr == 3; }

hipError_t launch_sweep2d(const PhysSel &s, const StageArgs &a, double *dqx, const PhysConst<double> &cd,
                          const PhysConst<float> &cf, hipStream_t st)
{
    if (a.nx < 1 || a.ny < 1 || !dqx) return hipErrorInvalidValue;
    if (s.phys == EU2) return gosc<EU2>(s, a, dqx, cd, cf, st);
    if (s.phys == SW2) return gosc<SW2>(s, a, dqx, cd, cf, st);
    return hipErrorInvalidValue;
}

// Tile shapes of the 2D stage kernel (256 threads each): a thread per cell
// plus spare threads for the tile's edge interfaces, TX*TY + max(TX, TY)
// <= 256, so every sweep is one pass.
struct TileShape {
    int tx, ty;
};
static const TileShape kTiles[NTILE2D] = {{16, 15}, {15, 16}, {32, 7}, {14, 17}};

template <int P, int CHAR, class T, int TX, int TY, bool CAPA = false>
static hipError_t go2t(const StageArgs &a0, const PhysConst<T> &c, hipStream_t st)
{
    StageArgs a = a0;
    a.ntx = (a.nx + TX - 1) / TX;
    a.nty = (a.ny + TY - 1) / TY;
    return launch_k(stage2d_kernel<P, CHAR, T, TX, TY, CAPA>, dim3(a.ntx * a.nty), dim3(256), st, t_ev_start,
                    t_ev_stop, a, c);
}

template <int P, int CHAR, class T>
static hipError_t go2(int tile, const StageArgs &a, const PhysConst<T> &c, hipStream_t st)
{
    if constexpr (P == EU2 || P == SW2) {
        // a capacity (flux1.f90:56-61) runs on the 16x15 tile only
        if (a.capa) return tile == 0 ? go2t<P, CHAR, T, 16, 15, true>(a, c, st) : hipErrorInvalidValue;
    }
    switch (tile) {
    case 1: return go2t<P, CHAR, T, 15, 16>(a, c, st);
    case 2: return go2t<P, CHAR, T, 32, 7>(a, c, st);
    case 3: return go2t<P, CHAR, T, 14, 17>(a, c, st);
    default: return go2t<P, CHAR, T, 16, 15>(a, c, st);
    }
}

template <int P>
static hipError_t go2c(const PhysSel &s, int tile, const StageArgs &a, const PhysConst<double> &cd,
                       const PhysConst<float> &cf, hipStream_t st)
{
    if (s.chr == 2)
        return s.fp32 ? go2<P, 2, float>(tile, a, cf, st) : go2<P, 2, double>(tile, a, cd, st);
    return s.fp32 ? go2<P, 0, float>(tile, a, cf, st) : go2<P, 0, double>(tile, a, cd, st);
}

hipError_t launch_stage2d(const PhysSel &s, int tile, const StageArgs &a, const PhysConst<double> &cd,
                          const PhysConst<float> &cf, hipStream_t st)
{
    if (tile < 0 || tile >= NTILE2D) tile = 0;
    switch (s.phys) {
    case EU2: return go2c<EU2>(s, tile, a, cd, cf, st);
    case SW2: return go2c<SW2>(s, tile, a, cd, cf, st);
    case ADV2:
        // component-wise only (the scalar's characteristic form is the same)
        return s.fp32 ? go2<ADV2, 0, float>(tile, a, cf, st) : go2<ADV2, 0, double>(tile, a, cd, st);
    }
    return hipErrorInvalidValue;
}

template <int P, int CHAR, class T, int TX, int TY>
static hipError_t go2r(const StageArgs &a, const Res2D *R, int k, const PhysConst<T> &c, hipStream_t st)
{
    hipLaunchKernelGGL((stage2d_res_kernel<P, CHAR, T, TX, TY>), dim3(a.ntx * a.nty), dim3(256), 0, st, a, R, k, c);
    return hipGetLastError();
}

template <int P, int CHAR, class T>
static hipError_t go2r_t(int tile, const StageArgs &a, const Res2D *R, int k, const PhysConst<T> &c, hipStream_t st)
{
    switch (tile) {
    case 1: return go2r<P, CHAR, T, 15, 16>(a, R, k, c, st);
    case 2: return go2r<P, CHAR, T, 32, 7>(a, R, k, c, st);
    case 3: return go2r<P, CHAR, T, 14, 17>(a, R, k, c, st);
    default: return go2r<P, CHAR, T, 16, 15>(a, R, k, c, st);
    }
}

template <int P>
static hipError_t go2r_p(const PhysSel &s, int tile, const StageArgs &a, const Res2D *R, int k,
                         const PhysConst<double> &cd, const PhysConst<float> &cf, hipStream_t st)
{
    if (s.chr == 2)
        return s.fp32 ? go2r_t<P, 2, float>(tile, a, R, k, cf, st) : go2r_t<P, 2, double>(tile, a, R, k, cd, st);
    return s.fp32 ? go2r_t<P, 0, float>(tile, a, R, k, cf, st) : go2r_t<P, 0, double>(tile, a, R, k, cd, st);
}

hipError_t launch_stage2d_res(const PhysSel &s, int tile, const StageArgs &a, const Res2D *R, int k,
                              const PhysConst<double> &cd, const PhysConst<float> &cf, hipStream_t st)
{
    if (tile < 0 || tile >= NTILE2D) tile = 0;
    if (a.ntx * a.nty < 1) return hipErrorInvalidValue;
    switch (s.phys) {
    case EU2: return go2r_p<EU2>(s, tile, a, R, k, cd, cf, st);
    case SW2: return go2r_p<SW2>(s, tile, a, R, k, cd, cf, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_res2d_ctl(Res2D *R, const unsigned char *tflag, int *tlist, int *tnf, unsigned long long *set,
                            int ntx, int nty, int do_post, int do_pre, hipStream_t st)
{
    if ((long long)ntx * nty > 16384) return hipErrorInvalidValue;
    hipLaunchKernelGGL(res2d_ctl_kernel, dim3(1), dim3(1024), 0, st, R, tflag, tlist, tnf, set, ntx, nty, do_post,
                       do_pre);
    return hipGetLastError();
}

void tile2d_shape(int tile, int *tx, int *ty)
{
    if (tile < 0 || tile >= NTILE2D) tile = 0;
    *tx = kTiles[tile].tx;
    *ty = kTiles[tile].ty;
}

hipError_t launch_bc(int ndim, int meqn, double *q, int nx, int ny, int LX, int LY, const int *bc,
                     hipStream_t st)
{
    if (ndim == 1) {
        hipLaunchKernelGGL(bc1_kernel, dim3(1), dim3(64), 0, st, q, meqn, nx, LX, bc[0], bc[1]);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(bc2x_kernel, dim3((6 * LY + 255) / 256), dim3(256), 0, st, q, meqn, nx, ny, LX,
                       LY, bc[0], bc[1]);
    hipLaunchKernelGGL(bc2y_kernel, dim3((6 * LX + 255) / 256), dim3(256), 0, st, q, meqn, nx, ny, LX,
                       LY, bc[2], bc[3]);
    return hipGetLastError();
}

template <int P, int IXY, class T>
static hipError_t gorp(int mx, int mbc, const double *ql, const double *qr, double *wave, double *sp,
                       double *amdq, double *apdq, const PhysConst<T> &c, hipStream_t st)
{
    int n = mx + 2 * mbc - 1;
    hipLaunchKernelGGL((rp_kernel<P, IXY, T>), dim3((n + 127) / 128), dim3(128), 0, st, mx, mbc, ql,
                       qr, wave, sp, amdq, apdq, c);
    return hipGetLastError();
}

template <int P, int IXY>
static hipError_t gorpp(const PhysSel &s, int mx, int mbc, const double *ql, const double *qr,
                        double *wave, double *sp, double *amdq, double *apdq,
                        const PhysConst<double> &cd, const PhysConst<float> &cf, hipStream_t st)
{
    return s.fp32 ? gorp<P, IXY, float>(mx, mbc, ql, qr, wave, sp, amdq, apdq, cf, st)
                  : gorp<P, IXY, double>(mx, mbc, ql, qr, wave, sp, amdq, apdq, cd, st);
}

hipError_t launch_rp(const PhysSel &s, int ixy, int mx, int mbc, const double *ql, const double *qr,
                     double *wave, double *sp, double *amdq, double *apdq,
                     const PhysConst<double> &cd, const PhysConst<float> &cf, hipStream_t st)
{
#define RPA s, mx, mbc, ql, qr, wave, sp, amdq, apdq, cd, cf, st
    switch (s.phys) {
    case ADV1: return gorpp<ADV1, 1>(RPA);
    case AC1: return gorpp<AC1, 1>(RPA);
    case EU1: return gorpp<EU1, 1>(RPA);
    case SW1: return gorpp<SW1, 1>(RPA);
    case EU2: return ixy == 1 ? gorpp<EU2, 1>(RPA) : gorpp<EU2, 2>(RPA);
    case SW2: return ixy == 1 ? gorpp<SW2, 1>(RPA) : gorpp<SW2, 2>(RPA);
    }
#undef RPA
    return hipErrorInvalidValue;
}

template <int P, int IXY, class T>
static hipError_t goev(int mx, int mbc, const double *q, double *evl, double *evr,
                       const PhysConst<T> &c, hipStream_t st)
{
    int n = mx + 2 * mbc - 1;
    hipLaunchKernelGGL((evec_kernel<P, IXY, T>), dim3((n + 127) / 128), dim3(128), 0, st, mx, mbc, q,
                       evl, evr, c);
    return hipGetLastError();
}

template <int P, int IXY>
static hipError_t goevp(const PhysSel &s, int mx, int mbc, const double *q, double *evl, double *evr,
                        const PhysConst<double> &cd, const PhysConst<float> &cf, hipStream_t st)
{
    return s.fp32 ? goev<P, IXY, float>(mx, mbc, q, evl, evr, cf, st)
                  : goev<P, IXY, double>(mx, mbc, q, evl, evr, cd, st);
}

hipError_t launch_evec(const PhysSel &s, int ixy, int mx, int mbc, const double *q, double *evl,
                       double *evr, const PhysConst<double> &cd, const PhysConst<float> &cf,
                       hipStream_t st)
{
#define EVA s, mx, mbc, q, evl, evr, cd, cf, st
    switch (s.phys) {
    case ADV1: return goevp<ADV1, 1>(EVA);
    case AC1: return goevp<AC1, 1>(EVA);
    case EU1: return goevp<EU1, 1>(EVA);
    case SW1: return goevp<SW1, 1>(EVA);
    case EU2: return ixy == 1 ? goevp<EU2, 1>(EVA) : goevp<EU2, 2>(EVA);
    case SW2: return ixy == 1 ? goevp<SW2, 1>(EVA) : goevp<SW2, 2>(EVA);
    }
#undef EVA
    return hipErrorInvalidValue;
}

template <int M, int CHAR, class T>
static hipError_t gow(int mx2, const double *q, const double *evl, const double *evr, double *ql,
                      double *qr, hipStream_t st)
{
    int n = mx2 - 5;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((weno_kernel<M, CHAR, T>), dim3((n + 127) / 128), dim3(128), 0, st, mx2, q, evl,
                       evr, ql, qr);
    return hipGetLastError();
}

template <int M>
static hipError_t gowm(int chr, int fp32, int mx2, const double *q, const double *evl,
                       const double *evr, double *ql, double *qr, hipStream_t st)
{
    if (chr == 2)
        return fp32 ? gow<M, 2, float>(mx2, q, evl, evr, ql, qr, st)
                    : gow<M, 2, double>(mx2, q, evl, evr, ql, qr, st);
    return fp32 ? gow<M, 0, float>(mx2, q, evl, evr, ql, qr, st)
                : gow<M, 0, double>(mx2, q, evl, evr, ql, qr, st);
}

hipError_t launch_weno(int meqn, int chr, int fp32, int mx2, const double *q, const double *evl,
                       const double *evr, double *ql, double *qr, hipStream_t st)
{
    switch (meqn) {
    case 1: return gowm<1>(chr, fp32, mx2, q, evl, evr, ql, qr, st);
    case 2: return gowm<2>(chr, fp32, mx2, q, evl, evr, ql, qr, st);
    case 3: return gowm<3>(chr, fp32, mx2, q, evl, evr, ql, qr, st);
    case 4: return gowm<4>(chr, fp32, mx2, q, evl, evr, ql, qr, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_sums(int ndim, int meqn, const double *q, int nx, int ny, int LX, int LY,
                       double *partial, int nblk, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(sums_partial, dim3(nblk, meqn), dim3(256), 0, st, ndim, q, nx, ny, LX, LY,
                       partial, nblk);
    hipLaunchKernelGGL(sums_final, dim3(meqn), dim3(64), 0, st, partial, nblk, out);
    return hipGetLastError();
}

hipError_t launch_flush_small(int ndim, int meqn, double *q, int nx, int ny, int LX, int LY,
                              hipStream_t st)
{
    long long total = (long long)nx * (ndim == 2 ? ny : 1) * meqn;
    int nblk = (int)std::min<long long>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(flush_small_kernel, dim3(nblk), dim3(256), 0, st, ndim, q, nx, ny, LX, LY, meqn);
    return hipGetLastError();
}

hipError_t launch_copy(double *dst, const double *src, size_t n, hipStream_t st)
{
    size_t n2 = n / 2;
    hipLaunchKernelGGL(copy_kernel, dim3(2048), dim3(256), 0, st, (double2 *)dst,
                       (const double2 *)src, n2);
    return hipGetLastError();
}

}  // namespace sc
