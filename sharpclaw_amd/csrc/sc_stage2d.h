// The fused 2D stage on one tile, shared by the single-grid stage kernel
// (sc_kernels.hip) and the batched 2D kernels (sc_batch2d.hip).
#pragma once
#include <type_traits>

#include "sc_stage.h"

namespace sc {

// ------------------------------------------------------ 2D fused stage

// 2D advection (rpn2ad1.f:45-50): the speed of the x- (d = 0) or y- (d = 1)
// problem of grid cell (i, j) -- its left / bottom edge velocity times the
// step's factor, i.e. auxl(i,ixy) of the slice.  Indices past the ghost
// cells (cells of a partial tile beyond the grid, whose results are not
// used) are clamped.
template <class T>
__device__ __forceinline__ PhysConst<T> adv_const(const PhysConst<T> &c, const StageArgs &a, int i, int j, int d)
{
    PhysConst<T> cu = c;
    i = min(i, a.nx + 3);
    j = min(j, a.ny + 3);
    cu.u = T(a.vscale * a.aux[(size_t)(i + 2) + (size_t)a.LX * (size_t)(j + 2) +
                             (size_t)d * (size_t)a.LX * (size_t)a.LY]);
    return cu;
}

// The CFL of one interface on a grid with a capacity (flux1.f90:126-133
// with dtdx(i) = dt/(dx aux(i,mcapa)), flux1.f90:56-61): max over its waves
// of s/kappa(right cell) and -s/kappa(left cell), in units of dt/dx.
template <int W, class T>
__device__ __forceinline__ T capa_speed(const T (&s)[W], double kl, double kr)
{
    const T il = T(1) / T(kl), ir = T(1) / T(kr);
    T m = T(0);
#pragma unroll
    for (int w = 0; w < W; w++) m = fmax(m, fmax(s[w] * ir, -(s[w] * il)));
    return m;
}

// kappa of grid cell (i, j), 0..n+1 (a.capa: one field in the q layout)
__device__ __forceinline__ double capa_at(const StageArgs &a, int i, int j)
{
    return a.capa[(size_t)(i + 2) + (size_t)a.LX * (size_t)(j + 2)];
}

// A block of 256 threads updates the TX x TY tile b (row-major over the
// a.ntx x a.nty tiles).  q of the tile plus 3-cell halos along x and y is
// staged in LDS (the corner blocks are not needed: every stencil is
// one-directional).  The thread of cell (cx, cy) solves the x-interface on
// its right and, later, the y-interface above it; the TX*TY-th and later
// threads solve the left-edge x-interfaces and the bottom-edge
// y-interfaces.  A cell's in-cell problem and dq then need from its
// neighbour only the interface's right state and apdq, which pass through
// LDS; its own interface's left state and amdq stay in registers.  (Two
// exchange arrays instead of four keep the LDS per block low enough for
// four blocks per CU.)
// CFL: the reference also sweeps the ghost rows j = 0, ny+1 and columns
// i = 0, nx+1 (flux2.f90:36,68); for boundary options 1-3 their states are
// copies (wall: with the tangential momentum negated) of the adjacent
// interior row/column, whose wave speeds they therefore repeat exactly, so
// the interior sweeps give the same maximum.  (For 2D advection the speeds
// are the aux velocities, which the ghost rows do not repeat: the host adds
// their maximum, which does not depend on q.)  With a capacity (a.capa:
// aux(:,:,mcapa) in the layout of one field of q) the update of a cell is
// scaled by 1/kappa of the cell in both directions, and each interface's CFL
// speeds by the kappa of the cell the wave enters (capa_speed); a ghost row
// or column repeats the speeds of the interior one it copies but with its own
// kappa, so the tile that holds that interior row or column adds them.
// Every tile takes the full path then (a.full).  CAPA selects the
// instantiation that reads a.capa (the others ignore it).
template <int P, int CHAR, class T, int TX, int TY, bool CAPA = false>
__device__ __forceinline__ void stage2d_tile(const StageArgs &a, const PhysConst<T> &c, int b)
{
    constexpr int M = Shape<P>::M, W = Shape<P>::W;
    constexpr int NT = 256;
    static_assert(TX * TY + (TX > TY ? TX : TY) <= NT, "a thread per cell plus the edge interfaces");
    constexpr int LXT = TX + 6, LYT = TY + 6;
    constexpr int NIX = (TX + 1) * TY, NIY = TX * (TY + 1);
    constexpr int NI = NIX > NIY ? NIX : NIY;
    // Euler and shallow water take the flux form (incell_by_flux): each
    // interface passes its float64 flux F to the cell on its right / above;
    // 2D advection passes the right state and apdq
    constexpr bool FF = incell_by_flux<P>();
    __shared__ T sq[M][LYT][LXT];
    __shared__ T sR[FF ? 1 : M][NI], sAp[FF ? 1 : M][NI];
    __shared__ double sF[FF ? M : 1][NI];
    const int t = threadIdx.x;
    const int nx = a.nx, ny = a.ny;
    const int i0 = 1 + (b % a.ntx) * TX, j0 = 1 + (b / a.ntx) * TY;

    // Each staged value is also compared with the tile's first interior cell
    // (i0, j0): when every staged cell of the tile and its halos equals it,
    // every x and y stencil of the tile is constant, its WENO5 corrections
    // and all its fluctuations are exactly zero, and the stage update is the
    // RK combination with dq = 0 (up to the sign of a zero).  The barrier
    // after staging reduces the test over the block (__syncthreads_and), so
    // the branch is block-uniform.
    const size_t plane = (size_t)a.LX * a.LY;
    T ref[M];
    {
        const size_t g0 = (size_t)(i0 + 2) + (size_t)a.LX * (size_t)(j0 + 2);
#pragma unroll
        for (int m = 0; m < M; m++) ref[m] = T(a.qin[g0 + m * plane]);
    }
    // A deep-constant tile (stages 2 and 3 of a step, a.dflag set): this
    // tile and its eight neighbours were constant tiles at the step's first
    // stage, so q^n is one value V on all nine, and so is the first stage's
    // output on them (V + 0).  Every cell this stage reads -- the tile with
    // its 3-cell halos, and for the third stage the second stage's values
    // there, which depend on first-stage values up to 6 cells from the tile,
    // within the nine tiles since every tile shape is at least 7 cells wide
    // -- then holds one value, as do q^n and the register on the tile: the
    // stage is the constant tile's update below, from one cell's values,
    // without staging the tile.  A missing neighbour beyond the grid counts
    // as constant on a side with extrapolation (its ghost cells copy the
    // tile's edge); periodic, wall and user boundaries do not qualify.  The
    // values computed are those the constant-tile branch computes.
    if constexpr (P != ADV2) {
        if (a.dflag && !a.full) {
            // threads 0..8 test one tile of the 3 x 3 block each (one round
            // trip to memory), the barrier combines them
            bool ok = true;
            if (t < 9) {
                const int x = b % a.ntx + t % 3 - 1, y = b / a.ntx + t / 3 - 1;
                if (x < 0) ok = a.bc[0] == 1;
                else if (x >= a.ntx) ok = a.bc[1] == 1;
                if (y < 0) ok = ok && a.bc[2] == 1;
                else if (y >= a.nty) ok = ok && a.bc[3] == 1;
                if (x >= 0 && x < a.ntx && y >= 0 && y < a.nty) ok = ok && a.dflag[y * a.ntx + x] == 0;
            }
            if (__syncthreads_and(ok)) {
                const size_t g0 = (size_t)(i0 + 2) + (size_t)a.LX * (size_t)(j0 + 2);
                double outv[M], qbov[M];
#pragma unroll
                for (int m = 0; m < M; m++) {
                    const double qv = a.qin[g0 + m * plane];
                    const double qo = op_reads_qold(a.op) ? a.qold[g0 + m * plane] : 0.0;
                    const double qbi = a.op == OP_SSP104_10 ? a.qbin[g0 + m * plane] : 0.0;
                    qbov[m] = 0.0;
                    combine(a.op, qo, qv, 0.0, qbi, outv[m], qbov[m]);
                }
                if (a.tflag && t == 0) a.tflag[b] = 0;
                if (t == 0 && a.smax) {
                    T sx, sy;
                    flat_speeds<P, T>(c, ref, sx, sy);
                    unsigned long long *w = a.smax + ((blockIdx.x + blockIdx.y) % CFL_SUB) * CFL_STRIDE;
                    if (sx > T(0)) atomicMax(w, (unsigned long long)__double_as_longlong((double)sx));
                    if (sy > T(0)) atomicMax(w + 1, (unsigned long long)__double_as_longlong((double)sy));
                }
                const bool own = t < TX * TY;
                const int ox = own ? t % TX : 0, oy = own ? t / TX : 0;
                if (own && i0 + ox <= nx && j0 + oy <= ny) {
                    const size_t gcell = (size_t)(i0 + ox + 2) + (size_t)a.LX * (size_t)(j0 + oy + 2);
#pragma unroll
                    for (int m = 0; m < M; m++) {
                        a.qout[gcell + m * plane] = outv[m];
                        if (a.op == OP_SSP104_5) a.qbout[gcell + m * plane] = qbov[m];
                    }
                }
                return;
            }
        }
    }

    // The RK combination's other input for this thread's cell, q^n: a float64
    // tile loads it now, so that its latency overlaps the staging (constant
    // tiles then only store); a float tile, whose registers are the tighter
    // limit on its occupancy, loads it and q itself at the update.
    constexpr bool PRE = std::is_same<T, double>::value;
    const bool owns_cell = t < TX * TY;
    const int cx = owns_cell ? t % TX : 0, cy = owns_cell ? t / TX : 0;
    const bool upd = owns_cell && i0 + cx <= nx && j0 + cy <= ny;
    const size_t gc = (size_t)(i0 + cx + 2) + (size_t)a.LX * (size_t)(j0 + cy + 2);
    const bool rq = op_reads_qold(a.op);
    double qo_c[M];
#pragma unroll
    for (int m = 0; m < M; m++) qo_c[m] = PRE && upd && rq ? a.qold[gc + m * plane] : 0.0;
    // q^n and q of this thread's cell at the update
    auto qold_at = [&](int m) { return PRE ? qo_c[m] : (rq ? a.qold[gc + m * plane] : 0.0); };
    auto q_at = [&](int m) { return PRE ? (double)sq[m][cy + 3][cx + 3] : a.qin[gc + m * plane]; };
    bool same = true;
    for (int k = t; k < LYT * LXT; k += NT) {
        int ly = k / LXT, lx = k - ly * LXT;
        bool cx = lx < 3 || lx >= TX + 3, cy = ly < 3 || ly >= TY + 3;
        if (cx && cy) continue;
        int i = min(max(i0 - 3 + lx, -2), nx + 3);
        int j = min(max(j0 - 3 + ly, -2), ny + 3);
        bool negx = false, negy = false;
        if (j >= 1 && j <= ny) {
            i = bc_src(i, nx, a.bc[0], a.bc[1], negx);
        } else {
            i = min(max(i, 1), nx);
            j = bc_src(j, ny, a.bc[2], a.bc[3], negy);
        }
        size_t g = (size_t)(i + 2) + (size_t)a.LX * (size_t)(j + 2);
#pragma unroll
        for (int m = 0; m < M; m++) {
            double val = a.qin[g + m * plane];
            if ((m == 1 && negx) || (m == 2 && negy)) val = -val;
            const T tv = T(val);
            sq[m][ly][lx] = tv;
            same = same && tv == ref[m];
        }
    }
    const bool flat = __syncthreads_and(same) != 0;
    if (a.tflag && t == 0) a.tflag[b] = flat ? 0 : 1;
    __shared__ double red[2][NT / 64];

    // A constant tile: every interface of it has the same left and right
    // state, the cell value, whose Roe waves are zero, and the stage update
    // is the RK combination with dq = 0 (a constant stencil's WENO5
    // corrections are zeros, at most of the other sign).  The wave speeds of
    // every interface are u - c, u, u + c of that state, so the tile's max
    // |s| is |u| + c in x and |v| + c in y, which thread 0 forms directly and
    // merges into the CFL words.  (The full path's Roe averages of two equal
    // states give the same speeds to rounding.)  2D advection's speeds vary
    // with aux, so it always takes the full path.
    if constexpr (P != ADV2) {
        if (flat && !a.full) {
            if (t == 0 && a.smax) {
                T sx, sy;
                flat_speeds<P, T>(c, ref, sx, sy);
                unsigned long long *w = a.smax + ((blockIdx.x + blockIdx.y) % CFL_SUB) * CFL_STRIDE;
                if (sx > T(0)) atomicMax(w, (unsigned long long)__double_as_longlong((double)sx));
                if (sy > T(0)) atomicMax(w + 1, (unsigned long long)__double_as_longlong((double)sy));
            }
            if (upd) {
#pragma unroll
                for (int m = 0; m < M; m++) {
                    const size_t gm = gc + m * plane;
                    double qbi = a.op == OP_SSP104_10 ? a.qbin[gm] : 0.0;
                    double out, qbo = 0.0;
                    combine(a.op, qold_at(m), q_at(m), 0.0, qbi, out, qbo);
                    a.qout[gm] = out;
                    if (a.op == OP_SSP104_5) a.qbout[gm] = qbo;
                }
            }
            return;
        }
    }

    const int e = t - TX * TY;  // edge-interface index of the spare threads
    T smx = T(0), smy = T(0);
    T wave[M * W], s[W];
    T lf[M], am[M];  // own interface: left state, amdq
    double Fo[M];    // own interface's flux (flux form)
    double dqx[M];

    // x-interfaces: interface ci of row r lies between tile cells ci-1, ci
    {
        int r = -1, ci = 0;
        if (owns_cell) {
            r = cy;
            ci = cx + 1;
        } else if (e < TY) {
            r = e;
            ci = 0;
        }
        if (r >= 0) {
            T v[6][M], rt[M], ap[M];
#pragma unroll
            for (int kk = 0; kk < 6; kk++)
#pragma unroll
                for (int m = 0; m < M; m++) v[kk][m] = sq[m][r + 3][ci + kk];
            reconstruct_iface<P, 1, CHAR, T>(c, v, lf, rt);
            if constexpr (FF)
                riemann_flux<P, 1, T>(c, lf, rt, wave, s, am, Fo);
            else if constexpr (P == ADV2)
                riemann<P, 1, T>(adv_const(c, a, i0 + ci, j0 + r, 0), lf, rt, wave, s, am, ap);
            else
                riemann<P, 1, T>(c, lf, rt, wave, s, am, ap);
            if (i0 + ci <= nx + 1 && j0 + r <= ny) {
                if constexpr (CAPA) {
                    // interface gi between cells gi-1, gi of row gj; ghost rows
                    // 0 and ny+1 copy row 1 / ny (extrapolation, wall) or
                    // ny / 1 (periodic)
                    const int gi = i0 + ci, gj = j0 + r;
                    smx = fmax(smx, capa_speed<W, T>(s, capa_at(a, gi - 1, gj), capa_at(a, gi, gj)));
                    if ((gj == 1 && a.bc[2] != 2) || (gj == ny && a.bc[2] == 2))
                        smx = fmax(smx, capa_speed<W, T>(s, capa_at(a, gi - 1, 0), capa_at(a, gi, 0)));
                    if ((gj == ny && a.bc[3] != 2) || (gj == 1 && a.bc[3] == 2))
                        smx = fmax(smx, capa_speed<W, T>(s, capa_at(a, gi - 1, ny + 1), capa_at(a, gi, ny + 1)));
                } else {
#pragma unroll
                    for (int w = 0; w < W; w++) smx = fmax(smx, fabs(s[w]));
                }
            }
            const int k = r * (TX + 1) + ci;
#pragma unroll
            for (int m = 0; m < M; m++) {
                if constexpr (FF) {
                    Fo[m] = Fo[m] + (double)am[m];
                    sF[m][k] = Fo[m];
                } else {
                    sR[m][k] = rt[m];
                    sAp[m][k] = ap[m];
                }
            }
        }
    }
    __syncthreads();
    if (owns_cell) {
        // in-cell problem of cell (cx, cy): ql = right state of the interface
        // on its left, qr = left state of its own interface
        const int kl = cy * (TX + 1) + cx;
        if constexpr (FF) {
            // X = F(i+1/2) - F(i-1/2); dtdx(i) = dt/(dx kappa(i)) with a capacity
            double dtdx = a.dtdx;
            if constexpr (CAPA) dtdx = a.dtdx / capa_at(a, min(i0 + cx, nx + 1), min(j0 + cy, ny + 1));
#pragma unroll
            for (int m = 0; m < M; m++) dqx[m] = -(dtdx * (Fo[m] - sF[m][kl]));
        } else {
            T ql_[M];
#pragma unroll
            for (int m = 0; m < M; m++) ql_[m] = sR[m][kl];
            T am2[M], ap2[M];
            if constexpr (P == ADV2)
                riemann<P, 1, T>(adv_const(c, a, i0 + cx, j0 + cy, 0), ql_, lf, wave, s, am2, ap2);
            else
                riemann<P, 1, T>(c, ql_, lf, wave, s, am2, ap2);
#pragma unroll
            for (int m = 0; m < M; m++) {
                T X = ((am[m] + sAp[m][kl]) + am2[m]) + ap2[m];
                dqx[m] = -(a.dtdx * (double)X);
            }
        }
    }
    __syncthreads();
    // y-interfaces: interface r of column ci lies between tile rows r-1, r
    {
        int r = -1, ci = 0;
        if (owns_cell) {
            r = cy + 1;
            ci = cx;
        } else if (e < TX) {
            r = 0;
            ci = e;
        }
        if (r >= 0) {
            T v[6][M], rt[M], ap[M];
#pragma unroll
            for (int kk = 0; kk < 6; kk++)
#pragma unroll
                for (int m = 0; m < M; m++) v[kk][m] = sq[m][r + kk][ci + 3];
            reconstruct_iface<P, 2, CHAR, T>(c, v, lf, rt);
            if constexpr (FF)
                riemann_flux<P, 2, T>(c, lf, rt, wave, s, am, Fo);
            else if constexpr (P == ADV2)
                riemann<P, 2, T>(adv_const(c, a, i0 + ci, j0 + r, 1), lf, rt, wave, s, am, ap);
            else
                riemann<P, 2, T>(c, lf, rt, wave, s, am, ap);
            if (i0 + ci <= nx && j0 + r <= ny + 1) {
                if constexpr (CAPA) {
                    // interface gj between cells gj-1, gj of column gi; ghost
                    // columns 0 and nx+1 as the ghost rows above
                    const int gi = i0 + ci, gj = j0 + r;
                    smy = fmax(smy, capa_speed<W, T>(s, capa_at(a, gi, gj - 1), capa_at(a, gi, gj)));
                    if ((gi == 1 && a.bc[0] != 2) || (gi == nx && a.bc[0] == 2))
                        smy = fmax(smy, capa_speed<W, T>(s, capa_at(a, 0, gj - 1), capa_at(a, 0, gj)));
                    if ((gi == nx && a.bc[1] != 2) || (gi == 1 && a.bc[1] == 2))
                        smy = fmax(smy, capa_speed<W, T>(s, capa_at(a, nx + 1, gj - 1), capa_at(a, nx + 1, gj)));
                } else {
#pragma unroll
                    for (int w = 0; w < W; w++) smy = fmax(smy, fabs(s[w]));
                }
            }
            const int k = r * TX + ci;
#pragma unroll
            for (int m = 0; m < M; m++) {
                if constexpr (FF) {
                    Fo[m] = Fo[m] + (double)am[m];
                    sF[m][k] = Fo[m];
                } else {
                    sR[m][k] = rt[m];
                    sAp[m][k] = ap[m];
                }
            }
        }
    }
    __syncthreads();
    if (owns_cell) {
        const int kb = cy * TX + cx;
        // Y[m] = ((amdq(j+1) + apdq(j)) + amdq2(j)) + apdq2(j)
        double Y[M];
        if constexpr (FF) {
#pragma unroll
            for (int m = 0; m < M; m++) Y[m] = Fo[m] - sF[m][kb];
        } else {
            T ql_[M];
#pragma unroll
            for (int m = 0; m < M; m++) ql_[m] = sR[m][kb];
            T am2[M], ap2[M];
            if constexpr (P == ADV2)
                riemann<P, 2, T>(adv_const(c, a, i0 + cx, j0 + cy, 1), ql_, lf, wave, s, am2, ap2);
            else
                riemann<P, 2, T>(c, ql_, lf, wave, s, am2, ap2);
#pragma unroll
            for (int m = 0; m < M; m++) Y[m] = (double)(((am[m] + sAp[m][kb]) + am2[m]) + ap2[m]);
        }
        if (upd) {
            double dtdy = a.dtdy;
            if constexpr (CAPA) dtdy = a.dtdy / capa_at(a, i0 + cx, j0 + cy);
#pragma unroll
            for (int m = 0; m < M; m++) {
                double r = dqx[m] + -(dtdy * Y[m]);
                size_t gm = gc + m * plane;
                double qbi = a.op == OP_SSP104_10 ? a.qbin[gm] : 0.0;
                double out, qbo = 0.0;
                combine(a.op, qold_at(m), q_at(m), r, qbi, out, qbo);
                a.qout[gm] = out;
                if (a.op == OP_SSP104_5) a.qbout[gm] = qbo;
            }
        }
    }
    if (a.smax) {
        merge_smax<T, NT / 64>(a.smax, smx, red[0]);
        merge_smax<T, NT / 64>(a.smax + 1, smy, red[1]);
    }
}

}  // namespace sc
