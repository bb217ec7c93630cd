// Building blocks shared by the fused stage kernels (sc_kernels.hip) and
// the batched persistent kernel (sc_batch.hip).
#pragma once
#include "sc_kernels.h"
#include "sc_weno.h"

namespace sc {

// --------------------------------------------------------------- helpers

__device__ __forceinline__ void combine(int op, double qo, double q, double r, double qbi,
                                        double &out, double &qbo)
{
    switch (op) {
    case OP_DQ: out = r; break;
    case OP_ADD1: out = q + r; break;
    case OP_SSP33_2: out = 0.75 * qo + 0.25 * (q + r); break;
    case OP_SSP33_3: out = (qo + 2.0 * (q + r)) / 3.0; break;
    case OP_SSP22_2: out = 0.5 * (qo + q + r); break;
    case OP_ADD_THIRD: out = q + (1.0 / 3.0) * r; break;
    case OP_SSP42_4: out = 0.25 * qo + 0.75 * q + 0.25 * r; break;
    case OP_ADD_SIXTH: out = q + (1.0 / 6.0) * r; break;
    case OP_SSP104_5: {
        double r2 = q + (1.0 / 6.0) * r;
        qbo = 0.04 * qo + 0.36 * r2;
        out = 0.6 * qo + 0.4 * r2;
        break;
    }
    default: out = qbi + 0.6 * q + 0.1 * r; break;  // OP_SSP104_10
    }
}

__device__ __forceinline__ bool op_reads_qold(int op)
{
    return op == OP_SSP33_2 || op == OP_SSP33_3 || op == OP_SSP22_2 || op == OP_SSP42_4 ||
           op == OP_SSP104_5;
}

// Source cell of (possibly ghost) cell j for one direction of length n
// (bc1.f90:40-107, bc2.f90:39-204); neg: component mnorm is negated.
__host__ __device__ __forceinline__ int bc_src(int j, int n, int bl, int br, bool &neg)
{
    neg = false;
    if (j < 1) {
        if (bl == 1) return 1;
        if (bl == 2) return n + j;
        if (bl == 3) { neg = true; return 1 - j; }
        return j;
    }
    if (j > n) {
        if (br == 1) return n;
        if (br == 2) return j - n;
        if (br == 3) { neg = true; return 2 * n + 1 - j; }
        return j;
    }
    return j;
}

template <class T> __device__ __forceinline__ T wave_max(T v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// max|s| of the block merged into a device word with one atomic per block
// (doubles >= 0 order like their bit patterns).  red: NW doubles of LDS.
template <class T, int NW>
__device__ __forceinline__ void merge_smax(unsigned long long *slot, T v, double *red)
{
    double d = (double)wave_max(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = d;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
#pragma unroll
        for (int k = 1; k < NW; k++) m = fmax(m, red[k]);
        if (m > 0.0) atomicMax(slot, (unsigned long long)__double_as_longlong(m));
    }
}

// ------------------------------------------------------ 1D fused stage

// One chunk of a 1D stage for a block of NT threads: interfaces
// i = c0..c0+NT-1 (thread t owns interface c0+t) and the NT-1 cells
// c0..c0+NT-2, whose updated stage values are written to a.qout.  The cells
// c0-3..c0+NT+1 are staged in sq (with boundary ghost cells synthesised);
// interface results pass to the neighbouring thread through sx.  Returns
// this thread's max |s| over interfaces 1..n+1 (flux1.f90:126-133).
// Leaves with a barrier, so sq/sx can be reused by the next chunk.
template <int P, int IXY, int CHAR, class T, int NT>
__device__ __forceinline__ T stage1d_chunk(const StageArgs &a, const PhysConst<T> &c, int c0,
                                           T (*sq)[NT + 5], T (*sx)[NT])
{
    constexpr int M = Shape<P>::M, W = Shape<P>::W;
    constexpr int NC = NT - 1;
    constexpr int NL = NT + 5;
    constexpr int MN = IXY == 1 ? 1 : 2;  // normal momentum for a wall
    const int t = threadIdx.x;
    const int n = a.nx;

    for (int k = t; k < NL; k += NT) {
        int j = min(max(c0 - 3 + k, -2), n + 3);
        bool neg;
        int js = bc_src(j, n, a.bc[0], a.bc[1], neg);
#pragma unroll
        for (int m = 0; m < M; m++) {
            double val = a.qin[(size_t)(js + 2) + (size_t)m * a.LX];
            if (m == MN && neg) val = -val;
            sq[m][k] = T(val);
        }
    }
    __syncthreads();

    const int i = c0 + t;  // interface i-1/2
    T v[6][M];
#pragma unroll
    for (int k = 0; k < 6; k++)
#pragma unroll
        for (int m = 0; m < M; m++) v[k][m] = sq[m][t + k];
    T lf[M], rt[M], wave[M * W], s[W], am[M], ap[M];
    reconstruct_iface<P, IXY, CHAR, T>(c, v, lf, rt);
    riemann<P, IXY, T>(c, lf, rt, wave, s, am, ap);
    T smax = T(0);
    if (i <= n + 1) {
#pragma unroll
        for (int w = 0; w < W; w++) smax = fmax(smax, fabs(s[w]));
    }
#pragma unroll
    for (int m = 0; m < M; m++) {
        sx[m][t] = lf[m];
        sx[M + m][t] = am[m];
    }
    __syncthreads();

    if (t < NC && i <= n) {
        // in-cell problem of cell i: left ql(i) = rt, right qr(i) from interface i+1/2
        T qrc[M], am2[M], ap2[M];
#pragma unroll
        for (int m = 0; m < M; m++) qrc[m] = sx[m][t + 1];
        riemann<P, IXY, T>(c, rt, qrc, wave, s, am2, ap2);
        const bool rq = op_reads_qold(a.op);
#pragma unroll
        for (int m = 0; m < M; m++) {
            // flux1.f90:150-152 (tfluct: the in-cell problem's total
            // fluctuation adq = amdq + apdq, tfluct*.f of the applications)
            // or flux1.f90:183-186 (the same problem through rp, swapped)
            T X = c.tfluct ? (ap[m] + (am2[m] + ap2[m])) + sx[M + m][t + 1]
                           : ((sx[M + m][t + 1] + ap[m]) + am2[m]) + ap2[m];
            double r = -(a.dtdx * (double)X);
            size_t g = (size_t)(i + 2) + (size_t)m * a.LX;
            double q = a.qin[g];
            double qo = rq ? a.qold[g] : 0.0;
            double qbi = a.op == OP_SSP104_10 ? a.qbin[g] : 0.0;
            double out, qbo = 0.0;
            combine(a.op, qo, q, r, qbi, out, qbo);
            a.qout[g] = out;
            if (a.op == OP_SSP104_5) a.qbout[g] = qbo;
        }
    }
    __syncthreads();
    return smax;
}

}  // namespace sc
