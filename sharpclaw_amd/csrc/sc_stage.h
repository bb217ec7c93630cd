// Building blocks shared by the fused stage kernels (sc_kernels.hip) and
// the batched persistent kernel (sc_batch.hip).
#pragma once
#include <type_traits>

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

// CFL words.  Each workgroup merges its max |s| into a device word with one
// atomicMax (doubles >= 0 order like their bit patterns).  Thousands of
// workgroups per launch on one address queue up behind each other (tens of
// microseconds per launch at 1024^2), so each word has CFL_SUB copies
// CFL_STRIDE words (256 bytes) apart and workgroup b merges into copy
// b mod CFL_SUB (b = blockIdx.x + blockIdx.y).  A CFL set holds the x word (offset 0) and the y word
// (offset 1) with their copies, CFL_WORDS words; a reader takes the max over
// the copies (cfl_word).
constexpr int CFL_SUB = 32, CFL_STRIDE = 32, CFL_WORDS = CFL_SUB * CFL_STRIDE;

// word w (0: x, 1: y) of a CFL set, as the bit pattern of a double >= 0
__host__ __device__ inline unsigned long long cfl_word(const unsigned long long *set, int w)
{
    unsigned long long m = 0;
    for (int k = 0; k < CFL_SUB; k++) {
        unsigned long long v = set[k * CFL_STRIDE + w];
        m = v > m ? v : m;
    }
    return m;
}

// max|s| of the block merged into word `slot` (x or y word of a CFL set)
// with one atomic per block.  red: NW doubles of LDS.
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
        if (m > 0.0)
            atomicMax(slot + ((blockIdx.x + blockIdx.y) % CFL_SUB) * CFL_STRIDE, (unsigned long long)__double_as_longlong(m));
    }
}

// ------------------------------------------- 1D reconstruction variants
//
// Reconstruction codes of the 1D stage (flux1.f90:65-117):
//   0 weno5, 2 weno5_char (lim_type 2, char_decomp 0/2; sc_weno.h)
//   1 weno5_wave, 3 weno5_trans (lim_type 2, char_decomp 1/3)
//   11 tvd2_wave, 12 tvd2_char (lim_type 1, char_decomp 1/2)
// weno5_wave, weno5_trans and tvd2_wave use data of the neighbouring
// interfaces (waves of the cell-average jumps, or projected jumps and
// eigenvectors); each block computes those once per interface into LDS.

constexpr int RK_WENO = 0, RK_WENO_WAVE = 1, RK_WENO_CHAR = 2, RK_WENO_TRANS = 3;
constexpr int RK_TVD_WAVE = 11, RK_TVD_CHAR = 12;

// "1.e-14" in reconstruct.f90:404,623 is a default-real literal
template <class T> __device__ __forceinline__ T tol14() { return T((double)1.0e-14f); }

// limiter functions of reconstruct.f90:626-654, 726-753 (the same five in
// tvd2): 1 minmod, 2 superbee, 3 van Leer, 4 monotonized centred,
// 5 Cada & Torrilhon
template <class T> __device__ __forceinline__ T tvd_limiter(int kind, T r)
{
    switch (kind) {
    case 1: return fmax(T(0), fmin(T(1), r));
    case 2: return fmax(fmax(T(0), fmin(T(1), T(2) * r)), fmin(T(2), r));
    case 3: {
        T ar = fabs(r);
        return fdiv(r + ar, T(1) + ar);
    }
    case 4: return fmax(T(0), fmin(fmin((T(1) + r) * T(0.5), T(2)), T(2) * r));
    default: {
        T pp = (T(2) + r) * T(1.0 / 3.0);
        T amax = fmax(fmax(-(T(1.0 / 3.0) * r), T(0)), fmin(fmin(T(2) * r, pp), T(2)));
        return fmax(T(0), fmin(pp, amax));
    }
    }
}

// words of T in an eigenvector object (stored to LDS for weno5_trans)
template <int P, int IXY, class T> constexpr int eig_words()
{
    return std::is_empty<Eig<P, IXY, T>>::value ? 0 : (int)(sizeof(Eig<P, IXY, T>) / sizeof(T));
}

// per-interface words of neighbour data in LDS for reconstruction RK
template <int P, int IXY, int RK, class T> constexpr int nbr_words()
{
    constexpr int M = Shape<P>::M, W = Shape<P>::W;
    return (RK == RK_WENO_WAVE || RK == RK_TVD_WAVE) ? M * W + W
           : RK == RK_WENO_TRANS                      ? M + eig_words<P, IXY, T>()
                                                      : 0;
}

// Neighbour data of interface j (cells j-1, j = ql, qr) into column k of
// sw (rows of NW words): the rp waves and speeds of the cell-average jump
// (flux1.f90:98-99, 84-85: rp with ql = qr = q1d), or for weno5_trans the
// jump projected with the interface's own left eigenvectors (gg,
// reconstruct.f90:269-277) and the eigenvector object itself.
template <int P, int IXY, int RK, class T>
__device__ __forceinline__ void nbr_data(const PhysConst<T> &c, const T *ql, const T *qr, T *sw, int NW,
                                         int k)
{
    constexpr int M = Shape<P>::M, W = Shape<P>::W;
    if constexpr (RK == RK_WENO_WAVE || RK == RK_TVD_WAVE) {
        T wave[M * W], s[W], am[M], ap[M];
        riemann<P, IXY, T>(c, ql, qr, wave, s, am, ap);
#pragma unroll
        for (int e = 0; e < M * W; e++) sw[e * NW + k] = wave[e];
#pragma unroll
        for (int w = 0; w < W; w++) sw[(M * W + w) * NW + k] = s[w];
    } else if constexpr (RK == RK_WENO_TRANS) {
        Eig<P, IXY, T> e;
        e.init(c, ql, qr);
        T d[M], g[M];
#pragma unroll
        for (int m = 0; m < M; m++) d[m] = qr[m] - ql[m];
        e.to_char(d, g);
#pragma unroll
        for (int m = 0; m < M; m++) sw[m * NW + k] = g[m];
        constexpr int NE = eig_words<P, IXY, T>();
        if constexpr (NE > 0) {
            T ew[NE];
            __builtin_memcpy(ew, &e, sizeof(ew));
#pragma unroll
            for (int f = 0; f < NE; f++) sw[(M + f) * NW + k] = ew[f];
        }
    }
}

// Whether a line chunk's exchange array sx (2M words per interface) lives in
// the neighbour-data array sw (nbr_words per interface, NT+4 interfaces):
// sw is dead once every interface is reconstructed, and sharing it keeps
// the LDS of a block low enough for more blocks per CU.
template <int P, int IXY, int RK, class T, int NT> constexpr bool sx_in_sw()
{
    return nbr_words<P, IXY, RK, T>() > 0 && nbr_words<P, IXY, RK, T>() * (NT + 4) >= 2 * Shape<P>::M * NT;
}

template <int P, int IXY, class T>
__device__ __forceinline__ Eig<P, IXY, T> eig_at(const T *sw, int NW, int k)
{
    constexpr int M = Shape<P>::M, NE = eig_words<P, IXY, T>();
    Eig<P, IXY, T> e;
    if constexpr (NE > 0) {
        T ew[NE];
#pragma unroll
        for (int f = 0; f < NE; f++) ew[f] = sw[(M + f) * NW + k];
        __builtin_memcpy(&e, ew, sizeof(ew));
    }
    return e;
}

// Both sides of interface i-1/2 (lf = qr(i-1), rt = ql(i)) for the
// reconstructions other than weno5 / weno5_char.  v: cells i-3..i+2;
// sw column kc holds interface i (kc-2..kc+2: i-2..i+2).
template <int P, int IXY, int RK, class T>
__device__ __forceinline__ void reconstruct_other(const PhysConst<T> &c, const T (&v)[6][Shape<P>::M],
                                                  const T *sw, int NW, int kc, T (&lf)[Shape<P>::M],
                                                  T (&rt)[Shape<P>::M])
{
    constexpr int M = Shape<P>::M, W = Shape<P>::W;
    if constexpr (RK == RK_WENO_WAVE || RK == RK_WENO_TRANS) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            T central = (-v[1][m] + T(7) * (v[2][m] + v[3][m]) - v[4][m]) * (T(1) / T(12));
            lf[m] = central;
            rt[m] = central;
        }
    } else {
#pragma unroll
        for (int m = 0; m < M; m++) {
            lf[m] = v[2][m];
            rt[m] = v[3][m];
        }
    }
    if constexpr (RK == RK_WENO_WAVE) {
        // reconstruct.f90:364-426: WENO on the projections of the waves of
        // interfaces i-2..i+2 onto the wave of interface i, per family
#pragma unroll
        for (int mw = 0; mw < W; mw++) {
            T wi[M];
#pragma unroll
            for (int m = 0; m < M; m++) wi[m] = sw[(m + M * mw) * NW + kc];
            T dd[5];
#pragma unroll
            for (int k = 0; k < 5; k++) {
                T acc = sw[(M * mw) * NW + kc - 2 + k] * wi[0];
#pragma unroll
                for (int m = 1; m < M; m++) acc = acc + sw[(m + M * mw) * NW + kc - 2 + k] * wi[m];
                dd[k] = acc;
            }
            T u1 = T(0), u2 = T(0), sc = T(0);
            if (dd[2] > tol14<T>()) {
                weno_pair(dd[0], dd[1], dd[2], dd[3], dd[4], u1, u2);
                sc = frcp(dd[2]);
            }
#pragma unroll
            for (int m = 0; m < M; m++) {
                lf[m] = lf[m] + u1 * wi[m] * sc;
                rt[m] = rt[m] + u2 * wi[m] * sc;
            }
        }
    } else if constexpr (RK == RK_WENO_TRANS) {
        // reconstruct.f90:269-345: jumps projected with their own
        // interface's left eigenvectors (gg), carried to interface i by
        // L(i) R(i+k), WENO per field, back with R(i)
        Eig<P, IXY, T> own = eig_at<P, IXY, T>(sw, NW, kc);
        T U1[M], U2[M];
#pragma unroll
        for (int mw = 0; mw < M; mw++) {
            T hh[5];
#pragma unroll
            for (int k = 0; k < 5; k++) {
                Eig<P, IXY, T> ek = eig_at<P, IXY, T>(sw, NW, kc - 2 + k);
                T e[M], col[M], w[M];
#pragma unroll
                for (int m = 0; m < M; m++) e[m] = m == mw ? T(1) : T(0);
                ek.from_char(e, col);
                own.to_char(col, w);
                hh[k] = sw[mw * NW + kc - 2 + k] * w[mw];
            }
            weno_pair(hh[0], hh[1], hh[2], hh[3], hh[4], U1[mw], U2[mw]);
        }
        T o1[M], o2[M];
        own.from_char(U1, o1);
        own.from_char(U2, o2);
#pragma unroll
        for (int m = 0; m < M; m++) {
            lf[m] = lf[m] + o1[m];
            rt[m] = rt[m] + o2[m];
        }
    } else if constexpr (RK == RK_TVD_WAVE) {
        // reconstruct.f90:698-760: each wave of interface i, limited by its
        // projection onto the upwind neighbour's wave, added to the left
        // state and taken from the right state
#pragma unroll
        for (int mw = 0; mw < W; mw++) {
            T wnorm2 = T(0), dotl = T(0), dotr = T(0), wi[M];
#pragma unroll
            for (int m = 0; m < M; m++) {
                wi[m] = sw[(m + M * mw) * NW + kc];
                T wl = sw[(m + M * mw) * NW + kc - 1], wr = sw[(m + M * mw) * NW + kc + 1];
                wnorm2 = wnorm2 + wi[m] * wi[m];
                dotl = dotl + wl * wi[m];
                dotr = dotr + wi[m] * wr;
            }
            if (wnorm2 != T(0)) {
                T s = sw[(M * W + mw) * NW + kc];
                T r = fdiv(s > T(0) ? dotl : dotr, wnorm2);
                T uu = T(0.5) * tvd_limiter<T>(c.mthlim[mw], r);
#pragma unroll
                for (int m = 0; m < M; m++) {
                    lf[m] = lf[m] + wi[m] * uu;
                    rt[m] = rt[m] - wi[m] * uu;
                }
            }
        }
    } else if constexpr (RK == RK_TVD_CHAR) {
        // reconstruct.f90:606-680: jumps on either side projected with the
        // eigenvectors of interface i, limited per field, back with R(i)
        Eig<P, IXY, T> e;
        e.init(c, v[2], v[3]);
        T dm[M], d0[M], dp[M], hm[M], h0[M], hp[M];
#pragma unroll
        for (int m = 0; m < M; m++) {
            dm[m] = v[2][m] - v[1][m];
            d0[m] = v[3][m] - v[2][m];
            dp[m] = v[4][m] - v[3][m];
        }
        e.to_char(dm, hm);
        e.to_char(d0, h0);
        e.to_char(dp, hp);
        T U1[M], U2[M];
#pragma unroll
        for (int m = 0; m < M; m++) {
            T r1 = fabs(hm[m]) > tol14<T>() ? fdiv(h0[m], hm[m]) : T(0);
            T r2 = fabs(h0[m]) > tol14<T>() ? fdiv(hp[m], h0[m]) : T(0);
            U1[m] = T(0.5) * tvd_limiter<T>(c.mthlim[m], r1) * hm[m];
            U2[m] = T(-0.5) * tvd_limiter<T>(c.mthlim[m], r2) * h0[m];
        }
        T o1[M], o2[M];
        e.from_char(U1, o1);
        e.from_char(U2, o2);
#pragma unroll
        for (int m = 0; m < M; m++) {
            lf[m] = lf[m] + o1[m];
            rt[m] = rt[m] + o2[m];
        }
    }
}

// ------------------------------------------------------ 1D fused stage

// One chunk of NLN lines of n cells each (a 1D grid, or rows or columns of
// a 2D grid as flux2.f90 passes them to flux1) for a block of NT * NLN
// threads: interfaces i = c0..c0+NT-1 of every line (thread t = p NLN + l
// owns interface c0+p of line l) and the NT-1 cells c0..c0+NT-2.
// load(l, j, v) gives the M values of cell j = -2..n+3 of line l (ghost
// cells resolved by the caller); the cells c0-3..c0+NT+1 are staged in sq
// and interface results pass to the neighbouring thread of the line
// through sx.  Both hold line l's word k at k NLN + l, so the NLN lines of
// one staged row are adjacent (the loads of a row of NLN neighbouring
// columns are contiguous).  Each cell i <= n of the chunk is handed to
// emit(l, i, X, qc) with its flux difference X = amdq(i+1) + apdq(i) +
// amdq2(i) + apdq2(i) (flux1.f90:183-186; 150-152 with tfluct) and its
// staged values qc.  capa: a single line's capacity aux(:,mcapa) in the
// line's q layout, or NULL.  Returns this thread's max |s| (max(s dtdx(i),
// -s dtdx(i-1)) / (dt/dx) with a capacity) over interfaces 1..n+1
// (flux1.f90:126-133).  Leaves with a barrier, so the LDS arrays can be
// reused by the next chunk.  CHAR is the reconstruction code (RK_*); the
// codes with neighbour data need sw, NLN blocks of nbr_words() rows of
// NT+4 words.  sq holds M (NT+5) NLN words, sx 2M NT NLN; when sx_in_sw()
// the caller passes sx = sw, and a barrier separates the last reads of
// the neighbour data from the first writes of sx.
template <int P, int IXY, int CHAR, class T, int NT, int NLN, class Load, class Emit>
__device__ __forceinline__ T line_chunk(const PhysConst<T> &c, int c0, int n, const double *capa, T *sq,
                                        T *sx, T *sw, Load load, Emit emit)
{
    constexpr int M = Shape<P>::M, W = Shape<P>::W;
    constexpr int NC = NT - 1;
    constexpr int NL = NT + 5;
    constexpr int NB = nbr_words<P, IXY, CHAR, T>();
    auto SQ = [&](int m, int k, int l) -> T & { return sq[(m * NL + k) * NLN + l]; };
    auto SX = [&](int m, int k, int l) -> T & { return sx[(m * NT + k) * NLN + l]; };

    for (int e = threadIdx.x; e < NL * NLN; e += NT * NLN) {
        const int k = e / NLN, l = e % NLN;
        T val[M];
        load(l, min(max(c0 - 3 + k, -2), n + 3), val);
#pragma unroll
        for (int m = 0; m < M; m++) SQ(m, k, l) = val[m];
    }
    __syncthreads();

    constexpr int NW = NT + 4;
    if constexpr (NB > 0) {
        // interfaces c0-2..c0+NT+1: interface c0-2+k has cells k, k+1
        for (int e = threadIdx.x; e < NW * NLN; e += NT * NLN) {
            const int k = e / NLN, l = e % NLN;
            T ql[M], qr[M];
#pragma unroll
            for (int m = 0; m < M; m++) {
                ql[m] = SQ(m, k, l);
                qr[m] = SQ(m, k + 1, l);
            }
            nbr_data<P, IXY, CHAR, T>(c, ql, qr, sw + l * NB * NW, NW, k);
        }
        __syncthreads();
    }

    const int t = threadIdx.x / NLN, ln = threadIdx.x % NLN;
    const int i = c0 + t;  // interface i-1/2
    T v[6][M];
#pragma unroll
    for (int k = 0; k < 6; k++)
#pragma unroll
        for (int m = 0; m < M; m++) v[k][m] = SQ(m, t + k, ln);
    // Euler and shallow water with weno5 / weno5_char (the reconstructions of
    // batches and the server) take the flux form of the stage (riemann_flux:
    // X(i) = F(i+1/2) - F(i-1/2), F = f(qr) + amdq of the interface, which is
    // the reference's amdq(i+1) + apdq(i) + amdq2(i) + apdq2(i) through Roe's
    // property); the other reconstructions keep apdq and the swapped Roe
    // solve of the in-cell problem, so their results stay the reference's
    // also where a state is unphysical (NaN at the Shu-Osher shock with
    // weno5_wave)
    constexpr bool BYFLUX = (P == EU1 || P == SW1) && (CHAR == RK_WENO || CHAR == RK_WENO_CHAR);
    // the flux form's F is float64: sx holds M NT NLN of them (aligned to 8)
    auto SXD = [&](int m, int k, int l) -> double & {
        return reinterpret_cast<double *>(sx)[(m * NT + k) * NLN + l];
    };
    T lf[M], rt[M], wave[M * W], s[W], am[M], ap[M];
    double Fo[M];
    if constexpr (CHAR == RK_WENO || CHAR == RK_WENO_CHAR)
        reconstruct_iface<P, IXY, CHAR, T>(c, v, lf, rt);
    else
        reconstruct_other<P, IXY, CHAR, T>(c, v, sw + ln * NB * NW, NW, t + 2, lf, rt);
    if constexpr (BYFLUX)
        riemann_flux<P, IXY, T>(c, lf, rt, wave, s, am, Fo);
    else
        riemann<P, IXY, T>(c, lf, rt, wave, s, am, ap);
    T smax = T(0);
    if constexpr (sx_in_sw<P, IXY, CHAR, T, NT>()) __syncthreads();  // sw is read: sx reuses it
    if (i <= n + 1) {
        if (capa) {
            // flux1.f90:128-132 with dtdx(i) = dt/(dx*capa(i)): s > 0 is
            // scaled by the right cell's capacity, s < 0 by the left one's
            const T kr = T(1) / T(capa[i + 2]), kl = T(1) / T(capa[i + 1]);
#pragma unroll
            for (int w = 0; w < W; w++) smax = fmax(smax, fmax(s[w] * kr, -(s[w] * kl)));
        } else {
#pragma unroll
            for (int w = 0; w < W; w++) smax = fmax(smax, fabs(s[w]));
        }
    }
#pragma unroll
    for (int m = 0; m < M; m++) {
        if constexpr (BYFLUX) {
            Fo[m] = Fo[m] + (double)am[m];
            SXD(m, t, ln) = Fo[m];
        } else {
            SX(m, t, ln) = lf[m];
            SX(M + m, t, ln) = am[m];
        }
    }
    __syncthreads();

    if (t < NC && i <= n) {
        // cell i between interfaces i-1/2 (this thread) and i+1/2
        double X[M];
        T qc[M];
        if constexpr (BYFLUX) {
            // the same with tfluct_solver = 1 (flux1.f90:150-152): the
            // shallow-water tfluct (tfluct1sw.f) is the in-cell problem's Roe
            // wave total, f(qr) - f(ql)
#pragma unroll
            for (int m = 0; m < M; m++) {
                X[m] = SXD(m, t + 1, ln) - Fo[m];
                qc[m] = SQ(m, t + 3, ln);
            }
        } else {
            // in-cell problem of cell i: left ql(i) = rt, right qr(i) from interface i+1/2
            T qrc[M], am2[M], ap2[M];
#pragma unroll
            for (int m = 0; m < M; m++) qrc[m] = SX(m, t + 1, ln);
            riemann<P, IXY, T>(c, rt, qrc, wave, s, am2, ap2);
#pragma unroll
            for (int m = 0; m < M; m++) {
                // flux1.f90:150-152 (tfluct: the in-cell problem's total
                // fluctuation adq = amdq + apdq, tfluct*.f of the applications)
                // or flux1.f90:183-186 (the same problem through rp, swapped)
                const T amn = SX(M + m, t + 1, ln);
                X[m] = c.tfluct ? (double)((ap[m] + (am2[m] + ap2[m])) + amn)
                                : (double)(((amn + ap[m]) + am2[m]) + ap2[m]);
                qc[m] = SQ(m, t + 3, ln);
            }
        }
        emit(ln, i, X, qc);
    }
    __syncthreads();
    return smax;
}

// One chunk of a 1D stage: line_chunk on the grid a.qin with the boundary
// ghosts of bc1.f90, updated stage values written to a.qout.  last: the
// last cell of the chunk to update (and last + 1 the last interface of its
// CFL maximum), at most a.nx; cells past it are left alone, and the values
// it reads stay within cells last + 3 of a.qin.
template <int P, int IXY, int CHAR, class T, int NT>
__device__ __forceinline__ T stage1d_chunk(const StageArgs &a, const PhysConst<T> &c, int c0,
                                           T (*sq)[NT + 5], T (*sx)[NT], T *sw = nullptr,
                                           int last = 0x7fffffff)
{
    constexpr int M = Shape<P>::M;
    constexpr int MN = IXY == 1 ? 1 : 2;  // normal momentum for a wall
    const int n = a.nx;
    auto load = [&](int, int j, T (&v)[M]) {
        bool neg;
        const int js = bc_src(j, n, a.bc[0], a.bc[1], neg);
#pragma unroll
        for (int m = 0; m < M; m++) {
            double val = a.qin[(size_t)(js + 2) + (size_t)m * a.LX];
            if (m == MN && neg) val = -val;
            v[m] = T(val);
        }
    };
    auto emit = [&](int, int i, const double (&X)[M], const T (&qc)[M]) {
        const bool rq = op_reads_qold(a.op);
        // dtdx(i) = dt/(dx*capa(i)) with a capacity (flux1.f90:56-61)
        const double dtdx = a.capa ? a.dtdx / a.capa[i + 2] : a.dtdx;
        // src(q,dq,...) before flux1 (step1d.f90): dq starts from the source
        double sv[M];
#pragma unroll
        for (int m = 0; m < M; m++) sv[m] = 0.0;
        if constexpr (P == SW1) {
            if (a.src == 1) {
                // apps/shallowWater/2dRadialDam/1dRadialSymm-qwave/src1.f90:24-28
                const double h = a.qin[(size_t)(i + 2)], hu = a.qin[(size_t)(i + 2) + a.LX];
                const double rad = a.xlo + (i - 0.5) * a.dxc;
                sv[0] = -1.0 * a.dt * hu / rad;
                sv[1] = -1.0 * a.dt * (hu * hu) / h / rad;
            }
        }
#pragma unroll
        for (int m = 0; m < M; m++) {
            double r = sv[m] - dtdx * X[m];
            size_t g = (size_t)(i + 2) + (size_t)m * a.LX;
            // float64 chunks hold q itself
            double q = std::is_same<T, double>::value ? (double)qc[m] : a.qin[g];
            double qo = rq ? a.qold[g] : 0.0;
            double qbi = a.op == OP_SSP104_10 ? a.qbin[g] : 0.0;
            double out, qbo = 0.0;
            combine(a.op, qo, q, r, qbi, out, qbo);
            a.qout[g] = out;
            if (a.op == OP_SSP104_5) a.qbout[g] = qbo;
        }
    };
    return line_chunk<P, IXY, CHAR, T, NT, 1>(c, c0, min(n, last), a.capa, &sq[0][0], &sx[0][0], sw, load, emit);
}

// ------------------------------------------------- 2D stage by slices

// The 2D stage as the reference's flux2 forms it (flux2.f90:36-92): x sweeps
// over the rows j = 0..ny+1, then y sweeps over the columns i = 0..nx+1,
// each a 1D line through line_chunk.  This is the path of the
// reconstructions whose interfaces need data of their neighbouring
// interfaces (weno5_wave, weno5_trans, tvd2_wave, tvd2_char), which the
// tile kernel does not stage.  The block takes the chunk starting at cell
// c0 of the NLN slices sl0..sl0+NLN-1 (slices past the last, n_other + 1,
// repeat its loads and write nothing).  The x sweep (IXY = 1) writes
// dqx = -(dtdx X) of the interior cells; the y sweep adds -(dtdy Y) and
// combines the stage.  Ghost cells follow bc2.f90: x ghosts of every row
// first, then the y ghosts of every column, so a corner cell takes both
// maps (and both negations with walls).  The ghost slices j = 0, ny+1 and
// i = 0, nx+1 feed only the CFL, as in the reference.
template <int P, int IXY, int CHAR, class T, int NT, int NLN>
__device__ __forceinline__ T sweep2d_chunk(const StageArgs &a, const PhysConst<T> &c, double *dqx, int c0,
                                           int sl0, T *sq, T *sx, T *sw)
{
    constexpr int M = Shape<P>::M;
    // slices: rows j (x sweeps) or columns i (y sweeps)
    const int n = IXY == 1 ? a.nx : a.ny;
    const int nother = IXY == 1 ? a.ny : a.nx;
    const size_t plane = (size_t)a.LX * (size_t)a.LY;
    auto load = [&](int l, int k, T (&v)[M]) {
        const int sl = min(sl0 + l, nother + 1);
        const int i = IXY == 1 ? k : sl, j = IXY == 1 ? sl : k;
        bool negx, negy;
        const int is = bc_src(i, a.nx, a.bc[0], a.bc[1], negx);
        const int js = bc_src(j, a.ny, a.bc[2], a.bc[3], negy);
        const size_t g = (size_t)(is + 2) + (size_t)a.LX * (size_t)(js + 2);
#pragma unroll
        for (int m = 0; m < M; m++) {
            double val = a.qin[g + (size_t)m * plane];
            if ((m == 1 && negx) || (m == 2 && negy)) val = -val;
            v[m] = T(val);
        }
    };
    auto emit = [&](int l, int k, const double (&X)[M], const T (&qc)[M]) {
        const int sl = sl0 + l;
        if (sl < 1 || sl > nother) return;
        const int i = IXY == 1 ? k : sl, j = IXY == 1 ? sl : k;
        const size_t g = (size_t)(i + 2) + (size_t)a.LX * (size_t)(j + 2);
        if constexpr (IXY == 1) {
#pragma unroll
            for (int m = 0; m < M; m++) dqx[g + (size_t)m * plane] = -(a.dtdx * X[m]);
        } else {
            const bool rq = op_reads_qold(a.op);
#pragma unroll
            for (int m = 0; m < M; m++) {
                const size_t gm = g + (size_t)m * plane;
                double r = dqx[gm] + -(a.dtdy * X[m]);
                double q = std::is_same<T, double>::value ? (double)qc[m] : a.qin[gm];
                double qo = rq ? a.qold[gm] : 0.0;
                double qbi = a.op == OP_SSP104_10 ? a.qbin[gm] : 0.0;
                double out, qbo = 0.0;
                combine(a.op, qo, q, r, qbi, out, qbo);
                a.qout[gm] = out;
                if (a.op == OP_SSP104_5) a.qbout[gm] = qbo;
            }
        }
    };
    return line_chunk<P, IXY, CHAR, T, NT, NLN>(c, c0, n, nullptr, sq, sx, sw, load, emit);
}

}  // namespace sc
