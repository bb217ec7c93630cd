// Device physics for gfx950: the Riemann-solver plugins (rp) and the
// eigenvector plugins (evec) of the reference, one interface at a time.
//
// A physics P fixes meqn (M) and mwaves (W).  rp<P,IXY>(c, lf, rt, ...)
// solves the Riemann problem whose left state is qr(i-1,:) and right state
// ql(i,:) (the reference's i-th problem, e.g. rp1eu.f:19-20); IXY selects the
// normal momentum for the 2D solvers (rpn2sw.f:60-66).  The arithmetic
// follows each reference solver's expressions, with shared reciprocals
// (sc_fastmath.h) in place of repeated divisions and with fused
// multiply-adds, so results agree with the reference to rounding, not bit
// for bit.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "sc_fastmath.h"

namespace sc {

// ADV2: 2D advection with the edge velocities in aux (apps/advection/2d/
// swirl/rpn2ad1.f); the stage kernel puts the interface's velocity in
// PhysConst::u before each Riemann solve
enum PhysId { ADV1 = 0, AC1 = 1, EU1 = 2, SW1 = 3, EU2 = 4, SW2 = 5, ADV2 = 6 };

template <int P> struct Shape;
template <> struct Shape<ADV1> { static constexpr int M = 1, W = 1, DIM = 1; };
template <> struct Shape<AC1> { static constexpr int M = 2, W = 2, DIM = 1; };
template <> struct Shape<EU1> { static constexpr int M = 3, W = 3, DIM = 1; };
template <> struct Shape<SW1> { static constexpr int M = 2, W = 2, DIM = 1; };
template <> struct Shape<EU2> { static constexpr int M = 4, W = 3, DIM = 2; };
template <> struct Shape<SW2> { static constexpr int M = 3, W = 3, DIM = 2; };
template <> struct Shape<ADV2> { static constexpr int M = 1, W = 1, DIM = 2; };

// physical constants in the compute precision
template <class T> struct PhysConst {
    T gamma, gamma1;  // /param/ (setprob.f: gamma1 = gamma - 1)
    T grav;           // /comrp/ grav
    T u;              // /comrp/ u (advection)
    T cc, zz;         // /cparam/ sound speed and impedance
    T hzz;            // 1/(2 zz)
    int efix;
    int tfluct;       // 1: in-cell total fluctuation (tfluct_solver = 1), 1D only
    int mthlim[4];    // TVD limiter per wave family (lim_type = 1)
};

template <class T> __device__ __forceinline__ T dsq(T x) { return x * x; }

// f(q) of direction IXY in float64 (defined with the in-cell total fluctuation)
template <int P, int IXY, class T>
__device__ __forceinline__ void normal_flux(const PhysConst<T> &c, const T *q, double *f);

// ---------------------------------------------------------------- rp

// apps/advection/1d/rp1.f:35-43
template <class T>
__device__ __forceinline__ void rp_adv(const PhysConst<T> &c, const T *lf, const T *rt, T *wave,
                                       T *s, T *amdq, T *apdq)
{
    T w = rt[0] - lf[0];
    wave[0] = w;
    s[0] = c.u;
    amdq[0] = fmin(c.u, T(0)) * w;
    apdq[0] = fmax(c.u, T(0)) * w;
}

// apps/acoustics/1d/homogeneous/rp1.f:50-76
template <class T>
__device__ __forceinline__ void rp_ac(const PhysConst<T> &c, const T *lf, const T *rt, T *wave,
                                      T *s, T *amdq, T *apdq)
{
    const T zz = c.zz, cc = c.cc;
    T d1 = rt[0] - lf[0], d2 = rt[1] - lf[1];
    T a1 = (-d1 + zz * d2) * c.hzz;
    T a2 = (d1 + zz * d2) * c.hzz;
    // wave(m, mw) stored at wave[m + M*mw]
    wave[0] = -(a1 * zz);
    wave[1] = a1;
    wave[2] = a2 * zz;
    wave[3] = a2;
    s[0] = -cc;
    s[1] = cc;
    amdq[0] = s[0] * wave[0];
    amdq[1] = s[0] * wave[1];
    apdq[0] = s[1] * wave[2];
    apdq[1] = s[1] * wave[3];
}

template <int M, int W, class T>
__device__ __forceinline__ void split_by_sign(const T *wave, const T *s, T *amdq, T *apdq)
{
#pragma unroll
    for (int m = 0; m < M; m++) {
        T am = T(0), ap = T(0);
#pragma unroll
        for (int w = 0; w < W; w++) {
            if (s[w] < T(0))
                am = am + s[w] * wave[m + M * w];
            else
                ap = ap + s[w] * wave[m + M * w];
        }
        amdq[m] = am;
        apdq[m] = ap;
    }
}

// amdq alone of split_by_sign (the flux form of the stage needs no apdq)
template <int M, int W, class T>
__device__ __forceinline__ void split_minus(const T *wave, const T *s, T *amdq)
{
#pragma unroll
    for (int m = 0; m < M; m++) {
        T am = T(0);
#pragma unroll
        for (int w = 0; w < W; w++)
            if (s[w] < T(0)) am = am + s[w] * wave[m + M * w];
        amdq[m] = am;
    }
}

template <int M, int W, class T>
__device__ __forceinline__ void apdq_total(const T *wave, const T *s, const T *amdq, T *apdq)
{
#pragma unroll
    for (int m = 0; m < M; m++) {
        T df = T(0);
#pragma unroll
        for (int w = 0; w < W; w++) df = df + s[w] * wave[m + M * w];
        apdq[m] = df - amdq[m];
    }
}

// The entropy fix of rp1eu.f:126-201 for a Roe solve of Euler (1D: MV < 0;
// 2D: the tangential momentum MV rides along, as in rp_eu2).  wave(m, w) is
// wave[m + M w]; irl, irr are 1/rho and pl, pr the pressures of the left and
// right states.  Its branches test the signs of u - c for the left state and
// for the state behind wave 1 (rp1eu.f:135-151), and of u + c for the state
// ahead of wave 3 and the right state (rp1eu.f:167-185).  Those signs are
// decided here from squares: u - c >= 0 iff c^2 >= 0 (the square root is a
// number), u >= 0 and u^2 >= c^2; for an intermediate state (rho1, m1, E1)
// with P1 = p1 rho1 = (gamma-1)(E1 rho1 - m1^2/2), u1 - c1 > 0 iff P1 >= 0,
// m1 rho1 > 0 and m1^2 > gamma P1 (u1 = m1/rho1, c1 = sqrt(gamma P1)/|rho1|),
// so neither the square roots nor the reciprocals of the intermediate states
// are formed unless the wave is a transonic rarefaction, where the
// reference's expressions give the speeds and the split fraction.  The
// decisions equal the reference's except where u - c is within rounding of
// zero, where both branches give the same amdq to rounding.
template <int M, int MU, int MV, int ME, class T>
__device__ __forceinline__ void euler_efix(const PhysConst<T> &c, const T *lf, const T *rt, T irl, T irr, T pl,
                                           T pr, const T *wave, const T *s, T *amdq)
{
    const T g = c.gamma, g1 = c.gamma1, h = T(0.5);
    // left state: s0 = u_l - c_l
    const T ul = lf[MU] * irl, xl = g * pl * irl;
    const bool s0_ge = xl >= T(0) && ul >= T(0) && ul * ul >= xl;
    const bool s0_lt = xl >= T(0) && !s0_ge;
    if (s0_ge && s[0] > T(0)) {
#pragma unroll
        for (int m = 0; m < M; m++) amdq[m] = T(0);
        return;
    }
    // the state behind wave 1: s1 = u1 - c1
    const T rho1 = lf[0] + wave[0], m1 = lf[MU] + wave[MU], e1 = lf[ME] + wave[ME];
    const T t1 = MV >= 0 ? lf[MV < 0 ? 0 : MV] + wave[MV < 0 ? 0 : MV] : T(0);
    const T q1 = m1 * m1 + t1 * t1;
    const T P1 = g1 * (e1 * rho1 - h * q1);
    const bool s1_gt = P1 >= T(0) && m1 * rho1 > T(0) && m1 * m1 > g * P1;
    T sfract;
    if (s0_lt && s1_gt) {  // transonic rarefaction in the 1-wave
        const T s0 = ul - fsqrt(xl);
        const T ir1 = frcp(rho1);
        const T p1 = g1 * (e1 - h * q1 * ir1);
        const T s1 = m1 * ir1 - fsqrt(g * p1 * ir1);
        sfract = s0 * fdiv(s1 - s[0], s1 - s0);
    } else if (s[0] < T(0)) {
        sfract = s[0];
    } else {
        sfract = T(0);
    }
#pragma unroll
    for (int m = 0; m < M; m++) amdq[m] = sfract * wave[m];
    if (s[1] >= T(0)) return;
#pragma unroll
    for (int m = 0; m < M; m++) amdq[m] = amdq[m] + s[1] * wave[M + m];
    // the right state (s3 = u_r + c_r) and the state ahead of wave 3 (s2 = u2 + c2)
    const T ur = rt[MU] * irr, xr = g * pr * irr;
    const bool s3_gt = xr >= T(0) && (ur > T(0) || xr > ur * ur);
    const T rho2 = rt[0] - wave[2 * M], m2 = rt[MU] - wave[2 * M + MU], e2 = rt[ME] - wave[2 * M + ME];
    const T t2 = MV >= 0 ? rt[MV < 0 ? 0 : MV] - wave[2 * M + (MV < 0 ? 0 : MV)] : T(0);
    const T q2 = m2 * m2 + t2 * t2;
    const T P2 = g1 * (e2 * rho2 - h * q2);
    const bool s2_lt = P2 >= T(0) && m2 * rho2 < T(0) && m2 * m2 > g * P2;
    if (s2_lt && s3_gt) {  // transonic rarefaction in the 3-wave
        const T s3 = ur + fsqrt(xr);
        const T ir2 = frcp(rho2);
        const T p2 = g1 * (e2 - h * q2 * ir2);
        const T s2 = m2 * ir2 + fsqrt(g * p2 * ir2);
        sfract = s2 * fdiv(s3 - s[2], s3 - s2);
    } else if (s[2] < T(0)) {
        sfract = s[2];
    } else {
        return;
    }
#pragma unroll
    for (int m = 0; m < M; m++) amdq[m] = amdq[m] + sfract * wave[2 * M + m];
}

// apps/euler/shockentropy/rp1eu.f:44-213 (efix: rp1eu.f:126-201)
// With AP = false only amdq is formed, and flf (when not NULL) receives the
// normal flux of the left state in float64 (the flux form of the stage,
// stage1d_chunk / stage2d_tile).
template <class T, bool AP = true>
__device__ __forceinline__ void rp_eu1(const PhysConst<T> &c, const T *lf, const T *rt, T *wave,
                                       T *s, T *amdq, T *apdq, double *flf = nullptr)
{
    const T g1 = c.gamma1, h = T(0.5);
    T rl = lf[0], ml = lf[1], el = lf[2];
    T rr = rt[0], mr = rt[1], er = rt[2];
    T isl = frsq(rl), isr = frsq(rr);  // 1/sqrt(rho)
    T irl = isl * isl, irr = isr * isr;  // 1/rho
    T pl = g1 * (el - h * (ml * ml) * irl);
    T pr = g1 * (er - h * (mr * mr) * irr);
    T irs = frcp(rl * isl + rr * isr);  // 1/(sqrt(rhol) + sqrt(rhor))
    T u = (ml * isl + mr * isr) * irs;
    T enth = ((el + pl) * isl + (er + pr) * isr) * irs;
    T a2 = g1 * (enth - h * (u * u));
    T ia = frsq(a2);
    T a = a2 * ia;
    T d1 = rr - rl, d2 = mr - ml, d3 = er - el;
    T b2 = g1 * (ia * ia) * ((enth - u * u) * d1 + u * d2 - d3);
    T b3 = (d2 + (a - u) * d1 - a * b2) * (h * ia);
    T b1 = d1 - b2 - b3;
    wave[0] = b1;
    wave[1] = b1 * (u - a);
    wave[2] = b1 * (enth - u * a);
    s[0] = u - a;
    wave[3] = b2;
    wave[4] = b2 * u;
    wave[5] = b2 * h * (u * u);
    s[1] = u;
    wave[6] = b3;
    wave[7] = b3 * (u + a);
    wave[8] = b3 * (enth + u * a);
    s[2] = u + a;
    if (!c.efix) {
        if constexpr (AP)
            split_by_sign<3, 3>(wave, s, amdq, apdq);
        else
            split_minus<3, 3>(wave, s, amdq);
    } else {
        euler_efix<3, 1, -1, 2>(c, lf, rt, irl, irr, pl, pr, wave, s, amdq);
        if constexpr (AP) apdq_total<3, 3>(wave, s, amdq, apdq);
    }
    if (flf) {
        if constexpr (std::is_same<T, double>::value) {
            const double un = ml * irl;
            flf[0] = ml;
            flf[1] = fma(ml, un, pl);
            flf[2] = (el + pl) * un;
        } else {
            normal_flux<EU1, 1, T>(c, lf, flf);
        }
    }
}

// Roe solver for 2D Euler: normal/tangential selection of rpn2sw.f:60-66,
// Roe averages, wave strengths and entropy fix of rp1eu.f extended by the
// tangential momentum; the entropy and shear jumps share wave 2 (speed u).
template <int IXY, class T, bool AP = true>
__device__ __forceinline__ void rp_eu2(const PhysConst<T> &c, const T *lf, const T *rt, T *wave,
                                       T *s, T *amdq, T *apdq, double *flf = nullptr)
{
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    const T g1 = c.gamma1, h = T(0.5);
    T rl = lf[0], ul = lf[mu], vl = lf[mv], el = lf[3];
    T rr = rt[0], ur = rt[mu], vr = rt[mv], er = rt[3];
    T isl = frsq(rl), isr = frsq(rr);  // 1/sqrt(rho)
    T irl = isl * isl, irr = isr * isr;
    T pl = g1 * (el - h * (ul * ul + vl * vl) * irl);
    T pr = g1 * (er - h * (ur * ur + vr * vr) * irr);
    T irs = frcp(rl * isl + rr * isr);
    T u = (ul * isl + ur * isr) * irs;
    T v = (vl * isl + vr * isr) * irs;
    T enth = ((el + pl) * isl + (er + pr) * isr) * irs;
    T u2v2 = u * u + v * v;
    T a2 = g1 * (enth - h * u2v2);
    T ia = frsq(a2);
    T a = a2 * ia;
    T d1 = rr - rl, d2 = ur - ul, d3 = vr - vl, d4 = er - el;
    T a3 = d3 - v * d1;
    T a2s = g1 * (ia * ia) * ((enth - u2v2) * d1 + u * d2 + v * d3 - d4);
    T a4 = (d2 + (a - u) * d1 - a * a2s) * (h * ia);
    T a1 = d1 - a2s - a4;
    // wave(m, w) at wave[m + 4*w]
    wave[0] = a1;
    wave[mu] = a1 * (u - a);
    wave[mv] = a1 * v;
    wave[3] = a1 * (enth - u * a);
    s[0] = u - a;
    wave[4] = a2s;
    wave[4 + mu] = a2s * u;
    wave[4 + mv] = a2s * v + a3;
    wave[7] = a2s * h * u2v2 + a3 * v;
    s[1] = u;
    wave[8] = a4;
    wave[8 + mu] = a4 * (u + a);
    wave[8 + mv] = a4 * v;
    wave[11] = a4 * (enth + u * a);
    s[2] = u + a;
    if (!c.efix) {
        if constexpr (AP)
            split_by_sign<4, 3>(wave, s, amdq, apdq);
        else
            split_minus<4, 3>(wave, s, amdq);
    } else {
        euler_efix<4, mu, mv, 3>(c, lf, rt, irl, irr, pl, pr, wave, s, amdq);
        if constexpr (AP) apdq_total<4, 3>(wave, s, amdq, apdq);
    }
    if (flf) {
        if constexpr (std::is_same<T, double>::value) {
            const double un = ul * irl;
            flf[0] = ul;
            flf[mu] = fma(ul, un, pl);
            flf[mv] = vl * un;
            flf[3] = (el + pl) * un;
        } else {
            normal_flux<EU2, IXY, T>(c, lf, flf);
        }
    }
}

// apps/shallowWater/2dRadialDam/1dRadialSymm-qwave/rp1sw.f:43-183.  The
// entropy fix (rp1sw.f:120-170) tests the signs of u - c for the left state
// and the state behind wave 1, and of u + c for the state ahead of wave 2
// and the right state; as in euler_efix they are decided from squares (for
// an intermediate state (h1, m1): u1 - c1 > 0 iff h1 > 0, m1 > 0 and
// m1^2 > g h1^3), and the speeds are formed only for a transonic
// rarefaction.  AP and flf as for rp_eu1.
template <class T, bool AP = true>
__device__ __forceinline__ void rp_sw1(const PhysConst<T> &c, const T *lf, const T *rt, T *wave,
                                       T *s, T *amdq, T *apdq, double *flf = nullptr)
{
    const T gr = c.grav, h = T(0.5);
    T hl = lf[0], hul = lf[1], hr = rt[0], hur = rt[1];
    T isl = frsq(hl), isr = frsq(hr);
    T ihl = isl * isl, ihr = isr * isr;
    T ubar = (hul * isl + hur * isr) * frcp(hl * isl + hr * isr);
    T cbar = fsqrt(h * gr * (hl + hr));
    T icb = frcp(cbar);
    T d1 = hr - hl, d2 = hur - hul;
    T a1 = h * (-d2 + (ubar + cbar) * d1) * icb;
    T a2 = h * (d2 - (ubar - cbar) * d1) * icb;
    wave[0] = a1;
    wave[1] = a1 * (ubar - cbar);
    s[0] = ubar - cbar;
    wave[2] = a2;
    wave[3] = a2 * (ubar + cbar);
    s[1] = ubar + cbar;
    if (!c.efix) {
        if constexpr (AP)
            split_by_sign<2, 2>(wave, s, amdq, apdq);
        else
            split_minus<2, 2>(wave, s, amdq);
    } else {
        do {
            // left state: s0 = u_l - c_l
            const T ul = hul * ihl, xl = gr * hl;
            const bool s0_ge = xl >= T(0) && ul >= T(0) && ul * ul >= xl;
            const bool s0_lt = xl >= T(0) && !s0_ge;
            if (s0_ge && s[0] > T(0)) {
                amdq[0] = amdq[1] = T(0);
                break;
            }
            // behind wave 1: s1 = u1 - c1
            const T hr1 = hl + wave[0], uhr1 = hul + wave[1];
            const bool s1_gt = hr1 > T(0) && uhr1 > T(0) && uhr1 * uhr1 > gr * (hr1 * hr1) * hr1;
            T sfract;
            if (s0_lt && s1_gt) {
                const T s0 = ul - fsqrt(xl);
                const T s1 = fdiv(uhr1, hr1) - fsqrt(gr * hr1);
                sfract = s0 * fdiv(s1 - s[0], s1 - s0);
            } else if (s[0] < T(0)) {
                sfract = s[0];
            } else {
                sfract = T(0);
            }
            amdq[0] = sfract * wave[0];
            amdq[1] = sfract * wave[1];
            // right state (s3 = u_r + c_r) and ahead of wave 2 (s2 = u2 + c2)
            const T ur = hur * ihr, xr = gr * hr;
            const bool s3_gt = xr >= T(0) && (ur > T(0) || xr > ur * ur);
            const T hl2 = hr - wave[2], uhl2 = hur - wave[3];
            const bool s2_lt = hl2 > T(0) && uhl2 < T(0) && uhl2 * uhl2 > gr * (hl2 * hl2) * hl2;
            if (s2_lt && s3_gt) {
                const T s3 = ur + fsqrt(xr);
                const T s2 = fdiv(uhl2, hl2) + fsqrt(gr * hl2);
                sfract = s2 * fdiv(s3 - s[1], s3 - s2);
            } else if (s[1] < T(0)) {
                sfract = s[1];
            } else {
                break;
            }
            amdq[0] = amdq[0] + sfract * wave[2];
            amdq[1] = amdq[1] + sfract * wave[3];
        } while (0);
        if constexpr (AP) apdq_total<2, 2>(wave, s, amdq, apdq);
    }
    if (flf) {
        if constexpr (std::is_same<T, double>::value) {
            flf[0] = hul;
            flf[1] = fma(hul, hul * ihl, 0.5 * gr * (hl * hl));
        } else {
            normal_flux<SW1, 1, T>(c, lf, flf);
        }
    }
}

// apps/shallowWater/2dRadialDam/rpn2sw.f:60-228
template <int IXY, class T, bool AP = true>
__device__ __forceinline__ void rp_sw2(const PhysConst<T> &c, const T *lf, const T *rt, T *wave,
                                       T *s, T *amdq, T *apdq, double *flf = nullptr)
{
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    const T gr = c.grav, h = T(0.5);
    T hh = (lf[0] + rt[0]) * h;
    T isl = frsq(lf[0]), isr = frsq(rt[0]);
    T irs = frcp(lf[0] * isl + rt[0] * isr);
    T u = (lf[mu] * isl + rt[mu] * isr) * irs;
    T v = (lf[mv] * isl + rt[mv] * isr) * irs;
    T ia = frsq(gr * hh);
    T a = gr * hh * ia;
    T d1 = rt[0] - lf[0], d2 = rt[mu] - lf[mu], d3 = rt[mv] - lf[mv];
    T a1 = ((u + a) * d1 - d2) * (h * ia);
    T a2 = -(v * d1) + d3;
    T a3 = (-((u - a) * d1) + d2) * (h * ia);
    wave[0] = a1;
    wave[mu] = a1 * (u - a);
    wave[mv] = a1 * v;
    s[0] = u - a;
    wave[3] = T(0);
    wave[3 + mu] = T(0);
    wave[3 + mv] = a2;
    s[1] = u;
    wave[6] = a3;
    wave[6 + mu] = a3 * (u + a);
    wave[6 + mv] = a3 * v;
    s[2] = u + a;
    if (!c.efix) {
        if constexpr (AP)
            split_by_sign<3, 3>(wave, s, amdq, apdq);
        else
            split_minus<3, 3>(wave, s, amdq);
    } else {
    do {
        T him1 = lf[0];
        T s0 = fdiv(lf[mu], him1) - fsqrt(gr * him1);
        if (s0 > T(0) && s[0] > T(0)) {
            amdq[0] = amdq[1] = amdq[2] = T(0);
            break;
        }
        T h1 = lf[0] + wave[0], hu1 = lf[mu] + wave[mu];
        T s1 = fdiv(hu1, h1) - fsqrt(gr * h1);
        T sfract;
        if (s0 < T(0) && s1 > T(0))
            sfract = s0 * fdiv(s1 - s[0], s1 - s0);
        else if (s[0] < T(0))
            sfract = s[0];
        else
            sfract = T(0);
#pragma unroll
        for (int m = 0; m < 3; m++) amdq[m] = sfract * wave[m];
        if (s[1] > T(0)) break;
#pragma unroll
        for (int m = 0; m < 3; m++) amdq[m] = amdq[m] + s[1] * wave[3 + m];
        T hi = rt[0];
        T s03 = fdiv(rt[mu], hi) + fsqrt(gr * hi);
        T h3 = rt[0] - wave[6], hu3 = rt[mu] - wave[6 + mu];
        T s3 = fdiv(hu3, h3) + fsqrt(gr * h3);
        if (s3 < T(0) && s03 > T(0))
            sfract = s3 * fdiv(s03 - s[2], s03 - s3);
        else if (s[2] < T(0))
            sfract = s[2];
        else
            break;
#pragma unroll
        for (int m = 0; m < 3; m++) amdq[m] = amdq[m] + sfract * wave[6 + m];
    } while (0);
    if constexpr (AP) apdq_total<3, 3>(wave, s, amdq, apdq);
    }
    if (flf) {
        if constexpr (std::is_same<T, double>::value) {
            const double ih = isl * isl, hn = lf[mu];
            const double un = hn * ih;
            flf[0] = hn;
            flf[mu] = fma(hn, un, 0.5 * gr * (lf[0] * lf[0]));
            flf[mv] = lf[mv] * un;
        } else {
            normal_flux<SW2, IXY, T>(c, lf, flf);
        }
    }
}

// ------------------------------------------- in-cell total fluctuation
//
// The update uses the in-cell problem (left ql(i), right qr(i)) only
// through amdq2(i) + apdq2(i) (flux1.f90:184-186).  For the Roe solvers of
// Euler and shallow water that sum is sum_w s_w wave_w = A_roe (qr - ql)
// = f(qr(i)) - f(ql(i)) (Roe's property; the entropy fix moves flux between
// amdq and apdq and keeps the sum: apdq = df - amdq, rp1eu.f / rpn2sw.f),
// the total fluctuation flux1.f90:146-148 names as the conservative one.
// It is formed here from the normal fluxes of the two states, in float64
// also when T is float.  The result agrees with the swapped Riemann solve to
// rounding; where a state is unphysical (negative density or pressure) the
// Roe solve gives NaN and the flux difference a finite number.

// normal flux f(q) of direction IXY in float64
template <int P, int IXY, class T>
__device__ __forceinline__ void normal_flux(const PhysConst<T> &c, const T *q, double *f)
{
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    if constexpr (P == EU2) {
        const double r = q[0], mn = q[mu], mt = q[mv], E = q[3];
        const double ir = frcp(r);
        const double un = mn * ir;
        const double p = double(c.gamma1) * (E - 0.5 * (mn * mn + mt * mt) * ir);
        f[0] = mn;
        f[mu] = mn * un + p;
        f[mv] = mt * un;
        f[3] = (E + p) * un;
    } else if constexpr (P == SW2) {
        const double h = q[0], hn = q[mu], ht = q[mv];
        const double un = hn * frcp(h);
        f[0] = hn;
        f[mu] = hn * un + 0.5 * double(c.grav) * (h * h);
        f[mv] = ht * un;
    } else if constexpr (P == EU1) {
        const double r = q[0], m = q[1], E = q[2];
        const double ir = frcp(r);
        const double u = m * ir;
        const double p = double(c.gamma1) * (E - 0.5 * (m * m) * ir);
        f[0] = m;
        f[1] = m * u + p;
        f[2] = (E + p) * u;
    } else {
        static_assert(P == SW1, "normal_flux: Euler and shallow water only");
        const double h = q[0], hu = q[1];
        f[0] = hu;
        f[1] = hu * (hu * frcp(h)) + 0.5 * double(c.grav) * (h * h);
    }
}

// 2D physics whose stage takes the flux form (stage2d_tile; the 1D stage
// decides per reconstruction, line_chunk).  With the in-cell total
// fluctuation f(qr(i)) - f(ql(i)) and, at interface i-1/2, apdq(i) =
// f(ql(i)) - f(qr(i-1)) - amdq(i) (Roe's property), the reference's
// X(i) = amdq(i+1) + apdq(i) + amdq2(i) + apdq2(i) (flux1.f90:183-186) is
// F(i+1/2) - F(i-1/2) with the interface flux F(i-1/2) = f(qr(i-1)) +
// amdq(i): one normal flux and amdq per interface, no apdq and no in-cell
// problem.  It agrees with the reference to rounding.
template <int P> constexpr bool incell_by_flux() { return P == EU2 || P == SW2; }

// adq = amdq2 + apdq2 of the in-cell problem (left ql, right qr) as
// f(qr) - f(ql)
template <int P, int IXY, class T>
__device__ __forceinline__ void incell_flux_diff(const PhysConst<T> &c, const T *ql, const T *qr,
                                                 double *adq)
{
    constexpr int M = Shape<P>::M;
    double fl[M], fr[M];
    normal_flux<P, IXY, T>(c, ql, fl);
    normal_flux<P, IXY, T>(c, qr, fr);
#pragma unroll
    for (int m = 0; m < M; m++) adq[m] = fr[m] - fl[m];
}

template <int P, int IXY, class T>
__device__ __forceinline__ void riemann(const PhysConst<T> &c, const T *lf, const T *rt, T *wave,
                                        T *s, T *amdq, T *apdq)
{
    // rpn2ad1.f:45-50 is rp1.f's formula with the speed auxl(i,ixy)
    if constexpr (P == ADV1 || P == ADV2) rp_adv(c, lf, rt, wave, s, amdq, apdq);
    else if constexpr (P == AC1) rp_ac(c, lf, rt, wave, s, amdq, apdq);
    else if constexpr (P == EU1) rp_eu1(c, lf, rt, wave, s, amdq, apdq);
    else if constexpr (P == SW1) rp_sw1(c, lf, rt, wave, s, amdq, apdq);
    else if constexpr (P == EU2) rp_eu2<IXY>(c, lf, rt, wave, s, amdq, apdq);
    else rp_sw2<IXY>(c, lf, rt, wave, s, amdq, apdq);
}

// The Riemann solve of the flux form of a stage (Euler and shallow water,
// incell_by_flux): wave speeds s, amdq, and the float64 normal flux flf of
// the left state; the numerical flux of the interface is flf + amdq.
template <int P, int IXY, class T>
__device__ __forceinline__ void riemann_flux(const PhysConst<T> &c, const T *lf, const T *rt, T *wave, T *s,
                                             T *amdq, double *flf)
{
    if constexpr (P == EU1) rp_eu1<T, false>(c, lf, rt, wave, s, amdq, nullptr, flf);
    else if constexpr (P == SW1) rp_sw1<T, false>(c, lf, rt, wave, s, amdq, nullptr, flf);
    else if constexpr (P == EU2) rp_eu2<IXY, T, false>(c, lf, rt, wave, s, amdq, nullptr, flf);
    else {
        static_assert(P == SW2, "riemann_flux: Euler and shallow water only");
        rp_sw2<IXY, T, false>(c, lf, rt, wave, s, amdq, nullptr, flf);
    }
}

// The largest |s| in x and in y of the Roe waves at an interface whose two
// states are both q (every interface of a constant tile): the speeds are
// u - c, u, u + c (v - c, v, v + c), so |u| + c and |v| + c.  2D Euler and
// shallow water.
template <int P, class T>
__device__ __forceinline__ void flat_speeds(const PhysConst<T> &c, const T *q, T &sx, T &sy)
{
    const T ir = frcp(q[0]);
    const T u = q[1] * ir, v = q[2] * ir;
    T cs;
    if constexpr (P == EU2) {
        const T p = c.gamma1 * (q[3] - T(0.5) * (q[1] * q[1] + q[2] * q[2]) * ir);
        cs = fsqrt(c.gamma * p * ir);
    } else {
        static_assert(P == SW2, "flat_speeds: 2D Euler and shallow water");
        cs = fsqrt(c.grav * q[0]);
    }
    sx = fabs(u) + cs;
    sy = fabs(v) + cs;
}

// --------------------------------------------------------------- evec
//
// Left (evl[p][m]) and right (evr[m][p]) eigenvectors at the interface
// between cells qa = q(i-1) and qb = q(i); stored row-major as
// evl[p*M + m], evr[m*M + p].

template <int P, int IXY, class T>
__device__ __forceinline__ void eigvec(const PhysConst<T> &c, const T *qa, const T *qb, T *evl,
                                       T *evr)
{
    constexpr int M = Shape<P>::M;
    const T h = T(0.5);
    if constexpr (P == ADV1 || P == ADV2) {  // apps/advection/1d/evec.f:74-77
        evl[0] = T(1);
        evr[0] = T(1);
    } else if constexpr (P == AC1) {  // apps/acoustics/1d/homogeneous/evec.f:32-42
        const T zz = c.zz;
        evr[0 * M + 0] = -zz;
        evr[0 * M + 1] = zz;
        evr[1 * M + 0] = T(1);
        evr[1 * M + 1] = T(1);
        T z = T(1) / (zz + zz);
        evl[0 * M + 0] = -z;
        evl[0 * M + 1] = zz * z;
        evl[1 * M + 0] = z;
        evl[1 * M + 1] = zz * z;
    } else if constexpr (P == EU1) {  // R of evec1eu.f:49-57, L = R^-1
        const T g1 = c.gamma1;
        T rl = qa[0], ml = qa[1], el = qa[2], rr = qb[0], mr = qb[1], er = qb[2];
        T rhsqrtl = sqrt(rl), rhsqrtr = sqrt(rr);
        T pl = g1 * (el - h * (ml * ml) / rl);
        T pr = g1 * (er - h * (mr * mr) / rr);
        T rhsq2 = rhsqrtl + rhsqrtr;
        T u = (ml / rhsqrtl + mr / rhsqrtr) / rhsq2;
        T enth = ((el + pl) / rhsqrtl + (er + pr) / rhsqrtr) / rhsq2;
        T a2 = g1 * (enth - h * (u * u));
        T a = sqrt(a2);
        evr[0] = T(1); evr[1] = T(1); evr[2] = T(1);
        evr[3] = u - a; evr[4] = u; evr[5] = u + a;
        evr[6] = enth - u * a; evr[7] = h * (u * u); evr[8] = enth + u * a;
        T b = g1, h2 = h / a2, bk = h * b * (u * u);
        evl[0] = (bk + u * a) * h2; evl[1] = -((b * u + a) * h2); evl[2] = b * h2;
        evl[3] = T(1) - bk / a2;   evl[4] = b * u / a2;          evl[5] = -(b / a2);
        evl[6] = (bk - u * a) * h2; evl[7] = -((b * u - a) * h2); evl[8] = b * h2;
    } else if constexpr (P == EU2) {  // fields (u-c, entropy, shear, u+c)
        constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
        const T g1 = c.gamma1;
        T rl = qa[0], ul = qa[mu], vl = qa[mv], el = qa[3];
        T rr = qb[0], ur = qb[mu], vr = qb[mv], er = qb[3];
        T rhsqrtl = sqrt(rl), rhsqrtr = sqrt(rr);
        T pl = g1 * (el - h * (ul * ul + vl * vl) / rl);
        T pr = g1 * (er - h * (ur * ur + vr * vr) / rr);
        T rhsq2 = rhsqrtl + rhsqrtr;
        T u = (ul / rhsqrtl + ur / rhsqrtr) / rhsq2;
        T v = (vl / rhsqrtl + vr / rhsqrtr) / rhsq2;
        T enth = ((el + pl) / rhsqrtl + (er + pr) / rhsqrtr) / rhsq2;
        T u2v2 = u * u + v * v;
        T a2 = g1 * (enth - h * u2v2);
        T a = sqrt(a2);
        // evr[m*4 + p]
        evr[0 * 4 + 0] = T(1); evr[mu * 4 + 0] = u - a; evr[mv * 4 + 0] = v; evr[3 * 4 + 0] = enth - u * a;
        evr[0 * 4 + 1] = T(1); evr[mu * 4 + 1] = u;     evr[mv * 4 + 1] = v; evr[3 * 4 + 1] = h * u2v2;
        evr[0 * 4 + 2] = T(0); evr[mu * 4 + 2] = T(0);  evr[mv * 4 + 2] = T(1); evr[3 * 4 + 2] = v;
        evr[0 * 4 + 3] = T(1); evr[mu * 4 + 3] = u + a; evr[mv * 4 + 3] = v; evr[3 * 4 + 3] = enth + u * a;
        T b = g1, h2 = h / a2, bk = h * b * u2v2;
        evl[0 * 4 + 0] = (bk + u * a) * h2; evl[0 * 4 + mu] = -((b * u + a) * h2);
        evl[0 * 4 + mv] = -(b * v * h2);    evl[0 * 4 + 3] = b * h2;
        evl[1 * 4 + 0] = T(1) - bk / a2;    evl[1 * 4 + mu] = b * u / a2;
        evl[1 * 4 + mv] = b * v / a2;       evl[1 * 4 + 3] = -(b / a2);
        evl[2 * 4 + 0] = -v;                evl[2 * 4 + mu] = T(0);
        evl[2 * 4 + mv] = T(1);             evl[2 * 4 + 3] = T(0);
        evl[3 * 4 + 0] = (bk - u * a) * h2; evl[3 * 4 + mu] = -((b * u - a) * h2);
        evl[3 * 4 + mv] = -(b * v * h2);    evl[3 * 4 + 3] = b * h2;
    } else if constexpr (P == SW1) {  // Roe averages of rp1sw.f:47-49
        T hl = qa[0], hul = qa[1], hr = qb[0], hur = qb[1];
        T sl = sqrt(hl), sr = sqrt(hr);
        T ubar = (hul / sl + hur / sr) / (sl + sr);
        T cbar = sqrt(h * c.grav * (hl + hr));
        evr[0] = T(1); evr[1] = T(1);
        evr[2] = ubar - cbar; evr[3] = ubar + cbar;
        T z = h / cbar;
        evl[0] = (ubar + cbar) * z; evl[1] = -z;
        evl[2] = -((ubar - cbar) * z); evl[3] = z;
    } else {  // SW2, Roe averages of rpn2sw.f:80-87; fields (u-c, shear, u+c)
        constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
        T hh = (qa[0] + qb[0]) * h;
        T hsqrtl = sqrt(qa[0]), hsqrtr = sqrt(qb[0]);
        T hsq2 = hsqrtl + hsqrtr;
        T u = (qa[mu] / hsqrtl + qb[mu] / hsqrtr) / hsq2;
        T v = (qa[mv] / hsqrtl + qb[mv] / hsqrtr) / hsq2;
        T a = sqrt(c.grav * hh);
        evr[0 * 3 + 0] = T(1); evr[mu * 3 + 0] = u - a; evr[mv * 3 + 0] = v;
        evr[0 * 3 + 1] = T(0); evr[mu * 3 + 1] = T(0);  evr[mv * 3 + 1] = T(1);
        evr[0 * 3 + 2] = T(1); evr[mu * 3 + 2] = u + a; evr[mv * 3 + 2] = v;
        T z = h / a;
        evl[0 * 3 + 0] = (u + a) * z; evl[0 * 3 + mu] = -z; evl[0 * 3 + mv] = T(0);
        evl[1 * 3 + 0] = -v;          evl[1 * 3 + mu] = T(0); evl[1 * 3 + mv] = T(1);
        evl[2 * 3 + 0] = -((u - a) * z); evl[2 * 3 + mu] = z; evl[2 * 3 + mv] = T(0);
    }
}

// -------------------------------------- structured eigen-decompositions
//
// Eig<P,IXY,T> holds the Roe state of one interface and applies L (to_char)
// and R (from_char) without forming the matrices: the eigenvectors of
// eigvec() have zeros, ones and shared sub-expressions that a dense M x M
// product would multiply out.  With b = gamma-1, q2 = |u|^2/2, K = b/c^2,
// h2 = 1/(2c^2) and jumps (d0, dn, dt, dE) in (rho, normal, tangential, E):
//   Y = q2 d0 - u dn - v dt + dE,  Z = u d0 - dn
//   w(u-c) = h2 (bY + cZ),  w(entropy) = d0 - K Y,  w(shear) = dt - v d0,
//   w(u+c) = h2 (bY - cZ)
// which are the rows of L = R^-1 in eigvec() regrouped.

template <int P, int IXY, class T> struct Eig;

template <int IXY, class T> struct Eig<ADV1, IXY, T> {
    __device__ __forceinline__ void init(const PhysConst<T> &, const T *, const T *) {}
    __device__ __forceinline__ void to_char(const T *d, T *w) const { w[0] = d[0]; }
    __device__ __forceinline__ void from_char(const T *c, T *o) const { o[0] = c[0]; }
};

template <int IXY, class T> struct Eig<ADV2, IXY, T> : Eig<ADV1, IXY, T> {};

template <int IXY, class T> struct Eig<AC1, IXY, T> {
    T zz, z;
    __device__ __forceinline__ void init(const PhysConst<T> &c, const T *, const T *)
    {
        zz = c.zz;
        z = c.hzz;
    }
    __device__ __forceinline__ void to_char(const T *d, T *w) const
    {
        w[0] = z * (zz * d[1] - d[0]);
        w[1] = z * (d[0] + zz * d[1]);
    }
    __device__ __forceinline__ void from_char(const T *c, T *o) const
    {
        o[0] = zz * (c[1] - c[0]);
        o[1] = c[0] + c[1];
    }
};

// Roe average of two cell states for Euler (u, v, H, c); v = 0 in 1D
template <int MU, int MV, int ME, class T>
__device__ __forceinline__ void roe_euler(const PhysConst<T> &c, const T *qa, const T *qb, T &u, T &v,
                                          T &H, T &a2)
{
    const T g1 = c.gamma1, h = T(0.5);
    T rl = qa[0], rr = qb[0];
    T isl = frsq(rl), isr = frsq(rr);  // 1/sqrt(rho)
    T irl = isl * isl, irr = isr * isr;
    T ul = qa[MU], ur = qb[MU];
    T vl = MV >= 0 ? qa[MV] : T(0), vr = MV >= 0 ? qb[MV] : T(0);
    T pl = g1 * (qa[ME] - h * (ul * ul + vl * vl) * irl);
    T pr = g1 * (qb[ME] - h * (ur * ur + vr * vr) * irr);
    T irs = frcp(rl * isl + rr * isr);
    u = (ul * isl + ur * isr) * irs;
    v = (vl * isl + vr * isr) * irs;
    H = ((qa[ME] + pl) * isl + (qb[ME] + pr) * isr) * irs;
    a2 = g1 * (H - h * (u * u + v * v));
}

// Euler: L applied to a jump d is (with Y = q2 d0 - u d1 + d2, Z = u d0 - d1,
// K = (gamma-1)/c^2) w0 = (K Y + Z / c) / 2, w1 = d0 - K Y, w2 = (K Y - Z / c) / 2;
// hK = K/2 and hi = 1/(2c) are kept so each costs one fma.
template <int IXY, class T> struct Eig<EU1, IXY, T> {
    T u, a, H, q2, hK, hi, K;
    __device__ __forceinline__ void init(const PhysConst<T> &c, const T *qa, const T *qb)
    {
        T v, a2;
        roe_euler<1, -1, 2>(c, qa, qb, u, v, H, a2);
        T ia = frsq(a2);
        a = a2 * ia;
        K = c.gamma1 * (ia * ia);
        hK = T(0.5) * K;
        hi = T(0.5) * ia;
        q2 = T(0.5) * (u * u);
    }
    __device__ __forceinline__ void to_char(const T *d, T *w) const
    {
        T Y = q2 * d[0] - u * d[1] + d[2];
        T Z = u * d[0] - d[1];
        T kY = hK * Y;
        w[0] = fma(hi, Z, kY);
        w[1] = d[0] - K * Y;
        w[2] = fma(-hi, Z, kY);
    }
    __device__ __forceinline__ void from_char(const T *c, T *o) const
    {
        T S = c[0] + c[1] + c[2], D = c[2] - c[0];
        o[0] = S;
        o[1] = u * S + a * D;
        o[2] = H * (c[0] + c[2]) + q2 * c[1] + (u * a) * D;
    }
};

template <int IXY, class T> struct Eig<EU2, IXY, T> {
    static constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    T u, v, a, H, q2, hK, hi, K;
    __device__ __forceinline__ void init(const PhysConst<T> &c, const T *qa, const T *qb)
    {
        T a2;
        roe_euler<mu, mv, 3>(c, qa, qb, u, v, H, a2);
        T ia = frsq(a2);
        a = a2 * ia;
        K = c.gamma1 * (ia * ia);
        hK = T(0.5) * K;
        hi = T(0.5) * ia;
        q2 = T(0.5) * (u * u + v * v);
    }
    __device__ __forceinline__ void to_char(const T *d, T *w) const
    {
        T Y = q2 * d[0] - u * d[mu] - v * d[mv] + d[3];
        T Z = u * d[0] - d[mu];
        T kY = hK * Y;
        w[0] = fma(hi, Z, kY);
        w[1] = d[0] - K * Y;
        w[2] = d[mv] - v * d[0];
        w[3] = fma(-hi, Z, kY);
    }
    __device__ __forceinline__ void from_char(const T *c, T *o) const
    {
        T S = c[0] + c[1] + c[3], D = c[3] - c[0];
        o[0] = S;
        o[mu] = u * S + a * D;
        o[mv] = v * S + c[2];
        o[3] = H * (c[0] + c[3]) + q2 * c[1] + v * c[2] + (u * a) * D;
    }
};

template <int IXY, class T> struct Eig<SW1, IXY, T> {
    T u, a, z;
    __device__ __forceinline__ void init(const PhysConst<T> &c, const T *qa, const T *qb)
    {
        T isl = frsq(qa[0]), isr = frsq(qb[0]);
        u = (qa[1] * isl + qb[1] * isr) * frcp(qa[0] * isl + qb[0] * isr);
        T ia = frsq(T(0.5) * c.grav * (qa[0] + qb[0]));
        a = T(0.5) * c.grav * (qa[0] + qb[0]) * ia;
        z = T(0.5) * ia;
    }
    __device__ __forceinline__ void to_char(const T *d, T *w) const
    {
        w[0] = z * ((u + a) * d[0] - d[1]);
        w[1] = z * (d[1] - (u - a) * d[0]);
    }
    __device__ __forceinline__ void from_char(const T *c, T *o) const
    {
        T S = c[0] + c[1];
        o[0] = S;
        o[1] = u * S + a * (c[1] - c[0]);
    }
};

template <int IXY, class T> struct Eig<SW2, IXY, T> {
    static constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    T u, v, a, z;
    __device__ __forceinline__ void init(const PhysConst<T> &c, const T *qa, const T *qb)
    {
        T isl = frsq(qa[0]), isr = frsq(qb[0]);
        T irs = frcp(qa[0] * isl + qb[0] * isr);
        u = (qa[mu] * isl + qb[mu] * isr) * irs;
        v = (qa[mv] * isl + qb[mv] * isr) * irs;
        T gh = c.grav * ((qa[0] + qb[0]) * T(0.5));
        T ia = frsq(gh);
        a = gh * ia;
        z = T(0.5) * ia;
    }
    __device__ __forceinline__ void to_char(const T *d, T *w) const
    {
        w[0] = z * ((u + a) * d[0] - d[mu]);
        w[1] = d[mv] - v * d[0];
        w[2] = z * (d[mu] - (u - a) * d[0]);
    }
    __device__ __forceinline__ void from_char(const T *c, T *o) const
    {
        T S = c[0] + c[2];
        o[0] = S;
        o[mu] = u * S + a * (c[2] - c[0]);
        o[mv] = v * S + c[1];
    }
};

}  // namespace sc
