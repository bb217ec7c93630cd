// WENO5 reconstruction on the device (src/reconstruct.f90:77-241).
//
// For an interface i-1/2 the stencil is the six cells v[k] = q(i-3+k),
// k = 0..5, with jumps d[k] = v[k+1] - v[k] (d[k] = dq(i-2+k) of the
// reference).  weno5 / weno5_char produce both sides of that interface:
//   lf = qr(i-1)  (m1 = 1 in the reference, jumps d0..d3)
//   rt = ql(i)    (m1 = 2, jumps d4..d1 mirrored)
// The characteristic version projects the jumps with the left eigenvectors
// at the same interface, reconstructs each field and adds R * correction to
// the 4th-order central value (reconstruct.f90:171-237).
#pragma once
#include "sc_physics.h"

namespace sc {

// epweno = 1.e-36 is a default-real literal (reconstruct.f90:6)
#define SC_EPWENO ((double)1.0e-36f)

// One side of reconstruct.f90:113-132 in float64: the nonlinear correction
// (without the central part) from the jumps at offsets intwo, inone, 0, ione.
__device__ __forceinline__ double weno_side(double im, double a2, double a1, double a0, double am1)
{
    double t1 = im * (a2 - a1);
    double t2 = im * (a1 - a0);
    double t3 = im * (a0 - am1);
    double tt1 = 13.0 * (t1 * t1) + 3.0 * dsq(a2 - 3.0 * a1);
    double tt2 = 13.0 * (t2 * t2) + 3.0 * dsq(a1 + a0);
    double tt3 = 13.0 * (t3 * t3) + 3.0 * dsq(3.0 * a0 - am1);
    tt1 = dsq(SC_EPWENO + tt1);
    tt2 = dsq(SC_EPWENO + tt2);
    tt3 = dsq(SC_EPWENO + tt3);
    double s1 = tt2 * tt3;
    double s2 = 6.0 * tt1 * tt3;
    double s3 = 3.0 * tt1 * tt2;
    double t0 = frcp(s1 + s2 + s3);
    s1 = s1 * t0;
    s3 = s3 * t0;
    return (s1 * (t2 - t1) + (0.5 * s3 - 0.25) * (t3 - t2)) * (1.0 / 3.0);
}

// The weights and correction of weno_side from the squared differences
// p1..p3 (= t1^2..t3^2), the second terms b1..b3 of the indicators and the
// differences f21 = t2 - t1, f32 = t3 - t2 (f15 = 1.5 f32, f12 = f32 / 12).
// The reference's tt_k = (eps + 13 p_k + 3 b_k^2)^2 enter only through
// ratios of products of two of them, so the common factor 9 is divided out:
// g_k = tt_k / 9 = (eps/3 + (13/3) p_k + b_k^2)^2 (eps/3 enters through the
// fma that squares b_k).  With u = g1 g2 and
// D = g2 g3 + 6 g1 g3 + 3 u (the reference's s1 + s2 + s3 over 81) the
// correction (s1 (t2-t1) + (s3/2 - 1/4)(t3-t2)) / 3 over (s1+s2+s3) is
// (g2 g3 f21 + 1.5 u f32) / (3 D) - f32 / 12.  EXACT = false takes the
// hardware estimate of 1/D (relative error below 2^-24) without refinement:
// enough for a correction rounded to float32.
template <bool EXACT = true>
__device__ __forceinline__ double weno_w(double p1, double p2, double p3, double b1, double b2,
                                         double b3, double f21, double f15, double f12)
{
    constexpr double c13 = 13.0 / 3.0, e3 = SC_EPWENO / 3.0;
    const double g1 = dsq(fma(c13, p1, fma(b1, b1, e3)));
    const double g2 = dsq(fma(c13, p2, fma(b2, b2, e3)));
    const double g3 = dsq(fma(c13, p3, fma(b3, b3, e3)));
    const double s1 = g2 * g3;
    const double u = g1 * g2;
    const double d = fma(3.0, u, fma(6.0 * g1, g3, s1));
    const double t0 = EXACT ? frcp(d) : __builtin_amdgcn_rcp(d);
    const double num = fma(u, f15, s1 * f21);
    return fma(num * t0, 1.0 / 3.0, -f12);
}

// Both sides of one field in float64 from its five jumps h0..h4 at the
// interface: the left side (m1 = 1) is weno_side(1, h0, h1, h2, h3), the
// right side (m1 = 2) weno_side(-1, h4, h3, h2, h1).  With e_k = h_k -
// h_(k+1) the left side's t1, t2, t3 are e0, e1, e2 and the right side's
// e3, e2, e1 (its im = -1 flips each sign), so the two share e1, e2, their
// squares and t3 - t2.
template <bool EXACT = true>
__device__ __forceinline__ void weno_pair(double h0, double h1, double h2, double h3, double h4,
                                          double &cl, double &cr)
{
    const double e0 = h0 - h1, e1 = h1 - h2, e2 = h2 - h3, e3 = h3 - h4;
    const double p0 = e0 * e0, p1 = e1 * e1, p2 = e2 * e2, p3 = e3 * e3;
    const double f32 = e2 - e1;
    const double f15 = 1.5 * f32, f12 = f32 * (1.0 / 12.0);
    cl = weno_w<EXACT>(p0, p1, p2, h0 - 3.0 * h1, h1 + h2, 3.0 * h2 - h3, e1 - e0, f15, f12);
    cr = weno_w<EXACT>(p3, p2, p1, h4 - 3.0 * h3, h3 + h2, 3.0 * h2 - h1, e2 - e3, -f15, -f12);
}

// Both sides for float32 jumps.  gfx950 issues unpacked float32 and float64
// VALU instructions at the same rate, so the float32 path forms its weights
// in float64 from the float32 jumps: float32 weights would need either the
// range of float64 (the products of squared indicators grow as jump^8 and
// leave float32's range for jumps outside ~[1e-4, 1e4]) or a ratio form with
// five reciprocals (transcendental, quarter rate) in place of one.  The
// correction is rounded to float32 once, so 1/D needs no refinement.
__device__ __forceinline__ void weno_pair(float h0, float h1, float h2, float h3, float h4, float &cl,
                                          float &cr)
{
    double l, r;
    weno_pair<false>((double)h0, (double)h1, (double)h2, (double)h3, (double)h4, l, r);
    cl = (float)l;
    cr = (float)r;
}

template <class T> __device__ __forceinline__ T weno_corr(T im, T a2, T a1, T a0, T am1);
template <> __device__ __forceinline__ double weno_corr<double>(double im, double a2, double a1,
                                                                double a0, double am1)
{
    return weno_side(im, a2, a1, a0, am1);
}
template <> __device__ __forceinline__ float weno_corr<float>(float im, float a2, float a1, float a0,
                                                              float am1)
{
    return (float)weno_side((double)im, (double)a2, (double)a1, (double)a0, (double)am1);
}

// The fourth-order central value (-q(i-2) + 7(q(i-1) + q(i)) - q(i+1)) / 12
// both sides of interface i-1/2 start from (reconstruct.f90:171-176)
template <class T> __device__ __forceinline__ T weno_central(T a, T b, T c, T d)
{
    return (-a + T(7) * (b + c) - d) * (T(1) / T(12));
}

// Both sides of interface i-1/2 from the six cells v[k][m] = q(i-3+k, m).
template <int P, int IXY, int CHAR, class T>
__device__ __forceinline__ void reconstruct_iface(const PhysConst<T> &c, const T (&v)[6][Shape<P>::M],
                                                  T (&lf)[Shape<P>::M], T (&rt)[Shape<P>::M])
{
    constexpr int M = Shape<P>::M;
    T d[5][M];
#pragma unroll
    for (int k = 0; k < 5; k++)
#pragma unroll
        for (int m = 0; m < M; m++) d[k][m] = v[k + 1][m] - v[k][m];
#pragma unroll
    for (int m = 0; m < M; m++) {
        T central = weno_central(v[1][m], v[2][m], v[3][m], v[4][m]);
        lf[m] = central;
        rt[m] = central;
    }
    if constexpr (CHAR == 0 || P == ADV1 || P == ADV2) {
        // component-wise (reconstruct.f90:94-140); with the advection
        // eigenvectors L = R = 1 the characteristic form is the same
#pragma unroll
        for (int m = 0; m < M; m++) {
            T cl, cr;
            weno_pair(d[0][m], d[1][m], d[2][m], d[3][m], d[4][m], cl, cr);
            lf[m] = cl + lf[m];
            rt[m] = cr + rt[m];
        }
    } else {
        // characteristic-wise (reconstruct.f90:178-237): project the five
        // jumps with L of this interface, reconstruct each field on both
        // sides, return with R
        Eig<P, IXY, T> e;
        e.init(c, v[2], v[3]);
        T hh[M][5];
#pragma unroll
        for (int k = 0; k < 5; k++) {
            T w[M];
            e.to_char(d[k], w);
#pragma unroll
            for (int p = 0; p < M; p++) hh[p][k] = w[p];
        }
        T cl[M], cr[M], ol[M], orr[M];
#pragma unroll
        for (int p = 0; p < M; p++)
            weno_pair(hh[p][0], hh[p][1], hh[p][2], hh[p][3], hh[p][4], cl[p], cr[p]);
        e.from_char(cl, ol);
        e.from_char(cr, orr);
#pragma unroll
        for (int m = 0; m < M; m++) {
            lf[m] = lf[m] + ol[m];
            rt[m] = rt[m] + orr[m];
        }
    }
}

}  // namespace sc
