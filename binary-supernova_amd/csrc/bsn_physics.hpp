// bsn_physics.hpp -- per-interface numerics of the Euler right-hand side on
// CDNA4 (gfx950): WENO5-Z reconstruction, cell pressure, interface energy and
// the HLLC flux.
//
// Two forms:
//   * exact FP64 (namespace ex): every operation of src/weno5.jl:14-61,
//     src/hllc.jl:17-208 and src/euler3d.jl:86-106 in the reference's order,
//     compiled with -ffp-contract=off, so results are the reference's bits.
//     What differs is only *where* values are computed: the cell pressure is
//     computed once per cell (the reference recomputes it six times per
//     interface, src/gpu_kernels.jl:69-74), and within one interface the
//     smoothness indicator and candidate polynomials that the left and right
//     reconstructions share (left beta1 == right beta0, left q1 == right q0,
//     left q2 == right q1, and the first term of left beta2 == that of right
//     beta1) are evaluated once.  Each shared value is the same expression on
//     the same operands, so the bits do not change.
//   * split FP32 (namespace sp): WENO5-Z weights and the HLLC flux in FP32
//     with hardware reciprocal and square root and FMA contraction; the cell
//     pressure (the E - rho v^2/2 cancellation) stays FP64 and is rounded to
//     FP32 only after it is formed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef BSN_NG
#define BSN_NG 3
#endif

namespace bsn {

// ---------------------------------------------------------------------------
// Julia Base.max / min / clamp semantics for Float64: NaN propagates, and
// -0.0 orders below +0.0.

__device__ __forceinline__ double jl_max(double x, double y) {
    double diff = x - y;
    double am = __builtin_signbit(diff) ? y : x;
    return (__builtin_isnan(x) || __builtin_isnan(y)) ? diff : am;
}
__device__ __forceinline__ double jl_min(double x, double y) {
    double diff = x - y;
    double am = __builtin_signbit(diff) ? x : y;
    return (__builtin_isnan(x) || __builtin_isnan(y)) ? diff : am;
}
__device__ __forceinline__ double jl_clamp(double x, double lo, double hi) {
    return x > hi ? hi : (x < lo ? lo : x);
}

namespace ex {

// src/euler3d.jl:86-90
__device__ __forceinline__ double cell_pressure(double r, double mx, double my, double mz,
                                                double E, double g) {
    r = jl_max(r, 1e-30);
    double KE = (0.5 * (mx * mx + my * my + mz * mz)) / r;
    return jl_max((g - 1.0) * (E - KE), 0.0);
}

// src/euler3d.jl:100-106
__device__ __forceinline__ double EL_from_P(double PL, double rL, double mxL, double myL,
                                            double mzL, double g) {
    double r_pos = jl_max(rL, 1e-30);
    double m2 = mxL * mxL + myL * myL + mzL * mzL;
    double KE_raw = (0.5 * m2) / r_pos;
    double KE = jl_min(KE_raw, (0.5 * r_pos) * 1e4);
    return PL / (g - 1.0) + KE;
}

// src/weno5.jl:14-21
__device__ __forceinline__ double wenoz_combine(double b0, double b1, double b2,
                                                double d0, double d1, double d2,
                                                double q0, double q1, double q2) {
    const double eps = 1e-36;
    double tau5 = fabs(b0 - b2);
    double t0 = tau5 / (eps + b0);
    double t1 = tau5 / (eps + b1);
    double t2 = tau5 / (eps + b2);
    double a0 = d0 * (1.0 + t0 * t0);
    double a1 = d1 * (1.0 + t1 * t1);
    double a2 = d2 * (1.0 + t2 * t2);
    double asum = a0 + a1 + a2;
    double w0 = a0 / asum, w1 = a1 / asum, w2 = a2 / asum;
    return w0 * q0 + w1 * q1 + w2 * q2;       // src/weno5.jl:40 / :60
}

// Left and right WENO5-Z values at the interface between a[2] and a[3]
// (src/weno5.jl:25-61 applied to a[0..4] and a[1..5]).
__device__ __forceinline__ void weno5_pair(double a0, double a1, double a2, double a3,
                                           double a4, double a5, double& vl, double& vr) {
    const double c13 = 13.0 / 12.0, c14 = 1.0 / 4.0;
    // candidate polynomials
    double qL0 = (1.0 / 3.0) * a0 + (-7.0 / 6.0) * a1 + (11.0 / 6.0) * a2;
    double qC  = (-1.0 / 6.0) * a1 + (5.0 / 6.0) * a2 + (1.0 / 3.0) * a3;   // left q1, right q0
    double qR  = (1.0 / 3.0) * a2 + (5.0 / 6.0) * a3 + (-1.0 / 6.0) * a4;   // left q2, right q1
    double qR2 = (11.0 / 6.0) * a3 + (-7.0 / 6.0) * a4 + (1.0 / 3.0) * a5;
    // smoothness indicators
    double s, t;
    s = a0 - 2.0 * a1 + a2;  t = a0 - 4.0 * a1 + 3.0 * a2;
    double bL0 = c13 * (s * s) + c14 * (t * t);
    s = a1 - 2.0 * a2 + a3;  t = a1 - a3;
    double bC = c13 * (s * s) + c14 * (t * t);                  // left beta1, right beta0
    s = a2 - 2.0 * a3 + a4;
    double h = c13 * (s * s);                                    // shared first term
    t = 3.0 * a2 - 4.0 * a3 + a4;
    double bL2 = h + c14 * (t * t);
    t = a2 - 4.0 * a3 + 3.0 * a4;
    double bR1 = h + c14 * (t * t);
    s = a3 - 2.0 * a4 + a5;  t = 3.0 * a3 - 4.0 * a4 + a5;
    double bR2 = c13 * (s * s) + c14 * (t * t);
    vl = wenoz_combine(bL0, bC, bL2, 1.0 / 10.0, 6.0 / 10.0, 3.0 / 10.0, qL0, qC, qR);
    vr = wenoz_combine(bC, bR1, bR2, 3.0 / 10.0, 6.0 / 10.0, 1.0 / 10.0, qC, qR, qR2);
}

// src/hllc.jl:17-165.  Inputs in (rho, m_normal, m_t1, m_t2, E) order.
__device__ __forceinline__ void hllc_normal(double rL, double mnL, double mt1L, double mt2L,
                                            double EL, double rR, double mnR, double mt1R,
                                            double mt2R, double ER, double g, double gfac,
                                            double F[5]) {
    if (rL <= 0.0) { rL = 1e-30; mnL = 0.0; mt1L = 0.0; mt2L = 0.0; EL = 0.0; }
    else           { rL = jl_max(rL, 1e-30); }
    if (rR <= 0.0) { rR = 1e-30; mnR = 0.0; mt1R = 0.0; mt2R = 0.0; ER = 0.0; }
    else           { rR = jl_max(rR, 1e-30); }

    double m2L = mnL * mnL + mt1L * mt1L + mt2L * mt2L;
    if (m2L > (rL * rL) * 1e4) {
        double s = (rL * 100.0) / sqrt(m2L);
        mnL *= s; mt1L *= s; mt2L *= s;
    }
    double m2R = mnR * mnR + mt1R * mt1R + mt2R * mt2R;
    if (m2R > (rR * rR) * 1e4) {
        double s = (rR * 100.0) / sqrt(m2R);
        mnR *= s; mt1R *= s; mt2R *= s;
    }
    double KE_L = jl_min((0.5 * (mnL * mnL + mt1L * mt1L + mt2L * mt2L)) / rL, (0.5 * rL) * 1e4);
    double KE_R = jl_min((0.5 * (mnR * mnR + mt1R * mt1R + mt2R * mt2R)) / rR, (0.5 * rR) * 1e4);
    EL = jl_max(EL, KE_L);
    ER = jl_max(ER, KE_R);

    double vnL = mnL / rL, vt1L = mt1L / rL, vt2L = mt2L / rL;
    double vnR = mnR / rR, vt1R = mt1R / rR, vt2R = mt2R / rR;
    double pL = (g - 1.0) * (EL - KE_L);
    double pR = (g - 1.0) * (ER - KE_R);
    double csL = sqrt((g * pL) / rL);
    double csR = sqrt((g * pR) / rR);

    double SL, SR;
    if (pL > 0.0 && pR > 0.0) {
        double r_avg = 0.5 * (rL + rR);
        double cs_avg = 0.5 * (csL + csR);
        double p_star = jl_max(0.0, 0.5 * (pL + pR) - ((0.5 * (vnR - vnL)) * r_avg) * cs_avg);
        double qL = p_star > pL ? sqrt(1.0 + gfac * (p_star / pL - 1.0)) : 1.0;
        double qR = p_star > pR ? sqrt(1.0 + gfac * (p_star / pR - 1.0)) : 1.0;
        SL = vnL - csL * qL;
        SR = vnR + csR * qR;
    } else {
        SL = jl_min(vnL - csL, vnR - csR);
        SR = jl_max(vnL + csL, vnR + csR);
    }

    double denom = rL * (SL - vnL) - rR * (SR - vnR);
    if (fabs(denom) < 1e-100) { F[0] = F[1] = F[2] = F[3] = F[4] = 0.0; return; }
    double S_star = ((pR - pL) + (rL * vnL) * (SL - vnL) - (rR * vnR) * (SR - vnR)) / denom;

    double FrL = mnL, FmnL = rL * (vnL * vnL) + pL, Fmt1L = mt1L * vnL, Fmt2L = mt2L * vnL;
    double FEL = (EL + pL) * vnL;
    if (SL >= 0.0) { F[0] = FrL; F[1] = FmnL; F[2] = Fmt1L; F[3] = Fmt2L; F[4] = FEL; return; }
    double FrR = mnR, FmnR = rR * (vnR * vnR) + pR, Fmt1R = mt1R * vnR, Fmt2R = mt2R * vnR;
    double FER = (ER + pR) * vnR;
    if (SR <= 0.0) { F[0] = FrR; F[1] = FmnR; F[2] = Fmt1R; F[3] = Fmt2R; F[4] = FER; return; }

    if (jl_max(rL, rR) > 3.0 * jl_min(rL, rR)) {
        double inv_dS = 1.0 / (SR - SL);
        double SLSR = SL * SR;
        F[0] = (SR * FrL - SL * FrR + SLSR * (rR - rL)) * inv_dS;
        F[1] = (SR * FmnL - SL * FmnR + SLSR * (mnR - mnL)) * inv_dS;
        F[2] = (SR * Fmt1L - SL * Fmt1R + SLSR * (mt1R - mt1L)) * inv_dS;
        F[3] = (SR * Fmt2L - SL * Fmt2R + SLSR * (mt2R - mt2L)) * inv_dS;
        F[4] = (SR * FEL - SL * FER + SLSR * (ER - EL)) * inv_dS;
        return;
    }
    if (S_star >= 0.0) {
        double dL = rL * (SL - vnL);
        double chi = dL / (SL - S_star);
        double EsL = fabs(dL) > 1e-100 ? chi * (EL / rL + (S_star - vnL) * (S_star + pL / dL)) : 0.0;
        F[0] = FrL + SL * (chi - rL);
        F[1] = FmnL + SL * (chi * S_star - mnL);
        F[2] = Fmt1L + SL * (chi * vt1L - mt1L);
        F[3] = Fmt2L + SL * (chi * vt2L - mt2L);
        F[4] = FEL + SL * (EsL - EL);
    } else {
        double dR = rR * (SR - vnR);
        double chi = dR / (SR - S_star);
        double EsR = fabs(dR) > 1e-100 ? chi * (ER / rR + (S_star - vnR) * (S_star + pR / dR)) : 0.0;
        F[0] = FrR + SR * (chi - rR);
        F[1] = FmnR + SR * (chi * S_star - mnR);
        F[2] = Fmt1R + SR * (chi * vt1R - mt1R);
        F[3] = Fmt2R + SR * (chi * vt2R - mt2R);
        F[4] = FER + SR * (EsR - ER);
    }
}

// One interface flux along direction DIR (0 = x, 1 = y, 2 = z) from the
// six-cell stencil c[q][s] (q = rho, mx, my, mz; s = 0..5 for cells i-2..i+3
// of the reference's src/gpu_kernels.jl:61-82) and cell pressures p[s].
template <int DIR>
__device__ __forceinline__ void interface_flux(const double (&c)[4][6], const double (&p)[6],
                                               double g, double gfac, double F[5]) {
    double L[4], R[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        weno5_pair(c[q][0], c[q][1], c[q][2], c[q][3], c[q][4], c[q][5], L[q], R[q]);
    double P_maxL = jl_max(jl_max(jl_max(jl_max(p[0], p[1]), p[2]), p[3]), p[4]);
    double P_maxR = jl_max(jl_max(jl_max(jl_max(p[1], p[2]), p[3]), p[4]), p[5]);
    double PLw, PRw;
    weno5_pair(p[0], p[1], p[2], p[3], p[4], p[5], PLw, PRw);
    double PL = jl_clamp(PLw, 0.0, P_maxL);
    double PR = jl_clamp(PRw, 0.0, P_maxR);
    double EL = EL_from_P(PL, L[0], L[1], L[2], L[3], g);
    double ER = EL_from_P(PR, R[0], R[1], R[2], R[3], g);
    double G[5];
    if (DIR == 0) {
        hllc_normal(L[0], L[1], L[2], L[3], EL, R[0], R[1], R[2], R[3], ER, g, gfac, G);
        F[0] = G[0]; F[1] = G[1]; F[2] = G[2]; F[3] = G[3]; F[4] = G[4];
    } else if (DIR == 1) {
        hllc_normal(L[0], L[2], L[1], L[3], EL, R[0], R[2], R[1], R[3], ER, g, gfac, G);
        F[0] = G[0]; F[1] = G[2]; F[2] = G[1]; F[3] = G[3]; F[4] = G[4];
    } else {
        hllc_normal(L[0], L[3], L[1], L[2], EL, R[0], R[3], R[1], R[2], ER, g, gfac, G);
        F[0] = G[0]; F[1] = G[2]; F[2] = G[3]; F[3] = G[1]; F[4] = G[4];
    }
}

}  // namespace ex

// ---------------------------------------------------------------------------
// Split precision.  FP32 with hardware reciprocal (v_rcp_f32, 1 ulp) and
// square root (v_sqrt_f32).  Every fused multiply-add is written out
// (__builtin_fma / __builtin_fmaf) and the compiler contracts nothing (the
// library is built with -ffp-contract=off): an interface's flux then has the
// same bits in every kernel that computes it, whatever code surrounds the
// call, which the fused and in-place step forms and the tile and walk shapes
// rely on.
namespace sp {

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// 1/x in FP64 from the hardware reciprocal and one Newton step (relative
// error a few ulp, not correctly rounded): the split form's cell pressure
// needs FP64 range and about 1e-15 relative accuracy, not IEEE division.
__device__ __forceinline__ double rcp64(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}

// Cell pressure for the FP32 stencils (src/euler3d.jl:86-90), formed in FP64
// so that E - rho v^2 / 2 does not cancel in FP32.
__device__ __forceinline__ double cell_pressure(double r, double mx, double my, double mz, double E,
                                                double g) {
    r = __builtin_fmax(r, 1e-30);
    const double m2 = __builtin_fma(mz, mz, __builtin_fma(my, my, mx * mx));
    const double KE = (0.5 * m2) * rcp64(r);
    return __builtin_fmax((g - 1.0) * (E - KE), 0.0);
}
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
// Hardware max / min (one v_max_f32 / v_min_f32 instead of a compare and a
// select).  They differ from the reference's max / min only in which operand
// a NaN yields, which the split form does not promise to reproduce.
__device__ __forceinline__ float fmaxj(float x, float y) { return __builtin_fmaxf(x, y); }
__device__ __forceinline__ float fminj(float x, float y) { return __builtin_fminf(x, y); }

// WENO5-Z on three-cell positions.  The interface between a[2] and a[3]
// reconstructs from four positions j = 0..3 of three consecutive cells
// (x, y, z) = (a[j], a[j+1], a[j+2]).  Every smoothness indicator and
// candidate polynomial of src/weno5.jl:25-61 is a function of one position's
// second difference s = x - 2y + z and centred difference tB = x - z:
//   t forms   tA = x - 4y + 3z = 2s - tB,   tC = 3x - 4y + z = 2s + tB,
//   beta      13/12 s^2 + 1/4 t^2 = 13/12 (s^2 + 3/13 t^2),
//   left  (d = 1/10, 6/10, 3/10): beta(A, j0), beta(B, j1), beta(C, j2),
//                                 q = qA(j0), qB(j1), qR(j2)
//   right (d = 3/10, 6/10, 1/10): beta(B, j1), beta(A, j2), beta(C, j3),
//                                 q = qB(j1), qR(j2), qD(j3)
// with qA = z + s/12 + tA/4 = 1/3 x - 7/6 y + 11/6 z,
//      qB = y + s/12 - tB/4 = -1/6 x + 5/6 y + 1/3 z,
//      qR = y - 5/12 s + tC/4 = 1/3 x + 5/6 y - 1/6 z,
//      qD = x + s/12 + tC/4 = 11/6 x - 7/6 y + 1/3 z.
// The WENO-Z weights are unchanged when every beta, tau and eps is scaled by
// one factor, so the betas here are the reference's divided by 13/12.
//
// On a uniform state s and t vanish, every candidate is the constant
// exactly, and the left and right values come out equal (see wenoz_pair), so
// a uniform state at rest has no mass or energy flux in FP32 either (an
// outflow boundary does not leak).  The combination stays a weighted average
// of the candidates: a form around one candidate would cancel where the
// weights single out a candidate much smaller than the others (a cold medium
// next to a shock).
//
// Each value is formed by explicit operations, so a value carried from one
// interface to the next along a walk (the indicator or candidate of a
// position the previous interfaces also used) has the bits a fresh evaluation
// gives, and every kernel reconstructs an interface identically.

struct Pos3 { float s, tB, ss; };

__device__ __forceinline__ Pos3 pos3(float x, float y, float z) {
    Pos3 P;
    P.s = __builtin_fmaf(-2.0f, y, x + z);
    P.tB = x - z;
    P.ss = P.s * P.s;
    return P;
}
__device__ __forceinline__ float tA(const Pos3& P) { return __builtin_fmaf(2.0f, P.s, -P.tB); }
__device__ __forceinline__ float tC(const Pos3& P) { return __builtin_fmaf(2.0f, P.s, P.tB); }
__device__ __forceinline__ float beta(const Pos3& P, float t) {
    return __builtin_fmaf((3.0f / 13.0f) * t, t, P.ss);
}
__device__ __forceinline__ float qA(const Pos3& P, float ta, float z) {
    return __builtin_fmaf(0.25f, ta, __builtin_fmaf(1.0f / 12.0f, P.s, z));
}
__device__ __forceinline__ float qB(const Pos3& P, float y) {
    return __builtin_fmaf(-0.25f, P.tB, __builtin_fmaf(1.0f / 12.0f, P.s, y));
}
__device__ __forceinline__ float qR(const Pos3& P, float tc, float y) {
    return __builtin_fmaf(0.25f, tc, __builtin_fmaf(-5.0f / 12.0f, P.s, y));
}
__device__ __forceinline__ float qD(const Pos3& P, float tc, float x) {
    return __builtin_fmaf(0.25f, tc, __builtin_fmaf(1.0f / 12.0f, P.s, x));
}

// Left and right WENO5-Z (p = 2) values from the five indicators and four
// candidate values of one interface.  eps per side is the reference's 1e-36,
// raised to 1e-17 tau where needed so that (tau / (eps + beta))^2 stays below
// 1e34 and cannot overflow FP32 (whenever the raise is active the weight of
// that stencil is already 1 - O(1e-34)).  One eps for both sides, raised by
// the larger tau, would let a shock on one side swamp the tiny indicators of
// a cold, nearly uniform medium on the other.  The weights d_k are scaled to
// (1, 6, 3) and (3, 6, 1); the normalisation removes the factor.
__device__ __forceinline__ void wenoz_pair(float bA0, float bB1, float bC2, float bA2, float bC3,
                                           float qa0, float qb1, float qr2, float qd3, float& vl,
                                           float& vr) {
    const float tauL = __builtin_fabsf(bA0 - bC2), tauR = __builtin_fabsf(bB1 - bC3);
    const float epsL = 1e-17f * __builtin_fmaxf(1e-19f, tauL);
    const float epsR = 1e-17f * __builtin_fmaxf(1e-19f, tauR);
    // Both sides sum (outer, outer) first, then the centre, so that on a
    // uniform state (weights 1, 6, 3 and 3, 6, 1, equal candidates) the two
    // sums are the same operations on the same values in swapped order: the
    // left and right values are equal and a uniform state at rest has no flux.
    {
        const float t0 = tauL * rcp(epsL + bA0), t1 = tauL * rcp(epsL + bB1);
        const float t2 = tauL * rcp(epsL + bC2);
        const float u0 = __builtin_fmaf(t0, t0, 1.0f);
        const float u1 = 6.0f * __builtin_fmaf(t1, t1, 1.0f);
        const float u2 = 3.0f * __builtin_fmaf(t2, t2, 1.0f);
        const float num = __builtin_fmaf(u1, qb1, u0 * qa0 + u2 * qr2);
        vl = num * rcp((u0 + u2) + u1);
    }
    {
        const float t0 = tauR * rcp(epsR + bB1), t1 = tauR * rcp(epsR + bA2);
        const float t2 = tauR * rcp(epsR + bC3);
        const float u0 = 3.0f * __builtin_fmaf(t0, t0, 1.0f);
        const float u1 = 6.0f * __builtin_fmaf(t1, t1, 1.0f);
        const float u2 = __builtin_fmaf(t2, t2, 1.0f);
        const float num = __builtin_fmaf(u1, qr2, u0 * qb1 + u2 * qd3);
        vr = num * rcp((u0 + u2) + u1);
    }
}

// What a walk carries from one interface to the next, per quantity: beta(A)
// and qA of the positions that become j = 0 and j = 1, and beta(C) of the one
// that becomes j = 2 (the previous interface's j = 1, j = 2 and j = 3).
struct WenoCarry { float A0, A1, C2, qA0, qA1; };

// The carry before the first interface of a walk (positions 0..2 of that
// interface, cells a[0..4]).
__device__ __forceinline__ void weno5_walk_start(float a0, float a1, float a2, float a3, float a4,
                                                 WenoCarry& w) {
    const Pos3 P0 = pos3(a0, a1, a2), P1 = pos3(a1, a2, a3), P2 = pos3(a2, a3, a4);
    const float ta0 = tA(P0), ta1 = tA(P1);
    w.A0 = beta(P0, ta0);
    w.qA0 = qA(P0, ta0, a2);
    w.A1 = beta(P1, ta1);
    w.qA1 = qA(P1, ta1, a3);
    w.C2 = beta(P2, tC(P2));
}

// One interface: with WALK position 0 and the C indicator of position 2 come
// from w (and w moves on to the next interface), otherwise all are formed here.
template <bool WALK>
__device__ __forceinline__ void weno5_pair(float a0, float a1, float a2, float a3, float a4,
                                           float a5, WenoCarry& w, float& vl, float& vr) {
    const Pos3 P1 = pos3(a1, a2, a3), P2 = pos3(a2, a3, a4), P3 = pos3(a3, a4, a5);
    const float tc2 = tC(P2), tc3 = tC(P3), ta2 = tA(P2);
    const float bB1 = beta(P1, P1.tB), bA2 = beta(P2, ta2), bC3 = beta(P3, tc3);
    float bA0, bC2, qa0;
    if (WALK) {
        bA0 = w.A0;
        qa0 = w.qA0;
        bC2 = w.C2;
        w.A0 = w.A1;
        w.qA0 = w.qA1;
        w.A1 = bA2;
        w.qA1 = qA(P2, ta2, a4);
        w.C2 = bC3;
    } else {
        const Pos3 P0 = pos3(a0, a1, a2);
        const float ta0 = tA(P0);
        bA0 = beta(P0, ta0);
        qa0 = qA(P0, ta0, a2);
        bC2 = beta(P2, tc2);
    }
    wenoz_pair(bA0, bB1, bC2, bA2, bC3, qa0, qB(P1, a2), qR(P2, tc2, a3), qD(P3, tc3, a3), vl, vr);
}

// The reference's HLLC (src/hllc.jl:17-165) in FP32: same branches and
// guards; 1e-100 thresholds become 1e-37 (the FP32 normal range).
// m2X, irX and KE_X are |m|^2, 1/max(rho, 1e-30) and the capped kinetic
// energy of each side as the interface energy rebuild formed them; they are
// formed again here only where a guard changes the state.
__device__ __forceinline__ void hllc_normal(float rL, float mnL, float mt1L, float mt2L, float EL,
                                            float m2L, float irL, float KE_L,
                                            float rR, float mnR, float mt1R, float mt2R, float ER,
                                            float m2R, float irR, float KE_R,
                                            float g, float gfac, float F[5]) {
    // rho <= 0 resets the side; 1/max(rho, 1e-30) is then already 1e30
    if (rL <= 0.0f) { rL = 1e-30f; mnL = 0.0f; mt1L = 0.0f; mt2L = 0.0f; EL = 0.0f; m2L = 0.0f; KE_L = 0.0f; }
    else            { rL = fmaxj(rL, 1e-30f); }
    if (rR <= 0.0f) { rR = 1e-30f; mnR = 0.0f; mt1R = 0.0f; mt2R = 0.0f; ER = 0.0f; m2R = 0.0f; KE_R = 0.0f; }
    else            { rR = fmaxj(rR, 1e-30f); }
    if (m2L > (rL * rL) * 1e4f) {
        float s = (rL * 100.0f) * __builtin_amdgcn_rsqf(m2L);
        mnL *= s; mt1L *= s; mt2L *= s;
        m2L = __builtin_fmaf(mt2L, mt2L, __builtin_fmaf(mt1L, mt1L, mnL * mnL));
        KE_L = fminj((0.5f * m2L) * irL, rL * 5e3f);
    }
    if (m2R > (rR * rR) * 1e4f) {
        float s = (rR * 100.0f) * __builtin_amdgcn_rsqf(m2R);
        mnR *= s; mt1R *= s; mt2R *= s;
        m2R = __builtin_fmaf(mt2R, mt2R, __builtin_fmaf(mt1R, mt1R, mnR * mnR));
        KE_R = fminj((0.5f * m2R) * irR, rR * 5e3f);
    }
    EL = fmaxj(EL, KE_L);
    ER = fmaxj(ER, KE_R);
    float vnL = mnL * irL;
    float vnR = mnR * irR;
    float pL = (g - 1.0f) * (EL - KE_L);
    float pR = (g - 1.0f) * (ER - KE_R);
    float csL = fsqrt((g * pL) * irL);
    float csR = fsqrt((g * pR) * irR);
    float SL, SR;
    if (pL > 0.0f && pR > 0.0f) {
        float r_avg = 0.5f * (rL + rR);
        float cs_avg = 0.5f * (csL + csR);
        float b = (0.5f * (vnR - vnL)) * r_avg;
        float p_star = fmaxj(0.0f, __builtin_fmaf(-b, cs_avg, 0.5f * (pL + pR)));
        float qL = p_star > pL ? fsqrt(__builtin_fmaf(gfac, __builtin_fmaf(p_star, rcp(pL), -1.0f), 1.0f)) : 1.0f;
        float qR = p_star > pR ? fsqrt(__builtin_fmaf(gfac, __builtin_fmaf(p_star, rcp(pR), -1.0f), 1.0f)) : 1.0f;
        SL = __builtin_fmaf(-csL, qL, vnL);
        SR = __builtin_fmaf(csR, qR, vnR);
    } else {
        SL = fminj(vnL - csL, vnR - csR);
        SR = fmaxj(vnL + csL, vnR + csR);
    }
    float dL = rL * (SL - vnL), dR = rR * (SR - vnR);
    float denom = dL - dR;
    if (fabsf(denom) < 1e-37f) { F[0] = F[1] = F[2] = F[3] = F[4] = 0.0f; return; }
    float FmnL = __builtin_fmaf(rL, vnL * vnL, pL), FEL = (EL + pL) * vnL;
    if (SL >= 0.0f) { F[0] = mnL; F[1] = FmnL; F[2] = mt1L * vnL; F[3] = mt2L * vnL; F[4] = FEL; return; }
    float FmnR = __builtin_fmaf(rR, vnR * vnR, pR), FER = (ER + pR) * vnR;
    if (SR <= 0.0f) { F[0] = mnR; F[1] = FmnR; F[2] = mt1R * vnR; F[3] = mt2R * vnR; F[4] = FER; return; }
    if (fmaxj(rL, rR) > 3.0f * fminj(rL, rR)) {
        float inv_dS = rcp(SR - SL);
        float SLSR = SL * SR;
        F[0] = __builtin_fmaf(SLSR, rR - rL, __builtin_fmaf(SR, mnL, -(SL * mnR))) * inv_dS;
        F[1] = __builtin_fmaf(SLSR, mnR - mnL, __builtin_fmaf(SR, FmnL, -(SL * FmnR))) * inv_dS;
        F[2] = __builtin_fmaf(SLSR, mt1R - mt1L, __builtin_fmaf(SR, mt1L * vnL, -(SL * (mt1R * vnR)))) * inv_dS;
        F[3] = __builtin_fmaf(SLSR, mt2R - mt2L, __builtin_fmaf(SR, mt2L * vnL, -(SL * (mt2R * vnR)))) * inv_dS;
        F[4] = __builtin_fmaf(SLSR, ER - EL, __builtin_fmaf(SR, FEL, -(SL * FER))) * inv_dS;
        return;
    }
    float S_star = __builtin_fmaf(-mnR, SR - vnR, __builtin_fmaf(mnL, SL - vnL, pR - pL)) * rcp(denom);
    // Star-state jumps written through del = (S* - vn) / (S - S*), so that
    // chi = rho (1 + del) and every jump U* - U vanishes exactly when S* = vn
    // (identical states at rest give zero mass and energy flux in FP32 too);
    // algebraically the reference's expressions (src/hllc.jl:140-165).  The
    // side (left star state if S* >= 0, else right) is chosen per lane first,
    // so a wave whose lanes differ evaluates the star formulas once.
    const bool lft = S_star >= 0.0f;
    const float r = lft ? rL : rR, mn = lft ? mnL : mnR, mt1 = lft ? mt1L : mt1R;
    const float mt2 = lft ? mt2L : mt2R, E = lft ? EL : ER, p = lft ? pL : pR;
    const float vn = lft ? vnL : vnR, S = lft ? SL : SR, d = lft ? dL : dR;
    const float Fmn = lft ? FmnL : FmnR, FE = lft ? FEL : FER;
    float del = (S_star - vn) * rcp(S - S_star);
    float chi = __builtin_fmaf(r, del, r);
    float dE = fabsf(d) > 1e-37f
                   ? __builtin_fmaf(E, del, (chi * (S_star - vn)) * __builtin_fmaf(p, rcp(d), S_star))
                   : -E;
    F[0] = __builtin_fmaf(S, r * del, mn);
    F[1] = __builtin_fmaf(S, __builtin_fmaf(chi, S_star, -mn), Fmn);
    F[2] = __builtin_fmaf(S, mt1 * del, mt1 * vn);
    F[3] = __builtin_fmaf(S, mt2 * del, mt2 * vn);
    F[4] = __builtin_fmaf(S, dE, FE);
}

// Interface flux along DIR from the six-cell stencil c[q][s], p[s] (as in
// the FP64 form).  WALK: the indicators of positions 0 and 2 come from the
// carry w[5] of the quantities (rho, m_x, m_y, m_z, P), which moves on.
template <int DIR, bool WALK = false>
__device__ __forceinline__ void interface_flux(const float (&c)[4][6], const float (&p)[6],
                                               float g, float gfac, float ig1, float F[5],
                                               WenoCarry (&w)[5]) {
    float L[4], R[4];
    // one quantity at a time: interleaving all five reconstructions raises
    // the register pressure past what the fused update kernel can hold
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        weno5_pair<WALK>(c[q][0], c[q][1], c[q][2], c[q][3], c[q][4], c[q][5], w[q], L[q], R[q]);
        __builtin_amdgcn_sched_barrier(0);
    }
    float P_maxL = fmaxj(fmaxj(fmaxj(fmaxj(p[0], p[1]), p[2]), p[3]), p[4]);
    float P_maxR = fmaxj(fmaxj(fmaxj(fmaxj(p[1], p[2]), p[3]), p[4]), p[5]);
    float PLw, PRw;
    weno5_pair<WALK>(p[0], p[1], p[2], p[3], p[4], p[5], w[4], PLw, PRw);
    float PL = fminj(fmaxj(PLw, 0.0f), P_maxL);
    float PR = fminj(fmaxj(PRw, 0.0f), P_maxR);
    float rpL = fmaxj(L[0], 1e-30f), rpR = fmaxj(R[0], 1e-30f);
    float irL = rcp(rpL), irR = rcp(rpR);
    float m2L = __builtin_fmaf(L[3], L[3], __builtin_fmaf(L[2], L[2], L[1] * L[1]));
    float m2R = __builtin_fmaf(R[3], R[3], __builtin_fmaf(R[2], R[2], R[1] * R[1]));
    float KEL = fminj((0.5f * m2L) * irL, rpL * 5e3f);
    float KER = fminj((0.5f * m2R) * irR, rpR * 5e3f);
    float EL = __builtin_fmaf(PL, ig1, KEL);
    float ER = __builtin_fmaf(PR, ig1, KER);
    float G[5];
    if (DIR == 0) {
        hllc_normal(L[0], L[1], L[2], L[3], EL, m2L, irL, KEL, R[0], R[1], R[2], R[3], ER, m2R, irR, KER,
                    g, gfac, G);
        F[0] = G[0]; F[1] = G[1]; F[2] = G[2]; F[3] = G[3]; F[4] = G[4];
    } else if (DIR == 1) {
        hllc_normal(L[0], L[2], L[1], L[3], EL, m2L, irL, KEL, R[0], R[2], R[1], R[3], ER, m2R, irR, KER,
                    g, gfac, G);
        F[0] = G[0]; F[1] = G[2]; F[2] = G[1]; F[3] = G[3]; F[4] = G[4];
    } else {
        hllc_normal(L[0], L[3], L[1], L[2], EL, m2L, irL, KEL, R[0], R[3], R[1], R[2], ER, m2R, irR, KER,
                    g, gfac, G);
        F[0] = G[0]; F[1] = G[2]; F[2] = G[3]; F[3] = G[1]; F[4] = G[4];
    }
}

template <int DIR>
__device__ __forceinline__ void interface_flux(const float (&c)[4][6], const float (&p)[6],
                                               float g, float gfac, float ig1, float F[5]) {
    WenoCarry w[5];
    interface_flux<DIR, false>(c, p, g, gfac, ig1, F, w);
}

// Starts a walk along DIR: the carry for the first interface from cells
// 0..4 of its stencil.
__device__ __forceinline__ void interface_walk_start(const float (&c)[4][6], const float (&p)[6],
                                                     WenoCarry (&w)[5]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) weno5_walk_start(c[q][0], c[q][1], c[q][2], c[q][3], c[q][4], w[q]);
    weno5_walk_start(p[0], p[1], p[2], p[3], p[4], w[4]);
}

}  // namespace sp
}  // namespace bsn
