// Device arithmetic policies, limiters and fluxes of the four dimensionally
// split schemes.
//
// Every function is written in "sweep order": state index 0 is the density
// (h or rho), 1 the momentum normal to the sweep, 2 the tangential momentum,
// 3 the energy (Euler).  A y sweep passes (h, hv, hu[, E]) and maps the result
// back, as computeFluxG does in GPUSimulators/cuda/SWE2D_KP07_dimsplit.cu:71-110.
//
// Two arithmetic policies share the code:
//   MathExact  IEEE division and square root, and (by the translation unit's
//              -ffp-contract=off) no fused multiply-add.  Each expression is
//              evaluated in the operand order of the CUDA source, so results
//              are bitwise equal to an IEEE float32 evaluation of the
//              reference arithmetic.
//   MathFast   hardware reciprocal / square root (v_rcp_f32, v_sqrt_f32, about
//              1 ulp) with shared reciprocals and contraction allowed.  Held
//              against the float32 oracle at relative L1 <= 1e-5.
#pragma once

#include <hip/hip_runtime.h>

namespace fvmi {

// Correctly rounded float32 division and square root in fewer instructions
// than the compiler's general IEEE sequences, for operands in the normal range
// (no scaling for denormal or overflowing intermediates, no special-case
// fix-up: a zero or infinite divisor gives NaN).
//  * reciprocal: v_rcp_f32 (1 ulp) refined by one Newton step with FMA,
//    y = y0 + y0 (1 - b y0), which rounds to the correctly rounded 1/b;
//  * quotient: q0 = a y, the exact remainder r = a - b q0 by FMA, and
//    q = q0 + r y: correctly rounded when y is (Markstein's theorem); a zero
//    quotient is q0 itself, whose sign is that of a b (q0 + r y would turn
//    -0 into +0, and the sign of a zero reaches copysign in the WAF limiter);
//  * square root: v_sqrt_f32 (1 ulp) and its two neighbours, the one whose
//    residual x - s s' changes sign chosen by two FMAs.
// Checked against the IEEE sequences on every mantissa of [1, 2) (reciprocal)
// and [1, 4) (square root) and on random quotients, and in every exact-
// arithmetic test, which compare the fields bitwise with the float32 oracle.
__device__ __forceinline__ float rcp_rn(float b) {
    const float y = __builtin_amdgcn_rcpf(b);
    return __builtin_fmaf(__builtin_fmaf(-b, y, 1.0f), y, y);
}
__device__ __forceinline__ float div_rn(float a, float b) {
    const float y = rcp_rn(b);
    const float q = a * y;
    const float qc = __builtin_fmaf(__builtin_fmaf(-b, q, a), y, q);
    return q == 0.0f ? q : qc;
}
__device__ __forceinline__ float sqrt_rn(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float dn = __uint_as_float(__float_as_uint(s) - 1u);
    const float up = __uint_as_float(__float_as_uint(s) + 1u);
    float r = (__builtin_fmaf(-dn, s, x) <= 0.0f) ? dn : s;
    r = (__builtin_fmaf(-up, s, x) > 0.0f) ? up : r;
    return r;
}

struct MathExact {
    static constexpr bool exact = true;
    __device__ static __forceinline__ float div(float a, float b) { return div_rn(a, b); }
    __device__ static __forceinline__ float sqrt(float x) { return sqrt_rn(x); }
    __device__ static __forceinline__ float rcp(float x) { return rcp_rn(x); }
};

struct MathFast {
    static constexpr bool exact = false;
    __device__ static __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    __device__ static __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    __device__ static __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
};

// ---------------------------------------------------------------------------
// Limiters: GPUSimulators/cuda/limiters.h, common.h
// ---------------------------------------------------------------------------

// minmodSlope, limiters.h:31-41, in the reference's operation order (exact
// policy).  The fast policy evaluates the same minmod as
// med3(med3(b, f, 0), c, 0) -- the inner median is 0 unless b and f share a
// sign, and then c shares it too -- in its scaled form (reconstruct in
// fv_substep.h, half_slope in fv_unsplit.h).
template <class M>
__device__ __forceinline__ float minmod_slope(float left, float center, float right, float theta) {
    static_assert(M::exact, "the fast policy uses the scaled med3 form");
    const float backward = (center - left) * theta;
    const float central = (right - left) * 0.5f;
    const float forward = (right - center) * theta;
    const float sb = __builtin_copysignf(1.0f, backward);
    const float sc = __builtin_copysignf(1.0f, central);
    const float sf = __builtin_copysignf(1.0f, forward);
    return 0.25f * sb * (sb + sc) * (sc + sf) * fminf(fminf(fabsf(backward), fabsf(central)), fabsf(forward));
}

// generalized_minmod, limiters.h:100-102
__device__ __forceinline__ float generalized_minmod(float r, float theta) {
    return fmaxf(0.0f, fminf(theta * r, fminf((1.0f + r) / 2.0f, theta)));
}

// desingularize, common.h:81-83
template <class M>
__device__ __forceinline__ float desingularize(float x, float eps) {
    return __builtin_copysignf(1.0f, x) *
           fmaxf(fabsf(x), fminf(M::div(x * x, 2.0f * eps) + 0.5f * eps, eps));
}

// limiterToWAFLimiter, SWECommon.h:109-111
__device__ __forceinline__ float limiter_to_waf(float r, float c) {
    return 1.0f - (1.0f - fabsf(c)) * r;
}

// ---------------------------------------------------------------------------
// Shallow water: GPUSimulators/cuda/SWECommon.h
// ---------------------------------------------------------------------------

// F_func, SWECommon.h:29-37
template <class M>
__device__ __forceinline__ void swe_F(const float q[3], float g, float F[3]) {
    if constexpr (M::exact) {
        F[0] = q[1];
        F[1] = M::div(q[1] * q[1], q[0]) + 0.5f * g * q[0] * q[0];
        F[2] = M::div(q[1] * q[2], q[0]);
    } else {
        const float u = q[1] * M::rcp(q[0]);
        F[0] = q[1];
        F[1] = q[1] * u + 0.5f * g * q[0] * q[0];
        F[2] = q[2] * u;
    }
}

// CentralUpwindFlux, SWECommon.h:228-241.  qm = left state, qp = right state.
template <class M>
__device__ __forceinline__ void swe_cu_flux(const float qm[3], const float qp[3], float g, float out[3]) {
    if constexpr (M::exact) {
        float Fp[3], Fm[3];
        swe_F<M>(qp, g, Fp);
        const float up = M::div(qp[1], qp[0]);
        const float cp = M::sqrt(g * qp[0]);
        swe_F<M>(qm, g, Fm);
        const float um = M::div(qm[1], qm[0]);
        const float cm = M::sqrt(g * qm[0]);
        const float am = fminf(fminf(um - cm, up - cp), 0.0f);
        const float ap = fmaxf(fmaxf(um + cm, up + cp), 0.0f);
        const float apam = ap * am;
        const float den = ap - am;
#pragma unroll
        for (int k = 0; k < 3; ++k) out[k] = M::div((ap * Fm[k] - am * Fp[k]) + apam * (qp[k] - qm[k]), den);
    } else {
        const float ip = M::rcp(qp[0]);
        const float im = M::rcp(qm[0]);
        const float up = qp[1] * ip, um = qm[1] * im;
        const float cp = M::sqrt(g * qp[0]), cm = M::sqrt(g * qm[0]);
        const float hg = 0.5f * g;
        const float Fp[3] = {qp[1], qp[1] * up + hg * qp[0] * qp[0], qp[2] * up};
        const float Fm[3] = {qm[1], qm[1] * um + hg * qm[0] * qm[0], qm[2] * um};
        const float am = fminf(fminf(um - cm, up - cp), 0.0f);
        const float ap = fmaxf(fmaxf(um + cm, up + cp), 0.0f);
        const float apam = ap * am;
        const float iden = M::rcp(ap - am);
#pragma unroll
        for (int k = 0; k < 3; ++k) out[k] = ((ap * Fm[k] - am * Fp[k]) + apam * (qp[k] - qm[k])) * iden;
    }
}

// HLL_flux (SWE), SWECommon.h:274-313
template <class M>
__device__ __forceinline__ void swe_hll_flux(const float ql[3], const float qr[3], float g, float out[3]) {
    const float h_l = ql[0], h_r = qr[0];
    float u_l, u_r;
    if constexpr (M::exact) {
        u_l = M::div(ql[1], h_l);
        u_r = M::div(qr[1], h_r);
    } else {
        u_l = ql[1] * M::rcp(h_l);
        u_r = qr[1] * M::rcp(h_r);
    }
    const float c_l = M::sqrt(g * h_l);
    const float c_r = M::sqrt(g * h_r);
    const float h_dag = 0.5f * (h_l + h_r) - M::div(0.25f * (u_r - u_l) * (h_l + h_r), c_l + c_r);
    const float q_l_tmp = M::sqrt(0.5f * M::div((h_dag + h_l) * h_dag, h_l * h_l));
    const float q_r_tmp = M::sqrt(0.5f * M::div((h_dag + h_r) * h_dag, h_r * h_r));
    const float q_l = (h_dag > h_l) ? q_l_tmp : 1.0f;
    const float q_r = (h_dag > h_r) ? q_r_tmp : 1.0f;
    const float S_l = u_l - c_l * q_l;
    const float S_r = u_r + c_r * q_r;
    float F_l[3], F_r[3];
    swe_F<M>(ql, g, F_l);
    swe_F<M>(qr, g, F_r);
    const float srsl = S_r * S_l;
    if constexpr (M::exact) {
        const float den = S_r - S_l;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float star = M::div(S_r * F_l[k] - S_l * F_r[k] + srsl * (qr[k] - ql[k]), den);
            out[k] = (S_l >= 0.0f) ? F_l[k] : ((S_r <= 0.0f) ? F_r[k] : star);
        }
    } else {
        const float iden = M::rcp(S_r - S_l);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float star = (S_r * F_l[k] - S_l * F_r[k] + srsl * (qr[k] - ql[k])) * iden;
            out[k] = (S_l >= 0.0f) ? F_l[k] : ((S_r <= 0.0f) ? F_r[k] : star);
        }
    }
}

// computeHStar, SWECommon.h:114-121
template <class M>
__device__ __forceinline__ float compute_h_star(float h_l, float h_r, float u_l, float u_r, float c_l, float c_r,
                                                float g) {
    (void)h_l;
    (void)h_r;
    const float h_tmp = 0.5f * (c_l + c_r) + 0.25f * (u_l - u_r);
    return M::div(h_tmp * h_tmp, g);
}

// ---------------------------------------------------------------------------
// WAF in pieces: per-cell quantities, h* per interface, and the flux from the
// two cells at the interface plus what it needs of the outer two.  A march
// computes each cell's quantities and each interface's h* once and shares
// them between the (up to four) fluxes that use them; the values are those
// swe_waf_flux computes for itself, so the exact policy stays bitwise equal.
// ---------------------------------------------------------------------------

struct WafCell {
    float h, hu, hv;  // state in sweep order (depth, normal, tangential momentum)
    float u, v, c;    // normal and tangential velocity, sqrt(g h)
    float ih;         // 1/h (fast policy only)
};

template <class M>
__device__ __forceinline__ WafCell waf_cell(const float q[3], float g, float sqrt_g) {
    WafCell p;
    p.h = q[0];
    p.hu = q[1];
    p.hv = q[2];
    if constexpr (M::exact) {
        p.u = M::div(q[1], q[0]);
        p.v = M::div(q[2], q[0]);
        p.c = M::sqrt(g * q[0]);
        p.ih = 0.0f;
    } else {
        const float s = __builtin_amdgcn_rsqf(q[0]);
        p.ih = s * s;
        p.u = q[1] * p.ih;
        p.v = q[2] * p.ih;
        p.c = sqrt_g * q[0] * s;
    }
    return p;
}

// the cell seen by the other sweep direction (normal and tangential swapped)
__device__ __forceinline__ WafCell waf_transpose(const WafCell& p) {
    WafCell t = p;
    t.hu = p.hv;
    t.hv = p.hu;
    t.u = p.v;
    t.v = p.u;
    return t;
}

// computeHStar (SWECommon.h:114-121) of the interface between l and r
template <class M>
__device__ __forceinline__ float waf_hstar(const WafCell& l, const WafCell& r, float g, float inv_g) {
    if constexpr (M::exact) {
        return compute_h_star<M>(l.h, r.h, l.u, r.u, l.c, r.c, g);
    } else {
        const float t = 0.5f * (l.c + r.c) + 0.25f * (l.u - r.u);
        return t * t * inv_g;
    }
}

// desingularize(x, 1e-6) of the WAF limiter ratios (SWECommon.h:170-176); fast
// form: the sign of x on the magnitude max(|x|, min(x^2/(2 eps) + eps/2, eps))
template <class M>
__device__ __forceinline__ float waf_desing(float x) {
    if constexpr (M::exact) {
        return desingularize<M>(x, 1.0e-6f);
    } else {
        const float eps = 1.0e-6f, inv_2eps = 0.5e6f;
        return __builtin_copysignf(
            __builtin_fmaxf(__builtin_fabsf(x), __builtin_fminf(x * x * inv_2eps + 0.5f * eps, eps)), x);
    }
}

// The limiter denominators of the interface between cells l and r with h* hd:
// desingularize(h* - h_l), desingularize(v_r - v_l), desingularize(h_r - h*)
// (SWECommon.h:170-176).  The upwind numerators of an interface's three
// ratios are the same quantities of the interface to its left (positive
// Courant number) or to its right, with the same operands in the same order,
// so each interface computes its own once and its neighbours take them.
struct WafDen {
    float d1, d2, d3;
};

template <class M>
__device__ __forceinline__ WafDen waf_den(const WafCell& l, const WafCell& r, float hd) {
    WafDen D;
    D.d1 = waf_desing<M>(hd - l.h);
    D.d2 = waf_desing<M>(r.v - l.v);
    D.d3 = waf_desing<M>(r.h - hd);
    return D;
}

// WAF_1D_flux (SWECommon.h:131-210) between cells l (= Q_l1) and r (= Q_r1)
// with h* h_dag, given the limiter denominators of this interface (D) and of
// the interfaces to its left (Dl) and right (Dr).  dtd = dt / d (fast policy).
template <class M>
__device__ __forceinline__ void waf_flux_core(const WafCell& L, const WafCell& R, const WafDen& Dl, const WafDen& D,
                                              const WafDen& Dr, float h_dag, float g, float hg, float d, float dt,
                                              float dtd, float out[3]) {
    const float h_l = L.h, h_r = R.h, u_l = L.u, u_r = R.u, v_l = L.v, v_r = R.v, c_l = L.c, c_r = R.c;
    const float q_l1[3] = {L.h, L.hu, L.hv};
    const float q_r1[3] = {R.h, R.hu, R.hv};
    if constexpr (M::exact) {
        const float q_l_tmp = M::div(M::sqrt(0.5f * ((h_dag + h_l) * h_dag)), h_l);
        const float q_r_tmp = M::div(M::sqrt(0.5f * ((h_dag + h_r) * h_dag)), h_r);
        const float q_l = (h_dag > h_l) ? q_l_tmp : 1.0f;
        const float q_r = (h_dag > h_r) ? q_r_tmp : 1.0f;
        const float S_l = u_l - c_l * q_l;
        const float S_r = u_r + c_r * q_r;
        const float S_star = M::div(S_l * h_r * (u_r - S_r) - S_r * h_l * (u_l - S_l),
                                    h_r * (u_r - S_r) - h_l * (u_l - S_l));
        const float s_l = M::div(h_l * (S_l - u_l), S_l - S_star);
        const float s_r = M::div(h_r * (S_r - u_r), S_r - S_star);
        const float Q_star_l[3] = {s_l * 1.0f, s_l * S_star, s_l * v_l};
        const float Q_star_r[3] = {s_r * 1.0f, s_r * S_star, s_r * v_r};
        float F_1[3], F_4[3], F_2[3], F_3[3];
        swe_F<M>(q_l1, g, F_1);
        swe_F<M>(q_r1, g, F_4);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            F_2[k] = F_1[k] + S_l * (Q_star_l[k] - q_l1[k]);
            F_3[k] = F_4[k] + S_r * (Q_star_r[k] - q_r1[k]);
        }
        const float c_1 = M::div(S_l * dt, d);
        const float c_2 = M::div(S_star * dt, d);
        const float c_3 = M::div(S_r * dt, d);
        // desingularize((c_1 > 0) ? h_dag_l - h_l2 : h_dag_r - h_r) / desingularize(h_dag - h_l), ...
        const float r_1 = M::div((c_1 > 0.0f) ? Dl.d1 : Dr.d1, D.d1);
        const float r_2 = M::div((c_2 > 0.0f) ? Dl.d2 : Dr.d2, D.d2);
        const float r_3 = M::div((c_3 > 0.0f) ? Dl.d3 : Dr.d3, D.d3);
        const float A_1 = __builtin_copysignf(1.0f, c_1) * limiter_to_waf(generalized_minmod(r_1, 1.9f), c_1);
        const float A_2 = __builtin_copysignf(1.0f, c_2) * limiter_to_waf(generalized_minmod(r_2, 1.9f), c_2);
        const float A_3 = __builtin_copysignf(1.0f, c_3) * limiter_to_waf(generalized_minmod(r_3, 1.9f), c_3);
#pragma unroll
        for (int k = 0; k < 3; ++k)
            out[k] = 0.5f * (F_1[k] + F_4[k]) -
                     0.5f * (A_1 * (F_2[k] - F_1[k]) + A_2 * (F_3[k] - F_2[k]) + A_3 * (F_4[k] - F_3[k]));
    } else {
        const float il = L.ih, ir = R.ih;
        const float q_l = (h_dag > h_l) ? M::sqrt(0.5f * ((h_dag + h_l) * h_dag)) * il : 1.0f;
        const float q_r = (h_dag > h_r) ? M::sqrt(0.5f * ((h_dag + h_r) * h_dag)) * ir : 1.0f;
        const float S_l = u_l - c_l * q_l;
        const float S_r = u_r + c_r * q_r;
        const float dl = h_l * (u_l - S_l), dr = h_r * (u_r - S_r);
        const float S_star = (S_l * dr - S_r * dl) * M::rcp(dr - dl);
        // shared reciprocal of (S_l - S_star)(S_r - S_star)
        const float el = S_l - S_star, er = S_r - S_star;
        const float ie = M::rcp(el * er);
        const float s_l = -dl * (er * ie);  // h_l (S_l - u_l) / (S_l - S_star)
        const float s_r = -dr * (el * ie);  // h_r (S_r - u_r) / (S_r - S_star)
        const float F_1[3] = {L.hu, L.hu * u_l + hg * h_l * h_l, L.hv * u_l};
        const float F_4[3] = {R.hu, R.hu * u_r + hg * h_r * h_r, R.hv * u_r};
        const float Qsl[3] = {s_l, s_l * S_star, s_l * v_l};
        const float Qsr[3] = {s_r, s_r * S_star, s_r * v_r};
        const float c_1 = S_l * dtd, c_2 = S_star * dtd, c_3 = S_r * dtd;
        auto gmm = [](float r) {
            return __builtin_fmaxf(0.0f, __builtin_fminf(1.9f * r, __builtin_fminf((1.0f + r) * 0.5f, 1.9f)));
        };
        const float r_1 = ((c_1 > 0.0f) ? Dl.d1 : Dr.d1) * M::rcp(D.d1);
        const float r_2 = ((c_2 > 0.0f) ? Dl.d2 : Dr.d2) * M::rcp(D.d2);
        const float r_3 = ((c_3 > 0.0f) ? Dl.d3 : Dr.d3) * M::rcp(D.d3);
        const float A_1 = __builtin_copysignf(1.0f, c_1) * (1.0f - (1.0f - __builtin_fabsf(c_1)) * gmm(r_1));
        const float A_2 = __builtin_copysignf(1.0f, c_2) * (1.0f - (1.0f - __builtin_fabsf(c_2)) * gmm(r_2));
        const float A_3 = __builtin_copysignf(1.0f, c_3) * (1.0f - (1.0f - __builtin_fabsf(c_3)) * gmm(r_3));
        // With a = F_2 - F_1 = S_l (Q*_l - Q_l1) and b = F_3 - F_4 = S_r (Q*_r - Q_r1),
        // A_1 (F_2 - F_1) + A_2 (F_3 - F_2) + A_3 (F_4 - F_3)
        //   = (A_1 - A_2) a + A_2 (F_4 - F_1) + (A_2 - A_3) b.
        const float w12 = 0.5f * (A_1 - A_2), w2 = 0.5f * A_2, w23 = 0.5f * (A_2 - A_3);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float a = S_l * (Qsl[k] - q_l1[k]);
            const float b = S_r * (Qsr[k] - q_r1[k]);
            out[k] = 0.5f * (F_1[k] + F_4[k]) - (w12 * a + w2 * (F_4[k] - F_1[k]) + w23 * b);
        }
    }
}

// writeCfl per-cell value, SWECommon.h:497-504
template <class M>
__device__ __forceinline__ float swe_cell_dt(const float q[3], float dx, float dy, float g) {
    if constexpr (M::exact) {
        const float h = q[0];
        const float u = M::div(q[1], h);
        const float v = M::div(q[2], h);
        const float max_u = M::div(dx, fabsf(u) + M::sqrt(g * h));
        const float max_v = M::div(dy, fabsf(v) + M::sqrt(g * h));
        return fminf(max_u, max_v);
    } else {
        const float ih = M::rcp(q[0]);
        const float c = M::sqrt(g * q[0]);
        return fminf(M::div(dx, fabsf(q[1] * ih) + c), M::div(dy, fabsf(q[2] * ih) + c));
    }
}

// ---------------------------------------------------------------------------
// Euler: GPUSimulators/cuda/EulerCommon.h
// ---------------------------------------------------------------------------

// pressure, EulerCommon.h:83-90
template <class M>
__device__ __forceinline__ float euler_pressure(const float q[4], float gamma) {
    return (gamma - 1.0f) * (q[3] - M::div(0.5f * (q[1] * q[1] + q[2] * q[2]), q[0]));
}

// F_func, EulerCommon.h:93-109.  P is a pressure.
template <class M>
__device__ __forceinline__ void euler_F(const float q[4], float P, float F[4]) {
    const float u = M::div(q[1], q[0]);
    F[0] = q[1];
    F[1] = q[1] * u + P;
    F[2] = q[2] * u;
    F[3] = u * (q[3] + P);
}

// HLL_flux (Euler), EulerCommon.h:117-160
template <class M>
__device__ __forceinline__ void euler_hll_flux(const float ql[4], const float qr[4], float gamma, float out[4]) {
    const float h_l = ql[0], h_r = qr[0];
    const float u_l = M::div(ql[1], h_l);
    const float u_r = M::div(qr[1], h_r);
    const float P_l = euler_pressure<M>(ql, gamma);
    const float P_r = euler_pressure<M>(qr, gamma);
    const float c_l = M::sqrt(M::div(gamma * P_l, ql[0]));
    const float c_r = M::sqrt(M::div(gamma * P_r, qr[0]));
    const float h_dag = 0.5f * (h_l + h_r) - M::div(0.25f * (u_r - u_l) * (h_l + h_r), c_l + c_r);
    const float q_l_tmp = M::sqrt(0.5f * M::div((h_dag + h_l) * h_dag, h_l * h_l));
    const float q_r_tmp = M::sqrt(0.5f * M::div((h_dag + h_r) * h_dag, h_r * h_r));
    const float q_l = (h_dag > h_l) ? q_l_tmp : 1.0f;
    const float q_r = (h_dag > h_r) ? q_r_tmp : 1.0f;
    const float S_l = u_l - c_l * q_l;
    const float S_r = u_r + c_r * q_r;
    float F_l[4], F_r[4];
    euler_F<M>(ql, P_l, F_l);
    euler_F<M>(qr, P_r, F_r);
    const float srsl = S_r * S_l;
    if constexpr (M::exact) {
        const float den = S_r - S_l;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float star = M::div(S_r * F_l[k] - S_l * F_r[k] + srsl * (qr[k] - ql[k]), den);
            out[k] = (S_l >= 0.0f) ? F_l[k] : ((S_r <= 0.0f) ? F_r[k] : star);
        }
    } else {
        const float iden = M::rcp(S_r - S_l);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float star = (S_r * F_l[k] - S_l * F_r[k] + srsl * (qr[k] - ql[k])) * iden;
            out[k] = (S_l >= 0.0f) ? F_l[k] : ((S_r <= 0.0f) ? F_r[k] : star);
        }
    }
}

// CentralUpwindFlux (Euler), EulerCommon.h:171-187
template <class M>
__device__ __forceinline__ void euler_cu_flux(const float qm[4], const float qp[4], float gamma, float out[4]) {
    const float Pp = euler_pressure<M>(qp, gamma);
    float Fp[4], Fm[4];
    euler_F<M>(qp, Pp, Fp);
    const float up = M::div(qp[1], qp[0]);
    const float cp = M::sqrt(M::div(gamma * Pp, qp[0]));
    const float Pm = euler_pressure<M>(qm, gamma);
    euler_F<M>(qm, Pm, Fm);
    const float um = M::div(qm[1], qm[0]);
    const float cm = M::sqrt(M::div(gamma * Pm, qm[0]));
    const float am = fminf(fminf(um - cm, up - cp), 0.0f);
    const float ap = fmaxf(fmaxf(um + cm, up + cp), 0.0f);
    const float apam = ap * am;
    if constexpr (M::exact) {
        const float den = ap - am;
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = M::div((ap * Fm[k] - am * Fp[k]) + apam * (qp[k] - qm[k]), den);
    } else {
        const float iden = M::rcp(ap - am);
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = ((ap * Fm[k] - am * Fp[k]) + apam * (qp[k] - qm[k])) * iden;
    }
}

// writeCfl per-cell value (Euler), EulerCommon.h:41-49: sqrt(gamma*rho) as written
template <class M>
__device__ __forceinline__ float euler_cell_dt(const float q[4], float dx, float dy, float gamma) {
    const float rho = q[0];
    const float u = M::div(q[1], rho);
    const float v = M::div(q[2], rho);
    const float max_u = M::div(dx, fabsf(u) + M::sqrt(gamma * rho));
    const float max_v = M::div(dy, fabsf(v) + M::sqrt(gamma * rho));
    return fminf(max_u, max_v);
}

}  // namespace fvmi
