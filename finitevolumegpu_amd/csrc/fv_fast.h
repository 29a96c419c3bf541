// Fast-arithmetic forms of the scheme fluxes (MathFast policy).
//
// Same mathematics as fv_math.h (and the reference expressions it cites), with
// the divisions and square roots regrouped so that each state costs one
// transcendental instruction where the reference needs several:
//   * one v_rsq per state gives 1/h = s*s and sqrt(g*h) = sqrt(g)*h*s;
//   * reciprocals of a pair share one v_rcp of the product;
//   * quotients with a common denominator reuse one reciprocal.
// Rounding differs from the exact policy by a few ulp per operation; the
// results are judged against the float32 oracle at relative L1 <= 1e-5.
#pragma once

#include <hip/hip_runtime.h>

namespace fvmi {
namespace fast {

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Predictor difference F(QL) - F(QR) of the SWE flux (SWECommon.h:29-37), one
// reciprocal for both faces.
__device__ __forceinline__ void swe_pred_diff(const float QL[3], const float QR[3], float hg, float D[3]) {
    const float r = rcp(QL[0] * QR[0]);
    const float uL = QL[1] * (QR[0] * r);
    const float uR = QR[1] * (QL[0] * r);
    D[0] = QL[1] - QR[1];
    D[1] = (QL[1] * uL - QR[1] * uR) + hg * (QL[0] * QL[0] - QR[0] * QR[0]);
    D[2] = QL[2] * uL - QR[2] * uR;
}

// Central-upwind flux (SWECommon.h:228-241)
__device__ __forceinline__ void swe_cu_flux(const float qm[3], const float qp[3], float hg, float sg, float out[3]) {
    const float sp = rsq(qp[0]), sm = rsq(qm[0]);
    const float up = qp[1] * (sp * sp), um = qm[1] * (sm * sm);
    const float cp = sg * qp[0] * sp, cm = sg * qm[0] * sm;
    const float Fp1 = qp[1] * up + hg * qp[0] * qp[0], Fp2 = qp[2] * up;
    const float Fm1 = qm[1] * um + hg * qm[0] * qm[0], Fm2 = qm[2] * um;
    const float am = __builtin_fminf(__builtin_fminf(um - cm, up - cp), 0.0f);
    const float ap = __builtin_fmaxf(__builtin_fmaxf(um + cm, up + cp), 0.0f);
    const float iden = rcp(ap - am);
    const float wm = ap * iden, wp = am * iden, wq = ap * am * iden;
    out[0] = (wm * qm[1] - wp * qp[1]) + wq * (qp[0] - qm[0]);
    out[1] = (wm * Fm1 - wp * Fp1) + wq * (qp[1] - qm[1]);
    out[2] = (wm * Fm2 - wp * Fp2) + wq * (qp[2] - qm[2]);
}

// HLL flux, SWE (SWECommon.h:274-313)
__device__ __forceinline__ void swe_hll_flux(const float ql[3], const float qr[3], float hg, float sg, float out[3]) {
    const float h_l = ql[0], h_r = qr[0];
    const float sl = rsq(h_l), sr = rsq(h_r);
    const float il = sl * sl, ir = sr * sr;
    const float u_l = ql[1] * il, u_r = qr[1] * ir;
    const float c_l = sg * h_l * sl, c_r = sg * h_r * sr;
    const float hs = h_l + h_r;
    const float h_dag = 0.5f * hs - 0.25f * (u_r - u_l) * hs * rcp(c_l + c_r);
    const float q_l_tmp = sqrt(0.5f * (h_dag + h_l) * h_dag) * il;
    const float q_r_tmp = sqrt(0.5f * (h_dag + h_r) * h_dag) * ir;
    const float q_l = (h_dag > h_l) ? q_l_tmp : 1.0f;
    const float q_r = (h_dag > h_r) ? q_r_tmp : 1.0f;
    const float S_l = u_l - c_l * q_l;
    const float S_r = u_r + c_r * q_r;
    const float F_l[3] = {ql[1], ql[1] * u_l + hg * h_l * h_l, ql[2] * u_l};
    const float F_r[3] = {qr[1], qr[1] * u_r + hg * h_r * h_r, qr[2] * u_r};
    const float iden = rcp(S_r - S_l);
    const float wl = S_r * iden, wr = S_l * iden, wq = S_r * S_l * iden;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float star = (wl * F_l[k] - wr * F_r[k]) + wq * (qr[k] - ql[k]);
        out[k] = (S_l >= 0.0f) ? F_l[k] : ((S_r <= 0.0f) ? F_r[k] : star);
    }
}

__device__ __forceinline__ float desing(float x, float eps, float inv_2eps) {
    return __builtin_copysignf(1.0f, x) *
           __builtin_fmaxf(__builtin_fabsf(x), __builtin_fminf(x * x * inv_2eps + 0.5f * eps, eps));
}

__device__ __forceinline__ float gen_minmod(float r, float theta) {
    return __builtin_fmaxf(0.0f, __builtin_fminf(theta * r, __builtin_fminf((1.0f + r) * 0.5f, theta)));
}

// WAF flux (SWECommon.h:131-210); dtdx = dt/dx, inv_g = 1/g
__device__ __forceinline__ void swe_waf_flux(const float q_l2[3], const float q_l1[3], const float q_r1[3],
                                             const float q_r2[3], float hg, float sg, float inv_g, float dtdx,
                                             float out[3]) {
    const float h_l = q_l1[0], h_r = q_r1[0], h_l2 = q_l2[0], h_r2 = q_r2[0];
    const float sl = rsq(h_l), sr = rsq(h_r), sl2 = rsq(h_l2), sr2 = rsq(h_r2);
    const float il = sl * sl, ir = sr * sr, il2 = sl2 * sl2, ir2 = sr2 * sr2;
    const float u_l = q_l1[1] * il, u_r = q_r1[1] * ir, u_l2 = q_l2[1] * il2, u_r2 = q_r2[1] * ir2;
    const float v_l = q_l1[2] * il, v_r = q_r1[2] * ir, v_l2 = q_l2[2] * il2, v_r2 = q_r2[2] * ir2;
    const float c_l = sg * h_l * sl, c_r = sg * h_r * sr, c_l2 = sg * h_l2 * sl2, c_r2 = sg * h_r2 * sr2;
    auto hstar = [&](float ul, float ur, float cl, float cr) {
        const float t = 0.5f * (cl + cr) + 0.25f * (ul - ur);
        return t * t * inv_g;
    };
    const float h_dag_l = hstar(u_l2, u_l, c_l2, c_l);
    const float h_dag = hstar(u_l, u_r, c_l, c_r);
    const float h_dag_r = hstar(u_r, u_r2, c_r, c_r2);
    const float q_l = (h_dag > h_l) ? sqrt(0.5f * ((h_dag + h_l) * h_dag)) * il : 1.0f;
    const float q_r = (h_dag > h_r) ? sqrt(0.5f * ((h_dag + h_r) * h_dag)) * ir : 1.0f;
    const float S_l = u_l - c_l * q_l;
    const float S_r = u_r + c_r * q_r;
    const float dl = h_l * (u_l - S_l), dr = h_r * (u_r - S_r);
    const float S_star = (S_l * dr - S_r * dl) * rcp(dr - dl);
    // shared reciprocal of (S_l - S_star)(S_r - S_star)
    const float el = S_l - S_star, er = S_r - S_star;
    const float ie = rcp(el * er);
    const float s_l = h_l * (S_l - u_l) * (er * ie);
    const float s_r = h_r * (S_r - u_r) * (el * ie);
    const float F_1[3] = {q_l1[1], q_l1[1] * u_l + hg * h_l * h_l, q_l1[2] * u_l};
    const float F_4[3] = {q_r1[1], q_r1[1] * u_r + hg * h_r * h_r, q_r1[2] * u_r};
    const float Qsl[3] = {s_l, s_l * S_star, s_l * v_l};
    const float Qsr[3] = {s_r, s_r * S_star, s_r * v_r};
    const float c_1 = S_l * dtdx, c_2 = S_star * dtdx, c_3 = S_r * dtdx;
    const float eps = 1.0e-6f, inv_2eps = 0.5e6f;
    const float n1 = desing((c_1 > 0.0f) ? (h_dag_l - h_l2) : (h_dag_r - h_r), eps, inv_2eps);
    const float n2 = desing((c_2 > 0.0f) ? (v_l - v_l2) : (v_r2 - v_r), eps, inv_2eps);
    const float n3 = desing((c_3 > 0.0f) ? (h_l - h_dag_l) : (h_r2 - h_dag_r), eps, inv_2eps);
    const float d1 = desing(h_dag - h_l, eps, inv_2eps);
    const float d2 = desing(v_r - v_l, eps, inv_2eps);
    const float d3 = desing(h_r - h_dag, eps, inv_2eps);
    const float r_1 = n1 * rcp(d1), r_2 = n2 * rcp(d2), r_3 = n3 * rcp(d3);
    const float A_1 = __builtin_copysignf(1.0f, c_1) * (1.0f - (1.0f - __builtin_fabsf(c_1)) * gen_minmod(r_1, 1.9f));
    const float A_2 = __builtin_copysignf(1.0f, c_2) * (1.0f - (1.0f - __builtin_fabsf(c_2)) * gen_minmod(r_2, 1.9f));
    const float A_3 = __builtin_copysignf(1.0f, c_3) * (1.0f - (1.0f - __builtin_fabsf(c_3)) * gen_minmod(r_3, 1.9f));
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float F_2 = F_1[k] + S_l * (Qsl[k] - q_l1[k]);
        const float F_3 = F_4[k] + S_r * (Qsr[k] - q_r1[k]);
        out[k] = 0.5f * (F_1[k] + F_4[k]) - 0.5f * (A_1 * (F_2 - F_1[k]) + A_2 * (F_3 - F_2) + A_3 * (F_4[k] - F_3));
    }
}

// 1 / (admissible dt) of one cell, SWE (SWECommon.h:497-504): max over the two
// directions of (|u| + sqrt(g h)) / d.
__device__ __forceinline__ float swe_cell_rate(const float q[3], float sg, float idx, float idy) {
    const float s = rsq(q[0]);
    const float ih = s * s;
    const float c = sg * q[0] * s;
    return __builtin_fmaxf((__builtin_fabsf(q[1] * ih) + c) * idx, (__builtin_fabsf(q[2] * ih) + c) * idy);
}

// ---- Euler (EulerCommon.h) -------------------------------------------------

// Predictor difference F(QL, gamma) - F(QR, gamma) (the reference's predictor
// passes gamma in the pressure slot, EE2D_KP07_dimsplit.cu:57-58).
__device__ __forceinline__ void euler_pred_diff(const float QL[4], const float QR[4], float gamma, float D[4]) {
    const float r = rcp(QL[0] * QR[0]);
    const float uL = QL[1] * (QR[0] * r);
    const float uR = QR[1] * (QL[0] * r);
    D[0] = QL[1] - QR[1];
    D[1] = QL[1] * uL - QR[1] * uR;  // the +gamma terms cancel
    D[2] = QL[2] * uL - QR[2] * uR;
    D[3] = uL * (QL[3] + gamma) - uR * (QR[3] + gamma);
}

struct EulerState {
    float inv, u, P, c;
};

__device__ __forceinline__ EulerState euler_state(const float q[4], float gamma) {
    EulerState s;
    s.inv = rcp(q[0]);
    s.u = q[1] * s.inv;
    s.P = (gamma - 1.0f) * (q[3] - 0.5f * (q[1] * q[1] + q[2] * q[2]) * s.inv);
    s.c = sqrt(gamma * s.P * s.inv);
    return s;
}

// HLL flux, Euler (EulerCommon.h:117-160)
__device__ __forceinline__ void euler_hll_flux(const float ql[4], const float qr[4], float gamma, float out[4]) {
    const EulerState L = euler_state(ql, gamma), R = euler_state(qr, gamma);
    const float h_l = ql[0], h_r = qr[0];
    const float hs = h_l + h_r;
    const float h_dag = 0.5f * hs - 0.25f * (R.u - L.u) * hs * rcp(L.c + R.c);
    const float q_l = (h_dag > h_l) ? sqrt(0.5f * (h_dag + h_l) * h_dag) * L.inv : 1.0f;
    const float q_r = (h_dag > h_r) ? sqrt(0.5f * (h_dag + h_r) * h_dag) * R.inv : 1.0f;
    const float S_l = L.u - L.c * q_l;
    const float S_r = R.u + R.c * q_r;
    const float F_l[4] = {ql[1], ql[1] * L.u + L.P, ql[2] * L.u, L.u * (ql[3] + L.P)};
    const float F_r[4] = {qr[1], qr[1] * R.u + R.P, qr[2] * R.u, R.u * (qr[3] + R.P)};
    const float iden = rcp(S_r - S_l);
    const float wl = S_r * iden, wr = S_l * iden, wq = S_r * S_l * iden;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float star = (wl * F_l[k] - wr * F_r[k]) + wq * (qr[k] - ql[k]);
        out[k] = (S_l >= 0.0f) ? F_l[k] : ((S_r <= 0.0f) ? F_r[k] : star);
    }
}

// Central-upwind flux, Euler (EulerCommon.h:171-187)
__device__ __forceinline__ void euler_cu_flux(const float qm[4], const float qp[4], float gamma, float out[4]) {
    const EulerState m = euler_state(qm, gamma), p = euler_state(qp, gamma);
    const float Fp[4] = {qp[1], qp[1] * p.u + p.P, qp[2] * p.u, p.u * (qp[3] + p.P)};
    const float Fm[4] = {qm[1], qm[1] * m.u + m.P, qm[2] * m.u, m.u * (qm[3] + m.P)};
    const float am = __builtin_fminf(__builtin_fminf(m.u - m.c, p.u - p.c), 0.0f);
    const float ap = __builtin_fmaxf(__builtin_fmaxf(m.u + m.c, p.u + p.c), 0.0f);
    const float iden = rcp(ap - am);
    const float wm = ap * iden, wp = am * iden, wq = ap * am * iden;
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = (wm * Fm[k] - wp * Fp[k]) + wq * (qp[k] - qm[k]);
}

// 1 / (admissible dt) of one cell, Euler (EulerCommon.h:41-49, sqrt(gamma*rho) as written)
__device__ __forceinline__ float euler_cell_rate(const float q[4], float gamma, float idx, float idy) {
    const float s = rsq(gamma * q[0]);
    const float c = gamma * q[0] * s;
    const float ih = gamma * s * s;  // 1/rho
    return __builtin_fmaxf((__builtin_fabsf(q[1] * ih) + c) * idx, (__builtin_fabsf(q[2] * ih) + c) * idy);
}

}  // namespace fast
}  // namespace fvmi
