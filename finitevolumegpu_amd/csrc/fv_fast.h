// Fast-arithmetic forms of the scheme fluxes (MathFast policy).
//
// Same mathematics as fv_math.h (and the reference expressions it cites), with
// the divisions and square roots regrouped so that each state costs one
// transcendental instruction where the reference needs several:
//   * one v_rsq per state gives 1/h = s*s and sqrt(g*h) = sqrt(g)*h*s;
//   * reciprocals of a pair share one v_rcp of the product;
//   * quotients with a common denominator reuse one reciprocal.
// Rounding differs from the exact policy by a few ulp per operation; the
// results are held to relative L1 <= 1e-5 of the float32 oracle.
#pragma once

#include <hip/hip_runtime.h>

namespace fvmi {
namespace fast {

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Predictor difference F(QL) - F(QR) of the SWE flux (SWECommon.h:29-37), one
// reciprocal for both faces.  The faces are h0 -/+ (theta/2) m0 in depth, so
// the pressure term g/2 (QL0^2 - QR0^2) = -g theta m0 h0.
__device__ __forceinline__ void swe_pred_diff(const float QL[3], const float QR[3], float h0, float m0, float g_theta,
                                              float D[3]) {
    const float r = rcp(QL[0] * QR[0]);
    const float uL = QL[1] * (QR[0] * r);
    const float uR = QR[1] * (QL[0] * r);
    D[0] = QL[1] - QR[1];
    D[1] = (QL[1] * uL - QR[1] * uR) - g_theta * (m0 * h0);
    D[2] = QL[2] * uL - QR[2] * uR;
}

// Weights of the HLL flux F = wl F_l - wr F_r + wq (q_r - q_l) with the
// reference's upwind selection (SWECommon.h:299-311, EulerCommon.h:146-159:
// F_l if S_l >= 0, else F_r if S_r <= 0, else the HLL average) folded into
// clamped speeds sl = min(S_l, 0) and sr = max(S_r, S_l, 1e-30):
//   wl = sr / (sr - sl), wr = sl / (sr - sl), wq = sr sl / (sr - sl).
// S_l >= 0 gives sl = 0 and sr > 0: F_l (also when S_r <= 0 <= S_l, where the
// reference tests S_l first).  S_l < 0 and S_r <= 0 give sr = 1e-30: F_r up to
// a relative 1e-30 / |S_l|.  Otherwise sl = S_l, sr = S_r: the HLL average.
__device__ __forceinline__ void hll_weights(float S_l, float S_r, float& wl, float& wr, float& wq) {
    const float sl = __builtin_fminf(S_l, 0.0f);
    const float sr = __builtin_fmaxf(__builtin_fmaxf(S_r, S_l), 1e-30f);
    const float iden = rcp(sr - sl);
    wl = sr * iden;
    wr = sl * iden;
    wq = sr * sl * iden;
}

// Central-upwind flux (SWECommon.h:228-241),
//   (ap F(qm) - am F(qp) + ap am (qp - qm)) / (ap - am)
//     = wm (F(qm) - am qm) - wp (F(qp) - ap qp),  wm = ap/(ap-am), wp = am/(ap-am),
// where F(q) - a q = (q1 - a q0, q1 (u - a) + g/2 q0^2, q2 (u - a)).
__device__ __forceinline__ void swe_cu_flux(const float qm[3], const float qp[3], float hg, float sg, float out[3]) {
    const float sp = rsq(qp[0]), sm = rsq(qm[0]);
    const float up = qp[1] * (sp * sp), um = qm[1] * (sm * sm);
    const float cp = sg * qp[0] * sp, cm = sg * qm[0] * sm;
    const float am = __builtin_fminf(__builtin_fminf(um - cm, up - cp), 0.0f);
    const float ap = __builtin_fmaxf(__builtin_fmaxf(um + cm, up + cp), 0.0f);
    const float iden = rcp(ap - am);
    const float wm = ap * iden, wp = am * iden;
    const float dm = um - am, dp = up - ap;
    const float A0 = qm[1] - am * qm[0], B0 = qp[1] - ap * qp[0];
    const float A1 = qm[1] * dm + hg * qm[0] * qm[0], B1 = qp[1] * dp + hg * qp[0] * qp[0];
    const float A2 = qm[2] * dm, B2 = qp[2] * dp;
    out[0] = wm * A0 - wp * B0;
    out[1] = wm * A1 - wp * B1;
    out[2] = wm * A2 - wp * B2;
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
    float wl, wr, wq;
    hll_weights(S_l, S_r, wl, wr, wq);
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = (wl * F_l[k] - wr * F_r[k]) + wq * (qr[k] - ql[k]);
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
    float wl, wr, wq;
    hll_weights(S_l, S_r, wl, wr, wq);
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = (wl * F_l[k] - wr * F_r[k]) + wq * (qr[k] - ql[k]);
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
