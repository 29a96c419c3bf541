// Fused dimensionally split substep: ghost fill + first sweep + second sweep
// (+ gravity) + store + CFL minimum, one pass over HBM.
//
// Replaces the tile kernels KP07DimsplitKernel (GPUSimulators/cuda/SWE2D_KP07_dimsplit.cu:123-205),
// HLL2Kernel (SWE2D_HLL2.cu:126-214), WAFKernel (SWE2D_WAF.cu:103-176) and the
// Euler KP07DimsplitKernel (EE2D_KP07_dimsplit.cu:126-247).
//
// Work decomposition (gfx950, 64-lane wavefronts).  A wavefront owns a
// "strip" of output columns x `tile_rows` output rows, C padded columns per
// lane: with C = 1 lane c holds padded column x0+c (x0 = 60*strip), lanes
// 2..61 are the strip's own 60 columns and lanes 0,1,62,63 its x halo; with
// C = 2 lane c holds columns x0+2c, x0+2c+1 (x0 = 124*strip), lanes 1..62 the
// strip's own 124 columns and lanes 0, 63 the halo, each row read and written
// with 8-byte accesses.  The wave marches down the strip one padded row at a
// time (tile_rows+4 rows including the 2+2 y halo):
//   * the x sweep of a row is done across lanes, neighbours exchanged with
//     cross-lane moves, no LDS and no barrier;
//   * the y sweep is done per lane with a register window of the last 3
//     (slope schemes) or 4 (WAF) rows, so every y flux is computed once.
// The reference's "X then Y" (step 0) / "Y then X" (step 1) order, its
// extended first sweep over the halo lines, and its read/write masks are
// preserved exactly (GPUSimulators/cuda/common.h:320-475).
//
// Ghost cells are not materialised: each load maps its padded index through the
// boundary rule (periodic wrap, mirror with sign, or the memory value) so the
// kernel reads the values that readBlock + handleReflectiveBoundary would have
// produced, in global form (correct for every grid size, unlike the tile code's
// nx%bw==1 case).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fv_fast.h"
#include "fv_math.h"

namespace fvmi {

enum : int { BC_DIRICHLET = 0, BC_NEUMANN = 1, BC_PERIODIC = 2, BC_REFLECTIVE = 3, BC_OUTFLOW = 4 };

constexpr int kLanes = 64;

// A wavefront marches a strip of the domain, C padded columns per lane.  The
// lanes at each end that hold the strip's 2-column x halo are not written:
// halo_lanes<C>() of them, leaving strip_cols<C>() output columns (60 for one
// column per lane, 124 for two).
template <int C>
__host__ __device__ constexpr int halo_lanes() {
    return (2 + C - 1) / C;
}
template <int C>
__host__ __device__ constexpr int strip_cols() {
    return (kLanes - 2 * halo_lanes<C>()) * C;
}

struct SubstepArgs {
    int nx, ny;
    float dx, dy, dt, g, theta, gamma;
    int step, bc;
    // interior sub-area [x0,x1) x [y0,y1) to update (0,0,nx,ny = whole domain)
    int x0, y0, x1, y1;
    const float* in[4];
    int in_pitch[4];  // bytes
    float* out[4];
    int out_pitch[4];
    float* cfl;  // two-slot CFL buffer: a step-1 launch minimises slot 1, every launch resets slot 1-step; may be null
    // Work decomposition: the sub-area is cut into strips of strip_cols columns and every
    // strip into `segs_per_strip` segments of `seg_rows` rows (the last one
    // shorter), one per wavefront.  Segment j starts at the same row in every
    // strip, so the waves of neighbouring strips (on the same XCD) read the
    // cache lines they share at about the same time.
    int seg_rows;
    int segs_per_strip;
    int strip_cols;  // output columns per strip (strip_cols<S::kC>() of the scheme launched)
    int strips_x;
    int rows_per_strip;  // y1 - y0
    int num_waves;       // strips_x * segs_per_strip
    int waves_per_block;
    int num_blocks;  // logical blocks = ceil(num_waves / waves_per_block)
    int xcd_remap;
    // Round-weighted plan (round_plan = 1, see round_segment): the grid is
    // kRounds blocks per CU; strips are cut into gps row groups of grows rows
    // (the last one shorter), and part k of a group goes to a wave of dispatch
    // round k.
    int round_plan;
    int xcds, cus_per_xcd;
    int gps, grows;
    // derived constants for the fast policy (set by the host launcher)
    float hg, sqrt_g, inv_g, inv_dx, inv_dy, dtdx, dtdy;
    float half_theta;  // theta / 2
    float inv_2theta;  // 1 / (2 theta), 0 for theta <= 0
    float g_theta;     // g theta
};

// Rows per segment of the equal-segment plan: at most waves_per_simd waves on
// every SIMD (segments per strip rounded down) and no segment shorter than
// min_seg rows, so that a segment's halo rows stay a small part of its work.
// A grid too small for that, where one wave per SIMD would leave each wave at
// most small_limit rows, is bound by the latency of one wave's march instead:
// its strips are cut into segments of that length (at least kMinSegRowsSmall
// rows), at most one wave per SIMD.
constexpr int kMinSegRowsSmall = 4;
__host__ __device__ inline int balanced_segment_rows(int rows, int strips, int cu_count, int waves_per_simd,
                                                     int min_seg, int small_limit) {
    int k = cu_count * 4 * waves_per_simd / strips;
    if (k < 1) k = 1;
    const int seg = (rows + k - 1) / k;
    if (seg >= min_seg) return seg;
    int k1 = cu_count * 4 / strips;
    if (k1 < 1) k1 = 1;
    const int seg1 = (rows + k1 - 1) / k1;
    if (seg1 >= min_seg || seg1 > small_limit) return min_seg;
    return seg1 < kMinSegRowsSmall ? kMinSegRowsSmall : seg1;
}

// the slope constants of the fast policy from g and theta
__host__ __device__ inline void set_slope_constants(SubstepArgs& a) {
    a.half_theta = 0.5f * a.theta;
    a.inv_2theta = a.theta > 0.0f ? 0.5f / a.theta : 0.0f;
    a.g_theta = a.g * a.theta;
}

struct AxisMap {
    int idx;
    bool mirrored;
};

// Memory index read for padded index k along an axis of n interior cells
// (common.h:247-288 and the mirror rules of :133-239).  Indices past the padded
// array are clamped as readBlock does (those lanes/rows are never stored).
// GC ghost cells a side: ghost k < GC mirrors cell 2*GC-1-k (bc*Reflective,
// common.h:133-239, for any gc), periodic wraps by n.
template <int GC>
__device__ __forceinline__ AxisMap map_axis_gc(int k, int n, int bc_lo, int bc_hi) {
    if (k < GC) {
        if (bc_lo == BC_PERIODIC) return {k + n, false};
        if (bc_lo == BC_REFLECTIVE) return {2 * GC - 1 - k, true};
        if (bc_lo == BC_OUTFLOW) return {GC, false};
        return {k, false};
    }
    if (k >= n + GC) {
        if (k > n + 2 * GC - 1) k = n + 2 * GC - 1;
        if (bc_hi == BC_PERIODIC) return {k - n, false};
        if (bc_hi == BC_REFLECTIVE) return {2 * (n + GC) - 1 - k, true};
        if (bc_hi == BC_OUTFLOW) return {n + GC - 1, false};
        return {k, false};
    }
    return {k, false};
}

__device__ __forceinline__ AxisMap map_axis(int k, int n, int bc_lo, int bc_hi) {
    return map_axis_gc<2>(k, n, bc_lo, bc_hi);
}

// The same map along y in affine, branch-free form for the row march (the row
// index is wave-uniform, so the selects stay scalar): a ghost row k < GC reads
// memory row lo_a*k + lo_b, a ghost row k >= n+GC (clamped to the padded
// array) hi_a*k + hi_b, an interior row itself.  A mirroring side flips the
// sign of the y momentum (sign-bit mask).
struct RowMap {
    int lo_a, lo_b, hi_a, hi_b;
    int n_gc;   // n + GC, the first ghost row of the high side
    int k_max;  // n + 2 GC - 1, the last padded row
    unsigned lo_sign, hi_sign;
};

template <int GC>
__device__ __forceinline__ void row_map_side(int bc, int n, bool hi, int& a, int& b, unsigned& sign) {
    a = 1;
    b = 0;
    sign = 0u;
    if (bc == BC_PERIODIC) {
        b = hi ? -n : n;
    } else if (bc == BC_REFLECTIVE) {
        a = -1;
        b = hi ? 2 * (n + GC) - 1 : 2 * GC - 1;
        sign = 0x80000000u;
    } else if (bc == BC_OUTFLOW) {
        a = 0;
        b = hi ? n + GC - 1 : GC;
    }
}

template <int GC>
__device__ __forceinline__ RowMap make_row_map(int n, int bc_lo, int bc_hi) {
    RowMap m;
    row_map_side<GC>(bc_lo, n, false, m.lo_a, m.lo_b, m.lo_sign);
    row_map_side<GC>(bc_hi, n, true, m.hi_a, m.hi_b, m.hi_sign);
    m.n_gc = n + GC;
    m.k_max = n + 2 * GC - 1;
    return m;
}

// memory row of padded row k (wave-uniform) and the y-momentum sign mask
template <int GC>
__device__ __forceinline__ int map_row(const RowMap& m, int k, unsigned& sign) {
    const bool lo = k < GC, hi = k >= m.n_gc;
    // every candidate is computed, then selected (no branches)
    const int lo_row = m.lo_a * k + m.lo_b;
    const int hi_row = m.hi_a * min(k, m.k_max) + m.hi_b;
    int row = hi ? hi_row : k;
    row = lo ? lo_row : row;
    unsigned sg = hi ? m.hi_sign : 0u;
    sg = lo ? m.lo_sign : sg;
    sign = __builtin_amdgcn_readfirstlane(sg);
    return __builtin_amdgcn_readfirstlane(row);
}

// Cross-lane moves: the value held by lane-1 / lane+1, as one DPP row move
// (wave_shr:1 / wave_shl:1).  Lane 0 (resp. 63) receives 0; those lanes are
// halo lanes whose results are never stored.
__device__ __forceinline__ float from_left(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_right(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}

// the WafCell of lane+1
__device__ __forceinline__ WafCell waf_shift_right(const WafCell& p) {
    WafCell r;
    r.h = from_right(p.h);
    r.hu = from_right(p.hu);
    r.hv = from_right(p.hv);
    r.u = from_right(p.u);
    r.v = from_right(p.v);
    r.c = from_right(p.c);
    r.ih = from_right(p.ih);
    return r;
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off, kLanes));
    return v;
}

// Rows loaded ahead of the row being processed.  Each wave keeps a ring of
// row slots in registers: the rows its y stencil still needs plus this many
// rows in flight, so a load is issued kRowPrefetch rows of work before its
// first use.
#ifndef FVMI_ROW_PREFETCH
#define FVMI_ROW_PREFETCH 2
#endif
constexpr int kRowPrefetch = FVMI_ROW_PREFETCH;
static_assert(kRowPrefetch >= 1 && kRowPrefetch <= 3, "row prefetch distance 1..3");

// ---------------------------------------------------------------------------
// Scheme policies
// ---------------------------------------------------------------------------

// Scheme policies.  pred_diff = F(QL) - F(QR) for the half-step predictor;
// flux_x / flux_y the interface fluxes; cell_dt (exact) or cell_rate = 1/dt
// (fast) the CFL quantity of an updated cell.
// kPinRows: the march pins each prefetched row at its first use (use_row);
// kCarryDiff: the y sweep of the fast policy carries one row's forward
// difference over as the next row's backward difference; kPrefetch: rows
// loaded ahead (kRowPrefetch unless a scheme sets its own).  Measured at 4096^2 (SWE) and 8192^2 (Euler):
// pinning takes the WAF substep from 210 to 197 us and HLL2 about 1 % faster,
// KP07_dimsplit even; the Euler kernel is 7 % (carry) and 20-25 % (pinned rows)
// slower with them, so it keeps neither, and 1.5 % faster with one row ahead
// instead of two.
// Columns per lane, rows loaded ahead and row pinning of the KP07_dimsplit and
// HLL2 marches (build-time choices; the defaults are the measured best).
// Measured at 4096^2 per substep, same box: one column, 3 rows ahead, pinned
// rows: KP07_dimsplit 113.0 us, HLL2 132.1; two columns (8-byte accesses):
// 1 row ahead 111.3-112.7 and 156-159 us, 2 rows ahead 132-146 and 153-176 us.
#ifndef FVMI_SWE_COLS
#define FVMI_SWE_COLS 1
#endif
#ifndef FVMI_SWE_PREFETCH
#define FVMI_SWE_PREFETCH 3
#endif
#ifndef FVMI_SWE_PIN
#define FVMI_SWE_PIN 1
#endif

template <class M>
struct SweBase {
    static constexpr int NV = 3;
    static constexpr bool kEuler = false;
    static constexpr bool kPinRows = FVMI_SWE_PIN != 0;
    static constexpr bool kCarryDiff = true;
    static constexpr int kPrefetch = kRowPrefetch;
    __device__ static void pred_diff(const float QL[3], const float QR[3], const SubstepArgs& a, float D[3]) {
        float FL[3], FR[3];
        swe_F<M>(QL, a.g, FL);
        swe_F<M>(QR, a.g, FR);
#pragma unroll
        for (int k = 0; k < 3; ++k) D[k] = FL[k] - FR[k];
    }
    // fast policy: faces QL/QR = c -/+ (theta/2) m of the cell c with limited
    // slope m (in units of theta)
    __device__ static void pred_diff_fast(const float QL[3], const float QR[3], const float c[3], const float m[3],
                                          const SubstepArgs& a, float D[3]) {
        fast::swe_pred_diff(QL, QR, c[0], m[0], a.g_theta, D);
    }
    __device__ static float cell_dt(const float q[3], const SubstepArgs& a) { return swe_cell_dt<M>(q, a.dx, a.dy, a.g); }
    __device__ static float cell_rate(const float q[3], const SubstepArgs& a) {
        return fast::swe_cell_rate(q, a.sqrt_g, a.inv_dx, a.inv_dy);
    }
};

template <class M>
struct SweKP07 : SweBase<M> {
    static constexpr int kPrefetch = FVMI_SWE_PREFETCH;
    static constexpr bool kSlopes = true;
    static constexpr int kC = FVMI_SWE_COLS;  // columns per lane
    __device__ static void flux(const float l[3], const float r[3], const SubstepArgs& a, float o[3]) {
        if constexpr (M::exact)
            swe_cu_flux<M>(l, r, a.g, o);
        else
            fast::swe_cu_flux(l, r, a.hg, a.sqrt_g, o);
    }
    __device__ static void flux_x(const float l[3], const float r[3], const SubstepArgs& a, float o[3]) { flux(l, r, a, o); }
    __device__ static void flux_y(const float l[3], const float r[3], const SubstepArgs& a, float o[3]) { flux(l, r, a, o); }
};

template <class M>
struct SweHLL2 : SweBase<M> {
    static constexpr int kPrefetch = FVMI_SWE_PREFETCH;
    static constexpr bool kSlopes = true;
    static constexpr int kC = FVMI_SWE_COLS;  // columns per lane
    __device__ static void flux(const float l[3], const float r[3], const SubstepArgs& a, float o[3]) {
        if constexpr (M::exact)
            swe_hll_flux<M>(l, r, a.g, o);
        else
            fast::swe_hll_flux(l, r, a.hg, a.sqrt_g, o);
    }
    __device__ static void flux_x(const float l[3], const float r[3], const SubstepArgs& a, float o[3]) { flux(l, r, a, o); }
    __device__ static void flux_y(const float l[3], const float r[3], const SubstepArgs& a, float o[3]) { flux(l, r, a, o); }
};

template <class M>
struct SweWAF : SweBase<M> {
    // the flux is assembled by the march from shared pieces (waf_cell,
    // waf_hstar, waf_flux_core in fv_math.h)
    static constexpr bool kSlopes = false;
    static constexpr int kC = 1;  // columns per lane (its march needs the registers of two)
};

template <class M>
struct EulerKP07 {
    static constexpr int NV = 4;
    static constexpr bool kSlopes = true;
    // columns per lane: two in fast arithmetic (8-byte accesses: 8192^2 step
    // 1.263 -> 1.157 ms); the exact arithmetic's division sequences need the
    // registers of the second column
    static constexpr int kC = M::exact ? 1 : 2;
    static constexpr bool kEuler = true;
    static constexpr bool kPinRows = false;
    static constexpr bool kCarryDiff = false;
    static constexpr int kPrefetch = 1;
    // The predictor passes gamma in the pressure slot: EE2D_KP07_dimsplit.cu:57-58, 102-103.
    __device__ static void pred_diff(const float QL[4], const float QR[4], const SubstepArgs& a, float D[4]) {
        float FL[4], FR[4];
        euler_F<M>(QL, a.gamma, FL);
        euler_F<M>(QR, a.gamma, FR);
#pragma unroll
        for (int k = 0; k < 4; ++k) D[k] = FL[k] - FR[k];
    }
    __device__ static void pred_diff_fast(const float QL[4], const float QR[4], const float c[4], const float m[4],
                                          const SubstepArgs& a, float D[4]) {
        (void)c;
        (void)m;
        fast::euler_pred_diff(QL, QR, a.gamma, D);
    }
    __device__ static void flux_x(const float l[4], const float r[4], const SubstepArgs& a, float o[4]) {
        if constexpr (M::exact)
            euler_hll_flux<M>(l, r, a.gamma, o);
        else
            fast::euler_hll_flux(l, r, a.gamma, o);
    }
    __device__ static void flux_y(const float l[4], const float r[4], const SubstepArgs& a, float o[4]) {
        if constexpr (M::exact)
            euler_cu_flux<M>(l, r, a.gamma, o);
        else
            fast::euler_cu_flux(l, r, a.gamma, o);
    }
    __device__ static float cell_dt(const float q[4], const SubstepArgs& a) {
        return euler_cell_dt<M>(q, a.dx, a.dy, a.gamma);
    }
    __device__ static float cell_rate(const float q[4], const SubstepArgs& a) {
        return fast::euler_cell_rate(q, a.gamma, a.inv_dx, a.inv_dy);
    }
};

// ---------------------------------------------------------------------------
// Sweep pieces
// ---------------------------------------------------------------------------

// y-sweep order: (rho, normal = y momentum, tangential = x momentum[, E])
template <int NV>
__device__ __forceinline__ void to_y(const float q[NV], float o[NV]) {
    o[0] = q[0];
    o[1] = q[2];
    o[2] = q[1];
    if constexpr (NV == 4) o[3] = q[3];
}

// Fast-policy reconstruction + half-step predictor of one cell from its
// backward and forward differences, b(k) = c - l and f(k) = r - c of variable
// k (callables, so that each difference is formed where it is used):
//   minmod(theta b, (b + f)/2, theta f) = theta minmod(b, (b + f)/(2 theta), f)
// for theta > 0, so each neighbour enters once (a cross-lane operand in the x
// sweep) and the y sweep reuses the forward difference of one row as the
// backward difference of the next.
template <class S, class B, class F>
__device__ __forceinline__ void reconstruct_fast(B b, F f, const float c[S::NV], const SubstepArgs& a, float kk,
                                                 float qbl[S::NV], float qbr[S::NV]) {
    constexpr int NV = S::NV;
    float QL[NV], QR[NV], D[NV], m[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const float bk = b(k), fk = f(k);
        m[k] = __builtin_amdgcn_fmed3f(__builtin_amdgcn_fmed3f(bk, fk, 0.0f), (bk + fk) * a.inv_2theta, 0.0f);
        QL[k] = __builtin_fmaf(-a.half_theta, m[k], c[k]);
        QR[k] = __builtin_fmaf(a.half_theta, m[k], c[k]);
    }
    S::pred_diff_fast(QL, QR, c, m, a, D);
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        qbl[k] = QL[k] + kk * D[k];
        qbr[k] = QR[k] + kk * D[k];
    }
}

// Reconstruction + half-step predictor of one cell (computeFluxF, SWE2D_KP07_dimsplit.cu:42-58):
// slope from (l, c, r); returns the predicted left-face (qbl) and right-face (qbr) states.
template <class S, class M>
__device__ __forceinline__ void reconstruct(const float l[S::NV], const float c[S::NV], const float r[S::NV],
                                            const SubstepArgs& a, float kk, float qbl[S::NV], float qbr[S::NV]) {
    constexpr int NV = S::NV;
    if constexpr (M::exact) {
        float QL[NV], QR[NV], D[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const float s = minmod_slope<M>(l[k], c[k], r[k], a.theta);
            QL[k] = c[k] - 0.5f * s;
            QR[k] = c[k] + 0.5f * s;
        }
        S::pred_diff(QL, QR, a, D);
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            qbl[k] = QL[k] + kk * D[k];
            qbr[k] = QR[k] + kk * D[k];
        }
    } else {
        reconstruct_fast<S>([&](int k) { return c[k] - l[k]; }, [&](int k) { return r[k] - c[k]; }, c, a, kk, qbl,
                            qbr);
    }
}

// evolveF/evolveG, common.h:450, 471; dtd = dt/d (fast policy only)
template <class M>
__device__ __forceinline__ float evolve(float q, float f_in, float f_out, float dt, float d, float dtd) {
    if constexpr (M::exact) {
        return q + M::div((f_in - f_out) * dt, d);
    } else {
        return q + (f_in - f_out) * dtd;
    }
}

// Neighbours across a row held C columns per lane: column c of this lane's
// left neighbour is column c - 1 here, or for c = 0 column C - 1 of lane - 1;
// the right one column c + 1, or for c = C - 1 column 0 of lane + 1.
template <int C, int NV>
__device__ __forceinline__ float left_of(const float (&x)[C][NV], int c, int k) {
    return c > 0 ? x[c - 1][k] : from_left(x[C - 1][k]);
}
template <int C, int NV>
__device__ __forceinline__ float right_of(const float (&x)[C][NV], int c, int k) {
    return c < C - 1 ? x[c + 1][k] : from_right(x[0][k]);
}

// x sweep of one padded row held across the 64 lanes, C columns per lane
// (valid on the strip's output columns).
template <class S, class M>
__device__ __forceinline__ void x_sweep_row(float (&q)[S::kC][S::NV], const SubstepArgs& a) {
    constexpr int NV = S::NV;
    constexpr int C = S::kC;
    float F[C][NV];
    if constexpr (S::kSlopes) {
        const float kk = M::exact ? M::div(a.dt, 2.0f * a.dx) : 0.5f * a.dtdx;
        float qbl[C][NV], qbr[C][NV];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            float l[NV], r[NV];
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                l[k] = left_of<C, NV>(q, c, k);
                r[k] = right_of<C, NV>(q, c, k);
            }
            reconstruct<S, M>(l, q[c], r, a, kk, qbl[c], qbr[c]);
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            float qbl_r[NV];
#pragma unroll
            for (int k = 0; k < NV; ++k) qbl_r[k] = right_of<C, NV>(qbl, c, k);
            S::flux_x(qbr[c], qbl_r, a, F[c]);  // interface (c, c+1)
        }
    } else {
        // WAF: each column's cell quantities, and the h* and limiter
        // denominators of its right interface, once; the neighbours' through DPP
        WafCell P[C], R[C];
        float hd[C];
        WafDen D[C];
#pragma unroll
        for (int c = 0; c < C; ++c) P[c] = waf_cell<M>(q[c], a.g, a.sqrt_g);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            R[c] = c < C - 1 ? P[c + 1] : waf_shift_right(P[0]);
            hd[c] = waf_hstar<M>(P[c], R[c], a.g, a.inv_g);
            D[c] = waf_den<M>(P[c], R[c], hd[c]);
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const WafDen Dl = c > 0 ? D[c - 1]
                                    : WafDen{from_left(D[C - 1].d1), from_left(D[C - 1].d2), from_left(D[C - 1].d3)};
            const WafDen Dr = c < C - 1 ? D[c + 1] : WafDen{from_right(D[0].d1), from_right(D[0].d2), from_right(D[0].d3)};
            waf_flux_core<M>(P[c], R[c], Dl, D[c], Dr, hd[c], a.g, a.hg, a.dx, a.dt, a.dtdx, F[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const float Fl = left_of<C, NV>(F, c, k);
            q[c][k] = evolve<M>(q[c][k], Fl, F[c][k], a.dt, a.dx, a.dtdx);
        }
    }
}

// ---------------------------------------------------------------------------
// The kernel
// ---------------------------------------------------------------------------

// Raw buffer descriptor of one padded field (dword3 = 32-bit data format,
// the value used for raw buffers on gfx9).  Reads past `bytes` return 0 and
// writes past it are dropped by the hardware.
//
// Lanes that must not write use the vector offset kMaskedOffset (2 GiB).  The
// host keeps every field below 2 GiB, so that offset is out of range whether
// or not the scalar (row) offset, itself below 2 GiB, is added to it.
constexpr int kMaskedOffset = static_cast<int>(0x80000000u);
__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const void* base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}

typedef unsigned fvmi_v2u __attribute__((ext_vector_type(2)));

// Per-wave constants of a strip of C columns per lane.  Lane L holds padded
// columns wstart + C L + c; halo_lanes<C>() lanes at each end hold the x halo.
//
// Loads (C = 2): one 8-byte load per field and row at pair_off; its words are
// the lane's two columns in order, or reversed (pair_rev) where both columns
// are mirrored ghost columns.  A lane whose two mapped columns are not such a
// pair at an 8-byte offset (zero-gradient ghosts, a periodic or mirrored
// ghost next to an interior column when nx is odd, the columns past the
// domain) loads each column with a 4-byte load instead (fix_off); waves
// without such a lane (wave_fix false) skip those loads.
// Stores (C = 2): an 8-byte store of the two columns where both are written
// and 8-byte aligned (store_off), else 4-byte stores of the written ones
// (single_off; only in waves with such a lane, wave_single).
template <int C>
struct Strip {
    int lane;
    int col[C];          // padded column of each column of this lane
    int row0;            // padded row of r = 0
    int r_end;           // rows r in [0, r_end)
    int lane_off[C];     // byte offset of each column after the boundary map
    unsigned xsign[C];   // sign-bit mask of the x momentum (mirrored ghost column)
    bool lane_out[C];    // column written
    int store_off;       // C = 1: the column, C = 2: the pair (8-byte store); kMaskedOffset if not
    int pair_off;        // C = 2: offset of the 8-byte load (kMaskedOffset for fix lanes)
    bool pair_rev;       // C = 2: the pair's words are (column 1, column 0)
    bool lane_fix;       // C = 2: this lane loads its columns one by one
    bool wave_fix;       // C = 2: some lane of the wave does
    int fix_off[2];      // C = 2: the 4-byte loads of a fix lane (kMaskedOffset elsewhere)
    int single_off[2];   // C = 2: 4-byte stores of written columns not in a full aligned pair
    bool wave_single;    // C = 2: some lane of the wave has such a column
    RowMap ymap;
    __amdgpu_buffer_rsrc_t rin[4], rout[4];
};

// x mirrors flip the x momentum, y mirrors the y momentum
template <int C, int NV>
__device__ __forceinline__ void apply_signs(const Strip<C>& st, unsigned ysign, float (&v)[C][NV]) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
        v[c][1] = __uint_as_float(__float_as_uint(v[c][1]) ^ st.xsign[c]);
        v[c][2] = __uint_as_float(__float_as_uint(v[c][2]) ^ ysign);
    }
}

// Issues the loads of padded row r and returns the sign-bit mask of its y
// momentum.  Pinned: the raw words are returned and use_row applies the mirror
// signs at the row's first use; otherwise they are applied here.
template <int C, int NV, bool Pinned, bool Edge = false>
__device__ __forceinline__ unsigned load_row_raw(const SubstepArgs& a, const Strip<C>& st, int r,
                                                 float (&v)[C][NV]) {
    // Row offsets are wave-uniform (SGPR soffset); the lane adds its column
    // offset (VGPR voffset): no per-lane address arithmetic.
    unsigned ysign;
    const int row = map_row<2>(st.ymap, __builtin_amdgcn_readfirstlane(st.row0 + r), ysign);
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const int soff = row * a.in_pitch[k];
        if constexpr (C == 1) {
            // the b32 buffer builtins move raw 32-bit words
            v[0][k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(st.rin[k], st.lane_off[0], soff, 0));
        } else {
            static_assert(C == 2, "one or two columns per lane");
            const fvmi_v2u x = __builtin_amdgcn_raw_buffer_load_b64(st.rin[k], st.pair_off, soff, 0);
            unsigned w0 = st.pair_rev ? x.y : x.x;
            unsigned w1 = st.pair_rev ? x.x : x.y;
            if constexpr (Edge) {
                const unsigned f0 = __builtin_amdgcn_raw_buffer_load_b32(st.rin[k], st.fix_off[0], soff, 0);
                const unsigned f1 = __builtin_amdgcn_raw_buffer_load_b32(st.rin[k], st.fix_off[1], soff, 0);
                w0 = st.lane_fix ? f0 : w0;
                w1 = st.lane_fix ? f1 : w1;
            }
            v[0][k] = __uint_as_float(w0);
            v[1][k] = __uint_as_float(w1);
        }
    }
    if constexpr (!Pinned) apply_signs<C, NV>(st, ysign, v);
    return ysign;
}

// First use of a row loaded by load_row_raw.  Pinned: an empty volatile
// statement takes the loaded words and hands them on; being volatile it stays
// behind every load and store issued before it in program order, so the
// instruction scheduler cannot hoist the row's consumers (and the wait for its
// data) above the later rows' loads, and each load keeps its distance of
// kRowPrefetch rows of work.  The mirror signs are applied after it.
template <int C, int NV, bool Pinned>
__device__ __forceinline__ void use_row(const Strip<C>& st, unsigned ysign, float (&v)[C][NV]) {
    if constexpr (Pinned) {
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
            for (int k = 0; k < NV; ++k) asm volatile("" : "+v"(v[c][k]));
        apply_signs<C, NV>(st, ysign, v);
    }
}

template <int C, int NV, bool Edge = false>
__device__ __forceinline__ void store_row(const SubstepArgs& a, const Strip<C>& st, int r, const float (&q)[C][NV]) {
    const int prow = __builtin_amdgcn_readfirstlane(st.row0 + r - 2);
    // every lane stores; the halo lanes' offset is out of range and the
    // hardware drops their writes (no branch around the stores)
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const int soff = prow * a.out_pitch[k];
        if constexpr (C == 1) {
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(q[0][k]), st.rout[k], st.store_off, soff, 0);
        } else {
            const fvmi_v2u x = {__float_as_uint(q[0][k]), __float_as_uint(q[1][k])};
            __builtin_amdgcn_raw_buffer_store_b64(x, st.rout[k], st.store_off, soff, 0);
        }
    }
    if constexpr (C == 2 && Edge) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const int soff = prow * a.out_pitch[k];
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(q[0][k]), st.rout[k], st.single_off[0], soff, 0);
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(q[1][k]), st.rout[k], st.single_off[1], soff, 0);
        }
    }
}

// CFL accumulator: the exact policy keeps min(dt) as the reference's writeCfl
// does; the fast policy keeps max(1/dt) (one reciprocal per strip, taken in
// finish_cfl).
template <class S, class M>
__device__ __forceinline__ void accumulate_cfl(const float (&q)[S::kC][S::NV], const SubstepArgs& a,
                                               const Strip<S::kC>& st, float& acc) {
#pragma unroll
    for (int c = 0; c < S::kC; ++c) {
        if constexpr (M::exact) {
            const float cd = S::cell_dt(q[c], a);
            acc = st.lane_out[c] ? fminf(acc, cd) : acc;
        } else {
            const float rate = S::cell_rate(q[c], a);
            acc = st.lane_out[c] ? fmaxf(acc, rate) : acc;
        }
    }
}

// One row step of the y sweep for slope schemes.  Loop-carried state (per
// column): qbt, the predicted top face of row r-2, and g, the flux through its
// bottom interface; both are replaced by the values of row r-1 (in place:
// reads precede writes).
template <class S, class M, bool XFirst, bool Edge>
__device__ __forceinline__ void slope_row(const SubstepArgs& a, const Strip<S::kC>& st, float kky,
                                          const float (&w0)[S::kC][S::NV], const float (&w1)[S::kC][S::NV],
                                          float (&w2)[S::kC][S::NV], float (&qbt)[S::kC][S::NV],
                                          float (&g)[S::kC][S::NV], float (&dq)[S::kC][S::NV], int r, bool emit,
                                          float& dt_min) {
    // w0, w1: rows r-2, r-1 (already x swept when XFirst); w2: raw row r on entry.
    // dq (fast policy): y-order difference row r-1 minus row r-2 on entry, row r
    // minus row r-1 on exit.
    constexpr int NV = S::NV;
    constexpr int C = S::kC;
    if constexpr (XFirst) x_sweep_row<S, M>(w2, a);
    float outq[C][NV];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        float y0[NV], y1[NV], y2[NV], qbb[NV], qbt_next[NV];
        to_y<NV>(w0[c], y0);
        to_y<NV>(w1[c], y1);
        to_y<NV>(w2[c], y2);
        if constexpr (M::exact || !S::kCarryDiff) {
            reconstruct<S, M>(y0, y1, y2, a, kky, qbb, qbt_next);  // faces of row r-1
        } else {
            float f[NV];
#pragma unroll
            for (int k = 0; k < NV; ++k) f[k] = y2[k] - y1[k];
            reconstruct_fast<S>([&](int k) { return dq[c][k]; }, [&](int k) { return f[k]; }, y1, a, kky, qbb,
                                qbt_next);  // faces of row r-1
#pragma unroll
            for (int k = 0; k < NV; ++k) dq[c][k] = f[k];
        }
        float G[NV];
        S::flux_y(qbt[c], qbb, a, G);  // interface (r-2, r-1)
        if (emit) {
#pragma unroll
            for (int k = 0; k < NV; ++k) y0[k] = evolve<M>(y0[k], g[c][k], G[k], a.dt, a.dy, a.dtdy);
            to_y<NV>(y0, outq[c]);  // row r-2 after the y sweep
        }
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            qbt[c][k] = qbt_next[k];
            g[c][k] = G[k];
        }
    }
    if (emit) {
        if constexpr (!XFirst) x_sweep_row<S, M>(outq, a);
        if constexpr (S::kEuler) {
            // gravity source, EE2D_KP07_dimsplit.cu:197-205
            if (a.g > 0.0f) {
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const float rho_v = outq[c][2];
                    outq[c][2] -= a.g * outq[c][0] * a.dt;
                    outq[c][3] -= a.g * rho_v * a.dt;
                }
            }
        }
        store_row<C, NV, Edge>(a, st, r, outq);
        // the CFL minimum is taken after the y-then-x substep only: it is what
        // computeDt reads after a full step (the last launch's values)
        if constexpr (!XFirst) accumulate_cfl<S, M>(outq, a, st, dt_min);
    }
}

// WAF: the y window holds 3 rows (r-2 .. r) and there is no slope state.
// Each ring slot carries its row's WafCells (sweep-y view), computed once when
// the row is ready; hd carries h* of the interface (r-2, r-1) and dd[0], dd[1]
// the limiter denominators of the interfaces (r-3, r-2) and (r-2, r-1); g is
// the flux through the bottom interface of row r-2.  All carried values are
// replaced in place.
template <class S, class M, bool XFirst, bool Edge>
__device__ __forceinline__ void waf_row(const SubstepArgs& a, const Strip<S::kC>& st, const float (&w1)[S::kC][3],
                                        float (&w3)[S::kC][3], const WafCell (&c1)[S::kC],
                                        const WafCell (&c2)[S::kC], WafCell (&c3)[S::kC], float (&hd)[S::kC],
                                        WafDen (&dd)[2][S::kC], float (&g)[S::kC][3], int r, bool emit,
                                        float& dt_min) {
    constexpr int C = S::kC;
    // w1: row r-2 (x swept when XFirst), w3: raw row r on entry
    if constexpr (XFirst) x_sweep_row<S, M>(w3, a);
    float outq[C][3];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        {
            float y3[3];
            to_y<3>(w3[c], y3);
            c3[c] = waf_cell<M>(y3, a.g, a.sqrt_g);
        }
        const float hd_r = waf_hstar<M>(c2[c], c3[c], a.g, a.inv_g);  // interface (r-1, r)
        const WafDen Dn = waf_den<M>(c2[c], c3[c], hd_r);
        float G[3];
        waf_flux_core<M>(c1[c], c2[c], dd[0][c], dd[1][c], Dn, hd[c], a.g, a.hg, a.dy, a.dt, a.dtdy,
                         G);  // interface (r-2, r-1)
        if (emit) {
            float y1[3];
            to_y<3>(w1[c], y1);
#pragma unroll
            for (int k = 0; k < 3; ++k) y1[k] = evolve<M>(y1[k], g[c][k], G[k], a.dt, a.dy, a.dtdy);
            to_y<3>(y1, outq[c]);
        }
        hd[c] = hd_r;
        dd[0][c] = dd[1][c];
        dd[1][c] = Dn;
#pragma unroll
        for (int k = 0; k < 3; ++k) g[c][k] = G[k];
    }
    if (emit) {
        if constexpr (!XFirst) x_sweep_row<S, M>(outq, a);
        store_row<C, 3, Edge>(a, st, r, outq);
        // the CFL minimum is taken after the y-then-x substep only: it is what
        // computeDt reads after a full step (the last launch's values)
        if constexpr (!XFirst) accumulate_cfl<S, M>(outq, a, st, dt_min);
    }
}

// Row march of a slope scheme over one segment.  Row m of the segment lives in
// ring slot m % N; the steady loop is unrolled N times so that every slot index
// is a compile-time constant (the ring stays in registers, nothing is copied).
// Edge: the wave has lanes whose columns need 4-byte loads or stores (see
// Strip); the other waves' marches carry no code for them.
template <class S, class M, bool XFirst, bool Edge>
__device__ __forceinline__ float march_slopes(const SubstepArgs& a, const Strip<S::kC>& st) {
    constexpr int NV = S::NV;
    constexpr int C = S::kC;
    constexpr int PF = S::kPrefetch;
    constexpr int N = PF + 3;  // rows r-2, r-1, r and PF rows in flight
    const float kky = M::exact ? M::div(a.dt, 2.0f * a.dy) : 0.5f * a.dtdy;
    float dt_min = M::exact ? __builtin_inff() : 0.0f;
    float ring[N][C][NV], P[C][NV], G[C][NV], dq[C][NV];
    unsigned ys[N];  // y-momentum sign mask of each slot's row
    const int last = st.r_end - 1;  // r_end >= 5
    // prologue: rows 0 .. N-1; faces of row 1; then step r = 3 (flux (1,2), no output)
#pragma unroll
    for (int m = 0; m < N; ++m) ys[m] = load_row_raw<C, NV, S::kPinRows, Edge>(a, st, min(m, last), ring[m]);
#pragma unroll
    for (int m = 0; m < 3; ++m) use_row<C, NV, S::kPinRows>(st, ys[m], ring[m]);
    if constexpr (XFirst) {
        x_sweep_row<S, M>(ring[0], a);
        x_sweep_row<S, M>(ring[1], a);
        x_sweep_row<S, M>(ring[2], a);
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        float y0[NV], y1[NV], y2[NV], qbb[NV];
        to_y<NV>(ring[0][c], y0);
        to_y<NV>(ring[1][c], y1);
        to_y<NV>(ring[2][c], y2);
        reconstruct<S, M>(y0, y1, y2, a, kky, qbb, P[c]);  // faces of row 1
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            dq[c][k] = y2[k] - y1[k];
            G[c][k] = 0.0f;
        }
    }
    ys[(3 + PF) % N] = load_row_raw<C, NV, S::kPinRows, Edge>(a, st, min(3 + PF, last), ring[(3 + PF) % N]);
    use_row<C, NV, S::kPinRows>(st, ys[3 % N], ring[3 % N]);
    slope_row<S, M, XFirst, Edge>(a, st, kky, ring[1], ring[2], ring[3 % N], P, G, dq, 3, false, dt_min);
    // steady state: r = 4, 4 + N, ...; the step of row rr loads row rr + PF into
    // the slot of row rr - 3, whose last use was the previous step.  Whole
    // groups of N rows run without branches; the remainder (< N rows) keeps
    // the same compile-time slots behind wave-uniform tests.
    auto step = [&](int u, int rr) {
        ys[(4 + u + PF) % N] =
            load_row_raw<C, NV, S::kPinRows, Edge>(a, st, min(rr + PF, last), ring[(4 + u + PF) % N]);
        use_row<C, NV, S::kPinRows>(st, ys[(4 + u) % N], ring[(4 + u) % N]);
        slope_row<S, M, XFirst, Edge>(a, st, kky, ring[(2 + u) % N], ring[(3 + u) % N], ring[(4 + u) % N], P, G, dq, rr,
                                true, dt_min);
    };
    int r = 4;
    for (; r + N <= st.r_end; r += N) {
#pragma unroll
        for (int u = 0; u < N; ++u) step(u, r + u);
    }
#pragma unroll
    for (int u = 0; u < N - 1; ++u)
        if (r + u < st.r_end) step(u, r + u);
    return dt_min;
}

// Row march of WAF (3-row window): row m and its WafCells in ring slot m % N.
template <class S, class M, bool XFirst, bool Edge>
__device__ __forceinline__ float march_waf(const SubstepArgs& a, const Strip<S::kC>& st) {
    constexpr int C = S::kC;
    constexpr int PF = S::kPrefetch;
    constexpr int N = PF + 3;  // rows r-2 .. r and PF rows in flight
    float dt_min = M::exact ? __builtin_inff() : 0.0f;
    float ring[N][C][3], G[C][3], hd[C];
    WafCell cell[N][C];
    WafDen dd[2][C];
    unsigned ys[N];  // y-momentum sign mask of each slot's row
    const int last = st.r_end - 1;  // r_end >= 5
    // prologue: rows 0 .. N-1; interfaces (0, 1) and (1, 2); row 3 + PF; then
    // step r = 3 (flux (1,2), no output)
#pragma unroll
    for (int m = 0; m < N; ++m) ys[m] = load_row_raw<C, 3, S::kPinRows, Edge>(a, st, min(m, last), ring[m]);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        use_row<C, 3, S::kPinRows>(st, ys[m], ring[m]);
        if constexpr (XFirst) x_sweep_row<S, M>(ring[m], a);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            float y[3];
            to_y<3>(ring[m][c], y);
            cell[m][c] = waf_cell<M>(y, a.g, a.sqrt_g);
        }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const float hd01 = waf_hstar<M>(cell[0][c], cell[1][c], a.g, a.inv_g);
        hd[c] = waf_hstar<M>(cell[1][c], cell[2][c], a.g, a.inv_g);
        dd[0][c] = waf_den<M>(cell[0][c], cell[1][c], hd01);
        dd[1][c] = waf_den<M>(cell[1][c], cell[2][c], hd[c]);
#pragma unroll
        for (int k = 0; k < 3; ++k) G[c][k] = 0.0f;
    }
    // row 3 + PF into the slot of row 0, whose last use was above
    ys[(3 + PF) % N] = load_row_raw<C, 3, S::kPinRows, Edge>(a, st, min(3 + PF, last), ring[(3 + PF) % N]);
    use_row<C, 3, S::kPinRows>(st, ys[3], ring[3]);
    waf_row<S, M, XFirst, Edge>(a, st, ring[1], ring[3], cell[1], cell[2], cell[3], hd, dd, G, 3, false, dt_min);
    // steady state: r = 4, 4 + N, ...; the step of row rr loads row rr + PF into
    // the slot of row rr - 3, whose last use was the previous step
    auto step = [&](int u, int rr) {
        ys[(4 + u + PF) % N] = load_row_raw<C, 3, S::kPinRows, Edge>(a, st, min(rr + PF, last), ring[(4 + u + PF) % N]);
        use_row<C, 3, S::kPinRows>(st, ys[(4 + u) % N], ring[(4 + u) % N]);
        waf_row<S, M, XFirst, Edge>(a, st, ring[(2 + u) % N], ring[(4 + u) % N], cell[(2 + u) % N], cell[(3 + u) % N],
                              cell[(4 + u) % N], hd, dd, G, rr, true, dt_min);
    };
    int r = 4;
    for (; r + N <= st.r_end; r += N) {
#pragma unroll
        for (int u = 0; u < N; ++u) step(u, r + u);
    }
#pragma unroll
    for (int u = 0; u < N - 1; ++u)
        if (r + u < st.r_end) step(u, r + u);
    return dt_min;
}

template <class S, class M>
__device__ __forceinline__ void setup_strip(const SubstepArgs& a, Strip<S::kC>& st, int lane, int strip_x, int row_in,
                                            int count) {
    constexpr int C = S::kC;
    constexpr int H = halo_lanes<C>();
    st.lane = lane;
    const int wstart = a.x0 + strip_cols<C>() * strip_x + 2 - H * C;  // padded column of lane 0, column 0
    st.row0 = __builtin_amdgcn_readfirstlane(a.y0 + row_in);         // padded row of r = 0
    st.r_end = __builtin_amdgcn_readfirstlane(count + 4);            // output rows r = 4 .. count+3
    st.ymap = make_row_map<2>(a.ny, (a.bc >> 16) & 0xF, (a.bc >> 24) & 0xF);
#pragma unroll
    for (int c = 0; c < C; ++c) {
        st.col[c] = wstart + C * lane + c;
        const AxisMap mx = map_axis(st.col[c], a.nx, a.bc & 0xF, (a.bc >> 8) & 0xF);
        st.lane_off[c] = mx.idx * (int)sizeof(float);
        st.xsign[c] = mx.mirrored ? 0x80000000u : 0u;
        st.lane_out[c] = lane >= H && lane < kLanes - H && st.col[c] < a.x1 + 2;
    }
    if constexpr (C == 1) {
        st.store_off = st.lane_out[0] ? st.col[0] * (int)sizeof(float) : kMaskedOffset;
    } else {
        const int o0 = st.lane_off[0], o1 = st.lane_off[1];
        const bool fwd = o1 == o0 + 4, rev = o1 == o0 - 4;
        const int base = rev ? o1 : o0;
        st.pair_rev = rev;
        st.lane_fix = !(fwd || rev) || (base & 7) != 0;
        st.pair_off = st.lane_fix ? kMaskedOffset : base;
        st.fix_off[0] = st.lane_fix ? o0 : kMaskedOffset;
        st.fix_off[1] = st.lane_fix ? o1 : kMaskedOffset;
        st.wave_fix = __builtin_amdgcn_ballot_w64(st.lane_fix) != 0;
        const int s0 = st.col[0] * (int)sizeof(float);
        const bool full = st.lane_out[0] && st.lane_out[1] && (s0 & 7) == 0;
        st.store_off = full ? s0 : kMaskedOffset;
        st.single_off[0] = (!full && st.lane_out[0]) ? s0 : kMaskedOffset;
        st.single_off[1] = (!full && st.lane_out[1]) ? s0 + 4 : kMaskedOffset;
        st.wave_single = __builtin_amdgcn_ballot_w64(!full && (st.lane_out[0] || st.lane_out[1])) != 0;
    }
#pragma unroll
    for (int k = 0; k < S::NV; ++k) {
        st.rin[k] = field_rsrc(a.in[k], (unsigned)(a.ny + 4) * (unsigned)a.in_pitch[k]);
        st.rout[k] = field_rsrc(a.out[k], (unsigned)(a.ny + 4) * (unsigned)a.out_pitch[k]);
    }
}

// The march of one segment: sweep order by step, and the edge variant only
// for waves that need it (a wave-uniform choice made once per segment).
template <class S, class M>
__device__ __forceinline__ float march(const SubstepArgs& a, const Strip<S::kC>& st) {
    if constexpr (S::kC == 2) {
        if (st.wave_fix || st.wave_single) {
            if constexpr (S::kSlopes)
                return (a.step == 0) ? march_slopes<S, M, true, true>(a, st) : march_slopes<S, M, false, true>(a, st);
            else
                return (a.step == 0) ? march_waf<S, M, true, true>(a, st) : march_waf<S, M, false, true>(a, st);
        }
    }
    if constexpr (S::kSlopes)
        return (a.step == 0) ? march_slopes<S, M, true, false>(a, st) : march_slopes<S, M, false, false>(a, st);
    else
        return (a.step == 0) ? march_waf<S, M, true, false>(a, st) : march_waf<S, M, false, false>(a, st);
}

// at least this many waves per SIMD must fit (the plan asks for four)
#ifndef FVMI_SUBSTEP_MIN_WAVES
#define FVMI_SUBSTEP_MIN_WAVES 4
#endif

// Round-weighted work split.  Workgroups are dispatched in index order, one
// per CU per round (workgroup p lands on XCD p % xcds, in round
// (p / xcds) / cus_per_xcd), and a SIMD's instruction arbiter favours its
// oldest wave, so with equal segments a SIMD's waves finish one after another
// and the last runs alone.  Here the waves of earlier rounds take longer parts
// of each row group (weights kRoundWeights).  The waves of one workgroup take
// the same rows of adjacent strips, so that each padded row is read as one
// run of waves_per_block * 240 bytes at about the same time (DRAM pages are
// opened once per row run instead of once per strip); consecutive CUs of an
// XCD take consecutive row groups of the same strips.
constexpr int kRounds = 4;
// With the banded mapping, per launch against 13:11:9:7 (same box): KP07_dimsplit
// 4096^2 113.5 / 114.1 us, HLL2 132.8 / 134.1, EE2D 8192^2 step 1.167 / 1.187 ms
// (fast) and 2.125 / 2.156 (exact); 12:11:10:9 and 10:10:10:10 were slower.
#ifndef FVMI_ROUND_WEIGHTS
#define FVMI_ROUND_WEIGHTS 14, 11, 9, 6
#endif
constexpr int kRoundWeights[kRounds] = {FVMI_ROUND_WEIGHTS};
constexpr int round_weight_sum() {
    int sum = 0;
    for (int k = 0; k < kRounds; ++k) sum += kRoundWeights[k];
    return sum;
}
constexpr int kRoundWeightSum = round_weight_sum();

// rows [lo, hi) of part k of a group of n rows
__host__ __device__ inline void round_part(int n, int k, int& lo, int& hi) {
    int c[kRounds + 1];
    c[0] = 0;
    for (int j = 0; j < kRounds - 1; ++j) c[j + 1] = c[j] + (n * kRoundWeights[j]) / kRoundWeightSum;
    c[kRounds] = n;
    lo = c[k];
    hi = c[k + 1];
}

// The segment of physical wave `wave` of workgroup `block` in the
// round-weighted split of `strips` strips of `rows` rows into `gps` groups of
// `grows` rows; false for an idle slot.
__host__ __device__ inline bool round_segment_of(int block, int wave, int xcds, int cus_per_xcd, int waves_per_block,
                                                 int strips, int rows, int gps, int grows, int& strip_x, int& row_in,
                                                 int& count) {
    const int x = block % xcds;
    const int q = block / xcds;
    const int k = q / cus_per_xcd;
    const int c = q - k * cus_per_xcd;
    if (k >= kRounds) return false;
    // CU slot T takes row group T % gps of the strip band T / gps (a band is
    // waves_per_block adjacent strips, one per wave)
    const int T = x * cus_per_xcd + c;
    const int bands = (strips + waves_per_block - 1) / waves_per_block;
    if (T >= bands * gps) return false;
    const int band = T / gps;
    strip_x = band * waves_per_block + wave;
    if (strip_x >= strips) return false;
    const int g0 = (T - band * gps) * grows;
    const int n = min(grows, rows - g0);
    if (n <= 0) return false;
    int lo, hi;
    round_part(n, k, lo, hi);
    row_in = g0 + lo;
    count = hi - lo;
    return count > 0;
}

// Whether the round-weighted split fits: row groups (gps per strip, grows rows)
// whose shortest part has at least `min_part` rows.
__host__ __device__ inline bool round_split_fits(int strips, int rows, int waves_per_block, int cu_count, int min_part,
                                                 int& gps, int& grows) {
    // one 4-wave workgroup per CU and round: a wave on every SIMD
    if (waves_per_block != 4 || cu_count <= 0 || cu_count % 8 != 0 || strips <= 0) return false;
    gps = cu_count / ((strips + waves_per_block - 1) / waves_per_block);
    if (gps <= 0) return false;
    grows = (rows + gps - 1) / gps;
    const int last = rows - (gps - 1) * grows;  // the shortest (last) group
    return (last * kRoundWeights[kRounds - 1]) / kRoundWeightSum >= min_part;
}

__host__ __device__ inline bool round_segment(const SubstepArgs& a, int block, int wave, int& strip_x, int& row_in,
                                              int& count) {
    return round_segment_of(block, wave, a.xcds, a.cus_per_xcd, a.waves_per_block, a.strips_x, a.rows_per_strip,
                            a.gps, a.grows, strip_x, row_in, count);
}

// The segment of physical wave `wave` of workgroup `block` (of `grid`) in the
// equal-segment plan: global wave gw takes segment gw % segs_per_strip of strip
// gw / segs_per_strip.  Hardware dispatch places workgroup i on XCD i % 8; with
// xcd_remap each XCD owns a contiguous range of logical blocks (neighbouring
// strips share L2).  (Giving a workgroup's waves adjacent strips instead, as
// the round-weighted split does, made WAF 19 % slower: 174 -> 207 us per
// 4096^2 substep.)
__host__ __device__ inline bool uniform_segment(const SubstepArgs& a, int block, int grid, int wave, int& strip_x,
                                                int& row_in, int& count) {
    int lb = block;
    if (a.xcd_remap) {
        const int per = grid >> 3;  // grid is a multiple of 8 when remapping
        lb = (block & 7) * per + (block >> 3);
    }
    const int gw = lb * a.waves_per_block + wave;
    if (lb >= a.num_blocks || gw >= a.num_waves) return false;
    strip_x = gw / a.segs_per_strip;
    row_in = (gw - strip_x * a.segs_per_strip) * a.seg_rows;
    count = min(a.seg_rows, a.rows_per_strip - row_in);
    return true;
}

// the segment of a physical wave in either plan
__host__ __device__ inline bool wave_segment(const SubstepArgs& a, int block, int grid, int wave, int& strip_x,
                                             int& row_in, int& count) {
    return a.round_plan ? round_segment(a, block, wave, strip_x, row_in, count)
                        : uniform_segment(a, block, grid, wave, strip_x, row_in, count);
}

template <class S, class M>
__global__ void __launch_bounds__(256, FVMI_SUBSTEP_MIN_WAVES) substep_kernel(SubstepArgs a) {
    const int lane = threadIdx.x & (kLanes - 1);
    // wave index made scalar so that every row address is computed in SGPRs
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (a.cfl != nullptr && blockIdx.x == 0 && threadIdx.x == 0) a.cfl[1 - (a.step & 1)] = __builtin_inff();
    int strip_x, row_in, count;
    if (!wave_segment(a, blockIdx.x, gridDim.x, wave, strip_x, row_in, count)) return;
    strip_x = __builtin_amdgcn_readfirstlane(strip_x);
    row_in = __builtin_amdgcn_readfirstlane(row_in);
    count = __builtin_amdgcn_readfirstlane(count);
    Strip<S::kC> st;
    setup_strip<S, M>(a, st, lane, strip_x, row_in, count);
    const float acc = march<S, M>(a, st);
    float dt_min = acc;
    if constexpr (!M::exact) dt_min = (acc > 0.0f) ? fast::rcp(acc) : __builtin_inff();
    if (a.cfl != nullptr && a.step == 1) {
        const float m = wave_min(dt_min);
        if (lane == 0 && m < __builtin_inff())
            atomicMin(reinterpret_cast<unsigned int*>(a.cfl + 1), __float_as_uint(m));
    }
}

}  // namespace fvmi
