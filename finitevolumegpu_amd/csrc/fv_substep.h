// Fused dimensionally split substep: ghost fill + first sweep + second sweep
// (+ gravity) + store + CFL minimum, one pass over HBM.
//
// Replaces the tile kernels KP07DimsplitKernel (GPUSimulators/cuda/SWE2D_KP07_dimsplit.cu:123-205),
// HLL2Kernel (SWE2D_HLL2.cu:126-214), WAFKernel (SWE2D_WAF.cu:103-176) and the
// Euler KP07DimsplitKernel (EE2D_KP07_dimsplit.cu:126-247).
//
// Work decomposition (gfx950, 64-lane wavefronts).  A wavefront owns a
// "strip": 60 output columns x `tile_rows` output rows.  Lane c holds padded
// column x0+c (x0 = 60*strip), so lanes 2..61 are the strip's own columns and
// lanes 0,1,62,63 its x halo.  The wave marches down the strip one padded row
// at a time (tile_rows+4 rows including the 2+2 y halo):
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
constexpr int kStripCols = 60;  // output columns per wavefront

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
    // Work decomposition: the sub-area is cut into 60-column strips and every
    // strip into `segs_per_strip` segments of `seg_rows` rows (the last one
    // shorter), one per wavefront.  Segment j starts at the same row in every
    // strip, so the waves of neighbouring strips (on the same XCD) read the
    // cache lines they share at about the same time.
    int seg_rows;
    int segs_per_strip;
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
template <class M>
struct SweBase {
    static constexpr int NV = 3;
    static constexpr bool kEuler = false;
    static constexpr bool kPinRows = true;
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
    static constexpr int kPrefetch = 3;  // 0.7 % faster than 2 at 4096^2
    static constexpr bool kSlopes = true;
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
    static constexpr int kPrefetch = 3;  // 1.1 % faster than 2 at 4096^2
    static constexpr bool kSlopes = true;
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
};

template <class M>
struct EulerKP07 {
    static constexpr int NV = 4;
    static constexpr bool kSlopes = true;
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
        return q + (f_in - f_out) * dt / d;
    } else {
        return q + (f_in - f_out) * dtd;
    }
}

// x sweep of one padded row held across the 64 lanes (valid on lanes 2..61).
template <class S, class M>
__device__ __forceinline__ void x_sweep_row(float q[S::NV], const SubstepArgs& a) {
    constexpr int NV = S::NV;
    float F[NV];
    if constexpr (S::kSlopes) {
        float l[NV], r[NV], qbl[NV], qbr[NV], qbl_r[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            l[k] = from_left(q[k]);
            r[k] = from_right(q[k]);
        }
        const float kk = M::exact ? a.dt / (2.0f * a.dx) : 0.5f * a.dtdx;
        reconstruct<S, M>(l, q, r, a, kk, qbl, qbr);
#pragma unroll
        for (int k = 0; k < NV; ++k) qbl_r[k] = from_right(qbl[k]);
        S::flux_x(qbr, qbl_r, a, F);  // interface (c, c+1)
    } else {
        // WAF: each lane's cell quantities, and the h* and limiter denominators
        // of its right interface, once; the neighbours' through DPP
        const WafCell P = waf_cell<M>(q, a.g, a.sqrt_g);
        const WafCell R = waf_shift_right(P);
        const float hd = waf_hstar<M>(P, R, a.g, a.inv_g);
        const WafDen D = waf_den<M>(P, R, hd);
        const WafDen Dl = {from_left(D.d1), from_left(D.d2), from_left(D.d3)};
        const WafDen Dr = {from_right(D.d1), from_right(D.d2), from_right(D.d3)};
        waf_flux_core<M>(P, R, Dl, D, Dr, hd, a.g, a.hg, a.dx, a.dt, a.dtdx, F);
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const float Fl = from_left(F[k]);
        q[k] = evolve<M>(q[k], Fl, F[k], a.dt, a.dx, a.dtdx);
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

// Per-wave constants of a strip.
struct Strip {
    int lane;
    int col;       // padded column of this lane
    int row0;      // padded row of r = 0
    int r_end;     // rows r in [0, r_end)
    int lane_off;    // byte offset of this lane's column in a row (after the boundary map)
    unsigned xsign;  // sign-bit mask of the x momentum (mirrored ghost column)
    bool lane_out;   // lane writes its column
    int store_off;   // byte offset of the column for writing lanes, kMaskedOffset for the others
    RowMap ymap;
    __amdgpu_buffer_rsrc_t rin[4], rout[4];
};

// x mirrors flip the x momentum, y mirrors the y momentum
template <int NV>
__device__ __forceinline__ void apply_signs(const Strip& st, unsigned ysign, float* v) {
    v[1] = __uint_as_float(__float_as_uint(v[1]) ^ st.xsign);
    v[2] = __uint_as_float(__float_as_uint(v[2]) ^ ysign);
}

// Issues the loads of padded row r and returns the sign-bit mask of its y
// momentum.  Pinned: the raw words are returned and use_row applies the mirror
// signs at the row's first use; otherwise they are applied here.
template <int NV, bool Pinned>
__device__ __forceinline__ unsigned load_row_raw(const SubstepArgs& a, const Strip& st, int r, float v[NV]) {
    // Row offsets are wave-uniform (SGPR soffset); the lane adds its column
    // offset (VGPR voffset): no per-lane address arithmetic.
    unsigned ysign;
    const int row = map_row<2>(st.ymap, __builtin_amdgcn_readfirstlane(st.row0 + r), ysign);
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        // the b32 buffer builtins move raw 32-bit words
        unsigned x = __builtin_amdgcn_raw_buffer_load_b32(st.rin[k], st.lane_off, row * a.in_pitch[k], 0);
        if constexpr (!Pinned) {
            if (k == 1) x ^= st.xsign;
            if (k == 2) x ^= ysign;
        }
        v[k] = __uint_as_float(x);
    }
    return ysign;
}

// First use of a row loaded by load_row_raw.  Pinned: an empty volatile
// statement takes the loaded words and hands them on; being volatile it stays
// behind every load and store issued before it in program order, so the
// instruction scheduler cannot hoist the row's consumers (and the wait for its
// data) above the later rows' loads, and each load keeps its distance of
// kRowPrefetch rows of work.  The mirror signs are applied after it.
template <int NV, bool Pinned>
__device__ __forceinline__ void use_row(const Strip& st, unsigned ysign, float (&v)[NV]) {
    if constexpr (Pinned) {
#pragma unroll
        for (int k = 0; k < NV; ++k) asm volatile("" : "+v"(v[k]));
        apply_signs<NV>(st, ysign, v);
    }
}

template <int NV>
__device__ __forceinline__ void store_row(const SubstepArgs& a, const Strip& st, int r, const float (&q)[NV]) {
    const int prow = __builtin_amdgcn_readfirstlane(st.row0 + r - 2);
    // every lane stores; the halo lanes' offset is out of range and the
    // hardware drops their writes (no branch around the stores)
#pragma unroll
    for (int k = 0; k < NV; ++k)
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(q[k]), st.rout[k], st.store_off, prow * a.out_pitch[k], 0);
}

// CFL accumulator: the exact policy keeps min(dt) as the reference's writeCfl
// does; the fast policy keeps max(1/dt) (one reciprocal per strip, taken in
// finish_cfl).
template <class S, class M>
__device__ __forceinline__ void accumulate_cfl(const float (&q)[S::NV], const SubstepArgs& a, bool take, float& acc) {
    if constexpr (M::exact) {
        const float cd = S::cell_dt(q, a);
        acc = take ? fminf(acc, cd) : acc;
    } else {
        const float rate = S::cell_rate(q, a);
        acc = take ? fmaxf(acc, rate) : acc;
    }
}

// One row step of the y sweep for slope schemes.  Loop-carried state: qbt, the
// predicted top face of row r-2, and g, the flux through its bottom interface;
// both are replaced by the values of row r-1 (in place: reads precede writes).
template <class S, class M, bool XFirst>
__device__ __forceinline__ void slope_row(const SubstepArgs& a, const Strip& st, float kky,
                                          const float (&w0)[S::NV], const float (&w1)[S::NV], float (&w2)[S::NV],
                                          float (&qbt)[S::NV], float (&g)[S::NV], float (&dq)[S::NV], int r,
                                          bool emit, float& dt_min) {
    // w0, w1: rows r-2, r-1 (already x swept when XFirst); w2: raw row r on entry.
    // dq (fast policy): y-order difference row r-1 minus row r-2 on entry, row r
    // minus row r-1 on exit.
    constexpr int NV = S::NV;
    if constexpr (XFirst) x_sweep_row<S, M>(w2, a);
    float y0[NV], y1[NV], y2[NV], qbb[NV], qbt_next[NV];
    to_y<NV>(w0, y0);
    to_y<NV>(w1, y1);
    to_y<NV>(w2, y2);
    if constexpr (M::exact || !S::kCarryDiff) {
        reconstruct<S, M>(y0, y1, y2, a, kky, qbb, qbt_next);  // faces of row r-1
    } else {
        float f[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) f[k] = y2[k] - y1[k];
        reconstruct_fast<S>([&](int k) { return dq[k]; }, [&](int k) { return f[k]; }, y1, a, kky, qbb,
                            qbt_next);  // faces of row r-1
#pragma unroll
        for (int k = 0; k < NV; ++k) dq[k] = f[k];
    }
    float G[NV];
    S::flux_y(qbt, qbb, a, G);  // interface (r-2, r-1)
    if (emit) {
        float outq[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) y0[k] = evolve<M>(y0[k], g[k], G[k], a.dt, a.dy, a.dtdy);
        to_y<NV>(y0, outq);  // row r-2 after the y sweep
        if constexpr (!XFirst) x_sweep_row<S, M>(outq, a);
        if constexpr (S::kEuler) {
            // gravity source, EE2D_KP07_dimsplit.cu:197-205
            if (a.g > 0.0f) {
                const float rho_v = outq[2];
                outq[2] -= a.g * outq[0] * a.dt;
                outq[3] -= a.g * rho_v * a.dt;
            }
        }
        store_row<NV>(a, st, r, outq);
        // the CFL minimum is taken after the y-then-x substep only: it is what
        // computeDt reads after a full step (the last launch's values)
        if constexpr (!XFirst) accumulate_cfl<S, M>(outq, a, st.lane_out, dt_min);
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        qbt[k] = qbt_next[k];
        g[k] = G[k];
    }
}

// WAF: the y window holds 3 rows (r-2 .. r) and there is no slope state.
// Each ring slot carries its row's WafCell (sweep-y view), computed once when
// the row is ready; hd carries h* of the interface (r-2, r-1) and dd[0], dd[1]
// the limiter denominators of the interfaces (r-3, r-2) and (r-2, r-1); g is
// the flux through the bottom interface of row r-2.  All carried values are
// replaced in place.
template <class S, class M, bool XFirst>
__device__ __forceinline__ void waf_row(const SubstepArgs& a, const Strip& st, const float (&w1)[3],
                                        float (&w3)[3], const WafCell& c1, const WafCell& c2, WafCell& c3,
                                        float& hd, WafDen (&dd)[2], float (&g)[3], int r, bool emit, float& dt_min) {
    // w1: row r-2 (x swept when XFirst), w3: raw row r on entry
    if constexpr (XFirst) x_sweep_row<S, M>(w3, a);
    {
        float y3[3];
        to_y<3>(w3, y3);
        c3 = waf_cell<M>(y3, a.g, a.sqrt_g);
    }
    const float hd_r = waf_hstar<M>(c2, c3, a.g, a.inv_g);  // interface (r-1, r)
    const WafDen Dn = waf_den<M>(c2, c3, hd_r);
    float G[3];
    waf_flux_core<M>(c1, c2, dd[0], dd[1], Dn, hd, a.g, a.hg, a.dy, a.dt, a.dtdy, G);  // interface (r-2, r-1)
    if (emit) {
        float y1[3], outq[3];
        to_y<3>(w1, y1);
#pragma unroll
        for (int k = 0; k < 3; ++k) y1[k] = evolve<M>(y1[k], g[k], G[k], a.dt, a.dy, a.dtdy);
        to_y<3>(y1, outq);
        if constexpr (!XFirst) x_sweep_row<S, M>(outq, a);
        store_row<3>(a, st, r, outq);
        // the CFL minimum is taken after the y-then-x substep only: it is what
        // computeDt reads after a full step (the last launch's values)
        if constexpr (!XFirst) accumulate_cfl<S, M>(outq, a, st.lane_out, dt_min);
    }
    hd = hd_r;
    dd[0] = dd[1];
    dd[1] = Dn;
#pragma unroll
    for (int k = 0; k < 3; ++k) g[k] = G[k];
}

template <int NV>
__device__ __forceinline__ void copyv(float (&d)[NV], const float (&s)[NV]) {
#pragma unroll
    for (int k = 0; k < NV; ++k) d[k] = s[k];
}

// Row march of a slope scheme over one segment.  Row m of the segment lives in
// ring slot m % N; the steady loop is unrolled N times so that every slot index
// is a compile-time constant (the ring stays in registers, nothing is copied).
template <class S, class M, bool XFirst>
__device__ __forceinline__ float march_slopes(const SubstepArgs& a, const Strip& st) {
    constexpr int NV = S::NV;
    constexpr int PF = S::kPrefetch;
    constexpr int N = PF + 3;  // rows r-2, r-1, r and PF rows in flight
    const float kky = M::exact ? a.dt / (2.0f * a.dy) : 0.5f * a.dtdy;
    float dt_min = M::exact ? __builtin_inff() : 0.0f;
    float ring[N][NV], P[NV], G[NV], dq[NV];
    unsigned ys[N];  // y-momentum sign mask of each slot's row
    const int last = st.r_end - 1;  // r_end >= 5
    // prologue: rows 0 .. N-1; faces of row 1; then step r = 3 (flux (1,2), no output)
#pragma unroll
    for (int m = 0; m < N; ++m) ys[m] = load_row_raw<NV, S::kPinRows>(a, st, min(m, last), ring[m]);
#pragma unroll
    for (int m = 0; m < 3; ++m) use_row<NV, S::kPinRows>(st, ys[m], ring[m]);
    if constexpr (XFirst) {
        x_sweep_row<S, M>(ring[0], a);
        x_sweep_row<S, M>(ring[1], a);
        x_sweep_row<S, M>(ring[2], a);
    }
    {
        float y0[NV], y1[NV], y2[NV], qbb[NV];
        to_y<NV>(ring[0], y0);
        to_y<NV>(ring[1], y1);
        to_y<NV>(ring[2], y2);
        reconstruct<S, M>(y0, y1, y2, a, kky, qbb, P);  // faces of row 1
#pragma unroll
        for (int k = 0; k < NV; ++k) dq[k] = y2[k] - y1[k];
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) G[k] = 0.0f;
    ys[(3 + PF) % N] = load_row_raw<NV, S::kPinRows>(a, st, min(3 + PF, last), ring[(3 + PF) % N]);
    use_row<NV, S::kPinRows>(st, ys[3 % N], ring[3 % N]);
    slope_row<S, M, XFirst>(a, st, kky, ring[1], ring[2], ring[3 % N], P, G, dq, 3, false, dt_min);
    // steady state: r = 4, 4 + N, ...; the step of row rr loads row rr + PF into
    // the slot of row rr - 3, whose last use was the previous step.  Whole
    // groups of N rows run without branches; the remainder (< N rows) keeps
    // the same compile-time slots behind wave-uniform tests.
    auto step = [&](int u, int rr) {
        ys[(4 + u + PF) % N] = load_row_raw<NV, S::kPinRows>(a, st, min(rr + PF, last), ring[(4 + u + PF) % N]);
        use_row<NV, S::kPinRows>(st, ys[(4 + u) % N], ring[(4 + u) % N]);
        slope_row<S, M, XFirst>(a, st, kky, ring[(2 + u) % N], ring[(3 + u) % N], ring[(4 + u) % N], P, G, dq, rr,
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

// Row march of WAF (3-row window): row m and its WafCell in ring slot m % N.
template <class S, class M, bool XFirst>
__device__ __forceinline__ float march_waf(const SubstepArgs& a, const Strip& st) {
    constexpr int PF = S::kPrefetch;
    constexpr int N = PF + 3;  // rows r-2 .. r and PF rows in flight
    float dt_min = M::exact ? __builtin_inff() : 0.0f;
    float ring[N][3], G[3], hd;
    WafCell cell[N];
    WafDen dd[2];
    unsigned ys[N];  // y-momentum sign mask of each slot's row
    const int last = st.r_end - 1;  // r_end >= 5
    // prologue: rows 0 .. N-1; interfaces (0, 1) and (1, 2); row 3 + PF; then
    // step r = 3 (flux (1,2), no output)
#pragma unroll
    for (int m = 0; m < N; ++m) ys[m] = load_row_raw<3, S::kPinRows>(a, st, min(m, last), ring[m]);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        use_row<3, S::kPinRows>(st, ys[m], ring[m]);
        if constexpr (XFirst) x_sweep_row<S, M>(ring[m], a);
        float y[3];
        to_y<3>(ring[m], y);
        cell[m] = waf_cell<M>(y, a.g, a.sqrt_g);
    }
    {
        const float hd01 = waf_hstar<M>(cell[0], cell[1], a.g, a.inv_g);
        hd = waf_hstar<M>(cell[1], cell[2], a.g, a.inv_g);
        dd[0] = waf_den<M>(cell[0], cell[1], hd01);
        dd[1] = waf_den<M>(cell[1], cell[2], hd);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) G[k] = 0.0f;
    // row 3 + PF into the slot of row 0, whose last use was above
    ys[(3 + PF) % N] = load_row_raw<3, S::kPinRows>(a, st, min(3 + PF, last), ring[(3 + PF) % N]);
    use_row<3, S::kPinRows>(st, ys[3], ring[3]);
    waf_row<S, M, XFirst>(a, st, ring[1], ring[3], cell[1], cell[2], cell[3], hd, dd, G, 3, false, dt_min);
    // steady state: r = 4, 4 + N, ...; the step of row rr loads row rr + PF into
    // the slot of row rr - 3, whose last use was the previous step
    auto step = [&](int u, int rr) {
        ys[(4 + u + PF) % N] = load_row_raw<3, S::kPinRows>(a, st, min(rr + PF, last), ring[(4 + u + PF) % N]);
        use_row<3, S::kPinRows>(st, ys[(4 + u) % N], ring[(4 + u) % N]);
        waf_row<S, M, XFirst>(a, st, ring[(2 + u) % N], ring[(4 + u) % N], cell[(2 + u) % N], cell[(3 + u) % N],
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
__device__ __forceinline__ void setup_strip(const SubstepArgs& a, Strip& st, int lane, int strip_x, int row_in,
                                            int count) {
    st.lane = lane;
    st.col = a.x0 + kStripCols * strip_x + lane;
    st.row0 = __builtin_amdgcn_readfirstlane(a.y0 + row_in);  // padded row of r = 0
    st.r_end = __builtin_amdgcn_readfirstlane(count + 4);     // output rows r = 4 .. count+3
    st.ymap = make_row_map<2>(a.ny, (a.bc >> 16) & 0xF, (a.bc >> 24) & 0xF);
    const AxisMap mx = map_axis(st.col, a.nx, a.bc & 0xF, (a.bc >> 8) & 0xF);
    st.lane_off = mx.idx * (int)sizeof(float);
    st.xsign = mx.mirrored ? 0x80000000u : 0u;
    st.lane_out = lane >= 2 && lane < 2 + kStripCols && st.col < a.x1 + 2;
    st.store_off = st.lane_out ? st.col * (int)sizeof(float) : kMaskedOffset;
#pragma unroll
    for (int k = 0; k < S::NV; ++k) {
        st.rin[k] = field_rsrc(a.in[k], (unsigned)(a.ny + 4) * (unsigned)a.in_pitch[k]);
        st.rout[k] = field_rsrc(a.out[k], (unsigned)(a.ny + 4) * (unsigned)a.out_pitch[k]);
    }
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
constexpr int kRoundWeights[kRounds] = {13, 11, 9, 7};
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
    Strip st;
    setup_strip<S, M>(a, st, lane, strip_x, row_in, count);
    float acc;
    if constexpr (S::kSlopes) {
        acc = (a.step == 0) ? march_slopes<S, M, true>(a, st) : march_slopes<S, M, false>(a, st);
    } else {
        acc = (a.step == 0) ? march_waf<S, M, true>(a, st) : march_waf<S, M, false>(a, st);
    }
    float dt_min = acc;
    if constexpr (!M::exact) dt_min = (acc > 0.0f) ? fast::rcp(acc) : __builtin_inff();
    if (a.cfl != nullptr && a.step == 1) {
        const float m = wave_min(dt_min);
        if (lane == 0 && m < __builtin_inff())
            atomicMin(reinterpret_cast<unsigned int*>(a.cfl + 1), __float_as_uint(m));
    }
}

}  // namespace fvmi
