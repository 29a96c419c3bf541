// The reference's other shallow-water schemes: Lax-Friedrichs, FORCE and
// first-order HLL (one ghost cell) and the unsplit second-order
// Kurganov-Petrova scheme with TVD RK2 (two ghost cells).
//
//   LxF   GPUSimulators/cuda/SWE2D_LxF.cu:93-166   x and y fluxes of the same
//         state (LxF_2D_flux, SWECommon.h:421-427), one update
//   HLL   SWE2D_HLL.cu:100-160, FORCE  SWE2D_FORCE.cu:88-140   first order,
//         x sweep of every row (halo rows included, evolveF common.h:443-456)
//         then y sweep of that state (evolveG :464-475), in one launch
//   KP07  SWE2D_KP07.cu:120-230   minmod slopes in x and y of the same state,
//         face values without predictor, central-upwind fluxes, one update;
//         the second stage of RK2 writes 0.5 (out_old + Q) (:205-214)
//
// Work split and addressing as the dimensionally split kernel (fv_substep.h):
// a wavefront owns a strip of 64 - 2*GC columns (lane c = padded column
// 62*s + c for GC = 1, 60*s + c for GC = 2; the GC lanes at each edge are the
// x halo) and marches down a segment of rows; x neighbours come through DPP
// row moves, y neighbours from a register ring of rows.
#pragma once

#include "fv_substep.h"

namespace fvmi {

enum UnsplitKind : int { KIND_LXF = 0, KIND_SPLIT1 = 1, KIND_KP07U = 2 };

struct UnsplitArgs {
    int nx, ny;
    float dx, dy, dt, g, theta;
    int bc;
    int stage2;    // KP07 with order 2, second stage: out = 0.5 (out_old + Q)
    int cfl_slot;  // this launch minimises cfl[cfl_slot] and resets the other slot
    const float* in[3];
    int in_pitch[3];
    float* out[3];
    int out_pitch[3];
    float* cfl;  // two floats, may be null
    int seg_rows;
    int segs_per_strip;
    int strips_x;
    int num_waves;
    int waves_per_block;
    int num_blocks;
    int xcd_remap;
    // round-weighted split (round_plan = 1; round_segment_of in fv_substep.h)
    int round_plan;
    int xcds, cus_per_xcd;
    int gps, grows;
    // fast-policy constants (host): 0.5 g, sqrt(g), dt/dx, dt/dy and the
    // centred fluxes' coefficients dx/(4dt), dy/(4dt) (LxF), dx/(2dt),
    // dy/(2dt), dt/(2dx), dt/(2dy) (FORCE)
    float hg, sqrt_g, inv_dx, inv_dy, dtdx, dtdy;
    float lxf_kx, lxf_ky, lf_kx, lf_ky, lw_kx, lw_ky;
    float half_theta, inv_2theta;  // theta/2 and 1/(2 theta) (0 for theta <= 0), KP07
};

// Half the limited slope, s/2, of a cell with neighbours l and r.  Fast form:
// minmod(theta b, (b + f)/2, theta f) = theta minmod(b, (b + f)/(2 theta), f)
// for theta > 0 (b, f the backward and forward differences).
template <class M>
__device__ __forceinline__ float half_slope(float l, float c, float r, const UnsplitArgs& a) {
    if constexpr (M::exact) {
        return 0.5f * minmod_slope<M>(l, c, r, a.theta);
    } else {
        const float b = c - l, f = r - c;
        return a.half_theta *
               __builtin_amdgcn_fmed3f(__builtin_amdgcn_fmed3f(b, f, 0.0f), (b + f) * a.inv_2theta, 0.0f);
    }
}

// ---------------------------------------------------------------------------
// Centred fluxes (SWECommon.h:404-473), in sweep order (h, normal, tangential)
// ---------------------------------------------------------------------------

// LxF_2D_flux, SWECommon.h:421-427: 0.5 (F_l + F_r) + dx/(4 dt) (Q_l - Q_r);
// k = dx/(4 dt) (exact: computed here in the reference's order)
template <class M>
__device__ __forceinline__ void lxf_2d_flux(const float ql[3], const float qr[3], float g, float d, float dt, float k,
                                            float out[3]) {
    float Fl[3], Fr[3];
    swe_F<M>(ql, g, Fl);
    swe_F<M>(qr, g, Fr);
    if constexpr (M::exact) k = M::div(d, 4.0f * dt);
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] = 0.5f * (Fl[i] + Fr[i]) + k * (ql[i] - qr[i]);
}

// FORCE_1D_flux, SWECommon.h:466-470: mean of LxF_1D_flux (:404-410) and the
// Richtmyer flux LxW2_1D_flux (:443-450).  klf = dx/(2dt), klw = dt/(2dx).
template <class M>
__device__ __forceinline__ void force_1d_flux(const float ql[3], const float qr[3], float g, float d, float dt,
                                              float klf, float klw, float out[3]) {
    float Fl[3], Fr[3], q[3], Flw[3];
    swe_F<M>(ql, g, Fl);
    swe_F<M>(qr, g, Fr);
    if constexpr (M::exact) {
        klf = M::div(d, 2.0f * dt);
        klw = M::div(dt, 2.0f * d);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) q[i] = 0.5f * (ql[i] + qr[i]) + klw * (Fl[i] - Fr[i]);
    swe_F<M>(q, g, Flw);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float lf = 0.5f * (Fl[i] + Fr[i]) + klf * (ql[i] - qr[i]);
        out[i] = 0.5f * (lf + Flw[i]);
    }
}

// ---------------------------------------------------------------------------
// Scheme policies
// ---------------------------------------------------------------------------

template <class M>
struct UnsplitBase {
    __device__ static float cell_dt(const float q[3], const UnsplitArgs& a) {
        return swe_cell_dt<M>(q, a.dx, a.dy, a.g);
    }
    __device__ static float cell_rate(const float q[3], const UnsplitArgs& a) {
        return fast::swe_cell_rate(q, a.sqrt_g, a.inv_dx, a.inv_dy);
    }
};

template <class M>
struct SweLxF : UnsplitBase<M> {
    static constexpr int GC = 1;
    static constexpr int kind = KIND_LXF;
    __device__ static void flux_x(const float l[3], const float r[3], const UnsplitArgs& a, float o[3]) {
        lxf_2d_flux<M>(l, r, a.g, a.dx, a.dt, a.lxf_kx, o);
    }
    __device__ static void flux_y(const float l[3], const float r[3], const UnsplitArgs& a, float o[3]) {
        lxf_2d_flux<M>(l, r, a.g, a.dy, a.dt, a.lxf_ky, o);
    }
};

template <class M>
struct SweHLL1 : UnsplitBase<M> {
    static constexpr int GC = 1;
    static constexpr int kind = KIND_SPLIT1;
    __device__ static void flux_x(const float l[3], const float r[3], const UnsplitArgs& a, float o[3]) {
        if constexpr (M::exact)
            swe_hll_flux<M>(l, r, a.g, o);
        else
            fast::swe_hll_flux(l, r, a.hg, a.sqrt_g, o);
    }
    __device__ static void flux_y(const float l[3], const float r[3], const UnsplitArgs& a, float o[3]) {
        flux_x(l, r, a, o);
    }
};

template <class M>
struct SweFORCE : UnsplitBase<M> {
    static constexpr int GC = 1;
    static constexpr int kind = KIND_SPLIT1;
    __device__ static void flux_x(const float l[3], const float r[3], const UnsplitArgs& a, float o[3]) {
        force_1d_flux<M>(l, r, a.g, a.dx, a.dt, a.lf_kx, a.lw_kx, o);
    }
    __device__ static void flux_y(const float l[3], const float r[3], const UnsplitArgs& a, float o[3]) {
        force_1d_flux<M>(l, r, a.g, a.dy, a.dt, a.lf_ky, a.lw_ky, o);
    }
};

template <class M>
struct SweKP07U : UnsplitBase<M> {
    static constexpr int GC = 2;
    static constexpr int kind = KIND_KP07U;
    __device__ static void flux(const float l[3], const float r[3], const UnsplitArgs& a, float o[3]) {
        if constexpr (M::exact)
            swe_cu_flux<M>(l, r, a.g, o);
        else
            fast::swe_cu_flux(l, r, a.hg, a.sqrt_g, o);
    }
};

// ---------------------------------------------------------------------------
// Strip addressing for GC ghost cells
// ---------------------------------------------------------------------------

template <int GC>
struct UStrip {
    int col;       // padded column of this lane
    int row0;      // padded row of r = 0
    int r_end;     // rows r in [0, r_end)
    int lane_off;    // byte offset of this lane's column (after the boundary map)
    unsigned xsign;  // sign-bit mask of the x momentum (mirrored ghost column)
    bool lane_out;
    int store_off;  // byte offset of the column for writing lanes, kMaskedOffset for the others
    RowMap ymap;
    __amdgpu_buffer_rsrc_t rin[3], rout[3];
};

template <int GC>
__device__ __forceinline__ void u_setup(const UnsplitArgs& a, UStrip<GC>& st, int lane, int strip_x, int row_in,
                                        int count) {
    constexpr int W = kLanes - 2 * GC;
    st.col = W * strip_x + lane;
    st.row0 = __builtin_amdgcn_readfirstlane(row_in);  // padded row of the first row read
    st.r_end = __builtin_amdgcn_readfirstlane(count + 2 * GC);
    st.ymap = make_row_map<GC>(a.ny, (a.bc >> 16) & 0xF, (a.bc >> 24) & 0xF);
    const AxisMap mx = map_axis_gc<GC>(st.col, a.nx, a.bc & 0xF, (a.bc >> 8) & 0xF);
    st.lane_off = mx.idx * (int)sizeof(float);
    st.xsign = mx.mirrored ? 0x80000000u : 0u;
    st.lane_out = lane >= GC && lane < GC + W && st.col < a.nx + GC;
    st.store_off = st.lane_out ? st.col * (int)sizeof(float) : kMaskedOffset;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        st.rin[k] = field_rsrc(a.in[k], (unsigned)(a.ny + 2 * GC) * (unsigned)a.in_pitch[k]);
        st.rout[k] = field_rsrc(a.out[k], (unsigned)(a.ny + 2 * GC) * (unsigned)a.out_pitch[k]);
    }
}

// Issues the loads of padded row r (raw words) and returns the sign-bit mask
// of its y momentum; u_use applies the mirror signs at the row's first use.
template <int GC>
__device__ __forceinline__ unsigned u_load(const UnsplitArgs& a, const UStrip<GC>& st, int r, float (&v)[3]) {
    unsigned ysign;
    const int row = map_row<GC>(st.ymap, __builtin_amdgcn_readfirstlane(st.row0 + r), ysign);
#pragma unroll
    for (int k = 0; k < 3; ++k)
        v[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(st.rin[k], st.lane_off, row * a.in_pitch[k], 0));
    return ysign;
}

// Values first used here: an empty volatile statement hands them on, so the
// scheduler keeps their consumers behind the loads issued before this point
// and the prefetch distance holds (see use_row in fv_substep.h).  Measured at
// 4096^2 against the same code without it: LxF 9 %, FORCE 7 %, HLL 4 % faster.
__device__ __forceinline__ void u_pin(float (&v)[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) asm volatile("" : "+v"(v[k]));
}

// First use of a row loaded by u_load: pinned, then x mirrors flip the x
// momentum and y mirrors the y momentum.
template <int GC>
__device__ __forceinline__ void u_use(const UStrip<GC>& st, unsigned ysign, float (&v)[3]) {
    u_pin(v);
    v[1] = __uint_as_float(__float_as_uint(v[1]) ^ st.xsign);
    v[2] = __uint_as_float(__float_as_uint(v[2]) ^ ysign);
}

// Output row `prow` (padded): the stored state, or for the second RK2 stage
// 0.5 (old + q) with `old` the output buffer's value (SWE2D_KP07.cu:205-209),
// read ahead by u_load_old.  Every lane stores; the halo lanes' offset is out
// of range (kMaskedOffset), so the hardware drops their writes (and returns 0
// for their reads).
template <int GC, bool Avg = false>
__device__ __forceinline__ void u_store(const UnsplitArgs& a, const UStrip<GC>& st, int prow, const float (&q)[3],
                                        const float (&old)[3]) {
    const int row = __builtin_amdgcn_readfirstlane(prow);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float v = Avg ? 0.5f * (old[k] + q[k]) : q[k];
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), st.rout[k], st.store_off, row * a.out_pitch[k], 0);
    }
}

// the output buffer's row `prow` (padded), for the second RK2 stage
template <int GC>
__device__ __forceinline__ void u_load_old(const UnsplitArgs& a, const UStrip<GC>& st, int prow, float (&old)[3]) {
    const int row = __builtin_amdgcn_readfirstlane(prow);
#pragma unroll
    for (int k = 0; k < 3; ++k)
        old[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(st.rout[k], st.store_off, row * a.out_pitch[k], 0));
}

template <class S, class M>
__device__ __forceinline__ void u_cfl(const float (&q)[3], const UnsplitArgs& a, bool take, float& acc) {
    if constexpr (M::exact) {
        const float cd = S::cell_dt(q, a);
        acc = take ? fminf(acc, cd) : acc;
    } else {
        const float rate = S::cell_rate(q, a);
        acc = take ? fmaxf(acc, rate) : acc;
    }
}

__device__ __forceinline__ void swap_tn(const float (&q)[3], float (&o)[3]) {
    o[0] = q[0];
    o[1] = q[2];
    o[2] = q[1];
}

// q + ((fl - fr) dt / dx + (gs - gn) dt / dy), the unsplit update of
// SWE2D_LxF.cu:147-152 and SWE2D_KP07.cu:190-195
template <class M>
__device__ __forceinline__ float u_update(float q, float fl, float fr, float gs, float gn, const UnsplitArgs& a) {
    if constexpr (M::exact)
        return q + (M::div((fl - fr) * a.dt, a.dx) + M::div((gs - gn) * a.dt, a.dy));
    else
        return q + ((fl - fr) * a.dtdx + (gs - gn) * a.dtdy);
}

// ---------------------------------------------------------------------------
// Row steps
// ---------------------------------------------------------------------------

// x fluxes of one row held across the lanes: F[k] through the interface
// (c, c+1) of this lane's column c.
template <class S, class M>
__device__ __forceinline__ void x_fluxes_first_order(const float (&q)[3], const UnsplitArgs& a, float (&F)[3]) {
    float r[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) r[k] = from_right(q[k]);
    S::flux_x(q, r, a, F);
}

// x fluxes of KP07: minmod slope along the row, face values Q -+ slope/2,
// central-upwind flux at (c, c+1) (SWE2D_KP07.cu:30-56, 103-115)
template <class S, class M>
__device__ __forceinline__ void x_fluxes_kp07(const float (&q)[3], const UnsplitArgs& a, float (&F)[3]) {
    float qm[3], qp[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float hs = half_slope<M>(from_left(q[k]), q[k], from_right(q[k]), a);
        qm[k] = q[k] + hs;               // right face of cell c
        qp[k] = from_right(q[k] - hs);   // left face of cell c + 1
    }
    S::flux(qm, qp, a, F);
}

// y flux between two rows (lower = row r, upper = row r+1): the flux
// functions see (h, hv, hu) and the result is swapped back
template <class S>
__device__ __forceinline__ void y_flux(const float (&lo)[3], const float (&hi)[3], const UnsplitArgs& a,
                                       float (&G)[3]) {
    float l[3], h[3], o[3];
    swap_tn(lo, l);
    swap_tn(hi, h);
    S::flux_y(l, h, a, o);
    swap_tn(o, G);
}

// One step of the one-ghost-cell march: w0 = row r-1 (x-evolved for the split
// schemes), w1 = raw row r on entry (x-evolved here for the split schemes);
// g = flux (r-2, r-1) on entry, (r-1, r) on return; emits row r-1.
template <class S, class M>
__device__ __forceinline__ void gc1_row(const UnsplitArgs& a, const UStrip<1>& st, const float (&w0)[3],
                                        float (&w1)[3], float (&g)[3], int r, bool emit, float& acc) {
    if constexpr (S::kind == KIND_SPLIT1) {
        float F[3];
        x_fluxes_first_order<S, M>(w1, a, F);
#pragma unroll
        for (int k = 0; k < 3; ++k) w1[k] = evolve<M>(w1[k], from_left(F[k]), F[k], a.dt, a.dx, a.dtdx);
    }
    float G[3];
    y_flux<S>(w0, w1, a, G);
    if (emit) {
        float out[3];
        if constexpr (S::kind == KIND_SPLIT1) {
#pragma unroll
            for (int k = 0; k < 3; ++k) out[k] = evolve<M>(w0[k], g[k], G[k], a.dt, a.dy, a.dtdy);
        } else {
            float F[3];
            x_fluxes_first_order<S, M>(w0, a, F);
#pragma unroll
            for (int k = 0; k < 3; ++k) out[k] = u_update<M>(w0[k], from_left(F[k]), F[k], g[k], G[k], a);
        }
        u_store<1>(a, st, st.row0 + r - 1, out, out);
        u_cfl<S, M>(out, a, st.lane_out, acc);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) g[k] = G[k];
}

// One step of the KP07 march: w0, w1, w2 = raw rows r-2, r-1, r; top = top
// face of row r-2 on entry (of r-1 on return); g = flux (r-3, r-2) on entry,
// (r-2, r-1) on return; emits row r-2 (Stage2: averaged with `old`, the output
// buffer's row r-2).
template <class S, class M, bool Stage2>
__device__ __forceinline__ void kp07u_row(const UnsplitArgs& a, const UStrip<2>& st, const float (&w0)[3],
                                          const float (&w1)[3], const float (&w2)[3], const float (&old)[3],
                                          float (&top)[3], float (&g)[3], int r, bool emit, float& acc) {
    float bot[3], top_next[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float hs = half_slope<M>(w0[k], w1[k], w2[k], a);
        top_next[k] = w1[k] + hs;
        bot[k] = w1[k] - hs;
    }
    float G[3];
    {
        float l[3], h[3], o[3];
        swap_tn(top, l);
        swap_tn(bot, h);
        S::flux(l, h, a, o);
        swap_tn(o, G);
    }
    if (emit) {
        float F[3], out[3];
        x_fluxes_kp07<S, M>(w0, a, F);
#pragma unroll
        for (int k = 0; k < 3; ++k) out[k] = u_update<M>(w0[k], from_left(F[k]), F[k], g[k], G[k], a);
        u_store<2, Stage2>(a, st, st.row0 + r - 2, out, old);
        u_cfl<S, M>(out, a, st.lane_out, acc);  // before the RK2 average, as writeCfl reads the tile
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        top[k] = top_next[k];
        g[k] = G[k];
    }
}

// ---------------------------------------------------------------------------
// Marches (row m of a segment in ring slot m % N; see march_slopes)
// ---------------------------------------------------------------------------

template <class S, class M>
__device__ __forceinline__ float march_gc1(const UnsplitArgs& a, const UStrip<1>& st) {
    constexpr int PF = kRowPrefetch;
    constexpr int N = PF + 2;  // rows r-1, r and PF rows in flight
    float acc = M::exact ? __builtin_inff() : 0.0f;
    float ring[N][3], G[3];
    unsigned ys[N];  // y-momentum sign mask of each slot's row
    const int last = st.r_end - 1;  // r_end >= 3
#pragma unroll
    for (int m = 0; m < N; ++m) ys[m] = u_load<1>(a, st, min(m, last), ring[m]);
    u_use<1>(st, ys[0], ring[0]);
    u_use<1>(st, ys[1], ring[1]);
    if constexpr (S::kind == KIND_SPLIT1) {
        float F[3];
        x_fluxes_first_order<S, M>(ring[0], a, F);
#pragma unroll
        for (int k = 0; k < 3; ++k) ring[0][k] = evolve<M>(ring[0][k], from_left(F[k]), F[k], a.dt, a.dx, a.dtdx);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) G[k] = 0.0f;
    // the prologue holds rows 0 .. 1+PF, all that step 1 and its prefetch need
    gc1_row<S, M>(a, st, ring[0], ring[1], G, 1, false, acc);  // flux (0, 1)
    auto step = [&](int u, int rr) {
        ys[(2 + u + PF) % N] = u_load<1>(a, st, min(rr + PF, last), ring[(2 + u + PF) % N]);
        u_use<1>(st, ys[(2 + u) % N], ring[(2 + u) % N]);
        gc1_row<S, M>(a, st, ring[(1 + u) % N], ring[(2 + u) % N], G, rr, true, acc);
    };
    int r = 2;
    for (; r + N <= st.r_end; r += N) {
#pragma unroll
        for (int u = 0; u < N; ++u) step(u, r + u);
    }
#pragma unroll
    for (int u = 0; u < N - 1; ++u)
        if (r + u < st.r_end) step(u, r + u);
    return acc;
}

// Stage2: the output buffer's row of step m (padded row row0 + m - 2) is read
// with input row m, into slot m % N of its own ring, PF steps before it is
// averaged and stored.  Only steps m >= 4 emit, so only their rows are read
// (the rows of steps 0 and 1 would lie above the buffer for the first segment).
template <class S, class M, bool Stage2>
__device__ __forceinline__ float march_kp07u(const UnsplitArgs& a, const UStrip<2>& st) {
    constexpr int PF = kRowPrefetch;
    constexpr int N = PF + 3;
    static_assert(3 + PF >= 4, "rows loaded in the steps belong to emitting steps");
    float acc = M::exact ? __builtin_inff() : 0.0f;
    float ring[N][3], old[N][3], P[3], G[3];
    unsigned ys[N];  // y-momentum sign mask of each slot's row
    const int last = st.r_end - 1;  // r_end >= 5
    auto load = [&](int m, int slot) {  // m >= 4
        ys[slot] = u_load<2>(a, st, min(m, last), ring[slot]);
        if constexpr (Stage2) u_load_old<2>(a, st, st.row0 + min(m, last) - 2, old[slot]);
    };
#pragma unroll
    for (int m = 0; m < N; ++m) {
        if (m >= 4)
            load(m, m);
        else
            ys[m] = u_load<2>(a, st, min(m, last), ring[m]);
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) u_use<2>(st, ys[m], ring[m]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {  // top face of row 1
        P[k] = ring[1][k] + half_slope<M>(ring[0][k], ring[1][k], ring[2][k], a);
        G[k] = 0.0f;
    }
    load(3 + PF, (3 + PF) % N);
    u_use<2>(st, ys[3 % N], ring[3 % N]);
    kp07u_row<S, M, Stage2>(a, st, ring[1], ring[2], ring[3 % N], old[3 % N], P, G, 3, false, acc);  // flux (1, 2)
    auto step = [&](int u, int rr) {
        load(rr + PF, (4 + u + PF) % N);
        u_use<2>(st, ys[(4 + u) % N], ring[(4 + u) % N]);
        if constexpr (Stage2) u_pin(old[(4 + u) % N]);
        kp07u_row<S, M, Stage2>(a, st, ring[(2 + u) % N], ring[(3 + u) % N], ring[(4 + u) % N], old[(4 + u) % N], P,
                                G, rr, true, acc);
    };
    int r = 4;
    for (; r + N <= st.r_end; r += N) {
#pragma unroll
        for (int u = 0; u < N; ++u) step(u, r + u);
    }
#pragma unroll
    for (int u = 0; u < N - 1; ++u)
        if (r + u < st.r_end) step(u, r + u);
    return acc;
}

// The segment of physical wave `wave` of workgroup `block` (of `grid`): the
// round-weighted split, or equal segments with the XCD remap of
// uniform_segment (fv_substep.h).
__host__ __device__ inline bool unsplit_segment(const UnsplitArgs& a, int block, int grid, int wave, int& strip_x,
                                                int& row_in, int& count) {
    if (a.round_plan)
        return round_segment_of(block, wave, a.xcds, a.cus_per_xcd, a.waves_per_block, a.strips_x, a.ny, a.gps,
                                a.grows, strip_x, row_in, count);
    int lb = block;
    if (a.xcd_remap) {
        const int per = grid >> 3;
        lb = (block & 7) * per + (block >> 3);
    }
    const int gw = lb * a.waves_per_block + wave;
    if (lb >= a.num_blocks || gw >= a.num_waves) return false;
    strip_x = gw / a.segs_per_strip;
    row_in = (gw - strip_x * a.segs_per_strip) * a.seg_rows;
    count = min(a.seg_rows, a.ny - row_in);
    return true;
}

template <class S, class M>
__global__ void __launch_bounds__(256, 4) unsplit_kernel(UnsplitArgs a) {
    constexpr int GC = S::GC;
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (a.cfl != nullptr && blockIdx.x == 0 && threadIdx.x == 0) a.cfl[1 - a.cfl_slot] = __builtin_inff();
    int strip_x, row_in, count;
    if (!unsplit_segment(a, blockIdx.x, gridDim.x, wave, strip_x, row_in, count)) return;
    strip_x = __builtin_amdgcn_readfirstlane(strip_x);
    row_in = __builtin_amdgcn_readfirstlane(row_in);
    count = __builtin_amdgcn_readfirstlane(count);
    UStrip<GC> st;
    u_setup<GC>(a, st, lane, strip_x, row_in, count);
    float acc;
    if constexpr (GC == 1)
        acc = march_gc1<S, M>(a, st);
    else
        acc = a.stage2 ? march_kp07u<S, M, true>(a, st) : march_kp07u<S, M, false>(a, st);
    float dt_min = acc;
    if constexpr (!M::exact) dt_min = (acc > 0.0f) ? fast::rcp(acc) : __builtin_inff();
    if (a.cfl != nullptr) {
        const float m = wave_min(dt_min);
        if (lane == 0 && m < __builtin_inff())
            atomicMin(reinterpret_cast<unsigned int*>(a.cfl + a.cfl_slot), __float_as_uint(m));
    }
}

// Strip-aligned segments (as plan_substep) for strips of 64 - 2*GC columns:
// the round-weighted split where `round_ok` and it fits, else equal segments,
// at most waves_per_simd waves on every SIMD.
inline int plan_unsplit(UnsplitArgs& a, int gc, int cu_count, bool round_ok, int waves_per_simd = 4,
                        int min_seg = 24, int min_part = 16) {
    const int W = kLanes - 2 * gc;
    a.strips_x = (a.nx + W - 1) / W;
    a.round_plan = 0;
    a.xcds = 8;
    a.cus_per_xcd = cu_count / 8;
    a.gps = a.grows = 0;
    int gps = 0, grows = 0;
    if (round_ok && a.seg_rows <= 0 && a.xcd_remap &&
        round_split_fits(a.strips_x, a.ny, a.waves_per_block, cu_count, min_part, gps, grows)) {
        a.round_plan = 1;
        a.gps = gps;
        a.grows = grows;
        a.seg_rows = grows / kRounds;  // nominal, for reporting
        a.segs_per_strip = gps * kRounds;
        a.num_waves = a.strips_x * gps * kRounds;
        a.num_blocks = kRounds * cu_count;
        return a.num_blocks;
    }
    // short segments only where one wave per SIMD leaves at most 4 rows a wave:
    // HLL 256^2 12.4 -> 8.5 us per step, but 512^2 (5 rows) 14.4 -> 16.0 us
    if (a.seg_rows <= 0) a.seg_rows = balanced_segment_rows(a.ny, a.strips_x, cu_count, waves_per_simd, min_seg, 4);
    a.segs_per_strip = (a.ny + a.seg_rows - 1) / a.seg_rows;
    a.num_waves = a.strips_x * a.segs_per_strip;
    a.num_blocks = (a.num_waves + a.waves_per_block - 1) / a.waves_per_block;
    int grid = a.num_blocks;
    if (a.xcd_remap) grid = ((grid + 7) / 8) * 8;
    return grid;
}

enum UnsplitSchemeId : int { USCHEME_LXF = 0, USCHEME_HLL = 1, USCHEME_FORCE = 2, USCHEME_KP07 = 3 };

// Schemes that take the round-weighted split.  Measured at 4096^2 against equal
// segments: HLL 5 % and the unsplit KP07 1 % faster, LxF even, FORCE 3 % slower.
inline bool unsplit_round_ok(int scheme) { return scheme == USCHEME_HLL || scheme == USCHEME_KP07; }

template <class M>
hipError_t launch_unsplit(int scheme, const UnsplitArgs& a, int grid, hipStream_t stream) {
    const dim3 block(kLanes * a.waves_per_block);
    switch (scheme) {
        case USCHEME_LXF:
            hipLaunchKernelGGL((unsplit_kernel<SweLxF<M>, M>), dim3(grid), block, 0, stream, a);
            break;
        case USCHEME_HLL:
            hipLaunchKernelGGL((unsplit_kernel<SweHLL1<M>, M>), dim3(grid), block, 0, stream, a);
            break;
        case USCHEME_FORCE:
            hipLaunchKernelGGL((unsplit_kernel<SweFORCE<M>, M>), dim3(grid), block, 0, stream, a);
            break;
        case USCHEME_KP07:
            hipLaunchKernelGGL((unsplit_kernel<SweKP07U<M>, M>), dim3(grid), block, 0, stream, a);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Defined in fv_kernels_exact.hip / fv_kernels_fast.hip.
hipError_t launch_unsplit_exact(int scheme, const UnsplitArgs& a, int grid, hipStream_t stream);
hipError_t launch_unsplit_fast(int scheme, const UnsplitArgs& a, int grid, hipStream_t stream);

}  // namespace fvmi
