// Batched ensemble: many independent grids of one scheme advanced together.
//
// Device-resident instance table (fvmi_instance_t, include/fvmi.h) holds each
// grid's fields, parameters and time-stepping state.  One full step of every
// active instance is four launches, all reading the table on the device (so a
// step sequence can be captured in a hipGraph and replayed with no host work):
//
//   prepare   per instance: active = t < t_end; every 100 steps dt = 0.5 *
//             cfl_min * cfl_scale (double, as BaseSimulator.simulate computes
//             it, Simulator.py:192-199); dt_step = float(min(dt, t_end - t));
//             CFL slots reset
//   substep 0 (x then y) and substep 1 (y then x): the single-grid march over a
//             work list of segments (instance, strip, first row, rows) shared
//             by all instances, so small and large grids fill the device
//             together; inactive instances' segments return at once.  The
//             list's length is read on the device, so re-planning between
//             graph replays needs no new graph
//   advance   per instance: t += dt_step (double), nt += 1
//
// Per-cell arithmetic is the single-grid kernel's, so an instance evolves
// bitwise as the plain Simulator.simulate() loop would evolve it alone.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/fvmi.h"
#include "fv_substep.h"

namespace fvmi {

struct EnsembleArgs {
    fvmi_instance_t* table;
    const int4* segs;     // (instance, strip_x, row_in, count); strips of strip_cols<S::kC>() columns
    const int* num_segs;  // device: valid entries of segs
    int step;
    int waves_per_block;
    int num_blocks;
    int xcd_remap;
};

// The table is read with vector loads (the kernel also writes it, so the
// compiler cannot use scalar loads); every value read from it is made
// wave-uniform explicitly, so the fields' buffer descriptors live in scalar
// registers (a descriptor in vector registers costs a loop over the distinct
// values around every buffer access).
__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uniform_f(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ float* uniform_ptr(float* p) {
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
    return reinterpret_cast<float*>(((unsigned long long)hi << 32) | lo);
}

// One segment of one instance: the single-grid march with the instance's
// parameters; returns after the CFL update (step 1).
template <class S, class M>
__device__ __forceinline__ void ensemble_segment(const EnsembleArgs& e, const int4 seg, int lane) {
    const int id = uniform_i(seg.x);
    const fvmi_instance_t& I = e.table[id];
    if (!uniform_i(I.active)) return;

    SubstepArgs a;
    a.nx = uniform_i(I.nx);
    a.ny = uniform_i(I.ny);
    a.dx = uniform_f(I.dx);
    a.dy = uniform_f(I.dy);
    a.dt = uniform_f(I.dt_step) * 0.5f;  // substep length, KP07_dimsplit.py:108-109
    a.g = uniform_f(I.g);
    a.theta = uniform_f(I.theta);
    a.gamma = uniform_f(I.gamma);
    a.step = e.step;
    a.bc = uniform_i(I.bc);
    a.x0 = 0;
    a.y0 = 0;
    a.x1 = a.nx;
    a.y1 = a.ny;
    const int pitch = uniform_i(I.pitch);
#pragma unroll
    for (int k = 0; k < S::NV; ++k) {
        // after a full step the state is back in buffer 0
        a.in[k] = uniform_ptr(I.field[e.step][k]);
        a.out[k] = uniform_ptr(I.field[e.step ^ 1][k]);
        a.in_pitch[k] = pitch;
        a.out_pitch[k] = pitch;
    }
    a.cfl = nullptr;
    a.hg = 0.5f * a.g;
    a.sqrt_g = uniform_f(I.sqrt_g);
    a.inv_g = uniform_f(I.inv_g);
    a.inv_dx = uniform_f(I.inv_dx);
    a.inv_dy = uniform_f(I.inv_dy);
    a.dtdx = a.dt / a.dx;
    a.dtdy = a.dt / a.dy;
    set_slope_constants(a);

    Strip<S::kC> st;
    setup_strip<S, M>(a, st, lane, __builtin_amdgcn_readfirstlane(seg.y), __builtin_amdgcn_readfirstlane(seg.z),
                      __builtin_amdgcn_readfirstlane(seg.w));
    const float acc = march<S, M>(a, st);
    if (e.step != 1) return;  // the CFL is taken after the y-then-x substep
    float dt_min = acc;
    if constexpr (!M::exact) dt_min = (acc > 0.0f) ? fast::rcp(acc) : __builtin_inff();
    const float m = wave_min(dt_min);
    if (lane == 0 && m < __builtin_inff())
        atomicMin(reinterpret_cast<unsigned int*>(&e.table[id].cfl[1]), __float_as_uint(m));
}

// A fixed grid of waves walks the work list (wave w takes segments w, w + W,
// ...; the list is sorted longest first), so the launch shape does not depend
// on the list and a captured graph survives re-planning.  Sized to the list's
// buffer, the grid gives every segment its own wave.
template <class S, class M>
__global__ void __launch_bounds__(256, 4) ensemble_substep_kernel(EnsembleArgs e) {
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int lb = blockIdx.x;
    if (e.xcd_remap) {
        const int per = gridDim.x >> 3;
        lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    }
    if (lb >= e.num_blocks) return;
    const int total = e.num_blocks * e.waves_per_block;
    const int n = __builtin_amdgcn_readfirstlane(*e.num_segs);
    for (int s = __builtin_amdgcn_readfirstlane(lb * e.waves_per_block + wave); s < n; s += total)
        ensemble_segment<S, M>(e, e.segs[s], lane);
}

}  // namespace fvmi
