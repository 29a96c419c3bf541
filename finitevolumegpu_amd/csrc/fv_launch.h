// Launch helpers shared by the two arithmetic-policy translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "fv_ensemble.h"
#include "fv_substep.h"

namespace fvmi {

enum SchemeId : int { SCHEME_KP07 = 0, SCHEME_HLL2 = 1, SCHEME_WAF = 2, SCHEME_EULER = 3 };

// Columns per lane of each scheme's march in each arithmetic policy (the kC
// of its policy struct) and the output columns of one of its strips.
inline int scheme_cols_per_lane(int scheme, int math) {
    static_assert(SweKP07<MathFast>::kC == FVMI_SWE_COLS && SweHLL2<MathFast>::kC == FVMI_SWE_COLS &&
                      SweWAF<MathFast>::kC == 1 && EulerKP07<MathFast>::kC == 2 &&
                      SweKP07<MathExact>::kC == FVMI_SWE_COLS && SweHLL2<MathExact>::kC == FVMI_SWE_COLS &&
                      SweWAF<MathExact>::kC == 1 && EulerKP07<MathExact>::kC == 1,
                  "scheme_cols_per_lane follows the policy structs");
    if (scheme == SCHEME_WAF) return 1;
    if (scheme == SCHEME_EULER) return math == FVMI_MATH_EXACT ? 1 : 2;
    return FVMI_SWE_COLS;
}
inline int scheme_strip_cols(int scheme, int math) {
    return scheme_cols_per_lane(scheme, math) == 2 ? strip_cols<2>() : strip_cols<1>();
}

// Fills the decomposition fields of `a` (sub-area, launch shape, device size)
// and returns the grid size to launch.  seg_rows == 0 picks the shape:
//   * the round-weighted plan (round_segment) when `round_ok`, the workgroups
//     have 4 waves, the XCD remap is on and every strip holds row groups long
//     enough for their shortest part to have `min_part` rows;
//   * otherwise the balanced segment length: at most `waves_per_simd` waves on
//     every SIMD of the device, but no segment shorter than `min_seg` rows (the
//     4-row prologue of a segment is then at most 1/6 of its work), except on
//     small grids (balanced_segment_rows).
inline int plan_substep(SubstepArgs& a, int cu_count, bool round_ok = false, int waves_per_simd = 4,
                        int min_seg = 24, int min_part = 16) {
    const int wx = a.x1 - a.x0;
    a.rows_per_strip = a.y1 - a.y0;
    a.strips_x = (wx + a.strip_cols - 1) / a.strip_cols;
    a.round_plan = 0;
    a.xcds = 8;
    a.cus_per_xcd = cu_count / 8;
    a.gps = a.grows = 0;
    int gps = 0, grows = 0;
    if (round_ok && a.seg_rows <= 0 && a.xcd_remap &&
        round_split_fits(a.strips_x, a.rows_per_strip, a.waves_per_block, cu_count, min_part, gps, grows)) {
        a.round_plan = 1;
        a.gps = gps;
        a.grows = grows;
        a.seg_rows = grows / kRounds;  // nominal, for reporting
        a.segs_per_strip = gps * kRounds;
        a.num_waves = a.strips_x * gps * kRounds;
        a.num_blocks = kRounds * cu_count;
        return a.num_blocks;
    }
    if (a.seg_rows <= 0) {
        // at most waves_per_simd waves on every SIMD (one more would run after
        // all the others), segments of at least min_seg rows; shorter ones on
        // grids too small to fill one wave per SIMD with them
        // (measured per full step: KP07_dimsplit 256^2 27.2 -> 14.8 us, 512^2
        // 29.6 -> 23.3 us; EE2D 512^2 46.4 -> 31.5 us)
        a.seg_rows = balanced_segment_rows(a.rows_per_strip, a.strips_x, cu_count, waves_per_simd, min_seg, min_seg);
    }
    a.segs_per_strip = (a.rows_per_strip + a.seg_rows - 1) / a.seg_rows;
    a.num_waves = a.strips_x * a.segs_per_strip;
    a.num_blocks = (a.num_waves + a.waves_per_block - 1) / a.waves_per_block;
    int grid = a.num_blocks;
    if (a.xcd_remap) grid = ((grid + 7) / 8) * 8;
    return grid;
}

// One scheme's kernel (each translation unit instantiates the ones it owns).
template <class S, class M>
hipError_t launch_substep_scheme(const SubstepArgs& a, int grid, hipStream_t stream) {
    hipLaunchKernelGGL((substep_kernel<S, M>), dim3(grid), dim3(kLanes * a.waves_per_block), 0, stream, a);
    return hipGetLastError();
}

template <class M>
hipError_t launch_substep(int scheme, SubstepArgs a, int grid, hipStream_t stream) {
    const dim3 block(kLanes * a.waves_per_block);
    switch (scheme) {
        case SCHEME_KP07:
            hipLaunchKernelGGL((substep_kernel<SweKP07<M>, M>), dim3(grid), block, 0, stream, a);
            break;
        case SCHEME_HLL2:
            hipLaunchKernelGGL((substep_kernel<SweHLL2<M>, M>), dim3(grid), block, 0, stream, a);
            break;
        case SCHEME_WAF:
            hipLaunchKernelGGL((substep_kernel<SweWAF<M>, M>), dim3(grid), block, 0, stream, a);
            break;
        case SCHEME_EULER:
            hipLaunchKernelGGL((substep_kernel<EulerKP07<M>, M>), dim3(grid), block, 0, stream, a);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class S, class M>
hipError_t launch_ensemble_scheme(const EnsembleArgs& e, int grid, hipStream_t stream) {
    hipLaunchKernelGGL((ensemble_substep_kernel<S, M>), dim3(grid), dim3(kLanes * e.waves_per_block), 0, stream, e);
    return hipGetLastError();
}

template <class M>
hipError_t launch_ensemble_substep(int scheme, EnsembleArgs e, int grid, hipStream_t stream) {
    const dim3 block(kLanes * e.waves_per_block);
    switch (scheme) {
        case SCHEME_KP07:
            hipLaunchKernelGGL((ensemble_substep_kernel<SweKP07<M>, M>), dim3(grid), block, 0, stream, e);
            break;
        case SCHEME_HLL2:
            hipLaunchKernelGGL((ensemble_substep_kernel<SweHLL2<M>, M>), dim3(grid), block, 0, stream, e);
            break;
        case SCHEME_WAF:
            hipLaunchKernelGGL((ensemble_substep_kernel<SweWAF<M>, M>), dim3(grid), block, 0, stream, e);
            break;
        case SCHEME_EULER:
            hipLaunchKernelGGL((ensemble_substep_kernel<EulerKP07<M>, M>), dim3(grid), block, 0, stream, e);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Defined in fv_kernels_exact.hip / fv_kernels_fast.hip.
hipError_t launch_ensemble_exact(int scheme, const EnsembleArgs& e, int grid, hipStream_t stream);
hipError_t launch_ensemble_fast(int scheme, const EnsembleArgs& e, int grid, hipStream_t stream);
hipError_t launch_substep_exact(int scheme, const SubstepArgs& a, int grid, hipStream_t stream);
hipError_t launch_substep_fast(int scheme, const SubstepArgs& a, int grid, hipStream_t stream);
// KP07_dimsplit and HLL2 in fast arithmetic, single grid and ensemble
// (fv_kernels_fast_ilp.hip, built with the ILP-maximising scheduler)
hipError_t launch_substep_fast_ilp(int scheme, const SubstepArgs& a, int grid, hipStream_t stream);
hipError_t launch_ensemble_fast_ilp(int scheme, const EnsembleArgs& e, int grid, hipStream_t stream);

}  // namespace fvmi
