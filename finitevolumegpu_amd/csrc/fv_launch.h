// Launch helpers shared by the two arithmetic-policy translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "fv_substep.h"

namespace fvmi {

enum SchemeId : int { SCHEME_KP07 = 0, SCHEME_HLL2 = 1, SCHEME_WAF = 2, SCHEME_EULER = 3 };

// Fills the decomposition fields of `a` and returns the grid size to launch.
inline int plan_substep(SubstepArgs& a) {
    const int wx = a.x1 - a.x0, wy = a.y1 - a.y0;
    a.strips_x = (wx + kStripCols - 1) / kStripCols;
    a.strips_y = (wy + a.tile_rows - 1) / a.tile_rows;
    a.block_cols = (a.strips_x + a.waves_per_block - 1) / a.waves_per_block;
    a.num_blocks = a.block_cols * a.strips_y;
    int grid = a.num_blocks;
    if (a.xcd_remap) grid = ((grid + 7) / 8) * 8;
    return grid;
}

template <class M>
hipError_t launch_substep(int scheme, SubstepArgs a, hipStream_t stream) {
    const int grid = plan_substep(a);
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

// Defined in fv_kernels_exact.hip / fv_kernels_fast.hip.
hipError_t launch_substep_exact(int scheme, const SubstepArgs& a, hipStream_t stream);
hipError_t launch_substep_fast(int scheme, const SubstepArgs& a, hipStream_t stream);

}  // namespace fvmi
