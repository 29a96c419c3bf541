// KP07_dimsplit and HLL2 substep kernels in fast arithmetic, built with the
// ILP-maximising instruction scheduler (3 % faster on the 4096^2 KP07 step;
// it raises register use, which WAF and the Euler kernel cannot afford, so
// those stay in fv_kernels_fast.hip with the default scheduler).
#include "fv_launch.h"

namespace fvmi {
hipError_t launch_substep_fast_ilp(int scheme, const SubstepArgs& a, int grid, hipStream_t stream) {
    switch (scheme) {
        case SCHEME_KP07:
            return launch_substep_scheme<SweKP07<MathFast>, MathFast>(a, grid, stream);
        case SCHEME_HLL2:
            return launch_substep_scheme<SweHLL2<MathFast>, MathFast>(a, grid, stream);
        default:
            return hipErrorInvalidValue;
    }
}
}  // namespace fvmi
