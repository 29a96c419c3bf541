// Fast arithmetic instantiation (hardware reciprocal / square root, contraction
// allowed) of the Euler kernels, built with the ILP-maximising instruction
// scheduler (1 % faster than the default one on the EE2D 8192^2 substep).  The
// shallow-water kernels live in fv_kernels_fast_ilp.hip.
#include "fv_launch.h"

namespace fvmi {
hipError_t launch_substep_fast(int scheme, const SubstepArgs& a, int grid, hipStream_t stream) {
    switch (scheme) {
        case SCHEME_KP07:
        case SCHEME_HLL2:
        case SCHEME_WAF:
            return launch_substep_fast_ilp(scheme, a, grid, stream);
        case SCHEME_EULER:
            return launch_substep_scheme<EulerKP07<MathFast>, MathFast>(a, grid, stream);
        default:
            return hipErrorInvalidValue;
    }
}
hipError_t launch_ensemble_fast(int scheme, const EnsembleArgs& e, int grid, hipStream_t stream) {
    switch (scheme) {
        case SCHEME_KP07:
        case SCHEME_HLL2:
        case SCHEME_WAF:
            return launch_ensemble_fast_ilp(scheme, e, grid, stream);
        case SCHEME_EULER:
            return launch_ensemble_scheme<EulerKP07<MathFast>, MathFast>(e, grid, stream);
        default:
            return hipErrorInvalidValue;
    }
}
}  // namespace fvmi
