// Exact arithmetic instantiation.  Built with -ffp-contract=off and IEEE
// division / square root, so every operation rounds as an IEEE float32
// evaluation of the reference expressions would.
#include "fv_launch.h"
#include "fv_unsplit.h"

namespace fvmi {
hipError_t launch_substep_exact(int scheme, const SubstepArgs& a, int grid, hipStream_t stream) {
    return launch_substep<MathExact>(scheme, a, grid, stream);
}
hipError_t launch_ensemble_exact(int scheme, const EnsembleArgs& e, int grid, hipStream_t stream) {
    return launch_ensemble_substep<MathExact>(scheme, e, grid, stream);
}
hipError_t launch_unsplit_exact(int scheme, const UnsplitArgs& a, int grid, hipStream_t stream) {
    return launch_unsplit<MathExact>(scheme, a, grid, stream);
}
}  // namespace fvmi
