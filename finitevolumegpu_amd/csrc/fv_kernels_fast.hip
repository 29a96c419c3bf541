// Fast arithmetic instantiation (hardware reciprocal / square root, contraction
// allowed).  The production path.
#include "fv_launch.h"
#include "fv_unsplit.h"

namespace fvmi {
hipError_t launch_substep_fast(int scheme, const SubstepArgs& a, int grid, hipStream_t stream) {
    return launch_substep<MathFast>(scheme, a, grid, stream);
}
hipError_t launch_ensemble_fast(int scheme, const EnsembleArgs& e, int grid, hipStream_t stream) {
    return launch_ensemble_substep<MathFast>(scheme, e, grid, stream);
}
hipError_t launch_unsplit_fast(int scheme, const UnsplitArgs& a, int grid, hipStream_t stream) {
    return launch_unsplit<MathFast>(scheme, a, grid, stream);
}
}  // namespace fvmi
