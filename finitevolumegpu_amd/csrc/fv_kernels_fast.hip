// Fast arithmetic instantiation (hardware reciprocal / square root, contraction
// allowed).  The production path.
#include "fv_launch.h"

namespace fvmi {
hipError_t launch_substep_fast(int scheme, const SubstepArgs& a, int grid, hipStream_t stream) {
    return launch_substep<MathFast>(scheme, a, grid, stream);
}
}  // namespace fvmi
