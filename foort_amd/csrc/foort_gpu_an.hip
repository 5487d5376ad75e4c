// foort_gpu_an.hip -- kernels with analytic Christoffel symbols (exact
// derivatives of the reference metrics by forward-mode dual numbers), in
// binary64 or mixed binary32/binary64.  These are not bit-compatible with the
// reference's finite differences, so this unit is compiled with the multiply-adds
// of each expression fused (-ffp-contract=on: the same source rounds the same way
// in every kernel).
#include "foort_kernels.hpp"

namespace foort {

KernelPair pick_analytic(int metric, int integ, int precision, bool every)
{
    if (precision == FOORT_PRECISION_MIXED)
        return pick_policy<RhsMixed>(metric, integ, every);
    return pick_policy<RhsAn64>(metric, integ, every);
}

} // namespace foort
