// tcx_tc.cpp -- TC/BITC terms (tc2x, tc3a1, tc3b1) on the context.
#include "tcx_terms.hpp"

namespace tcx {

void Context::tc2x(const int*, const double*, double*, int, unsigned)
{
    throw Error(TCX_ERR_UNSUPPORTED, "tc2x: not implemented yet");
}

void Context::tc3_local(int, const int*, const double*, double*, int, unsigned)
{
    throw Error(TCX_ERR_UNSUPPORTED, "tc3a1/tc3b1: not implemented yet");
}

void Context::ensure_density() {}

}  // namespace tcx
