// Hex8 element stiffness on MFMA (placeholder until the BF16 hi/lo kernel lands).
#include "gf_internal.hpp"

using namespace gf;

int gf::stiffness_bf16x3(gf_quad, const double*, double*, hipStream_t)
{
    set_error("GF_BF16X3 stiffness not available in this build");
    return GF_EINVAL;
}
