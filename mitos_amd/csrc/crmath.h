// Double-precision log and cos for the Box-Muller transform of the noise.
#pragma once

#include "mitos_common.h"

namespace mitos {

__device__ __forceinline__ double cr_log(double x) { return log(x); }
__device__ __forceinline__ double cr_cos(double x) { return cos(x); }

}  // namespace mitos
