// LDS-resident transform kernels with 24 complex values per thread (lengths <= 12288).
#define IRLS2D_TRANSFORMS_ONLY
#include "instances.hpp"

namespace irls2d {
IRLS2D_FFT_KERNELS(24, )
}  // namespace irls2d
