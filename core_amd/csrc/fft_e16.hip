// LDS-resident transform kernels with 16 complex values per thread (lengths <= 8192).
#define IRLS2D_TRANSFORMS_ONLY
#include "instances.hpp"

namespace irls2d {
IRLS2D_FFT_KERNELS(16, )
}  // namespace irls2d
