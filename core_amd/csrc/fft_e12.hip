// LDS-resident transform kernels with 12 complex values per thread on up to 1024 threads
// (lengths <= 12288; the default above 8192, where a workgroup's LDS leaves room for one per CU).
#define IRLS2D_TRANSFORMS_ONLY
#include "instances.hpp"

namespace irls2d {
IRLS2D_FFT_KERNELS(12, )
}  // namespace irls2d
