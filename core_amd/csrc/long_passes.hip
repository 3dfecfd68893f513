// Two-pass (four-step) transform kernels for lines longer than 12288.
#define IRLS2D_TRANSFORMS_ONLY
#include "instances.hpp"

namespace irls2d {
IRLS2D_LONG_KERNELS()
}  // namespace irls2d
