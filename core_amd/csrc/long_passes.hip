// Four-step transform kernels: the two passes of lines longer than 12288, the three slab passes
// of the column transform on the row-major field.
#define IRLS2D_TRANSFORMS_ONLY
#include "instances.hpp"

namespace irls2d {
IRLS2D_LONG_KERNELS()
IRLS2D_SLAB_KERNELS()
}  // namespace irls2d
