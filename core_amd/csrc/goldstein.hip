// Goldstein filter kernels for patch sizes 16, 32 and 64 per axis.
#include "goldstein.hpp"

namespace irls2d {
#define IRLS2D_GOLD_PAIR(PY, PX)                            \
  template __global__ void k_gold_patch<PY, PX>(GoldArgs); \
  template __global__ void k_gold_fold<PY, PX>(GoldArgs);
IRLS2D_GOLD_PAIR(16, 16)
IRLS2D_GOLD_PAIR(16, 32)
IRLS2D_GOLD_PAIR(16, 64)
IRLS2D_GOLD_PAIR(32, 16)
IRLS2D_GOLD_PAIR(32, 32)
IRLS2D_GOLD_PAIR(32, 64)
IRLS2D_GOLD_PAIR(64, 16)
IRLS2D_GOLD_PAIR(64, 32)
IRLS2D_GOLD_PAIR(64, 64)
}  // namespace irls2d
