// The transform kernels are instantiated in translation units of their own (fft_e12.hip,
// fft_e16.hip, fft_e24.hip, long_passes.hip) so that the build compiles them in parallel; the host
// translation unit (irls2d.hip) only declares these instantiations.
#pragma once

#include "kernels.hpp"
#include "long_dct.hpp"
#include "slab_dct.hpp"

#define IRLS2D_FFT_KERNELS(E, X)                                                                  \
  X template __global__ void k_row_dct2<false, E>(Ctx, float *, float *, const float *, int);     \
  X template __global__ void k_row_dct2<true, E>(Ctx, float *, float *, const float *, int);      \
  X template __global__ void k_col<E, false>(Ctx, float *, int);                                  \
  X template __global__ void k_col<E, true>(Ctx, float *, int);                                   \
  X template __global__ void k_row_dct3<E>(Ctx, const float *, float *, int, int);

#define IRLS2D_LONG_KERNELS(X)                                                                    \
  X template __global__ void k_long_p1<LONG_DCT2, LONG_E>(Ctx, int, float *, const float *, int); \
  X template __global__ void k_long_p1<LONG_DCT2_UPDATE, LONG_E>(Ctx, int, float *, const float *, int); \
  X template __global__ void k_long_p1<LONG_DCT3, LONG_E>(Ctx, int, float *, const float *, int); \
  X template __global__ void k_long_p1<LONG_DCT3_SCALE, LONG_E>(Ctx, int, float *, const float *, int); \
  X template __global__ void k_long_p2<LONG_DCT2, LONG_E>(Ctx, int, float *, int, int);           \
  X template __global__ void k_long_p2<LONG_DCT3, LONG_E>(Ctx, int, float *, int, int);           \
  X template __global__ void k_lb_pack<LONG_DCT2>(Ctx, int, float *, const float *, int);         \
  X template __global__ void k_lb_pack<LONG_DCT2_UPDATE>(Ctx, int, float *, const float *, int);  \
  X template __global__ void k_lb_pack<LONG_DCT3>(Ctx, int, float *, const float *, int);         \
  X template __global__ void k_lb_pack<LONG_DCT3_SCALE>(Ctx, int, float *, const float *, int);   \
  X template __global__ void k_lb_fft13<1, LONG_E>(Ctx, int, int);                                \
  X template __global__ void k_lb_fft13<3, LONG_E>(Ctx, int, int);                                \
  X template __global__ void k_lb_fft2<LONG_E>(Ctx, int, int);                                    \
  X template __global__ void k_lb_post<LONG_DCT2>(Ctx, int, float *, int, int);                   \
  X template __global__ void k_lb_post<LONG_DCT3>(Ctx, int, float *, int, int);

#define IRLS2D_SLAB_KERNELS(X)                                                                    \
  X template __global__ void k_slab1<SLAB_E>(Ctx, const float *, int);                            \
  X template __global__ void k_slab2<SLAB_E>(Ctx, int);                                           \
  X template __global__ void k_slab3<SLAB_E>(Ctx, float *, int);                                 \
  X template __global__ void k_slab1_f<16>(Ctx, const float *, int);                              \
  X template __global__ void k_slab1_f<32>(Ctx, const float *, int);                              \
  X template __global__ void k_slab1_f<64>(Ctx, const float *, int);                              \
  X template __global__ void k_slab1_f<128>(Ctx, const float *, int);                             \
  X template __global__ void k_slab3_f<16>(Ctx, float *, int);                                    \
  X template __global__ void k_slab3_f<32>(Ctx, float *, int);                                    \
  X template __global__ void k_slab3_f<64>(Ctx, float *, int);                                    \
  X template __global__ void k_slab3_f<128>(Ctx, float *, int);                                   \
  X template __global__ void k_slab2_f<16>(Ctx, int);                                             \
  X template __global__ void k_slab2_f<32>(Ctx, int);                                             \
  X template __global__ void k_slab2_f<64>(Ctx, int);
