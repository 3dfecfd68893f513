// NaN-aware separable Gaussian filter (gfx950): insardev/insardev/utils_gaussian.py
// nanconvolve2d_gaussian_pytorch (:14-116), the reference's GPU path.
//
// C channels per pixel (1: real, 2: complex as interleaved real/imaginary parts, each filtered on
// its own as the reference does).  Per channel: data NaN -> 0 with weight 0, else weight 1 or the
// caller's weight (NaN -> 0); numerator = G * (data weight), denominator = G * weight, with G the
// separable Gaussian applied along y then x with replicate borders; out = numerator /
// (denominator + 1e-10) where denominator > lim, NaN elsewhere (lim = threshold without a weight
// array, 1e-10 with one).  k_gauss_taps builds both normalised 1-D kernels (one block, float64
// sums in a fixed order), k_gauss_v the y pass into a scratch of 2C floats per pixel, k_gauss_h
// the x pass and the division.  Consecutive threads take consecutive x, so every tap of either
// pass is a coalesced row access (the overlapping reads of neighbouring taps hit the caches).
#pragma once

#include "common.hpp"

namespace irls2d {

constexpr int GAUSS_THREADS = 256;

struct GaussArgs {
  const float *data;    // ny x nx x C
  const float *weight;  // ny x nx, or null
  float *tmp;           // ny x nx x 2C: y-pass numerators then denominators
  float *out;           // ny x nx x C
  float *ky, *kx;       // taps, 2 ry + 1 and 2 rx + 1
  long long ny, nx;
  int ry, rx;
  float sy, sx;
  float lim;
};

// exp(-x^2 / 2 s^2) for x = -r .. r, normalised to sum 1 (utils_gaussian.py:51-55), both axes.
__global__ __launch_bounds__(GAUSS_THREADS) void k_gauss_taps(GaussArgs g) {
  __shared__ double part[GAUSS_THREADS];
  for (int axis = 0; axis < 2; ++axis) {
    const int r = axis ? g.rx : g.ry;
    const double s = axis ? g.sx : g.sy;
    float *k = axis ? g.kx : g.ky;
    double acc = 0.0;
    for (int i = threadIdx.x; i <= 2 * r; i += blockDim.x) {
      const double x = (double)(i - r) / s;
      acc += exp(-0.5 * x * x);
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    double tot = 0.0;
    for (int t = 0; t < (int)blockDim.x; ++t) tot += part[t];  // fixed order
    for (int i = threadIdx.x; i <= 2 * r; i += blockDim.x) {
      const double x = (double)(i - r) / s;
      k[i] = (float)(exp(-0.5 * x * x) / tot);
    }
    __syncthreads();
  }
}

template <int C>
__global__ __launch_bounds__(GAUSS_THREADS) void k_gauss_v(GaussArgs g) {
  const long long n = g.ny * g.nx;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const long long y = i / g.nx, x = i - y * g.nx;
    float num[C], den[C];
#pragma unroll
    for (int c = 0; c < C; ++c) num[c] = den[c] = 0.f;
    for (int a = -g.ry; a <= g.ry; ++a) {
      const long long yy = min(max(y + a, 0LL), g.ny - 1);
      const long long j = yy * g.nx + x;
      const float k = g.ky[a + g.ry];
      float w = 1.f;
      if (g.weight) {
        w = g.weight[j];
        if (isnan(w)) w = 0.f;
      }
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const float d = g.data[j * C + c];
        const bool ok = !isnan(d);
        const float wc = ok ? w : 0.f;
        num[c] += k * (ok ? d * wc : 0.f);
        den[c] += k * wc;
      }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
      g.tmp[i * 2 * C + c] = num[c];
      g.tmp[i * 2 * C + C + c] = den[c];
    }
  }
}

template <int C>
__global__ __launch_bounds__(GAUSS_THREADS) void k_gauss_h(GaussArgs g) {
  const long long n = g.ny * g.nx;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const long long y = i / g.nx, x = i - y * g.nx;
    float num[C], den[C];
#pragma unroll
    for (int c = 0; c < C; ++c) num[c] = den[c] = 0.f;
    for (int b = -g.rx; b <= g.rx; ++b) {
      const long long xx = min(max(x + b, 0LL), g.nx - 1);
      const long long j = y * g.nx + xx;
      const float k = g.kx[b + g.rx];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        num[c] += k * g.tmp[j * 2 * C + c];
        den[c] += k * g.tmp[j * 2 * C + C + c];
      }
    }
#pragma unroll
    for (int c = 0; c < C; ++c)
      g.out[i * C + c] = den[c] > g.lim ? num[c] / (den[c] + 1e-10f) : __builtin_nanf("");
  }
}

}  // namespace irls2d
