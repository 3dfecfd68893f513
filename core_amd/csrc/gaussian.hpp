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
// the x pass and the division (radii beyond k_gauss_fused's).  Consecutive threads take
// consecutive x, so every tap of either pass is a coalesced row access.
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

// Both passes in one kernel, for vertical radii up to RM: a block takes GAUSS_THREADS consecutive
// input columns (its TX = GAUSS_THREADS - 2 rx output columns and their x halo) and marches down a
// band of GAUSS_TY output rows, two at a time.  Each thread keeps the weighted data and weights of
// its column for the last 2 RM + 2 input rows in registers (a ring whose slot indices are
// compile-time: the row loop is unrolled by the ring length; taps beyond the runtime radius are
// zero), forms the y pass of both output rows and puts them in LDS (one vector per column and row);
// after one barrier the output threads form the x pass of both rows.  The LDS rows are
// double-buffered, so one barrier per row pair suffices.  HBM traffic: the input once (plus the x
// halo and 2 RM rows per band), the output once; the taps and the order of every sum are those of
// k_gauss_v / k_gauss_h, so the results are identical.
constexpr int GAUSS_TY = 64;

template <int C> struct GaussVec;
template <> struct GaussVec<1> { using T = float2; };
template <> struct GaussVec<2> { using T = float4; };

template <int C, int RM>
__global__ __launch_bounds__(GAUSS_THREADS) void k_gauss_fused(GaussArgs g) {
  static_assert(GAUSS_TY % 2 == 0, "row pairs");
  constexpr int NW = 2 * RM + 1, U = NW + 1;  // U: even ring length, the window of a row pair
  using V = typename GaussVec<C>::T;
  __shared__ V row[2][2][GAUSS_THREADS];  // [buffer][row of the pair][column]
  const int t = threadIdx.x;
  const int rx = g.rx, ry = g.ry;
  const int TX = GAUSS_THREADS - 2 * rx;
  const long long nstrip = (g.nx + TX - 1) / TX;
  const long long strip = blockIdx.x % nstrip, band = blockIdx.x / nstrip;
  const long long x0 = strip * TX, y0 = band * GAUSS_TY;
  const long long yend = min(y0 + GAUSS_TY, g.ny);
  const long long xi = min(max(x0 - rx + t, 0LL), g.nx - 1);  // replicate border
  // vertical taps on the ring's 2 RM + 1 offsets
  float ky[NW];
#pragma unroll
  for (int j = 0; j < NW; ++j) {
    const int a = j - RM;
    ky[j] = (a >= -ry && a <= ry) ? g.ky[a + ry] : 0.f;
  }
  // weighted data and weight of input row y0 - RM + i of this thread's column (slot i mod U)
  auto load = [&](long long i, float (&c)[2 * C]) {
    const long long j = min(max(y0 - RM + i, 0LL), g.ny - 1) * g.nx + xi;
    float w = 1.f;
    if (g.weight) {
      w = g.weight[j];
      if (isnan(w)) w = 0.f;
    }
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const float d = g.data[j * C + q];
      const bool ok = !isnan(d);
      const float wc = ok ? w : 0.f;
      c[q] = ok ? d * wc : 0.f;
      c[C + q] = wc;
    }
  };
  auto put = [&](V &dst, const float (&v)[2 * C]) {
    if constexpr (C == 1) dst = make_float2(v[0], v[1]);
    else dst = make_float4(v[0], v[1], v[2], v[3]);
  };
  auto get = [&](const V &src, float (&v)[2 * C]) {
    if constexpr (C == 1) {
      v[0] = src.x;
      v[1] = src.y;
    } else {
      v[0] = src.x;
      v[1] = src.y;
      v[2] = src.z;
      v[3] = src.w;
    }
  };
  float ring[U][2 * C];
#pragma unroll
  for (int i = 0; i < NW - 1; ++i) load(i, ring[i]);
  for (long long g0 = 0; y0 + g0 < yend; g0 += U) {
#pragma unroll
    for (int jj = 0; jj < U; jj += 2) {
      // output rows y0 + n and y0 + n + 1 take input rows n .. n + NW (slots (jj + j) mod U)
      const long long n = g0 + jj, yo = y0 + n;
      if (yo >= yend) break;
      load(n + NW - 1, ring[(jj + NW - 1) % U]);
      load(n + NW, ring[(jj + NW) % U]);
      float v0[2 * C], v1[2 * C];
#pragma unroll
      for (int q = 0; q < 2 * C; ++q) v0[q] = v1[q] = 0.f;
#pragma unroll
      for (int j = 0; j < NW; ++j)
#pragma unroll
        for (int q = 0; q < 2 * C; ++q) {
          v0[q] += ky[j] * ring[(jj + j) % U][q];
          v1[q] += ky[j] * ring[(jj + j + 1) % U][q];
        }
      const int b = (int)((n >> 1) & 1);
      put(row[b][0][t], v0);
      put(row[b][1][t], v1);
      __syncthreads();
      const long long xo = x0 + t - rx;
      if (t >= rx && t < rx + TX && xo < g.nx) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          if (yo + h >= yend) break;
          float num[C], den[C];
#pragma unroll
          for (int q = 0; q < C; ++q) num[q] = den[q] = 0.f;
          for (int bb = -rx; bb <= rx; ++bb) {
            const float k = g.kx[bb + rx];
            float e[2 * C];
            get(row[b][h][t + bb], e);
#pragma unroll
            for (int q = 0; q < C; ++q) {
              num[q] += k * e[q];
              den[q] += k * e[C + q];
            }
          }
          const long long i = (yo + h) * g.nx + xo;
#pragma unroll
          for (int q = 0; q < C; ++q)
            g.out[i * C + q] = den[q] > g.lim ? num[q] / (den[q] + 1e-10f) : __builtin_nanf("");
        }
      }
    }
  }
}

}  // namespace irls2d
