// Goldstein adaptive phase filter (gfx950): the patch filter and the overlap-add of
// insardev/insardev/utils_goldstein.py (goldstein_pytorch :157-330, goldstein_numpy :12-154).
//
// Patches of PY x PX complex pixels at steps PY/2, PX/2 from the origin (pixels beyond the image
// and NaN pixels read as 0, NaN correlation as 0).  k_gold_patch runs 256 / max(PY, PX) patches
// per workgroup: thread l of a patch loads patch row l straight from HBM (with its partial valid
// count and sum of w * corr) and transforms it in registers (Dft<PX>); an LDS transpose (row pitch
// PX + 1, conflict-free) hands the columns to thread l for Dft<PY>, the power-spectrum weight
// max(|F|, 1e-10)^alpha with alpha = 1 - sum(w corr) / sum(w), and the first inverse step (as the
// forward DFT of the conjugate); a second transpose and Dft<PX> finish the inverse.  The result,
// times the triangle window w, goes to a scratch of npatch x PY x PX, zero for an invalid patch
// (fewer than min_valid valid pixels).  k_gold_fold sums, per output pixel, the (at most four)
// patches covering it in raster order, and divides by the number of valid ones (mode 0, the
// reference's GPU path) or by the sum of their window weights (mode 1, its CPU path); a NaN input
// pixel, or one that no valid patch (mode 0) / no positive weight (mode 1) covers, is NaN.
#pragma once

#include "fft_lds.hpp"

namespace irls2d {

constexpr int GOLD_THREADS = 256;

struct GoldArgs {
  const float2 *phase;  // ny x nx complex64, rows of nx
  const float *corr;    // ny x nx
  float2 *out;          // ny x nx complex64
  float2 *patches;      // npy x npx patches of PY x PX
  int *pvalid;          // npy x npx
  long long ny, nx;
  int npy, npx;
  int min_valid;
  int mode;             // 0: divide by the valid-patch count, 1: by the window-weight sum
  double wsum;          // sum of the float32 window
};

// One axis of the triangle window (utils_goldstein.py:62-68): 1 - |i - (h - 1)| / (h - 1) for
// i < h = P / 2, mirrored for i >= h; float64 as in the reference.
template <int P>
__device__ __forceinline__ double gold_w1(int i) {
  constexpr int h = P / 2;
  const int j = i < h ? i : P - 1 - i;
  return 1.0 - fabs((double)j - (double)(h - 1)) / (double)(h - 1);
}

// The float32 window value at (r, c): the float64 product rounded once, as np.block(...).astype.
template <int PY, int PX>
__device__ __forceinline__ float gold_w(int r, int c) {
  return (float)(gold_w1<PY>(r) * gold_w1<PX>(c));
}

template <int PY, int PX>
__global__ __launch_bounds__(GOLD_THREADS) void k_gold_patch(GoldArgs g) {
  constexpr int PM = PY > PX ? PY : PX;
  constexpr int NP = GOLD_THREADS / PM;
  constexpr int RP = PX + 1;
  extern __shared__ float2 lds[];
  __shared__ double cw[NP][PY];
  __shared__ int cv[NP][PY];
  const int lp = threadIdx.x / PM, l = threadIdx.x % PM;
  const long long npatch = (long long)g.npy * g.npx;
  const long long pid = (long long)blockIdx.x * NP + lp;
  const bool live = pid < npatch;
  const int iy = live ? (int)(pid / g.npx) : 0, ix = live ? (int)(pid % g.npx) : 0;
  const long long y0 = (long long)iy * (PY / 2), x0 = (long long)ix * (PX / 2);
  float2 *buf = lds + lp * PY * RP;
  // rows: load, statistics, forward DFT along x
  if (l < PY) {
    float2 v[PX];
    double acc = 0.0;
    int nv = 0;
    const long long y = y0 + l;
    const bool yin = live && y < g.ny;
#pragma unroll
    for (int c = 0; c < PX; ++c) {
      const long long x = x0 + c;
      float2 z = make_float2(0.f, 0.f);
      float cr = 0.f;
      if (yin && x < g.nx) {
        z = g.phase[y * g.nx + x];
        cr = g.corr[y * g.nx + x];
        if (isnan(z.x) || isnan(z.y)) z = make_float2(0.f, 0.f);
        else ++nv;
        if (isnan(cr)) cr = 0.f;
      }
      v[c] = z;
      acc += (double)gold_w<PY, PX>(l, c) * (double)cr;
    }
    cw[lp][l] = acc;
    cv[lp][l] = nv;
    Dft<PX>::run(v);
#pragma unroll
    for (int c = 0; c < PX; ++c) buf[l * RP + c] = v[c];
  }
  __syncthreads();
  double sw = 0.0;
  int nvalid = 0;
#pragma unroll
  for (int r = 0; r < PY; ++r) {
    sw += cw[lp][r];
    nvalid += cv[lp][r];
  }
  const bool pv = live && nvalid >= g.min_valid;
  const float alpha = (float)(1.0 - sw / g.wsum);
  // columns: forward DFT along y, spectrum weight, conjugate, DFT along y (first inverse step)
  if (l < PX) {
    float2 v[PY];
#pragma unroll
    for (int r = 0; r < PY; ++r) v[r] = buf[r * RP + l];
    Dft<PY>::run(v);
#pragma unroll
    for (int r = 0; r < PY; ++r) {
      const float m = fmaxf(sqrtf(v[r].x * v[r].x + v[r].y * v[r].y), 1e-10f);
      v[r] = cconj(cscale(v[r], powf(m, alpha)));
    }
    Dft<PY>::run(v);
#pragma unroll
    for (int r = 0; r < PY; ++r) buf[r * RP + l] = v[r];
  }
  __syncthreads();
  // rows: DFT along x of the conjugate, conjugate and 1 / (PY PX) = inverse FFT2; times w
  if (l < PY && live) {
    float2 v[PX];
#pragma unroll
    for (int c = 0; c < PX; ++c) v[c] = buf[l * RP + c];
    Dft<PX>::run(v);
    constexpr float inv = 1.0f / (float)(PY * PX);
    float2 *dst = g.patches + (pid * PY + l) * PX;
#pragma unroll
    for (int c = 0; c < PX; c += 2) {
      float2 a = make_float2(0.f, 0.f), b = make_float2(0.f, 0.f);
      if (pv) {
        a = cscale(cconj(v[c]), inv * gold_w<PY, PX>(l, c));
        b = cscale(cconj(v[c + 1]), inv * gold_w<PY, PX>(l, c + 1));
      }
      *(float4 *)(dst + c) = make_float4(a.x, a.y, b.x, b.y);
    }
    if (l == 0) g.pvalid[pid] = pv ? 1 : 0;
  }
}

template <int PY, int PX>
__global__ __launch_bounds__(GOLD_THREADS) void k_gold_fold(GoldArgs g) {
  const long long n = g.ny * g.nx;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const long long y = i / g.nx, x = i - y * g.nx;
    constexpr int sy = PY / 2, sx = PX / 2;
    const int iy1 = (int)(y / sy), ix1 = (int)(x / sx);
    float2 sum = make_float2(0.f, 0.f);
    float cnt = 0.f;
#pragma unroll
    for (int a = 1; a >= 0; --a) {
      const int iy = iy1 - a;
      if (iy < 0 || iy >= g.npy) continue;
#pragma unroll
      for (int b = 1; b >= 0; --b) {
        const int ix = ix1 - b;
        if (ix < 0 || ix >= g.npx) continue;
        const long long p = (long long)iy * g.npx + ix;
        if (!g.pvalid[p]) continue;
        const int r = (int)(y - (long long)iy * sy), c = (int)(x - (long long)ix * sx);
        sum = cadd(sum, g.patches[(p * PY + r) * PX + c]);
        cnt += g.mode ? gold_w<PY, PX>(r, c) : 1.f;
      }
    }
    const float2 z = g.phase[i];
    float2 o;
    if (isnan(z.x) || isnan(z.y) || cnt < 1e-6f) {
      o = make_float2(__builtin_nanf(""), __builtin_nanf(""));
    } else {
      const float d = g.mode ? fmaxf(cnt, 1e-10f) : fmaxf(cnt, 1.f);
      o = make_float2(sum.x / d, sum.y / d);
    }
    g.out[i] = o;
  }
}

}  // namespace irls2d
