// Shared device/host types and device helpers for the IRLS unwrap solver (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace irls2d {

// Elementwise/stencil tile: 64 columns x 64 rows, 256 threads (4 wave64s); lane = column,
// wave w handles rows w, w+4, ... so that each wave reads 256-byte row segments.
constexpr int TILE = 64;
constexpr int TILE_THREADS = 256;
// Transpose tile
constexpr int TT = 64;
// LDS FFT: at most EPT complex values per thread per pass, <= 1024 threads per workgroup.
constexpr int EPT = 16;
constexpr int MAX_FFT_THREADS = 1024;
constexpr int MAX_FFT_LEN = EPT * MAX_FFT_THREADS;  // 16384
constexpr int MAX_PASS = 16;
constexpr int MAX_NV = 3;  // values per block partial

struct FftPlan {
  int L;       // complex length
  int npass;   // number of Stockham passes
  int tw_off;  // float2 offset of exp(-2 pi i m / L), m in [0, L)
  int pad;
  int radix[MAX_PASS];
};

struct DctPlan {
  int N;          // DCT length
  int M;          // FFT length actually run (N, or >= 2N-1 for Bluestein)
  int blue;       // 1 = Bluestein chirp-z
  int chirp_off;  // N entries exp(-i pi n^2 / N)          (Bluestein)
  int bhat_off;   // M entries FFT_M(conj chirp, wrapped)/M (Bluestein)
  int mk_off;     // N entries exp(-i pi k / (2N))          (Makhoul post-twiddle)
  float invN;     // 1/N
  int pad;
  FftPlan fft;
};

struct ImgDesc {
  long long off;  // element offset of the image in every field arena
  int H, W;
  int rowplan;    // DctPlan index for length W
  int colplan;    // DctPlan index for length H
  int cosH_off;   // double offset of cos(pi i / H), i < H
  int cosW_off;   // double offset of cos(pi j / W), j < W
};

// Per-image solver state.  Written only by "last block" finalizers, read by later kernels.
struct ImgState {
  double rz;        // r.z of the current PCG iteration
  double alpha;
  double beta;
  double best_res;  // best IRLS residual so far
  double mean;      // masked mean of u
  double vcount;    // number of valid pixels
  int irls_active;
  int cg_active;
  int n_cg;         // PCG iterations in the current IRLS step
  int n_cg_total;
  int n_irls;
  int improved_it;  // IRLS iteration whose u became u_best
  int restore_it;   // IRLS iteration that restored u_best after a non-finite u
  int status;
  unsigned counter; // arrival counter of the last-block reductions
  int pad;
};

struct Fields {
  float *tx, *ty, *wx, *wy;     // wrapped target gradients, static edge weights
  float *u, *u_prev, *u_best;   // solution (the caller's output arena), copies
  float *wxi, *wyi, *b;         // IRLS edge weights, right-hand side
  float *r, *z, *p0, *p1, *Ap;  // PCG vectors (p ping-pong)
  float *T, *Tt;                // DCT row-pass output and its transpose
  unsigned char *valid;
};

struct Geo {  // block start of each image per launch geometry, [n + 1] each
  const int *tile, *rowpair, *colpair, *trans;
};

struct Ctx {
  const ImgDesc *img;
  ImgState *st;
  const DctPlan *plans;
  const float2 *tab;
  const double *dtab;
  double *partials;
  Geo geo;
  Fields f;
  int n;
  float eps_sq;
  double tol, cg_tol;
};

#ifdef __HIPCC__

__device__ __forceinline__ int find_img(const int *start, int n, int blk) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (start[mid] <= blk) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-wide sum of NV doubles; every thread receives the totals.  All threads must call.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double sh[NV][MAX_FFT_THREADS / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) sh[q][wv] = v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += sh[q][w];  // fixed order: deterministic
    v[q] = s;
  }
  __syncthreads();
}

// Two-stage deterministic reduction over the nblk blocks of one image.  Each block stores its
// partial; the block that arrives last sums all partials of the image in a fixed order and
// returns true with the totals in `tot` (all its threads).  Other blocks return false.
template <int NV>
__device__ __forceinline__ bool reduce_last(double (&v)[NV], double *partials, unsigned *counter,
                                            int first_blk, int nblk, double (&tot)[NV]) {
  __shared__ int is_last;
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) partials[(long long)blockIdx.x * NV + q] = v[q];
    __threadfence();
    unsigned prev = atomicAdd(counter, 1u);
    is_last = (prev == (unsigned)(nblk - 1));
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = 0.0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
#pragma unroll
    for (int q = 0; q < NV; ++q)
      acc[q] += __hip_atomic_load(&partials[(long long)(first_blk + i) * NV + q], __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
  }
  block_sum<NV>(acc);
#pragma unroll
  for (int q = 0; q < NV; ++q) tot[q] = acc[q];
  if (threadIdx.x == 0) atomicExch(counter, 0u);
  return true;
}

__device__ __forceinline__ float clampf_nan(float v, float lo, float hi) {
  // torch.clamp semantics: NaN passes through
  return v < lo ? lo : (v > hi ? hi : v);
}

__device__ __forceinline__ double clampd_nan(double v, double lo, double hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

#endif  // __HIPCC__

}  // namespace irls2d
