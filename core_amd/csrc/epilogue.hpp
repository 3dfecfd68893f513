// Device epilogue of one solved image (insardev/insardev/utils_unwrap2d.py:1659-1746), run in
// host-orchestrated steps so that the few decisions with numpy semantics (the median of the cycle
// counts, the size ordering of the components) are taken on the host from small arrays:
//   k_ep_scan    non-finite count, valid count, min / max of k = rint((u - psi) / 2 pi)
//   k_ep_hist    histogram of k                                   -> host: median (np.median)
//   k_ep_shift   u -= k_median 2 pi on valid pixels, NaN elsewhere; union-find init
//   k_ccl_merge  4-connected union of valid neighbours (atomic-min linking: roots become the
//                smallest pitched index of their component, i.e. the raster-first pixel)
//   k_ccl_flat   every pixel points at its root; per-block root counts -> host: offsets
//   k_ccl_number roots numbered in raster order (the numbering of scipy.ndimage.label and
//                cv2.connectedComponents)
//   k_ep_sums    per-component float64 sums of sin/cos of psi and of the shifted u, and counts
//                -> host: circular-mean adjustment and size relabel (numpy, as the host epilogue)
//   k_ep_apply   u = float(double(u) + adj[label]), conncomp = new_label[label]
// Pixel (y, x) of an image is element y P + x of its pitched view; labels use that index.
#pragma once

#include "common.hpp"

namespace irls2d {

constexpr int EP_THREADS = 256;
constexpr int EP_PER_THREAD = 16;
constexpr int EP_BLOCK_PX = EP_THREADS * EP_PER_THREAD;  // pitched indices per block
constexpr int EP_HASH = 128;                              // LDS slots of the per-block label table
constexpr int EP_LDS_BINS = 8192;                         // histogram bins kept in LDS

// Cycle counts beyond +-EP_KMAX saturate there (a diverged but finite u); the host then takes
// the median from the float32 values themselves instead of from the histogram.
constexpr int EP_KMAX = 1 << 30;

// k as the reference computes it in float32: round((u - psi) / (2 pi)), half to even, saturated
// to [-EP_KMAX, EP_KMAX] (a float outside the int range has no defined conversion).
__device__ __forceinline__ int ep_k(float u, float psi) {
  const float two_pi = 6.28318530717958647692f;
  const float k = rintf((u - psi) / two_pi);
  return k >= (float)EP_KMAX ? EP_KMAX : (k <= -(float)EP_KMAX ? -EP_KMAX : (int)k);
}

template <typename T>
__device__ __forceinline__ T ep_wave_reduce_min(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
template <typename T>
__device__ __forceinline__ T ep_wave_reduce_max(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}
template <typename T>
__device__ __forceinline__ T ep_wave_reduce_add(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// stats: [0] non-finite u on valid pixels, [1] min k, [2] max k, [3] valid pixels.
__global__ __launch_bounds__(EP_THREADS) void k_ep_scan(int H, int W, int P, const float *__restrict__ psi,
                                                        const float *__restrict__ u, int *stats) {
  const int n_idx = H * P;
  int nonfin = 0, kmin = 0x7fffffff, kmax = -0x7fffffff - 1, nvalid = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n_idx; i += (long long)gridDim.x * blockDim.x) {
    if ((int)(i % P) >= W) continue;
    const float p = psi[i];
    if (isnan(p)) continue;
    ++nvalid;
    const float v = u[i];
    if (!isfinite(v)) { ++nonfin; continue; }
    const int k = ep_k(v, p);
    kmin = min(kmin, k);
    kmax = max(kmax, k);
  }
  nonfin = ep_wave_reduce_add(nonfin);
  nvalid = ep_wave_reduce_add(nvalid);
  kmin = ep_wave_reduce_min(kmin);
  kmax = ep_wave_reduce_max(kmax);
  if ((threadIdx.x & 63) == 0) {
    if (nonfin) atomicAdd(&stats[0], nonfin);
    if (kmin != 0x7fffffff) atomicMin(&stats[1], kmin);
    if (kmax != -0x7fffffff - 1) atomicMax(&stats[2], kmax);
    if (nvalid) atomicAdd(&stats[3], nvalid);
  }
}

// hist[k - kmin] += 1 for every valid pixel (all finite by then) whose k lies in [kmin, kmin + nb).
__global__ __launch_bounds__(EP_THREADS) void k_ep_hist(int H, int W, int P, const float *__restrict__ psi,
                                                        const float *__restrict__ u, int kmin, int nb, int *hist) {
  __shared__ int lh[EP_LDS_BINS];
  const bool local = nb <= EP_LDS_BINS;
  if (local)
    for (int b = threadIdx.x; b < nb; b += blockDim.x) lh[b] = 0;
  __syncthreads();
  const int n_idx = H * P;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n_idx; i += (long long)gridDim.x * blockDim.x) {
    if ((int)(i % P) >= W) continue;
    const float p = psi[i];
    if (isnan(p)) continue;
    const int b = ep_k(u[i], p) - kmin;
    if (b < 0 || b >= nb) continue;  // outside the caller's range: not counted
    if (local) atomicAdd(&lh[b], 1);
    else atomicAdd(&hist[b], 1);
  }
  __syncthreads();
  if (local)
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
      if (lh[b]) atomicAdd(&hist[b], lh[b]);
}

// u -= shift on valid pixels (shift = k_median * 2 pi in float32), NaN elsewhere; L[i] = i on
// valid pixels, -1 elsewhere (padding included).
__global__ __launch_bounds__(EP_THREADS) void k_ep_shift(int H, int W, int P, const float *__restrict__ psi,
                                                         float *__restrict__ u, float shift, int *__restrict__ L) {
  const int n_idx = H * P;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n_idx; i += (long long)gridDim.x * blockDim.x) {
    const bool in = (int)(i % P) < W;
    const bool valid = in && !isnan(psi[i]);
    if (in) u[i] = valid ? u[i] - shift : __builtin_nanf("");
    L[i] = valid ? (int)i : -1;
  }
}

__device__ __forceinline__ int ccl_load(const int *L, int x) {
  return __hip_atomic_load(L + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int ccl_find(const int *L, int x) {
  int p = ccl_load(L, x);
  while (p != x) {
    x = p;
    p = ccl_load(L, x);
  }
  return x;
}

// Links the trees of a and b: the larger root is pointed at the smaller one (atomicMin keeps
// every parent index below its child's, so the final root is the component's smallest index).
__device__ __forceinline__ void ccl_unite(int *L, int a, int b) {
  while (true) {
    a = ccl_find(L, a);
    b = ccl_find(L, b);
    if (a == b) return;
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    const int old = atomicMin(&L[a], b);
    if (old == a) return;
    a = old;
  }
}

__global__ __launch_bounds__(EP_THREADS) void k_ccl_merge(int H, int W, int P, int *L) {
  const int n_idx = H * P;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n_idx; i += (long long)gridDim.x * blockDim.x) {
    if (ccl_load(L, (int)i) < 0) continue;
    const int x = (int)(i % P);
    if (x + 1 < W && ccl_load(L, (int)i + 1) >= 0) ccl_unite(L, (int)i, (int)i + 1);
    if (i + P < n_idx && ccl_load(L, (int)(i + P)) >= 0) ccl_unite(L, (int)i, (int)(i + P));
  }
}

// L[i] = root(i) for valid pixels; counts[b] = roots among the EP_BLOCK_PX indices of block b.
__global__ __launch_bounds__(EP_THREADS) void k_ccl_flat(int H, int P, int *L, int *counts) {
  const int n_idx = H * P;
  const long long base = (long long)blockIdx.x * EP_BLOCK_PX;
  int roots = 0;
  for (int q = 0; q < EP_PER_THREAD; ++q) {
    const long long i = base + threadIdx.x + (long long)q * EP_THREADS;
    if (i >= n_idx) break;
    const int p = L[i];
    if (p < 0) continue;
    const int r = ccl_find(L, (int)i);
    if (r != p) L[i] = r;
    roots += (r == (int)i);
  }
  roots = ep_wave_reduce_add(roots);
  __shared__ int sh[EP_THREADS / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = roots;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < EP_THREADS / 64; ++w) s += sh[w];
    counts[blockIdx.x] = s;
  }
}

// R[i] = raster-order number (1-based) of root i: offsets[b] + rank of i among the roots of
// block b (an LDS scan over the block's indices in increasing order).
__global__ __launch_bounds__(EP_THREADS) void k_ccl_number(int H, int P, const int *__restrict__ L,
                                                           const int *__restrict__ offsets, int *__restrict__ R) {
  const int n_idx = H * P;
  const long long base = (long long)blockIdx.x * EP_BLOCK_PX;
  // thread t owns the contiguous indices base + t * EP_PER_THREAD .. + EP_PER_THREAD - 1
  int flags = 0, cnt = 0;
  for (int q = 0; q < EP_PER_THREAD; ++q) {
    const long long i = base + (long long)threadIdx.x * EP_PER_THREAD + q;
    const bool root = i < n_idx && L[i] == (int)i;
    flags |= (root ? 1 : 0) << q;
    cnt += root;
  }
  __shared__ int sc[EP_THREADS];
  sc[threadIdx.x] = cnt;
  __syncthreads();
  for (int o = 1; o < EP_THREADS; o <<= 1) {  // inclusive Hillis-Steele scan
    const int v = threadIdx.x >= o ? sc[threadIdx.x - o] : 0;
    __syncthreads();
    sc[threadIdx.x] += v;
    __syncthreads();
  }
  int rank = offsets[blockIdx.x] + sc[threadIdx.x] - cnt;
  for (int q = 0; q < EP_PER_THREAD; ++q) {
    if (flags & (1 << q)) R[base + (long long)threadIdx.x * EP_PER_THREAD + q] = ++rank;
  }
}

// Per-component sums in float64 (float32 sin / cos, as the reference's accumulator) and counts:
// sums[4 c + 0..3] = sin psi, cos psi, sin u, cos u; cnt[c].  Each block aggregates its pixels in
// an LDS table keyed by component (label 0 = invalid is skipped) and adds the table to the
// global arrays; labels that do not fit the table go to the global arrays directly.
__global__ __launch_bounds__(EP_THREADS) void k_ep_sums(int H, int W, int P, const float *__restrict__ psi,
                                                        const float *__restrict__ u, const int *__restrict__ L,
                                                        const int *__restrict__ R, double *sums,
                                                        unsigned long long *cnt) {
  __shared__ int key[EP_HASH];
  __shared__ double acc[EP_HASH][4];
  __shared__ unsigned int num[EP_HASH];
  for (int s = threadIdx.x; s < EP_HASH; s += blockDim.x) {
    key[s] = 0;
    num[s] = 0;
    acc[s][0] = acc[s][1] = acc[s][2] = acc[s][3] = 0.0;
  }
  __syncthreads();
  const int n_idx = H * P;
  const long long base = (long long)blockIdx.x * EP_BLOCK_PX;
  // a thread walks EP_PER_THREAD consecutive indices and flushes a run whenever the label changes
  int run_lab = 0;
  unsigned int run_n = 0;
  double run[4] = {0.0, 0.0, 0.0, 0.0};
  auto flush = [&]() {
    if (run_lab <= 0 || run_n == 0) return;
    int h = (int)(((unsigned)run_lab * 2654435761u) >> 25) & (EP_HASH - 1);
    for (int probe = 0; probe < EP_HASH; ++probe) {
      const int k = atomicCAS(&key[h], 0, run_lab);
      if (k == 0 || k == run_lab) {
        atomicAdd(&acc[h][0], run[0]);
        atomicAdd(&acc[h][1], run[1]);
        atomicAdd(&acc[h][2], run[2]);
        atomicAdd(&acc[h][3], run[3]);
        atomicAdd(&num[h], run_n);
        run_n = 0;
        return;
      }
      h = (h + 1) & (EP_HASH - 1);
    }
    double *g = sums + 4LL * run_lab;  // table full
    atomicAdd(g + 0, run[0]);
    atomicAdd(g + 1, run[1]);
    atomicAdd(g + 2, run[2]);
    atomicAdd(g + 3, run[3]);
    atomicAdd(cnt + run_lab, (unsigned long long)run_n);
    run_n = 0;
  };
  for (int q = 0; q < EP_PER_THREAD; ++q) {
    const long long i = base + (long long)threadIdx.x * EP_PER_THREAD + q;
    if (i >= n_idx || (int)(i % P) >= W) continue;
    const int root = L[i];
    if (root < 0) continue;
    const int lab = R[root];
    if (lab != run_lab) {
      flush();
      run_lab = lab;
      run[0] = run[1] = run[2] = run[3] = 0.0;
    }
    const float p = psi[i], v = u[i];
    run[0] += (double)sinf(p);
    run[1] += (double)cosf(p);
    run[2] += (double)sinf(v);
    run[3] += (double)cosf(v);
    ++run_n;
  }
  flush();
  __syncthreads();
  for (int s = threadIdx.x; s < EP_HASH; s += blockDim.x) {
    const int k = key[s];
    if (k > 0 && num[s]) {
      double *g = sums + 4LL * k;
      atomicAdd(g + 0, acc[s][0]);
      atomicAdd(g + 1, acc[s][1]);
      atomicAdd(g + 2, acc[s][2]);
      atomicAdd(g + 3, acc[s][3]);
      atomicAdd(cnt + k, (unsigned long long)num[s]);
    }
  }
}

// u = float(double(u) + adj[label]) and conncomp = new_label[label] on valid pixels; conncomp 0
// elsewhere (row padding untouched).
__global__ __launch_bounds__(EP_THREADS) void k_ep_apply(int H, int W, int P, float *__restrict__ u,
                                                         const int *__restrict__ L, const int *__restrict__ R,
                                                         const double *__restrict__ adj,
                                                         const unsigned short *__restrict__ relabel,
                                                         unsigned short *__restrict__ cc) {
  const int n_idx = H * P;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n_idx; i += (long long)gridDim.x * blockDim.x) {
    if ((int)(i % P) >= W) continue;
    const int root = L[i];
    if (root < 0) {
      cc[i] = 0;
      continue;
    }
    const int lab = R[root];
    u[i] = (float)((double)u[i] + adj[lab]);
    cc[i] = relabel[lab];
  }
}

}  // namespace irls2d
