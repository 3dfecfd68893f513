// Shared device/host types and device helpers for the IRLS unwrap solver (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace irls2d {

// Elementwise/stencil tile: 1024 columns x 16 rows, 256 threads (4 wave64s).  Wave w owns the
// 256-column strip j0 + 256w and walks down its 16 rows.  Scalar kernels give lane l the columns
// strip + l + 64u (u < 4); the float4 stencils give it the four columns strip + 4l .. 4l + 3.
constexpr int TILE_W = 1024;
constexpr int TILE_H = 16;
constexpr int STRIP = 256;
constexpr int TILE_THREADS = 256;
constexpr int NUM_XCD = 8;  // workgroups are dispatched round-robin over the 8 XCDs of an MI355X
// Transpose tile
constexpr int TT = 64;
// LDS FFT: at most E complex values per thread per pass (E = 12, 16 or 24, a template parameter
// of the transform kernels).  E = 16 and 24 run <= 512 threads per workgroup (256 VGPRs per lane,
// no spills at E = 24); E = 12 runs <= 1024, which doubles the waves of a workgroup whose LDS
// (lengths above 8192) leaves room for only one per CU.
constexpr int MAX_EPT = 24;
constexpr int MAX_FFT_THREADS = 512;
constexpr int MAX_FFT_THREADS_WIDE = 1024;
template <int E>
constexpr int fft_threads_max() { return E == 12 ? MAX_FFT_THREADS_WIDE : MAX_FFT_THREADS; }
// Compile-time boolean tag for generic lambdas.
template <bool B>
struct BoolC {
  static constexpr bool value = B;
};

// Items per thread whose global loads k_row_dct3 issues together (input step, output step).
constexpr int DCT3_U = 8;
constexpr int DCT3_UO = 8;
// Values per thread whose Bluestein tables (chirp, transformed chirp) are loaded together.
constexpr int BLUE_U = 8;
// Elements per thread whose global loads k_col issues together (load and store steps; half that
// many coefficient pairs in its middle step).
constexpr int COL_U = 8;
constexpr int MAX_FFT_LEN = MAX_EPT * MAX_FFT_THREADS;  // 12288
constexpr int MAX_PASS = 16;
constexpr int MAX_NV = 3;  // values per block partial
// Split twiddle table in LDS: 64 "lo" entries then up to MAX_FFT_LEN/64 "hi" entries.
constexpr int TW_LO = 64;
constexpr int TW_LDS = TW_LO + MAX_FFT_LEN / 64;
// Padded LDS length of a sequence of L complex values (one pad slot per 16).
constexpr int lds_seq_len(int L) { return L + (L >> 4) + 1; }
// Lines longer than MAX_FFT_LEN: four-step transform N = N1 N2 in two global passes, each a batch
// of B short FFTs in LDS (B * factor <= LONG_E * LONG_THREADS).
constexpr int LONG_THREADS = 512;
constexpr int LONG_E = 16;
constexpr int LONG_MAX_FACTOR = 1024;
constexpr int LONG_BATCH_ELEMS = LONG_E * LONG_THREADS;  // 8192
constexpr int LONG_MAX_BATCH = 32;
// LDS of a long pass in float2: twiddles of the factor, then factor x (B + 1) values.
constexpr int LONG_LDS = TW_LDS + LONG_BATCH_ELEMS + LONG_MAX_FACTOR;

// Column pass on the row-major field (no transposes): slabs of SLAB_CP column pairs, the column
// length H = N1 N2 transformed as a four-step FFT in three passes, each a batch of short FFTs in
// LDS (slots x factor <= SLAB_E * threads, at most SLAB_MAX_THREADS threads for the batched-LDS
// kernels).
constexpr int SLAB_CP = 32;
constexpr int SLAB_E = 16;
constexpr int SLAB_MAX_THREADS = 1024;

// One Stockham pass, precomputed on the host so the device does no integer division.
struct PassDesc {
  int R;        // radix
  int Ns;       // product of the earlier radices
  int nb;       // L / R butterflies
  int twmul;    // L / (Ns R): twiddle index step
  int rstride;  // padded LDS stride between a butterfly's reads when nb % 16 == 0, else 0
  int wstride;  // padded LDS stride between its writes when Ns % 16 == 0 (or Ns == 1, R | 16), else 0
  int nsmask;   // Ns - 1 when Ns is a power of two, else -1
  float inv_ns; // 1 / Ns for the float quotient of the mixed-radix passes
};

struct FftPlan {
  int L;       // complex length
  int npass;   // number of Stockham passes
  int tw_off;  // float2 offset of the split twiddle table: exp(-2 pi i b / L), b < 64, then
               // exp(-2 pi i 64 a / L), a < ceil(L / 64)
  int ct_len;  // L when the passes are the default power-of-two schedule that has a compile-time
               // specialisation (1024 .. 8192), else 0
  PassDesc pass[MAX_PASS];
};

struct DctPlan {
  int N;          // DCT length
  int M;          // FFT length actually run (N, or >= 2N-1 for Bluestein)
  int blue;       // 1 = Bluestein chirp-z
  int chirp_off;  // N entries exp(-i pi n^2 / N)          (Bluestein)
  int bhat_off;   // M entries FFT_M(conj chirp, wrapped)/M (Bluestein)
  int mk_off;     // N entries exp(-i pi k / (2N))          (Makhoul post-twiddle)
  float invN;     // 1/N
  int ept;        // elements per thread the plan asks for (16 or 24)
  FftPlan fft;
};

// DCT of a line longer than MAX_FFT_LEN: the complex FFT of length N = N1 N2 (Makhoul-packed
// pair of real lines) as N2 FFTs of length N1 over n1 of x[N2 n1 + n2] (pass 1), the twiddle
// exp(-2 pi i n2 k1 / N), then N1 FFTs of length N2 over n2 (pass 2); the result for k1 + N1 k2
// sits at row k1, column k2 of the scratch line.  Pass 1 blocks take a set of columns closed
// under the reflection that pairs their inputs, pass 2 blocks a set of rows closed under the one
// that pairs their outputs, so every global element is read and written once, coalesced.
//
// blue = 1: N has no such split (a prime factor above 5, or factors above 1024).  The length-N FFT
// of the packed sequence then runs as Bluestein's chirp convolution of length M = N1 N2 >= 2N - 1
// (5-smooth, both factors <= 1024): a pack pass writes the sequence to the scratch line Sa, three
// four-step passes (k_lb_fft1 / k_lb_fft2 / k_lb_fft3) turn it into its spectrum in place through
// the scratch Sb of M values, and a post pass forms the DCT coefficients or samples.
struct LongPlan {
  int N, N1, N2;   // line length; the FFT factors (of N, or of M for Bluestein)
  int f1, f2;      // DctPlan indices whose FftPlans transform lengths N1 and N2
  int tw_off;      // split twiddle table of length N (M for Bluestein): exp(-2 pi i b / L), b < 64, then exp(-2 pi i 64 a / L)
  int mk_off;      // N entries exp(-i pi k / 2N)
  int lb1, lb2;    // log2 of the FFTs per block (slots) in pass 1 and pass 2
  int nb1, nb2;    // blocks per line pair in pass 1 and pass 2 (Bluestein: fft1 / fft3, and fft2)
  int spitch;      // float2 elements per line pair in the scratch arena
  float invN;
  float inv_n2;    // 1 / N2 for float quotients
  int blue;        // 1: Bluestein (below)
  int M;           // Bluestein convolution length N1 N2
  int chirp_off;   // Bluestein: N entries exp(-i pi n^2 / N)
  int bhat_off;    // Bluestein: M entries FFT_M(conj chirp, wrapped) / M
  int sb_off;      // Bluestein: float2 offset of Sb inside a line pair's scratch (Sa first)
  int nbp;         // Bluestein: pack / post blocks per line pair
};

// Column DCT pair of length N = N1 N2 on slabs of the row-major field (k_slab1/2/3).
struct SlabPlan {
  int N, N1, N2;
  int f1, f2;    // DctPlan indices whose FftPlans transform lengths N1 and N2
  int tw_off;    // split twiddle table of length N: exp(-2 pi i b / N), b < 64, then exp(-2 pi i 64 a / N)
  int mk_off;    // N entries exp(-i pi k / 2N)
  int lb1;       // log2 of the slots of a K1 / K3 block: 2^(lb1 - 5) values of n2 (or m2) x SLAB_CP pairs
  int nb1;       // K1 / K3 blocks per slab
  int nb2;       // K2 blocks per slab: one per pair {k1, N1 - k1}, k1 = 0 .. N1 / 2
  float invN;
};

struct ImgDesc {
  long long off;   // element offset of the image in every field arena (multiple of 64)
  long long offT;  // element offset of the image in the transposed arena (multiple of 64)
  long long offS;  // float2 offset of the image in the long-transform scratch arena
  int H, W;
  int P;           // row pitch in elements: W rounded up to a multiple of 64
  int PT;          // row pitch of the transposed image: H rounded up to a multiple of 64
  int rowplan;     // DctPlan index for length W (-1 when W > MAX_FFT_LEN)
  int colplan;     // DctPlan index for length H (-1 when H > MAX_FFT_LEN)
  int rowlong;     // LongPlan index for length W when W > MAX_FFT_LEN, else -1
  int collong;     // LongPlan index for length H when H > MAX_FFT_LEN, else -1
  int cosH_off;    // double offset of cos(pi i / H), i < H
  int cosW_off;    // double offset of cos(pi j / W), j < W
  int colslab;     // SlabPlan index when the column pass runs on slabs (offT is then its scratch), else -1
  int nslab;       // slabs of the image: ceil(W / (2 SLAB_CP))
  int rowpart0;    // first block-partial slot of the image's row-pair blocks (short rows)
  int coldirect;   // 1: the column DCTs read column pairs straight from the row-major field
  int lbpart;      // Bluestein long rows: first slot of the pack / post block partials (partials2)
  int coltri;      // 1: the column half of the preconditioner runs as tridiagonal solves (k_col_trc)
  int tri_nb;      // tridiagonal solves: bands of TRC_CH chunks the columns are cut into (1: one pass)
  long long tri_off;  // double offset of the image's band carries (6 x tri_nb x W) when tri_nb > 1
  long long tri_dc;   // double offset of the image's column-0 input (H floats) for the singular solve
  int xdefer;      // 1: the PCG step's x += alpha p runs in the next step's k_cg_pap (or k_cg_xcommit)
                   // instead of the row DCT-II pass (images whose row transform is bandwidth-bound)
  int pad_;
};

// Per-image solver state.  Written by k_state_init and k_finalize (one block per image), read by
// the other kernels to decide whether to run.
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
  int xpend;  // xdefer images: x += alpha p of the last PCG step is still to be applied
  int pad1;
};

struct Fields {
  float *tx, *ty, *wx, *wy;     // wrapped target gradients, static edge weights
  float *u, *u_prev, *u_best;   // solution (the caller's output arena), copies
  float *wxi, *wyi, *b;         // IRLS edge weights, right-hand side
  float *r, *z, *p0, *p1, *Ap;  // PCG vectors (p ping-pong)
  float *T, *Tt;                // DCT row-pass output and its transpose (or the slab scratch)
  unsigned char *valid;
};

// Block start of each image per launch geometry, [n + 1] each.  The transform passes run one
// launch per class of images that share a kernel instantiation (elements per thread, direct or
// transposed columns, slab split); rowpair / colpair / slab1 / slab2 then point at that class's
// table, in which the images of other classes have no blocks.
struct Geo {
  const int *tile, *rowpair, *colpair, *trans;
  const int *rowL1, *rowL2, *colL1, *colL2;  // long-line passes (images with W or H > MAX_FFT_LEN)
  const int *slab1, *slab2;                  // slab column passes: K1 / K3, and K2
  // Bluestein long lines: pack / post, fft1 / fft3 and fft2 blocks, rows then columns
  const int *browP, *brow1, *brow2, *bcolP, *bcol1, *bcol2;
  // tridiagonal column solves, TRC_COLS columns x one band per block: the final solves per chunk
  // length (4, 6, 8, 10, 12 rows), the band transfers, the band carry chains, column 0
  const int *coltri[9], *tritr, *tricarry, *tridc;  // coltri: one table per chunk length class
};

struct Ctx {
  const ImgDesc *img;
  ImgState *st;
  const DctPlan *plans;
  const LongPlan *lplans;
  const SlabPlan *splans;
  const float2 *tab;
  const double *dtab;
  double *partials;
  double *partials2;  // block partials of the long-line passes (their own block numbering)
  float2 *S;          // long-transform scratch arena
  double *tri;        // band carries of the tridiagonal column solves
  Geo geo;
  Fields f;
  unsigned char *ws_base;  // workspace base (host side: class tables are offsets from it)
  int n;
  float eps_sq;
  double tol, cg_tol;
};

#ifdef __HIPCC__

// Logical block index such that each XCD (which receives hardware blocks b, b + 8, b + 16, ...)
// works on one contiguous range of logical blocks: neighbouring tiles then share that XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int G) {
  const int q = G / NUM_XCD, r = G % NUM_XCD;
  const int x = b % NUM_XCD, i = b / NUM_XCD;
  return x * q + min(x, r) + i;
}

// Row of the row-pass output field T that holds Makhoul position n of a slab-pass image: the
// positions N2 n1 + n2 sit at row n2 N1 + n1, so that the strided subsequences the four-step
// passes transform are runs of consecutive rows.
__device__ __forceinline__ int slab_trow(int n, const SlabPlan &L) { return (n % L.N2) * L.N1 + n / L.N2; }

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

// Block-wide sum of NV doubles; every thread receives the totals.  All threads must call.  MW:
// the most waves a workgroup of the calling kernel has (the array is sized for it only: a larger
// one made the compiler spill in the 8192-point row kernel).
template <int NV, int MW = MAX_FFT_THREADS / 64>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double sh[NV][MW];
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

// First stage of a deterministic two-stage reduction: the block total of each value goes to
// partials[blockIdx.x * MAX_NV + q]; k_finalize sums an image's partials in block order.
template <int NV, int MW = MAX_FFT_THREADS / 64>
__device__ __forceinline__ void block_partial(double (&v)[NV], double *partials, int slot) {
  block_sum<NV, MW>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) partials[(long long)slot * MAX_NV + q] = v[q];
  }
}

template <int NV, int MW = MAX_FFT_THREADS / 64>
__device__ __forceinline__ void block_partial(double (&v)[NV], double *partials) {
  block_partial<NV, MW>(v, partials, blockIdx.x);
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
