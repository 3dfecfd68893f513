// DCT-II / DCT-III of lines longer than the LDS-resident limit (MAX_FFT_LEN), as a four-step
// FFT in two global passes (gfx950).
//
// A pair of real lines (a, b) is packed as one complex Makhoul sequence v of length N = N1 N2
// (v[n] = x[2n] for 2n < N, v[N - 1 - n] = x[2n + 1]), exactly as the LDS-resident transform does.
// Viewing n = N2 n1 + n2 and k = k1 + N1 k2:
//   pass 1: for each column n2, the length-N1 DFT over n1, times exp(-2 pi i n2 k1 / N),
//           stored at row k1, column n2 of the complex scratch line;
//   pass 2: for each row k1, the length-N2 DFT over n2, giving the spectrum at k1 + N1 k2.
// Each workgroup runs a batch of B = 2^lb short FFTs in LDS, interleaved (element i of FFT s at
// buf[i (B + 1) + s]), so consecutive lanes touch consecutive LDS words in every pass.
//
// The inputs and outputs of the DCT come in pairs: the Makhoul packing pairs x[2n] and x[2n+1]
// (positions n and N - 1 - n), the DCT-II post-step and the DCT-III pre-step pair k and N - k.
// Both pairings map a row or column index to a reflection of it, so a block takes C
// representatives and their reflections (B = 2C slots); each global element is then read and
// written by exactly one block, in runs of consecutive addresses.
#pragma once

#include "fft_lds.hpp"

namespace irls2d {

enum LongMode { LONG_DCT2 = 0, LONG_DCT2_UPDATE = 1, LONG_DCT3 = 2, LONG_DCT3_SCALE = 3 };

// One Stockham pass of radix R over B = 2^lb interleaved sequences of length L (BS = B + 1).
template <int R, int E>
__device__ __forceinline__ void batch_pass(float2 *buf, int BS, int lb, const float2 *tws, const PassDesc &pd) {
  constexpr int MB = (E + R - 1) / R;
  const int T = blockDim.x, tid = threadIdx.x;
  const int nb = pd.nb, Ns = pd.Ns;
  const int nbf = nb << lb, smask = (1 << lb) - 1;
  float2 v[MB][R];
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int g = tid + q * T;
    if (g < nbf) {
      const int s = g & smask, j = g >> lb;
#pragma unroll
      for (int t = 0; t < R; ++t) v[q][t] = buf[__umul24(j + __umul24(t, nb), BS) + s];
      if (Ns > 1) stockham_twiddle<R>(v[q], tws, pd, stockham_k(j, pd));
      Dft<R>::run(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int g = tid + q * T;
    if (g < nbf) {
      const int s = g & smask, j = g >> lb;
      const int k = (Ns == 1) ? 0 : stockham_k(j, pd);
      const int base = __umul24(j - k, R) + k;
#pragma unroll
      for (int t = 0; t < R; ++t) buf[__umul24(base + __umul24(t, Ns), BS) + s] = v[q][t];
    }
  }
  __syncthreads();
}

// Compile-time form of batch_pass for a constant length, radix, stride product and batch (2^LB
// sequences): every index, quotient and twiddle step is a constant.  Same arithmetic.
template <int R, int NS, int L, int LB, int E>
__device__ __forceinline__ void batch_pass_ct(float2 *buf, const float2 *tws) {
  constexpr int B = 1 << LB, BS = B + 1, NB = L / R, NBF = NB << LB, MB = (E + R - 1) / R;
  constexpr int TWMUL = L / (NS * R);
  const int T = blockDim.x, tid = threadIdx.x;
  float2 v[MB][R];
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int g = tid + q * T;
    if (g < NBF) {
      const int s = g & (B - 1), j = g >> LB;
#pragma unroll
      for (int t = 0; t < R; ++t) v[q][t] = buf[(j + t * NB) * BS + s];
      if (NS > 1) twiddle_inputs<R>(v[q], tws, (j % NS) * TWMUL);
      Dft<R>::run(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int g = tid + q * T;
    if (g < NBF) {
      const int s = g & (B - 1), j = g >> LB;
      const int k = NS == 1 ? 0 : j % NS;
      const int base = (j - k) * R + k;
#pragma unroll
      for (int t = 0; t < R; ++t) buf[(base + t * NS) * BS + s] = v[q][t];
    }
  }
  __syncthreads();
}

// Forward FFTs of all B sequences.  Requires B * L <= E * blockDim.x.  The default plan of
// 200-point lines in batches of 32 (40000 = 200 x 200) runs the compile-time passes.
template <int E>
__device__ __forceinline__ void batch_fft(float2 *buf, int BS, int lb, const float2 *tws, const FftPlan *P) {
  if (P->L == 200 && lb == 5 && P->npass == 3 && P->pass[0].R == 8 && P->pass[1].R == 5 && P->pass[2].R == 5) {
    batch_pass_ct<8, 1, 200, 5, E>(buf, tws);
    batch_pass_ct<5, 8, 200, 5, E>(buf, tws);
    batch_pass_ct<5, 40, 200, 5, E>(buf, tws);
    return;
  }
  const int np = P->npass;
  for (int s = 0; s < np; ++s) {
    const PassDesc pd = P->pass[s];
    switch (pd.R) {
      case 16: batch_pass<16, E>(buf, BS, lb, tws, pd); break;
      case 8: batch_pass<8, E>(buf, BS, lb, tws, pd); break;
      case 4: batch_pass<4, E>(buf, BS, lb, tws, pd); break;
      case 2: batch_pass<2, E>(buf, BS, lb, tws, pd); break;
      case 3: batch_pass<3, E>(buf, BS, lb, tws, pd); break;
      default: batch_pass<5, E>(buf, BS, lb, tws, pd); break;
    }
  }
}

// Slot s of a block -> index along the batched dimension (length L).  Slots s < C hold the
// representatives r = blk C + s, slots C + s their reflections refl(r) = (o - r) mod L with
// o = L - 1 (pair n, N - 1 - n) or o = 0 (pair k, N - k).  Returns -1 for an empty slot (past the
// last representative, or the reflection of a self-reflected index).
__device__ __forceinline__ int slot_index(int s, int C, int blk, int L, int o) {
  const int nrep = (o == 0) ? L / 2 + 1 : (L + 1) / 2;
  const int r = blk * C + (s < C ? s : s - C);
  if (r >= nrep) return -1;
  if (s < C) return r;
  const int x = (o == 0) ? (r == 0 ? 0 : L - r) : L - 1 - r;
  return x == r ? -1 : x;
}

// The slot holding refl(index of slot s): the partner slot, or s itself for a self-reflected index.
__device__ __forceinline__ int slot_partner(int s, int C, int idx, int L, int o) {
  const int x = (o == 0) ? (idx == 0 ? 0 : L - idx) : L - 1 - idx;
  if (x == idx) return s;
  return s < C ? s + C : s - C;
}

// Split-table twiddle exp(-2 pi i m / N) from global memory (tw: 64 low entries, then high).
__device__ __forceinline__ float2 long_twiddle(const float2 *__restrict__ tw, int m) {
  return cmul(tw[TW_LO + (m >> 6)], tw[m & 63]);
}

static_assert(LONG_THREADS % LONG_MAX_BATCH == 0, "a thread must keep its slot across rows");

struct LongLine {
  const LongPlan *L;
  int img, lp, blk;  // image, line pair, block within the line pair
  int gb;            // logical block index of the launch (partials slot)
  long long la;      // element offset of line a (line b = la + pitch)
  long long ta, tb;  // offsets of lines a and b in the row-pass output T (its slab row order for
                     // an image whose column pass runs on slabs, else la and la + pitch)
  long long pitch;
  bool hasb;
  float2 *S;         // scratch line of this pair
};

// The line pair and block of this workgroup in a launch over `geo`, whose images have `which`
// blocks per line pair: 1 = nb1, 2 = nb2, 0 = nbp (Bluestein pack / post).  `local` (optional)
// receives the block's index among its image's blocks.  Logical blocks are numbered XCD-major
// (xcd_block): the blocks of one line pair, whose row segments share cache lines at their ends,
// then run on one XCD at about the same time and fetch those lines once into its L2.
__device__ __forceinline__ bool long_line_geo(const Ctx &c, int dir, const int *geo, int which, int guard,
                                              LongLine &ln, int *local_out = nullptr) {
  ln.gb = xcd_block(blockIdx.x, gridDim.x);
  ln.img = find_img(geo, c.n, ln.gb);
  const ImgDesc &d = c.img[ln.img];
  if (guard_off(&c.st[ln.img], guard)) return false;
  ln.L = c.lplans + (dir ? d.collong : d.rowlong);
  const int nbp = which == 1 ? ln.L->nb1 : (which == 2 ? ln.L->nb2 : ln.L->nbp);
  const int local = ln.gb - geo[ln.img];
  if (local_out) *local_out = local;
  ln.lp = local / nbp;
  ln.blk = local - ln.lp * nbp;
  const int nlines = dir ? d.W : d.H;
  ln.pitch = dir ? d.PT : d.P;
  ln.la = (dir ? d.offT : d.off) + (long long)(2 * ln.lp) * ln.pitch;
  ln.hasb = 2 * ln.lp + 1 < nlines;
  ln.ta = ln.la;
  ln.tb = ln.la + ln.pitch;
  if (!dir && d.colslab >= 0) {
    const SlabPlan &S = c.splans[d.colslab];
    ln.ta = d.off + (long long)slab_trow(makhoul_pos(2 * ln.lp, d.H), S) * d.P;
    ln.tb = ln.hasb ? d.off + (long long)slab_trow(makhoul_pos(2 * ln.lp + 1, d.H), S) * d.P : ln.ta;
  }
  ln.S = c.S + d.offS + (long long)ln.lp * ln.L->spitch;
  return true;
}

__device__ __forceinline__ bool long_line(const Ctx &c, int dir, int pass, int guard, LongLine &ln) {
  const int *geo = dir ? (pass == 1 ? c.geo.colL1 : c.geo.colL2) : (pass == 1 ? c.geo.rowL1 : c.geo.rowL2);
  return long_line_geo(c, dir, geo, pass, guard, ln);
}

__device__ __forceinline__ void load_tws(float2 *tws, const float2 *__restrict__ tab, const FftPlan *F) {
  const int n = TW_LO + ((F->L + 63) >> 6);
  for (int i = threadIdx.x; i < n; i += blockDim.x) tws[i] = tab[F->tw_off + i];
}

// float2 load of elements m, m + 1 of a line (m even; m + 1 may be the pitch padding).
__device__ __forceinline__ float2 ld2(const float *p, long long m) { return *(const float2 *)(p + m); }

// Stores elements m and, when `two`, m + 1 (never the padding after the last element).
__device__ __forceinline__ void st2(float *p, long long m, float2 v, bool two) {
  if (two) *(float2 *)(p + m) = v;
  else p[m] = v.x;
}

// Pass 1.  LONG_DCT2: src lines -> Makhoul packing.  LONG_DCT2_UPDATE: the PCG update
// x += alpha p, r -= alpha Ap (utils_unwrap2d.py:1444-1458) on the loaded elements first, with the
// partial sums of r.r and of non-finite x.  LONG_DCT3(_SCALE): DCT-III pre-step from the
// coefficients at k and N - k; _SCALE first divides them by -lambda (the column pass of the
// preconditioner, utils_unwrap2d.py:1395-1398) with lambda's DC entry giving 0.
// Global loads are issued in chunks of LONG_U rows before any of them is used, so that each
// thread keeps several requests in flight.
constexpr int LONG_U = 4;

template <int MODE, int E>
__global__ __launch_bounds__(LONG_THREADS) void k_long_p1(Ctx c, int dir, float *src, const float *__restrict__ p,
                                                          int guard) {
  extern __shared__ float2 lds[];
  LongLine ln;
  if (!long_line(c, dir, 1, guard, ln)) return;
  const LongPlan &L = *ln.L;
  const int N = L.N, N1 = L.N1, N2 = L.N2, lb = L.lb1, B = 1 << lb, C = B >> 1, BS = B + 1;
  const FftPlan *F = &c.plans[L.f1].fft;
  float2 *tws = lds, *buf = lds + TW_LDS;
  load_tws(tws, c.tab, F);
  constexpr bool dct2 = (MODE == LONG_DCT2 || MODE == LONG_DCT2_UPDATE);
  constexpr int U = LONG_U;
  const int o = dct2 ? N2 - 1 : 0;
  const float *a = src + ln.ta;  // the DCT-III modes read T (or Tt) lines
  const float *b = src + ln.tb;
  const int T = blockDim.x;
  double acc[2] = {0.0, 0.0};
  // blockDim.x is a multiple of B, so a thread keeps its slot (one FFT of the batch) and walks
  // rows by T / B.  Only rows n1 <= N1 / 2 hold canonical elements (2n < N, or k <= N - k); the
  // others are filled through the partner.
  const int s = threadIdx.x & (B - 1), rstep = T >> lb;
  const int col = slot_index(s, C, ln.blk, N2, o);
  const int sp = col >= 0 ? slot_partner(s, C, col, N2, o) : 0;
  const int n1_end = col >= 0 ? N1 / 2 + 1 : 0;
  if (dct2) {
    const float alpha = (MODE == LONG_DCT2_UPDATE) ? (float)c.st[ln.img].alpha : 0.f;
    const bool upx = !c.img[ln.img].xdefer;  // xdefer: x is updated by the next k_cg_pap instead
    float *x = c.f.u;
    const float *Ap = c.f.Ap;
    for (int n0 = threadIdx.x >> lb; n0 < n1_end; n0 += U * rstep) {
      float2 v[U][2], xv[U][2], pv[U][2], av[U][2];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int n1 = n0 + u * rstep;
        const int n = N2 * n1 + col;
        const long long m = 2 * (long long)n;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          v[u][h] = xv[u][h] = pv[u][h] = av[u][h] = make_float2(0.f, 0.f);
          if (n1 < n1_end && 2 * n < N && (h == 0 || ln.hasb)) {
            const long long e = ln.la + (long long)h * ln.pitch + m;
            v[u][h] = ld2(src, e);
            if (MODE == LONG_DCT2_UPDATE) {
              if (upx) {
                xv[u][h] = ld2(x, e);
                pv[u][h] = ld2(p, e);
              }
              av[u][h] = ld2(Ap, e);
            }
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int n1 = n0 + u * rstep;
        const int n = N2 * n1 + col;
        if (!(n1 < n1_end && 2 * n < N)) continue;
        const long long m = 2 * (long long)n;
        const bool two = 2 * n + 1 < N;
        if (MODE == LONG_DCT2_UPDATE) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            if (h == 1 && !ln.hasb) break;
            const long long e = ln.la + (long long)h * ln.pitch + m;
            const float2 rv = make_float2(v[u][h].x - alpha * av[u][h].x, v[u][h].y - alpha * av[u][h].y);
            if (upx) {
              const float2 xn = make_float2(xv[u][h].x + alpha * pv[u][h].x, xv[u][h].y + alpha * pv[u][h].y);
              st2(x, e, xn, two);
              acc[1] += (isfinite(xn.x) ? 0.0 : 1.0) + ((two && !isfinite(xn.y)) ? 1.0 : 0.0);
            }
            st2(src, e, rv, two);
            acc[0] += (double)rv.x * rv.x + (two ? (double)rv.y * rv.y : 0.0);
            v[u][h] = rv;
          }
        }
        buf[n1 * BS + s] = make_float2(v[u][0].x, v[u][1].x);
        if (two) buf[(N1 - 1 - n1) * BS + sp] = make_float2(v[u][0].y, v[u][1].y);
      }
    }
  } else {
    const float2 *mk = c.tab + L.mk_off;
    const ImgDesc &d = c.img[ln.img];
    const double *cosH = c.dtab + d.cosH_off;
    const int ja = 2 * ln.lp;
    // lines are image columns ja, ja + 1 of length H = N; lambda = 2cos(pi k/H) + 2cos(pi j/W) - 4
    const double cwa = (MODE == LONG_DCT3_SCALE) ? 2.0 * c.dtab[d.cosW_off + ja] : 0.0;
    const double cwb = (MODE == LONG_DCT3_SCALE && ln.hasb) ? 2.0 * c.dtab[d.cosW_off + ja + 1] : 0.0;
    for (int n0 = threadIdx.x >> lb; n0 < n1_end; n0 += U * rstep) {
      float X[U][4];
      float2 wk[U], wn[U];
      double ck[U], cn[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int n1 = n0 + u * rstep;
        const int k = N2 * n1 + col;
        const int nk = k ? N - k : 0;
        X[u][0] = X[u][1] = X[u][2] = X[u][3] = 0.f;
        wk[u] = wn[u] = make_float2(1.f, 0.f);
        ck[u] = cn[u] = 0.0;
        if (n1 < n1_end && (k == 0 || 2 * k <= N)) {
          X[u][0] = a[k];
          X[u][1] = ln.hasb ? b[k] : 0.f;
          X[u][2] = k ? a[nk] : 0.f;
          X[u][3] = (k && ln.hasb) ? b[nk] : 0.f;
          wk[u] = mk[k];
          wn[u] = mk[nk];
          if (MODE == LONG_DCT3_SCALE) {
            ck[u] = 2.0 * cosH[k];
            cn[u] = 2.0 * cosH[nk];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int n1 = n0 + u * rstep;
        const int k = N2 * n1 + col;
        if (!(n1 < n1_end && (k == 0 || 2 * k <= N))) continue;
        const int nk = k ? N - k : 0;
        float Xak = X[u][0], Xbk = X[u][1], Xan = X[u][2], Xbn = X[u][3];
        if (MODE == LONG_DCT3_SCALE) {
          Xak = (k == 0 && ja == 0) ? 0.f : -(Xak * __builtin_amdgcn_rcpf((float)(ck[u] + cwa - 4.0)));
          Xbk = ln.hasb ? -(Xbk * __builtin_amdgcn_rcpf((float)(ck[u] + cwb - 4.0))) : 0.f;
          if (k) {
            Xan = -(Xan * __builtin_amdgcn_rcpf((float)(cn[u] + cwa - 4.0)));
            Xbn = ln.hasb ? -(Xbn * __builtin_amdgcn_rcpf((float)(cn[u] + cwb - 4.0))) : 0.f;
          }
        }
        buf[n1 * BS + s] = dct3_pre_w(wk[u], Xak, Xbk, Xan, Xbn);
        if (nk != k) {
          const int n1p = (col == 0) ? N1 - n1 : N1 - 1 - n1;  // row of nk = N2 n1' + refl(col)
          buf[n1p * BS + sp] = dct3_pre_w(wn[u], Xan, Xbn, Xak, Xbk);
        }
      }
    }
  }
  __syncthreads();
  batch_fft<E>(buf, BS, lb, tws, F);
  const float2 *tw = c.tab + L.tw_off;
  for (int k0 = col >= 0 ? (int)(threadIdx.x >> lb) : N1; k0 < N1; k0 += U * rstep) {
    float2 v[U], w[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k1 = k0 + u * rstep;
      v[u] = w[u] = make_float2(1.f, 0.f);
      if (k1 < N1) {
        v[u] = buf[k1 * BS + s];
        if (k1 && col) w[u] = long_twiddle(tw, k1 * col);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k1 = k0 + u * rstep;
      if (k1 < N1) ln.S[k1 * N2 + col] = (k1 && col) ? cmul(v[u], w[u]) : v[u];
    }
  }
  if (MODE == LONG_DCT2_UPDATE) block_partial<2>(acc, c.partials2, ln.gb);
}

// Pass 2.  LONG_DCT2: DCT-II coefficients from the spectrum at k and N - k -> dst lines.
// LONG_DCT3: inverse Makhoul packing -> dst lines, with the partial sums of r.dst when `dot`.
template <int MODE, int E>
__global__ __launch_bounds__(LONG_THREADS) void k_long_p2(Ctx c, int dir, float *__restrict__ dst, int dot,
                                                          int guard) {
  extern __shared__ float2 lds[];
  LongLine ln;
  if (!long_line(c, dir, 2, guard, ln)) return;
  const LongPlan &L = *ln.L;
  const int N = L.N, N1 = L.N1, N2 = L.N2, lb = L.lb2, B = 1 << lb, C = B >> 1, BS = B + 1;
  const FftPlan *F = &c.plans[L.f2].fft;
  float2 *tws = lds, *buf = lds + TW_LDS;
  load_tws(tws, c.tab, F);
  constexpr bool dct2 = (MODE == LONG_DCT2);
  constexpr int U = LONG_U;
  const int o = dct2 ? 0 : N1 - 1;
  const int T = blockDim.x;
  // rows of the scratch line, each a contiguous run of N2 values
  for (int i0 = threadIdx.x; i0 < (N2 << lb); i0 += U * T) {
    float2 v[U];
    int pos[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      // unconditional loads: an index past the batch or an empty slot reads row 0 (not stored)
      const int idx0 = i0 + u * T;
      const bool in = idx0 < (N2 << lb);
      const int idx = in ? idx0 : 0;
      int q = (int)((float)idx * L.inv_n2);
      int i = idx - q * N2;
      if (i < 0) { --q; i += N2; }
      if (i >= N2) { ++q; i -= N2; }
      const int row = slot_index(q, C, ln.blk, N1, o);
      v[u] = ln.S[(row >= 0 ? row : 0) * N2 + i];
      pos[u] = (in && row >= 0) ? i * BS + q : -1;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (pos[u] >= 0) buf[pos[u]] = v[u];
  }
  __syncthreads();
  batch_fft<E>(buf, BS, lb, tws, F);
  // the DCT-II writes T (or Tt) lines, the DCT-III the output field in row order
  float *da = dst + (dct2 ? ln.ta : ln.la), *db = dst + (dct2 ? ln.tb : ln.la + ln.pitch);
  double acc[1] = {0.0};
  const int s = threadIdx.x & (B - 1), rstep = T >> lb;
  const int k1 = slot_index(s, C, ln.blk, N1, o);
  const int sp = k1 >= 0 ? slot_partner(s, C, k1, N1, o) : 0;
  // canonical elements (k <= N - k, or 2n < N) all have k2 <= N2 / 2
  const int k2_end = k1 >= 0 ? N2 / 2 + 1 : 0;
  const float2 *mk = c.tab + L.mk_off;
  for (int q0 = threadIdx.x >> lb; q0 < k2_end; q0 += U * rstep) {
    if (dct2) {
      float2 wk[U], wn[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k2 = q0 + u * rstep;
        const int kr = k1 + N1 * k2;
        const int k = (k2 < k2_end && (kr == 0 || 2 * kr <= N)) ? kr : 0;  // unconditional loads
        wk[u] = mk[k];
        wn[u] = mk[k ? N - k : 0];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k2 = q0 + u * rstep;
        const int k = k1 + N1 * k2;
        if (!(k2 < k2_end && (k == 0 || 2 * k <= N))) continue;
        const int nk = k ? N - k : 0;
        // nk = k1' + N1 k2' with k1' = refl(k1) and k2' = -k2 mod N2 (k1 = 0) or N2 - 1 - k2
        const int k2p = k1 ? N2 - 1 - k2 : (k2 ? N2 - k2 : 0);
        const float2 Zk = buf[k2 * BS + s], Zn = buf[k2p * BS + sp];
        float Xak, Xbk, Xan, Xbn;
        dct2_coef_w(Zk, Zn, wk[u], wn[u], Xak, Xbk, Xan, Xbn);
        da[k] = Xak;
        if (ln.hasb) db[k] = Xbk;
        if (nk != k) {
          da[nk] = Xan;
          if (ln.hasb) db[nk] = Xbn;
        }
      }
    } else {
      float2 ra[U], rb[U];
      if (dot) {
        // unconditional loads: an element past the canonical ones reads element k1 (2 k1 < N), a
        // missing line b reads line a; neither is used
        const float *r = c.f.r + ln.la;
        const float *rbp = ln.hasb ? r + ln.pitch : r;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int k2 = q0 + u * rstep;
          const int n = k1 + N1 * k2;
          const long long m = 2 * (long long)((k2 < k2_end && 2 * n < N) ? n : k1);
          ra[u] = ld2(r, m);
          rb[u] = ld2(rbp, m);
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k2 = q0 + u * rstep;
        const int n = k1 + N1 * k2;
        if (!(k2 < k2_end && 2 * n < N)) continue;  // written with its partner N - 1 - n
        const bool two = 2 * n + 1 < N;
        const float2 w0 = buf[k2 * BS + s];
        const float2 w1 = two ? buf[(N2 - 1 - k2) * BS + sp] : make_float2(0.f, 0.f);
        const float2 va = make_float2(w0.x * L.invN, w1.x * L.invN);
        const float2 vb = make_float2(-w0.y * L.invN, -w1.y * L.invN);
        const long long m = 2 * (long long)n;
        st2(da, m, va, two);
        if (ln.hasb) st2(db, m, vb, two);
        if (dot) {
          acc[0] += (double)ra[u].x * va.x + (two ? (double)ra[u].y * va.y : 0.0);
          if (ln.hasb) acc[0] += (double)rb[u].x * vb.x + (two ? (double)rb[u].y * vb.y : 0.0);
        }
      }
    }
  }
  if (!dct2 && dot) block_partial<1>(acc, c.partials2, ln.gb);
}

// ---------------------------------------------------------------------------------------------
// Long lines without a smooth two-factor split (LongPlan::blue): the length-N FFT of the packed
// sequence u as Bluestein's chirp convolution, U[m] = w[m] conj(sum_k conj(Y[k] H[k]) e^{-2 pi i k m / M})
// with w[n] = exp(-i pi n^2 / N), Y = FFT_M(u w zero-padded), H = FFT_M(conj w wrapped) / M, computed
// four-step with n = N2 n1 + n2 (forward) and k = k1 + N1 k2, m = m2 + N2 m1 (inverse):
//   k_lb_fft1  per column n2: FFT over n1 of u[n] w[n], times exp(-2 pi i n2 k1 / M) -> Sb[k1 N2 + n2]
//   k_lb_fft2  per row k1: FFT over n2 -> Y[k1 + N1 k2]; conj(Y H); FFT over k2 -> m2, times
//              exp(-2 pi i k1 m2 / M) -> Sb[k1 N2 + m2]
//   k_lb_fft3  per column m2: FFT over k1 -> m1; U[m] = w[m] conj(.) for m < N -> Sa[m]
// around them k_lb_pack writes u to Sa (the Makhoul packing of both lines, or the DCT-III
// pre-step) and k_lb_post forms the DCT-II coefficients (from U[k], U[N - k]) or the DCT-III
// samples (from U at each sample's Makhoul position).  Sa holds N values per line pair, Sb M.
// ---------------------------------------------------------------------------------------------

constexpr int LB_CHUNK = 4096;  // sequence elements (or coefficient pairs) per pack / post block

template <int MODE>
__global__ __launch_bounds__(LONG_THREADS) void k_lb_pack(Ctx c, int dir, float *src, const float *__restrict__ p,
                                                          int guard) {
  LongLine ln;
  int local = 0;
  if (!long_line_geo(c, dir, dir ? c.geo.bcolP : c.geo.browP, 0, guard, ln, &local)) return;
  const LongPlan &L = *ln.L;
  const int N = L.N, T = blockDim.x;
  float2 *__restrict__ Sa = ln.S;
  constexpr bool dct2 = (MODE == LONG_DCT2 || MODE == LONG_DCT2_UPDATE);
  constexpr int U = LONG_U;
  double acc[2] = {0.0, 0.0};
  if (dct2) {
    // element m of both lines -> Makhoul slot makhoul_pos(m); the update first
    const float alpha = (MODE == LONG_DCT2_UPDATE) ? (float)c.st[ln.img].alpha : 0.f;
    const bool upx = !c.img[ln.img].xdefer;  // xdefer: x is updated by the next k_cg_pap instead
    float *x = c.f.u;
    const float *Ap = c.f.Ap;
    const int e0 = ln.blk * LB_CHUNK, e1 = min(e0 + LB_CHUNK, N);
    for (int m0 = e0 + (int)threadIdx.x; m0 < e1; m0 += U * T) {
      float v[U][2], xv[U][2], pv[U][2], av[U][2];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int m = m0 + u * T;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          v[u][h] = xv[u][h] = pv[u][h] = av[u][h] = 0.f;
          if (m < e1 && (h == 0 || ln.hasb)) {
            const long long e = ln.la + (long long)h * ln.pitch + m;
            v[u][h] = src[e];
            if (MODE == LONG_DCT2_UPDATE) {
              xv[u][h] = x[e];
              pv[u][h] = p[e];
              av[u][h] = Ap[e];
            }
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int m = m0 + u * T;
        if (m >= e1) continue;
        if (MODE == LONG_DCT2_UPDATE) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            if (h == 1 && !ln.hasb) break;
            const long long e = ln.la + (long long)h * ln.pitch + m;
            const float xn = xv[u][h] + alpha * pv[u][h];
            const float rn = v[u][h] - alpha * av[u][h];
            x[e] = xn;
            src[e] = rn;
            acc[0] += (double)rn * rn;
            acc[1] += isfinite(xn) ? 0.0 : 1.0;
            v[u][h] = rn;
          }
        }
        Sa[makhoul_pos(m, N)] = make_float2(v[u][0], ln.hasb ? v[u][1] : 0.f);
      }
    }
  } else {
    // DCT-III pre-step of the pairs (k, N - k), k = 0 .. N / 2; _SCALE first divides by -lambda
    const float *a = src + ln.ta;
    const float *b = src + ln.tb;
    const float2 *mk = c.tab + L.mk_off;
    const ImgDesc &d = c.img[ln.img];
    const double *cosH = c.dtab + d.cosH_off;
    const int ja = 2 * ln.lp;
    const double cwa = (MODE == LONG_DCT3_SCALE) ? 2.0 * c.dtab[d.cosW_off + ja] : 0.0;
    const double cwb = (MODE == LONG_DCT3_SCALE && ln.hasb) ? 2.0 * c.dtab[d.cosW_off + ja + 1] : 0.0;
    const int e0 = ln.blk * LB_CHUNK, e1 = min(e0 + LB_CHUNK, N / 2 + 1);
    for (int k = e0 + (int)threadIdx.x; k < e1; k += T) {
      const int nk = k ? N - k : 0;
      float Xak = a[k], Xbk = ln.hasb ? b[k] : 0.f;
      float Xan = k ? a[nk] : 0.f, Xbn = (k && ln.hasb) ? b[nk] : 0.f;
      if (MODE == LONG_DCT3_SCALE) {
        const double ck = 2.0 * cosH[k], cn = 2.0 * cosH[nk];
        Xak = (k == 0 && ja == 0) ? 0.f : -(Xak * __builtin_amdgcn_rcpf((float)(ck + cwa - 4.0)));
        Xbk = ln.hasb ? -(Xbk * __builtin_amdgcn_rcpf((float)(ck + cwb - 4.0))) : 0.f;
        if (k) {
          Xan = -(Xan * __builtin_amdgcn_rcpf((float)(cn + cwa - 4.0)));
          Xbn = ln.hasb ? -(Xbn * __builtin_amdgcn_rcpf((float)(cn + cwb - 4.0))) : 0.f;
        }
      }
      Sa[k] = dct3_pre_w(mk[k], Xak, Xbk, Xan, Xbn);
      if (nk != k) Sa[nk] = dct3_pre_w(mk[nk], Xan, Xbn, Xak, Xbk);
    }
  }
  if (MODE == LONG_DCT2_UPDATE) block_partial<2>(acc, c.partials2, c.img[ln.img].lbpart + local);
}

// k_lb_fft1 (STAGE 1) and k_lb_fft3 (STAGE 3): a block transforms B = 2^lb1 columns of the
// N1 x N2 matrix over its N1 rows; a thread keeps one column slot and walks the rows.
template <int STAGE, int E>
__global__ __launch_bounds__(LONG_THREADS) void k_lb_fft13(Ctx c, int dir, int guard) {
  extern __shared__ float2 lds[];
  LongLine ln;
  if (!long_line_geo(c, dir, dir ? c.geo.bcol1 : c.geo.brow1, 1, guard, ln)) return;
  const LongPlan &L = *ln.L;
  const int N = L.N, N1 = L.N1, N2 = L.N2, lb = L.lb1, B = 1 << lb, BS = B + 1;
  const FftPlan *F = &c.plans[L.f1].fft;
  float2 *tws = lds, *buf = lds + TW_LDS;
  load_tws(tws, c.tab, F);
  float2 *__restrict__ Sa = ln.S;
  float2 *__restrict__ Sb = ln.S + L.sb_off;
  const float2 *__restrict__ ch = c.tab + L.chirp_off;
  const int s = threadIdx.x & (B - 1), rstep = blockDim.x >> lb;
  const int col = ln.blk * B + s;  // n2 (stage 1) or m2 (stage 3)
  const bool live = col < N2;
  for (int r = threadIdx.x >> lb; r < N1; r += rstep) {
    float2 v = make_float2(0.f, 0.f);
    if (live) {
      if (STAGE == 1) {
        const int n = N2 * r + col;
        if (n < N) v = cmul(Sa[n], ch[n]);
      } else {
        v = Sb[(long long)r * N2 + col];
      }
    }
    buf[r * BS + s] = v;
  }
  __syncthreads();
  batch_fft<E>(buf, BS, lb, tws, F);
  const float2 *tw = c.tab + L.tw_off;
  for (int r = threadIdx.x >> lb; r < N1; r += rstep) {
    if (!live) continue;
    const float2 v = buf[r * BS + s];
    if (STAGE == 1) {
      Sb[(long long)r * N2 + col] = (r && col) ? cmul(v, long_twiddle(tw, r * col)) : v;  // r = k1
    } else {
      const int m = col + N2 * r;  // r = m1
      if (m < N) Sa[m] = cmul(ch[m], cconj(v));
    }
  }
}

// k_lb_fft2: a block takes B = 2^lb2 rows k1 of the N1 x N2 matrix: FFT over the row, product with
// the transformed chirp, conjugate, FFT over the row again, twiddle, in place.
template <int E>
__global__ __launch_bounds__(LONG_THREADS) void k_lb_fft2(Ctx c, int dir, int guard) {
  extern __shared__ float2 lds[];
  LongLine ln;
  if (!long_line_geo(c, dir, dir ? c.geo.bcol2 : c.geo.brow2, 2, guard, ln)) return;
  const LongPlan &L = *ln.L;
  const int N1 = L.N1, N2 = L.N2, lb = L.lb2, B = 1 << lb, BS = B + 1;
  const FftPlan *F = &c.plans[L.f2].fft;
  float2 *tws = lds, *buf = lds + TW_LDS;
  load_tws(tws, c.tab, F);
  float2 *__restrict__ Sb = ln.S + L.sb_off;
  const float2 *__restrict__ bh = c.tab + L.bhat_off;
  const int T = blockDim.x, nel = N2 << lb;
  // element idx = q N2 + i: row k1 = blk B + q, column i (contiguous runs along the rows)
  auto rowcol = [&](int idx, int &q, int &i) {
    q = (int)((float)idx * L.inv_n2);
    i = idx - q * N2;
    if (i < 0) { --q; i += N2; }
    if (i >= N2) { ++q; i -= N2; }
  };
  for (int idx = threadIdx.x; idx < nel; idx += T) {
    int q, i;
    rowcol(idx, q, i);
    const int k1 = ln.blk * B + q;
    buf[i * BS + q] = k1 < N1 ? Sb[(long long)k1 * N2 + i] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  batch_fft<E>(buf, BS, lb, tws, F);
  for (int idx = threadIdx.x; idx < nel; idx += T) {
    const int q = idx & (B - 1), k2 = idx >> lb;
    const int k1 = ln.blk * B + q;
    if (k1 < N1) buf[k2 * BS + q] = cconj(cmul(buf[k2 * BS + q], bh[k1 + N1 * k2]));
  }
  __syncthreads();
  batch_fft<E>(buf, BS, lb, tws, F);
  const float2 *tw = c.tab + L.tw_off;
  for (int idx = threadIdx.x; idx < nel; idx += T) {
    int q, i;
    rowcol(idx, q, i);  // i = m2
    const int k1 = ln.blk * B + q;
    if (k1 >= N1) continue;
    const float2 v = buf[i * BS + q];
    Sb[(long long)k1 * N2 + i] = (k1 && i) ? cmul(v, long_twiddle(tw, k1 * i)) : v;
  }
}

// Post pass.  LONG_DCT2: coefficients at k and N - k from U[k], U[N - k] -> dst lines (T / Tt
// order).  LONG_DCT3: sample m of both lines from U at its Makhoul position -> dst lines (row
// order), with the partial sums of r.dst when `dot`.
template <int MODE>
__global__ __launch_bounds__(LONG_THREADS) void k_lb_post(Ctx c, int dir, float *__restrict__ dst, int dot,
                                                          int guard) {
  LongLine ln;
  int local = 0;
  if (!long_line_geo(c, dir, dir ? c.geo.bcolP : c.geo.browP, 0, guard, ln, &local)) return;
  const LongPlan &L = *ln.L;
  const int N = L.N, T = blockDim.x;
  const float2 *__restrict__ Sa = ln.S;
  double acc[1] = {0.0};
  if (MODE == LONG_DCT2) {
    float *da = dst + ln.ta, *db = dst + ln.tb;
    const float2 *mk = c.tab + L.mk_off;
    const int e0 = ln.blk * LB_CHUNK, e1 = min(e0 + LB_CHUNK, N / 2 + 1);
    for (int k = e0 + (int)threadIdx.x; k < e1; k += T) {
      const int nk = k ? N - k : 0;
      float Xak, Xbk, Xan, Xbn;
      dct2_coef_w(Sa[k], Sa[nk], mk[k], mk[nk], Xak, Xbk, Xan, Xbn);
      da[k] = Xak;
      if (ln.hasb) db[k] = Xbk;
      if (nk != k) {
        da[nk] = Xan;
        if (ln.hasb) db[nk] = Xbn;
      }
    }
  } else {
    float *da = dst + ln.la, *db = dst + ln.la + ln.pitch;
    const float *r = c.f.r + ln.la;
    const int e0 = ln.blk * LB_CHUNK, e1 = min(e0 + LB_CHUNK, N);
    for (int m = e0 + (int)threadIdx.x; m < e1; m += T) {
      const float2 w = Sa[makhoul_pos(m, N)];
      const float va = w.x * L.invN, vb = -w.y * L.invN;
      da[m] = va;
      if (ln.hasb) db[m] = vb;
      if (dot) acc[0] += (double)r[m] * va + (ln.hasb ? (double)r[ln.pitch + m] * vb : 0.0);
    }
  }
  if (MODE == LONG_DCT3 && dot) block_partial<1>(acc, c.partials2, c.img[ln.img].lbpart + local);
}

}  // namespace irls2d
