// Column pass of the preconditioner straight on the row-major field, without transposes (gfx950).
//
// Columns (ja, ja + 1) form one complex Makhoul sequence of length N = H (v[n] = x[2n] for 2n < N,
// v[N - 1 - n] = x[2n + 1]), as in the LDS-resident column transform; an image's columns are taken
// in slabs of SLAB_CP such pairs, so that every row access is a 256-byte run.  The column DCT-II,
// the division by -lambda and the DCT-III are two forward FFTs of length N = N1 N2, written as
// four-step FFTs and split over three global passes through a scratch copy S:
//   K1 (k_slab1): n = N2 n1 + n2.  For each n2, the length-N1 DFT over n1 of the packed input,
//       times exp(-2 pi i n2 k1 / N), stored at S row k1 N2 + n2.
//   K2 (k_slab2): for each k1, the length-N2 DFT over n2 gives the spectrum Z at k = k1 + N1 k2.
//       A block holds k1 and N1 - k1, so Z_k and Z_{N-k} are both at hand: DCT-II coefficients,
//       division by -lambda, DCT-III pre-step c_k.  Then the first step of the second FFT, with
//       k = k1 + N1 k2 and m = m2 + N2 m1: the length-N2 DFT over k2, times exp(-2 pi i k1 m2 / N),
//       stored back at S row k1 N2 + m2.
//   K3 (k_slab3): for each m2, the length-N1 DFT over k1 gives the DCT-III sequence at
//       n = m2 + N2 m1.
// The row passes keep T of such an image in slab order: Makhoul position n = N2 n1 + n2 at row
// n2 N1 + n1 (slab_trow), so K1 reads and K3 writes runs of consecutive rows.  S is row-major
// with the pitch of T (row q holds the pairs of every slab), so the blocks of neighbouring slabs,
// which run together, write and read whole rows between them.
// Two kernel families: for power-of-two factors (N1 = 16 .. 128, N2 = 16 .. 64) each short FFT
// of length 16 R is two register stages with one LDS exchange between them (reg_stage1/2, all
// strides and radices compile-time); for the other factors the short FFTs run batched in LDS,
// interleaved over the slots of the block (batch_fft of the long-line transform).
#pragma once

#include "long_dct.hpp"

namespace irls2d {

struct SlabBlk {
  const ImgDesc *d;
  const SlabPlan *L;
  int img, grp, slab;
  float2 *S;     // scratch row 0 of this slab
  long long ss;  // float2 between consecutive scratch rows (nslab x SLAB_CP: pitch P floats)
};

__device__ __forceinline__ bool slab_block(const Ctx &c, const int *geo, int guard, SlabBlk &b) {
  b.img = find_img(geo, c.n, blockIdx.x);
  b.d = c.img + b.img;
  if (guard_off(&c.st[b.img], guard)) return false;
  b.L = c.splans + b.d->colslab;
  const int local = blockIdx.x - geo[b.img];
  const int ns = b.d->nslab;
  b.grp = local / ns;
  b.slab = local - b.grp * ns;
  b.S = (float2 *)(c.f.Tt + b.d->offT) + (long long)b.slab * SLAB_CP;
  b.ss = (long long)ns * SLAB_CP;
  return true;
}

constexpr int SLAB_U = 4;  // global loads in flight per thread before their LDS stores

// K1: field T (row DCT-II output) -> S.
template <int E>
__global__ __launch_bounds__(SLAB_MAX_THREADS) void k_slab1(Ctx c, const float *__restrict__ src, int guard) {
  extern __shared__ float2 lds[];
  SlabBlk b;
  if (!slab_block(c, c.geo.slab1, guard, b)) return;
  const SlabPlan &L = *b.L;
  const int N1 = L.N1, N2 = L.N2, lb = L.lb1, B = 1 << lb, BS = B + 1;
  const int W = b.d->W;
  const long long P = b.d->P;
  const FftPlan *F = &c.plans[L.f1].fft;
  float2 *tws = lds, *buf = lds + TW_LDS;
  load_tws(tws, c.tab, F);
  const float *base = src + b.d->off;
  const int T = blockDim.x, tot = N1 << lb;
  const int n2_0 = b.grp << (lb - 5);
  const int jslab = 2 * b.slab * SLAB_CP;
  for (int e0 = threadIdx.x; e0 < tot; e0 += SLAB_U * T) {
    float2 v[SLAB_U];
#pragma unroll
    for (int u = 0; u < SLAB_U; ++u) {
      const int e = e0 + u * T;
      v[u] = make_float2(0.f, 0.f);
      if (e < tot) {
        const int s = e & (B - 1), n1 = e >> lb;
        const int n2 = n2_0 + (s >> 5), ja = jslab + 2 * (s & (SLAB_CP - 1));
        if (n2 < N2 && ja < W) {
          const float *p = base + (long long)(n2 * N1 + n1) * P + ja;  // slab_trow(N2 n1 + n2)
          v[u] = ja + 1 < W ? *(const float2 *)p : make_float2(p[0], 0.f);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < SLAB_U; ++u) {
      const int e = e0 + u * T;
      if (e < tot) buf[(e >> lb) * BS + (e & (B - 1))] = v[u];
    }
  }
  __syncthreads();
  batch_fft<E>(buf, BS, lb, tws, F);
  const float2 *tw = c.tab + L.tw_off;
  for (int e = threadIdx.x; e < tot; e += T) {
    const int s = e & (B - 1), k1 = e >> lb;
    const int n2 = n2_0 + (s >> 5), cp = s & (SLAB_CP - 1);
    if (n2 >= N2) continue;
    float2 v = buf[k1 * BS + s];
    if (k1 && n2) v = cmul(v, long_twiddle(tw, k1 * n2));
    b.S[(k1 * N2 + n2) * b.ss + cp] = v;
  }
}

// K2: S -> S in place (the rows of k1 and N1 - k1).
template <int E>
__global__ __launch_bounds__(SLAB_MAX_THREADS) void k_slab2(Ctx c, int guard) {
  extern __shared__ float2 lds[];
  SlabBlk b;
  if (!slab_block(c, c.geo.slab2, guard, b)) return;
  const SlabPlan &L = *b.L;
  const int N = L.N, N1 = L.N1, N2 = L.N2;
  constexpr int lb = 6, B = 2 * SLAB_CP, BS = B + 1;  // slots: set (k1a, then k1b) x pair
  const int k1a = b.grp, k1b = b.grp ? N1 - b.grp : 0;
  const bool two = k1b != k1a;
  const int W = b.d->W;
  const FftPlan *F = &c.plans[L.f2].fft;
  float2 *tws = lds, *buf = lds + TW_LDS;
  load_tws(tws, c.tab, F);
  const int T = blockDim.x, tot = N2 << lb;
  const float2 *Sa = b.S + (long long)k1a * N2 * b.ss;
  const float2 *Sb = b.S + (long long)k1b * N2 * b.ss;
  for (int e0 = threadIdx.x; e0 < tot; e0 += SLAB_U * T) {
    float2 v[SLAB_U];
#pragma unroll
    for (int u = 0; u < SLAB_U; ++u) {
      const int e = e0 + u * T;
      v[u] = make_float2(0.f, 0.f);
      if (e < tot) {
        const int s = e & (B - 1), n2 = e >> lb, cp = s & (SLAB_CP - 1);
        if (s < SLAB_CP) v[u] = Sa[n2 * b.ss + cp];
        else if (two) v[u] = Sb[n2 * b.ss + cp];
      }
    }
#pragma unroll
    for (int u = 0; u < SLAB_U; ++u) {
      const int e = e0 + u * T;
      if (e < tot) buf[(e >> lb) * BS + (e & (B - 1))] = v[u];
    }
  }
  __syncthreads();
  batch_fft<E>(buf, BS, lb, tws, F);
  // Pairs (k, N - k): with two sets, (k1a, k2) and (k1b, N2 - 1 - k2); in the single set of
  // k1 = 0, (k2, -k2 mod N2); in that of k1 = N1 / 2, (k2, N2 - 1 - k2).
  const int npairs = two ? N2 : (k1a == 0 ? N2 / 2 + 1 : (N2 + 1) / 2);
  const float2 *mk = c.tab + L.mk_off;
  const double *cosH = c.dtab + b.d->cosH_off;
  for (int e = threadIdx.x; e < (npairs << 5); e += T) {
    const int cp = e & (SLAB_CP - 1), k2 = e >> 5;
    const int k2p = (two || k1a) ? N2 - 1 - k2 : (k2 ? N2 - k2 : 0);
    const int sp = two ? SLAB_CP + cp : cp;
    const int k = k1a + N1 * k2;
    const int nk = k ? N - k : 0;
    const int ja = 2 * (b.slab * SLAB_CP + cp);
    float Xak = 0.f, Xbk = 0.f, Xan = 0.f, Xbn = 0.f;
    const float2 wk = mk[k], wn = mk[nk];
    if (ja < W) {
      dct2_coef_w(buf[k2 * BS + cp], buf[k2p * BS + sp], wk, wn, Xak, Xbk, Xan, Xbn);
      // X / -lambda, lambda = 2cos(pi k/H) + 2cos(pi j/W) - 4 rounded to float from float64
      // (utils_unwrap2d.py:1284-1289, :1395-1398); the DC coefficient becomes 0
      const bool hasb = ja + 1 < W;
      const double cwa = 2.0 * c.dtab[b.d->cosW_off + ja];
      const double cwb = hasb ? 2.0 * c.dtab[b.d->cosW_off + ja + 1] : 0.0;
      const double ck = 2.0 * cosH[k], cn = 2.0 * cosH[nk];
      Xak = (k == 0 && ja == 0) ? 0.f : -(Xak * __builtin_amdgcn_rcpf((float)(ck + cwa - 4.0)));
      Xbk = hasb ? -(Xbk * __builtin_amdgcn_rcpf((float)(ck + cwb - 4.0))) : 0.f;
      Xan = (nk == 0 && ja == 0) ? 0.f : -(Xan * __builtin_amdgcn_rcpf((float)(cn + cwa - 4.0)));
      Xbn = hasb ? -(Xbn * __builtin_amdgcn_rcpf((float)(cn + cwb - 4.0))) : 0.f;
    }
    buf[k2 * BS + cp] = dct3_pre_w(wk, Xak, Xbk, k ? Xan : 0.f, k ? Xbn : 0.f);
    if (nk != k) buf[k2p * BS + sp] = dct3_pre_w(wn, Xan, Xbn, Xak, Xbk);
  }
  __syncthreads();
  batch_fft<E>(buf, BS, lb, tws, F);
  const float2 *tw = c.tab + L.tw_off;
  float2 *Da = b.S + (long long)k1a * N2 * b.ss;
  float2 *Db = b.S + (long long)k1b * N2 * b.ss;
  for (int e = threadIdx.x; e < tot; e += T) {
    const int s = e & (B - 1), m2 = e >> lb, cp = s & (SLAB_CP - 1);
    const bool setb = s >= SLAB_CP;
    if (setb && !two) continue;
    const int k1 = setb ? k1b : k1a;
    float2 v = buf[m2 * BS + s];
    if (k1 && m2) v = cmul(v, long_twiddle(tw, k1 * m2));
    (setb ? Db : Da)[m2 * b.ss + cp] = v;
  }
}

// K3: S -> field (the column pass output, read by the row DCT-III).
template <int E>
__global__ __launch_bounds__(SLAB_MAX_THREADS) void k_slab3(Ctx c, float *__restrict__ dst, int guard) {
  extern __shared__ float2 lds[];
  SlabBlk b;
  if (!slab_block(c, c.geo.slab1, guard, b)) return;
  const SlabPlan &L = *b.L;
  const int N1 = L.N1, N2 = L.N2, lb = L.lb1, B = 1 << lb, BS = B + 1;
  const int W = b.d->W;
  const long long P = b.d->P;
  const FftPlan *F = &c.plans[L.f1].fft;
  float2 *tws = lds, *buf = lds + TW_LDS;
  load_tws(tws, c.tab, F);
  const int T = blockDim.x, tot = N1 << lb;
  const int m2_0 = b.grp << (lb - 5);
  for (int e0 = threadIdx.x; e0 < tot; e0 += SLAB_U * T) {
    float2 v[SLAB_U];
#pragma unroll
    for (int u = 0; u < SLAB_U; ++u) {
      const int e = e0 + u * T;
      v[u] = make_float2(0.f, 0.f);
      if (e < tot) {
        const int s = e & (B - 1), k1 = e >> lb;
        const int m2 = m2_0 + (s >> 5);
        if (m2 < N2) v[u] = b.S[(k1 * N2 + m2) * b.ss + (s & (SLAB_CP - 1))];
      }
    }
#pragma unroll
    for (int u = 0; u < SLAB_U; ++u) {
      const int e = e0 + u * T;
      if (e < tot) buf[(e >> lb) * BS + (e & (B - 1))] = v[u];
    }
  }
  __syncthreads();
  batch_fft<E>(buf, BS, lb, tws, F);
  float *base = dst + b.d->off;
  const int jslab = 2 * b.slab * SLAB_CP;
  const float invN = L.invN;
  for (int e = threadIdx.x; e < tot; e += T) {
    const int s = e & (B - 1), m1 = e >> lb;
    const int m2 = m2_0 + (s >> 5), ja = jslab + 2 * (s & (SLAB_CP - 1));
    if (m2 >= N2 || ja >= W) continue;
    const float2 y = buf[m1 * BS + s];
    float *p = base + (long long)(m2 * N1 + m1) * P + ja;  // slab_trow(m2 + N2 m1)
    if (ja + 1 < W) *(float2 *)p = make_float2(y.x * invN, -y.y * invN);
    else p[0] = y.x * invN;
  }
}

// ---------------------------------------------------------------------------------------------
// Power-of-two factors: two-stage register FFTs
// ---------------------------------------------------------------------------------------------

// Split twiddle table of length N (exp(-2 pi i b / N), b < 64, then exp(-2 pi i 64 a / N)) into LDS.
__device__ __forceinline__ void load_tws_len(float2 *tws, const float2 *__restrict__ tab, int N) {
  const int n = TW_LO + ((N + 63) >> 6);
  for (int i = threadIdx.x; i < n; i += blockDim.x) tws[i] = tab[i];
}

// Length L = 16 R FFT, element n = t + R j of a sequence held by thread t < R of its slot s as
// v[j].  Stage 1: DFT-16 over j, times exp(-2 pi i t k1 / L) (from the length-N table, index
// step tstep = N / L), to LDS at ((k1 R + t) SB + s).  Needs a barrier before stage 2.
template <int L>
__device__ __forceinline__ void reg_stage1(float2 (&v)[16], int t, const float2 *tws, int tstep, float2 *buf,
                                           int SB, int s) {
  constexpr int R = L / 16;
  Dft<16>::run(v);
  if (R == 1) return;
  if (t) twiddle_inputs<16>(v, tws, t * tstep);
#pragma unroll
  for (int k1 = 0; k1 < 16; ++k1) buf[(k1 * R + t) * SB + s] = v[k1];
}

// Stage 2: thread q < R of slot s: DFT-R over t of the stage-1 values of k1 = q Q + i (Q = 16 / R);
// o[i R + k2] = X[q Q + i + 16 k2].
template <int L>
__device__ __forceinline__ void reg_stage2(float2 (&o)[16], const float2 (&v)[16], int q, const float2 *buf,
                                           int SB, int s) {
  constexpr int R = L / 16, Q = 16 / R;
  if (R == 1) {
#pragma unroll
    for (int i = 0; i < 16; ++i) o[i] = v[i];
    return;
  }
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    float2 w[R];
#pragma unroll
    for (int t = 0; t < R; ++t) w[t] = buf[((q * Q + i) * R + t) * SB + s];
    Dft<R>::run(w);
#pragma unroll
    for (int k2 = 0; k2 < R; ++k2) o[i * R + k2] = w[k2];
  }
}

// K1 for N1 = 16 R: 256 threads, G = 8 / R values of n2 x SLAB_CP pairs per block.
template <int N1>
__global__ __launch_bounds__(256) void k_slab1_f(Ctx c, const float *__restrict__ src, int guard) {
  constexpr int R = N1 / 16, Q = 16 / R, G = 8 / R, SB = SLAB_CP * G;
  extern __shared__ float2 lds[];
  SlabBlk b;
  if (!slab_block(c, c.geo.slab1, guard, b)) return;
  const SlabPlan &L = *b.L;
  const int N = L.N, N2 = L.N2, W = b.d->W;
  const long long P = b.d->P;
  float2 *tws = lds, *buf = lds + TW_LDS;
  load_tws_len(tws, c.tab + L.tw_off, N);
  const int s = threadIdx.x & (SB - 1), t = threadIdx.x / SB;
  const int cp = s & (SLAB_CP - 1);
  const int n2 = b.grp * G + (s >> 5);
  const int ja = 2 * (b.slab * SLAB_CP + cp);
  float2 v[16];
  if (n2 < N2 && ja < W) {
    const float *p = src + b.d->off + (long long)(n2 * N1 + t) * P + ja;  // slab_trow(N2 n1 + n2), n1 = t + R j
    const bool two = ja + 1 < W;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const float *q = p + (long long)(R * j) * P;
      v[j] = two ? *(const float2 *)q : make_float2(q[0], 0.f);
    }
  } else {
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  reg_stage1<N1>(v, t, tws, N / N1, buf, SB, s);
  __syncthreads();
  float2 o[16];
  reg_stage2<N1>(o, v, t, buf, SB, s);
  if (n2 >= N2) return;
  float2 *S = b.S + n2 * b.ss + cp;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
#pragma unroll
    for (int k2 = 0; k2 < R; ++k2) {
      const int k1 = t * Q + i + 16 * k2;
      float2 y = o[i * R + k2];
      if (k1 && n2) y = cmul(y, twiddle(tws, k1 * n2));
      S[(long long)k1 * N2 * b.ss] = y;
    }
  }
}

// K3 for N1 = 16 R.
template <int N1>
__global__ __launch_bounds__(256) void k_slab3_f(Ctx c, float *__restrict__ dst, int guard) {
  constexpr int R = N1 / 16, Q = 16 / R, G = 8 / R, SB = SLAB_CP * G;
  extern __shared__ float2 lds[];
  SlabBlk b;
  if (!slab_block(c, c.geo.slab1, guard, b)) return;
  const SlabPlan &L = *b.L;
  const int N = L.N, N2 = L.N2, W = b.d->W;
  const long long P = b.d->P;
  float2 *tws = lds, *buf = lds + TW_LDS;
  load_tws_len(tws, c.tab + L.tw_off, N);
  const int s = threadIdx.x & (SB - 1), t = threadIdx.x / SB;
  const int cp = s & (SLAB_CP - 1);
  const int m2 = b.grp * G + (s >> 5);
  const int ja = 2 * (b.slab * SLAB_CP + cp);
  float2 v[16];
  if (m2 < N2) {
    const float2 *S = b.S + (long long)(t * N2 + m2) * b.ss + cp;  // k1 = t + R j
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = S[(long long)(R * j) * N2 * b.ss];
  } else {
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  reg_stage1<N1>(v, t, tws, N / N1, buf, SB, s);
  __syncthreads();
  float2 o[16];
  reg_stage2<N1>(o, v, t, buf, SB, s);
  if (m2 >= N2 || ja >= W) return;
  const float invN = L.invN;
  float *p = dst + b.d->off + (long long)(m2 * N1) * P + ja;  // slab_trow(m2 + N2 m1) = m2 N1 + m1
  const bool two = ja + 1 < W;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
#pragma unroll
    for (int k2 = 0; k2 < R; ++k2) {
      const int m1 = t * Q + i + 16 * k2;
      const float2 y = o[i * R + k2];
      float *q = p + (long long)m1 * P;
      if (two) *(float2 *)q = make_float2(y.x * invN, -y.y * invN);
      else q[0] = y.x * invN;
    }
  }
}

// K2 for N2 = 16 R: 64 R threads, slots = set (k1a, k1b) x SLAB_CP pairs.
// Per-block LDS tables (built once, before the first barrier): rt[j] = exp(-2 pi i j / N2) for the
// twiddles inside both length-N2 FFTs, ft[set][m] = exp(-2 pi i k1 m / N) for the last one, and at
// k = k1 + N1 kk of both sets the Makhoul post-twiddle exp(-i pi k / 2N) and 2 s_k, the doubled
// column half of -lambda.  Each thread forms the scaled DCT-II coefficients of both lines at its
// own k only; the partner's (at N - k) come through LDS, so every coefficient and reciprocal is
// computed once.
template <int N2>
__global__ __launch_bounds__(256) void k_slab2_f(Ctx c, int guard) {
  constexpr int R = N2 / 16, Q = 16 / R, SB = 2 * SLAB_CP;
  extern __shared__ float2 lds[];
  SlabBlk b;
  if (!slab_block(c, c.geo.slab2, guard, b)) return;
  const SlabPlan &L = *b.L;
  const int N = L.N, N1 = L.N1, W = b.d->W;
  float2 *buf = lds + TW_LDS;  // (the first TW_LDS slots: the twiddle table of K1 / K3, unused here)
  float2 *mkT = buf + N2 * SB, *rt = mkT + 2 * N2, *ft = rt + N2;
  float *sT = (float *)(ft + 2 * N2);
  const int s = threadIdx.x & (SB - 1), t = threadIdx.x / SB;
  const int cp = s & (SLAB_CP - 1), sig = s >> 5;
  const int k1a = b.grp, k1b = b.grp ? N1 - b.grp : 0;
  const bool two = k1a != k1b, live = sig == 0 || two;
  const int k1 = sig ? k1b : k1a;
  // element n2 = t + R j of this slot's sequence: S row k1 N2 + n2 (rows ss float2 apart)
  const float2 *Sl = b.S + ((long long)k1 * N2 + t) * b.ss + cp;
  const unsigned rs = (unsigned)(R * b.ss);
  float2 v[16];
  if (live) {
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = Sl[j * rs];
  } else {
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = make_float2(0.f, 0.f);
  }
  // s_k = 4 sin^2(pi k / 2H) = 2 - 2 cos(pi k / H) from the float64 cosine, rounded once; the
  // float32 sum -lambda_kj = s_k + s_j of two positive terms keeps the float32-rounded float64
  // eigenvalue to about one unit in the last place.  Both are stored doubled (exactly), which
  // folds the DCT-III pre-step's factor 1/2 into the reciprocal.
  const int tstep = N / N2;
  {
    // exp(-2 pi i m / N) from sincospif: N is a power of two here, so 2 m / N = 2 m invN is exact
    // in float32 (no division; no table read before the first barrier)
    const float invN = L.invN;
    auto expi = [&](int m) {
      float sn, cs;
      sincospif((float)(2 * m) * invN, &sn, &cs);
      return make_float2(cs, -sn);
    };
    const float2 *mk = c.tab + L.mk_off;
    const double *cosH = c.dtab + b.d->cosH_off;
    // 4 N2 threads: the first 2 N2 build the per-set tables, the next N2 the twiddles rt, so that
    // no thread evaluates two sincospif
    const int e = threadIdx.x;
    if (e < 2 * N2) {
      const int kk = e & (N2 - 1), k1e = e < N2 ? k1a : k1b;
      const int k = k1e + N1 * kk;
      mkT[e] = mk[k];
      sT[e] = (float)(4.0 - 4.0 * cosH[k]);
      ft[e] = expi(k1e * kk);
    } else if (e < 3 * N2) {
      rt[e - 2 * N2] = expi((e - 2 * N2) * tstep);
    }
  }
  __syncthreads();
  // first FFT, stage 1: DFT-16 over j, times exp(-2 pi i t k' / N2), exchange
  Dft<16>::run(v);
  if (R > 1) {
    if (t) {
#pragma unroll
      for (int k = 1; k < 16; ++k) v[k] = cmul(v[k], rt[t * k]);
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) buf[(k * R + t) * SB + s] = v[k];
  }
  __syncthreads();
  float2 o[16];
  reg_stage2<N2>(o, v, t, buf, SB, s);
  // spectrum Z at k = k1 + N1 kk, kk = t Q + i + 16 k2, in natural order for the partner reads
  __syncthreads();
#pragma unroll
  for (int i = 0; i < Q; ++i)
#pragma unroll
    for (int k2 = 0; k2 < R; ++k2) buf[(t * Q + i + 16 * k2) * SB + s] = o[i * R + k2];
  __syncthreads();
  // partner of kk: the other set at N2 - 1 - kk; in the single set of k1 = 0 at -kk mod N2, of
  // k1 = N1 / 2 at N2 - 1 - kk
  const int ja = 2 * (b.slab * SLAB_CP + cp);
  const bool hasa = live && ja < W, hasb = hasa && ja + 1 < W;
  // a column pair past the image's edge (no line a) computes finite values that no one reads
  const float swa = hasa ? (float)(4.0 - 4.0 * c.dtab[b.d->cosW_off + ja]) : 1.f;
  const float swb = hasb ? (float)(4.0 - 4.0 * c.dtab[b.d->cosW_off + ja + 1]) : 0.f;
  const int sp = two ? (s ^ SLAB_CP) : s;
  const int ts = sig ? N2 : 0;
  const bool refl0 = !two && k1a == 0;
  auto partner = [&](int kk) { return refl0 ? (kk ? N2 - kk : 0) : N2 - 1 - kk; };
  // own DCT-II coefficients of both lines at k, divided by -2 lambda
#pragma unroll
  for (int i = 0; i < Q; ++i) {
#pragma unroll
    for (int k2 = 0; k2 < R; ++k2) {
      const int kk = t * Q + i + 16 * k2;
      const float2 Zk = o[i * R + k2], Zn = buf[partner(kk) * SB + sp];
      const float2 wk = mkT[ts + kk];
      const float sk = sT[ts + kk];
      const float xa = wk.x * (Zk.x + Zn.x) - wk.y * (Zk.y - Zn.y);
      const float xb = wk.x * (Zk.y + Zn.y) + wk.y * (Zk.x - Zn.x);
      // line b past the edge must be zero: it shares the complex sequence with line a
      o[i * R + k2] = make_float2(xa * __builtin_amdgcn_rcpf(sk + swa), hasb ? xb * __builtin_amdgcn_rcpf(sk + swb) : 0.f);
    }
  }
  // the image's DC coefficient (k = 0 of column 0, eigenvalue 0): element kk = 0 of thread t = 0
  if (ja == 0 && k1 == 0 && t == 0) o[0].x = 0.f;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < Q; ++i)
#pragma unroll
    for (int k2 = 0; k2 < R; ++k2) buf[(t * Q + i + 16 * k2) * SB + s] = o[i * R + k2];
  __syncthreads();
  // DCT-III pre-step at k from the scaled coefficients at k and N - k (none at k = 0)
#pragma unroll
  for (int i = 0; i < Q; ++i) {
#pragma unroll
    for (int k2 = 0; k2 < R; ++k2) {
      const int kk = t * Q + i + 16 * k2;
      const float2 y = o[i * R + k2];
      float2 yn = buf[partner(kk) * SB + sp];
      if (k1 == 0 && kk == 0) yn = make_float2(0.f, 0.f);
      const float2 wk = mkT[ts + kk];
      const float cx = wk.x, cy = -wk.y;
      const float vax = cx * y.x + cy * yn.x, vay = cy * y.x - cx * yn.x;
      const float vbx = cx * y.y + cy * yn.y, vby = cy * y.y - cx * yn.y;
      o[i * R + k2] = make_float2(vax - vby, -(vay + vbx));
    }
  }
  // second FFT over kk = a + 16 b (a = t Q + i): DFT-R over b, times exp(-2 pi i a m' / N2),
  // exchange, DFT-16 over a: X[m' + R m''] with m' = t
  __syncthreads();
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    float2 w[R];
#pragma unroll
    for (int bb = 0; bb < R; ++bb) w[bb] = o[i * R + bb];
    Dft<R>::run(w);
    const int a = t * Q + i;
#pragma unroll
    for (int m = 0; m < R; ++m) buf[(a * R + m) * SB + s] = m ? cmul(w[m], rt[a * m]) : w[m];
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 16; ++a) v[a] = buf[(a * R + t) * SB + s];
  Dft<16>::run(v);
  if (!live) return;
  float2 *So = const_cast<float2 *>(Sl);
  // ft = exp(0) = (1, -0) where k1 = 0 or m = 0: the product keeps the value (up to the sign of a
  // zero), so no select
#pragma unroll
  for (int mm = 0; mm < 16; ++mm) So[mm * rs] = cmul(v[mm], ft[ts + t + R * mm]);
}

}  // namespace irls2d
