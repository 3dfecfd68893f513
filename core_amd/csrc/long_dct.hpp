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

// Forward FFTs of all B sequences.  Requires B * L <= E * blockDim.x.
template <int E>
__device__ __forceinline__ void batch_fft(float2 *buf, int BS, int lb, const float2 *tws, const FftPlan *P) {
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

struct LongLine {
  const LongPlan *L;
  int img, lp, blk;  // image, line pair, block within the line pair
  long long la;      // element offset of line a (line b = la + pitch)
  long long pitch;
  bool hasb;
  float2 *S;         // scratch line of this pair
};

__device__ __forceinline__ bool long_line(const Ctx &c, int dir, int pass, int guard, LongLine &ln) {
  const int *geo = dir ? (pass == 1 ? c.geo.colL1 : c.geo.colL2) : (pass == 1 ? c.geo.rowL1 : c.geo.rowL2);
  ln.img = find_img(geo, c.n, blockIdx.x);
  const ImgDesc &d = c.img[ln.img];
  if (guard_off(&c.st[ln.img], guard)) return false;
  ln.L = c.lplans + (dir ? d.collong : d.rowlong);
  const int nbp = pass == 1 ? ln.L->nb1 : ln.L->nb2;
  const int local = blockIdx.x - geo[ln.img];
  ln.lp = local / nbp;
  ln.blk = local - ln.lp * nbp;
  const int nlines = dir ? d.W : d.H;
  ln.pitch = dir ? d.PT : d.P;
  ln.la = (dir ? d.offT : d.off) + (long long)(2 * ln.lp) * ln.pitch;
  ln.hasb = 2 * ln.lp + 1 < nlines;
  ln.S = c.S + d.offS + (long long)ln.lp * ln.L->spitch;
  return true;
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
  const int o = dct2 ? N2 - 1 : 0;
  const float *a = src + ln.la;
  const float *b = src + ln.la + ln.pitch;
  const int T = blockDim.x;
  double acc[2] = {0.0, 0.0};
  for (int idx = threadIdx.x; idx < (N1 << lb); idx += T) {
    const int s = idx & (B - 1), n1 = idx >> lb;
    const int col = slot_index(s, C, ln.blk, N2, o);
    if (col < 0) continue;
    const int sp = slot_partner(s, C, col, N2, o);
    const int n = N2 * n1 + col;
    if (dct2) {
      if (2 * n >= N) continue;  // filled by the partner element N - 1 - n
      const long long m = 2 * (long long)n;
      const bool two = 2 * n + 1 < N;
      float2 va = ld2(a, m), vb = ln.hasb ? ld2(b, m) : make_float2(0.f, 0.f);
      if (MODE == LONG_DCT2_UPDATE) {
        const ImgState *st = &c.st[ln.img];
        const float alpha = (float)st->alpha;
        float *x = c.f.u;
        const float *Ap = c.f.Ap;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          if (h == 1 && !ln.hasb) break;
          const long long e = ln.la + (long long)h * ln.pitch + m;
          const float2 xv = ld2(x, e), pv = ld2(p, e), av = ld2(Ap, e);
          float2 rv = h ? vb : va;
          const float2 xn = make_float2(xv.x + alpha * pv.x, xv.y + alpha * pv.y);
          rv = make_float2(rv.x - alpha * av.x, rv.y - alpha * av.y);
          st2(x, e, xn, two);
          st2(src, e, rv, two);
          acc[0] += (double)rv.x * rv.x + (two ? (double)rv.y * rv.y : 0.0);
          acc[1] += (isfinite(xn.x) ? 0.0 : 1.0) + ((two && !isfinite(xn.y)) ? 1.0 : 0.0);
          if (h) vb = rv; else va = rv;
        }
      }
      buf[n1 * BS + s] = make_float2(va.x, vb.x);
      if (two) buf[(N1 - 1 - n1) * BS + sp] = make_float2(va.y, vb.y);
    } else {
      const int k = n;
      const int nk = k ? N - k : 0;
      if (nk < k) continue;  // done by the partner
      float Xak = a[k], Xbk = ln.hasb ? b[k] : 0.f;
      float Xan = k ? a[nk] : 0.f, Xbn = (k && ln.hasb) ? b[nk] : 0.f;
      if (MODE == LONG_DCT3_SCALE) {
        // lines are image columns ja, ja + 1 of length H = N; lambda = 2cos(pi k/H) + 2cos(pi j/W) - 4
        const ImgDesc &d = c.img[ln.img];
        const double *cosH = c.dtab + d.cosH_off;
        const int ja = 2 * ln.lp;
        const double cwa = 2.0 * c.dtab[d.cosW_off + ja];
        const double cwb = ln.hasb ? 2.0 * c.dtab[d.cosW_off + ja + 1] : 0.0;
        const double chk = 2.0 * cosH[k], chn = 2.0 * cosH[nk];
        Xak = (k == 0 && ja == 0) ? 0.f : -(Xak * __builtin_amdgcn_rcpf((float)(chk + cwa - 4.0)));
        Xbk = ln.hasb ? -(Xbk * __builtin_amdgcn_rcpf((float)(chk + cwb - 4.0))) : 0.f;
        if (k) {
          Xan = -(Xan * __builtin_amdgcn_rcpf((float)(chn + cwa - 4.0)));
          Xbn = ln.hasb ? -(Xbn * __builtin_amdgcn_rcpf((float)(chn + cwb - 4.0))) : 0.f;
        }
      }
      const float2 *mk = c.tab + L.mk_off;
      buf[n1 * BS + s] = dct3_pre_w(mk[k], Xak, Xbk, Xan, Xbn);
      if (nk != k) {
        const int n1p = (col == 0) ? N1 - n1 : N1 - 1 - n1;  // row of nk = N2 n1' + refl(col)
        buf[n1p * BS + sp] = dct3_pre_w(mk[nk], Xan, Xbn, Xak, Xbk);
      }
    }
  }
  __syncthreads();
  batch_fft<E>(buf, BS, lb, tws, F);
  const float2 *tw = c.tab + L.tw_off;
  for (int idx = threadIdx.x; idx < (N1 << lb); idx += T) {
    const int s = idx & (B - 1), k1 = idx >> lb;
    const int col = slot_index(s, C, ln.blk, N2, o);
    if (col < 0) continue;
    const float2 v = buf[k1 * BS + s];
    ln.S[k1 * N2 + col] = (k1 && col) ? cmul(v, long_twiddle(tw, k1 * col)) : v;
  }
  if (MODE == LONG_DCT2_UPDATE) block_partial<2>(acc, c.partials2);
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
  const int o = dct2 ? 0 : N1 - 1;
  const int T = blockDim.x;
  // rows of the scratch line, each a contiguous run of N2 values
  for (int idx = threadIdx.x; idx < (N2 << lb); idx += T) {
    int s = (int)((float)idx * L.inv_n2);
    int i = idx - s * N2;
    if (i < 0) { --s; i += N2; }
    if (i >= N2) { ++s; i -= N2; }
    const int k1 = slot_index(s, C, ln.blk, N1, o);
    if (k1 < 0) continue;
    buf[i * BS + s] = ln.S[k1 * N2 + i];
  }
  __syncthreads();
  batch_fft<E>(buf, BS, lb, tws, F);
  float *da = dst + ln.la, *db = dst + ln.la + ln.pitch;
  double acc[1] = {0.0};
  for (int idx = threadIdx.x; idx < (N2 << lb); idx += T) {
    const int s = idx & (B - 1), k2 = idx >> lb;
    const int k1 = slot_index(s, C, ln.blk, N1, o);
    if (k1 < 0) continue;
    const int sp = slot_partner(s, C, k1, N1, o);
    const int k = k1 + N1 * k2;
    if (dct2) {
      const int nk = k ? N - k : 0;
      if (nk < k) continue;
      // nk = k1' + N1 k2' with k1' = refl(k1) and k2' = -k2 mod N2 (k1 = 0) or N2 - 1 - k2
      const int k2p = k1 ? N2 - 1 - k2 : (k2 ? N2 - k2 : 0);
      const float2 Zk = buf[k2 * BS + s], Zn = buf[k2p * BS + sp];
      const float2 *mk = c.tab + L.mk_off;
      float Xak, Xbk, Xan, Xbn;
      dct2_coef_w(Zk, Zn, mk[k], mk[nk], Xak, Xbk, Xan, Xbn);
      da[k] = Xak;
      if (ln.hasb) db[k] = Xbk;
      if (nk != k) {
        da[nk] = Xan;
        if (ln.hasb) db[nk] = Xbn;
      }
    } else {
      const int n = k;
      if (2 * n >= N) continue;  // written with its partner N - 1 - n
      const bool two = 2 * n + 1 < N;
      const float2 w0 = buf[k2 * BS + s];
      const float2 w1 = two ? buf[(N2 - 1 - k2) * BS + sp] : make_float2(0.f, 0.f);
      const float2 va = make_float2(w0.x * L.invN, w1.x * L.invN);
      const float2 vb = make_float2(-w0.y * L.invN, -w1.y * L.invN);
      const long long m = 2 * (long long)n;
      st2(da, m, va, two);
      if (ln.hasb) st2(db, m, vb, two);
      if (dot) {
        const float *r = c.f.r + ln.la;
        const float2 ra = ld2(r, m);
        acc[0] += (double)ra.x * va.x + (two ? (double)ra.y * va.y : 0.0);
        if (ln.hasb) {
          const float2 rb = ld2(r + ln.pitch, m);
          acc[0] += (double)rb.x * vb.x + (two ? (double)rb.y * vb.y : 0.0);
        }
      }
    }
  }
  if (!dct2 && dot) block_partial<1>(acc, c.partials2);
}

}  // namespace irls2d
