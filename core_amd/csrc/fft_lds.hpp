// LDS-resident mixed-radix FFT and the paired-line DCT-II / DCT-III built on it (gfx950).
//
// One workgroup transforms two real lines at once: they are packed as the real and imaginary
// parts of one complex sequence (the even/odd Makhoul reordering, J. Makhoul, "A fast cosine
// transform in one and two dimensions", IEEE TASSP 28(1), 1980), transformed by one complex FFT
// of length N, and separated with the Hermitian symmetry of real spectra.  The complex FFT is a
// Stockham auto-sort FFT (radices 16, 8, 4, 2, 3, 5) working in place in LDS: every pass reads all
// of its butterfly inputs into registers, synchronises, then writes.  Lengths whose prime factors
// exceed 5 use Bluestein's chirp-z algorithm with a 5-smooth length M >= 2N - 1.
//
// LDS layout: element i of the sequence lives at pidx(i) = i + i/16 (one pad slot every 16
// elements), which spreads the stride-R writes of the early Stockham passes over all banks.
// Twiddles exp(-2 pi i m / L) come from a split table in LDS: lo[m mod 64] * hi[m div 64].
//
// DCT conventions (the torch-dct package's norm=None pair the reference calls at
// insardev/insardev/utils_unwrap2d.py:1334): DCT-II X_k = 2 sum_n x_n cos(pi k (2n+1) / 2N) and
// its exact inverse.
#pragma once

#include "common.hpp"

namespace irls2d {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 mul_negi(float2 a) { return make_float2(a.y, -a.x); }  // -i * a

__device__ __forceinline__ int pidx(int i) { return i + (i >> 4); }

template <int R> struct Dft;

template <> struct Dft<1> {
  static __device__ __forceinline__ void run(float2 *) {}
};

template <> struct Dft<2> {
  static __device__ __forceinline__ void run(float2 *v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <> struct Dft<3> {
  static __device__ __forceinline__ void run(float2 *v) {
    const float s = 0.86602540378443864676f;  // sin(2 pi / 3)
    float2 s12 = cadd(v[1], v[2]), d12 = csub(v[1], v[2]);
    float2 y0 = cadd(v[0], s12);
    float2 t = make_float2(v[0].x - 0.5f * s12.x, v[0].y - 0.5f * s12.y);
    float2 u = mul_negi(cscale(d12, s));
    v[0] = y0;
    v[1] = cadd(t, u);
    v[2] = csub(t, u);
  }
};

template <> struct Dft<4> {
  static __device__ __forceinline__ void run(float2 *v) {
    float2 a02 = cadd(v[0], v[2]), s02 = csub(v[0], v[2]);
    float2 a13 = cadd(v[1], v[3]), s13 = mul_negi(csub(v[1], v[3]));
    v[0] = cadd(a02, a13);
    v[2] = csub(a02, a13);
    v[1] = cadd(s02, s13);
    v[3] = csub(s02, s13);
  }
};

template <> struct Dft<5> {
  static __device__ __forceinline__ void run(float2 *v) {
    const float c1 = 0.30901699437494742410f;   // cos(2 pi / 5)
    const float c2 = -0.80901699437494742410f;  // cos(4 pi / 5)
    const float s1 = 0.95105651629515357212f;   // sin(2 pi / 5)
    const float s2 = 0.58778525229247312917f;   // sin(4 pi / 5)
    float2 b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
    float2 d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
    float2 y0 = cadd(v[0], cadd(b1, b2));
    float2 t1 = make_float2(v[0].x + c1 * b1.x + c2 * b2.x, v[0].y + c1 * b1.y + c2 * b2.y);
    float2 t2 = make_float2(v[0].x + c2 * b1.x + c1 * b2.x, v[0].y + c2 * b1.y + c1 * b2.y);
    float2 u1 = mul_negi(make_float2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y));
    float2 u2 = mul_negi(make_float2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y));
    v[0] = y0;
    v[1] = cadd(t1, u1);
    v[4] = csub(t1, u1);
    v[2] = cadd(t2, u2);
    v[3] = csub(t2, u2);
  }
};

// multiply by exp(-2 pi i k / 8) for k = 1, 3
__device__ __forceinline__ float2 w8_1(float2 a) {
  const float h = 0.70710678118654752440f;
  return make_float2((a.x + a.y) * h, (a.y - a.x) * h);
}
__device__ __forceinline__ float2 w8_3(float2 a) {
  const float h = 0.70710678118654752440f;
  return make_float2((a.y - a.x) * h, -(a.x + a.y) * h);
}

template <> struct Dft<8> {
  static __device__ __forceinline__ void run(float2 *v) {
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4>::run(e);
    Dft<4>::run(o);
    o[1] = w8_1(o[1]);
    o[2] = mul_negi(o[2]);
    o[3] = w8_3(o[3]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = cadd(e[k], o[k]);
      v[k + 4] = csub(e[k], o[k]);
    }
  }
};

// 16-point DFT as 4 x 4: inner DFT4 over m of x[q + 4m], twiddle exp(-2 pi i q k1 / 16), outer DFT4
// over q; X[k1 + 4 k2].
template <> struct Dft<16> {
  static __device__ __forceinline__ void run(float2 *v) {
    const float c1 = 0.92387953251128675613f;  // cos(pi/8)
    const float s1 = 0.38268343236508977173f;  // sin(pi/8)
    float2 b[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      float2 a[4] = {v[q], v[q + 4], v[q + 8], v[q + 12]};
      Dft<4>::run(a);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) b[q][k1] = a[k1];
    }
    // twiddles exp(-2 pi i q k1 / 16), q k1 in {1, 2, 3, 4, 6, 9}
    b[1][1] = cmul(b[1][1], make_float2(c1, -s1));
    b[1][2] = w8_1(b[1][2]);
    b[1][3] = cmul(b[1][3], make_float2(s1, -c1));
    b[2][1] = w8_1(b[2][1]);
    b[2][2] = mul_negi(b[2][2]);
    b[2][3] = w8_3(b[2][3]);
    b[3][1] = cmul(b[3][1], make_float2(s1, -c1));
    b[3][2] = w8_3(b[3][2]);
    b[3][3] = cmul(b[3][3], make_float2(-c1, s1));
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      float2 a[4] = {b[0][k1], b[1][k1], b[2][k1], b[3][k1]};
      Dft<4>::run(a);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = a[k2];
    }
  }
};

// exp(-2 pi i m / 64) for a compile-time m (first-quadrant cosines, folded by symmetry).
__device__ __forceinline__ float2 w64(int m) {
  constexpr float C[17] = {1.00000000000000000000f, 0.99518472667219692873f, 0.98078528040323043058f,
                           0.95694033573220882438f, 0.92387953251128673848f, 0.88192126434835504956f,
                           0.83146961230254523567f, 0.77301045336273699338f, 0.70710678118654757274f,
                           0.63439328416364548779f, 0.55557023301960228867f, 0.47139673682599780857f,
                           0.38268343236508983729f, 0.29028467725446233105f, 0.19509032201612833135f,
                           0.09801714032956077016f, 0.0f};
  m &= 63;
  const float c = m <= 16 ? C[m] : m <= 32 ? -C[32 - m] : m <= 48 ? -C[m - 32] : C[64 - m];
  const float s = m <= 16 ? C[16 - m] : m <= 32 ? C[m - 16] : m <= 48 ? -C[48 - m] : -C[m - 48];
  return make_float2(c, -s);
}

// 32-point DFT: two 16-point DFTs of the even and odd samples, radix-2 combine.
template <> struct Dft<32> {
  static __device__ __forceinline__ void run(float2 *v) {
    float2 e[16], o[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      e[i] = v[2 * i];
      o[i] = v[2 * i + 1];
    }
    Dft<16>::run(e);
    Dft<16>::run(o);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const float2 t = k == 0 ? o[0] : (k == 8 ? mul_negi(o[8]) : cmul(o[k], w64(2 * k)));
      v[k] = cadd(e[k], t);
      v[k + 16] = csub(e[k], t);
    }
  }
};

// 64-point DFT: four 16-point DFTs of x[4m + q], twiddles exp(-2 pi i q k1 / 64), radix-4 combine:
// X[k1 + 16 k2].
template <> struct Dft<64> {
  static __device__ __forceinline__ void run(float2 *v) {
    float2 e[4][16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int m = 0; m < 16; ++m) e[q][m] = v[4 * m + q];
      Dft<16>::run(e[q]);
    }
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) {
      float2 a[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) a[q] = (q == 0 || k1 == 0) ? e[q][k1] : cmul(e[q][k1], w64(q * k1));
      Dft<4>::run(a);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 16 * k2] = a[k2];
    }
  }
};

// Twiddle exp(-2 pi i m / L) from the split LDS table (lo: 64 entries, hi: ceil(L/64) entries).
__device__ __forceinline__ float2 twiddle(const float2 *tws, int m) {
  return cmul(tws[TW_LO + (m >> 6)], tws[m & 63]);
}

__device__ __forceinline__ float2 csqr(float2 a) {
  return make_float2(a.x * a.x - a.y * a.y, 2.f * a.x * a.y);
}

// Butterfly j of a Stockham pass works on k = j mod Ns (float quotient when Ns is no power of two).
__device__ __forceinline__ int stockham_k(int j, const PassDesc &pd) {
  if (pd.nsmask >= 0) return j & pd.nsmask;
  const int qt = (int)((float)j * pd.inv_ns);
  int k = j - __umul24(qt, pd.Ns);
  if (k < 0) k += pd.Ns;
  if (k >= pd.Ns) k -= pd.Ns;
  return k;
}

// Multiplies butterfly input t by exp(-2 pi i t m1 / L) from the split table.
template <int R>
__device__ __forceinline__ void twiddle_inputs(float2 *v, const float2 *tws, int m1) {
  float2 w[R];
  w[1] = twiddle(tws, m1);
#pragma unroll
  for (int t = 2; t < R; ++t) {
    // restart from the table every 8 powers, otherwise a short product tree
    w[t] = (t % 8 == 0) ? twiddle(tws, __umul24(t, m1)) : ((t % 2 == 0) ? csqr(w[t / 2]) : cmul(w[t - 1], w[1]));
  }
#pragma unroll
  for (int t = 1; t < R; ++t) v[t] = cmul(v[t], w[t]);
}

// Multiplies butterfly input t by exp(-2 pi i t k / (Ns R)) (no-op in the first pass).
template <int R>
__device__ __forceinline__ void stockham_twiddle(float2 *v, const float2 *tws, const PassDesc &pd, int k) {
  if (pd.Ns == 1) return;
  twiddle_inputs<R>(v, tws, __umul24(k, pd.twmul));
}

// One Stockham pass of radix R over a length-L sequence in LDS, Ns = product of earlier radices.
// Butterfly j reads x[j + t L/R] (t < R), twiddles by exp(-2 pi i t k / (Ns R)) with k = j mod Ns,
// and writes y[(j - k) R + k + t Ns].  Index arithmetic uses 24-bit multiplies and, when the
// host marked the strides as aligned to the 16-element padding, running LDS offsets.
template <int R, int E>
__device__ __forceinline__ void stockham_pass(float2 *buf, const float2 *tws, const PassDesc &pd) {
  constexpr int MB = (E + R - 1) / R;
  const int T = blockDim.x, tid = threadIdx.x;
  const int nb = pd.nb, Ns = pd.Ns;
  float2 v[MB][R];
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int j = tid + q * T;
    if (j < nb) {
      if (pd.rstride) {
        int a = pidx(j);
#pragma unroll
        for (int t = 0; t < R; ++t) {
          v[q][t] = buf[a];
          a += pd.rstride;
        }
      } else {
#pragma unroll
        for (int t = 0; t < R; ++t) v[q][t] = buf[pidx(j + __umul24(t, nb))];
      }
      if (Ns > 1) stockham_twiddle<R>(v[q], tws, pd, stockham_k(j, pd));
      Dft<R>::run(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int j = tid + q * T;
    if (j < nb) {
      const int k = (Ns == 1) ? 0 : stockham_k(j, pd);
      const int base = __umul24(j - k, R) + k;
      if (pd.wstride) {
        int a = pidx(base);
#pragma unroll
        for (int t = 0; t < R; ++t) {
          buf[a] = v[q][t];
          a += pd.wstride;
        }
      } else {
#pragma unroll
        for (int t = 0; t < R; ++t) buf[pidx(base + __umul24(t, Ns))] = v[q][t];
      }
    }
  }
  __syncthreads();
}

// Compile-time form of stockham_pass for the default power-of-two schedules (radix-16 passes,
// then one radix-8/4/2 pass): every stride, mask and twiddle step is a constant, so LDS accesses
// are one base address plus immediate offsets.  Requires NB % 16 == 0 and NS == 1 or NS % 16 == 0.
template <int R, int NS, int L, int E>
__device__ __forceinline__ void stockham_pass_ct(float2 *buf, const float2 *tws) {
  constexpr int NB = L / R;
  constexpr int MB = (E + R - 1) / R;
  constexpr int RS = NB + NB / 16;                 // padded distance of x[j] and x[j + NB]
  constexpr int WS = NS == 1 ? 1 : NS + NS / 16;   // padded distance of y[b] and y[b + NS]
  constexpr int TWMUL = L / (NS * R);
  static_assert(NB % 16 == 0 && (NS == 1 ? 16 % R == 0 : NS % 16 == 0), "no constant strides");
  const int T = blockDim.x, tid = threadIdx.x;
  float2 v[MB][R];
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int j = tid + q * T;
    if (j < NB) {
      const float2 *p = buf + pidx(j);
#pragma unroll
      for (int t = 0; t < R; ++t) v[q][t] = p[t * RS];
      if (NS > 1) twiddle_inputs<R>(v[q], tws, (j & (NS - 1)) * TWMUL);
      Dft<R>::run(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int j = tid + q * T;
    if (j < NB) {
      const int k = NS == 1 ? 0 : (j & (NS - 1));
      float2 *p = buf + pidx((j - k) * R + k);
#pragma unroll
      for (int t = 0; t < R; ++t) p[t * WS] = v[q][t];
    }
  }
  __syncthreads();
}

// Default schedule of L = 2^a * f (f = 1, 3 or 5): radix-16 passes, one radix-8/4/2 pass for the
// rest of 2^a, then one radix-f pass (the order radices() produces), all with constant strides.
template <int L, int E>
__device__ __forceinline__ void fft_lds_ct(float2 *buf, const float2 *tws) {
  constexpr int A = __builtin_ctz(L);
  constexpr int F = L >> A;
  constexpr int N16 = A / 4, REM = A % 4;
  static_assert((F == 1 || F == 3 || F == 5) && N16 >= 2 && N16 <= 3, "lengths 2^a * {1, 3, 5}, 1024 .. 12288");
  constexpr int NS_REM = N16 == 3 ? 4096 : 256;
  stockham_pass_ct<16, 1, L, E>(buf, tws);
  stockham_pass_ct<16, 16, L, E>(buf, tws);
  if constexpr (N16 >= 3) stockham_pass_ct<16, 256, L, E>(buf, tws);
  if constexpr (REM) stockham_pass_ct<(1 << REM), NS_REM, L, E>(buf, tws);
  if constexpr (F > 1) stockham_pass_ct<F, (1 << A), L, E>(buf, tws);
}

// Forward complex FFT of length P->L in LDS (in place).  Requires L <= E * blockDim.x.
template <int E>
__device__ __forceinline__ void fft_lds(float2 *buf, const float2 *tws, const FftPlan *P) {
  switch (P->ct_len) {
    case 1024: fft_lds_ct<1024, E>(buf, tws); return;
    case 1280: fft_lds_ct<1280, E>(buf, tws); return;
    case 1536: fft_lds_ct<1536, E>(buf, tws); return;
    case 2048: fft_lds_ct<2048, E>(buf, tws); return;
    case 2560: fft_lds_ct<2560, E>(buf, tws); return;
    case 3072: fft_lds_ct<3072, E>(buf, tws); return;
    case 4096: fft_lds_ct<4096, E>(buf, tws); return;
    case 5120: fft_lds_ct<5120, E>(buf, tws); return;
    case 6144: fft_lds_ct<6144, E>(buf, tws); return;
    case 8192: fft_lds_ct<8192, E>(buf, tws); return;
    default: break;
  }
  if constexpr (E >= 24 || E == 12) {  // 12 x 1024 and 24 x 512 threads reach 12288
    switch (P->ct_len) {
      case 10240: fft_lds_ct<10240, E>(buf, tws); return;
      case 12288: fft_lds_ct<12288, E>(buf, tws); return;
      default: break;
    }
  }
  const int np = P->npass;
  for (int s = 0; s < np; ++s) {
    const PassDesc pd = P->pass[s];
    switch (pd.R) {
      case 16: stockham_pass<16, E>(buf, tws, pd); break;
      case 8: stockham_pass<8, E>(buf, tws, pd); break;
      case 4: stockham_pass<4, E>(buf, tws, pd); break;
      case 2: stockham_pass<2, E>(buf, tws, pd); break;
      case 3: stockham_pass<3, E>(buf, tws, pd); break;
      default: stockham_pass<5, E>(buf, tws, pd); break;
    }
  }
}

// Workgroup context of one DCT length: LDS buffer, LDS twiddles, plan, global tables.
struct DctCtx {
  float2 *buf;
  float2 *tws;
  const DctPlan *P;
  const float2 *tab;
};

// Points into the dynamic LDS block (twiddles first, then the padded sequence) and loads the
// twiddles.  The first __syncthreads of the caller's transform orders the loads before use.
__device__ __forceinline__ DctCtx dct_begin(float2 *lds, const DctPlan *P, const float2 *__restrict__ tab) {
  DctCtx d;
  d.tws = lds;
  d.buf = lds + TW_LDS;
  d.P = P;
  d.tab = tab;
  const int n = TW_LO + ((P->fft.L + 63) >> 6);
  for (int i = threadIdx.x; i < n; i += blockDim.x) d.tws[i] = tab[P->fft.tw_off + i];
  return d;
}

// Length-N forward DFT of the sequence in buf[0, N).  For Bluestein the caller has already
// multiplied element n by chirp[n] and zeroed buf[N, M); the spectrum is then read via dct_Z.
template <int E>
__device__ __forceinline__ void dct_fft(const DctCtx &d) {
  fft_lds<E>(d.buf, d.tws, &d.P->fft);
  if (d.P->blue) {
    // pointwise product with the transformed chirp; a thread loads BU of its values before it uses
    // any (one memory latency per batch; index clamped so that every load is unconditional).  The
    // 12-element kernels take all of theirs in one batch (M <= 12 T).
    constexpr int BU = E == 12 ? 12 : BLUE_U;
    const float2 *__restrict__ bh = d.tab + d.P->bhat_off;
    const unsigned M = d.P->M, T = blockDim.x;
    for (unsigned k0 = threadIdx.x; k0 < M; k0 += BU * T) {
      float2 b[BU];
#pragma unroll
      for (int u = 0; u < BU; ++u) b[u] = bh[min(k0 + u * T, M - 1)];
#pragma unroll
      for (int u = 0; u < BU; ++u) {
        const unsigned k = k0 + u * T;
        if (k < M) {
          const int pk = pidx(k);
          d.buf[pk] = cconj(cmul(d.buf[pk], b[u]));
        }
      }
    }
    __syncthreads();
    fft_lds<E>(d.buf, d.tws, &d.P->fft);
  }
}

__device__ __forceinline__ float2 dct_Z(const DctCtx &d, int k) {
  float2 v = d.buf[pidx(k)];
  return d.P->blue ? cmul(d.tab[d.P->chirp_off + k], cconj(v)) : v;
}

__device__ __forceinline__ int makhoul_pos(int m, int N) { return (m & 1) ? (N - 1 - (m >> 1)) : (m >> 1); }

// Element m of the two lines into its Makhoul position in LDS (chirp-weighted for Bluestein).
__device__ __forceinline__ void dct2_pair_put(const DctCtx &d, int m, float a, float b) {
  const int n = makhoul_pos(m, d.P->N);
  float2 z = make_float2(a, b);
  if (d.P->blue) z = cmul(z, d.tab[d.P->chirp_off + n]);
  d.buf[pidx(n)] = z;
}

// After every element went through dct2_pair_put: zero the Bluestein padding and transform.
template <int E>
__device__ __forceinline__ void dct2_pair_run(const DctCtx &d) {
  if (d.P->blue)
    for (int n = d.P->N + threadIdx.x; n < d.P->M; n += blockDim.x) d.buf[pidx(n)] = make_float2(0.f, 0.f);
  __syncthreads();
  dct_fft<E>(d);
}

// DCT-II of two real lines a, b (b may be null = zeros) of length N into LDS spectrum form.
template <int E>
__device__ __forceinline__ void dct2_pair_load(const DctCtx &d, const float *a, const float *b) {
  const int N = d.P->N, T = blockDim.x;
#pragma unroll 4
  for (int m = threadIdx.x; m < N; m += T) dct2_pair_put(d, m, a[m], b ? b[m] : 0.f);
  dct2_pair_run<E>(d);
}

// DCT-II coefficients of both lines at k and nk = (N - k) mod N from the spectrum values Z_k, Z_nk
// and the post-twiddles wk = exp(-i pi k / 2N), wn = exp(-i pi nk / 2N).
__device__ __forceinline__ void dct2_coef_w(float2 Zk, float2 Zn, float2 wk, float2 wn, float &Xak, float &Xbk,
                                            float &Xan, float &Xbn) {
  Xak = wk.x * (Zk.x + Zn.x) - wk.y * (Zk.y - Zn.y);
  Xbk = wk.x * (Zk.y + Zn.y) + wk.y * (Zk.x - Zn.x);
  Xan = wn.x * (Zn.x + Zk.x) - wn.y * (Zn.y - Zk.y);
  Xbn = wn.x * (Zn.y + Zk.y) + wn.y * (Zn.x - Zk.x);
}

// After dct2_pair_run: DCT-II coefficients of both lines at k and at nk = (N - k) mod N, from the
// two spectrum values Z_k, Z_nk (Hermitian separation + Makhoul post-twiddle).
__device__ __forceinline__ void dct2_coef_pair(const DctCtx &d, int k, int nk, float &Xak, float &Xbk,
                                               float &Xan, float &Xbn) {
  const float2 Zk = dct_Z(d, k);
  const float2 Zn = dct_Z(d, nk);
  dct2_coef_w(Zk, Zn, d.tab[d.P->mk_off + k], d.tab[d.P->mk_off + nk], Xak, Xbk, Xan, Xbn);
}

// DCT-III input step at index k: from X_k and X_{N-k} (zero for k = 0) of both lines, the value
// conj(V_a + i V_b) (chirp-weighted for Bluestein) whose forward DFT gives N * conj(output).
// w = exp(-i pi k / 2N) = (cos t, -sin t), t = pi k / 2N.
__device__ __forceinline__ float2 dct3_pre_w(float2 w, float Xak, float Xbk, float Xan, float Xbn) {
  const float ya = 0.5f * Xak, yb = 0.5f * Xbk, yan = 0.5f * Xan, ybn = 0.5f * Xbn;
  const float cx = w.x, cy = -w.y;
  const float vax = cx * ya + cy * yan, vay = cy * ya - cx * yan;
  const float vbx = cx * yb + cy * ybn, vby = cy * yb - cx * ybn;
  return make_float2(vax - vby, -(vay + vbx));
}

__device__ __forceinline__ float2 dct3_pre(const DctCtx &d, int k, float Xak, float Xbk, float Xan, float Xbn) {
  float2 cz = dct3_pre_w(d.tab[d.P->mk_off + k], Xak, Xbk, Xan, Xbn);
  if (d.P->blue) cz = cmul(cz, d.tab[d.P->chirp_off + k]);
  return cz;
}

// The DCT-III input value of slot k given its post-twiddle w = exp(-i pi k / 2N), both lines'
// coefficients at k and at N - k.
__device__ __forceinline__ void dct3_put(const DctCtx &d, int k, float2 w, float Xak, float Xbk, float Xan,
                                         float Xbn) {
  float2 cz = dct3_pre_w(w, Xak, Xbk, Xan, Xbn);
  if (d.P->blue) cz = cmul(cz, d.tab[d.P->chirp_off + k]);
  d.buf[pidx(k)] = cz;
}

// Stores the DCT-III input values of the pair (k, nk) given both lines' coefficients there.
// Each pair of LDS slots is owned by one thread, so this can run in place over the spectrum.
__device__ __forceinline__ void dct3_put_pair(const DctCtx &d, int k, int nk, float Xak, float Xbk, float Xan,
                                              float Xbn) {
  if (k == 0) {
    d.buf[pidx(0)] = dct3_pre(d, 0, Xak, Xbk, 0.f, 0.f);
  } else {
    d.buf[pidx(k)] = dct3_pre(d, k, Xak, Xbk, Xan, Xbn);
    if (nk != k) d.buf[pidx(nk)] = dct3_pre(d, nk, Xan, Xbn, Xak, Xbk);
  }
}

// After every pair went through dct3_put_pair: zero the Bluestein padding and transform;
// dct3_pair_value(m) then gives sample m of both lines.
template <int E>
__device__ __forceinline__ void dct3_run(const DctCtx &d) {
  if (d.P->blue)
    for (int n = d.P->N + threadIdx.x; n < d.P->M; n += blockDim.x) d.buf[pidx(n)] = make_float2(0.f, 0.f);
  __syncthreads();
  dct_fft<E>(d);
}

__device__ __forceinline__ void dct3_pair_value(const DctCtx &d, int m, float &a, float &b) {
  const int n = makhoul_pos(m, d.P->N);
  const float2 w = dct_Z(d, n);
  a = w.x * d.P->invN;
  b = -w.y * d.P->invN;
}

}  // namespace irls2d
