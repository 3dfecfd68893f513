// LDS-resident mixed-radix FFT and the paired-line DCT-II / DCT-III built on it (gfx950).
//
// One workgroup transforms two real lines at once: they are packed as the real and imaginary
// parts of one complex sequence (the even/odd Makhoul reordering, J. Makhoul, "A fast cosine
// transform in one and two dimensions", IEEE TASSP 28(1), 1980), transformed by one complex FFT
// of length N, and separated with the Hermitian symmetry of real spectra.  The complex FFT is a
// Stockham auto-sort FFT (radices 8, 4, 2, 3, 5) working in place in LDS: every pass reads all of
// its butterfly inputs into registers, synchronises, then writes.  Lengths whose prime factors
// exceed 5 use Bluestein's chirp-z algorithm with a 5-smooth length M >= 2N - 1.
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

template <int R> struct Dft;

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

template <> struct Dft<8> {
  static __device__ __forceinline__ void run(float2 *v) {
    const float h = 0.70710678118654752440f;
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4>::run(e);
    Dft<4>::run(o);
    // o_k *= exp(-2 pi i k / 8)
    o[1] = make_float2((o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h);
    o[2] = mul_negi(o[2]);
    o[3] = make_float2((o[3].y - o[3].x) * h, -(o[3].x + o[3].y) * h);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = cadd(e[k], o[k]);
      v[k + 4] = csub(e[k], o[k]);
    }
  }
};

// One Stockham pass of radix R over a length-L sequence in LDS, Ns = product of earlier radices.
// Butterfly j reads x[j + t L/R] (t < R), twiddles by exp(-2 pi i t k / (Ns R)) with k = j mod Ns,
// and writes y[(j - k) R + k + t Ns].
template <int R>
__device__ __forceinline__ void stockham_pass(float2 *buf, int L, int Ns, const float2 *__restrict__ tw) {
  constexpr int MB = (EPT + R - 1) / R;
  const int T = blockDim.x, tid = threadIdx.x;
  const int nb = L / R;
  const int twmul = L / (Ns * R);
  float2 v[MB][R];
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int j = tid + q * T;
    if (j < nb) {
      const int k = j % Ns;
#pragma unroll
      for (int t = 0; t < R; ++t) v[q][t] = buf[j + t * nb];
      if (Ns > 1) {
#pragma unroll
        for (int t = 1; t < R; ++t) v[q][t] = cmul(v[q][t], tw[(t * k) * twmul]);
      }
      Dft<R>::run(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int j = tid + q * T;
    if (j < nb) {
      const int k = j % Ns;
      const int base = (j - k) * R + k;
#pragma unroll
      for (int t = 0; t < R; ++t) buf[base + t * Ns] = v[q][t];
    }
  }
  __syncthreads();
}

// Forward complex FFT of length P->L in LDS (in place).  Requires L <= EPT * blockDim.x.
__device__ __forceinline__ void fft_lds(float2 *buf, const FftPlan *P, const float2 *__restrict__ tab) {
  const int L = P->L, np = P->npass;
  const float2 *tw = tab + P->tw_off;
  int Ns = 1;
  for (int s = 0; s < np; ++s) {
    const int R = P->radix[s];
    switch (R) {
      case 8: stockham_pass<8>(buf, L, Ns, tw); break;
      case 4: stockham_pass<4>(buf, L, Ns, tw); break;
      case 2: stockham_pass<2>(buf, L, Ns, tw); break;
      case 3: stockham_pass<3>(buf, L, Ns, tw); break;
      default: stockham_pass<5>(buf, L, Ns, tw); break;
    }
    Ns *= R;
  }
}

// Length-N forward DFT of the sequence in buf[0, N).  For Bluestein the caller has already
// multiplied element n by chirp[n] and zeroed buf[N, M); the spectrum is then read via dct_Z.
__device__ __forceinline__ void dct_fft(float2 *buf, const DctPlan *P, const float2 *__restrict__ tab) {
  fft_lds(buf, &P->fft, tab);
  if (P->blue) {
    const float2 *bh = tab + P->bhat_off;
    const int M = P->M;
    for (int k = threadIdx.x; k < M; k += blockDim.x) buf[k] = cconj(cmul(buf[k], bh[k]));
    __syncthreads();
    fft_lds(buf, &P->fft, tab);
  }
}

__device__ __forceinline__ float2 dct_Z(const float2 *buf, const DctPlan *P, const float2 *__restrict__ chirp, int k) {
  float2 v = buf[k];
  return P->blue ? cmul(chirp[k], cconj(v)) : v;
}

__device__ __forceinline__ int makhoul_pos(int m, int N) { return (m & 1) ? (N - 1 - (m >> 1)) : (m >> 1); }

// Element m of the two lines into its Makhoul position in LDS (chirp-weighted for Bluestein).
__device__ __forceinline__ void dct2_pair_put(float2 *buf, const DctPlan *P, const float2 *__restrict__ tab,
                                              int m, float a, float b) {
  const int n = makhoul_pos(m, P->N);
  float2 z = make_float2(a, b);
  if (P->blue) z = cmul(z, tab[P->chirp_off + n]);
  buf[n] = z;
}

// After every element went through dct2_pair_put: zero the Bluestein padding and transform.
__device__ __forceinline__ void dct2_pair_run(float2 *buf, const DctPlan *P, const float2 *__restrict__ tab) {
  if (P->blue)
    for (int n = P->N + threadIdx.x; n < P->M; n += blockDim.x) buf[n] = make_float2(0.f, 0.f);
  __syncthreads();
  dct_fft(buf, P, tab);
}

// DCT-II of two real lines a, b (b may be null = zeros) of length N into LDS spectrum form.
__device__ __forceinline__ void dct2_pair_load(float2 *buf, const DctPlan *P, const float2 *__restrict__ tab,
                                               const float *a, const float *b) {
  const int N = P->N, T = blockDim.x;
  for (int m = threadIdx.x; m < N; m += T) dct2_pair_put(buf, P, tab, m, a[m], b ? b[m] : 0.f);
  dct2_pair_run(buf, P, tab);
}

// After dct2_pair_load: DCT-II coefficient k of both lines.
__device__ __forceinline__ void dct2_pair_coef(const float2 *buf, const DctPlan *P, const float2 *__restrict__ tab,
                                               int k, float &Xa, float &Xb) {
  const int N = P->N;
  const float2 *chirp = tab + P->chirp_off;
  const float2 Zk = dct_Z(buf, P, chirp, k);
  const float2 Zn = dct_Z(buf, P, chirp, k ? N - k : 0);
  const float2 w = tab[P->mk_off + k];
  Xa = w.x * (Zk.x + Zn.x) - w.y * (Zk.y - Zn.y);
  Xb = w.x * (Zk.y + Zn.y) + w.y * (Zk.x - Zn.x);
}

// Inverse of the DCT-II for two lines whose coefficients sit in buf[k] = (Xa_k, Xb_k), k < N.
// After the call, dct3_pair_value(m) gives sample m of both lines.
__device__ __forceinline__ void dct3_pair(float2 *buf, const DctPlan *P, const float2 *__restrict__ tab) {
  const int N = P->N, M = P->M, T = blockDim.x, tid = threadIdx.x;
  const float2 *chirp = tab + P->chirp_off;
  const float2 *mk = tab + P->mk_off;
  const bool blue = P->blue;
  float2 reg[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int k = tid + q * T;
    if (k < N) {
      const float2 Xk = buf[k];
      const float2 Xn = k ? buf[N - k] : make_float2(0.f, 0.f);
      const float ya = 0.5f * Xk.x, yb = 0.5f * Xk.y, yan = 0.5f * Xn.x, ybn = 0.5f * Xn.y;
      const float2 w = mk[k];  // (cos t, -sin t), t = pi k / 2N
      const float cx = w.x, cy = -w.y;
      const float vax = cx * ya + cy * yan, vay = cy * ya - cx * yan;
      const float vbx = cx * yb + cy * ybn, vby = cy * yb - cx * ybn;
      float2 cz = make_float2(vax - vby, -(vay + vbx));  // conj(Va + i Vb)
      if (blue) cz = cmul(cz, chirp[k]);
      reg[q] = cz;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int k = tid + q * T;
    if (k < N) buf[k] = reg[q];
  }
  if (blue)
    for (int n = N + tid; n < M; n += T) buf[n] = make_float2(0.f, 0.f);
  __syncthreads();
  dct_fft(buf, P, tab);
}

__device__ __forceinline__ void dct3_pair_value(const float2 *buf, const DctPlan *P, const float2 *__restrict__ tab,
                                                int m, float &a, float &b) {
  const int n = makhoul_pos(m, P->N);
  const float2 w = dct_Z(buf, P, tab + P->chirp_off, n);
  a = w.x * P->invN;
  b = -w.y * P->invN;
}

}  // namespace irls2d
