// Solver kernels.  Every kernel covers all images of a batch in one launch: blockIdx.x is mapped
// to (image, local block) through the per-geometry block-start tables in Ctx::geo.  Kernels that
// end a reduction finish it in the last-arriving block of each image ("finalizer"), which is the
// only writer of that image's ImgState fields; later kernels read the state to decide whether to
// run, so the whole fixed-count iteration can be enqueued (or captured in a hipGraph) without any
// host synchronisation while keeping the reference's early exits.
#pragma once

#include "fft_lds.hpp"

namespace irls2d {

enum Guard { GUARD_NONE = 0, GUARD_CG = 1, GUARD_IRLS = 2 };
enum DotMode { DOT_NONE = 0, DOT_RZ_INIT = 1, DOT_RZ = 2 };

__device__ __forceinline__ bool guard_off(const ImgState *st, int guard) {
  return (guard == GUARD_CG && !st->cg_active) || (guard == GUARD_IRLS && !st->irls_active);
}

struct TilePos {
  int img, i0, j0, local, nblk;
};

__device__ __forceinline__ TilePos tile_pos(const Ctx &c) {
  TilePos t;
  t.img = find_img(c.geo.tile, c.n, blockIdx.x);
  const int s = c.geo.tile[t.img];
  t.nblk = c.geo.tile[t.img + 1] - s;
  t.local = blockIdx.x - s;
  const int W = c.img[t.img].W;
  const int ntx = (W + TILE - 1) / TILE;
  t.i0 = (t.local / ntx) * TILE;
  t.j0 = (t.local % ntx) * TILE;
  return t;
}

// Calls f(i, j, idx) for every pixel of the tile owned by this thread.
template <class F>
__device__ __forceinline__ void for_tile(const ImgDesc &d, const TilePos &t, F &&f) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int j = t.j0 + lane;
  if (j >= d.W) return;
  const int iend = min(t.i0 + TILE, d.H);
  for (int i = t.i0 + wv; i < iend; i += TILE_THREADS / 64) {
    const long long idx = d.off + (long long)i * d.W + j;
    f(i, j, idx);
  }
}

// (A x)_ij with A = D^T W D (forward differences, Neumann borders), accumulated in the order of
// utils_unwrap2d.py:1378-1387: 0 - dx_{j-1} + dx_j - dy_{i-1} + dy_i, then negated.
template <class LX>
__device__ __forceinline__ float apply_A(const float *__restrict__ wxi, const float *__restrict__ wyi,
                                         long long idx, int i, int j, int H, int W, float xc, LX &&X) {
  float s = 0.f;
  if (j > 0) s = -((xc - X(idx - 1)) * wxi[idx - 1]);
  if (j < W - 1) s += (X(idx + 1) - xc) * wxi[idx];
  if (i > 0) s -= (xc - X(idx - W)) * wyi[idx - W];
  if (i < H - 1) s += (X(idx + W) - xc) * wyi[idx];
  return -s;
}

// ---------------------------------------------------------------------------------------------
// Set-up
// ---------------------------------------------------------------------------------------------

__global__ void k_state_init(Ctx c) {
  for (int i = threadIdx.x; i < c.n; i += blockDim.x) {
    ImgState s;
    s.rz = 0; s.alpha = 0; s.beta = 0;
    s.best_res = __builtin_huge_val();
    s.mean = 0; s.vcount = 0;
    s.irls_active = 1; s.cg_active = 0;
    s.n_cg = 0; s.n_cg_total = 0; s.n_irls = 0;
    s.improved_it = -2; s.restore_it = -2; s.status = 0;
    s.counter = 0u; s.pad = 0;
    c.st[i] = s;
  }
}

// utils_unwrap2d.py:1226-1241 (fill, clip) and :1255-1275 (targets, edge weights).
__device__ __forceinline__ void load_pw(const float *__restrict__ phase, const float *__restrict__ weight,
                                        long long idx, float &phi, float &w, bool &valid) {
  const float ph = phase[idx];
  valid = !isnan(ph);
  phi = valid ? ph : 0.f;
  if (weight) {
    float ww = weight[idx];
    if (isnan(ww) || !valid) ww = 0.f;
    w = fminf(fmaxf(ww, 0.01f), 1.0f);
  } else {
    w = valid ? 1.f : 0.f;
  }
}

__device__ __forceinline__ float wrap_atan2(float d) { return atan2f(sinf(d), cosf(d)); }

__global__ __launch_bounds__(TILE_THREADS) void k_setup(Ctx c, const float *__restrict__ phase,
                                                         const float *__restrict__ weight) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  double acc[1] = {0.0};
  for_tile(d, t, [&](int i, int j, long long idx) {
    float phi, w; bool v;
    load_pw(phase, weight, idx, phi, w, v);
    float tx = 0.f, ty = 0.f, wx = 0.f, wy = 0.f;
    if (j < d.W - 1) {
      float phr, wr; bool vr;
      load_pw(phase, weight, idx + 1, phr, wr, vr);
      tx = wrap_atan2(phr - phi);
      wx = ((w + wr) / 2.f) * ((v && vr) ? 1.f : 0.f);
    }
    if (i < d.H - 1) {
      float phd, wd; bool vd;
      load_pw(phase, weight, idx + d.W, phd, wd, vd);
      ty = wrap_atan2(phd - phi);
      wy = ((w + wd) / 2.f) * ((v && vd) ? 1.f : 0.f);
    }
    c.f.tx[idx] = tx; c.f.ty[idx] = ty; c.f.wx[idx] = wx; c.f.wy[idx] = wy;
    c.f.valid[idx] = v ? 1 : 0;
    acc[0] += v ? 1.0 : 0.0;
  });
  double tot[1];
  if (reduce_last<1>(acc, c.partials, &c.st[t.img].counter, c.geo.tile[t.img], t.nblk, tot)) {
    if (threadIdx.x == 0) c.st[t.img].vcount = tot[0];
  }
}

// rho = D^T W t with the static edge weights (utils_unwrap2d.py:1476-1483).
__global__ __launch_bounds__(TILE_THREADS) void k_rhs_static(Ctx c, float *__restrict__ out) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  const float *tx = c.f.tx, *ty = c.f.ty, *wx = c.f.wx, *wy = c.f.wy;
  for_tile(d, t, [&](int i, int j, long long idx) {
    float s = 0.f;
    if (j > 0) s += wx[idx - 1] * tx[idx - 1];
    if (j < d.W - 1) s -= wx[idx] * tx[idx];
    if (i > 0) s += wy[idx - d.W] * ty[idx - d.W];
    if (i < d.H - 1) s -= wy[idx] * ty[idx];
    out[idx] = s;
  });
}

// Masked mean of nan_to_num(u) after the DCT start (utils_unwrap2d.py:1498-1513).
__global__ __launch_bounds__(TILE_THREADS) void k_center_reduce(Ctx c) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  double acc[1] = {0.0};
  for_tile(d, t, [&](int i, int j, long long idx) {
    float u = c.f.u[idx];
    if (!isfinite(u)) u = 0.f;
    if (c.f.valid[idx]) acc[0] += (double)u;
  });
  double tot[1];
  if (reduce_last<1>(acc, c.partials, &c.st[t.img].counter, c.geo.tile[t.img], t.nblk, tot)) {
    if (threadIdx.x == 0) {
      ImgState *st = &c.st[t.img];
      st->mean = st->vcount > 0 ? tot[0] / st->vcount : 0.0;
    }
  }
}

__global__ __launch_bounds__(TILE_THREADS) void k_center_apply(Ctx c) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  const ImgState *st = &c.st[t.img];
  const float m = (float)st->mean;
  const bool has = st->vcount > 0;
  for_tile(d, t, [&](int i, int j, long long idx) {
    float u = c.f.u[idx];
    if (!isfinite(u)) u = 0.f;
    if (has) u = u - m;
    c.f.u[idx] = u;
    c.f.u_best[idx] = u;
  });
}

// ---------------------------------------------------------------------------------------------
// IRLS step (utils_unwrap2d.py:1533-1581)
// ---------------------------------------------------------------------------------------------

// u_prev = u (before centring) and the masked sum of u.
__global__ __launch_bounds__(TILE_THREADS) void k_irls_mean(Ctx c) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  ImgState *st = &c.st[t.img];
  if (!st->irls_active) return;
  double acc[1] = {0.0};
  for_tile(d, t, [&](int i, int j, long long idx) {
    const float u = c.f.u[idx];
    c.f.u_prev[idx] = u;
    if (c.f.valid[idx]) acc[0] += (double)u;
  });
  double tot[1];
  if (reduce_last<1>(acc, c.partials, &st->counter, c.geo.tile[t.img], t.nblk, tot)) {
    if (threadIdx.x == 0) st->mean = st->vcount > 0 ? tot[0] / st->vcount : 0.0;
  }
}

// Centre u, residuals r = D u - t, residual norm, IRLS weights and right-hand side b.
__global__ __launch_bounds__(TILE_THREADS) void k_irls_update(Ctx c, int it) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  ImgState *st = &c.st[t.img];
  if (!st->irls_active) return;
  const float m = st->vcount > 0 ? (float)st->mean : 0.f;
  const float eps_sq = c.eps_sq;
  const float *u = c.f.u, *tx = c.f.tx, *ty = c.f.ty, *wx = c.f.wx, *wy = c.f.wy;
  double acc[2] = {0.0, 0.0};  // residual, non-finite count
  // Neighbouring tiles read this tile's u, so the centred u goes to z (free between PCG solves)
  // and k_irls_commit moves it into u.
  for_tile(d, t, [&](int i, int j, long long idx) {
    const float uc = u[idx] - m;
    float rx = 0.f, ry = 0.f;
    float wxi_r = 0.f, wyi_d = 0.f;
    if (j < d.W - 1) rx = (u[idx + 1] - m) - uc;
    rx = rx - tx[idx];
    if (i < d.H - 1) ry = (u[idx + d.W] - m) - uc;
    ry = ry - ty[idx];
    wxi_r = clampf_nan(wx[idx] / sqrtf(rx * rx + eps_sq), 1e-6f, 1e6f);
    wyi_d = clampf_nan(wy[idx] / sqrtf(ry * ry + eps_sq), 1e-6f, 1e6f);
    float bb = 0.f;
    if (j > 0) {
      const float rxl = (uc - (u[idx - 1] - m)) - tx[idx - 1];
      const float wl = clampf_nan(wx[idx - 1] / sqrtf(rxl * rxl + eps_sq), 1e-6f, 1e6f);
      bb += wl * tx[idx - 1];
    }
    if (j < d.W - 1) bb -= wxi_r * tx[idx];
    if (i > 0) {
      const float ryu = (uc - (u[idx - d.W] - m)) - ty[idx - d.W];
      const float wu = clampf_nan(wy[idx - d.W] / sqrtf(ryu * ryu + eps_sq), 1e-6f, 1e6f);
      bb += wu * ty[idx - d.W];
    }
    if (i < d.H - 1) bb -= wyi_d * ty[idx];
    c.f.wxi[idx] = wxi_r;
    c.f.wyi[idx] = wyi_d;
    c.f.b[idx] = bb;
    acc[0] += (double)rx * rx + (double)ry * ry;
    acc[1] += isfinite(uc) ? 0.0 : 1.0;
    c.f.z[idx] = uc;
  });
  double tot[2];
  if (reduce_last<2>(acc, c.partials, &st->counter, c.geo.tile[t.img], t.nblk, tot)) {
    if (threadIdx.x == 0) {
      const double res = tot[0];
      if (res < st->best_res && tot[1] == 0.0) {
        st->best_res = res;
        st->improved_it = it;
      }
    }
  }
}

// u = centred u (from z); u_best = u when this iteration improved the residual.
__global__ __launch_bounds__(TILE_THREADS) void k_irls_commit(Ctx c, int it) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  const ImgState *st = &c.st[t.img];
  if (!st->irls_active) return;
  const bool imp = st->improved_it == it;
  for_tile(d, t, [&](int i, int j, long long idx) {
    const float v = c.f.z[idx];
    c.f.u[idx] = v;
    if (imp) c.f.u_best[idx] = v;
  });
}

// PCG start: r = b - A x (x = u); a non-finite r returns x unchanged (utils_unwrap2d.py:1414-1419).
__global__ __launch_bounds__(TILE_THREADS) void k_cg_init(Ctx c) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  ImgState *st = &c.st[t.img];
  if (!st->irls_active) return;
  const float *u = c.f.u;
  double acc[1] = {0.0};
  for_tile(d, t, [&](int i, int j, long long idx) {
    const float xc = u[idx];
    const float ax = apply_A(c.f.wxi, c.f.wyi, idx, i, j, d.H, d.W, xc, [&](long long k) { return u[k]; });
    const float r = c.f.b[idx] - ax;
    c.f.r[idx] = r;
    acc[0] += isfinite(r) ? 0.0 : 1.0;
  });
  double tot[1];
  if (reduce_last<1>(acc, c.partials, &st->counter, c.geo.tile[t.img], t.nblk, tot)) {
    if (threadIdx.x == 0) {
      st->cg_active = tot[0] == 0.0 ? 1 : 0;
      st->n_cg = 0;
    }
  }
}

// p = z (first iteration) or z + beta p; Ap = A p; p.Ap  (utils_unwrap2d.py:1429-1440, :1473)
__global__ __launch_bounds__(TILE_THREADS) void k_cg_pap(Ctx c, const float *__restrict__ p_old, float *__restrict__ p_new) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  ImgState *st = &c.st[t.img];
  if (!st->cg_active) return;
  const bool first = st->n_cg == 0;
  const float beta = (float)st->beta;
  const float *z = c.f.z;
  auto P = [&](long long k) -> float {
    return first ? z[k] : __fadd_rn(__fmul_rn(p_old[k], beta), z[k]);
  };
  double acc[1] = {0.0};
  for_tile(d, t, [&](int i, int j, long long idx) {
    const float pc = P(idx);
    const float ap = apply_A(c.f.wxi, c.f.wyi, idx, i, j, d.H, d.W, pc, P);
    p_new[idx] = pc;
    c.f.Ap[idx] = ap;
    acc[0] += (double)pc * (double)ap;
  });
  double tot[1];
  if (reduce_last<1>(acc, c.partials, &st->counter, c.geo.tile[t.img], t.nblk, tot)) {
    if (threadIdx.x == 0) {
      const double pAp = tot[0];
      st->n_cg += 1;
      st->n_cg_total += 1;
      if (!(fabs(pAp) >= 1e-15) || !isfinite(pAp)) {
        st->cg_active = 0;
      } else {
        st->alpha = clampd_nan(st->rz / pAp, -1e6, 1e6);
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Preconditioner z = IDCT2(-DCT2(r) / lambda), DC zeroed (utils_unwrap2d.py:1390-1403), as
// row DCT-II -> transpose -> column DCT-II * (-1/lambda) * column DCT-III -> transpose ->
// row DCT-III.  The first row pass can also apply the PCG update x += a p, r -= a Ap
// (utils_unwrap2d.py:1444-1458) to the two rows it transforms.
// ---------------------------------------------------------------------------------------------

template <bool FUSE_UPDATE>
__global__ __launch_bounds__(MAX_FFT_THREADS) void k_row_dct2(Ctx c, float *in_, float *__restrict__ out,
                                                               const float *__restrict__ p, int guard) {
  extern __shared__ float2 lds[];
  const int img = find_img(c.geo.rowpair, c.n, blockIdx.x);
  const ImgDesc d = c.img[img];
  ImgState *st = &c.st[img];
  if (guard_off(st, guard)) return;
  const int q = blockIdx.x - c.geo.rowpair[img];
  const int nblk = c.geo.rowpair[img + 1] - c.geo.rowpair[img];
  const int ra = 2 * q;
  const bool hasb = ra + 1 < d.H;
  const int W = d.W;
  const long long oa = d.off + (long long)ra * W;
  double acc[2] = {0.0, 0.0};  // r.r, non-finite x
  const DctPlan *P = c.plans + d.rowplan;
  if (FUSE_UPDATE) {
    // x += a p, r -= a Ap on both rows; the new r goes straight into the LDS transform buffer
    float *x = c.f.u;
    const float *Ap = c.f.Ap;
    const float alpha = (float)st->alpha;
    for (int m = threadIdx.x; m < W; m += blockDim.x) {
      float rv[2] = {0.f, 0.f};
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (h == 1 && !hasb) break;
        const long long k = oa + (long long)h * W + m;
        const float xn = x[k] + alpha * p[k];
        const float rn = in_[k] - alpha * Ap[k];
        x[k] = xn;
        in_[k] = rn;
        rv[h] = rn;
        acc[0] += (double)rn * rn;
        acc[1] += isfinite(xn) ? 0.0 : 1.0;
      }
      dct2_pair_put(lds, P, c.tab, m, rv[0], rv[1]);
    }
    dct2_pair_run(lds, P, c.tab);
  } else {
    dct2_pair_load(lds, P, c.tab, in_ + oa, hasb ? in_ + oa + W : nullptr);
  }
  for (int k = threadIdx.x; k < W; k += blockDim.x) {
    float Xa, Xb;
    dct2_pair_coef(lds, P, c.tab, k, Xa, Xb);
    out[oa + k] = Xa;
    if (hasb) out[oa + W + k] = Xb;
  }
  if (FUSE_UPDATE) {
    double tot[2];
    if (reduce_last<2>(acc, c.partials, &st->counter, c.geo.rowpair[img], nblk, tot)) {
      if (threadIdx.x == 0) {
        if (tot[1] != 0.0) {
          st->cg_active = 0;
        } else {
          const double rn = sqrt(tot[0]);
          if (rn < c.cg_tol || !isfinite(rn)) st->cg_active = 0;
        }
      }
    }
  }
}

// T (H x W) -> Tt (W x H) when dir == 0, Tt -> T when dir == 1.
__global__ __launch_bounds__(256) void k_transpose(Ctx c, const float *__restrict__ src, float *__restrict__ dst,
                                                   int dir, int guard) {
  __shared__ float tile[TT][TT + 1];
  const int img = find_img(c.geo.trans, c.n, blockIdx.x);
  const ImgDesc d = c.img[img];
  if (guard_off(&c.st[img], guard)) return;
  const int local = blockIdx.x - c.geo.trans[img];
  const int ntx = (d.W + TT - 1) / TT;
  const int ti = (local / ntx) * TT, tj = (local % ntx) * TT;
  const int x = threadIdx.x & 63, y0 = threadIdx.x >> 6;
  const int H = d.H, W = d.W;
  const float *s = src + d.off;
  float *o = dst + d.off;
  if (dir == 0) {
    for (int y = y0; y < TT; y += 4) {
      const int i = ti + y, j = tj + x;
      if (i < H && j < W) tile[y][x] = s[(long long)i * W + j];
    }
    __syncthreads();
    for (int y = y0; y < TT; y += 4) {
      const int j = tj + y, i = ti + x;
      if (i < H && j < W) o[(long long)j * H + i] = tile[x][y];
    }
  } else {
    for (int y = y0; y < TT; y += 4) {
      const int j = tj + y, i = ti + x;
      if (i < H && j < W) tile[y][x] = s[(long long)j * H + i];
    }
    __syncthreads();
    for (int y = y0; y < TT; y += 4) {
      const int i = ti + y, j = tj + x;
      if (i < H && j < W) o[(long long)i * W + j] = tile[x][y];
    }
  }
}

// Column pass on the transposed field: DCT-II along H, divide by -lambda (float32 lambda from
// float64 cosines, utils_unwrap2d.py:1284-1289, :1395-1398), zero the DC term, inverse DCT.
__global__ __launch_bounds__(MAX_FFT_THREADS) void k_col(Ctx c, float *__restrict__ Tt, int guard) {
  extern __shared__ float2 lds[];
  const int img = find_img(c.geo.colpair, c.n, blockIdx.x);
  const ImgDesc d = c.img[img];
  if (guard_off(&c.st[img], guard)) return;
  const int q = blockIdx.x - c.geo.colpair[img];
  const int ja = 2 * q;
  const bool hasb = ja + 1 < d.W;
  const int H = d.H;
  float *la = Tt + d.off + (long long)ja * H;
  const DctPlan *P = c.plans + d.colplan;
  dct2_pair_load(lds, P, c.tab, la, hasb ? la + H : nullptr);
  const double *cosH = c.dtab + d.cosH_off;
  const double cwa = 2.0 * c.dtab[d.cosW_off + ja];
  const double cwb = hasb ? 2.0 * c.dtab[d.cosW_off + ja + 1] : 0.0;
  float2 reg[EPT];
  const int T = blockDim.x;
#pragma unroll
  for (int qq = 0; qq < EPT; ++qq) {
    const int k = threadIdx.x + qq * T;
    if (k < H) {
      float Xa, Xb;
      dct2_pair_coef(lds, P, c.tab, k, Xa, Xb);
      const double ch = 2.0 * cosH[k];
      float la_ = (float)(ch + cwa - 4.0);
      float lb_ = (float)(ch + cwb - 4.0);
      if (k == 0 && ja == 0) la_ = 1.0f;
      float ya = -(Xa / la_);
      float yb = hasb ? -(Xb / lb_) : 0.f;
      if (k == 0 && ja == 0) ya = 0.f;
      reg[qq] = make_float2(ya, yb);
    }
  }
  __syncthreads();
#pragma unroll
  for (int qq = 0; qq < EPT; ++qq) {
    const int k = threadIdx.x + qq * T;
    if (k < H) lds[k] = reg[qq];
  }
  __syncthreads();
  dct3_pair(lds, P, c.tab);
  for (int m = threadIdx.x; m < H; m += T) {
    float a, b;
    dct3_pair_value(lds, P, c.tab, m, a, b);
    la[m] = a;
    if (hasb) la[H + m] = b;
  }
}

// Row DCT-III: T -> out.  DOT_RZ_INIT: rz = r.z;  DOT_RZ: the PCG beta step
// (utils_unwrap2d.py:1463-1474): beta = clamp(r.z / rz), rz = r.z, stop when |rz| < 1e-15.
__global__ __launch_bounds__(MAX_FFT_THREADS) void k_row_dct3(Ctx c, const float *__restrict__ T_,
                                                               float *__restrict__ out, int dot, int guard) {
  extern __shared__ float2 lds[];
  const int img = find_img(c.geo.rowpair, c.n, blockIdx.x);
  const ImgDesc d = c.img[img];
  ImgState *st = &c.st[img];
  if (guard_off(st, guard)) return;
  const int q = blockIdx.x - c.geo.rowpair[img];
  const int nblk = c.geo.rowpair[img + 1] - c.geo.rowpair[img];
  const int ra = 2 * q;
  const bool hasb = ra + 1 < d.H;
  const int W = d.W, T = blockDim.x;
  const long long oa = d.off + (long long)ra * W;
  const DctPlan *P = c.plans + d.rowplan;
  for (int k = threadIdx.x; k < W; k += T)
    lds[k] = make_float2(T_[oa + k], hasb ? T_[oa + W + k] : 0.f);
  __syncthreads();
  dct3_pair(lds, P, c.tab);
  double acc[1] = {0.0};
  const float *r = c.f.r;
  for (int m = threadIdx.x; m < W; m += T) {
    float a, b;
    dct3_pair_value(lds, P, c.tab, m, a, b);
    out[oa + m] = a;
    if (dot) acc[0] += (double)r[oa + m] * (double)a;
    if (hasb) {
      out[oa + W + m] = b;
      if (dot) acc[0] += (double)r[oa + W + m] * (double)b;
    }
  }
  if (dot) {
    double tot[1];
    if (reduce_last<1>(acc, c.partials, &st->counter, c.geo.rowpair[img], nblk, tot)) {
      if (threadIdx.x == 0) {
        const double rz_new = tot[0];
        if (dot == DOT_RZ_INIT) {
          st->rz = rz_new;
        } else if (fabs(st->rz) < 1e-15) {
          st->cg_active = 0;
        } else {
          st->beta = clampd_nan(rz_new / st->rz, -1e6, 1e6);
          st->rz = rz_new;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// End of an IRLS step (utils_unwrap2d.py:1599-1626)
// ---------------------------------------------------------------------------------------------

__global__ __launch_bounds__(TILE_THREADS) void k_irls_check(Ctx c, int it) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  ImgState *st = &c.st[t.img];
  if (!st->irls_active) return;
  double acc[3] = {0.0, 0.0, 0.0};  // non-finite u, |u_prev - u|^2, |u|^2
  for_tile(d, t, [&](int i, int j, long long idx) {
    const float u = c.f.u[idx];
    const float df = c.f.u_prev[idx] - u;
    acc[0] += isfinite(u) ? 0.0 : 1.0;
    acc[1] += (double)df * df;
    acc[2] += (double)u * u;
  });
  double tot[3];
  if (reduce_last<3>(acc, c.partials, &st->counter, c.geo.tile[t.img], t.nblk, tot)) {
    if (threadIdx.x == 0) {
      st->cg_active = 0;
      st->n_irls = it + 1;
      if (tot[0] != 0.0) {
        st->restore_it = it;
        st->irls_active = 0;
        st->status |= 1;
      } else {
        const double rel = sqrt(tot[1]) / (sqrt(tot[2]) + 1e-10);
        if (rel < c.tol) {
          st->irls_active = 0;
          st->status |= 2;
        }
      }
    }
  }
}

// u = u_best after a non-finite IRLS step.
__global__ __launch_bounds__(TILE_THREADS) void k_restore(Ctx c, int it) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  if (c.st[t.img].restore_it != it) return;
  for_tile(d, t, [&](int i, int j, long long idx) { c.f.u[idx] = c.f.u_best[idx]; });
}

__global__ void k_info(Ctx c, int32_t *info) {
  for (int i = threadIdx.x; i < c.n; i += blockDim.x) {
    const ImgState &s = c.st[i];
    info[i * 4 + 0] = s.status;
    info[i * 4 + 1] = s.n_irls;
    info[i * 4 + 2] = s.n_cg_total;
    info[i * 4 + 3] = (int32_t)s.vcount;
  }
}

}  // namespace irls2d
