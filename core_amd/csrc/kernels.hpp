// Solver kernels.  Every kernel covers all images of a batch in one launch: blockIdx.x is mapped
// to (image, local block) through the per-geometry block-start tables in Ctx::geo.  Reductions
// are two-stage and deterministic: each block writes its partial sums, then k_finalize (one block
// per image) adds them in block order and applies the reference's scalar logic (alpha, beta, the
// PCG and IRLS stopping tests).  The per-image ImgState flags written there decide whether later
// kernels run, so a fixed-count iteration can be enqueued (or captured in a hipGraph) without
// host synchronisation while keeping the reference's early exits.
#pragma once

#include "fft_lds.hpp"

namespace irls2d {

enum Guard { GUARD_NONE = 0, GUARD_CG = 1, GUARD_IRLS = 2 };
enum DotMode { DOT_NONE = 0, DOT_RZ_INIT = 1, DOT_RZ = 2 };
enum FinKind {
  FIN_VCOUNT = 0,   // k_setup          -> valid pixel count
  FIN_CENTER,       // k_center_reduce  -> mean of the DCT start
  FIN_IRLS_MEAN,    // k_irls_mean      -> masked mean
  FIN_IRLS_UPDATE,  // k_irls_update    -> residual, u_best tracking
  FIN_CG_INIT,      // k_cg_init        -> non-finite r0 test
  FIN_PAP,          // k_cg_pap         -> alpha or pAp break
  FIN_UPDATE,       // k_row_dct2<true> -> x finite / |r| tests
  FIN_RZ_INIT,      // k_row_dct3 DOT_RZ_INIT -> rz
  FIN_RZ,           // k_row_dct3 DOT_RZ      -> beta or rz break
  FIN_IRLS_CHECK,   // k_irls_check     -> restore / relative change
  FIN_GSUM,         // k_adj_gsum       -> mean of the incoming gradient over valid pixels
};

__device__ __forceinline__ bool guard_off(const ImgState *st, int guard) {
  return (guard == GUARD_CG && !st->cg_active) || (guard == GUARD_IRLS && !st->irls_active);
}

struct TilePos {
  int img, i0, j0, blk;  // blk: logical block index (partials slot)
};

__device__ __forceinline__ TilePos tile_pos(const Ctx &c) {
  TilePos t;
  t.blk = xcd_block(blockIdx.x, gridDim.x);
  t.img = find_img(c.geo.tile, c.n, t.blk);
  const int local = t.blk - c.geo.tile[t.img];
  const int W = c.img[t.img].W;
  const int ntx = (W + TILE_W - 1) / TILE_W;
  t.i0 = (local / ntx) * TILE_H;
  t.j0 = (local % ntx) * TILE_W;
  return t;
}

// Calls f(i, j, idx) for every pixel of the tile owned by this thread.
template <class F>
__device__ __forceinline__ void for_tile(const ImgDesc &d, const TilePos &t, F &&f) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int jw = t.j0 + STRIP * wv;
  if (jw >= d.W) return;
  const int iend = min(t.i0 + TILE_H, d.H);
  for (int i = t.i0; i < iend; ++i) {
    const long long row = d.off + (long long)i * d.P;
#pragma unroll
    for (int u = 0; u < STRIP / 64; ++u) {
      const int j = jw + lane + 64 * u;
      if (j < d.W) f(i, j, row + j);
    }
  }
}

// (A x)_ij with A = D^T W D (forward differences, Neumann borders), accumulated in the order of
// utils_unwrap2d.py:1378-1387: 0 - dx_{j-1} + dx_j - dy_{i-1} + dy_i, then negated.
// Same operator on four consecutive pixels of a row held in registers: x[k] for columns
// c0 + k, xl / xr the pixels left of c0 and right of c0 + 3, xu / xd the rows above and below,
// wxl the x-edge weight left of c0.  Columns >= W give 0.
__device__ __forceinline__ float4 apply_A4(int i, int c0, int H, int W, float4 x, float xl, float xr, float4 xu,
                                           float4 xd, float4 wx, float wxl, float4 wyu, float4 wy) {
  const float xs[6] = {xl, x.x, x.y, x.z, x.w, xr};
  const float wxs[5] = {wxl, wx.x, wx.y, wx.z, wx.w};
  const float us[4] = {xu.x, xu.y, xu.z, xu.w}, ds[4] = {xd.x, xd.y, xd.z, xd.w};
  const float wus[4] = {wyu.x, wyu.y, wyu.z, wyu.w}, wds[4] = {wy.x, wy.y, wy.z, wy.w};
  float a[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int j = c0 + k;
    const float xc = xs[k + 1];
    float s = 0.f;
    if (j > 0) s = -((xc - xs[k]) * wxs[k]);
    if (j < W - 1) s += (xs[k + 2] - xc) * wxs[k + 1];
    if (i > 0) s -= (xc - us[k]) * wus[k];
    if (i < H - 1) s += (ds[k] - xc) * wds[k];
    a[k] = (j < W) ? -s : 0.f;
  }
  return make_float4(a[0], a[1], a[2], a[3]);
}

#ifndef IRLS2D_TRANSFORMS_ONLY  // the transform-only translation units need none of these
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
    s.xpend = 0; s.pad1 = 0;
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
      load_pw(phase, weight, idx + d.P, phd, wd, vd);
      ty = wrap_atan2(phd - phi);
      wy = ((w + wd) / 2.f) * ((v && vd) ? 1.f : 0.f);
    }
    c.f.tx[idx] = tx; c.f.ty[idx] = ty; c.f.wx[idx] = wx; c.f.wy[idx] = wy;
    c.f.valid[idx] = v ? 1 : 0;
    acc[0] += v ? 1.0 : 0.0;
  });
  block_partial<1>(acc, c.partials, t.blk);
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
    if (i > 0) s += wy[idx - d.P] * ty[idx - d.P];
    if (i < d.H - 1) s -= wy[idx] * ty[idx];
    out[idx] = s;
  });
}

// Masked sum of nan_to_num(u) after the DCT start (utils_unwrap2d.py:1498-1513).
__global__ __launch_bounds__(TILE_THREADS) void k_center_reduce(Ctx c) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  double acc[1] = {0.0};
  for_tile(d, t, [&](int i, int j, long long idx) {
    float u = c.f.u[idx];
    if (!isfinite(u)) u = 0.f;
    if (c.f.valid[idx]) acc[0] += (double)u;
  });
  block_partial<1>(acc, c.partials, t.blk);
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
  if (!c.st[t.img].irls_active) return;
  double acc[1] = {0.0};
  for_tile(d, t, [&](int i, int j, long long idx) {
    const float u = c.f.u[idx];
    c.f.u_prev[idx] = u;
    if (c.f.valid[idx]) acc[0] += (double)u;
  });
  block_partial<1>(acc, c.partials, t.blk);
}

// Centre u, residuals r = D u - t, residual norm, IRLS weights and right-hand side b.
// Neighbouring tiles read this tile's u, so the centred u goes to z (free between PCG solves)
// and k_irls_commit moves it into u.
__global__ __launch_bounds__(TILE_THREADS) void k_irls_update(Ctx c) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  const ImgState *st = &c.st[t.img];
  if (!st->irls_active) return;
  const float m = st->vcount > 0 ? (float)st->mean : 0.f;
  const float eps_sq = c.eps_sq;
  const float *u = c.f.u, *tx = c.f.tx, *ty = c.f.ty, *wx = c.f.wx, *wy = c.f.wy;
  double acc[2] = {0.0, 0.0};  // residual, non-finite count
  for_tile(d, t, [&](int i, int j, long long idx) {
    const float uc = u[idx] - m;
    float rx = 0.f, ry = 0.f;
    if (j < d.W - 1) rx = (u[idx + 1] - m) - uc;
    rx = rx - tx[idx];
    if (i < d.H - 1) ry = (u[idx + d.P] - m) - uc;
    ry = ry - ty[idx];
    const float wxi_r = clampf_nan(wx[idx] / sqrtf(rx * rx + eps_sq), 1e-6f, 1e6f);
    const float wyi_d = clampf_nan(wy[idx] / sqrtf(ry * ry + eps_sq), 1e-6f, 1e6f);
    float bb = 0.f;
    if (j > 0) {
      const float rxl = (uc - (u[idx - 1] - m)) - tx[idx - 1];
      const float wl = clampf_nan(wx[idx - 1] / sqrtf(rxl * rxl + eps_sq), 1e-6f, 1e6f);
      bb += wl * tx[idx - 1];
    }
    if (j < d.W - 1) bb -= wxi_r * tx[idx];
    if (i > 0) {
      const float ryu = (uc - (u[idx - d.P] - m)) - ty[idx - d.P];
      const float wu = clampf_nan(wy[idx - d.P] / sqrtf(ryu * ryu + eps_sq), 1e-6f, 1e6f);
      bb += wu * ty[idx - d.P];
    }
    if (i < d.H - 1) bb -= wyi_d * ty[idx];
    c.f.wxi[idx] = wxi_r;
    c.f.wyi[idx] = wyi_d;
    c.f.b[idx] = bb;
    c.f.z[idx] = uc;
    acc[0] += (double)rx * rx + (double)ry * ry;
    acc[1] += isfinite(uc) ? 0.0 : 1.0;
  });
  block_partial<2>(acc, c.partials, t.blk);
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

// float4 access to 16-byte aligned rows (pitch and strip offsets are multiples of 4).
__device__ __forceinline__ float4 ld4(const float *__restrict__ a, long long g) {
  return *reinterpret_cast<const float4 *>(a + g);
}
__device__ __forceinline__ void st4(float *__restrict__ a, long long g, float4 v) {
  *reinterpret_cast<float4 *>(a + g) = v;
}
__device__ __forceinline__ float4 mask4(float4 v, int c0, int W) {
  return make_float4(c0 < W ? v.x : 0.f, c0 + 1 < W ? v.y : 0.f, c0 + 2 < W ? v.z : 0.f, c0 + 3 < W ? v.w : 0.f);
}

// Row-marching driver of the float4 stencils: wave w walks the strip j0 + 256 w of its tile, one
// row at a time; lane l owns columns c0 = strip + 4l .. c0 + 3.  X(i, aux) loads the operand at
// (i, c0..c0+3) (zero outside the image) and may leave four more values of that row in aux, which
// E receives with the row; XE(i, j) loads one operand value (zero outside), E(i, g, x, Ax, aux)
// consumes row i.  Lanes with c0 >= W own no pixel: they touch no memory (c0 < W keeps c0 + 3
// inside the row pitch, a multiple of 64) and only feed zeros to the shuffles.
// Three rows of the operand and the y-weights between the current row and the one behind stay in
// registers; the left/right neighbours come from the adjacent lanes, at the strip edges from loads
// by lanes 0 and 63 one row ahead.  Those edge columns are the outer columns of the neighbouring
// strips, which the neighbouring waves (of this block and of the tiles left and right, run at the
// same time on the same XCD) read in step with this one, so they are L2 hits.  For the same
// reason even tile rows march down and odd ones up: the halo rows two vertically adjacent tiles
// share are then read by both near their starts, or both near their ends.
template <bool UP, class A, class LX, class LXE, class EMIT>
__device__ __forceinline__ void march_dir(const Ctx &c, const ImgDesc &d, const TilePos &t, LX &&X, LXE &&XE,
                                          EMIT &&E) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int jw = t.j0 + STRIP * wv;
  const int c0 = jw + 4 * lane;
  const int H = d.H, W = d.W;
  const int iend = min(t.i0 + TILE_H, H);
  const float *__restrict__ wxi = c.f.wxi;
  const float *__restrict__ wyi = c.f.wyi;
  const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
  const bool own = c0 < W;
  constexpr int step = UP ? -1 : 1;
  const int ifirst = UP ? iend - 1 : t.i0;
  const int ilast = UP ? t.i0 : iend - 1;
  const bool edge_lane = lane == 0 || lane == 63;
  const int je = lane == 0 ? jw - 1 : jw + STRIP;
  // operand / x-weight left of the strip in row i, for lanes 0 (left) and 63 (right)
  auto edges = [&](int i, float &ex, float &ew) {
    ex = 0.f;
    ew = 0.f;
    if (edge_lane) {
      ex = XE(i, je);
      if (lane == 0 && jw > 0) ew = wxi[d.off + (long long)i * d.P + jw - 1];
    }
  };
  // xb: the row behind the march (above when marching down), xc: the current row; wb: the
  // y-weights between them (row i - 1 down, row i up)
  A aux_b, aux_c, aux_a;  // aux of the rows behind (unused), current, ahead
  float4 xb = X(ifirst - step, aux_b), xc = X(ifirst, aux_c);
  const int iwb = UP ? ifirst : ifirst - 1;
  float4 wb = (own && iwb >= 0) ? ld4(wyi, d.off + (long long)iwb * d.P + c0) : z4;
  float ex, ew;
  edges(ifirst, ex, ew);
  for (int i = ifirst;; i += step) {
    const long long g = d.off + (long long)i * d.P + c0;
    const int ia = i + step;
    const float4 xa = X(ia, aux_a);
    const int iwf = UP ? i - 1 : i;
    const float4 wx = own ? ld4(wxi, g) : z4;
    const float4 wf = (own && iwf >= 0) ? ld4(wyi, d.off + (long long)iwf * d.P + c0) : z4;
    float exa = 0.f, ewa = 0.f;
    if (i != ilast) edges(ia, exa, ewa);
    float xl = __shfl_up(xc.w, 1, 64);
    float wxl = __shfl_up(wx.w, 1, 64);
    float xr = __shfl_down(xc.x, 1, 64);
    if (lane == 0) {
      xl = ex;
      wxl = ew;
    }
    if (lane == 63) xr = ex;
    const float4 ax = UP ? apply_A4(i, c0, H, W, xc, xl, xr, xa, xb, wx, wxl, wf, wb)
                         : apply_A4(i, c0, H, W, xc, xl, xr, xb, xa, wx, wxl, wb, wf);
    if (own) E(i, g, xc, ax, aux_c);
    if (i == ilast) break;
    xb = xc;
    xc = xa;
    aux_c = aux_a;
    wb = wf;
    ex = exa;
    ew = ewa;
  }
}

template <class A = float4, class LX, class LXE, class EMIT>
__device__ __forceinline__ void march_rows(const Ctx &c, const ImgDesc &d, const TilePos &t, LX &&X, LXE &&XE,
                                           EMIT &&E) {
  if (t.j0 + STRIP * (int)(threadIdx.x >> 6) >= d.W) return;
  if ((t.i0 / TILE_H) & 1) march_dir<true, A>(c, d, t, X, XE, E);
  else march_dir<false, A>(c, d, t, X, XE, E);
}

// PCG start: r = b - A x (x = u); a non-finite r returns x unchanged (utils_unwrap2d.py:1414-1419).
__global__ __launch_bounds__(TILE_THREADS) void k_cg_init(Ctx c) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  double acc[1] = {0.0};
  if (c.st[t.img].irls_active) {
    const float *__restrict__ u = c.f.u;
    const float *__restrict__ b = c.f.b;
    float *__restrict__ r = c.f.r;
    const int c0 = t.j0 + STRIP * (threadIdx.x >> 6) + 4 * (threadIdx.x & 63);
    auto X = [&](int i, float4 &) -> float4 {
      if (i < 0 || i >= d.H || c0 >= d.W) return make_float4(0.f, 0.f, 0.f, 0.f);
      return mask4(ld4(u, d.off + (long long)i * d.P + c0), c0, d.W);
    };
    auto XE = [&](int i, int j) -> float {
      return (j >= 0 && j < d.W) ? u[d.off + (long long)i * d.P + j] : 0.f;
    };
    march_rows(c, d, t, X, XE, [&](int i, long long g, float4 x, float4 ax, float4) {
      const float4 bb = ld4(b, g);
      const float4 rr = make_float4(bb.x - ax.x, bb.y - ax.y, bb.z - ax.z, bb.w - ax.w);
      st4(r, g, rr);
      const float rv[4] = {rr.x, rr.y, rr.z, rr.w};
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (c0 + k < d.W) acc[0] += isfinite(rv[k]) ? 0.0 : 1.0;
    });
  } else {
    return;
  }
  block_partial<1>(acc, c.partials, t.blk);
}

// p = z (first iteration) or z + beta p; Ap = A p; p.Ap  (utils_unwrap2d.py:1429-1440, :1473).
// float4 loads along rows, vertical neighbours kept in registers while marching down the strip.
// For an xdefer image with a pending update, first the previous step's x += alpha p_old
// (utils_unwrap2d.py:1444-1449), with the count of non-finite x: this pass reads p_old anyway, so
// the update costs 8 B/px here instead of 12 in the row DCT-II pass.
__global__ __launch_bounds__(TILE_THREADS) void k_cg_pap(Ctx c, const float *__restrict__ p_old,
                                                          float *__restrict__ p_new) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  const ImgState *st = &c.st[t.img];
  if (!st->cg_active) return;
  const bool first = st->n_cg == 0;
  const float beta = (float)st->beta;
  const bool upd = d.xdefer && st->xpend;
  const float alpha = (float)st->alpha;
  float *__restrict__ x = c.f.u;
  const float *__restrict__ z = c.f.z;
  float *__restrict__ Ap = c.f.Ap;
  const int c0 = t.j0 + STRIP * (threadIdx.x >> 6) + 4 * (threadIdx.x & 63);
  auto pn = [&](float zv, float pv) { return first ? zv : __fadd_rn(__fmul_rn(pv, beta), zv); };
  // the row's p_old and, with a pending update, its x travel with the row to E (one load each)
  struct Aux {
    float4 p, x;
  };
  auto X = [&](int i, Aux &aux) -> float4 {
    const float4 z0 = make_float4(0.f, 0.f, 0.f, 0.f);
    aux.p = aux.x = z0;
    if (i < 0 || i >= d.H || c0 >= d.W) return z0;
    const long long g = d.off + (long long)i * d.P + c0;
    const float4 z4 = ld4(z, g);
    if (first) return mask4(z4, c0, d.W);
    const float4 p4 = ld4(p_old, g);
    if (upd && i >= t.i0 && i < t.i0 + TILE_H) {  // rows of the tile (not its halo rows)
      aux.p = p4;
      aux.x = ld4(x, g);
    }
    return mask4(make_float4(pn(z4.x, p4.x), pn(z4.y, p4.y), pn(z4.z, p4.z), pn(z4.w, p4.w)), c0, d.W);
  };
  auto XE = [&](int i, int j) -> float {
    if (j < 0 || j >= d.W) return 0.f;
    const long long g = d.off + (long long)i * d.P + j;
    return first ? z[g] : pn(z[g], p_old[g]);
  };
  double acc[2] = {0.0, 0.0};  // p.Ap, non-finite x
  march_rows<Aux>(c, d, t, X, XE, [&](int i, long long g, float4 pv, float4 ax, const Aux &aux) {
    st4(p_new, g, pv);
    st4(Ap, g, ax);
    acc[0] += (double)pv.x * ax.x + (double)pv.y * ax.y + (double)pv.z * ax.z + (double)pv.w * ax.w;
    if (upd) {
      const float4 po = aux.p, xo = aux.x;
      const float4 xn = make_float4(xo.x + alpha * po.x, xo.y + alpha * po.y, xo.z + alpha * po.z, xo.w + alpha * po.w);
      st4(x, g, xn);
      const float xs[4] = {xn.x, xn.y, xn.z, xn.w};
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (c0 + k < d.W) acc[1] += isfinite(xs[k]) ? 0.0 : 1.0;
    }
  });
  block_partial<2>(acc, c.partials, t.blk);
}

// The update x += alpha p still pending after the last PCG step of an xdefer image
// (utils_unwrap2d.py:1444-1449), from the p buffer that step wrote.
__global__ __launch_bounds__(TILE_THREADS) void k_cg_xcommit(Ctx c) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  const ImgState *st = &c.st[t.img];
  if (!(d.xdefer && st->xpend && st->irls_active)) return;
  const float alpha = (float)st->alpha;
  // step i (0-based) writes p1 when i is even, p0 when odd; n_cg = i + 1
  const float *__restrict__ p = ((st->n_cg - 1) & 1) ? c.f.p0 : c.f.p1;
  float *__restrict__ x = c.f.u;
  for_tile(d, t, [&](int i, int j, long long idx) { x[idx] = x[idx] + alpha * p[idx]; });
}

#endif  // IRLS2D_TRANSFORMS_ONLY
// ---------------------------------------------------------------------------------------------
// Preconditioner z = IDCT2(-DCT2(r) / lambda), DC zeroed (utils_unwrap2d.py:1390-1403), as
// row DCT-II -> transpose -> column DCT-II * (-1/lambda) * column DCT-III -> transpose ->
// row DCT-III.  The first row pass can also apply the PCG update x += a p, r -= a Ap
// (utils_unwrap2d.py:1444-1458) to the two rows it transforms.
// ---------------------------------------------------------------------------------------------

// Element offsets in T of the row-pass output of image rows ra and ra + 1: the rows themselves,
// or for an image whose column pass runs on slabs, the rows of their Makhoul positions in the
// slab order (slab_trow).
__device__ __forceinline__ void t_rows(const Ctx &c, const ImgDesc &d, int ra, long long &ta, long long &tb) {
  if (d.colslab < 0) {
    ta = d.off + (long long)ra * d.P;
    tb = ta + d.P;
    return;
  }
  const SlabPlan &L = c.splans[d.colslab];
  ta = d.off + (long long)slab_trow(makhoul_pos(ra, d.H), L) * d.P;
  tb = ra + 1 < d.H ? d.off + (long long)slab_trow(makhoul_pos(ra + 1, d.H), L) * d.P : ta;
}

// The steps of k_col for a column pair of length H, templated on Bluestein.  Each issues the global
// loads of a batch of U items per thread (data, chirps, float64 cosines, post-twiddles; indices
// clamped so that every load is unconditional) before it uses any of them; the arithmetic is that
// of dct2_pair_put / dct2_coef_pair / dct3_put_pair / dct3_pair_value.
//
// Load: element m of both lines (line b absent: zeros) into its Makhoul slot.  DIRECT reads one
// float2 (or one float without line b) per row of the row-major field at pitch `stride`; otherwise
// lines a and b are contiguous lines `stride` apart (rows of a field, or of the transposed field).
template <int U, bool BLUE, bool DIRECT>
__device__ __forceinline__ void lines_load(const DctCtx &D, const float *__restrict__ la, long long stride, bool hasb,
                                         int H) {
  const unsigned T = blockDim.x, Hu = H;
  const float2 *__restrict__ ch = D.tab + D.P->chirp_off;
  const float *__restrict__ lb = DIRECT ? la : (hasb ? la + stride : la);
  for (unsigned m0 = threadIdx.x; m0 < Hu; m0 += U * T) {
    float2 v[U], cm[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const unsigned m = min(m0 + u * T, Hu - 1);
      if (DIRECT) {
        const float *p = la + (long long)m * stride;
        v[u] = hasb ? *(const float2 *)p : make_float2(p[0], 0.f);
      } else {
        v[u] = make_float2(la[m], lb[m]);
      }
      if (BLUE) cm[u] = ch[makhoul_pos(m, H)];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const unsigned m = m0 + u * T;
      if (m < Hu) {
        float2 z = make_float2(v[u].x, hasb ? v[u].y : 0.f);
        if (BLUE) z = cmul(z, cm[u]);
        D.buf[pidx(makhoul_pos(m, H))] = z;
      }
    }
  }
}

// Middle: for k = 0 .. H/2 and nk = (H - k) mod H, the DCT-II coefficients of both lines, divided by
// -lambda (lambda rounded to float from float64 cosines; the hardware reciprocal, within 2 ulp of
// the correctly rounded division; DC of line a = 0 for the first column), and the DCT-III input
// values written back in place (each pair of slots is owned by one thread).
template <int U, bool BLUE>
__device__ __forceinline__ void col_middle(const DctCtx &D, const double *__restrict__ cosH, double cwa, double cwb,
                                           bool hasb, bool first_col, int H) {
  const unsigned T = blockDim.x, K = H / 2 + 1;
  const float2 *__restrict__ mk = D.tab + D.P->mk_off;
  const float2 *__restrict__ ch = D.tab + D.P->chirp_off;
  auto scale = [&](unsigned k, double ck, float &Xa, float &Xb) {
    const double c2 = 2.0 * ck;
    const float lam_a = (float)(c2 + cwa - 4.0);
    const float lam_b = (float)(c2 + cwb - 4.0);
    Xa = (k == 0 && first_col) ? 0.f : -(Xa * __builtin_amdgcn_rcpf(lam_a));
    Xb = hasb ? -(Xb * __builtin_amdgcn_rcpf(lam_b)) : 0.f;
  };
  for (unsigned k0 = threadIdx.x; k0 < K; k0 += U * T) {
    float2 wk[U], wn[U], ck[U], cn[U];
    double hk[U], hn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const unsigned k = min(k0 + u * T, K - 1), nk = k ? H - k : 0;
      wk[u] = mk[k];
      wn[u] = mk[nk];
      hk[u] = cosH[k];
      hn[u] = cosH[nk];
      if (BLUE) {
        ck[u] = ch[k];
        cn[u] = ch[nk];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const unsigned k = k0 + u * T;
      if (k >= K) continue;
      const unsigned nk = k ? H - k : 0;
      float2 Zk = D.buf[pidx(k)], Zn = D.buf[pidx(nk)];
      if (BLUE) {
        Zk = cmul(ck[u], cconj(Zk));
        Zn = cmul(cn[u], cconj(Zn));
      }
      float Xak, Xbk, Xan, Xbn;
      dct2_coef_w(Zk, Zn, wk[u], wn[u], Xak, Xbk, Xan, Xbn);
      scale(k, hk[u], Xak, Xbk);
      scale(nk, hn[u], Xan, Xbn);
      if (k == 0) {
        float2 z = dct3_pre_w(wk[u], Xak, Xbk, 0.f, 0.f);
        if (BLUE) z = cmul(z, ck[u]);
        D.buf[pidx(0)] = z;
      } else {
        float2 z = dct3_pre_w(wk[u], Xak, Xbk, Xan, Xbn);
        if (BLUE) z = cmul(z, ck[u]);
        D.buf[pidx(k)] = z;
        if (nk != k) {
          z = dct3_pre_w(wn[u], Xan, Xbn, Xak, Xbk);
          if (BLUE) z = cmul(z, cn[u]);
          D.buf[pidx(nk)] = z;
        }
      }
    }
  }
}

// Store: sample m of both lines (dct3_pair_value) back to the field.
template <int U, bool BLUE, bool DIRECT>
__device__ __forceinline__ void col_store(const DctCtx &D, float *__restrict__ la, long long stride, bool hasb,
                                          int H) {
  const unsigned T = blockDim.x, Hu = H;
  const float2 *__restrict__ ch = D.tab + D.P->chirp_off;
  const float s = D.P->invN;
  for (unsigned m0 = threadIdx.x; m0 < Hu; m0 += U * T) {
    float2 cm[U];
    if (BLUE) {
#pragma unroll
      for (int u = 0; u < U; ++u) cm[u] = ch[makhoul_pos(min(m0 + u * T, Hu - 1), H)];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const unsigned m = m0 + u * T;
      if (m >= Hu) continue;
      float2 w = D.buf[pidx(makhoul_pos(m, H))];
      if (BLUE) w = cmul(cm[u], cconj(w));
      const float a = w.x * s, b = -w.y * s;
      if (DIRECT) {
        float *p = la + (long long)m * stride;
        if (hasb) *(float2 *)p = make_float2(a, b);
        else p[0] = a;
      } else {
        la[m] = a;
        if (hasb) la[stride + m] = b;
      }
    }
  }
}

// Coefficient store of the row DCT-II: for k = 0 .. W/2 and nk = (W - k) mod W, the DCT-II
// coefficients of both lines (dct2_coef_pair) into the rows ta, tb of T; the post-twiddles and
// chirps of a batch of U pairs are loaded before any is used.
template <int U, bool BLUE>
__device__ __forceinline__ void dct2_coef_store(const DctCtx &D, float *__restrict__ out, long long ta, long long tb,
                                                bool hasb, int W) {
  const unsigned T = blockDim.x, K = W / 2 + 1;
  const float2 *__restrict__ mk = D.tab + D.P->mk_off;
  const float2 *__restrict__ ch = D.tab + D.P->chirp_off;
  for (unsigned k0 = threadIdx.x; k0 < K; k0 += U * T) {
    float2 wk[U], wn[U], ck[U], cn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const unsigned k = min(k0 + u * T, K - 1), nk = k ? W - k : 0;
      wk[u] = mk[k];
      wn[u] = mk[nk];
      if (BLUE) {
        ck[u] = ch[k];
        cn[u] = ch[nk];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const unsigned k = k0 + u * T;
      if (k >= K) continue;
      const unsigned nk = k ? W - k : 0;
      float2 Zk = D.buf[pidx(k)], Zn = D.buf[pidx(nk)];
      if (BLUE) {
        Zk = cmul(ck[u], cconj(Zk));
        Zn = cmul(cn[u], cconj(Zn));
      }
      float Xak, Xbk, Xan, Xbn;
      dct2_coef_w(Zk, Zn, wk[u], wn[u], Xak, Xbk, Xan, Xbn);
      out[ta + k] = Xak;
      if (hasb) out[tb + k] = Xbk;
      if (nk != k) {
        out[ta + nk] = Xan;
        if (hasb) out[tb + nk] = Xbn;
      }
    }
  }
}

template <bool FUSE_UPDATE, int E>
__global__ __launch_bounds__(fft_threads_max<E>()) void k_row_dct2(Ctx c, float *in_, float *__restrict__ out,
                                                               const float *__restrict__ p, int guard) {
  extern __shared__ float2 lds[];
  const int img = find_img(c.geo.rowpair, c.n, blockIdx.x);
  const ImgDesc d = c.img[img];
  const ImgState *st = &c.st[img];
  if (guard_off(st, guard)) return;
  const int ra = 2 * (blockIdx.x - c.geo.rowpair[img]);
  const bool hasb = ra + 1 < d.H;
  const int W = d.W;
  const long long oa = d.off + (long long)ra * d.P;
  long long ta, tb;
  t_rows(c, d, ra, ta, tb);
  const DctCtx D = dct_begin(lds, c.plans + d.rowplan, c.tab);
  double acc[2] = {0.0, 0.0};  // r.r, non-finite x
  if (FUSE_UPDATE) {
    // x += a p, r -= a Ap on both rows; the new r goes straight into the LDS transform buffer.
    // Loads of a chunk are all issued before its stores so that they overlap in flight.
    float *x = c.f.u;
    const float *Ap = c.f.Ap;
    const float alpha = (float)st->alpha;
    const bool upx = !d.xdefer;  // xdefer: x is updated by the next k_cg_pap instead
    const int T = blockDim.x;
    // (BLUE: the chirps of the batch's Makhoul slots are loaded with it)
    // (DEFER: x is updated by the next k_cg_pap; two loads per element instead of four, so twice
    // as many elements are in flight per batch)
    auto update = [&](auto blue_tag, auto defer_tag) {
      constexpr bool BLUE = decltype(blue_tag)::value;
      constexpr bool DEFER = decltype(defer_tag)::value;
      const float2 *__restrict__ ch = D.tab + D.P->chirp_off;
      // Bluestein rows of 4000-6000 points on 896-1024 threads: 6 per batch is all of them
      constexpr int U = DEFER ? 8 : (BLUE ? 6 : 4);
      for (int m0 = threadIdx.x; m0 < W; m0 += U * T) {
        float xv[U][2], pv[U][2], rv[U][2], av[U][2];
        float2 cm[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int m = m0 + u * T;
          if (BLUE) cm[u] = ch[makhoul_pos(min(m, W - 1), W)];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            xv[u][h] = pv[u][h] = rv[u][h] = av[u][h] = 0.f;
            if (m < W && (h == 0 || hasb)) {
              const long long k = oa + (long long)h * d.P + m;
              if (!DEFER) {
                xv[u][h] = x[k];
                pv[u][h] = p[k];
              }
              rv[u][h] = in_[k];
              av[u][h] = Ap[k];
            }
          }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int m = m0 + u * T;
          if (m >= W) break;
          float rn2[2];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const float xn = xv[u][h] + alpha * pv[u][h];
            const float rn = rv[u][h] - alpha * av[u][h];
            rn2[h] = rn;
            if (h == 0 || hasb) {
              const long long k = oa + (long long)h * d.P + m;
              if (!DEFER) {
                x[k] = xn;
                acc[1] += isfinite(xn) ? 0.0 : 1.0;
              }
              in_[k] = rn;
              acc[0] += (double)rn * rn;
            }
          }
          float2 z = make_float2(rn2[0], hasb ? rn2[1] : 0.f);
          if (BLUE) z = cmul(z, cm[u]);
          D.buf[pidx(makhoul_pos(m, W))] = z;
        }
      }
    };
    if (D.P->blue) {
      if (upx) update(BoolC<true>{}, BoolC<false>{});
      else update(BoolC<true>{}, BoolC<true>{});
    } else {
      if (upx) update(BoolC<false>{}, BoolC<false>{});
      else update(BoolC<false>{}, BoolC<true>{});
    }
  } else {
    if (D.P->blue) lines_load<COL_U, true, false>(D, in_ + oa, d.P, hasb, W);
    else lines_load<COL_U, false, false>(D, in_ + oa, d.P, hasb, W);
  }
  dct2_pair_run<E>(D);
  if (D.P->blue) dct2_coef_store<COL_U / 2, true>(D, out, ta, tb, hasb, W);
  else dct2_coef_store<COL_U / 2, false>(D, out, ta, tb, hasb, W);
  if (FUSE_UPDATE) block_partial<2, fft_threads_max<E>() / 64>(acc, c.partials, d.rowpart0 + ra / 2);
}

#ifndef IRLS2D_TRANSFORMS_ONLY
// T (H x W, pitch P) -> Tt (W x H, pitch PT) when dir == 0, Tt -> T when dir == 1.
// 64 x 64 tiles moved as float4: both pitches are multiples of 64 and offsets multiples of 64, so
// every float4 of a tile row lies inside the row's pitch; elements in the pitch padding are read
// and written but never used.  The LDS tile has a 65-float row pitch: the four scalar reads that
// gather an output float4 then hit distinct banks.
__global__ __launch_bounds__(256) void k_transpose(Ctx c, const float *__restrict__ src, float *__restrict__ dst,
                                                   int dir, int guard) {
  __shared__ float tile[TT][TT + 1];
  const int img = find_img(c.geo.trans, c.n, blockIdx.x);
  const ImgDesc d = c.img[img];
  if (guard_off(&c.st[img], guard)) return;
  const int local = blockIdx.x - c.geo.trans[img];
  const int ntx = (d.W + TT - 1) / TT;
  // tile origin in image coordinates: rows ti.., columns tj..
  const int ti = (local / ntx) * TT, tj = (local % ntx) * TT;
  // source rows (length `sl` valid of pitch `sp`) and destination rows
  const int srows = dir == 0 ? d.H : d.W, scols = dir == 0 ? d.W : d.H;
  const int s0 = dir == 0 ? ti : tj, c0 = dir == 0 ? tj : ti;  // tile origin in source coordinates
  const long long sp = dir == 0 ? d.P : d.PT, dp = dir == 0 ? d.PT : d.P;
  const float *s = src + (dir == 0 ? d.off : d.offT);
  float *o = dst + (dir == 0 ? d.offT : d.off);
  const int q4 = threadIdx.x & 15, r0 = threadIdx.x >> 4;  // float4 column, row within a 16-row slab
#pragma unroll
  for (int k = 0; k < TT / 16; ++k) {
    const int r = r0 + 16 * k;
    if (s0 + r < srows && c0 + 4 * q4 < scols) {
      const float4 v = *(const float4 *)(s + (long long)(s0 + r) * sp + c0 + 4 * q4);
      tile[r][4 * q4 + 0] = v.x;
      tile[r][4 * q4 + 1] = v.y;
      tile[r][4 * q4 + 2] = v.z;
      tile[r][4 * q4 + 3] = v.w;
    }
  }
  __syncthreads();
  // destination row j (= source column c0 + j) takes source rows s0 .. s0 + 63 of that column
#pragma unroll
  for (int k = 0; k < TT / 16; ++k) {
    const int j = r0 + 16 * k;
    if (c0 + j < scols && s0 + 4 * q4 < srows) {
      const float4 v = make_float4(tile[4 * q4 + 0][j], tile[4 * q4 + 1][j], tile[4 * q4 + 2][j], tile[4 * q4 + 3][j]);
      *(float4 *)(o + (long long)(c0 + j) * dp + s0 + 4 * q4) = v;
    }
  }
}

#endif  // IRLS2D_TRANSFORMS_ONLY
// Column pass on the transposed field: DCT-II along H, divide by -lambda (float32 lambda from
// float64 cosines, utils_unwrap2d.py:1284-1289, :1395-1398), zero the DC term, inverse DCT.
// DIRECT: the column pair is read from and written back to the row-major field (one float2 per
// row) instead of the transposed copy; blocks are remapped so that each XCD takes a contiguous
// range of column pairs, whose rows then share that XCD's L2 lines.
template <int E, bool DIRECT>
__global__ __launch_bounds__(fft_threads_max<E>()) void k_col(Ctx c, float *__restrict__ Tt, int guard) {
  extern __shared__ float2 lds[];
  const int blk = DIRECT ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int img = find_img(c.geo.colpair, c.n, blk);
  const ImgDesc d = c.img[img];
  if (guard_off(&c.st[img], guard)) return;
  const int ja = 2 * (blk - c.geo.colpair[img]);
  const bool hasb = ja + 1 < d.W;
  const int H = d.H;
  float *la = DIRECT ? Tt + d.off + ja : Tt + d.offT + (long long)ja * d.PT;
  const long long stride = DIRECT ? d.P : d.PT;
  const DctCtx D = dct_begin(lds, c.plans + d.colplan, c.tab);
  const bool blue = D.P->blue;
  if (blue) lines_load<COL_U, true, DIRECT>(D, la, stride, hasb, H);
  else lines_load<COL_U, false, DIRECT>(D, la, stride, hasb, H);
  dct2_pair_run<E>(D);
  const double *cosH = c.dtab + d.cosH_off;
  const double cwa = 2.0 * c.dtab[d.cosW_off + ja];
  const double cwb = hasb ? 2.0 * c.dtab[d.cosW_off + ja + 1] : 0.0;
  if (blue) col_middle<COL_U / 2, true>(D, cosH, cwa, cwb, hasb, ja == 0, H);
  else col_middle<COL_U / 2, false>(D, cosH, cwa, cwb, hasb, ja == 0, H);
  dct3_run<E>(D);
  if (blue) col_store<COL_U, true, DIRECT>(D, la, stride, hasb, H);
  else col_store<COL_U, false, DIRECT>(D, la, stride, hasb, H);
}

// Input step of k_row_dct3 over the pairs (j, W - j), 0 < j < (W + 1) / 2, U per thread per
// batch: a thread issues every load of a batch (both partners of both lines, the post-twiddle and,
// for Bluestein, both chirps) before it writes any slot, so one memory latency covers the batch.
// Rows are uniform bases with 32-bit offsets; the post-twiddle of slot W - j is
// exp(-i pi (W - j) / 2W) = (-w.y, -w.x) of w = exp(-i pi j / 2W).
template <int U, bool BLUE>
__device__ __forceinline__ void dct3_input_pairs(const DctCtx &D, const float *__restrict__ Ta,
                                                 const float *__restrict__ Tb, bool hasb, int W) {
  const unsigned J = (W + 1) / 2, T = blockDim.x;
  const float2 *__restrict__ mk = D.tab + D.P->mk_off;
  const float2 *__restrict__ ch = D.tab + D.P->chirp_off;
  for (unsigned j0 = threadIdx.x; j0 < J; j0 += U * T) {
    float xa[U], xb[U], ya[U], yb[U];
    float2 wk[U], cj[U], cn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const unsigned j = min(max(j0 + u * T, 1u), J - 1);
      xa[u] = Ta[j];
      ya[u] = Ta[W - j];
      xb[u] = Tb[j];
      yb[u] = Tb[W - j];
      wk[u] = mk[j];
      if (BLUE) {
        cj[u] = ch[j];
        cn[u] = ch[W - j];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const unsigned j = j0 + u * T;
      if (j >= 1 && j < J) {
        const float xbu = hasb ? xb[u] : 0.f, ybu = hasb ? yb[u] : 0.f;
        float2 zk = dct3_pre_w(wk[u], xa[u], xbu, ya[u], ybu);
        float2 zn = dct3_pre_w(make_float2(-wk[u].y, -wk[u].x), ya[u], ybu, xa[u], xbu);
        if (BLUE) {
          zk = cmul(zk, cj[u]);
          zn = cmul(zn, cn[u]);
        }
        D.buf[pidx(j)] = zk;
        D.buf[pidx(W - j)] = zn;
      }
    }
  }
}

// Output step of k_row_dct3 for even W: samples 2i and 2i + 1 are the Makhoul slots i and
// W - 1 - i (ascending and descending across the lanes: no bank conflicts), stored as float2.  The
// r of the first batch is in (rav, rbv) when `first` is set (loaded before the transform); other
// batches load their own.  For Bluestein the chirps of a batch are loaded together before any slot
// is read.
template <int U, bool BLUE>
__device__ __forceinline__ void dct3_output_pairs(const DctCtx &D, float2 *__restrict__ oa2, float2 *__restrict__ ob2,
                                                  const float2 *__restrict__ ra2, const float2 *__restrict__ rb2,
                                                  float2 (&rav)[U], float2 (&rbv)[U], bool first, bool hasb, int dot,
                                                  int W, double &acc) {
  const unsigned J2 = W / 2, T = blockDim.x;
  const float s = D.P->invN;
  const float2 *__restrict__ ch = D.tab + D.P->chirp_off;
  for (unsigned i0 = threadIdx.x; i0 < J2; i0 += U * T) {
    float2 c0[U], c1[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const unsigned i = min(i0 + u * T, J2 - 1);
      if (dot && (i0 != threadIdx.x || !first)) {
        rav[u] = ra2[i];
        rbv[u] = rb2[i];
      }
      if (BLUE) {
        c0[u] = ch[i];
        c1[u] = ch[W - 1 - i];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const unsigned i = i0 + u * T;
      if (i >= J2) continue;
      float2 w0 = D.buf[pidx(i)], w1 = D.buf[pidx(W - 1 - i)];
      if (BLUE) {
        w0 = cmul(c0[u], cconj(w0));
        w1 = cmul(c1[u], cconj(w1));
      }
      const float a0 = w0.x * s, b0 = -w0.y * s, a1 = w1.x * s, b1 = -w1.y * s;
      oa2[i] = make_float2(a0, a1);
      if (hasb) ob2[i] = make_float2(b0, b1);
      if (dot)
        acc += ((double)rav[u].x * (double)a0 + (double)rav[u].y * (double)a1) +
               (hasb ? (double)rbv[u].x * (double)b0 + (double)rbv[u].y * (double)b1 : 0.0);
    }
  }
}

// Row DCT-III: T -> out, with the partial sums of r.out when `dot` is set.
// E = 16 (lengths <= 8192): two workgroups of 8 waves per CU fit the LDS, so at most 128 VGPRs;
// E = 12: one of 16 waves, the same cap; E = 24: one of 8 waves.
template <int E>
__global__ __launch_bounds__(fft_threads_max<E>()) __attribute__((amdgpu_waves_per_eu(E == 24 ? 2 : 4))) void k_row_dct3(
    Ctx c, const float *__restrict__ T_, float *__restrict__ out, int dot, int guard) {
  extern __shared__ float2 lds[];
  const int img = find_img(c.geo.rowpair, c.n, blockIdx.x);
  const ImgDesc d = c.img[img];
  if (guard_off(&c.st[img], guard)) return;
  const int ra = 2 * (blockIdx.x - c.geo.rowpair[img]);
  const bool hasb = ra + 1 < d.H;
  const int W = d.W, T = blockDim.x;
  const long long oa = d.off + (long long)ra * d.P;
  long long ta, tb;
  t_rows(c, d, ra, ta, tb);
  const DctCtx D = dct_begin(lds, c.plans + d.rowplan, c.tab);
  const bool blue = D.P->blue;
  const float *r = c.f.r;
  // slot 0 and, for even W, the self-paired slot W / 2; then the pairs
  const float *__restrict__ Ta = T_ + ta;
  const float *__restrict__ Tb = hasb ? T_ + tb : Ta;
  if (threadIdx.x == 0) {
    const float2 *mk = D.tab + D.P->mk_off;
    dct3_put(D, 0, mk[0], Ta[0], hasb ? Tb[0] : 0.f, 0.f, 0.f);
    if (!(W & 1)) {
      const float h = Ta[W / 2], hb = hasb ? Tb[W / 2] : 0.f;
      dct3_put(D, W / 2, mk[W / 2], h, hb, h, hb);
    }
  }
  if (W > 2) {
    if (blue) dct3_input_pairs<DCT3_U / 2, true>(D, Ta, Tb, hasb, W);
    else dct3_input_pairs<DCT3_U, false>(D, Ta, Tb, hasb, W);
  }
  // float2 output for even W and 8-byte aligned rows; r of the first output batch (not Bluestein:
  // its output step needs the registers for the chirps) is issued here, after the input step (a wait at the input loop's head would otherwise cover it) and before the
  // transform, under which its latency hides (index clamped so that every load is unconditional; a
  // missing second line reads the first)
  const bool vec = !(W & 1) && !((reinterpret_cast<uintptr_t>(out) | reinterpret_cast<uintptr_t>(r)) & 7);
  const unsigned J2 = W / 2;
  const float2 *__restrict__ ra2 = reinterpret_cast<const float2 *>(r + oa);
  const float2 *__restrict__ rb2 = hasb ? reinterpret_cast<const float2 *>(r + oa + d.P) : ra2;
  float2 rav[DCT3_UO], rbv[DCT3_UO];
  if (vec && dot && !blue) {
#pragma unroll
    for (int u = 0; u < DCT3_UO; ++u) {
      const unsigned i = min(threadIdx.x + u * T, J2 - 1);
      rav[u] = ra2[i];
      rbv[u] = rb2[i];
    }
  }
  dct3_run<E>(D);
  double acc[1] = {0.0};
  if (vec) {
    float2 *__restrict__ oa2 = reinterpret_cast<float2 *>(out + oa);
    float2 *__restrict__ ob2 = reinterpret_cast<float2 *>(out + oa + d.P);
    if (blue) {
      float2 rav2[DCT3_UO / 2], rbv2[DCT3_UO / 2];
      dct3_output_pairs<DCT3_UO / 2, true>(D, oa2, ob2, ra2, rb2, rav2, rbv2, false, hasb, dot, W, acc[0]);
    } else {
      dct3_output_pairs<DCT3_UO, false>(D, oa2, ob2, ra2, rb2, rav, rbv, true, hasb, dot, W, acc[0]);
    }
  } else {
    constexpr int U = 4;
    for (int m0 = threadIdx.x; m0 < W; m0 += U * T) {
      float rs[U], rt[U];
      if (dot) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int m = m0 + u * T;
          rs[u] = (m < W) ? r[oa + m] : 0.f;
          rt[u] = (m < W && hasb) ? r[oa + d.P + m] : 0.f;
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int m = m0 + u * T;
        if (m >= W) break;
        float a, b;
        dct3_pair_value(D, m, a, b);
        out[oa + m] = a;
        if (hasb) out[oa + d.P + m] = b;
        if (dot) acc[0] += (double)rs[u] * (double)a + (double)rt[u] * (double)b;
      }
    }
  }
  if (dot) block_partial<1, fft_threads_max<E>() / 64>(acc, c.partials, d.rowpart0 + ra / 2);
}

#ifndef IRLS2D_TRANSFORMS_ONLY
// ---------------------------------------------------------------------------------------------
// Column half of the preconditioner as tridiagonal solves (utils_unwrap2d.py:1390-1403 restated)
// ---------------------------------------------------------------------------------------------
// The preconditioner z = IDCT2(-DCT2(r) / lambda) with DC zeroed is the pseudo-inverse of the
// Neumann five-point Laplacian L_W (+) L_H (lambda_kl = -(mu_k + nu_l), mu_k = 4 sin^2(pi k / 2W),
// nu_l = 4 sin^2(pi l / 2H): the DCT-II diagonalises the second-difference matrix with
// reflecting ends).  After the row DCT-II, column k of the row spectrum holds c_k and the column
// DCT pair with the division by -lambda equals the solve (L_H + mu_k I) x = c_k, L_H =
// tridiag(-1, 2, -1) with corner entries 1: the same operator, applied without transforms.  It is
// solved by LU in float64, in place on the row-major spectrum field; column 0 (mu_0 = 0, singular)
// takes the pseudo-inverse with the mean removed from input and output, which is the DC zeroing.
//
// The LU pivots of L_H + mu I have a closed form: with cosh(theta) = 1 + mu / 2,
//   a_i = 1 / d_i = cosh((i + 1/2) theta) / cosh((i + 3/2) theta) = e^-theta (1 + t_i) / (1 + e^-2theta t_i)
//   for i < H - 1, t_i = e^-(2i+1)theta, and the last (reflecting) row
//   1 / d_{H-1} = (1 + t_{H-1}) / ((e^theta - 1)(1 - e^-2H theta))
// (for H = 1 this is 1 / mu)
// (from d_0 = 1 + mu and d_i = 2 + mu - 1 / d_{i-1} by the identity 2 cosh a cosh b = cosh(a + b) +
// cosh(a - b)); no cancellation, unlike the recurrence, whose pivots lose digits for small mu.
// Forward: w_i = a_i (c_i + w_{i-1}); backward: x_i = w_i + a_i x_{i+1}.  Once t_i < 2^-60 in every
// lane of a wave, a_i = e^-theta to double precision and the closed form is skipped.
//
// Both recurrences are first-order linear, so a run of rows [s, e) maps its incoming carry
// linearly: w_{e-1} = w0_{e-1} + (prod_{s<=j<e} a_j) w_{s-1} and x_s = x0_s + (prod a_j) x_e, with
// w0 / x0 the results from a zero carry.  k_col_trc gives each thread TRC_CPT adjacent columns of
// a chunk of L consecutive rows in registers (vector row loads and stores: a wave instruction moves
// eight row segments of TRC_COLS floats); a block covers TRC_COLS columns and a band of TRC_CH
// chunks.  Sweep F runs from a zero carry (the chunk's end value and pivot product), the TRC_CH
// transfers of each column are chained in LDS (TRC_PER chunks per thread, then an affine-map scan
// over TRC_SEG lanes), sweep B runs backward on the true w, a second chain gives each chunk its
// incoming x, and sweep C corrects x and writes it; the field is read and written once (8 B/px).  Columns taller than one band of TRC_CH x TRC_L rows are cut
// into bands: k_col_trc<L, true> stores each band's transfers with zero outer carries (end w,
// pivot product, x at the band's first row and that x's response to a unit incoming w),
// k_col_trc_carry chains the bands of each column, and k_col_trc<L, false> solves every band from
// its carries (12 B/px).  Column 0 runs through the same sweeps with meaningless values; the final
// pass keeps its input, and k_col_trc_dc writes the singular solve over it.

constexpr int TRI_U = 16;  // rows per load batch of the singular column
constexpr int TRC_CPT = 2;       // columns per thread (float2 row accesses)
constexpr int TRC_COLS = 16;     // columns per block: 8 lanes of TRC_CPT columns
constexpr int TRC_CH = 64;       // chunks per column and band
constexpr int TRC_THREADS = TRC_COLS / TRC_CPT * TRC_CH;  // 512 = two waves per SIMD per block
// chaining: TRC_PER chunks composed per thread, then a TRC_SEG-lane scan
constexpr int TRC_PER = TRC_COLS * TRC_CH / TRC_THREADS;
constexpr int TRC_SEG = TRC_CH / TRC_PER;
static_assert(TRC_SEG <= 64 && (TRC_SEG & (TRC_SEG - 1)) == 0, "scan segments within a wave");
// the thread's TRC_CPT columns of one row as one vector access
typedef float TrcVec __attribute__((ext_vector_type(TRC_CPT)));
__device__ __forceinline__ TrcVec trc_vec(const float (&v)[TRC_CPT]) {
  TrcVec r;
#pragma unroll
  for (int i = 0; i < TRC_CPT; ++i) r[i] = v[i];
  return r;
}
constexpr int TRC_L = 24;        // rows per chunk of a band: bands of 1536 rows
// one-band columns: chunks of TRC_L_MIN, TRC_L_MIN + TRC_L_STEP, ... TRC_L rows, the shortest that
// covers the column in TRC_CH chunks (a kernel instantiation and a launch per length)
constexpr int TRC_L_MIN = 8;
constexpr int TRC_L_STEP = 4;
constexpr int TRC_NL = (TRC_L - TRC_L_MIN) / TRC_L_STEP + 1;
static_assert(TRC_NL <= 9, "Geo::coltri holds one table per chunk length class");
constexpr double TRI_TSMALL = 0x1p-60;
// dynamic LDS bytes of k_col_trc: pivot constants (TRC_NPIV per column), then 2 (3 with the band
// transfers) arrays of TRC_COLS x (TRC_CH + 1) doubles
constexpr int TRC_NPIV = 6;
constexpr size_t trc_lds_bytes(bool transfer) {
  return (TRC_NPIV * TRC_COLS + (transfer ? 3 : 2) * TRC_COLS * (TRC_CH + 1)) * sizeof(double);
}

struct TriPivots {
  double em, ep, q, eq2, theta, inv_last;  // e^-theta, e^theta, e^-2theta, e^2theta
  __device__ __forceinline__ double t_at(int i) const { return exp(-(2.0 * i + 1.0) * theta); }
  // 1 / d_i for i < H - 1: e^-theta (1 + t) / (1 + e^-2theta t), 1 + e^-2theta t in [1, 2), by the
  // hardware reciprocal and two Newton steps
  __device__ __forceinline__ double inv(double t) const {
    const double x = fma(q, t, 1.0);
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return em * (1.0 + t) * r;
  }
};

// x = L_H^+ (c - mean c) with mean(x) = 0, for column 0 (input src, output column 0 of base), by one
// wave.  With m = mean c,
// G_i = sum_{j<=i} c_j and X_i = -sum_{j<i} G_j (X_0 = 0), the solution of x_{i+1} = x_i - g_i,
// g_i = sum_{j<=i} (c_j - m), is x_i = X_i + m i (i + 1) / 2 - mean, where
// sum_i X_i = -sum_j (H - 1 - j) G_j.  Lane l owns rows [l L, l L + L): one pass sums its rows
// locally (G and the weighted sums from G = 0), wave scans give each lane G and X at its first row
// and the totals, a second pass writes its rows.
__device__ __forceinline__ double wave_scan_excl(double v, int lane) {
  double inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  return inc - v;
}

__device__ void tri_singular(const float *__restrict__ src, float *__restrict__ base, unsigned P, int H, int lane) {
  const int L = (H + 63) / 64;
  const int r0 = min(lane * L, H), r1 = min(r0 + L, H);
  const int len = r1 - r0;
  double sc = 0.0, sg = 0.0, sw = 0.0, g = 0.0;
  for (int i0 = r0; i0 < r1; i0 += TRI_U) {
    float v[TRI_U];
#pragma unroll
    for (int u = 0; u < TRI_U; ++u) v[u] = src[min(i0 + u, H - 1)];
#pragma unroll
    for (int u = 0; u < TRI_U; ++u) {
      const int i = i0 + u;
      if (i < r1) {
        sc += v[u];
        g += v[u];
        sg += g;
        sw += (double)(H - 1 - i) * g;
      }
    }
  }
  const double gstart = wave_scan_excl(sc, lane);
  const double m = wave_sum(sc) / H;
  const double xstart = -wave_scan_excl(len * gstart + sg, lane);
  // sum_{j in rows} (H - 1 - j) = len (H - 1) - (r0 + r1 - 1) len / 2
  const double wsum = (double)len * (H - 1) - 0.5 * (double)(r0 + r1 - 1) * len;
  const double sumx = -wave_sum(gstart * wsum + sw);
  const double mean = (sumx + m * ((double)(H - 1) * H * (H + 1) / 6.0)) / H;
  double G = gstart, X = xstart;
  for (int i0 = r0; i0 < r1; i0 += TRI_U) {
    float v[TRI_U];
#pragma unroll
    for (int u = 0; u < TRI_U; ++u) v[u] = src[min(i0 + u, H - 1)];
#pragma unroll
    for (int u = 0; u < TRI_U; ++u) {
      const int i = i0 + u;
      if (i < r1) {
        base[(size_t)i * P] = (float)(X + m * (0.5 * (double)i * (i + 1)) - mean);
        G += v[u];
        X -= G;
      }
    }
  }
}

// x^N for a compile-time N >= 1 by repeated squaring.
template <int N>
__device__ __forceinline__ double ipow(double x) {
  if constexpr (N == 1) return x;
  else {
    const double h = ipow<N / 2>(x);
    return (N & 1) ? h * h * x : h * h;
  }
}

// A pointer the compiler can keep in scalar registers (every lane of the wave holds the same one).
template <typename F>
__device__ __forceinline__ F *uniform_ptr(F *p) {
  const unsigned long long a = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a), hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
  return (F *)(((unsigned long long)hi << 32) | lo);
}

// Pivots of column k (mu_k = 4 sin^2(pi k / 2W)) for a column of H rows.
__device__ __forceinline__ TriPivots tri_pivots(int k, int W, int H) {
  const double sn = sin(M_PI * (double)k / (2.0 * W));
  const double mu = 4.0 * sn * sn;
  const double ch = 1.0 + 0.5 * mu, sh = sqrt(mu * (1.0 + 0.25 * mu));
  TriPivots tp;
  tp.ep = ch + sh;
  tp.em = 1.0 / (ch + sh);
  tp.q = tp.em * tp.em;
  tp.eq2 = (ch + sh) * (ch + sh);
  const double ep1 = 0.5 * mu + sh;  // e^theta - 1
  tp.theta = log1p(ep1);
  // 1 - e^-theta t_{H-1} = 1 - e^-2H theta
  tp.inv_last = (1.0 + tp.t_at(H - 1)) / (ep1 * -expm1(-2.0 * H * tp.theta));
  return tp;
}

// Inclusive scan of the affine maps x -> a x + b over groups of TRC_SEG lanes: lane l of a group
// ends with the composition f_l o ... o f_0 of its own map and those of the lanes below it.
__device__ __forceinline__ void affine_scan_seg(double &a, double &b, int l) {
#pragma unroll
  for (int dd = 1; dd < TRC_SEG; dd <<= 1) {
    const double ap = __shfl_up(a, dd, TRC_SEG), bp = __shfl_up(b, dd, TRC_SEG);
    if (l >= dd) {
      b = fma(a, bp, b);
      a *= ap;
    }
  }
}

// Block of TRC_COLS columns C g .. C g + C - 1 (C = TRC_COLS) and one band of TRC_CH chunks of L
// rows (the image's logical blocks, numbered XCD-major: g fastest, then the band).  Thread t holds
// columns C g + TRC_CPT (t mod 8) + i (i < TRC_CPT) of chunk t / 8.  Band carries, [6][nb][W]
// doubles from tri_off: the band's end w, pivot product, first-row x and unit response (written by
// TRANSFER), its incoming w and x (written by k_col_trc_carry).  Columns from W to the next
// multiple of C lie in the row padding (the pitch is a multiple of 64): they are solved with
// meaningless values that no reader uses.
template <int L, bool TRANSFER>
__global__ __launch_bounds__(TRC_THREADS, 4) void k_col_trc(Ctx c, float *__restrict__ T, int guard) {
  const int *start = TRANSFER ? c.geo.tritr : c.geo.coltri[(L - TRC_L_MIN) / TRC_L_STEP];
  const int gb = xcd_block(blockIdx.x, gridDim.x);  // neighbouring column groups share row lines
  const int img = find_img(start, c.n, gb);
  const ImgDesc d = c.img[img];
  if (guard_off(&c.st[img], guard)) return;
  const int H = d.H, W = d.W;
  const unsigned P = d.P;
  float *__restrict__ base = T + d.off;
  const int bi = gb - start[img];
  const int ngrp = (W + TRC_COLS - 1) / TRC_COLS;
  const int grp = bi % ngrp, band = bi / ngrp;
  const int nb = d.tri_nb;
  const int tid = (int)threadIdx.x, qd = tid & 7, ch = tid >> 3;
  const int k0 = grp * TRC_COLS + TRC_CPT * qd;  // the thread's first column
  const int r0 = band * TRC_CH * L, rend = min(H, r0 + TRC_CH * L);
  const int s = r0 + ch * L;
  const int len = max(0, min(L, rend - s));
  const int lastu = max(len - 1, 0);
  const int sc = min(s, H - 1);
  double *__restrict__ tri = c.tri + d.tri_off;
  const size_t bandW = (size_t)nb * W;

  // pivots of the block's columns, computed once per column by the first wave while the other
  // waves' loads are in flight
  // LDS (dynamic: the block needs more than 64 KiB, which static __shared__ arrays did not give
  // correctly on this platform): pivots of the block's columns, then the chained transfers per
  // column and chunk (end w or x, pivot product, and for TRANSFER the unit response)
  extern __shared__ double trc_lds[];
  double(*sPiv)[TRC_COLS] = reinterpret_cast<double(*)[TRC_COLS]>(trc_lds);
  double(*sW)[TRC_CH + 1] = reinterpret_cast<double(*)[TRC_CH + 1]>(trc_lds + TRC_NPIV * TRC_COLS);
  double(*sP)[TRC_CH + 1] = sW + TRC_COLS;
  double(*sR)[TRC_CH + 1] = sP + TRC_COLS;  // TRANSFER only
  if (tid < TRC_COLS) {
    const TriPivots t = tri_pivots(min(grp * TRC_COLS + tid, W - 1), W, H);
    sPiv[0][tid] = t.em;
    sPiv[1][tid] = t.q;
    sPiv[2][tid] = t.eq2;
    sPiv[3][tid] = t.inv_last;
    sPiv[4][tid] = t.theta;
    sPiv[5][tid] = t.ep;
  }

  // the chunk's rows (float4 of the thread's four columns; the row pitch is a multiple of 64 and
  // k0 of 4, so every access is 16-byte aligned and inside the row).  A short chunk re-reads its
  // last row, an empty one row H - 1.
  TrcVec *__restrict__ rowp = reinterpret_cast<TrcVec *>(base + (size_t)sc * P + k0);
  const size_t P4 = P / TRC_CPT;
  const bool full = !__ballot(!(len == L && s + L < H));
  float v[L][TRC_CPT];
#pragma unroll
  for (int u = 0; u < L; ++u) {
    const TrcVec q = rowp[(size_t)min(u, lastu) * P4];
#pragma unroll
    for (int i = 0; i < TRC_CPT; ++i) v[u][i] = q[i];
  }
  // column 0's input for k_col_trc_dc (the sweeps below overwrite it)
  if (!TRANSFER && k0 == 0) {
    float *__restrict__ dc = (float *)(c.tri + d.tri_dc);
#pragma unroll
    for (int u = 0; u < L; ++u)
      if (u < len) dc[s + u] = v[u][0];
  }
  __syncthreads();
  // the pivots of the thread's four columns, copied to registers once (the sweeps read no LDS)
  double em[TRC_CPT], pq[TRC_CPT], peq2[TRC_CPT], plast[TRC_CPT], pth[TRC_CPT], ep[TRC_CPT];
#pragma unroll
  for (int i = 0; i < TRC_CPT; ++i) {
    em[i] = sPiv[0][TRC_CPT * qd + i];
    pq[i] = sPiv[1][TRC_CPT * qd + i];
    peq2[i] = sPiv[2][TRC_CPT * qd + i];
    plast[i] = sPiv[3][TRC_CPT * qd + i];
    pth[i] = sPiv[4][TRC_CPT * qd + i];
    ep[i] = sPiv[5][TRC_CPT * qd + i];
  }
  // every chunk of the wave complete, without row H - 1, and t_s <= 2^-60 in every column of
  // every lane: a_i = e^-theta on the whole chunk, branch-free sweeps; otherwise every row is tested
  bool cl = true;
#pragma unroll
  for (int i = 0; i < TRC_CPT; ++i) cl = cl && pth[i] * (2.0 * sc + 1.0) > 60.0 * M_LN2;
  // wave-uniform, kept in a scalar register: the per-row branches below test it without VALU work
  const bool conv = __builtin_amdgcn_readfirstlane((int)(full && !__ballot(!cl))) != 0;

  // chaining layout: thread t takes column t / TRC_SEG and chunks TRC_PER (t mod TRC_SEG) .. + TRC_PER - 1
  const int cc = tid / TRC_SEG, sg = tid % TRC_SEG;
  const int kc = min(grp * TRC_COLS + cc, W - 1);
  const bool actc = grp * TRC_COLS + cc < W && grp * TRC_COLS + cc > 0;

  // The sweeps branch per wave and row (converged or general); every barrier and chaining step
  // below is reached by all threads of the block.
  {
    double Pc[TRC_CPT], Rc[TRC_CPT];
#pragma unroll
    for (int i = 0; i < TRC_CPT; ++i) {
      Pc[i] = 1.0;
      Rc[i] = 0.0;
    }
    // Three sweeps, each reading what the previous one stored in v:
    //   F   w0 from a zero carry (stored), with the chunk's pivot product and, for TRANSFER, R;
    //   B   x0 from a zero carry (stored) on the true w = w0 + P_u w_in, x0 at the chunk's first
    //       row returned.  P_u = a_s ... a_u = e^-(u+1)theta (1 + t_s) / (1 + t_{s+u+1}) below row
    //       H - 1 (the closed-form pivots telescope); its last factor is the reciprocal the pivot
    //       of row u needs anyway, and converged rows have P_u = e^-(u+1)theta;
    //   C   x = x0 + (a_u ... a_{e-1}) x_in, to the field.
    // Rows outer, the four columns interleaved, a scheduling barrier per row.  The pivot of a row
    // is e^-theta in a converged wave, else the closed form with t advanced per row; rows past a
    // short chunk's end are skipped.
    double tf[TRC_CPT], tb[TRC_CPT];
    auto t_reset = [&]() {
#pragma unroll
      for (int i = 0; i < TRC_CPT; ++i) {
        tf[i] = 0.0;
        tb[i] = 0.0;
        if (!conv) {
          tf[i] = exp(-(2.0 * sc + 1.0) * pth[i]);
          tb[i] = exp(-(2.0 * (sc + lastu) + 1.0) * pth[i]);
        }
      }
    };
    // closed-form pivot of row u of column i and r = 1 / (1 + t_{s+u+1}) (row H - 1: its own
    // pivot, r unused); fw: forward order (t_i -> t_{i+1}), else backward
    auto gpiv = [&](int u, int i, bool fw, double &r) -> double {
      double &t = fw ? tf[i] : tb[i];
      double a = plast[i];
      r = 1.0;
      if (s + u != H - 1) {
        const double x = fma(pq[i], t, 1.0);
        double y = __builtin_amdgcn_rcp(x);
        y = fma(y, fma(-x, y, 1.0), y);
        y = fma(y, fma(-x, y, 1.0), y);
        r = y;
        a = em[i] * (1.0 + t) * y;
      }
      t *= fw ? pq[i] : peq2[i];
      return a;
    };
    auto sweep_f = [&](double (&w)[TRC_CPT]) {
      t_reset();
#pragma unroll
      for (int u = 0; u < L; ++u) {
        if (conv) {
#pragma unroll
          for (int i = 0; i < TRC_CPT; ++i) {
            w[i] = em[i] * ((double)v[u][i] + w[i]);
            v[u][i] = (float)w[i];
            if (TRANSFER) {  // R needs the running product; otherwise e^-L theta follows the loop
              Rc[i] = fma(em[i] * Pc[i], Pc[i], Rc[i]);
              Pc[i] *= em[i];
            }
          }
        } else if (u < len) {
#pragma unroll
          for (int i = 0; i < TRC_CPT; ++i) {
            double r;
            const double a = gpiv(u, i, true, r);
            w[i] = a * ((double)v[u][i] + w[i]);
            v[u][i] = (float)w[i];
            if (TRANSFER) Rc[i] = fma(a * Pc[i], Pc[i], Rc[i]);
            Pc[i] *= a;
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      if (!TRANSFER && conv) {  // a converged chunk is full: the product of its L pivots e^-theta
#pragma unroll
        for (int i = 0; i < TRC_CPT; ++i) Pc[i] = ipow<L>(em[i]);
      }
    };
    auto sweep_b = [&](const double (&win)[TRC_CPT], double (&x)[TRC_CPT]) {
      t_reset();
      // ge = w_in (1 + t_s) e^-(u+1)theta at the current row u (converged: (1 + t_s) = 1)
      double ge[TRC_CPT];
#pragma unroll
      for (int i = 0; i < TRC_CPT; ++i)
        ge[i] = conv ? Pc[i] * win[i] : win[i] * (1.0 + tf[i]) * exp(-(double)len * pth[i]);
#pragma unroll
      for (int u = L - 1; u >= 0; --u) {
        if (conv) {
#pragma unroll
          for (int i = 0; i < TRC_CPT; ++i) {
            x[i] = fma(em[i], x[i], (double)v[u][i] + ge[i]);
            v[u][i] = (float)x[i];
            ge[i] *= ep[i];
          }
        } else if (u < len) {
#pragma unroll
          for (int i = 0; i < TRC_CPT; ++i) {
            double r;
            const double a = gpiv(u, i, false, r);
            const double g = s + u == H - 1 ? Pc[i] * win[i] : ge[i] * r;
            x[i] = fma(a, x[i], (double)v[u][i] + g);
            v[u][i] = (float)x[i];
            ge[i] *= ep[i];
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    };
    auto sweep_c = [&](double (&qr)[TRC_CPT]) {
      t_reset();
#pragma unroll
      for (int u = L - 1; u >= 0; --u) {
        if (conv) {
#pragma unroll
          for (int i = 0; i < TRC_CPT; ++i) {
            qr[i] *= em[i];
            v[u][i] += (float)qr[i];
          }
          rowp[(size_t)u * P4] = trc_vec(v[u]);
        } else if (u < len) {
#pragma unroll
          for (int i = 0; i < TRC_CPT; ++i) {
            double r;
            qr[i] *= gpiv(u, i, false, r);
            v[u][i] = (float)((double)v[u][i] + qr[i]);
          }
          rowp[(size_t)u * P4] = trc_vec(v[u]);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    };

    // forward transfers, chained per column from the band's incoming w
    double w[TRC_CPT];
#pragma unroll
    for (int i = 0; i < TRC_CPT; ++i) w[i] = 0.0;
    sweep_f(w);
#pragma unroll
    for (int i = 0; i < TRC_CPT; ++i) {
      sW[TRC_CPT * qd + i][ch] = w[i];
      sP[TRC_CPT * qd + i][ch] = Pc[i];
      if (TRANSFER) sR[TRC_CPT * qd + i][ch] = Rc[i];
    }
    __syncthreads();
    {
      // the thread's TRC_PER chunks composed in order, the segments scanned, then the incoming w
      // of each of its chunks
      double a4[TRC_PER], b4[TRC_PER];
#pragma unroll
      for (int j = 0; j < TRC_PER; ++j) {
        a4[j] = sP[cc][TRC_PER * sg + j];
        b4[j] = sW[cc][TRC_PER * sg + j];
      }
      double A = a4[0], B = b4[0];
#pragma unroll
      for (int j = 1; j < TRC_PER; ++j) {
        B = fma(a4[j], B, b4[j]);
        A *= a4[j];
      }
      affine_scan_seg(A, B, sg);
      const double wext = (!TRANSFER && nb > 1) ? tri[4 * bandW + (size_t)band * W + kc] : 0.0;
      const double Ap = __shfl_up(A, 1, TRC_SEG), Bp = __shfl_up(B, 1, TRC_SEG);
      double wv = sg == 0 ? wext : fma(Ap, wext, Bp);
      double wu = sg == 0 ? 1.0 : Ap;  // response to a unit incoming w of the band (TRANSFER)
#pragma unroll
      for (int j = 0; j < TRC_PER; ++j) {
        sW[cc][TRC_PER * sg + j] = wv;
        if (TRANSFER) sR[cc][TRC_PER * sg + j] *= wu;
        wv = fma(a4[j], wv, b4[j]);
        wu *= a4[j];
      }
      if (TRANSFER && sg == TRC_SEG - 1 && actc) {
        tri[(size_t)band * W + kc] = wv;          // w at the band's last row (zero incoming w)
        tri[bandW + (size_t)band * W + kc] = wu;  // pivot product of the band
      }
    }
    __syncthreads();
    // the chunk's incoming w, then its backward transfer on the true w
#pragma unroll
    for (int i = 0; i < TRC_CPT; ++i) w[i] = sW[TRC_CPT * qd + i][ch];
    double x[TRC_CPT];
#pragma unroll
    for (int i = 0; i < TRC_CPT; ++i) x[i] = 0.0;
    sweep_b(w, x);
    __syncthreads();  // every thread has read its incoming w: sW now takes the x transfers
#pragma unroll
    for (int i = 0; i < TRC_CPT; ++i) sW[TRC_CPT * qd + i][ch] = x[i];
    __syncthreads();
    {
      // backward chain: thread (column, segment 31 - sg) composes its four chunks from the last,
      // the scan runs from the band's last segment upwards; the unit response shares the products
      const int g4 = TRC_PER * (TRC_SEG - 1 - sg);
      double a4[TRC_PER], b4[TRC_PER], u4[TRC_PER];
#pragma unroll
      for (int j = 0; j < TRC_PER; ++j) {
        a4[j] = sP[cc][g4 + j];
        b4[j] = sW[cc][g4 + j];
        u4[j] = TRANSFER ? sR[cc][g4 + j] : 0.0;
      }
      double A = a4[TRC_PER - 1], B = b4[TRC_PER - 1], U = u4[TRC_PER - 1];
#pragma unroll
      for (int j = TRC_PER - 2; j >= 0; --j) {
        B = fma(a4[j], B, b4[j]);
        if (TRANSFER) U = fma(a4[j], U, u4[j]);
        A *= a4[j];
      }
#pragma unroll
      for (int dd = 1; dd < TRC_SEG; dd <<= 1) {
        const double ap = __shfl_up(A, dd, TRC_SEG), bp = __shfl_up(B, dd, TRC_SEG);
        double up = 0.0;
        if (TRANSFER) up = __shfl_up(U, dd, TRC_SEG);
        if (sg >= dd) {
          B = fma(A, bp, B);
          if (TRANSFER) U = fma(A, up, U);
          A *= ap;
        }
      }
      const double xext = (!TRANSFER && nb > 1) ? tri[5 * bandW + (size_t)band * W + kc] : 0.0;
      const double Ap = __shfl_up(A, 1, TRC_SEG), Bp = __shfl_up(B, 1, TRC_SEG);
      double xv = sg == 0 ? xext : fma(Ap, xext, Bp);
      double xu = 0.0;
      if (TRANSFER) {
        const double Up = __shfl_up(U, 1, TRC_SEG);
        xu = sg == 0 ? 0.0 : Up;
      }
#pragma unroll
      for (int j = TRC_PER - 1; j >= 0; --j) {
        sW[cc][g4 + j] = xv;
        xv = fma(a4[j], xv, b4[j]);
        if (TRANSFER) xu = fma(a4[j], xu, u4[j]);
      }
      if (TRANSFER && sg == TRC_SEG - 1 && actc) {
        tri[2 * bandW + (size_t)band * W + kc] = xv;  // x at the band's first row (zero carries)
        tri[3 * bandW + (size_t)band * W + kc] = xu;  // its response to a unit incoming w
      }
    }
    if (TRANSFER) return;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < TRC_CPT; ++i) x[i] = sW[TRC_CPT * qd + i][ch];
    sweep_c(x);
  }
}

// Column 0 (the singular solve), one wave per image, after the final k_col_trc launches.
__global__ __launch_bounds__(64) void k_col_trc_dc(Ctx c, float *__restrict__ T, int guard) {
  const int img = find_img(c.geo.tridc, c.n, blockIdx.x);
  const ImgDesc d = c.img[img];
  if (guard_off(&c.st[img], guard)) return;
  tri_singular((const float *)(c.tri + d.tri_dc), T + d.off, (unsigned)d.P, d.H, (int)threadIdx.x);
}

// Chains the bands of each column k >= 1 (256 columns per block): the incoming w of band b + 1 is
// w_end(b) + prod(b) w_in(b), the incoming x of band b - 1 is x0(b) + prod(b) x_in(b) + unit(b) w_in(b).
__global__ __launch_bounds__(256) void k_col_trc_carry(Ctx c, int guard) {
  const int img = find_img(c.geo.tricarry, c.n, blockIdx.x);
  const ImgDesc d = c.img[img];
  if (guard_off(&c.st[img], guard)) return;
  const int W = d.W, nb = d.tri_nb;
  const int k = 1 + (int)(blockIdx.x - c.geo.tricarry[img]) * 256 + (int)threadIdx.x;
  if (k >= W) return;
  double *__restrict__ tri = c.tri + d.tri_off;
  const size_t bandW = (size_t)nb * W;
  const double *__restrict__ wend = tri + k;
  const double *__restrict__ prod = tri + bandW + k;
  const double *__restrict__ x0 = tri + 2 * bandW + k;
  const double *__restrict__ unit = tri + 3 * bandW + k;
  double *__restrict__ win = tri + 4 * bandW + k;
  double *__restrict__ xin = tri + 5 * bandW + k;
  double wv = 0.0;
  for (int b = 0; b < nb; ++b) {
    win[(size_t)b * W] = wv;
    wv = fma(prod[(size_t)b * W], wv, wend[(size_t)b * W]);
  }
  double xv = 0.0;
  for (int b = nb - 1; b >= 0; --b) {
    xin[(size_t)b * W] = xv;
    xv = fma(prod[(size_t)b * W], xv, fma(unit[(size_t)b * W], win[(size_t)b * W], x0[(size_t)b * W]));
  }
}

// ---------------------------------------------------------------------------------------------
// End of an IRLS step (utils_unwrap2d.py:1599-1626)
// ---------------------------------------------------------------------------------------------

__global__ __launch_bounds__(TILE_THREADS) void k_irls_check(Ctx c) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  if (!c.st[t.img].irls_active) return;
  double acc[3] = {0.0, 0.0, 0.0};  // non-finite u, |u_prev - u|^2, |u|^2
  for_tile(d, t, [&](int i, int j, long long idx) {
    const float u = c.f.u[idx];
    const float df = c.f.u_prev[idx] - u;
    acc[0] += isfinite(u) ? 0.0 : 1.0;
    acc[1] += (double)df * df;
    acc[2] += (double)u * u;
  });
  block_partial<3>(acc, c.partials, t.blk);
}

// u = u_best after a non-finite IRLS step.
__global__ __launch_bounds__(TILE_THREADS) void k_restore(Ctx c, int it) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  if (c.st[t.img].restore_it != it) return;
  for_tile(d, t, [&](int i, int j, long long idx) { c.f.u[idx] = c.f.u_best[idx]; });
}

// ---------------------------------------------------------------------------------------------
// Second reduction stage + scalar logic, one block per image.
// ---------------------------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_finalize(Ctx c, int kind, int it) {
  const int img = blockIdx.x;
  ImgState *st = &c.st[img];
  int guard = GUARD_NONE;
  if (kind == FIN_IRLS_MEAN || kind == FIN_IRLS_UPDATE || kind == FIN_CG_INIT || kind == FIN_IRLS_CHECK)
    guard = GUARD_IRLS;
  else if (kind == FIN_PAP || kind == FIN_UPDATE || kind == FIN_RZ_INIT || kind == FIN_RZ)
    guard = GUARD_CG;
  if (guard_off(st, guard)) return;
  const bool rowgeo = (kind == FIN_UPDATE || kind == FIN_RZ_INIT || kind == FIN_RZ);
  // rows longer than MAX_FFT_LEN: the partials come from the long passes (update: pass 1 of the
  // DCT-II, r.z: pass 2 of the DCT-III) in their own region; short rows: one slot per row pair
  // from the image's rowpart0
  const ImgDesc &dsc = c.img[img];
  const bool lng = rowgeo && dsc.rowlong >= 0;
  const double *parts = lng ? c.partials2 : c.partials;
  int b0, b1;
  if (rowgeo && !lng) {
    b0 = dsc.rowpart0;
    b1 = b0 + (dsc.H + 1) / 2;
  } else if (lng && c.lplans[dsc.rowlong].blue) {
    // Bluestein long rows: the pack (update) and post (r.z) passes share one block numbering
    b0 = dsc.lbpart;
    b1 = b0 + ((dsc.H + 1) / 2) * c.lplans[dsc.rowlong].nbp;
  } else {
    const int *geo = rowgeo ? (kind == FIN_UPDATE ? c.geo.rowL1 : c.geo.rowL2) : c.geo.tile;
    b0 = geo[img];
    b1 = geo[img + 1];
  }
  int nv = 1;
  if (kind == FIN_IRLS_UPDATE || kind == FIN_UPDATE || kind == FIN_PAP) nv = 2;
  if (kind == FIN_IRLS_CHECK) nv = 3;
  double acc[MAX_NV] = {0.0, 0.0, 0.0};
  // partials b0 + tid, + 256, + 512, ... summed in that order; loaded FIN_U at a time so that
  // their L2 latencies overlap
  constexpr int FIN_U = 16;
  for (int b = b0 + threadIdx.x; b < b1; b += FIN_U * blockDim.x) {
    double v[FIN_U][MAX_NV];
#pragma unroll
    for (int u = 0; u < FIN_U; ++u) {
      const int bb = b + u * blockDim.x;
#pragma unroll
      for (int q = 0; q < MAX_NV; ++q) v[u][q] = (bb < b1 && q < nv) ? parts[(long long)bb * MAX_NV + q] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < FIN_U; ++u) {
      if (b + u * (int)blockDim.x >= b1) break;
#pragma unroll
      for (int q = 0; q < MAX_NV; ++q)
        if (q < nv) acc[q] += v[u][q];
    }
  }
  block_sum<MAX_NV>(acc);
  if (threadIdx.x != 0) return;
  switch (kind) {
    case FIN_VCOUNT:
      st->vcount = acc[0];
      break;
    case FIN_CENTER:
    case FIN_IRLS_MEAN:
    case FIN_GSUM:
      st->mean = st->vcount > 0 ? acc[0] / st->vcount : 0.0;
      break;
    case FIN_IRLS_UPDATE:
      // utils_unwrap2d.py:1555-1558: u_best when the residual improved and u is finite
      if (acc[0] < st->best_res && acc[1] == 0.0) {
        st->best_res = acc[0];
        st->improved_it = it;
      }
      break;
    case FIN_CG_INIT:
      st->cg_active = acc[0] == 0.0 ? 1 : 0;
      st->n_cg = 0;
      st->xpend = 0;
      break;
    case FIN_PAP: {
      if (st->xpend) {
        // k_cg_pap applied the previous step's x update: a non-finite x ends PCG after that step,
        // as the reference's test right after its update does (utils_unwrap2d.py:1451-1453)
        st->xpend = 0;
        if (acc[1] != 0.0) {
          st->cg_active = 0;
          break;
        }
      }
      const double pAp = acc[0];
      st->n_cg += 1;
      st->n_cg_total += 1;
      if (!(fabs(pAp) >= 1e-15) || !isfinite(pAp)) {
        st->cg_active = 0;
      } else {
        st->alpha = clampd_nan(st->rz / pAp, -1e6, 1e6);
        st->xpend = dsc.xdefer;
      }
      break;
    }
    case FIN_UPDATE:
      if (acc[1] != 0.0) {
        st->cg_active = 0;
      } else {
        const double rn = sqrt(acc[0]);
        if (rn < c.cg_tol || !isfinite(rn)) st->cg_active = 0;
      }
      break;
    case FIN_RZ_INIT:
      st->rz = acc[0];
      break;
    case FIN_RZ:
      if (fabs(st->rz) < 1e-15) {
        st->cg_active = 0;
      } else {
        st->beta = clampd_nan(acc[0] / st->rz, -1e6, 1e6);
        st->rz = acc[0];
      }
      break;
    case FIN_IRLS_CHECK:
      st->cg_active = 0;
      st->xpend = 0;  // k_cg_xcommit of this step has run
      st->n_irls = it + 1;
      if (acc[0] != 0.0) {
        st->restore_it = it;
        st->irls_active = 0;
        st->status |= 1;
      } else {
        const double rel = sqrt(acc[1]) / (sqrt(acc[2]) + 1e-10);
        if (rel < c.tol) {
          st->irls_active = 0;
          st->status |= 2;
        }
      }
      break;
    default:
      break;
  }
}

// ---------------------------------------------------------------------------------------------
// Implicit adjoint of the IRLS fixed point D^T (w_hat(r) r) = 0, r = D u - t (no reference
// counterpart).  With H = D^T W' D, W'_e = w_e eps^2 / (r_e^2 + eps^2)^(3/2) on unclamped edges and
// the clamp value on clamped ones: H lam = C^T g (C = masked-mean centring),
// dL/dpsi = valid (H lam), dL/dw_e = -(D lam)_e r_e / sqrt(r_e^2 + eps^2) (0 when clamped),
// chained through the edge average (w_i + w_j)/2 and the clip(coh, 0.01, 1).
// ---------------------------------------------------------------------------------------------

// sum of the incoming gradient over all pixels
__global__ __launch_bounds__(TILE_THREADS) void k_adj_gsum(Ctx c, const float *__restrict__ g) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  double acc[1] = {0.0};
  for_tile(d, t, [&](int i, int j, long long idx) { acc[0] += (double)g[idx]; });
  block_partial<1>(acc, c.partials, t.blk);
}

// W' (-> wxi, wyi), s' = d(w_hat r)/dw (-> u_prev, u_best), b = C^T g (-> b), lambda = 0 (-> lam)
__global__ __launch_bounds__(TILE_THREADS) void k_adj_prep(Ctx c, const float *__restrict__ u,
                                                            const float *__restrict__ g, float *__restrict__ lam) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  const float gm = (float)c.st[t.img].mean;
  const float eps_sq = c.eps_sq;
  for_tile(d, t, [&](int i, int j, long long idx) {
    float wpx = 0.f, spx = 0.f, wpy = 0.f, spy = 0.f;
    const float uc = u[idx];
    if (j < d.W - 1) {
      const float r = (u[idx + 1] - uc) - c.f.tx[idx];
      const float q = r * r + eps_sq, sq = sqrtf(q), w = c.f.wx[idx];
      const float what = w / sq;
      const bool clamped = !(what >= 1e-6f && what <= 1e6f);
      wpx = clamped ? clampf_nan(what, 1e-6f, 1e6f) : w * eps_sq / (q * sq);
      spx = clamped ? 0.f : r / sq;
    }
    if (i < d.H - 1) {
      const float r = (u[idx + d.P] - uc) - c.f.ty[idx];
      const float q = r * r + eps_sq, sq = sqrtf(q), w = c.f.wy[idx];
      const float what = w / sq;
      const bool clamped = !(what >= 1e-6f && what <= 1e6f);
      wpy = clamped ? clampf_nan(what, 1e-6f, 1e6f) : w * eps_sq / (q * sq);
      spy = clamped ? 0.f : r / sq;
    }
    c.f.wxi[idx] = wpx;
    c.f.wyi[idx] = wpy;
    c.f.u_prev[idx] = spx;
    c.f.u_best[idx] = spy;
    c.f.b[idx] = g[idx] - (c.f.valid[idx] ? gm : 0.f);
    lam[idx] = 0.f;
  });
}

// dL/dpsi into Tout and dL/dweight (optional) from D lam and s'.  The fixed point is
// translation-equivariant (u*(psi + delta) = u*(psi) + delta for delta vanishing on invalid pixels:
// every residual D u - t is unchanged), so H lam = C^T g exactly and dL/dpsi = valid C^T g = valid b;
// it is written from b directly instead of from the float32 product H lam.
__global__ __launch_bounds__(TILE_THREADS) void k_adj_grads(Ctx c, const float *__restrict__ lam,
                                                             const float *__restrict__ weight,
                                                             float *__restrict__ Tout, float *__restrict__ gw) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  const float *sx = c.f.u_prev, *sy = c.f.u_best;
  const unsigned char *v = c.f.valid;
  for_tile(d, t, [&](int i, int j, long long idx) {
    const bool vc = v[idx] != 0;
    Tout[idx] = vc ? c.f.b[idx] : 0.f;
    if (gw) {
      const float lc = lam[idx];
      float acc = 0.f;
      if (vc) {
        if (j < d.W - 1 && v[idx + 1]) acc += (lam[idx + 1] - lc) * sx[idx];
        if (j > 0 && v[idx - 1]) acc += (lc - lam[idx - 1]) * sx[idx - 1];
        if (i < d.H - 1 && v[idx + d.P]) acc += (lam[idx + d.P] - lc) * sy[idx];
        if (i > 0 && v[idx - d.P]) acc += (lc - lam[idx - d.P]) * sy[idx - d.P];
      }
      const float raw = weight ? weight[idx] : 0.f;
      const bool pass = vc && weight && raw > 0.01f && raw < 1.0f;  // clip(coh, 0.01, 1) derivative
      gw[idx] = pass ? -0.5f * acc : 0.f;
    }
  });
}

__global__ __launch_bounds__(TILE_THREADS) void k_copy(Ctx c, const float *__restrict__ src, float *__restrict__ dst) {
  const TilePos t = tile_pos(c);
  const ImgDesc d = c.img[t.img];
  for_tile(d, t, [&](int i, int j, long long idx) { dst[idx] = src[idx]; });
}

// Test hook: p(0, 0) = +inf in every image whose PCG is active (one block per image).
__global__ void k_inject(Ctx c, float *p) {
  if (threadIdx.x == 0 && c.st[blockIdx.x].cg_active) p[c.img[blockIdx.x].off] = __builtin_huge_valf();
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

#endif  // IRLS2D_TRANSFORMS_ONLY
}  // namespace irls2d
