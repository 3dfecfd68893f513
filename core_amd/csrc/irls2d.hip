// Host side of the MI355X IRLS unwrap solver: transform plans, workspace layout, the launch
// sequence of one (batched) solve, hipGraph capture/replay, and the C ABI of include/irls2d.h.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "irls2d.h"
#include "epilogue.hpp"
#include "gaussian.hpp"
#include "goldstein.hpp"
#include "instances.hpp"

namespace irls2d {
// defined in fft_e12.hip, fft_e16.hip, fft_e24.hip and long_passes.hip
IRLS2D_FFT_KERNELS(12, extern)
IRLS2D_FFT_KERNELS(16, extern)
IRLS2D_FFT_KERNELS(24, extern)
IRLS2D_LONG_KERNELS(extern)
IRLS2D_SLAB_KERNELS(extern)
// defined in goldstein.hip
#define IRLS2D_GOLD_EXTERN(PY, PX)                                 \
  extern template __global__ void k_gold_patch<PY, PX>(GoldArgs); \
  extern template __global__ void k_gold_fold<PY, PX>(GoldArgs);
IRLS2D_GOLD_EXTERN(16, 16)
IRLS2D_GOLD_EXTERN(16, 32)
IRLS2D_GOLD_EXTERN(16, 64)
IRLS2D_GOLD_EXTERN(32, 16)
IRLS2D_GOLD_EXTERN(32, 32)
IRLS2D_GOLD_EXTERN(32, 64)
IRLS2D_GOLD_EXTERN(64, 16)
IRLS2D_GOLD_EXTERN(64, 32)
IRLS2D_GOLD_EXTERN(64, 64)
}  // namespace irls2d

namespace irls2d {

static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(IRLS2D_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)

// ---------------------------------------------------------------------------------------------
// Plans and tables (double precision on the host, stored as float2 / double)
// ---------------------------------------------------------------------------------------------

using cd = std::complex<double>;

static bool smooth235(int64_t n) {
  if (n < 1) return false;
  for (int p : {2, 3, 5})
    while (n % p == 0) n /= p;
  return n == 1;
}

static std::vector<int> radices(int L) {
  std::vector<int> r;
  while (L % 16 == 0) { r.push_back(16); L /= 16; }
  while (L % 8 == 0) { r.push_back(8); L /= 8; }
  while (L % 4 == 0) { r.push_back(4); L /= 4; }
  while (L % 2 == 0) { r.push_back(2); L /= 2; }
  while (L % 3 == 0) { r.push_back(3); L /= 3; }
  while (L % 5 == 0) { r.push_back(5); L /= 5; }
  return r;  // L == 1 for 5-smooth lengths
}

// Recursive mixed-radix DFT in double (host only; builds the Bluestein kernel spectrum).
static void host_fft(std::vector<cd> &a) {
  const size_t L = a.size();
  if (L <= 1) return;
  size_t p = 2;
  while (L % p) ++p;
  const size_t m = L / p;
  std::vector<std::vector<cd>> sub(p, std::vector<cd>(m));
  for (size_t r = 0; r < p; ++r)
    for (size_t k = 0; k < m; ++k) sub[r][k] = a[k * p + r];
  for (auto &s : sub) host_fft(s);
  for (size_t k = 0; k < L; ++k) {
    cd acc = 0;
    for (size_t r = 0; r < p; ++r) {
      const double ang = -2.0 * M_PI * (double)((r * k) % L) / (double)L;
      acc += sub[r][k % m] * cd(std::cos(ang), std::sin(ang));
    }
    a[k] = acc;
  }
}

struct HostDct {
  DctPlan plan;                 // offsets relative to `tab` below
  std::vector<float2> tab;
};

static std::mutex g_plan_mu;
static std::map<int, std::shared_ptr<HostDct>> g_plans;

// Plan choices set by an autotuner (irls2d_plan_override): FFT length (N, or the Bluestein length),
// Stockham radix order and elements per thread, per DCT length.
struct PlanOverride {
  int fft_len;
  std::vector<int> rad;
  int ept;
};
static std::map<int, PlanOverride> g_over;
// Column pass strategy: 0 = automatic, 1 = transpose + column kernel, 2 = column kernel on the
// row-major field, 3 = four-step slab passes on the row-major field (irls2d_set_column_mode).
static int g_col_mode = 0;
// Automatic mode: slab passes for 5-smooth column lengths from this length on.
static int g_slab_min = 256;

// FFT lengths with a compile-time Stockham schedule (fft_lds_ct): 2^a * {1, 3, 5}, 1024 .. 12288.
static bool has_ct(int64_t M) {
  if (M < 1024 || M > MAX_FFT_LEN) return false;
  int64_t f = M;
  while (f % 2 == 0) f /= 2;
  return f == 1 || f == 3 || f == 5;
}

// Bluestein length: the 5-smooth M >= 2N - 1 (within 30%) with the lowest M * (passes + 1).
static int64_t bluestein_len(int64_t N) {
  const int64_t lo = 2 * N - 1;
  int64_t best = -1;
  double best_cost = 0;
  const int64_t hi = std::max<int64_t>(lo, std::min<int64_t>(lo + lo * 3 / 10 + 8, MAX_FFT_LEN));
  for (int64_t M = lo; M <= hi; ++M) {
    if (!smooth235(M)) continue;
    const double cost = (double)M * (double)(radices((int)std::min<int64_t>(M, 1 << 30)).size() + 1);
    if (best < 0 || cost < best_cost) { best = M; best_cost = cost; }
  }
  if (best < 0) {  // beyond the LDS limit: first 5-smooth length (reported as unsupported)
    best = lo;
    while (!smooth235(best)) ++best;
  }
  return best;
}

// Four-step split of a 5-smooth length N > MAX_FFT_LEN: N = N1 N2, both <= LONG_MAX_FACTOR, the
// most balanced such pair (N1 >= N2).
static bool long_split(int64_t N, int &N1, int &N2) {
  if (N <= MAX_FFT_LEN || !smooth235(N)) return false;
  int64_t best = -1;
  for (int64_t b = 2; b <= LONG_MAX_FACTOR; ++b) {
    if (N % b) continue;
    const int64_t a = N / b;
    if (a > LONG_MAX_FACTOR || a < b) continue;
    if (best < 0 || a < best) best = a;
  }
  if (best < 0) return false;
  N1 = (int)best;
  N2 = (int)(N / best);
  return true;
}

// The most balanced split M = M1 M2 (M1 >= M2, both <= LONG_MAX_FACTOR) of a 5-smooth M, or false.
static bool factor_split(int64_t M, int &M1, int &M2) {
  if (!smooth235(M)) return false;
  int64_t best = -1;
  for (int64_t b = 2; b <= LONG_MAX_FACTOR; ++b) {
    if (M % b) continue;
    const int64_t a = M / b;
    if (a > LONG_MAX_FACTOR || a < b) continue;
    if (best < 0 || a < best) best = a;
  }
  if (best < 0) return false;
  M1 = (int)best;
  M2 = (int)(M / best);
  return true;
}

// Bluestein length of a line longer than MAX_FFT_LEN without a smooth split: the 5-smooth
// M >= 2N - 1 (within 30%) that splits into two factors <= LONG_MAX_FACTOR, lowest
// M * (passes of M1 + passes of M2); 0 when there is none (N above about 2^19).
static int64_t long_blue_len(int64_t N, int &M1, int &M2) {
  const int64_t lo = 2 * N - 1;
  int64_t best = 0;
  double best_cost = 0;
  for (int64_t M = lo; M <= lo + lo * 3 / 10 + 8 && M <= (int64_t)LONG_MAX_FACTOR * LONG_MAX_FACTOR; ++M) {
    int a, b;
    if (!factor_split(M, a, b)) continue;
    const double cost = (double)M * (double)(radices(a).size() + radices(b).size());
    if (!best || cost < best_cost) {
      best = M;
      best_cost = cost;
      M1 = a;
      M2 = b;
    }
  }
  return best;
}

// Lines whose DCTs run as two four-step passes: longer than MAX_FFT_LEN, or with a Bluestein
// length beyond it (prime factors above 5 from 6145 on).
static bool long_line(int64_t N) {
  return N > MAX_FFT_LEN || (!smooth235(N) && bluestein_len(N) > MAX_FFT_LEN);
}

int supported(int64_t N) {
  if (N < 1) return 0;
  if (long_line(N)) {
    int a, b;
    return (long_split(N, a, b) || long_blue_len(N, a, b) > 0) ? 1 : 0;
  }
  return 1;
}

// 16 values per thread up to 8192 points (two workgroups of 8 waves per CU); above, 12 on up to
// 1024 threads (one workgroup of 16 waves: the LDS leaves room for one).
static int default_ept(int M) { return M <= 16 * MAX_FFT_THREADS ? 16 : 12; }
static int ept_threads_max(int E) { return E == 12 ? MAX_FFT_THREADS_WIDE : MAX_FFT_THREADS; }

static std::shared_ptr<HostDct> get_plan(int N) {
  std::lock_guard<std::mutex> lk(g_plan_mu);
  auto it = g_plans.find(N);
  if (it != g_plans.end()) return it->second;
  auto h = std::make_shared<HostDct>();
  DctPlan &P = h->plan;
  std::memset(&P, 0, sizeof(P));
  P.N = N;
  auto ov = g_over.find(N);
  if (ov != g_over.end()) {
    P.M = ov->second.fft_len;
    P.blue = P.M != N ? 1 : 0;
  } else {
    P.blue = smooth235(N) ? 0 : 1;
    P.M = P.blue ? (int)bluestein_len(N) : N;
  }
  P.invN = (float)(1.0 / (double)N);
  P.ept = (ov != g_over.end() && ov->second.ept) ? ov->second.ept : default_ept(P.M);
  auto push = [&](const std::vector<cd> &v) {
    int off = (int)h->tab.size();
    for (const cd &z : v) h->tab.push_back(make_float2((float)z.real(), (float)z.imag()));
    return off;
  };
  const int M = P.M;
  // split twiddle table: exp(-2 pi i b / M) for b < 64, then exp(-2 pi i 64 a / M), a < ceil(M/64)
  std::vector<cd> tw(TW_LO + (M + 63) / 64);
  for (int b = 0; b < TW_LO; ++b) {
    const double ang = -2.0 * M_PI * (double)b / (double)M;
    tw[b] = cd(std::cos(ang), std::sin(ang));
  }
  for (int a = 0; a < (M + 63) / 64; ++a) {
    const double ang = -2.0 * M_PI * (double)(64 * (int64_t)a) / (double)M;
    tw[TW_LO + a] = cd(std::cos(ang), std::sin(ang));
  }
  P.fft.L = M;
  P.fft.tw_off = push(tw);
  const std::vector<int> rad = (ov != g_over.end() && !ov->second.rad.empty()) ? ov->second.rad : radices(M);
  P.fft.npass = (int)rad.size();
  P.fft.ct_len = (has_ct(M) && rad == radices(M)) ? M : 0;
  int Ns = 1;
  for (size_t s = 0; s < rad.size(); ++s) {
    PassDesc &pd = P.fft.pass[s];
    const int R = rad[s];
    pd.R = R;
    pd.Ns = Ns;
    pd.nb = M / R;
    pd.twmul = M / (Ns * R);
    pd.rstride = (pd.nb % 16 == 0) ? pd.nb + pd.nb / 16 : 0;
    pd.wstride = (Ns % 16 == 0) ? Ns + Ns / 16 : ((Ns == 1 && 16 % R == 0) ? 1 : 0);
    pd.nsmask = ((Ns & (Ns - 1)) == 0) ? Ns - 1 : -1;
    pd.inv_ns = (float)(1.0 / (double)Ns);
    Ns *= R;
  }
  std::vector<cd> mk(N);
  for (int k = 0; k < N; ++k) {
    const double ang = -M_PI * (double)k / (2.0 * (double)N);
    mk[k] = cd(std::cos(ang), std::sin(ang));
  }
  P.mk_off = push(mk);
  if (P.blue) {
    std::vector<cd> ch(N);
    for (int64_t n = 0; n < N; ++n) {
      const int64_t q = (n * n) % (2 * (int64_t)N);
      const double ang = -M_PI * (double)q / (double)N;
      ch[n] = cd(std::cos(ang), std::sin(ang));
    }
    P.chirp_off = push(ch);
    std::vector<cd> b(M, cd(0, 0));
    for (int n = 0; n < N; ++n) b[n] = std::conj(ch[n]);
    for (int n = 1; n < N; ++n) b[M - n] = std::conj(ch[n]);
    host_fft(b);
    for (auto &z : b) z /= (double)M;
    P.bhat_off = push(b);
  }
  g_plans[N] = h;
  return h;
}

struct HostLong {
  LongPlan plan;                // tw_off / mk_off relative to `tab`; f1, f2 set per layout
  std::vector<float2> tab;
};
static std::map<int, std::shared_ptr<HostLong>> g_lplans;

// FFTs per block of a long pass: the largest power of two <= LONG_MAX_BATCH whose batch of
// length-L sequences fits LONG_BATCH_ELEMS.
static int batch_log2(int L) {
  int lb = 0;
  while ((2 << lb) <= LONG_MAX_BATCH && (int64_t)(2 << lb) * L <= LONG_BATCH_ELEMS) ++lb;
  return lb;
}

static std::shared_ptr<HostLong> get_long_plan(int N) {
  std::lock_guard<std::mutex> lk(g_plan_mu);
  auto it = g_lplans.find(N);
  if (it != g_lplans.end()) return it->second;
  auto h = std::make_shared<HostLong>();
  LongPlan &P = h->plan;
  std::memset(&P, 0, sizeof(P));
  P.N = N;
  P.invN = (float)(1.0 / (double)N);
  // the FFT length of the four-step passes: N itself, or the Bluestein convolution length
  int64_t L = N;
  if (!long_split(N, P.N1, P.N2)) {
    L = long_blue_len(N, P.N1, P.N2);
    P.blue = 1;
    P.M = (int)L;
  }
  P.inv_n2 = (float)(1.0 / (double)P.N2);
  auto push = [&](double ang) { h->tab.push_back(make_float2((float)std::cos(ang), (float)std::sin(ang))); };
  P.tw_off = 0;
  for (int b = 0; b < TW_LO; ++b) push(-2.0 * M_PI * (double)b / (double)L);
  for (int64_t a = 0; a < (L + 63) / 64; ++a) push(-2.0 * M_PI * (double)(64 * a) / (double)L);
  P.mk_off = (int)h->tab.size();
  for (int k = 0; k < N; ++k) push(-M_PI * (double)k / (2.0 * (double)N));
  P.lb1 = batch_log2(P.N1);
  P.lb2 = batch_log2(P.N2);
  if (P.blue) {
    // chirp w[n] = exp(-i pi n^2 / N) and H = FFT_M(conj w wrapped) / M, from float64
    std::vector<cd> ch(N);
    for (int64_t n = 0; n < N; ++n) ch[n] = std::polar(1.0, -M_PI * (double)((n * n) % (2 * (int64_t)N)) / (double)N);
    P.chirp_off = (int)h->tab.size();
    for (const cd &z : ch) h->tab.push_back(make_float2((float)z.real(), (float)z.imag()));
    std::vector<cd> b(L, cd(0, 0));
    for (int64_t n = 0; n < N; ++n) b[n] = std::conj(ch[n]);
    for (int64_t n = 1; n < N; ++n) b[L - n] = std::conj(ch[n]);
    host_fft(b);
    P.bhat_off = (int)h->tab.size();
    for (const cd &z : b) h->tab.push_back(make_float2((float)(z.real() / (double)L), (float)(z.imag() / (double)L)));
    // fft1 / fft3: 2^lb1 columns per block; fft2: 2^lb2 rows per block; pack / post: LB_CHUNK
    P.nb1 = (P.N2 + (1 << P.lb1) - 1) >> P.lb1;
    P.nb2 = (P.N1 + (1 << P.lb2) - 1) >> P.lb2;
    P.nbp = (N + LB_CHUNK - 1) / LB_CHUNK;
    P.sb_off = (N + 31) / 32 * 32;
    P.spitch = P.sb_off + (int)((L + 31) / 32 * 32);
  } else {
    // representatives per block = half the slots; the reflection k -> N - k has the most of them
    const int C1 = (1 << P.lb1) / 2, C2 = (1 << P.lb2) / 2;
    P.nb1 = (P.N2 / 2 + 1 + C1 - 1) / C1;
    P.nb2 = (P.N1 / 2 + 1 + C2 - 1) / C2;
    P.spitch = (N + 31) / 32 * 32;
  }
  g_lplans[N] = h;
  return h;
}

// Split of a 5-smooth column length for the slab passes: N = N1 N2 such that a K2 block (2 SLAB_CP
// slots of length N2) and a K1 block (at least SLAB_CP slots of length N1) fit SLAB_E values per
// thread of at most SLAB_MAX_THREADS; the split with the smallest larger block, then the larger
// N2.  Lengths above MAX_FFT_LEN qualify too (e.g. 40000 = 250 x 160).
static bool slab_split(int64_t N, int &N1, int &N2) {
  if (N < 4 || N > (int64_t)LONG_MAX_FACTOR * LONG_MAX_FACTOR || !smooth235(N)) return false;
  const int64_t cap = (int64_t)SLAB_E * SLAB_MAX_THREADS;
  int64_t best = -1, bestN2 = 0;
  for (int64_t b = 2; b <= N / 2; ++b) {
    if (N % b) continue;
    const int64_t a = N / b;
    const int64_t foot = std::max<int64_t>(2 * SLAB_CP * b, SLAB_CP * a);
    if (foot > cap) continue;
    if (best < 0 || foot < best || (foot == best && b > bestN2)) {
      best = foot;
      bestN2 = b;
    }
  }
  if (best < 0) return false;
  N2 = (int)bestN2;
  N1 = (int)(N / bestN2);
  return true;
}

// Factors with register-stage kernels (k_slab1_f / k_slab2_f / k_slab3_f).
static bool slab_fast(int N1, int N2) {
  return (N1 == 16 || N1 == 32 || N1 == 64 || N1 == 128) && (N2 == 16 || N2 == 32 || N2 == 64);
}

struct HostSlab {
  SlabPlan plan;  // tw_off / mk_off relative to `tab`; f1, f2 set per layout
  std::vector<float2> tab;
  int threads;    // workgroup size the plan's blocks need
};
static std::map<int, std::shared_ptr<HostSlab>> g_splans;

static std::shared_ptr<HostSlab> get_slab_plan(int N) {
  std::lock_guard<std::mutex> lk(g_plan_mu);
  auto it = g_splans.find(N);
  if (it != g_splans.end()) return it->second;
  int N1 = 0, N2 = 0;
  std::shared_ptr<HostSlab> h;
  if (slab_split(N, N1, N2)) {
    h = std::make_shared<HostSlab>();
    SlabPlan &P = h->plan;
    std::memset(&P, 0, sizeof(P));
    P.N = N;
    P.N1 = N1;
    P.N2 = N2;
    P.invN = (float)(1.0 / (double)N);
    P.tw_off = 0;
    for (int b = 0; b < TW_LO; ++b) {
      const double ang = -2.0 * M_PI * (double)b / (double)N;
      h->tab.push_back(make_float2((float)std::cos(ang), (float)std::sin(ang)));
    }
    for (int a = 0; a < (N + 63) / 64; ++a) {
      const double ang = -2.0 * M_PI * (double)(64 * (int64_t)a) / (double)N;
      h->tab.push_back(make_float2((float)std::cos(ang), (float)std::sin(ang)));
    }
    P.mk_off = (int)h->tab.size();
    for (int k = 0; k < N; ++k) {
      const double ang = -M_PI * (double)k / (2.0 * (double)N);
      h->tab.push_back(make_float2((float)std::cos(ang), (float)std::sin(ang)));
    }
    const int foot = std::max(2 * SLAB_CP * N2, SLAB_CP * N1);
    h->threads = std::max(64, ((foot + SLAB_E - 1) / SLAB_E + 63) / 64 * 64);
    int lg = 0;
    if (slab_fast(N1, N2)) {
      // the register-stage kernels: 8 / (N1 / 16) values of n2 per K1 / K3 block of 256 threads
      h->threads = std::max(h->threads, 256);
      while ((1 << lg) < 128 / N1) ++lg;
    } else {
      // K1 / K3: as many n2 values per block as fill those threads, at most N2
      while ((2 << lg) <= N2 && (int64_t)SLAB_CP * (2 << lg) * N1 <= (int64_t)SLAB_E * h->threads) ++lg;
    }
    P.lb1 = 5 + lg;
    P.nb1 = (N2 + (1 << lg) - 1) >> lg;
    P.nb2 = N1 / 2 + 1;
  }
  g_splans[N] = h;
  return h;
}

// ---------------------------------------------------------------------------------------------
// Batch layout
// ---------------------------------------------------------------------------------------------

constexpr int NFIELD = 15;  // tx ty wx wy u_prev u_best wxi wyi b r z p0 p1 Ap T (Tt has its own arena)
// Field stride padding: 149 x 256 B (an odd number of 256 B units, not a multiple of 4 KiB), so
// that the same pixel of different fields does not land on the same HBM channel and bank.
constexpr int64_t FIELD_PAD = 149 * 64;

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Fixed block-start tables of a layout; the class tables (LaunchClass::geo / geo2) follow them.
enum { G_TILE = 0, G_TRANS, G_ROWL1, G_ROWL2, G_COLL1, G_COLL2, G_BROWP, G_BROW1, G_BROW2, G_BCOLP, G_BCOL1,
       G_BCOL2, G_TRI, G_TRI_END = G_TRI + TRC_NL - 1, G_TRIT, G_TRIC, G_TRIDC, G_FIXED };

// Images whose transform pass shares one kernel instantiation: one launch per class and pass.
// Every choice is a function of the image's own shape (and the global mode / plan overrides), so an
// image's arithmetic does not depend on the other images of its batch.
struct LaunchClass {
  int E = 16;          // elements per thread (row / column DCT classes)
  bool direct = false; // column class: pairs read straight from the row-major field
  int f1 = 0, f2 = 0;  // slab class: register-stage split N1 x N2 (0 = batched-LDS kernels)
  int geo = -1;        // index of the class's block-start table (rows, columns, slab K1 / K3)
  int geo2 = -1;       // slab class: table of the K2 blocks
  int threads = 64;    // workgroup size
  size_t lds = 0;      // dynamic LDS bytes
  int maxM = 1;        // longest FFT among the class's images
};

struct Layout {
  int n = 0;
  std::vector<int64_t> H, W, off, pitch, offT;
  int64_t total = 0, totalT = 0;
  int64_t fstride = 0;  // elements between consecutive workspace fields
  // metadata (host images of the device region)
  std::vector<ImgDesc> desc;
  std::vector<std::vector<int>> geo;  // G_FIXED fixed tables, then the class tables
  std::vector<LaunchClass> rowcls, colcls, slabcls;
  std::vector<DctPlan> plans;
  std::vector<LongPlan> lplans;
  std::vector<SlabPlan> splans;
  std::vector<float2> tab;
  std::vector<double> dtab;
  int max_blocks = 0, max_blocks2 = 0;
  int64_t totalS = 0;  // float2 elements of the long-transform scratch arena
  int64_t totalTri = 0;  // doubles of the band carries of the tridiagonal column solves
  size_t long_lds = 0;
  // byte offsets inside the workspace
  size_t o_fields = 0, o_Tt = 0, o_S = 0, o_tri = 0, o_valid = 0, o_meta = 0, o_desc = 0, o_state = 0, o_plans = 0,
         o_lplans = 0, o_splans = 0, o_tab = 0, o_dtab = 0, o_partials = 0, o_partials2 = 0, bytes = 0,
         meta_bytes = 0;
  std::vector<size_t> o_geo;
  uint64_t sig = 0;
  std::vector<unsigned char> meta;  // host image of the metadata region (packed once per layout)
};

static uint64_t shape_sig(int n, const int64_t *H, const int64_t *W) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) { h ^= v; h *= 1099511628211ull; };
  mix((uint64_t)n);
  for (int i = 0; i < n; ++i) { mix((uint64_t)H[i]); mix((uint64_t)W[i]); }
  return h;
}

static int64_t pitch_of(int64_t w) { return (w + 63) / 64 * 64; }

// Image i occupies rows of pitch_of(W[i]) elements starting at off[i] (a multiple of 64).
static int64_t arena_layout(int n, const int64_t *H, const int64_t *W, int64_t *off, int64_t *pitch) {
  int64_t tot = 0;
  for (int i = 0; i < n; ++i) {
    if (H[i] < 1 || W[i] < 1) return -1;
    if (off) off[i] = tot;
    if (pitch) pitch[i] = pitch_of(W[i]);
    tot += H[i] * pitch_of(W[i]);
  }
  return tot;
}

static int build_layout(int n, const int64_t *H, const int64_t *W, Layout &Lo) {
  if (n < 1) return fail(IRLS2D_EINVAL, "n must be >= 1");
  if (!H || !W) return fail(IRLS2D_EINVAL, "null shape arrays");
  Lo = Layout();
  Lo.n = n;
  Lo.H.assign(H, H + n);
  Lo.W.assign(W, W + n);
  Lo.off.resize(n);
  Lo.pitch.resize(n);
  Lo.offT.resize(n);
  Lo.total = arena_layout(n, H, W, Lo.off.data(), Lo.pitch.data());
  if (Lo.total < 0) return fail(IRLS2D_EINVAL, "every image needs H >= 1 and W >= 1");
  for (int i = 0; i < n; ++i) {
    if (H[i] > (int64_t)1 << 30 || W[i] > (int64_t)1 << 30) return fail(IRLS2D_EINVAL, "dimension too large");
    if (!supported(H[i]) || !supported(W[i]))
      return fail(IRLS2D_EUNSUPPORTED, "grid " + std::to_string(H[i]) + "x" + std::to_string(W[i]) +
                                           ": a side is longer than the two-pass transforms reach (about 2^19)");
  }
  Lo.sig = shape_sig(n, H, W);
  std::map<int, int> plan_idx, long_idx, cos_off;
  auto add_plan = [&](int N) -> int {
    auto it = plan_idx.find(N);
    if (it != plan_idx.end()) return it->second;
    auto h = get_plan(N);
    DctPlan P = h->plan;
    const int base = (int)Lo.tab.size();
    Lo.tab.insert(Lo.tab.end(), h->tab.begin(), h->tab.end());
    P.fft.tw_off += base;
    P.mk_off += base;
    if (P.blue) { P.chirp_off += base; P.bhat_off += base; }
    Lo.plans.push_back(P);
    const int idx = (int)Lo.plans.size() - 1;
    plan_idx[N] = idx;
    return idx;
  };
  auto add_long = [&](int N) -> int {
    auto it = long_idx.find(N);
    if (it != long_idx.end()) return it->second;
    auto h = get_long_plan(N);
    LongPlan P = h->plan;
    const int base = (int)Lo.tab.size();
    Lo.tab.insert(Lo.tab.end(), h->tab.begin(), h->tab.end());
    P.tw_off += base;
    P.mk_off += base;
    if (P.blue) {
      P.chirp_off += base;
      P.bhat_off += base;
    }
    P.f1 = add_plan(P.N1);
    P.f2 = add_plan(P.N2);
    Lo.lplans.push_back(P);
    const int idx = (int)Lo.lplans.size() - 1;
    long_idx[N] = idx;
    const size_t l1 = (size_t)(TW_LDS + P.N1 * ((1 << P.lb1) + 1)) * sizeof(float2);
    const size_t l2 = (size_t)(TW_LDS + P.N2 * ((1 << P.lb2) + 1)) * sizeof(float2);
    Lo.long_lds = std::max(Lo.long_lds, std::max(l1, l2));
    return idx;
  };
  // slab column plan of length N, or -1 when N has no slab split or its factors' plans are not
  // plain FFTs of their own length (a Bluestein override)
  std::map<int, int> slab_idx, slab_threads_of;
  auto add_slab = [&](int N) -> int {
    auto it = slab_idx.find(N);
    if (it != slab_idx.end()) return it->second;
    auto h = get_slab_plan(N);
    int idx = -1;
    if (h) {
      SlabPlan P = h->plan;
      P.f1 = add_plan(P.N1);
      P.f2 = add_plan(P.N2);
      if (!Lo.plans[P.f1].blue && !Lo.plans[P.f2].blue) {
        const int base = (int)Lo.tab.size();
        Lo.tab.insert(Lo.tab.end(), h->tab.begin(), h->tab.end());
        P.tw_off += base;
        P.mk_off += base;
        Lo.splans.push_back(P);
        idx = (int)Lo.splans.size() - 1;
        slab_threads_of[idx] = h->threads;
      }
    }
    slab_idx[N] = idx;
    return idx;
  };
  auto add_cos = [&](int N) -> int {
    auto it = cos_off.find(N);
    if (it != cos_off.end()) return it->second;
    const int o = (int)Lo.dtab.size();
    for (int i = 0; i < N; ++i) Lo.dtab.push_back(std::cos(M_PI * (double)i / (double)N));
    cos_off[N] = o;
    return o;
  };
  Lo.geo.assign(G_FIXED, std::vector<int>(1, 0));
  // Automatic mode takes the slab passes only with the register-stage kernels (power-of-two splits);
  // other column lengths keep the transposed or direct path (on MI355X the batched-LDS slab kernels
  // measured 0-9% slower than the transposes at 6000^2, 3000 x 4000 and 1500 x 5000, and 4% slower
  // at 40000^2).
  auto auto_slab = [&](int64_t h, int64_t w, int &f1, int &f2) -> bool {
    return h >= g_slab_min && h <= MAX_FFT_LEN && w <= MAX_FFT_LEN && slab_split(h, f1, f2) && slab_fast(f1, f2);
  };
  // class lookup (creating the class and its empty block tables on first use)
  auto new_table = [&]() -> int {
    Lo.geo.emplace_back(1, 0);
    return (int)Lo.geo.size() - 1;
  };
  auto row_class = [&](int E) -> int {
    for (size_t k = 0; k < Lo.rowcls.size(); ++k)
      if (Lo.rowcls[k].E == E) return (int)k;
    LaunchClass L;
    L.E = E;
    L.geo = new_table();
    Lo.rowcls.push_back(L);
    return (int)Lo.rowcls.size() - 1;
  };
  auto col_class = [&](int E, bool direct) -> int {
    for (size_t k = 0; k < Lo.colcls.size(); ++k)
      if (Lo.colcls[k].E == E && Lo.colcls[k].direct == direct) return (int)k;
    LaunchClass L;
    L.E = E;
    L.direct = direct;
    L.geo = new_table();
    Lo.colcls.push_back(L);
    return (int)Lo.colcls.size() - 1;
  };
  auto slab_class = [&](int f1, int f2) -> int {
    for (size_t k = 0; k < Lo.slabcls.size(); ++k)
      if (Lo.slabcls[k].f1 == f1 && Lo.slabcls[k].f2 == f2) return (int)k;
    LaunchClass L;
    L.f1 = f1;
    L.f2 = f2;
    L.geo = new_table();
    L.geo2 = new_table();
    Lo.slabcls.push_back(L);
    return (int)Lo.slabcls.size() - 1;
  };
  // per image: (table, blocks) pairs added to the tables; every other table gets 0 blocks
  std::vector<std::vector<std::pair<int, int64_t>>> adds(n);
  int rowparts = 0;
  for (int i = 0; i < n; ++i) {
    const int h = (int)H[i], w = (int)W[i];
    ImgDesc d;
    std::memset(&d, 0, sizeof(d));
    d.off = Lo.off[i];
    d.H = h;
    d.W = w;
    d.P = (int)Lo.pitch[i];
    d.PT = (int)pitch_of(h);
    const bool lw = long_line(w), lh = long_line(h);
    d.rowplan = lw ? -1 : add_plan(w);
    d.colplan = lh ? -1 : add_plan(h);
    d.rowlong = lw ? add_long(w) : -1;
    // PCG x update deferred to the next k_cg_pap where the row DCT-II pass is bandwidth-bound (a
    // Cooley-Tukey length, LDS-resident or two-pass): 12 B/px fewer there, 8 more in k_cg_pap.
    // Bluestein rows are VALU-bound and keep the update.
    d.xdefer = lw ? !Lo.lplans[d.rowlong].blue : !Lo.plans[d.rowplan].blue;
    d.cosH_off = add_cos(h);
    d.cosW_off = add_cos(w);
    // slab column pass: automatic mode on this image's own register split; mode 3 on every
    // 5-smooth column length, long sides included
    int sf1 = 0, sf2 = 0;
    const bool want_slab = g_col_mode == 3 || (g_col_mode == 0 && auto_slab(h, w, sf1, sf2));
    d.colslab = want_slab ? add_slab(h) : -1;
    d.nslab = (w + 2 * SLAB_CP - 1) / (2 * SLAB_CP);
    const bool slab = d.colslab >= 0;
    // tridiagonal column solves (8 B/px of memory traffic, 12 with bands) for every image off the
    // slab path in automatic mode and in mode 4: cheaper than a column DCT pair in every measured
    // case on MI355X (Bluestein columns of the 48-burst stack, long columns of 40000^2, and the
    // transposes + LDS column DCTs of 8192^2: 397 vs 490 us per PCG iteration); in mode 3 for the
    // heights without a slab split that would need Bluestein or a long two-pass transform.  Column
    // DCTs in modes 1 and 2: on the transposed copy, or direct on the row-major field.
    const bool hblue = !lh && Lo.plans[d.colplan].blue;
    d.coltri = !slab && (g_col_mode == 4 || g_col_mode == 0 || (g_col_mode == 3 && (hblue || lh)));
    d.collong = (lh && !slab && !d.coltri) ? add_long(h) : -1;
    const bool colk = !slab && !lh && !d.coltri;
    d.coldirect = colk && g_col_mode == 2;
    // transposed copy (W x PT), or the slab scratch (H rows of nslab x SLAB_CP float2 = P floats)
    Lo.offT[i] = Lo.totalT;
    d.offT = Lo.totalT;
    Lo.totalT += slab ? (int64_t)h * d.P : (d.coltri || d.coldirect) ? 0 : (int64_t)w * d.PT;
    int64_t sz = 0;
    if (lw) {
      const LongPlan &P = Lo.lplans[d.rowlong];
      const int64_t lp = (h + 1) / 2;
      sz = std::max<int64_t>(sz, lp * P.spitch);
      if (P.blue) {
        adds[i].push_back({G_BROWP, lp * P.nbp});
        adds[i].push_back({G_BROW1, lp * P.nb1});
        adds[i].push_back({G_BROW2, lp * P.nb2});
      } else {
        adds[i].push_back({G_ROWL1, lp * P.nb1});
        adds[i].push_back({G_ROWL2, lp * P.nb2});
      }
    }
    if (d.collong >= 0) {
      const LongPlan &P = Lo.lplans[d.collong];
      const int64_t lp = (w + 1) / 2;
      sz = std::max<int64_t>(sz, lp * P.spitch);
      if (P.blue) {
        adds[i].push_back({G_BCOLP, lp * P.nbp});
        adds[i].push_back({G_BCOL1, lp * P.nb1});
        adds[i].push_back({G_BCOL2, lp * P.nb2});
      } else {
        adds[i].push_back({G_COLL1, lp * P.nb1});
        adds[i].push_back({G_COLL2, lp * P.nb2});
      }
    }
    d.offS = Lo.totalS;
    Lo.totalS += (sz + 31) / 32 * 32;
    adds[i].push_back({G_TILE, ((int64_t)(h + TILE_H - 1) / TILE_H) * ((w + TILE_W - 1) / TILE_W)});
    if (!slab && !d.coldirect && !d.coltri)
      adds[i].push_back({G_TRANS, ((int64_t)(h + TT - 1) / TT) * ((w + TT - 1) / TT)});
    if (d.coltri) {
      // one band of TRC_CH chunks of the shortest chunk length that covers the column; taller
      // columns in bands of TRC_CH x TRC_L rows with carries (transfer, chain and final launches)
      const int64_t ngrp = (w + TRC_COLS - 1) / TRC_COLS;
      d.tri_nb = (int)((h + TRC_CH * TRC_L - 1) / (TRC_CH * TRC_L));
      const int64_t lmin = (h + TRC_CH - 1) / TRC_CH;
      const int cls = d.tri_nb > 1 ? TRC_NL - 1
                                   : (int)std::max<int64_t>(0, (lmin - TRC_L_MIN + TRC_L_STEP - 1) / TRC_L_STEP);
      // column 0's input (H floats), then the band carries
      d.tri_dc = Lo.totalTri;
      Lo.totalTri += (h + 1) / 2;
      d.tri_off = Lo.totalTri;
      adds[i].push_back({G_TRI + cls, ngrp * d.tri_nb});
      adds[i].push_back({G_TRIDC, 1});
      if (d.tri_nb > 1) {
        Lo.totalTri += 6 * (int64_t)d.tri_nb * w;
        adds[i].push_back({G_TRIT, ngrp * d.tri_nb});
        adds[i].push_back({G_TRIC, (w - 1 + 255) / 256});
      }
    }
    if (!lw) {
      const DctPlan &R = Lo.plans[d.rowplan];
      LaunchClass &C = Lo.rowcls[row_class(R.ept)];
      C.maxM = std::max(C.maxM, R.M);
      adds[i].push_back({C.geo, (h + 1) / 2});
      d.rowpart0 = rowparts;
      rowparts += (h + 1) / 2;
    }
    if (colk) {
      const DctPlan &Cp = Lo.plans[d.colplan];
      LaunchClass &C = Lo.colcls[col_class(Cp.ept, d.coldirect)];
      C.maxM = std::max(C.maxM, Cp.M);
      adds[i].push_back({C.geo, (w + 1) / 2});
    }
    if (slab) {
      const SlabPlan &P = Lo.splans[d.colslab];
      const bool fast = slab_fast(P.N1, P.N2);
      LaunchClass &C = Lo.slabcls[slab_class(fast ? P.N1 : 0, fast ? P.N2 : 0)];
      C.threads = std::max(C.threads, fast ? std::max(256, 4 * P.N2) : slab_threads_of[d.colslab]);
      const int e1 = P.N1 * ((1 << P.lb1) + 1), e2 = P.N2 * (2 * SLAB_CP + 1);
      C.lds = std::max(C.lds, (size_t)(TW_LDS + std::max(e1, e2)) * sizeof(float2));
      adds[i].push_back({C.geo, (int64_t)P.nb1 * d.nslab});
      adds[i].push_back({C.geo2, (int64_t)P.nb2 * d.nslab});
    }
    Lo.desc.push_back(d);
  }
  // block-start tables: image i's blocks follow those of images 0 .. i-1 in every table
  for (auto &g : Lo.geo) g.assign(1, 0);
  for (int i = 0; i < n; ++i) {
    std::vector<int64_t> add(Lo.geo.size(), 0);
    for (auto &pr : adds[i]) add[pr.first] += pr.second;
    for (size_t g = 0; g < Lo.geo.size(); ++g) {
      const int64_t v = (int64_t)Lo.geo[g].back() + add[g];
      if (v > INT32_MAX) return fail(IRLS2D_EUNSUPPORTED, "too many blocks in one launch");
      Lo.geo[g].push_back((int)v);
    }
  }
  Lo.max_blocks = std::max(std::max(Lo.geo[G_TILE].back(), rowparts), 1);
  // partials2: the two-pass long kernels by block index, then the Bluestein long rows' pack / post
  // blocks from each image's lbpart
  Lo.max_blocks2 = std::max(std::max(Lo.geo[G_ROWL1].back(), Lo.geo[G_ROWL2].back()),
                            std::max(Lo.geo[G_COLL1].back(), Lo.geo[G_COLL2].back()));
  for (int i = 0; i < n; ++i) Lo.desc[i].lbpart = Lo.max_blocks2 + Lo.geo[G_BROWP][i];
  Lo.max_blocks2 += Lo.geo[G_BROWP].back();
  // launch shapes per class: the workgroup covers the class's longest transform
  auto shape_of = [&](LaunchClass &C) {
    int t = (C.maxM + C.E - 1) / C.E;
    t = (t + 63) / 64 * 64;
    C.threads = std::max(64, std::min(ept_threads_max(C.E), t));
    C.lds = (size_t)(TW_LDS + lds_seq_len(C.maxM)) * sizeof(float2);
  };
  for (auto &C : Lo.rowcls) shape_of(C);
  for (auto &C : Lo.colcls) shape_of(C);
  for (auto &C : Lo.slabcls)
    if (C.f1) {
      // register-stage kernels: K1 / K3 on 256 threads, K2 on 4 N2; exchange buffer, then the K2
      // tables (5 N2 float2 twiddles and 2 N2 floats)
      C.lds = (size_t)(TW_LDS + std::max(C.f1 * SLAB_CP * (128 / C.f1), C.f2 * 2 * SLAB_CP + 6 * C.f2)) *
              sizeof(float2);
    }
  // workspace map
  size_t o = 0;
  Lo.fstride = Lo.total + FIELD_PAD;
  Lo.o_fields = o;
  o = align_up(o + (size_t)NFIELD * Lo.fstride * sizeof(float), 256);
  Lo.o_Tt = o;
  o = align_up(o + (size_t)(Lo.totalT + FIELD_PAD) * sizeof(float), 256);
  Lo.o_S = o;
  o = align_up(o + (size_t)Lo.totalS * sizeof(float2), 256);
  Lo.o_tri = o;
  o = align_up(o + (size_t)Lo.totalTri * sizeof(double), 256);
  Lo.o_valid = o;
  o = align_up(o + (size_t)Lo.total, 256);
  Lo.o_state = o;
  o = align_up(o + n * sizeof(ImgState), 256);
  Lo.o_partials = o;
  o = align_up(o + (size_t)Lo.max_blocks * MAX_NV * sizeof(double), 256);
  Lo.o_partials2 = o;
  o = align_up(o + (size_t)Lo.max_blocks2 * MAX_NV * sizeof(double), 256);
  Lo.o_meta = o;  // everything below is uploaded from the host in one copy
  Lo.o_desc = o;
  o = align_up(o + n * sizeof(ImgDesc), 256);
  Lo.o_geo.resize(Lo.geo.size());
  for (size_t g = 0; g < Lo.geo.size(); ++g) {
    Lo.o_geo[g] = o;
    o = align_up(o + (n + 1) * sizeof(int), 256);
  }
  Lo.o_plans = o;
  o = align_up(o + Lo.plans.size() * sizeof(DctPlan), 256);
  Lo.o_lplans = o;
  o = align_up(o + Lo.lplans.size() * sizeof(LongPlan), 256);
  Lo.o_splans = o;
  o = align_up(o + Lo.splans.size() * sizeof(SlabPlan), 256);
  Lo.o_tab = o;
  o = align_up(o + Lo.tab.size() * sizeof(float2), 256);
  Lo.o_dtab = o;
  o = align_up(o + Lo.dtab.size() * sizeof(double), 256);
  Lo.bytes = o;
  Lo.meta_bytes = o - Lo.o_meta;
  return IRLS2D_OK;
}

static std::vector<unsigned char> pack_meta(const Layout &Lo) {
  std::vector<unsigned char> buf(Lo.meta_bytes, 0);
  auto put = [&](size_t off, const void *src, size_t nbytes) {
    if (nbytes) std::memcpy(buf.data() + (off - Lo.o_meta), src, nbytes);
  };
  put(Lo.o_desc, Lo.desc.data(), Lo.desc.size() * sizeof(ImgDesc));
  for (size_t g = 0; g < Lo.geo.size(); ++g) put(Lo.o_geo[g], Lo.geo[g].data(), Lo.geo[g].size() * sizeof(int));
  put(Lo.o_plans, Lo.plans.data(), Lo.plans.size() * sizeof(DctPlan));
  put(Lo.o_lplans, Lo.lplans.data(), Lo.lplans.size() * sizeof(LongPlan));
  put(Lo.o_splans, Lo.splans.data(), Lo.splans.size() * sizeof(SlabPlan));
  put(Lo.o_tab, Lo.tab.data(), Lo.tab.size() * sizeof(float2));
  put(Lo.o_dtab, Lo.dtab.data(), Lo.dtab.size() * sizeof(double));
  return buf;
}

// ---------------------------------------------------------------------------------------------
// Launch sequence
// ---------------------------------------------------------------------------------------------

struct Params {
  int max_iter, cg_max_iter;
  double tol, cg_tol, epsilon;
};

// Optional per-kernel timing (IRLS2D_FLAG_TIME_KERNELS): a hipEvent pair brackets every launch
// on the launch stream; durations are accumulated per kernel name after the solve.
struct KTimer {
  std::vector<std::tuple<const char *, hipEvent_t, hipEvent_t>> pending;
};
static thread_local KTimer *t_timer = nullptr;
static std::mutex g_stats_mu;
static std::map<std::string, std::pair<double, long long>> g_kstats;

#define LAUNCH(name, kern, grid, block, lds, stream, ...)                                   \
  do {                                                                                      \
    if (t_timer) {                                                                          \
      hipEvent_t e0_ = nullptr, e1_ = nullptr;                                              \
      (void)hipEventCreate(&e0_);                                                           \
      (void)hipEventCreate(&e1_);                                                           \
      /* events attached to the dispatch packet: kernel start and end timestamps */         \
      hipExtLaunchKernelGGL(kern, grid, block, lds, stream, e0_, e1_, 0, __VA_ARGS__);      \
      t_timer->pending.emplace_back(name, e0_, e1_);                                        \
    } else {                                                                                \
      hipLaunchKernelGGL(kern, grid, block, lds, stream, __VA_ARGS__);                      \
    }                                                                                       \
  } while (0)

static int collect_timer(KTimer &T, hipStream_t s) {
  HIP_TRY(hipStreamSynchronize(s));
  std::lock_guard<std::mutex> lk(g_stats_mu);
  for (auto &e : T.pending) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, std::get<1>(e), std::get<2>(e)));
    auto &acc = g_kstats[std::get<0>(e)];
    acc.first += ms;
    acc.second += 1;
    (void)hipEventDestroy(std::get<1>(e));
    (void)hipEventDestroy(std::get<2>(e));
  }
  T.pending.clear();
  return IRLS2D_OK;
}


static Ctx make_ctx(const Layout &Lo, unsigned char *ws, float *u, const Params &p) {
  Ctx c;
  std::memset(&c, 0, sizeof(c));
  c.img = (const ImgDesc *)(ws + Lo.o_desc);
  c.st = (ImgState *)(ws + Lo.o_state);
  c.plans = (const DctPlan *)(ws + Lo.o_plans);
  c.lplans = (const LongPlan *)(ws + Lo.o_lplans);
  c.splans = (const SlabPlan *)(ws + Lo.o_splans);
  c.partials2 = (double *)(ws + Lo.o_partials2);
  c.S = (float2 *)(ws + Lo.o_S);
  c.tab = (const float2 *)(ws + Lo.o_tab);
  c.dtab = (const double *)(ws + Lo.o_dtab);
  c.partials = (double *)(ws + Lo.o_partials);
  c.geo.tile = (const int *)(ws + Lo.o_geo[G_TILE]);
  c.geo.trans = (const int *)(ws + Lo.o_geo[G_TRANS]);
  c.geo.rowL1 = (const int *)(ws + Lo.o_geo[G_ROWL1]);
  c.geo.rowL2 = (const int *)(ws + Lo.o_geo[G_ROWL2]);
  c.geo.colL1 = (const int *)(ws + Lo.o_geo[G_COLL1]);
  c.geo.colL2 = (const int *)(ws + Lo.o_geo[G_COLL2]);
  c.geo.browP = (const int *)(ws + Lo.o_geo[G_BROWP]);
  c.geo.brow1 = (const int *)(ws + Lo.o_geo[G_BROW1]);
  c.geo.brow2 = (const int *)(ws + Lo.o_geo[G_BROW2]);
  c.geo.bcolP = (const int *)(ws + Lo.o_geo[G_BCOLP]);
  c.geo.bcol1 = (const int *)(ws + Lo.o_geo[G_BCOL1]);
  c.geo.bcol2 = (const int *)(ws + Lo.o_geo[G_BCOL2]);
  for (int q = 0; q < TRC_NL; ++q) c.geo.coltri[q] = (const int *)(ws + Lo.o_geo[G_TRI + q]);
  c.geo.tritr = (const int *)(ws + Lo.o_geo[G_TRIT]);
  c.geo.tricarry = (const int *)(ws + Lo.o_geo[G_TRIC]);
  c.geo.tridc = (const int *)(ws + Lo.o_geo[G_TRIDC]);
  c.tri = (double *)(ws + Lo.o_tri);
  // class tables: set per launch (class_ctx)
  c.geo.rowpair = c.geo.colpair = c.geo.slab1 = c.geo.slab2 = nullptr;
  c.ws_base = ws;
  float *F = (float *)(ws + Lo.o_fields);
  float *fld[NFIELD];
  for (int k = 0; k < NFIELD; ++k) fld[k] = F + (size_t)k * Lo.fstride;
  c.f.tx = fld[0]; c.f.ty = fld[1]; c.f.wx = fld[2]; c.f.wy = fld[3];
  c.f.u_prev = fld[4]; c.f.u_best = fld[5]; c.f.wxi = fld[6]; c.f.wyi = fld[7];
  c.f.b = fld[8]; c.f.r = fld[9]; c.f.z = fld[10]; c.f.p0 = fld[11]; c.f.p1 = fld[12];
  c.f.Ap = fld[13]; c.f.T = fld[14];
  c.f.Tt = (float *)(ws + Lo.o_Tt);
  c.f.u = u;
  c.f.valid = ws + Lo.o_valid;
  c.n = Lo.n;
  c.eps_sq = (float)(p.epsilon * p.epsilon);
  c.tol = p.tol;
  c.cg_tol = p.cg_tol;
  return c;
}

static void fin(const Layout &Lo, const Ctx &c, int kind, int it, hipStream_t s) {
  LAUNCH("k_finalize", k_finalize, dim3(Lo.n), dim3(256), 0, s, c, kind, it);
}

// Transform launches, dispatched on the elements-per-thread template parameter.
// Long lines (W or H > MAX_FFT_LEN) run the two four-step passes instead; images of both kinds
// can share a batch, each launch covering the images of its kind.
static void long_pass1(const Layout &Lo, const Ctx &c, int dir, int mode, float *src, const float *p, int guard,
                       hipStream_t s) {
  const dim3 g(Lo.geo[dir ? G_COLL1 : G_ROWL1].back()), b(LONG_THREADS);
  const size_t l = Lo.long_lds;
  switch (mode) {
    case LONG_DCT2:
      LAUNCH(dir ? "k_long_p1_dct2_cols" : "k_long_p1_dct2_rows", (k_long_p1<LONG_DCT2, LONG_E>), g, b, l, s, c,
             dir, src, p, guard);
      break;
    case LONG_DCT2_UPDATE:
      LAUNCH("k_long_p1_dct2_update", (k_long_p1<LONG_DCT2_UPDATE, LONG_E>), g, b, l, s, c, dir, src, p, guard);
      break;
    case LONG_DCT3:
      LAUNCH(dir ? "k_long_p1_dct3_cols" : "k_long_p1_dct3_rows", (k_long_p1<LONG_DCT3, LONG_E>), g, b, l, s, c,
             dir, src, p, guard);
      break;
    default:
      LAUNCH("k_long_p1_dct3_scale", (k_long_p1<LONG_DCT3_SCALE, LONG_E>), g, b, l, s, c, dir, src, p, guard);
      break;
  }
}

static void long_pass2(const Layout &Lo, const Ctx &c, int dir, int mode, float *dst, int dot, int guard,
                       hipStream_t s) {
  const dim3 g(Lo.geo[dir ? G_COLL2 : G_ROWL2].back()), b(LONG_THREADS);
  if (mode == LONG_DCT2)
    LAUNCH(dir ? "k_long_p2_dct2_cols" : "k_long_p2_dct2_rows", (k_long_p2<LONG_DCT2, LONG_E>), g, b, Lo.long_lds,
           s, c, dir, dst, dot, guard);
  else
    LAUNCH(dir ? "k_long_p2_dct3_cols" : "k_long_p2_dct3_rows", (k_long_p2<LONG_DCT3, LONG_E>), g, b, Lo.long_lds,
           s, c, dir, dst, dot, guard);
}

// Bluestein long lines: the three four-step passes of the chirp convolution (dir 0 rows, 1 columns).
static void lb_ffts(const Layout &Lo, const Ctx &c, int dir, int guard, hipStream_t s) {
  const dim3 b(LONG_THREADS), g1(Lo.geo[dir ? G_BCOL1 : G_BROW1].back()), g2(Lo.geo[dir ? G_BCOL2 : G_BROW2].back());
  LAUNCH("k_lb_fft1", (k_lb_fft13<1, LONG_E>), g1, b, Lo.long_lds, s, c, dir, guard);
  LAUNCH("k_lb_fft2", (k_lb_fft2<LONG_E>), g2, b, Lo.long_lds, s, c, dir, guard);
  LAUNCH("k_lb_fft3", (k_lb_fft13<3, LONG_E>), g1, b, Lo.long_lds, s, c, dir, guard);
}

// Bluestein long DCT-II of rows (dir 0: src field rows -> dst T rows, with the PCG update when
// `fuse`) or columns (dir 1: Tt -> Tt).
static void lb_dct2(const Layout &Lo, const Ctx &c, int dir, bool fuse, float *src, float *dst, const float *p,
                    int guard, hipStream_t s) {
  const dim3 gp(Lo.geo[dir ? G_BCOLP : G_BROWP].back()), b(LONG_THREADS);
  if (gp.x == 0) return;
  if (fuse) LAUNCH("k_lb_pack_update", (k_lb_pack<LONG_DCT2_UPDATE>), gp, b, 0, s, c, dir, src, p, guard);
  else LAUNCH("k_lb_pack_dct2", (k_lb_pack<LONG_DCT2>), gp, b, 0, s, c, dir, src, p, guard);
  lb_ffts(Lo, c, dir, guard, s);
  LAUNCH("k_lb_post_dct2", (k_lb_post<LONG_DCT2>), gp, b, 0, s, c, dir, dst, 0, guard);
}

// Bluestein long DCT-III (dir 0: T rows -> dst field rows, r.dst partials when `dot`; dir 1: Tt ->
// Tt with the -1/lambda scaling of the preconditioner first).
static void lb_dct3(const Layout &Lo, const Ctx &c, int dir, const float *src, float *dst, int dot, int guard,
                    hipStream_t s) {
  const dim3 gp(Lo.geo[dir ? G_BCOLP : G_BROWP].back()), b(LONG_THREADS);
  if (gp.x == 0) return;
  float *sp = const_cast<float *>(src);
  if (dir) LAUNCH("k_lb_pack_scale", (k_lb_pack<LONG_DCT3_SCALE>), gp, b, 0, s, c, dir, sp, nullptr, guard);
  else LAUNCH("k_lb_pack_dct3", (k_lb_pack<LONG_DCT3>), gp, b, 0, s, c, dir, sp, nullptr, guard);
  lb_ffts(Lo, c, dir, guard, s);
  LAUNCH("k_lb_post_dct3", (k_lb_post<LONG_DCT3>), gp, b, 0, s, c, dir, dst, dot, guard);
}

// Context of one class launch: the class's block-start tables in place of rowpair / colpair /
// slab1 / slab2.
static Ctx class_ctx(const Layout &Lo, const Ctx &c, const LaunchClass &C) {
  Ctx k = c;
  const int *t = (const int *)(c.ws_base + Lo.o_geo[C.geo]);
  k.geo.rowpair = k.geo.colpair = k.geo.slab1 = t;
  k.geo.slab2 = C.geo2 >= 0 ? (const int *)(c.ws_base + Lo.o_geo[C.geo2]) : nullptr;
  return k;
}

static void row_dct2(const Layout &Lo, const Ctx &c, bool fuse, float *src, float *dst, const float *p, int guard,
                     hipStream_t s) {
  if (Lo.geo[G_ROWL1].back() > 0) {
    long_pass1(Lo, c, 0, fuse ? LONG_DCT2_UPDATE : LONG_DCT2, src, p, guard, s);
    long_pass2(Lo, c, 0, LONG_DCT2, dst, 0, guard, s);
  }
  lb_dct2(Lo, c, 0, fuse, src, dst, p, guard, s);
  for (const LaunchClass &C : Lo.rowcls) {
    const int nb = Lo.geo[C.geo].back();
    if (nb == 0) continue;
    const Ctx k = class_ctx(Lo, c, C);
    const dim3 g(nb), b(C.threads);
    if (fuse) {
      if (C.E == 16) LAUNCH("k_row_dct2_update", (k_row_dct2<true, 16>), g, b, C.lds, s, k, src, dst, p, guard);
      else if (C.E == 12) LAUNCH("k_row_dct2_update", (k_row_dct2<true, 12>), g, b, C.lds, s, k, src, dst, p, guard);
      else LAUNCH("k_row_dct2_update", (k_row_dct2<true, 24>), g, b, C.lds, s, k, src, dst, p, guard);
    } else {
      if (C.E == 16) LAUNCH("k_row_dct2", (k_row_dct2<false, 16>), g, b, C.lds, s, k, src, dst, p, guard);
      else if (C.E == 12) LAUNCH("k_row_dct2", (k_row_dct2<false, 12>), g, b, C.lds, s, k, src, dst, p, guard);
      else LAUNCH("k_row_dct2", (k_row_dct2<false, 24>), g, b, C.lds, s, k, src, dst, p, guard);
    }
  }
}

static void row_dct3(const Layout &Lo, const Ctx &c, const float *src, float *dst, int dot, int guard, hipStream_t s) {
  if (Lo.geo[G_ROWL1].back() > 0) {
    long_pass1(Lo, c, 0, LONG_DCT3, const_cast<float *>(src), nullptr, guard, s);
    long_pass2(Lo, c, 0, LONG_DCT3, dst, dot, guard, s);
  }
  lb_dct3(Lo, c, 0, src, dst, dot, guard, s);
  for (const LaunchClass &C : Lo.rowcls) {
    const int nb = Lo.geo[C.geo].back();
    if (nb == 0) continue;
    const Ctx k = class_ctx(Lo, c, C);
    const dim3 g(nb), b(C.threads);
    if (C.E == 16) LAUNCH("k_row_dct3", (k_row_dct3<16>), g, b, C.lds, s, k, src, dst, dot, guard);
    else if (C.E == 12) LAUNCH("k_row_dct3", (k_row_dct3<12>), g, b, C.lds, s, k, src, dst, dot, guard);
    else LAUNCH("k_row_dct3", (k_row_dct3<24>), g, b, C.lds, s, k, src, dst, dot, guard);
  }
}

// The final tridiagonal launches, chunk length class Q (L = TRC_L_MIN + Q TRC_L_STEP).
template <int Q>
static void tri_final_launch(const Layout &Lo, const Ctx &c, int guard, hipStream_t s) {
  if (Lo.geo[G_TRI + Q].back() > 0)
    LAUNCH("k_col_trc", (k_col_trc<TRC_L_MIN + Q * TRC_L_STEP, false>), dim3(Lo.geo[G_TRI + Q].back()),
           dim3(TRC_THREADS), trc_lds_bytes(false), s, c, c.f.T, guard);
}
template <int... Q>
static void tri_final_launches(const Layout &Lo, const Ctx &c, int guard, hipStream_t s, std::integer_sequence<int, Q...>) {
  (tri_final_launch<Q>(Lo, c, guard, s), ...);
}

static void col_pass(const Layout &Lo, const Ctx &c, int guard, hipStream_t s) {
  for (const LaunchClass &C : Lo.slabcls) {
    const int nb1 = Lo.geo[C.geo].back(), nb2 = Lo.geo[C.geo2].back();
    if (nb1 == 0) continue;
    const Ctx k = class_ctx(Lo, c, C);
    const dim3 g1(nb1), g2(nb2);
    const size_t l = C.lds;
    const float *T = c.f.T;
    if (C.f1) {
      const dim3 b1(256), b2(4 * C.f2);
      switch (C.f1) {
        case 16: LAUNCH("k_slab1_f", k_slab1_f<16>, g1, b1, l, s, k, T, guard); break;
        case 32: LAUNCH("k_slab1_f", k_slab1_f<32>, g1, b1, l, s, k, T, guard); break;
        case 64: LAUNCH("k_slab1_f", k_slab1_f<64>, g1, b1, l, s, k, T, guard); break;
        default: LAUNCH("k_slab1_f", k_slab1_f<128>, g1, b1, l, s, k, T, guard); break;
      }
      switch (C.f2) {
        case 16: LAUNCH("k_slab2_f", k_slab2_f<16>, g2, b2, l, s, k, guard); break;
        case 32: LAUNCH("k_slab2_f", k_slab2_f<32>, g2, b2, l, s, k, guard); break;
        default: LAUNCH("k_slab2_f", k_slab2_f<64>, g2, b2, l, s, k, guard); break;
      }
      switch (C.f1) {
        case 16: LAUNCH("k_slab3_f", k_slab3_f<16>, g1, b1, l, s, k, c.f.T, guard); break;
        case 32: LAUNCH("k_slab3_f", k_slab3_f<32>, g1, b1, l, s, k, c.f.T, guard); break;
        case 64: LAUNCH("k_slab3_f", k_slab3_f<64>, g1, b1, l, s, k, c.f.T, guard); break;
        default: LAUNCH("k_slab3_f", k_slab3_f<128>, g1, b1, l, s, k, c.f.T, guard); break;
      }
    } else {
      const dim3 b(C.threads);
      LAUNCH("k_slab1", (k_slab1<SLAB_E>), g1, b, l, s, k, T, guard);
      LAUNCH("k_slab2", (k_slab2<SLAB_E>), g2, b, l, s, k, guard);
      LAUNCH("k_slab3", (k_slab3<SLAB_E>), g1, b, l, s, k, c.f.T, guard);
    }
  }
  if (Lo.geo[G_COLL1].back() > 0) {
    long_pass1(Lo, c, 1, LONG_DCT2, c.f.Tt, nullptr, guard, s);
    long_pass2(Lo, c, 1, LONG_DCT2, c.f.Tt, 0, guard, s);
    long_pass1(Lo, c, 1, LONG_DCT3_SCALE, c.f.Tt, nullptr, guard, s);
    long_pass2(Lo, c, 1, LONG_DCT3, c.f.Tt, 0, guard, s);
  }
  lb_dct2(Lo, c, 1, false, c.f.Tt, c.f.Tt, nullptr, guard, s);
  lb_dct3(Lo, c, 1, c.f.Tt, c.f.Tt, 0, guard, s);
  if (Lo.geo[G_TRIT].back() > 0) {
    LAUNCH("k_col_trc_band", (k_col_trc<TRC_L, true>), dim3(Lo.geo[G_TRIT].back()), dim3(TRC_THREADS),
           trc_lds_bytes(true), s, c,
           c.f.T, guard);
    LAUNCH("k_col_trc_carry", k_col_trc_carry, dim3(Lo.geo[G_TRIC].back()), dim3(256), 0, s, c, guard);
  }
  // one launch per chunk length present in the batch
  tri_final_launches(Lo, c, guard, s, std::make_integer_sequence<int, TRC_NL>{});
  if (Lo.geo[G_TRIDC].back() > 0)
    LAUNCH("k_col_trc_dc", k_col_trc_dc, dim3(Lo.geo[G_TRIDC].back()), dim3(64), 0, s, c, c.f.T, guard);
  for (const LaunchClass &C : Lo.colcls) {
    const int nb = Lo.geo[C.geo].back();
    if (nb == 0) continue;
    const Ctx k = class_ctx(Lo, c, C);
    const dim3 g(nb), b(C.threads);
    if (C.direct) {
      if (C.E == 16) LAUNCH("k_col_direct", (k_col<16, true>), g, b, C.lds, s, k, c.f.T, guard);
      else if (C.E == 12) LAUNCH("k_col_direct", (k_col<12, true>), g, b, C.lds, s, k, c.f.T, guard);
      else LAUNCH("k_col_direct", (k_col<24, true>), g, b, C.lds, s, k, c.f.T, guard);
    } else {
      if (C.E == 16) LAUNCH("k_col", (k_col<16, false>), g, b, C.lds, s, k, c.f.Tt, guard);
      else if (C.E == 12) LAUNCH("k_col", (k_col<12, false>), g, b, C.lds, s, k, c.f.Tt, guard);
      else LAUNCH("k_col", (k_col<24, false>), g, b, C.lds, s, k, c.f.Tt, guard);
    }
  }
}

// T <-> Tt for the images whose columns run on the transposed field (transposed column DCTs and
// long columns).
static void transpose(const Layout &Lo, const Ctx &c, int dir, int guard, hipStream_t s) {
  if (Lo.geo[G_TRANS].back() == 0) return;
  const float *src = dir == 0 ? c.f.T : c.f.Tt;
  float *dst = dir == 0 ? c.f.Tt : c.f.T;
  LAUNCH("k_transpose", k_transpose, dim3(Lo.geo[G_TRANS].back()), dim3(256), 0, s, c, src, dst, dir, guard);
}

// z = M^{-1} src (utils_unwrap2d.py:1390-1403); `dot` adds r.z (DOT_RZ_INIT / DOT_RZ).
static void precond(const Layout &Lo, const Ctx &c, float *src, float *dst, int dot, int guard, hipStream_t s) {
  row_dct2(Lo, c, false, src, c.f.T, nullptr, guard, s);
  transpose(Lo, c, 0, guard, s);
  col_pass(Lo, c, guard, s);
  transpose(Lo, c, 1, guard, s);
  row_dct3(Lo, c, c.f.T, dst, dot, guard, s);
  if (dot == DOT_RZ_INIT) fin(Lo, c, FIN_RZ_INIT, 0, s);
  if (dot == DOT_RZ) fin(Lo, c, FIN_RZ, 0, s);
}

// PCG on A = D^T W D with W from c.f.wxi / c.f.wyi, right-hand side c.f.b, x = c.f.u updated in
// place (utils_unwrap2d.py:1405-1474): r = b - A x, z = M^{-1} r, then cg_max_iter iterations.
// Test hook (irls2d_debug_inject): in PCG iteration g_inject_cg of IRLS iteration g_inject_irls,
// p(0, 0) of every active image becomes +inf after alpha is formed, so that the update x += alpha p
// is non-finite.  -1 = off; a solve then runs eagerly even when a graph was asked for.
static int g_inject_irls = -1, g_inject_cg = -1;

static void pcg(const Layout &Lo, const Ctx &c, int cg_max_iter, int it, hipStream_t s, int inject_cg = -1) {
  const int gt = Lo.geo[G_TILE].back();
  const dim3 tb(TILE_THREADS);
  LAUNCH("k_cg_init", k_cg_init, dim3(gt), tb, 0, s, c);
  fin(Lo, c, FIN_CG_INIT, it, s);
  precond(Lo, c, c.f.r, c.f.z, DOT_RZ_INIT, GUARD_CG, s);
  for (int i = 0; i < cg_max_iter; ++i) {
    float *pold = (i & 1) ? c.f.p1 : c.f.p0;
    float *pnew = (i & 1) ? c.f.p0 : c.f.p1;
    LAUNCH("k_cg_pap", k_cg_pap, dim3(gt), tb, 0, s, c, (const float *)pold, pnew);
    fin(Lo, c, FIN_PAP, it, s);
    if (i == inject_cg) LAUNCH("k_inject", k_inject, dim3(Lo.n), dim3(64), 0, s, c, pnew);
    row_dct2(Lo, c, true, c.f.r, c.f.T, pnew, GUARD_CG, s);
    fin(Lo, c, FIN_UPDATE, it, s);
    transpose(Lo, c, 0, GUARD_CG, s);
    col_pass(Lo, c, GUARD_CG, s);
    transpose(Lo, c, 1, GUARD_CG, s);
    row_dct3(Lo, c, c.f.T, c.f.z, DOT_RZ, GUARD_CG, s);
    fin(Lo, c, FIN_RZ, it, s);
  }
  // the x updates still pending after the last step (xdefer images)
  if (cg_max_iter > 0) LAUNCH("k_cg_xcommit", k_cg_xcommit, dim3(gt), tb, 0, s, c);
}

static void enqueue(const Layout &Lo, const Ctx &c, const float *phase, const float *weight, int32_t *info,
                    const Params &p, hipStream_t s) {
  const int gt = Lo.geo[G_TILE].back();
  const dim3 tb(TILE_THREADS);
  LAUNCH("k_state_init", k_state_init, dim3(1), dim3(64), 0, s, c);
  LAUNCH("k_setup", k_setup, dim3(gt), tb, 0, s, c, phase, weight);
  fin(Lo, c, FIN_VCOUNT, 0, s);
  LAUNCH("k_rhs_static", k_rhs_static, dim3(gt), tb, 0, s, c, c.f.r);
  precond(Lo, c, c.f.r, c.f.u, DOT_NONE, GUARD_NONE, s);
  LAUNCH("k_center_reduce", k_center_reduce, dim3(gt), tb, 0, s, c);
  fin(Lo, c, FIN_CENTER, 0, s);
  LAUNCH("k_center_apply", k_center_apply, dim3(gt), tb, 0, s, c);
  for (int it = 0; it < p.max_iter; ++it) {
    LAUNCH("k_irls_mean", k_irls_mean, dim3(gt), tb, 0, s, c);
    fin(Lo, c, FIN_IRLS_MEAN, it, s);
    LAUNCH("k_irls_update", k_irls_update, dim3(gt), tb, 0, s, c);
    fin(Lo, c, FIN_IRLS_UPDATE, it, s);
    LAUNCH("k_irls_commit", k_irls_commit, dim3(gt), tb, 0, s, c, it);
    pcg(Lo, c, p.cg_max_iter, it, s, it == g_inject_irls ? g_inject_cg : -1);
    LAUNCH("k_irls_check", k_irls_check, dim3(gt), tb, 0, s, c);
    fin(Lo, c, FIN_IRLS_CHECK, it, s);
    LAUNCH("k_restore", k_restore, dim3(gt), tb, 0, s, c, it);
  }
  if (info) LAUNCH("k_info", k_info, dim3(1), dim3(64), 0, s, c, info);
}

// ---------------------------------------------------------------------------------------------
// Caches: layouts per shape signature, graphs per call signature.  The metadata region of a
// workspace (descriptors, plans, tables) is uploaded on every call: the library never assumes
// that a workspace still holds what an earlier call wrote, because a freed workspace's address can
// come back for another one.
// ---------------------------------------------------------------------------------------------

static std::mutex g_mu;
static std::map<uint64_t, std::shared_ptr<Layout>> g_layouts;

struct GraphKey {
  void *ws;
  const float *phase, *weight;
  float *out;
  int32_t *info;
  uint64_t sig;
  int max_iter, cg_max_iter;
  double tol, cg_tol, eps;
  bool operator<(const GraphKey &o) const {
    return std::tie(ws, phase, weight, out, info, sig, max_iter, cg_max_iter, tol, cg_tol, eps) <
           std::tie(o.ws, o.phase, o.weight, o.out, o.info, o.sig, o.max_iter, o.cg_max_iter, o.tol, o.cg_tol,
                    o.eps);
  }
};
// Instantiated graphs, least recently used first out beyond MAX_GRAPHS (each holds a few thousand
// kernel nodes; callers that replay with new arena addresses every call would otherwise grow it).
constexpr size_t MAX_GRAPHS = 8;
struct GraphEntry {
  hipGraphExec_t exec;
  uint64_t last_use;
  int dev;
};
static std::map<GraphKey, GraphEntry> g_graphs;
static uint64_t g_graph_clock = 0;
static std::map<int, hipStream_t> g_capture_stream;  // per device

// Pinned staging of the metadata upload, two slots per device used in turn: a slot is refilled
// only after the copy that last read it has completed (its event), so the host can run one call
// ahead of the device without a stream synchronisation per call.
struct Staging {
  void *buf[2] = {nullptr, nullptr};
  size_t cap[2] = {0, 0};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool pending[2] = {false, false};
  int next = 0;
};
static std::map<int, Staging> g_staging;

static int get_layout(int n, const int64_t *H, const int64_t *W, std::shared_ptr<Layout> &out) {
  const uint64_t sig = shape_sig(n, H, W);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_layouts.find(sig);
    if (it != g_layouts.end() && it->second->n == n &&
        std::equal(H, H + n, it->second->H.begin()) && std::equal(W, W + n, it->second->W.begin())) {
      out = it->second;
      return IRLS2D_OK;
    }
  }
  auto Lo = std::make_shared<Layout>();
  int rc = build_layout(n, H, W, *Lo);
  if (rc) return rc;
  Lo->meta = pack_meta(*Lo);
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_layouts.size() > 64) g_layouts.clear();
  g_layouts[sig] = Lo;
  out = Lo;
  return IRLS2D_OK;
}

static bool g_attr_done[64];

// every k_col_trc instantiation: the final launches of each chunk length, then the band pass
template <int... Q>
static std::vector<const void *> tri_kernels(std::integer_sequence<int, Q...>) {
  return {(const void *)k_col_trc<TRC_L_MIN + Q * TRC_L_STEP, false>..., (const void *)k_col_trc<TRC_L, true>};
}

static int set_attrs() {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev >= 0 && dev < 64 && g_attr_done[dev]) return IRLS2D_OK;
  const int lds = (TW_LDS + lds_seq_len(MAX_FFT_LEN)) * (int)sizeof(float2);  // largest transform
  const void *fns[] = {(const void *)k_row_dct2<false, 16>, (const void *)k_row_dct2<false, 24>,
                       (const void *)k_row_dct2<true, 16>,  (const void *)k_row_dct2<true, 24>,
                       (const void *)k_col<16, false>,      (const void *)k_col<24, false>,
                       (const void *)k_col<16, true>,       (const void *)k_col<24, true>,
                       (const void *)k_row_dct3<16>,        (const void *)k_row_dct3<24>,
                       (const void *)k_row_dct2<false, 12>, (const void *)k_row_dct2<true, 12>,
                       (const void *)k_col<12, false>,      (const void *)k_col<12, true>,
                       (const void *)k_row_dct3<12>};
  for (const void *f : fns) HIP_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  const int llds = LONG_LDS * (int)sizeof(float2);
  const void *lfns[] = {(const void *)k_long_p1<LONG_DCT2, LONG_E>, (const void *)k_long_p1<LONG_DCT2_UPDATE, LONG_E>,
                        (const void *)k_long_p1<LONG_DCT3, LONG_E>, (const void *)k_long_p1<LONG_DCT3_SCALE, LONG_E>,
                        (const void *)k_long_p2<LONG_DCT2, LONG_E>, (const void *)k_long_p2<LONG_DCT3, LONG_E>,
                        (const void *)k_lb_fft13<1, LONG_E>, (const void *)k_lb_fft13<3, LONG_E>,
                        (const void *)k_lb_fft2<LONG_E>};
  for (const void *f : lfns) HIP_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, llds));
  const int slds = (TW_LDS + SLAB_E * SLAB_MAX_THREADS + 2 * SLAB_MAX_THREADS) * (int)sizeof(float2);
  const void *sfns[] = {(const void *)k_slab1<SLAB_E>, (const void *)k_slab2<SLAB_E>, (const void *)k_slab3<SLAB_E>};
  for (const void *f : sfns) HIP_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, slds));
  for (const void *f : tri_kernels(std::make_integer_sequence<int, TRC_NL>{}))
    HIP_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)trc_lds_bytes(true)));
  if (dev >= 0 && dev < 64) g_attr_done[dev] = true;
  return IRLS2D_OK;
}

// Uploads the plans, tables and descriptors of a layout into the workspace's metadata region,
// stream-ordered on `s` (through the pinned staging slots of the current device).
static int upload_meta(const Layout &Lo, void *workspace, hipStream_t s) {
  if (Lo.meta.empty()) return IRLS2D_OK;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  Staging &S = g_staging[dev];
  const int k = S.next;
  S.next ^= 1;
  if (!S.ev[k]) HIP_TRY(hipEventCreateWithFlags(&S.ev[k], hipEventDisableTiming));
  if (S.pending[k]) {
    HIP_TRY(hipEventSynchronize(S.ev[k]));  // the copy that last read this slot has finished
    S.pending[k] = false;
  }
  if (S.cap[k] < Lo.meta.size()) {
    if (S.buf[k]) HIP_TRY(hipHostFree(S.buf[k]));
    S.buf[k] = nullptr;
    S.cap[k] = 0;
    const size_t cap = align_up(Lo.meta.size(), (size_t)1 << 20);
    HIP_TRY(hipHostMalloc(&S.buf[k], cap, hipHostMallocDefault));
    S.cap[k] = cap;
  }
  std::memcpy(S.buf[k], Lo.meta.data(), Lo.meta.size());
  HIP_TRY(hipMemcpyAsync((unsigned char *)workspace + Lo.o_meta, S.buf[k], Lo.meta.size(), hipMemcpyHostToDevice, s));
  HIP_TRY(hipEventRecord(S.ev[k], s));
  S.pending[k] = true;
  return IRLS2D_OK;
}

int solve(int n, const int64_t *H, const int64_t *W, const float *phase, const float *weight, float *out,
          int max_iter, double tol, int cg_max_iter, double cg_tol, double epsilon, void *workspace,
          size_t ws_bytes, int32_t *info, void *stream, unsigned flags) {
  g_err.clear();
  if (!phase || !out || !workspace) return fail(IRLS2D_EINVAL, "null phase, output or workspace pointer");
  if (max_iter < 0 || cg_max_iter < 0) return fail(IRLS2D_EINVAL, "iteration counts must be >= 0");
  if (!(epsilon >= 0) || !std::isfinite(epsilon)) return fail(IRLS2D_EINVAL, "epsilon must be finite and >= 0");
  if (std::isnan(tol) || std::isnan(cg_tol)) return fail(IRLS2D_EINVAL, "tol and cg_tol must not be NaN");
  std::shared_ptr<Layout> Lo;
  int rc = get_layout(n, H, W, Lo);
  if (rc) return rc;
  if (ws_bytes < Lo->bytes)
    return fail(IRLS2D_EWORKSPACE, "workspace has " + std::to_string(ws_bytes) + " bytes, needs " +
                                       std::to_string(Lo->bytes));
  rc = set_attrs();
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  unsigned char *ws = (unsigned char *)workspace;
  rc = upload_meta(*Lo, workspace, s);
  if (rc) return rc;
  Params p{max_iter, cg_max_iter, tol, cg_tol, epsilon};
  Ctx c = make_ctx(*Lo, ws, out, p);
  if (!(flags & IRLS2D_FLAG_GRAPH) || g_inject_irls >= 0) {
    KTimer timer;
    if (flags & IRLS2D_FLAG_TIME_KERNELS) t_timer = &timer;
    enqueue(*Lo, c, phase, weight, info, p, s);
    t_timer = nullptr;
    HIP_TRY(hipGetLastError());
    if (flags & IRLS2D_FLAG_TIME_KERNELS) return collect_timer(timer, s);
    return IRLS2D_OK;
  }
  // Graph path: capture once on a private stream, replay with event ordering against `s`.
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipStream_t cs;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_capture_stream.find(dev);
    if (it == g_capture_stream.end()) {
      hipStream_t ns;
      HIP_TRY(hipStreamCreateWithFlags(&ns, hipStreamNonBlocking));
      g_capture_stream[dev] = ns;
      cs = ns;
    } else {
      cs = it->second;
    }
  }
  GraphKey key{workspace, phase, weight, out, info, Lo->sig, max_iter, cg_max_iter, tol, cg_tol, epsilon};
  hipGraphExec_t exec = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_graphs.find(key);
    if (it != g_graphs.end()) {
      exec = it->second.exec;
      it->second.last_use = ++g_graph_clock;
    }
  }
  if (!exec) {
    hipGraph_t graph;
    HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    enqueue(*Lo, c, phase, weight, info, p, cs);
    hipError_t le = hipGetLastError();
    HIP_TRY(hipStreamEndCapture(cs, &graph));
    if (le != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return fail(IRLS2D_EHIP, std::string("launch during capture: ") + hipGetErrorString(le));
    }
    HIP_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    HIP_TRY(hipGraphDestroy(graph));
    std::lock_guard<std::mutex> lk(g_mu);
    while (g_graphs.size() >= MAX_GRAPHS) {
      auto lru = g_graphs.begin();
      for (auto it = g_graphs.begin(); it != g_graphs.end(); ++it)
        if (it->second.last_use < lru->second.last_use) lru = it;
      // an evicted graph may still run on its capture stream: let it finish first
      auto cs_it = g_capture_stream.find(lru->second.dev);
      if (cs_it != g_capture_stream.end()) HIP_TRY(hipStreamSynchronize(cs_it->second));
      (void)hipGraphExecDestroy(lru->second.exec);
      g_graphs.erase(lru);
    }
    g_graphs[key] = GraphEntry{exec, ++g_graph_clock, dev};
  }
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreateWithFlags(&e0, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&e1, hipEventDisableTiming));
  HIP_TRY(hipEventRecord(e0, s));
  HIP_TRY(hipStreamWaitEvent(cs, e0, 0));
  HIP_TRY(hipGraphLaunch(exec, cs));
  HIP_TRY(hipEventRecord(e1, cs));
  HIP_TRY(hipStreamWaitEvent(s, e1, 0));
  HIP_TRY(hipEventDestroy(e0));
  HIP_TRY(hipEventDestroy(e1));
  return IRLS2D_OK;
}

// ---------------------------------------------------------------------------------------------
// Solve without a caller workspace (irls2d_solve_static): the workspace is the library's static
// device array `scratch`, one per GPU the library runs on, held until the process ends.  Host code
// never passes the array's name to the device: the address comes from hipGetSymbolAddress, per
// device, on that device's first call.
// ---------------------------------------------------------------------------------------------

static __device__ __align__(256) unsigned char scratch[IRLS2D_STATIC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static std::map<int, void *> g_scratch_ptr;  // device -> address of `scratch` on that device

static int scratch_address(void **out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(IRLS2D_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    it = g_scratch_ptr.emplace(dev, p).first;
  }
  *out = it->second;
  return IRLS2D_OK;
}

int solve_static(int n, const int64_t *H, const int64_t *W, const float *phase, const float *weight, float *out,
                 int max_iter, double tol, int cg_max_iter, double cg_tol, double epsilon, int32_t *info,
                 void *stream) {
  g_err.clear();
  if (!phase || !out) return fail(IRLS2D_EINVAL, "null phase or output pointer");
  if (max_iter < 0 || cg_max_iter < 0) return fail(IRLS2D_EINVAL, "iteration counts must be >= 0");
  if (!(epsilon >= 0) || !std::isfinite(epsilon)) return fail(IRLS2D_EINVAL, "epsilon must be finite and >= 0");
  if (std::isnan(tol) || std::isnan(cg_tol)) return fail(IRLS2D_EINVAL, "tol and cg_tol must not be NaN");
  std::shared_ptr<Layout> Lo;
  int rc = get_layout(n, H, W, Lo);
  if (rc) return rc;
  if (Lo->bytes > (size_t)IRLS2D_STATIC_SCRATCH_BYTES)
    return fail(IRLS2D_EWORKSPACE, "these shapes need a workspace of " + std::to_string(Lo->bytes) +
                                       " bytes; the static scratch holds " +
                                       std::to_string((size_t)IRLS2D_STATIC_SCRATCH_BYTES));
  void *ws = nullptr;
  rc = scratch_address(&ws);
  if (rc) return rc;
  rc = set_attrs();
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  // metadata straight from the layout's host image (pageable; the stream is synchronised after
  // the copy, so nothing is staged in memory the library would have to allocate)
  if (!Lo->meta.empty()) {
    HIP_TRY(hipMemcpyAsync((unsigned char *)ws + Lo->o_meta, Lo->meta.data(), Lo->meta.size(), hipMemcpyHostToDevice,
                           s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  Params p{max_iter, cg_max_iter, tol, cg_tol, epsilon};
  Ctx c = make_ctx(*Lo, (unsigned char *)ws, out, p);
  enqueue(*Lo, c, phase, weight, info, p, s);
  HIP_TRY(hipGetLastError());
  return IRLS2D_OK;
}

static void enqueue_adjoint(const Layout &Lo, const Ctx &c, const float *phase, const float *weight,
                            const float *u, const float *g, float *gphase, float *gweight, int32_t *info,
                            int cg_max_iter, hipStream_t s) {
  const int gt = Lo.geo[G_TILE].back();
  const dim3 tb(TILE_THREADS);
  LAUNCH("k_state_init", k_state_init, dim3(1), dim3(64), 0, s, c);
  LAUNCH("k_setup", k_setup, dim3(gt), tb, 0, s, c, phase, weight);
  fin(Lo, c, FIN_VCOUNT, 0, s);
  LAUNCH("k_adj_gsum", k_adj_gsum, dim3(gt), tb, 0, s, c, g);
  fin(Lo, c, FIN_GSUM, 0, s);
  LAUNCH("k_adj_prep", k_adj_prep, dim3(gt), tb, 0, s, c, u, g, gphase);
  pcg(Lo, c, cg_max_iter, 0, s);  // lambda accumulates in gphase (c.f.u)
  LAUNCH("k_adj_grads", k_adj_grads, dim3(gt), tb, 0, s, c, (const float *)gphase, weight, c.f.T, gweight);
  LAUNCH("k_copy", k_copy, dim3(gt), tb, 0, s, c, (const float *)c.f.T, gphase);
  if (info) LAUNCH("k_info", k_info, dim3(1), dim3(64), 0, s, c, info);
}

int adjoint(int n, const int64_t *H, const int64_t *W, const float *phase, const float *weight, const float *u,
            const float *g, float *gphase, float *gweight, double epsilon, int cg_max_iter, double cg_tol,
            void *workspace, size_t ws_bytes, int32_t *info, void *stream, unsigned flags) {
  g_err.clear();
  if (!phase || !u || !g || !gphase || !workspace) return fail(IRLS2D_EINVAL, "null pointer argument");
  if (cg_max_iter < 0) return fail(IRLS2D_EINVAL, "cg_max_iter must be >= 0");
  if (!(epsilon >= 0) || !std::isfinite(epsilon)) return fail(IRLS2D_EINVAL, "epsilon must be finite and >= 0");
  std::shared_ptr<Layout> Lo;
  int rc = get_layout(n, H, W, Lo);
  if (rc) return rc;
  if (ws_bytes < Lo->bytes) return fail(IRLS2D_EWORKSPACE, "workspace too small");
  rc = set_attrs();
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  unsigned char *ws = (unsigned char *)workspace;
  rc = upload_meta(*Lo, workspace, s);
  if (rc) return rc;
  Params p{0, cg_max_iter, 0.0, cg_tol, epsilon};
  Ctx c = make_ctx(*Lo, ws, gphase, p);
  KTimer timer;
  if (flags & IRLS2D_FLAG_TIME_KERNELS) t_timer = &timer;
  enqueue_adjoint(*Lo, c, phase, weight, u, g, gphase, gweight, info, cg_max_iter, s);
  t_timer = nullptr;
  HIP_TRY(hipGetLastError());
  if (flags & IRLS2D_FLAG_TIME_KERNELS) return collect_timer(timer, s);
  return IRLS2D_OK;
}

int precondition(int n, const int64_t *H, const int64_t *W, const float *r, float *z, void *workspace,
                 size_t ws_bytes, void *stream) {
  g_err.clear();
  if (!r || !z || !workspace) return fail(IRLS2D_EINVAL, "null pointer argument");
  if (r == z) return fail(IRLS2D_EINVAL, "r and z must not alias");
  std::shared_ptr<Layout> Lo;
  int rc = get_layout(n, H, W, Lo);
  if (rc) return rc;
  if (ws_bytes < Lo->bytes) return fail(IRLS2D_EWORKSPACE, "workspace too small");
  rc = set_attrs();
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  rc = upload_meta(*Lo, workspace, s);
  if (rc) return rc;
  Params p{0, 0, 0.0, 0.0, 0.0};
  Ctx c = make_ctx(*Lo, (unsigned char *)workspace, z, p);
  // the row DCT-II only reads its source
  precond(*Lo, c, const_cast<float *>(r), z, DOT_NONE, GUARD_NONE, s);
  HIP_TRY(hipGetLastError());
  return IRLS2D_OK;
}

// Drops every cached layout, uploaded table set and graph (plans changed).
static void drop_caches() {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &g : g_graphs) {
    auto cs = g_capture_stream.find(g.second.dev);
    if (cs != g_capture_stream.end()) (void)hipStreamSynchronize(cs->second);
    (void)hipGraphExecDestroy(g.second.exec);
  }
  g_graphs.clear();
  g_layouts.clear();
}

static bool valid_radix(int r) { return r == 2 || r == 3 || r == 4 || r == 5 || r == 8 || r == 16; }

int plan_override(int64_t N, int64_t fft_len, const int *rad, int nrad, int ept) {
  g_err.clear();
  if (N < 1 || N > MAX_FFT_LEN) return fail(IRLS2D_EINVAL, "plan overrides apply to lengths 1 .. 12288");
  if (fft_len == 0) {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    g_over.erase((int)N);
    g_plans.erase((int)N);
  } else {
    if (fft_len < N || fft_len > MAX_FFT_LEN || !smooth235(fft_len))
      return fail(IRLS2D_EINVAL, "fft_len must be 5-smooth, >= N and <= 12288");
    if (fft_len != N && fft_len < 2 * N - 1) return fail(IRLS2D_EINVAL, "a Bluestein length must be >= 2N - 1");
    if (fft_len == N && !smooth235(N)) return fail(IRLS2D_EINVAL, "N has a prime factor > 5: needs Bluestein");
    if (ept != 0 && ept != 12 && ept != 16 && ept != 24) return fail(IRLS2D_EINVAL, "ept must be 0, 12, 16 or 24");
    if (ept == 16 && fft_len > 16 * MAX_FFT_THREADS) return fail(IRLS2D_EINVAL, "ept 16 covers lengths <= 8192");
    PlanOverride o;
    o.fft_len = (int)fft_len;
    o.ept = ept;
    if (nrad > 0) {
      if (!rad || nrad > MAX_PASS) return fail(IRLS2D_EINVAL, "bad radix list");
      int64_t prod = 1;
      for (int i = 0; i < nrad; ++i) {
        if (!valid_radix(rad[i])) return fail(IRLS2D_EINVAL, "radices must be 2, 3, 4, 5, 8 or 16");
        prod *= rad[i];
      }
      if (prod != fft_len) return fail(IRLS2D_EINVAL, "the radices must multiply to fft_len");
      o.rad.assign(rad, rad + nrad);
    } else if (radices((int)fft_len).size() > MAX_PASS) {
      return fail(IRLS2D_EINVAL, "too many passes");
    }
    std::lock_guard<std::mutex> lk(g_plan_mu);
    g_over[(int)N] = o;
    g_plans.erase((int)N);
  }
  drop_caches();
  return IRLS2D_OK;
}

int plan_describe(int64_t N, int64_t *fft_len, int *rad, int *nrad, int *ept) {
  g_err.clear();
  if (N < 1 || N > MAX_FFT_LEN || !supported(N)) return fail(IRLS2D_EINVAL, "no LDS-resident plan for this length");
  auto h = get_plan((int)N);
  if (fft_len) *fft_len = h->plan.M;
  if (ept) *ept = h->plan.ept;
  if (nrad) *nrad = h->plan.fft.npass;
  if (rad)
    for (int i = 0; i < h->plan.fft.npass; ++i) rad[i] = h->plan.fft.pass[i].R;
  return IRLS2D_OK;
}

int set_column_mode(int mode) {
  g_err.clear();
  if (mode < 0 || mode > 4)
    return fail(IRLS2D_EINVAL,
                "column mode must be 0 (auto), 1 (transpose), 2 (direct), 3 (slab) or 4 (tridiagonal)");
  {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    g_col_mode = mode;
  }
  drop_caches();
  return IRLS2D_OK;
}

int plan_reset() {
  {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    g_over.clear();
    g_plans.clear();
  }
  drop_caches();
  return IRLS2D_OK;
}

// ---------------------------------------------------------------------------------------------
// Goldstein filter (utils_goldstein.py): patch filter + overlap-add
// ---------------------------------------------------------------------------------------------

static bool gold_psize_ok(int p) { return p == 16 || p == 32 || p == 64; }

// Patch grid per axis: floor((n - s) / s) + 1 patches for mode 1 (goldstein_numpy,
// utils_goldstein.py:53-57, :85-86), ceil(...) + 1 for mode 0 (goldstein_pytorch, :224-230).
static int64_t gold_npatch(int64_t n, int p, int mode) {
  const int64_t s = p / 2;
  return (mode == 1 ? (n - s) / s : (n - s + s - 1) / s) + 1;
}

static int gold_check(int64_t ny, int64_t nx, int py, int px) {
  g_err.clear();
  if (!gold_psize_ok(py) || !gold_psize_ok(px))
    return fail(IRLS2D_EUNSUPPORTED, "goldstein patch sizes must be 16, 32 or 64");
  if (ny < py / 2 || nx < px / 2) return fail(IRLS2D_EINVAL, "goldstein: image smaller than half a patch");
  if (ny * nx >= ((int64_t)1 << 40)) return fail(IRLS2D_EINVAL, "goldstein: image too large");
  return IRLS2D_OK;
}

static size_t gold_ws_bytes(int64_t ny, int64_t nx, int py, int px) {
  const int64_t np = gold_npatch(ny, py, 0) * gold_npatch(nx, px, 0);  // the larger grid
  return align_up((size_t)np * py * px * sizeof(float2), 256) + align_up((size_t)np * sizeof(int), 256);
}

int goldstein_workspace_size(int64_t ny, int64_t nx, int py, int px, size_t *bytes) {
  int rc = gold_check(ny, nx, py, px);
  if (rc) return rc;
  if (!bytes) return fail(IRLS2D_EINVAL, "null bytes pointer");
  *bytes = gold_ws_bytes(ny, nx, py, px);
  return IRLS2D_OK;
}

template <int PY, int PX>
static int gold_launch(const GoldArgs &g, hipStream_t s) {
  constexpr int PM = PY > PX ? PY : PX;
  constexpr int NP = GOLD_THREADS / PM;
  const size_t lds = (size_t)NP * PY * (PX + 1) * sizeof(float2);
  HIP_TRY(hipFuncSetAttribute((const void *)k_gold_patch<PY, PX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds));
  const long long np = (long long)g.npy * g.npx;
  LAUNCH("k_gold_patch", (k_gold_patch<PY, PX>), dim3((unsigned)((np + NP - 1) / NP)), dim3(GOLD_THREADS), lds, s, g);
  const long long n = g.ny * g.nx;
  const unsigned nb = (unsigned)std::min<long long>((n + GOLD_THREADS - 1) / GOLD_THREADS, 1 << 20);
  LAUNCH("k_gold_fold", (k_gold_fold<PY, PX>), dim3(nb), dim3(GOLD_THREADS), 0, s, g);
  HIP_TRY(hipGetLastError());
  return IRLS2D_OK;
}

int goldstein(int64_t ny, int64_t nx, const void *phase, const float *corr, void *out, int py, int px,
              double threshold, int mode, void *workspace, size_t ws_bytes, void *stream) {
  int rc = gold_check(ny, nx, py, px);
  if (rc) return rc;
  if (!phase || !corr || !out || !workspace) return fail(IRLS2D_EINVAL, "null pointer argument");
  if (mode != 0 && mode != 1) return fail(IRLS2D_EINVAL, "goldstein mode must be 0 (count) or 1 (weight)");
  if (!(threshold >= 0.0) || threshold > 1.0) return fail(IRLS2D_EINVAL, "goldstein threshold must be in [0, 1]");
  if (ws_bytes < gold_ws_bytes(ny, nx, py, px)) return fail(IRLS2D_EWORKSPACE, "goldstein workspace too small");
  GoldArgs g;
  g.phase = (const float2 *)phase;
  g.corr = corr;
  g.out = (float2 *)out;
  const int64_t npy = gold_npatch(ny, py, mode), npx = gold_npatch(nx, px, mode);
  g.patches = (float2 *)workspace;
  g.pvalid = (int *)((unsigned char *)workspace +
                     align_up((size_t)gold_npatch(ny, py, 0) * gold_npatch(nx, px, 0) * py * px * sizeof(float2), 256));
  g.ny = ny;
  g.nx = nx;
  g.npy = (int)npy;
  g.npx = (int)npx;
  g.min_valid = (int)((double)py * (double)px * threshold);  // Python int(): truncation
  g.mode = mode;
  double ws = 0.0;
  for (int r = 0; r < py; ++r)
    for (int c = 0; c < px; ++c) {
      auto w1 = [](int i, int p) {
        const int h = p / 2, j = i < h ? i : p - 1 - i;
        return 1.0 - std::fabs((double)j - (double)(h - 1)) / (double)(h - 1);
      };
      ws += (double)(float)(w1(r, py) * w1(c, px));
    }
  g.wsum = ws;
  hipStream_t s = (hipStream_t)stream;
#define IRLS2D_GOLD_CASE(PY, PX) \
  if (py == PY && px == PX) return gold_launch<PY, PX>(g, s);
  IRLS2D_GOLD_CASE(16, 16)
  IRLS2D_GOLD_CASE(16, 32)
  IRLS2D_GOLD_CASE(16, 64)
  IRLS2D_GOLD_CASE(32, 16)
  IRLS2D_GOLD_CASE(32, 32)
  IRLS2D_GOLD_CASE(32, 64)
  IRLS2D_GOLD_CASE(64, 16)
  IRLS2D_GOLD_CASE(64, 32)
  IRLS2D_GOLD_CASE(64, 64)
  return fail(IRLS2D_EUNSUPPORTED, "goldstein patch size");
}

// ---------------------------------------------------------------------------------------------
// NaN-aware Gaussian filter (utils_gaussian.py nanconvolve2d_gaussian_pytorch)
// ---------------------------------------------------------------------------------------------

static int gauss_check(int64_t ny, int64_t nx, int channels, double sy, double sx, double truncate) {
  g_err.clear();
  if (ny < 1 || nx < 1 || ny * nx >= ((int64_t)1 << 40)) return fail(IRLS2D_EINVAL, "gaussian: bad image shape");
  if (channels != 1 && channels != 2) return fail(IRLS2D_EINVAL, "gaussian: channels must be 1 (real) or 2 (complex)");
  if (!(sy > 0) || !(sx > 0) || !(truncate > 0) || !std::isfinite(sy * truncate) || !std::isfinite(sx * truncate))
    return fail(IRLS2D_EINVAL, "gaussian: sigma and truncate must be positive and finite");
  if (sy * truncate > (1 << 20) || sx * truncate > (1 << 20)) return fail(IRLS2D_EINVAL, "gaussian: kernel too long");
  return IRLS2D_OK;
}

// kernel radius: size 2 ceil(s truncate) + 1 (utils_gaussian.py:51-52)
static int gauss_radius(double s, double truncate) { return (int)std::ceil(s * truncate); }

static size_t gauss_ws_bytes(int64_t ny, int64_t nx, int channels, int ry, int rx) {
  return align_up((size_t)ny * nx * 2 * channels * sizeof(float), 256) +
         align_up((size_t)(2 * ry + 1) * sizeof(float), 256) + align_up((size_t)(2 * rx + 1) * sizeof(float), 256);
}

int gaussian_workspace_size(int64_t ny, int64_t nx, int channels, double sy, double sx, double truncate,
                            size_t *bytes) {
  int rc = gauss_check(ny, nx, channels, sy, sx, truncate);
  if (rc) return rc;
  if (!bytes) return fail(IRLS2D_EINVAL, "null bytes pointer");
  *bytes = gauss_ws_bytes(ny, nx, channels, gauss_radius(sy, truncate), gauss_radius(sx, truncate));
  return IRLS2D_OK;
}

int gaussian(int64_t ny, int64_t nx, int channels, const float *data, const float *weight, float *out, double sy,
             double sx, double truncate, double threshold, void *workspace, size_t ws_bytes, void *stream) {
  int rc = gauss_check(ny, nx, channels, sy, sx, truncate);
  if (rc) return rc;
  if (!data || !out || !workspace) return fail(IRLS2D_EINVAL, "null pointer argument");
  const int ry = gauss_radius(sy, truncate), rx = gauss_radius(sx, truncate);
  if (ws_bytes < gauss_ws_bytes(ny, nx, channels, ry, rx)) return fail(IRLS2D_EWORKSPACE, "gaussian workspace too small");
  GaussArgs g;
  g.data = data;
  g.weight = weight;
  g.out = out;
  unsigned char *w = (unsigned char *)workspace;
  g.tmp = (float *)w;
  w += align_up((size_t)ny * nx * 2 * channels * sizeof(float), 256);
  g.ky = (float *)w;
  w += align_up((size_t)(2 * ry + 1) * sizeof(float), 256);
  g.kx = (float *)w;
  g.ny = ny;
  g.nx = nx;
  g.ry = ry;
  g.rx = rx;
  g.sy = (float)sy;
  g.sx = (float)sx;
  // the comparison of the float32 denominator with the Python threshold happens in float32
  g.lim = weight ? 1e-10f : (float)threshold;
  hipStream_t s = (hipStream_t)stream;
  LAUNCH("k_gauss_taps", k_gauss_taps, dim3(1), dim3(GAUSS_THREADS), 0, s, g);
  if (ry <= 8 && rx <= 64) {
    // one pass (no scratch): strips of GAUSS_THREADS - 2 rx output columns, bands of GAUSS_TY rows
    const long long nstrip = (nx + GAUSS_THREADS - 2 * rx - 1) / (GAUSS_THREADS - 2 * rx);
    const long long nband = (ny + GAUSS_TY - 1) / GAUSS_TY;
    if (nstrip * nband >= ((long long)1 << 31)) return fail(IRLS2D_EINVAL, "gaussian: image too large");
    const dim3 gf((unsigned)(nstrip * nband));
    if (channels == 1) LAUNCH("k_gauss_fused", (k_gauss_fused<1, 8>), gf, dim3(GAUSS_THREADS), 0, s, g);
    else LAUNCH("k_gauss_fused", (k_gauss_fused<2, 8>), gf, dim3(GAUSS_THREADS), 0, s, g);
    HIP_TRY(hipGetLastError());
    return IRLS2D_OK;
  }
  const long long n = ny * nx;
  const dim3 grid((unsigned)std::min<long long>((n + GAUSS_THREADS - 1) / GAUSS_THREADS, 1 << 20));
  if (channels == 1) {
    LAUNCH("k_gauss_v", k_gauss_v<1>, grid, dim3(GAUSS_THREADS), 0, s, g);
    LAUNCH("k_gauss_h", k_gauss_h<1>, grid, dim3(GAUSS_THREADS), 0, s, g);
  } else {
    LAUNCH("k_gauss_v", k_gauss_v<2>, grid, dim3(GAUSS_THREADS), 0, s, g);
    LAUNCH("k_gauss_h", k_gauss_h<2>, grid, dim3(GAUSS_THREADS), 0, s, g);
  }
  HIP_TRY(hipGetLastError());
  return IRLS2D_OK;
}

// ---------------------------------------------------------------------------------------------
// Device epilogue steps (one image, pitched view; the host takes the median and size order)
// ---------------------------------------------------------------------------------------------

static int ep_check(int64_t H, int64_t W, int64_t P) {
  g_err.clear();
  if (H < 1 || W < 1 || P < W) return fail(IRLS2D_EINVAL, "bad image shape or pitch");
  if (H * P >= ((int64_t)1 << 31)) return fail(IRLS2D_EUNSUPPORTED, "image too large for the device epilogue");
  return IRLS2D_OK;
}

static dim3 ep_grid_stride(int64_t n) {
  return dim3((unsigned)std::min<int64_t>((n + EP_THREADS - 1) / EP_THREADS, 256 * 64));
}

int64_t ep_blocks(int64_t H, int64_t P) { return (H * P + EP_BLOCK_PX - 1) / EP_BLOCK_PX; }

int ep_scan(int64_t H, int64_t W, int64_t P, const float *psi, const float *u, int32_t *stats, hipStream_t s) {
  int rc = ep_check(H, W, P);
  if (rc) return rc;
  hipLaunchKernelGGL(k_ep_scan, ep_grid_stride(H * P), dim3(EP_THREADS), 0, s, (int)H, (int)W, (int)P, psi, u, stats);
  HIP_TRY(hipGetLastError());
  return IRLS2D_OK;
}

int ep_hist(int64_t H, int64_t W, int64_t P, const float *psi, const float *u, int kmin, int nb, int32_t *hist,
            hipStream_t s) {
  int rc = ep_check(H, W, P);
  if (rc) return rc;
  if (nb < 1) return fail(IRLS2D_EINVAL, "empty histogram");
  hipLaunchKernelGGL(k_ep_hist, ep_grid_stride(H * P), dim3(EP_THREADS), 0, s, (int)H, (int)W, (int)P, psi, u, kmin,
                     nb, hist);
  HIP_TRY(hipGetLastError());
  return IRLS2D_OK;
}

int ep_label(int64_t H, int64_t W, int64_t P, const float *psi, float *u, float shift, int32_t *L, int32_t *counts,
             hipStream_t s) {
  int rc = ep_check(H, W, P);
  if (rc) return rc;
  const dim3 g = ep_grid_stride(H * P), b(EP_THREADS);
  hipLaunchKernelGGL(k_ep_shift, g, b, 0, s, (int)H, (int)W, (int)P, psi, u, shift, L);
  hipLaunchKernelGGL(k_ccl_merge, g, b, 0, s, (int)H, (int)W, (int)P, L);
  hipLaunchKernelGGL(k_ccl_flat, dim3((unsigned)ep_blocks(H, P)), b, 0, s, (int)H, (int)P, L, counts);
  HIP_TRY(hipGetLastError());
  return IRLS2D_OK;
}

int ep_number(int64_t H, int64_t P, const int32_t *L, const int32_t *offsets, int32_t *R, hipStream_t s) {
  int rc = ep_check(H, 1, P);
  if (rc) return rc;
  hipLaunchKernelGGL(k_ccl_number, dim3((unsigned)ep_blocks(H, P)), dim3(EP_THREADS), 0, s, (int)H, (int)P, L,
                     offsets, R);
  HIP_TRY(hipGetLastError());
  return IRLS2D_OK;
}

int ep_sums(int64_t H, int64_t W, int64_t P, const float *psi, const float *u, const int32_t *L, const int32_t *R,
            double *sums, uint64_t *cnt, hipStream_t s) {
  int rc = ep_check(H, W, P);
  if (rc) return rc;
  hipLaunchKernelGGL(k_ep_sums, dim3((unsigned)ep_blocks(H, P)), dim3(EP_THREADS), 0, s, (int)H, (int)W, (int)P, psi,
                     u, L, R, sums, (unsigned long long *)cnt);
  HIP_TRY(hipGetLastError());
  return IRLS2D_OK;
}

int ep_apply(int64_t H, int64_t W, int64_t P, float *u, const int32_t *L, const int32_t *R, const double *adj,
             const uint16_t *relabel, uint16_t *cc, hipStream_t s) {
  int rc = ep_check(H, W, P);
  if (rc) return rc;
  hipLaunchKernelGGL(k_ep_apply, ep_grid_stride(H * P), dim3(EP_THREADS), 0, s, (int)H, (int)W, (int)P, u, L, R, adj,
                     (const unsigned short *)relabel, (unsigned short *)cc);
  HIP_TRY(hipGetLastError());
  return IRLS2D_OK;
}

int release(void *workspace) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto it = g_graphs.begin(); it != g_graphs.end();) {
    if (!workspace || it->first.ws == workspace) {
      auto cs = g_capture_stream.find(it->second.dev);
      if (cs != g_capture_stream.end()) (void)hipStreamSynchronize(cs->second);
      (void)hipGraphExecDestroy(it->second.exec);
      it = g_graphs.erase(it);
    } else {
      ++it;
    }
  }
  return IRLS2D_OK;
}

}  // namespace irls2d

// ---------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------

extern "C" {

int irls2d_version(void) { return 100; }

const char *irls2d_last_error(void) { return irls2d::g_err.c_str(); }

int irls2d_supported_length(int64_t n) { return irls2d::supported(n); }

int64_t irls2d_arena_layout(int n, const int64_t *H, const int64_t *W, int64_t *offsets, int64_t *pitches) {
  if (n < 1 || !H || !W) return -1;
  return irls2d::arena_layout(n, H, W, offsets, pitches);
}

int irls2d_workspace_size(int n, const int64_t *H, const int64_t *W, size_t *bytes) {
  if (!bytes) return irls2d::fail(IRLS2D_EINVAL, "null bytes pointer");
  std::shared_ptr<irls2d::Layout> Lo;
  int rc = irls2d::get_layout(n, H, W, Lo);
  if (rc) return rc;
  *bytes = Lo->bytes;
  return IRLS2D_OK;
}

int irls2d_solve(int n, const int64_t *H, const int64_t *W, const float *phase, const float *weight,
                 float *unwrapped, int max_iter, double tol, int cg_max_iter, double cg_tol, double epsilon,
                 void *workspace, size_t workspace_bytes, int32_t *info, void *stream, unsigned flags) {
  return irls2d::solve(n, H, W, phase, weight, unwrapped, max_iter, tol, cg_max_iter, cg_tol, epsilon, workspace,
                       workspace_bytes, info, stream, flags);
}

int irls2d_release(void *workspace) { return irls2d::release(workspace); }

int irls2d_debug_inject(int irls_iter, int cg_iter) {
  if (irls_iter < 0 || cg_iter < 0) irls_iter = cg_iter = -1;
  irls2d::g_inject_irls = irls_iter;
  irls2d::g_inject_cg = cg_iter;
  return IRLS2D_OK;
}

int irls2d_solve_static(int n, const int64_t *H, const int64_t *W, const float *phase, const float *weight,
                        float *unwrapped, int max_iter, double tol, int cg_max_iter, double cg_tol, double epsilon,
                        int32_t *info, void *stream) {
  return irls2d::solve_static(n, H, W, phase, weight, unwrapped, max_iter, tol, cg_max_iter, cg_tol, epsilon, info,
                              stream);
}

int64_t irls2d_ep_blocks(int64_t H, int64_t P) { return irls2d::ep_blocks(H, P); }

int irls2d_ep_scan(int64_t H, int64_t W, int64_t P, const float *phase, const float *u, int32_t *stats, void *stream) {
  return irls2d::ep_scan(H, W, P, phase, u, stats, (hipStream_t)stream);
}

int irls2d_ep_hist(int64_t H, int64_t W, int64_t P, const float *phase, const float *u, int32_t kmin, int32_t nbins,
                   int32_t *hist, void *stream) {
  return irls2d::ep_hist(H, W, P, phase, u, kmin, nbins, hist, (hipStream_t)stream);
}

int irls2d_ep_label(int64_t H, int64_t W, int64_t P, const float *phase, float *u, float shift, int32_t *parent,
                    int32_t *block_roots, void *stream) {
  return irls2d::ep_label(H, W, P, phase, u, shift, parent, block_roots, (hipStream_t)stream);
}

int irls2d_ep_number(int64_t H, int64_t P, const int32_t *parent, const int32_t *block_offsets, int32_t *root_label,
                     void *stream) {
  return irls2d::ep_number(H, P, parent, block_offsets, root_label, (hipStream_t)stream);
}

int irls2d_ep_sums(int64_t H, int64_t W, int64_t P, const float *phase, const float *u, const int32_t *parent,
                   const int32_t *root_label, double *sums, uint64_t *counts, void *stream) {
  return irls2d::ep_sums(H, W, P, phase, u, parent, root_label, sums, counts, (hipStream_t)stream);
}

int irls2d_ep_apply(int64_t H, int64_t W, int64_t P, float *u, const int32_t *parent, const int32_t *root_label,
                    const double *adjust, const uint16_t *relabel, uint16_t *conncomp, void *stream) {
  return irls2d::ep_apply(H, W, P, u, parent, root_label, adjust, relabel, conncomp, (hipStream_t)stream);
}

int irls2d_plan_override(int64_t n, int64_t fft_len, const int *radices, int nradices, int ept) {
  return irls2d::plan_override(n, fft_len, radices, nradices, ept);
}

int irls2d_plan_describe(int64_t n, int64_t *fft_len, int *radices, int *nradices, int *ept) {
  return irls2d::plan_describe(n, fft_len, radices, nradices, ept);
}

int irls2d_plan_reset(void) { return irls2d::plan_reset(); }

int irls2d_set_column_mode(int mode) { return irls2d::set_column_mode(mode); }

int irls2d_gaussian_workspace_size(int64_t ny, int64_t nx, int channels, double sigma_y, double sigma_x,
                                   double truncate, size_t *bytes) {
  return irls2d::gaussian_workspace_size(ny, nx, channels, sigma_y, sigma_x, truncate, bytes);
}

int irls2d_gaussian(int64_t ny, int64_t nx, int channels, const float *data, const float *weight, float *out,
                    double sigma_y, double sigma_x, double truncate, double threshold, void *workspace,
                    size_t workspace_bytes, void *stream) {
  return irls2d::gaussian(ny, nx, channels, data, weight, out, sigma_y, sigma_x, truncate, threshold, workspace,
                          workspace_bytes, stream);
}

int irls2d_goldstein_workspace_size(int64_t ny, int64_t nx, int psize_y, int psize_x, size_t *bytes) {
  return irls2d::goldstein_workspace_size(ny, nx, psize_y, psize_x, bytes);
}

int irls2d_goldstein(int64_t ny, int64_t nx, const void *phase, const float *corr, void *out, int psize_y,
                     int psize_x, double threshold, int mode, void *workspace, size_t workspace_bytes,
                     void *stream) {
  return irls2d::goldstein(ny, nx, phase, corr, out, psize_y, psize_x, threshold, mode, workspace, workspace_bytes,
                           stream);
}

int irls2d_describe_columns(int n, const int64_t *H, const int64_t *W, int32_t *strategy) {
  if (!strategy) return irls2d::fail(IRLS2D_EINVAL, "null strategy pointer");
  std::shared_ptr<irls2d::Layout> Lo;
  int rc = irls2d::get_layout(n, H, W, Lo);
  if (rc) return rc;
  for (int i = 0; i < n; ++i) {
    const irls2d::ImgDesc &d = Lo->desc[i];
    if (d.colslab >= 0) {
      const irls2d::SlabPlan &P = Lo->splans[d.colslab];
      strategy[i] = irls2d::slab_fast(P.N1, P.N2) ? IRLS2D_COLUMNS_SLAB_REGISTER : IRLS2D_COLUMNS_SLAB;
    } else if (d.collong >= 0) {
      strategy[i] = IRLS2D_COLUMNS_LONG;
    } else if (d.coltri) {
      strategy[i] = IRLS2D_COLUMNS_TRIDIAG;
    } else {
      strategy[i] = d.coldirect ? IRLS2D_COLUMNS_DIRECT : IRLS2D_COLUMNS_TRANSPOSE;
    }
  }
  return IRLS2D_OK;
}

int irls2d_precondition(int n, const int64_t *H, const int64_t *W, const float *r, float *z, void *workspace,
                        size_t workspace_bytes, void *stream) {
  return irls2d::precondition(n, H, W, r, z, workspace, workspace_bytes, stream);
}

int irls2d_adjoint(int n, const int64_t *H, const int64_t *W, const float *phase, const float *weight,
                   const float *unwrapped, const float *grad_unwrapped, float *grad_phase, float *grad_weight,
                   double epsilon, int cg_max_iter, double cg_tol, void *workspace, size_t workspace_bytes,
                   int32_t *info, void *stream, unsigned flags) {
  return irls2d::adjoint(n, H, W, phase, weight, unwrapped, grad_unwrapped, grad_phase, grad_weight, epsilon,
                         cg_max_iter, cg_tol, workspace, workspace_bytes, info, stream, flags);
}

int irls2d_kernel_stats(int max, char *names, double *total_ms, int64_t *launches) {
  std::lock_guard<std::mutex> lk(irls2d::g_stats_mu);
  int k = 0;
  for (const auto &e : irls2d::g_kstats) {
    if (k < max) {
      if (names) {
        std::strncpy(names + (size_t)k * 64, e.first.c_str(), 63);
        names[(size_t)k * 64 + 63] = 0;
      }
      if (total_ms) total_ms[k] = e.second.first;
      if (launches) launches[k] = e.second.second;
    }
    ++k;
  }
  return k;
}

void irls2d_kernel_stats_reset(void) {
  std::lock_guard<std::mutex> lk(irls2d::g_stats_mu);
  irls2d::g_kstats.clear();
}

}  // extern "C"
