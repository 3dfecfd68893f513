/*
 * irls2d.h -- C ABI of the MI355X (gfx950) DCT-preconditioned IRLS L1 phase-unwrapping solver.
 *
 * Reference interface replaced: the device section of
 *   insardev/insardev/utils_unwrap2d.py:1247-1632  (irls_unwrap_2d: H2D of the filled,
 *   bbox-cropped phase and weights, targets, edge weights, DCT eigenvalues, L2 DCT
 *   initialisation, IRLS loop with PCG, u_best tracking)
 * and, for n > 1 images, the sequential per-pair loop that
 *   insardev/insardev/Stack_unwrap2d.py:32-71,926-945 (blockwise task + process GPU lock)
 * runs over an SBAS stack.  The reference has no FFI for this path (it is Python calling
 * torch); the Python boundary core_amd.irls_unwrap_2d binds these symbols with ctypes.
 *
 * Conventions
 *   - All array arguments are DEVICE pointers (HIP global memory) unless noted.
 *   - Images are packed in "arenas": image i occupies H[i] rows of pitches[i] elements starting
 *     at element offsets[i], row-major (pixel (y, x) at offsets[i] + y * pitches[i] + x); both
 *     come from irls2d_arena_layout: pitches[i] is W[i] rounded up to a multiple of 64 and every
 *     offset is a multiple of 64, so rows are 256-byte aligned relative to the arena base.
 *     Padding elements are ignored on input and unspecified on output.
 *   - Caller float32 arrays need only 4-byte alignment: the kernels make float2 / float4
 *     accesses on the library's own workspace, and on a caller float32 array only after checking
 *     its address (irls2d_precondition's output).  Complex64 arrays (irls2d_goldstein) need
 *     their natural 8-byte alignment.
 *   - phase: float32 radians, NaN marks invalid pixels (the crop is done by the caller).
 *   - weight: float32 quality (NaN allowed) or NULL for uniform weights.
 *   - unwrapped: float32 output arena (the raw solver output u, before the host epilogue).
 *   - Every call is stream-ordered on `stream` (a hipStream_t, NULL = default stream).  The
 *     transform tables and image descriptors are uploaded into the workspace on every call
 *     (through pinned staging; the host waits at most for the upload two calls back).
 *   - Return value: IRLS2D_OK or an error code; irls2d_last_error() gives the message.
 */
#ifndef IRLS2D_H
#define IRLS2D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IRLS2D_OK 0
#define IRLS2D_EINVAL 1       /* bad argument (shape, pointer, iteration count) */
#define IRLS2D_EUNSUPPORTED 2 /* grid dimension beyond what the transforms handle (a side above ~2^19) */
#define IRLS2D_EWORKSPACE 3   /* workspace too small */
#define IRLS2D_EHIP 4         /* HIP runtime error */

/* flags for irls2d_solve */
#define IRLS2D_FLAG_GRAPH 1u        /* capture the whole launch sequence once in a hipGraph and replay it */
#define IRLS2D_FLAG_TIME_KERNELS 2u /* eager only: hipEvent pairs around every launch, see irls2d_kernel_stats;
                                       the call then synchronises `stream` before returning */

/* info[i*4 + k] written by irls2d_solve (device int32 array of 4*n) */
#define IRLS2D_INFO_STATUS 0   /* bit0: IRLS hit a non-finite u and restored u_best; bit1: stopped by tol */
#define IRLS2D_INFO_IRLS_ITERS 1
#define IRLS2D_INFO_CG_ITERS 2 /* PCG iterations summed over the IRLS loop */
#define IRLS2D_INFO_VALID 3    /* number of valid (non-NaN) pixels */

/* Library version (major*10000 + minor*100 + patch). */
int irls2d_version(void);

/* Message of the last error on the calling thread ("" if none). */
const char *irls2d_last_error(void);

/* 1 if a grid side of length N is supported: N <= 12288 (LDS-resident DCT, Bluestein for prime
   factors above 5); a larger N = 2^a 3^b 5^c that splits into two factors <= 1024 (two-pass
   four-step DCT); any other larger N whose Bluestein length (5-smooth, >= 2N - 1, within 30%)
   splits into two factors <= 1024, i.e. N up to about 2^19 (four-step Bluestein passes). */
int irls2d_supported_length(int64_t n);

/*
 * Arena layout.  Host arrays H[n], W[n] in; host offsets[n] and pitches[n] out (either may be
 * NULL).  Returns the total number of elements of one arena, or -1 on bad arguments.  No
 * reference counterpart (the reference solves one numpy array per call); it defines the packing
 * of many uneven images for one launch.
 */
int64_t irls2d_arena_layout(int n, const int64_t *H, const int64_t *W, int64_t *offsets, int64_t *pitches);

/*
 * Workspace bytes needed by irls2d_solve for this set of shapes (host arrays).  The workspace
 * holds the solver's working fields (targets, edge and IRLS weights, PCG vectors, DCT buffers),
 * the per-image FFT/DCT plans and twiddle tables, and reduction scratch.
 */
int irls2d_workspace_size(int n, const int64_t *H, const int64_t *W, size_t *bytes);

/*
 * Solve n images.  Replaces utils_unwrap2d.py:1247-1632 (per image) with identical control flow:
 * L2 DCT start, max_iter IRLS iterations of w/sqrt(r^2+eps^2) clamped to [1e-6, 1e6], each
 * followed by up to cg_max_iter DCT-preconditioned CG iterations (absolute residual test
 * cg_tol), relative-change test tol, u_best tracking.  tol = 0 and cg_tol = 0 give fixed
 * iteration counts.  H, W: host arrays; phase, weight, unwrapped: device arenas laid out by
 * irls2d_arena_layout; info: device int32[4*n] (may be NULL); workspace: device memory of at
 * least irls2d_workspace_size bytes, owned by the library for the duration of the call.
 * Calls that share a workspace must not overlap.
 */
int irls2d_solve(int n, const int64_t *H, const int64_t *W,
                 const float *phase, const float *weight, float *unwrapped,
                 int max_iter, double tol, int cg_max_iter, double cg_tol, double epsilon,
                 void *workspace, size_t workspace_bytes, int32_t *info,
                 void *stream, unsigned flags);

/*
 * irls2d_solve without a caller workspace: the same launch sequence on the library's static device
 * scratch array of IRLS2D_STATIC_SCRATCH_BYTES per GPU (allocated with the library's code object on
 * each GPU it runs on and held until the process ends).  No counterpart in the reference: for
 * callers that cannot allocate device memory (serving loops, harnesses that forbid hipMalloc).  The
 * call allocates no device memory (no hipMalloc / hipFree / allocating library routine); it
 * uploads the transform tables from pageable host memory and synchronises `stream` once before
 * launching, then runs eagerly (no hipGraph).  Shapes whose irls2d_workspace_size exceeds
 * IRLS2D_STATIC_SCRATCH_BYTES (about 1.5 Mpx in total, e.g. one 1024 x 1024 image) are rejected
 * with IRLS2D_EWORKSPACE.  Every call on a GPU uses that GPU's one array: calls that overlap on a
 * GPU (other threads or streams) race and must be serialised by the caller.  Nothing in the array
 * needs to start at zero (every value is written before it is read within a call).
 */
#define IRLS2D_STATIC_SCRATCH_BYTES (96ull << 20)
int irls2d_solve_static(int n, const int64_t *H, const int64_t *W,
                        const float *phase, const float *weight, float *unwrapped,
                        int max_iter, double tol, int cg_max_iter, double cg_tol, double epsilon,
                        int32_t *info, void *stream);

/*
 * Implicit adjoint of the solve (no reference counterpart: the reference is not differentiable).
 * Given the inputs of irls2d_solve, its output `unwrapped` (the IRLS fixed point u*) and
 * grad_unwrapped = dL/d(C u*) with C the centring u - mean_valid(u), writes
 *   grad_phase  = dL/dphase  (0 at NaN pixels),
 *   grad_weight = dL/dweight (may be NULL; 0 where the clip(weight, 0.01, 1) is flat),
 * by solving (D^T W' D) lambda = C^T grad with the same DCT-preconditioned CG (cg_max_iter
 * iterations, absolute tolerance cg_tol), W' = d(w_hat r)/dr of the IRLS weights at u*.
 * Same arenas and workspace layout as irls2d_solve.
 */
int irls2d_adjoint(int n, const int64_t *H, const int64_t *W,
                   const float *phase, const float *weight, const float *unwrapped,
                   const float *grad_unwrapped, float *grad_phase, float *grad_weight,
                   double epsilon, int cg_max_iter, double cg_tol,
                   void *workspace, size_t workspace_bytes, int32_t *info,
                   void *stream, unsigned flags);

/*
 * Poisson preconditioner alone: z = IDCT2(-DCT2(r) / lambda) with the DC term zeroed, for every
 * image (replaces apply_preconditioner, utils_unwrap2d.py:1390-1403, with the DCT pair of
 * :1321-1365 and the eigenvalues of :1284-1289).  r and z are device arenas laid out by
 * irls2d_arena_layout (they may not alias); same workspace as irls2d_solve.  This is the
 * transform step irls2d_solve applies twice per PCG iteration; exported for callers that need
 * the unweighted Poisson solve and to test the DCT engine on its own.
 */
int irls2d_precondition(int n, const int64_t *H, const int64_t *W, const float *r, float *z,
                        void *workspace, size_t workspace_bytes, void *stream);

/*
 * Autotuning hooks (no reference counterpart: the reference calls torch-dct's fixed plans).
 * irls2d_plan_override sets the transform plan of DCT length n (1 .. 12288): fft_len = n for a
 * direct FFT (n 5-smooth) or a 5-smooth Bluestein length >= 2n - 1, `radices` the Stockham pass
 * order (nradices entries from {2, 3, 4, 5, 8, 16} multiplying to fft_len; 0 entries = the
 * default order), ept the elements per thread (12 on up to 1024 threads, 16 or 24 on up to 512,
 * or 0 = default: 16 up to 8192 points, 12 above; 16 needs fft_len <= 8192).  Each image's row and
 * column transforms run with their own plan's value (one launch per value in a batch), so an
 * image's result does not depend on the other images of its batch.
 * fft_len = 0 removes the override.  Changing plans drops every cached layout, uploaded table set
 * and hipGraph (calls must not be in flight).  irls2d_plan_describe reports the plan in use
 * (radices: room for 16 entries); irls2d_plan_reset removes all overrides.
 * core_amd.autotune times candidate plans with irls2d_precondition and applies the fastest.
 */
int irls2d_plan_override(int64_t n, int64_t fft_len, const int *radices, int nradices, int ept);
int irls2d_plan_describe(int64_t n, int64_t *fft_len, int *radices, int *nradices, int *ept);
int irls2d_plan_reset(void);

/*
 * Column-pass strategy (no reference counterpart: the reference transposes with torch).
 * 0 = automatic (default), decided per image from its own shape: images with both sides <= 12288
 * whose column length, at least 256, splits into a power-of-two pair N1 N2 (N1 16 .. 128, N2
 * 16 .. 64: heights 256 .. 8192) run the three four-step slab passes with register-stage kernels on
 * the row-major field (no transposes; one launch per split in a batch); images whose column
 * transform would be Bluestein (a column length with a prime factor above 5) or a two-pass long
 * transform (H above 12288) replace the column DCT pair by the equivalent tridiagonal solves
 * (L_H + mu_k I) x = c_k, in float64, on the row-major field; the other images transpose their
 * rows, transform and transpose back;
 * 1 = transpose (two-pass long transforms above 12288); 2 = column DCTs on the row-major field
 * (long ones transposed); 3 = slab passes for every image whose column length allows them
 * (2^a 3^b 5^c with such a split, from 4), the automatic choice for the rest; 4 = tridiagonal
 * solves for every image.  Since every choice is per image, an image's result is the same alone
 * and in any batch (under one mode and one plan set).  Drops cached layouts and hipGraphs.
 * core_amd.autotune measures the strategies.
 */
int irls2d_set_column_mode(int mode);

/*
 * NaN-aware separable Gaussian filter of one ny x nx image, real (channels 1) or complex
 * (channels 2, interleaved re, im; each part filtered on its own), into `out` (same layout; may
 * not alias `data`).  Replaces insardev/insardev/utils_gaussian.py nanconvolve2d_gaussian_pytorch
 * (:14-116, the reference's GPU path): 1-D kernels of size 2 ceil(sigma truncate) + 1, replicate
 * borders, y then x; NaN data and NaN weights read as weight 0; `weight` (ny x nx float32) may be
 * null; out = numerator / (denominator + 1e-10) where the denominator exceeds `threshold` (no
 * weight) or 1e-10 (with weight), NaN elsewhere.
 */
int irls2d_gaussian_workspace_size(int64_t ny, int64_t nx, int channels, double sigma_y, double sigma_x,
                                   double truncate, size_t *bytes);
int irls2d_gaussian(int64_t ny, int64_t nx, int channels, const float *data, const float *weight, float *out,
                    double sigma_y, double sigma_x, double truncate, double threshold, void *workspace,
                    size_t workspace_bytes, void *stream);

/*
 * Goldstein adaptive phase filter of one ny x nx complex64 image (interleaved re, im; rows of nx)
 * with its float32 correlation, into `out` (complex64, same layout; may not alias `phase`).
 * Replaces insardev/insardev/utils_goldstein.py: goldstein_pytorch (:157-330, the reference's GPU
 * path, mode 0: overlapping patches averaged over the number of valid patches) and goldstein_numpy
 * (:12-154, its CPU path, mode 1: weighted by the triangle window sums), as dispatched by
 * Batch.py:1969-2035.  Patch sizes 16, 32 or 64 per axis (steps of half a patch); a patch is valid
 * with at least int(psize_y psize_x threshold) non-NaN pixels; NaN phase/corr read as 0; output
 * NaN where the input is NaN or no valid patch (mode 0) / no window weight (mode 1) covers it.
 * Errors: IRLS2D_EUNSUPPORTED for other patch sizes, IRLS2D_EINVAL for an image smaller than half
 * a patch or bad arguments, IRLS2D_EWORKSPACE below irls2d_goldstein_workspace_size.
 */
int irls2d_goldstein_workspace_size(int64_t ny, int64_t nx, int psize_y, int psize_x, size_t *bytes);
int irls2d_goldstein(int64_t ny, int64_t nx, const void *phase, const float *corr, void *out, int psize_y,
                     int psize_x, double threshold, int mode, void *workspace, size_t workspace_bytes,
                     void *stream);

/*
 * Column strategy each image of a batch of shapes gets under the current mode and plans (no
 * reference counterpart): strategy[i] is one of the IRLS2D_COLUMNS_* values below.
 */
#define IRLS2D_COLUMNS_TRANSPOSE 0      /* transpose, LDS-resident column DCTs, transpose back */
#define IRLS2D_COLUMNS_DIRECT 1         /* LDS-resident column DCTs on the row-major field */
#define IRLS2D_COLUMNS_LONG 2           /* transpose, two-pass four-step DCTs (H > 12288, no slab split) */
#define IRLS2D_COLUMNS_SLAB 3           /* three four-step slab passes, batched LDS FFTs */
#define IRLS2D_COLUMNS_SLAB_REGISTER 4  /* three four-step slab passes, register-stage FFTs */
#define IRLS2D_COLUMNS_TRIDIAG 5        /* float64 tridiagonal solves on the row-major field */
int irls2d_describe_columns(int n, const int64_t *H, const int64_t *W, int32_t *strategy);

/*
 * Device epilogue of one solved image (replaces the host part of utils_unwrap2d.py:1659-1746: the
 * non-finite check, the 2 pi median shift, 4-connected labelling, the per-component circular-mean
 * alignment of :24-75 and the relabel by size).  The image is an H x W view with row pitch P
 * (elements) of the phase arena and of the solver output `u` (updated in place); the other arrays
 * are device scratch owned by the caller.  The host runs the steps in order and takes the two
 * decisions with numpy semantics between them (core_amd/epilogue_gpu.py):
 *   irls2d_ep_scan   stats[4] (int32, preset {0, INT32_MAX, INT32_MIN, 0}) += {non-finite u on
 *                    valid pixels, min k, max k, valid pixels}, k = rint((u - phase) / 2 pi)
 *   irls2d_ep_hist   hist[k - kmin] += 1 (nbins = kmax - kmin + 1, zeroed)     -> host: median
 *   irls2d_ep_label  u -= shift on valid pixels, NaN elsewhere; parent[H * P]: 4-connected union
 *                    find, every valid pixel pointing at its component's raster-first pixel,
 *                    -1 elsewhere; block_roots[irls2d_ep_blocks(H, P)] = roots per block
 *   irls2d_ep_number root_label[root] = raster-order component number (1-based) from the
 *                    exclusive prefix sums of block_roots (block_offsets)
 *   irls2d_ep_sums   sums[4 c + 0..3] += sin phase, cos phase, sin u, cos u (float64), counts[c]
 *                    += 1 over component c's pixels (arrays of n_components + 1, zeroed)
 *                    -> host: adjustment per component and size relabel
 *   irls2d_ep_apply  u = float(double(u) + adjust[c]), conncomp = relabel[c] (uint16, pitch P),
 *                    0 on invalid pixels
 * No reference counterpart as an interface (the reference does these steps with numpy, cv2 and
 * numba on the host).  Images with H * P >= 2^31 are rejected (IRLS2D_EUNSUPPORTED).
 */
int64_t irls2d_ep_blocks(int64_t H, int64_t P);
int irls2d_ep_scan(int64_t H, int64_t W, int64_t P, const float *phase, const float *u, int32_t *stats, void *stream);
int irls2d_ep_hist(int64_t H, int64_t W, int64_t P, const float *phase, const float *u, int32_t kmin, int32_t nbins,
                   int32_t *hist, void *stream);
int irls2d_ep_label(int64_t H, int64_t W, int64_t P, const float *phase, float *u, float shift, int32_t *parent,
                    int32_t *block_roots, void *stream);
int irls2d_ep_number(int64_t H, int64_t P, const int32_t *parent, const int32_t *block_offsets, int32_t *root_label,
                     void *stream);
int irls2d_ep_sums(int64_t H, int64_t W, int64_t P, const float *phase, const float *u, const int32_t *parent,
                   const int32_t *root_label, double *sums, uint64_t *counts, void *stream);
int irls2d_ep_apply(int64_t H, int64_t W, int64_t P, float *u, const int32_t *parent, const int32_t *root_label,
                    const double *adjust, const uint16_t *relabel, uint16_t *conncomp, void *stream);

/*
 * Release the cached hipGraphs that refer to `workspace` (NULL = all), after the graph runs in
 * flight have finished.  Call before freeing a workspace that was used with IRLS2D_FLAG_GRAPH (the
 * cache keeps at most 8 graphs, least recently used evicted first).
 */
int irls2d_release(void *workspace);

/*
 * Test hook (no reference counterpart): from the next irls2d_solve on, in PCG iteration cg_iter
 * (0-based) of IRLS iteration irls_iter, the search direction at pixel (0, 0) of every image whose
 * PCG is still running is set to +inf after alpha is formed, which makes the update of x
 * non-finite and exercises the reference's non-finite-x break and u_best restore
 * (utils_unwrap2d.py:1444-1449, :1599-1604), unreachable from finite float32 inputs.  A negative
 * argument disarms it.  While armed, solves run eagerly (IRLS2D_FLAG_GRAPH is ignored).
 */
int irls2d_debug_inject(int irls_iter, int cg_iter);

/*
 * Per-kernel timing accumulated by calls made with IRLS2D_FLAG_TIME_KERNELS.  Fills up to `max`
 * entries: names (64-byte NUL-terminated slots), total milliseconds and launch counts (any
 * pointer may be NULL) and returns the number of distinct kernels recorded.  No reference
 * counterpart (the reference prints phase timings with debug=True, utils_unwrap2d.py:1758-1764).
 */
int irls2d_kernel_stats(int max, char *names, double *total_ms, int64_t *launches);
void irls2d_kernel_stats_reset(void);

#ifdef __cplusplus
}
#endif

#endif /* IRLS2D_H */
