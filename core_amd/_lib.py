"""ctypes binding of the gfx950 solver library (include/irls2d.h).

torch is imported before the library is opened so that the library's libamdhip64 dependency
resolves to the HIP runtime torch already loaded (same soname): device pointers and stream
handles then come from one runtime.  There is no fallback: if the library is missing or
fails to load, every solve raises.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (must precede the library load, see above)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libirls2d.so")

OK, EINVAL, EUNSUPPORTED, EWORKSPACE, EHIP = 0, 1, 2, 3, 4
FLAG_GRAPH = 1
FLAG_TIME_KERNELS = 2
# IRLS2D_STATIC_SCRATCH_BYTES of include/irls2d.h: the device scratch of irls2d_solve_static
STATIC_SCRATCH_BYTES = 96 << 20

EXPORTS = (
    "irls2d_version",
    "irls2d_last_error",
    "irls2d_supported_length",
    "irls2d_arena_layout",
    "irls2d_workspace_size",
    "irls2d_solve",
    "irls2d_solve_static",
    "irls2d_adjoint",
    "irls2d_precondition",
    "irls2d_plan_override",
    "irls2d_plan_describe",
    "irls2d_plan_reset",
    "irls2d_set_column_mode",
    "irls2d_describe_columns",
    "irls2d_gaussian_workspace_size",
    "irls2d_gaussian",
    "irls2d_goldstein_workspace_size",
    "irls2d_goldstein",
    "irls2d_ep_blocks",
    "irls2d_ep_scan",
    "irls2d_ep_hist",
    "irls2d_ep_label",
    "irls2d_ep_number",
    "irls2d_ep_sums",
    "irls2d_ep_apply",
    "irls2d_release",
    "irls2d_debug_inject",
    "irls2d_kernel_stats",
    "irls2d_kernel_stats_reset",
)

_lib = None


class SolverError(RuntimeError):
    pass


def load():
    """Open libirls2d.so (once) and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SolverError(f"{LIB_PATH} is missing: build it with `python -m core_amd.build` "
                          "(this package has no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    i64p = ctypes.POINTER(ctypes.c_int64)
    L.irls2d_version.restype = ctypes.c_int
    L.irls2d_last_error.restype = ctypes.c_char_p
    L.irls2d_supported_length.argtypes = [ctypes.c_int64]
    L.irls2d_supported_length.restype = ctypes.c_int
    L.irls2d_arena_layout.argtypes = [ctypes.c_int, i64p, i64p, i64p, i64p]
    L.irls2d_arena_layout.restype = ctypes.c_int64
    L.irls2d_workspace_size.argtypes = [ctypes.c_int, i64p, i64p, ctypes.POINTER(ctypes.c_size_t)]
    L.irls2d_workspace_size.restype = ctypes.c_int
    L.irls2d_solve.argtypes = [
        ctypes.c_int, i64p, i64p,
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
        ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint,
    ]
    L.irls2d_solve.restype = ctypes.c_int
    L.irls2d_solve_static.argtypes = [
        ctypes.c_int, i64p, i64p,
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
        ctypes.c_void_p, ctypes.c_void_p,
    ]
    L.irls2d_solve_static.restype = ctypes.c_int
    L.irls2d_adjoint.argtypes = [
        ctypes.c_int, i64p, i64p,
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.c_double, ctypes.c_int, ctypes.c_double,
        ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint,
    ]
    L.irls2d_adjoint.restype = ctypes.c_int
    L.irls2d_precondition.argtypes = [ctypes.c_int, i64p, i64p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    L.irls2d_precondition.restype = ctypes.c_int
    i32p = ctypes.POINTER(ctypes.c_int)
    L.irls2d_plan_override.argtypes = [ctypes.c_int64, ctypes.c_int64, i32p, ctypes.c_int, ctypes.c_int]
    L.irls2d_plan_override.restype = ctypes.c_int
    L.irls2d_plan_describe.argtypes = [ctypes.c_int64, i64p, i32p, i32p, i32p]
    L.irls2d_plan_describe.restype = ctypes.c_int
    L.irls2d_plan_reset.argtypes = []
    L.irls2d_plan_reset.restype = ctypes.c_int
    L.irls2d_set_column_mode.argtypes = [ctypes.c_int]
    L.irls2d_set_column_mode.restype = ctypes.c_int
    L.irls2d_describe_columns.argtypes = [ctypes.c_int, i64p, i64p, ctypes.POINTER(ctypes.c_int32)]
    L.irls2d_describe_columns.restype = ctypes.c_int
    L.irls2d_gaussian_workspace_size.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_double,
                                                 ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_size_t)]
    L.irls2d_gaussian_workspace_size.restype = ctypes.c_int
    L.irls2d_gaussian.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    L.irls2d_gaussian.restype = ctypes.c_int
    L.irls2d_goldstein_workspace_size.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                                  ctypes.POINTER(ctypes.c_size_t)]
    L.irls2d_goldstein_workspace_size.restype = ctypes.c_int
    L.irls2d_goldstein.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_void_p,
                                   ctypes.c_size_t, ctypes.c_void_p]
    L.irls2d_goldstein.restype = ctypes.c_int
    L.irls2d_debug_inject.argtypes = [ctypes.c_int, ctypes.c_int]
    L.irls2d_debug_inject.restype = ctypes.c_int
    L.irls2d_release.argtypes = [ctypes.c_void_p]
    L.irls2d_release.restype = ctypes.c_int
    L.irls2d_kernel_stats.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), i64p]
    L.irls2d_kernel_stats.restype = ctypes.c_int
    L.irls2d_kernel_stats_reset.argtypes = []
    L.irls2d_kernel_stats_reset.restype = None
    _lib = L
    return L


def check(rc: int, what: str):
    if rc != OK:
        msg = load().irls2d_last_error().decode(errors="replace")
        if rc == EINVAL:
            raise ValueError(f"{what}: {msg}")
        raise SolverError(f"{what} failed (code {rc}): {msg}")


def shapes_arrays(shapes):
    n = len(shapes)
    H = (ctypes.c_int64 * n)(*[int(h) for h, _ in shapes])
    W = (ctypes.c_int64 * n)(*[int(w) for _, w in shapes])
    return n, H, W


def arena_layout(shapes):
    """(offsets, pitches, total): image i is H rows of pitches[i] elements at offsets[i]."""
    L = load()
    n, H, W = shapes_arrays(shapes)
    off = (ctypes.c_int64 * n)()
    pitch = (ctypes.c_int64 * n)()
    total = L.irls2d_arena_layout(n, H, W, off, pitch)
    if total < 0:
        raise ValueError(f"bad shapes {shapes}")
    return [int(o) for o in off], [int(p) for p in pitch], int(total)


def workspace_size(shapes) -> int:
    L = load()
    n, H, W = shapes_arrays(shapes)
    b = ctypes.c_size_t(0)
    check(L.irls2d_workspace_size(n, H, W, ctypes.byref(b)), "irls2d_workspace_size")
    return int(b.value)


def supported_length(n: int) -> bool:
    return bool(load().irls2d_supported_length(int(n)))


def kernel_stats():
    """{kernel name: (total ms, launches)} recorded by solves run with FLAG_TIME_KERNELS."""
    L = load()
    n = L.irls2d_kernel_stats(0, None, None, None)
    names = ctypes.create_string_buffer(64 * max(n, 1))
    tot = (ctypes.c_double * max(n, 1))()
    cnt = (ctypes.c_int64 * max(n, 1))()
    n = L.irls2d_kernel_stats(n, names, tot, cnt)
    out = {}
    for k in range(n):
        nm = names.raw[64 * k:64 * (k + 1)].split(b"\0", 1)[0].decode()
        out[nm] = (float(tot[k]), int(cnt[k]))
    return out


def kernel_stats_reset():
    load().irls2d_kernel_stats_reset()


def plan_describe(n: int):
    """(fft_len, radices, ept) of the transform plan used for DCT length n."""
    L = load()
    fl = ctypes.c_int64()
    rad = (ctypes.c_int * 16)()
    nr = ctypes.c_int()
    ept = ctypes.c_int()
    check(L.irls2d_plan_describe(int(n), ctypes.byref(fl), rad, ctypes.byref(nr), ctypes.byref(ept)),
          "irls2d_plan_describe")
    return int(fl.value), [int(rad[i]) for i in range(nr.value)], int(ept.value)


def plan_override(n: int, fft_len: int, radices=(), ept: int = 0):
    """Set (fft_len > 0) or remove (fft_len = 0) the transform plan of DCT length n."""
    L = load()
    rad = (ctypes.c_int * max(1, len(radices)))(*[int(r) for r in radices])
    check(L.irls2d_plan_override(int(n), int(fft_len), rad, len(radices), int(ept)), "irls2d_plan_override")


def plan_reset():
    check(load().irls2d_plan_reset(), "irls2d_plan_reset")


COLUMN_AUTO, COLUMN_TRANSPOSE, COLUMN_DIRECT, COLUMN_SLAB, COLUMN_TRIDIAG = 0, 1, 2, 3, 4


COLUMN_STRATEGIES = ("transpose", "direct", "long", "slab", "slab_register", "tridiag")


def describe_columns(shapes):
    """Column strategy of each image of a batch of (H, W) shapes (a name from COLUMN_STRATEGIES)."""
    n, H, W = shapes_arrays(shapes)
    out = (ctypes.c_int32 * n)()
    check(load().irls2d_describe_columns(n, H, W, out), "irls2d_describe_columns")
    return [COLUMN_STRATEGIES[v] for v in out]


def set_column_mode(mode: int):
    """0 = automatic, 1 = transpose + column kernel, 2 = column kernel on the row-major field,
    3 = four-step slab passes on the row-major field wherever the column length allows them,
    4 = tridiagonal column solves for every image."""
    check(load().irls2d_set_column_mode(int(mode)), "irls2d_set_column_mode")
