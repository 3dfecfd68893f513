"""CPU tests of the boundary: the C ABI loads and exports every symbol of include/irls2d.h, its
host-only entry points (layout, workspace size, argument errors) behave, and the host epilogue of
the package matches the oracle's restatement of the reference epilogue."""
import ctypes
import os
import re
import warnings

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "irls2d.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(irls2d_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    import core_amd._lib as L
    return L


def test_library_exports_every_header_symbol(lib):
    L = lib.load()
    names = header_functions()
    assert len(names) >= 9
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(lib.EXPORTS)
    assert L.irls2d_version() >= 100


def test_static_scratch_size_matches_header(lib):
    """The Python constant of irls2d_solve_static's scratch is the header's, and it holds a
    1024 x 1024 solve but not a 2048 x 2048 one."""
    text = open(os.path.join(ROOT, "include", "irls2d.h")).read()
    m = re.search(r"#define IRLS2D_STATIC_SCRATCH_BYTES \((\d+)ull << (\d+)\)", text)
    assert m and int(m.group(1)) << int(m.group(2)) == lib.STATIC_SCRATCH_BYTES
    assert lib.workspace_size([(1024, 1024)]) <= lib.STATIC_SCRATCH_BYTES < lib.workspace_size([(2048, 2048)])


def test_column_strategy_is_per_image(lib):
    """Each image's column strategy follows from its own shape, the same alone and in any batch:
    power-of-two-split heights take the register slab passes (one launch per split), the others
    the tridiagonal column solves."""
    shapes = [(1024, 4000), (1030, 5000), (1000, 4500), (2048, 300), (96, 130), (103, 64)]
    alone = [lib.describe_columns([sh])[0] for sh in shapes]
    assert alone == ["slab_register", "tridiag", "tridiag", "slab_register", "tridiag", "tridiag"]
    assert lib.describe_columns(shapes) == alone
    assert lib.describe_columns(shapes[::-1]) == alone[::-1]


def test_arena_layout_pitch_and_alignment(lib):
    shapes = [(10, 20), (3, 65), (1, 1), (7, 128)]
    off, pitch, total = lib.arena_layout(shapes)
    assert pitch == [64, 128, 64, 128]
    assert off[0] == 0 and all(o % 64 == 0 for o in off)
    ends = [o + h * p for o, p, (h, w) in zip(off, pitch, shapes)]
    assert all(e <= o2 for e, o2 in zip(ends, off[1:]))
    assert total == ends[-1]


def test_workspace_size_grows_with_pixels(lib):
    a = lib.workspace_size([(64, 64)])
    b = lib.workspace_size([(128, 128)])
    c = lib.workspace_size([(64, 64), (128, 128)])
    assert 0 < a < b < c
    # 15 float fields + transposed buffer + valid bytes at least
    assert b >= 16 * 128 * 128 * 4


def test_supported_lengths(lib):
    assert lib.supported_length(8192) and lib.supported_length(12288)
    assert lib.supported_length(4999)            # Bluestein
    # prime factors above 5 between 6145 and 12288 (Bluestein length beyond the LDS limit): the
    # four-step Bluestein passes
    assert all(lib.supported_length(n) for n in (6145, 6151, 7001, 10007, 12007, 12287))
    assert lib.supported_length(12289)           # beyond the LDS-resident lengths: four-step Bluestein
    assert lib.supported_length(20011)           # prime
    assert lib.supported_length(40000)           # four-step 200 x 200
    assert not lib.supported_length(600011)      # the Bluestein length would exceed 1024 x 1024
    assert lib.supported_length(12500) and lib.supported_length(15625) and lib.supported_length(13122)
    assert lib.supported_length(2 * 12289)       # long and not 5-smooth: Bluestein
    assert not lib.supported_length(2 ** 21)     # no split into two factors <= 1024
    assert lib.supported_length(2 ** 20)
    assert not lib.supported_length(0)


def test_workspace_size_long_sides(lib):
    # a long side adds a complex scratch line per line pair (8 B per pair element = 4 B per pixel)
    short = lib.workspace_size([(24, 12000)])
    long_ = lib.workspace_size([(24, 12500)])
    assert long_ > short * 12500 / 12000 + 12 * 12500 * 8 * 0.99
    # long rows and columns: 15 float32 fields and the long-row scratch (4 B/px); the columns run as
    # tridiagonal solves in place (no transposed copy); with transposed two-pass columns the
    # transposed copy and the column scratch come on top
    px = 12500 * 12800
    both = lib.workspace_size([(12500, 12800)])
    assert 16 * px * 4 <= both < 17 * px * 4
    try:
        lib.set_column_mode(lib.COLUMN_TRANSPOSE)
        assert lib.workspace_size([(12500, 12800)]) >= 17 * px * 4
    finally:
        lib.set_column_mode(lib.COLUMN_AUTO)


def test_argument_errors_without_gpu(lib):
    L = lib.load()
    n, H, W = lib.shapes_arrays([(16, 16)])
    # null pointers -> EINVAL before any HIP call
    rc = L.irls2d_solve(n, H, W, None, None, None, 10, 0.0, 50, 0.0, 1e-2, None, 0, None, None, 0)
    assert rc == lib.EINVAL and b"null" in L.irls2d_last_error()
    with pytest.raises(ValueError):
        lib.check(rc, "irls2d_solve")
    nb, Hb, Wb = lib.shapes_arrays([(0, 16)])
    b = ctypes.c_size_t(0)
    assert L.irls2d_workspace_size(nb, Hb, Wb, ctypes.byref(b)) == lib.EINVAL
    nu, Hu, Wu = lib.shapes_arrays([(16, 600011)])   # beyond the four-step Bluestein lengths
    assert L.irls2d_workspace_size(nu, Hu, Wu, ctypes.byref(b)) == lib.EUNSUPPORTED
    nb2, Hb2, Wb2 = lib.shapes_arrays([(16, 20011)])  # a long prime side is supported
    assert L.irls2d_workspace_size(nb2, Hb2, Wb2, ctypes.byref(b)) == lib.OK
    assert L.irls2d_arena_layout(0, H, W, None, None) == -1


def test_resolve_device_rules():
    from core_amd.unwrap2d import resolve_device
    for d in ("cpu", "mps", "tpu"):
        with pytest.raises(ValueError):
            resolve_device(d)
    with pytest.raises(ValueError):
        resolve_device(3.0)


def test_crop_box_matches_oracle_bbox():
    from core_amd.postprocess import crop_box
    from oracle import post
    r = np.random.default_rng(0)
    for _ in range(20):
        a = r.normal(size=(13, 17)).astype(np.float32)
        a[r.uniform(size=a.shape) < 0.5] = np.nan
        a[:r.integers(0, 4)] = np.nan
        a[:, -r.integers(1, 4):] = np.nan
        nm, box = crop_box(a)
        if np.isnan(a).all():
            assert box is None
        else:
            assert box == post.bbox(np.isnan(a))


@pytest.mark.parametrize("seed", range(6))
def test_epilogue_matches_oracle(seed):
    from core_amd.postprocess import epilogue
    from oracle import post
    r = np.random.default_rng(seed)
    H, W = 37, 53
    psi = r.uniform(-np.pi, np.pi, size=(H, W)).astype(np.float32)
    psi[r.uniform(size=(H, W)) < 0.35] = np.nan           # many small components and ties
    nm = np.isnan(psi)
    u = (np.where(nm, 0, psi) + 2 * np.pi * r.integers(-3, 4) + r.normal(0, 0.4, size=(H, W))).astype(np.float32)
    for size in (1, 4, 30):
        a = epilogue(u.copy(), psi, nm, size)
        b = post.epilogue(u.copy(), psi, nm, size)
        np.testing.assert_array_equal(a[1], b[1])
        np.testing.assert_allclose(a[0], b[0], atol=2e-6, equal_nan=True)


def test_epilogue_nonfinite_warns():
    from core_amd.postprocess import epilogue
    psi = np.zeros((4, 4), np.float32)
    u = np.zeros((4, 4), np.float32)
    u[1, 1] = np.inf
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        assert epilogue(u, psi, np.isnan(psi), 30) is None
    assert any(issubclass(x.category, RuntimeWarning) for x in w)


def test_synthetic_generator_is_deterministic_and_banded():
    from core_amd import synthetic as S
    wl = S.Workload(70, 90, 5, ramp_max=40.0, fault=True, mask_frac=0.1)
    a = S.generate(wl, band_rows=2048)
    b = S.generate(wl, band_rows=7)           # banded draws = one whole-field draw
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    psi, coh, true = a
    assert np.nanmax(np.abs(psi)) <= np.pi + 1e-6
    assert np.isnan(psi).any() and (coh[np.isnan(psi)] == 0).all()
    assert S.stack_shapes()[:2] == [(1419, 5954), (1131, 4636)]


def test_plan_override_validation_and_roundtrip(lib):
    try:
        fl, rad, ept = lib.plan_describe(8192)
        assert fl == 8192 and np.prod(rad) == 8192 and ept == 16
        lib.plan_override(8192, 8192, [8, 8, 8, 8, 2], 24)
        assert lib.plan_describe(8192) == (8192, [8, 8, 8, 8, 2], 24)
        lib.plan_override(8192, 8192, [16, 16, 16, 2], 12)   # 12 per thread on up to 1024 threads
        assert lib.plan_describe(8192) == (8192, [16, 16, 16, 2], 12)
        assert lib.plan_describe(12288)[2] == 12              # the default above 8192
        b0 = lib.workspace_size([(16, 4999)])
        lib.plan_override(4999, 12288, [], 0)               # a longer Bluestein length
        assert lib.plan_describe(4999)[0] == 12288
        assert lib.workspace_size([(16, 4999)]) > b0        # bigger chirp-kernel table
        for bad in [(8192, 8192, [16, 16, 16], 0),           # product != fft_len
                    (8192, 8192, [16, 16, 16, 2], 32),       # ept
                    (8192, 8192, [16, 16, 7, 2], 0),         # radix
                    (4999, 9000, [], 0),                     # Bluestein length < 2N - 1 ... and not 5-smooth
                    (4999, 4999, [], 0),                     # prime length without Bluestein
                    (12289, 30000, [], 0),                   # beyond the LDS-resident lengths
                    (8192, 12288, [], 16)]:                  # ept 16 with a length > 8192
            with pytest.raises(ValueError):
                lib.plan_override(*bad)
        lib.plan_override(8192, 0)                          # remove one override
        assert lib.plan_describe(8192) == (8192, [16, 16, 16, 2], 16)
    finally:
        lib.plan_reset()
    assert lib.plan_describe(4999)[0] == 10240   # default: lowest M * (passes + 1)


def test_autotune_candidates_are_valid(lib):
    from core_amd import autotune as A
    try:
        for n in [40, 127, 1331, 1500, 2048, 4999, 5000, 6000, 8192, 12288]:
            cands = A.candidates(n)
            assert cands, n
            for fl, rad, ept in cands:
                assert int(np.prod(rad)) == fl and (fl == n or fl >= 2 * n - 1) and fl <= 12288
                lib.plan_override(n, fl, rad, ept)   # the library accepts every candidate
                assert lib.plan_describe(n) == (fl, rad, ept)
    finally:
        lib.plan_reset()


def test_column_strategies(lib):
    """Which column pass each image gets, from its own shape.  Automatic mode: the register-stage
    slab kernels for heights from 256 on whose split is a power-of-two pair (one launch per split);
    tridiagonal solves for every other height (the batched-LDS slab kernels and the transposed
    column DCTs are slower).
    Slab mode: every 5-smooth height within the block limits (register kernels where the split
    allows them)."""
    d = lib.describe_columns
    try:
        lib.set_column_mode(lib.COLUMN_AUTO)
        assert d([(8192, 8192)]) == ["slab_register"]
        assert d([(8192, 70), (8192, 3)]) == ["slab_register"] * 2
        assert d([(256, 256)]) == ["slab_register"]
        assert d([(8192, 70), (1000, 7)]) == ["slab_register", "tridiag"]  # 1000: no power-of-two split
        assert d([(8192, 70), (2048, 9), (8192, 5)]) == ["slab_register"] * 3  # two splits, two launches
        assert d([(2048, 9), (8192, 70)]) == ["slab_register"] * 2
        assert d([(6000, 6000)]) == ["tridiag"]
        assert d([(96, 130)]) == ["tridiag"]                       # below the slab threshold
        assert d([(127, 211)]) == ["tridiag"]                      # Bluestein columns
        assert d([(12800, 8)]) == ["tridiag"]                      # long columns: tridiagonal solves
        assert d([(1000, 12500)]) == ["tridiag"]
        assert d([(8192, 70), (127, 211), (12800, 8)]) == ["slab_register", "tridiag", "tridiag"]
        lib.set_column_mode(lib.COLUMN_SLAB)
        assert d([(64, 48), (96, 130)]) == ["slab", "slab"]
        assert d([(8192, 70), (1000, 7)]) == ["slab_register", "slab"]  # 1000 = 40 x 25: batched LDS
        assert d([(12800, 8)]) == ["slab"]                         # 12800 = 160 x 80
        assert d([(40000, 64)]) == ["slab"]                        # 40000 = 250 x 160
        assert d([(177147, 8)]) == ["tridiag"]                     # 3^11: no split within the block limits
        assert d([(1000, 12500)]) == ["slab"]                      # long rows write T in slab order too
        assert d([(127, 211)]) == ["tridiag"]                      # no slab split: automatic choice
        lib.set_column_mode(lib.COLUMN_TRANSPOSE)
        assert d([(8192, 8192), (12800, 8), (13001, 9)]) == ["transpose", "long", "long"]
        lib.set_column_mode(lib.COLUMN_DIRECT)
        assert d([(127, 211), (8192, 70), (12800, 8)]) == ["direct", "direct", "long"]
        lib.set_column_mode(lib.COLUMN_TRIDIAG)
        assert d([(127, 211), (8192, 70), (96, 130), (12800, 8)]) == ["tridiag"] * 4
        lib.set_column_mode(lib.COLUMN_AUTO)
        # columns with a prime factor above 5, above 12288, or short: tridiagonal solves
        assert d([(13001, 9), (7001, 9), (12800, 8), (9, 7001)]) == ["tridiag"] * 4
    finally:
        lib.set_column_mode(lib.COLUMN_AUTO)


def test_column_mode_validation(lib):
    with pytest.raises(ValueError):
        lib.set_column_mode(5)
    with pytest.raises(ValueError):
        lib.set_column_mode(-1)
    lib.set_column_mode(lib.COLUMN_TRIDIAG)
    lib.set_column_mode(lib.COLUMN_SLAB)
    lib.set_column_mode(lib.COLUMN_DIRECT)
    lib.set_column_mode(lib.COLUMN_AUTO)


def test_memory_groups_split_in_order(lib):
    from core_amd.unwrap2d import memory_groups
    shapes = [(100, 200), (300, 50), (64, 64), (500, 500), (10, 10)]
    one = [lib.workspace_size([s]) + 3 * s[0] * ((s[1] + 63) // 64 * 64) * 4 for s in shapes]
    assert memory_groups(shapes, 10 ** 12) == [[0, 1, 2, 3, 4]]
    g = memory_groups(shapes, one[3] - 1)      # image 3 alone exceeds the budget: its own group
    assert [i for grp in g for i in grp] == list(range(5))
    assert [3] in g
    for grp in g:
        assert len(grp) == 1 or sum(one[i] for i in grp) <= one[3] - 1
