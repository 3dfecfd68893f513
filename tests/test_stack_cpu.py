"""CPU tests of the stack adapters (core_amd/stack.py): the small-component filter against a
brute-force restatement of Stack_unwrap2d.py:343-375, and the structure of the Dataset adapters
with a minimal xarray stand-in (xarray is not installed here) and the device solve replaced by a
recording stub."""

import numpy as np
import pytest

from core_amd import stack


def brute_filter(u, cc, conncomp_size):
    out = u.copy()
    if np.isnan(u).all():
        return out, 0
    min_size = max(conncomp_size, 4)
    kept = 0
    for lab in range(1, int(cc.max()) + 1):
        n = int((cc == lab).sum())
        if n == 0:
            continue
        if n >= min_size:
            kept += 1
        else:
            out[cc == lab] = np.nan
    return out, kept


def test_filter_small_components_examples():
    cc = np.zeros((6, 8), np.uint16)
    cc[0:2, 0:5] = 1          # 10 px
    cc[3:5, 0:2] = 3          # 4 px
    cc[5, 5:8] = 2            # 3 px
    u = np.arange(48, dtype=np.float32).reshape(6, 8)
    out, kept = stack.filter_small_components(u, cc, conncomp_size=4)
    assert kept == 2
    assert np.isnan(out[5, 5:8]).all()                          # label 2 (3 px) dropped
    assert not np.isnan(out[cc == 1]).any() and not np.isnan(out[cc == 3]).any()
    assert np.array_equal(out[cc == 0], u[cc == 0])             # label 0 pixels kept as they are
    out1, kept1 = stack.filter_small_components(u, cc, conncomp_size=1)  # min size is 4 anyway
    assert kept1 == 2 and np.isnan(out1[5, 5]).all()
    allnan = np.full((3, 3), np.nan, np.float32)
    assert stack.filter_small_components(allnan, np.zeros((3, 3)), 5)[1] == 0


@pytest.mark.parametrize("seed", range(5))
def test_filter_small_components_matches_brute_force(seed):
    r = np.random.default_rng(seed)
    cc = r.integers(0, 9, size=(20, 25)).astype(np.uint16)
    u = r.normal(size=(20, 25)).astype(np.float32)
    u[r.uniform(size=u.shape) < 0.05] = np.nan
    for size in (1, 4, 40, 60, 100):
        a = stack.filter_small_components(u, cc, size)
        b = brute_filter(u, cc, size)
        assert a[1] == b[1]
        np.testing.assert_array_equal(a[0], b[0])


# --- a minimal xarray stand-in: just what the adapters use -------------------------------------

class FakeDataArray:
    def __init__(self, data, dims=(), coords=None, attrs=None):
        self.values = np.asarray(data)
        self.data = self.values
        self.dims = tuple(dims)
        self.coords = dict(coords or {})
        self.attrs = dict(attrs or {})

    def __getattr__(self, k):
        if k in ("y", "x") and k in self.coords:
            return self.coords[k]
        raise AttributeError(k)

    def assign_coords(self, **kw):
        c = dict(self.coords)
        for k, (dim, v) in kw.items():
            c[k] = FakeDataArray(v, (dim,))
        return FakeDataArray(self.values, self.dims, c, self.attrs)

    def chunk(self, spec):
        raise AssertionError("numpy-backed input must not be rechunked")


class FakeDataset:
    def __init__(self, data_vars, attrs=None):
        self._v = dict(data_vars)
        self.data_vars = list(self._v)
        self.attrs = dict(attrs or {})

    def __getitem__(self, k):
        return self._v[k]

    def items(self):
        return self._v.items()


class FakeXr:
    DataArray = FakeDataArray
    Dataset = FakeDataset


def make_burst(n_pairs=3, H=6, W=7, seed=0):
    r = np.random.default_rng(seed)
    y = FakeDataArray(np.arange(H) * 10.0, ("y",))
    x = FakeDataArray(np.arange(W) * 20.0, ("x",))
    ref = FakeDataArray(np.array(["2020-01-0%d" % i for i in range(n_pairs)]), ("pair",))
    ph = FakeDataArray(r.uniform(-3, 3, size=(n_pairs, H, W)).astype(np.float32), ("pair", "y", "x"),
                       {"y": y, "x": x, "ref": ref})
    co = FakeDataArray(r.uniform(0, 1, size=(n_pairs, H, W)).astype(np.float32), ("pair", "y", "x"),
                       {"y": y, "x": x, "ref": ref})
    aux = FakeDataArray(np.ones(n_pairs), ("pair",))
    return FakeDataset({"VV": ph, "BPR": aux}, {"subswath": "IW1"}), FakeDataset({"VV": co}, {})


@pytest.fixture
def stub_solver(monkeypatch):
    calls = []

    def fake_unwrap_stack(phase, weight=None, **kw):
        calls.append((phase.shape, None if weight is None else weight.shape, kw))
        cc = np.ones(phase.shape, np.uint16)
        cc[..., :, -2:] = 2                       # two components of 2 x H px each
        return phase.astype(np.float32) + 1.0, cc

    monkeypatch.setattr(stack, "unwrap_stack", fake_unwrap_stack)
    return calls


def test_unwrap2d_irls_structure(stub_solver):
    ph, co = make_burst()
    unw, cc = stack.unwrap2d_irls({"b1": ph}, {"b1": co}, device="cuda", xr=FakeXr)
    assert list(unw) == ["b1"] and list(cc) == ["b1"]
    u = unw["b1"]["VV"]
    c = cc["b1"]["VV"]
    assert u.dims == ("pair", "y", "x") and u.values.dtype == np.float32 and u.attrs["units"] == "radians"
    assert c.values.dtype == np.uint16 and "units" not in c.attrs
    assert u.coords["y"] is ph["VV"].coords["y"] and "ref" in u.coords and "ref" in c.coords
    assert unw["b1"].attrs == {"subswath": "IW1"} and unw["b1"].data_vars == ["VV"]   # BPR: no y/x
    np.testing.assert_array_equal(u.values, ph["VV"].values + 1.0)
    (shape, wshape, kw), = stub_solver
    assert shape == (3, 6, 7) and wshape == (3, 6, 7)
    # unwrap2d_irls defaults (Stack_unwrap2d.py:753-755)
    assert (kw["max_iter"], kw["tol"], kw["cg_max_iter"], kw["cg_tol"], kw["epsilon"], kw["conncomp_size"]) == \
        (50, 1e-2, 10, 1e-3, 1e-2, 30)


def test_unwrap2d_errors_and_modes(stub_solver):
    ph, co = make_burst()
    with pytest.raises(ValueError):
        stack.unwrap2d({"b": ph}, conncomp_size=4, conncomp_linksize=5, xr=FakeXr)
    unw, cc = stack.unwrap2d({"b": ph}, {"b": co}, conncomp=True, conncomp_size=4, xr=FakeXr)
    assert stub_solver[-1][2]["conncomp_size"] == 4
    out = stack.unwrap2d({"b": ph}, conncomp_size=10, conncomp_linksize=5, xr=FakeXr)
    # label 2 covers 12 px >= 10: both components survive and each pair is linked
    from core_amd.linking import link_components
    u_stub = ph["VV"].values.astype(np.float32) + 1.0
    cc_stub = np.ones(u_stub.shape, np.uint16)
    cc_stub[..., :, -2:] = 2
    for p in range(u_stub.shape[0]):
        assert np.array_equal(out["b"]["VV"].values[p], link_components(u_stub[p], cc_stub[p], 10, None, 5, 30))
    assert not np.isnan(out["b"]["VV"].values).any()
    out = stack.unwrap2d({"b": ph}, conncomp_size=13, conncomp_linksize=5, xr=FakeXr)
    assert np.isnan(out["b"]["VV"].values[..., -2:]).all()      # the 12 px component is dropped
    with pytest.raises(TypeError):
        stack.unwrap2d_dataset({"not": "a dataset"}, xr=FakeXr)
    res = stack.unwrap2d_dataset(ph, co, conncomp=True, conncomp_size=4, xr=FakeXr)
    assert isinstance(res, tuple) and res[0]["VV"].dims == ("pair", "y", "x")


def test_chunk_windows_follow_map_overlap_boundary_none():
    # 10 x 9 grid, chunks (4, 4): rows [0,4) [4,8) [8,10), cols [0,4) [4,8) [8,9)
    wins = stack.chunk_windows(10, 9, (4, 4), overlap=2)
    assert len(wins) == 9
    core, ext = wins[0]
    assert core == (0, 4, 0, 4) and ext == (0, 6, 0, 6)        # no margin towards the array edge
    core, ext = wins[4]
    assert core == (4, 8, 4, 8) and ext == (2, 10, 2, 9)        # clipped to the array
    # every pixel belongs to exactly one core window
    cover = np.zeros((10, 9), int)
    for (y0, y1, x0, x1), _ in wins:
        cover[y0:y1, x0:x1] += 1
    assert (cover == 1).all()
    # overlap=None still overlaps by 1 pixel; float overlap is a fraction of the chunk
    assert stack.chunk_windows(10, 9, (4, 4), None)[4][1] == (3, 9, 3, 9)
    assert stack.chunk_windows(10, 9, (4, 4), 0.5)[4][1] == (2, 10, 2, 9)
    assert stack.chunk_windows(10, 9, (4, 4), (1, 3))[4][1] == (3, 9, 1, 9)


def test_unwrap_chunked_assembles_trimmed_tiles(monkeypatch):
    """With the solver replaced by a stub that tags each tile, every output pixel comes from the
    core of its own chunk's extended tile at the right offset."""
    calls = []

    def fake_batch(tiles, weights=None, **kw):
        calls.append(kw)
        return [(t + 1000.0 * i, np.zeros(t.shape, np.uint16)) for i, t in enumerate(tiles)]

    monkeypatch.setattr(stack, "irls_unwrap_2d_batch", fake_batch)
    ph = np.arange(2 * 10 * 9, dtype=np.float32).reshape(2, 10, 9)
    out = stack.unwrap_chunked(ph, chunks=(4, 4), overlap=2)
    wins = stack.chunk_windows(10, 9, (4, 4), 2)
    for p in range(2):
        for t, (core, _) in enumerate(wins):
            y0, y1, x0, x1 = core
            np.testing.assert_array_equal(out[p, y0:y1, x0:x1], ph[p, y0:y1, x0:x1] + 1000.0 * (p * len(wins) + t))
    assert len(calls) == 1 and (calls[0]["max_iter"], calls[0]["tol"], calls[0]["cg_max_iter"]) == (50, 1e-2, 10)
