"""Component linking (the default `unwrap2d(conncomp=False)` path, Stack_unwrap2d.py:293-454):
the package's implementation (core_amd/linking.py) against the brute-force restatement
(oracle/linking.py), and both against what the mathematics fixes.

Pins: the offset program's optimum equals exhaustive enumeration on small graphs (and the closed
form k_j = k_i - k_e on consistent trees); nearest boundary points equal the brute-force minimum
over all pixel pairs; hulls equal scipy's; the two-component closed form of the reference's
formulas (k_e = round((mean_i - mean_j) / 2 pi), k_i - k_j = k_e, shift +k 2 pi); a planted sign
flip is caught.  OpenCV's contour polylines (nearest points between pixel centres) and OR-tools'
choice among tied optima are not reproduced: parity unpinned for those.
"""
import numpy as np
import pytest
from scipy.spatial import ConvexHull

from core_amd import linking as P
from oracle import linking as O


def _tree_edges(rng, n):
    edges = []
    for j in range(1, n):
        i = int(rng.integers(0, j))
        edges.append((i, j, int(rng.integers(-2, 3)), int(rng.integers(1, 1000))))
    return edges


def test_offsets_closed_form_on_consistent_trees():
    rng = np.random.default_rng(0)
    for _ in range(20):
        n = int(rng.integers(2, 6))
        edges = _tree_edges(rng, n)
        k = [0] * n
        for i, j, ke, _w in edges:        # parents come first: k_j = k_i - k_e
            k[j] = k[i] - ke
        if max(abs(v) for v in k) > 3:
            continue
        kb, cost, _ = O.offsets_bruteforce(n, edges)
        assert cost == 0 and kb == k
        assert P.solve_offsets(n, edges) == k


def test_offsets_milp_equals_enumeration():
    """Random graphs with cycles of inconsistent offsets (|k_e| <= 1, so every optimum lies in the
    enumerated window of +-6): the MILP reaches the enumerated optimum, and the same offsets where
    that optimum is unique."""
    rng = np.random.default_rng(1)
    checked = 0
    for _ in range(60):
        n = int(rng.integers(2, 6))
        edges = [(i, j, int(np.clip(ke, -1, 1)), w) for i, j, ke, w in _tree_edges(rng, n)]
        for _e in range(int(rng.integers(0, 4))):
            i, j = rng.choice(n, 2, replace=False)
            edges.append((int(i), int(j), int(rng.integers(-1, 2)), int(rng.integers(1, 1000))))
        kb, cost, n_opt = O.offsets_bruteforce(n, edges)
        km = P.solve_offsets(n, edges)
        assert sum(w * abs(km[i] - km[j] - ke) for i, j, ke, w in edges) == cost
        if n_opt == 1:
            assert km == kb
            checked += 1
    assert checked > 20


def test_offsets_infeasible_returns_none():
    # |k_i - k_j - k_e| <= 200 cannot hold with k in [-100, 100] when |k_e| > 400
    assert P.solve_offsets(2, [(0, 1, 450, 10)]) is None


def _blobs(rng, H, W, n):
    """Labels of n rectangles separated by gaps, numbered by size (1 = largest)."""
    lab = np.zeros((H, W), np.int32)
    rects = []
    for _ in range(200):
        if len(rects) == n:
            break
        h, w = rng.integers(4, H // 2), rng.integers(4, W // 2)
        y, x = rng.integers(0, H - h), rng.integers(0, W - w)
        if np.any(lab[max(0, y - 2):y + h + 2, max(0, x - 2):x + w + 2]):
            continue
        rects.append((h * w, y, x, h, w))
        lab[y:y + h, x:x + w] = -1
    lab[:] = 0
    for k, (_, y, x, h, w) in enumerate(sorted(rects, key=lambda r: -r[0])):
        lab[y:y + h, x:x + w] = k + 1
    return lab


def test_nearest_boundary_points_and_hulls():
    rng = np.random.default_rng(2)
    for _ in range(10):
        lab = _blobs(rng, 60, 70, 4)
        n = int(lab.max())
        comps = [(i, int((lab == i).sum())) for i in range(1, n + 1)]
        phase = rng.normal(size=lab.shape).astype(np.float32)
        conns = P.find_connections(lab, phase, comps)
        for a, b, pa, pb, d, *_ in conns:
            A = np.argwhere(lab == comps[a][0]).astype(float)
            B = np.argwhere(lab == comps[b][0]).astype(float)
            dmin = np.sqrt(((A[:, None] - B[None]) ** 2).sum(-1)).min()
            assert abs(d - dmin) < 1e-12 and abs(np.hypot(pa[0] - pb[0], pa[1] - pb[1]) - dmin) < 1e-12
        for i in range(1, n + 1):
            pts = O.boundary_pixels(lab, i)[:, ::-1]
            if len(np.unique(pts, axis=0)) >= 3:
                ref = ConvexHull(pts)
                assert set(map(tuple, O.hull(pts))) == set(map(tuple, pts[ref.vertices]))
                assert set(map(tuple, P._hull(pts))) == set(map(tuple, pts[ref.vertices]))


def _two_components(offset_cycles):
    H, W = 40, 60
    y, x = np.mgrid[0:H, 0:W]
    u = (0.05 * x + 0.02 * y).astype(np.float32)
    lab = np.zeros((H, W), np.int32)
    lab[:, :25] = 1
    lab[:, 30:] = 2           # a 5-column gap
    u[:, 30:] += np.float32(2 * np.pi * offset_cycles)
    u[lab == 0] = np.nan
    return u, lab


@pytest.mark.parametrize("m", [1, -2, 0])
def test_two_component_closed_form(m):
    """Component 2 carries an extra 2 pi m: the reference's formulas give k_e = round((mean_1 -
    mean_2) / 2 pi) = -m over the two windows at the nearest boundary points, the program sets
    k_1 - k_2 = k_e with k_1 = 0, so component 2 is shifted by +2 pi m (the offset doubles: the
    reference's sign convention, restated as written)."""
    u, lab = _two_components(m)
    out = P.link_components(u, lab, conncomp_size=10)
    exp = u.copy()
    exp[lab == 2] += m * 2 * np.pi
    assert np.array_equal(out, exp, equal_nan=True)
    assert np.array_equal(O.link(u, lab, conncomp_size=10), exp, equal_nan=True)


def test_planted_sign_flip_is_caught(monkeypatch):
    u, lab = _two_components(1)
    exp = P.link_components(u, lab, conncomp_size=10)
    orig = O.connections

    def flipped(*a, **k):
        return [(i, j, pi, pj, d, -ke, c, -dl) for i, j, pi, pj, d, ke, c, dl in orig(*a, **k)]
    monkeypatch.setattr(O, "connections", flipped)
    assert not np.array_equal(O.link(u, lab, conncomp_size=10), exp, equal_nan=True)


@pytest.mark.parametrize("seed", range(6))
def test_link_matches_bruteforce_restatement(seed):
    """Multi-component grids with random 2 pi offsets, small components, a gap limit and a
    blocking component between others: the package equals the brute-force restatement."""
    rng = np.random.default_rng(10 + seed)
    lab = _blobs(rng, 64, 80, 5)
    n = int(lab.max())
    y, x = np.mgrid[0:64, 0:80]
    u = (0.04 * x - 0.03 * y + 0.2 * rng.normal(size=lab.shape)).astype(np.float32)
    for i in range(1, n + 1):
        u[lab == i] += np.float32(2 * np.pi * int(rng.integers(-2, 3)))
    u[lab == 0] = np.nan
    for size, gap in [(10, None), (60, None), (10, 12.0)]:
        a = P.link_components(u, lab, conncomp_size=size, conncomp_gap=gap, conncomp_linksize=5)
        info = {}
        b = O.link(u, lab, conncomp_size=size, gap=gap, linksize=5, info=info)
        if info.get('n_opt', 1) == 1:     # ties: OR-tools and HiGHS may pick different optima
            assert np.array_equal(a, b, equal_nan=True), (seed, size, gap)
        else:
            assert np.array_equal(np.isnan(a), np.isnan(b)), (seed, size, gap)
    # fewer than two components: filter only
    one = np.where(lab == 1, u, np.nan).astype(np.float32)
    assert np.array_equal(P.link_components(one, np.where(lab == 1, 1, 0), 10), one, equal_nan=True)
