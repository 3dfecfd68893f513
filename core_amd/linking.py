"""Linking of unwrapped components by integer 2 pi offsets: the default path of the reference's
`unwrap2d(conncomp=False)` (insardev/insardev/Stack_unwrap2d.py:293-454, `_link_components`).

For one 2-D grid with the size-ordered IRLS labels (1 = largest):
  1. components below max(conncomp_size, 4) pixels become NaN (Stack_unwrap2d.py:343-375);
  2. connections between the remaining components (utils_unwrap2d.py:195-444,
     `find_component_connections_fast`): per component, its `conncomp_linkcount` nearest
     neighbours by convex-hull distance; per neighbour pair, the nearest points of their
     boundaries; a pair is kept when the segment between those points touches no third
     component's convex hull (and is no longer than `conncomp_gap`); pairs are taken in order of
     distance, at most `conncomp_linkcount` per component; each gets the phase offset of the
     means of its two `conncomp_linksize`-square windows (pixels of the own label only),
     k = round(delta / 2 pi), confidence = max(0, 1 - 2 |delta / 2 pi - k|);
  3. the integer program of utils_unwrap2d.py:886-1040 (`connect_components_ilp_fast`):
     minimise sum_e W_e |k_i - k_j - k_e| over integers k in [-100, 100] with k_0 = 0 and
     |k_i - k_j - k_e| <= 200, W_e = int(1000 conf_e sqrt(min(size_i, size_j)) / (sqrt(d_e) + 1));
     components not reachable from component 0 through connections keep k = 0; every component
     is shifted by k 2 pi.

Boundaries are the component's pixels with a 4-neighbour outside it (the pixel centres the
reference's OpenCV contours run through); distances are between those pixel sets.  The integer
program is solved exactly with scipy's HiGHS MILP (the reference uses OR-tools CP-SAT).  Between
equally good offset sets, and for nearest points that the contour polylines place between pixel
centres, the two may choose differently.
"""
from __future__ import annotations

import numpy as np
import scipy.ndimage as ndi
from scipy.optimize import Bounds, LinearConstraint, milp
from scipy.spatial import ConvexHull, cKDTree

__all__ = ["link_components", "find_connections", "solve_offsets"]

TWO_PI = 2 * np.pi


def _boundary_points(labels, comp_label, sl):
    """(row, col) of the component's pixels that have a 4-neighbour outside it."""
    m = labels[sl] == comp_label
    p = np.pad(m, 1)
    inner = p[1:-1, 1:-1] & p[:-2, 1:-1] & p[2:, 1:-1] & p[1:-1, :-2] & p[1:-1, 2:]
    r, c = np.nonzero(m & ~inner)
    return np.stack([r + sl[0].start, c + sl[1].start], axis=1).astype(np.float64)


def _hull(points):
    """Convex hull vertices (counter-clockwise) of (x, y) points; the points themselves (at most
    two distinct ones) when they are collinear."""
    u = np.unique(points, axis=0)
    if len(u) >= 3:
        try:
            h = ConvexHull(u)
            return u[h.vertices]
        except Exception:  # collinear: the two extreme points
            pass
    if len(u) <= 2:
        return u
    d = u - u[0]
    axis = d[np.argmax(np.abs(d).sum(1))]
    t = d @ axis
    return u[[int(np.argmin(t)), int(np.argmax(t))]]


def _edges(poly):
    n = len(poly)
    if n == 1:
        return [(poly[0], poly[0])]
    if n == 2:
        return [(poly[0], poly[1])]
    return [(poly[i], poly[(i + 1) % n]) for i in range(n)]


def _seg_point_dist(a, b, p):
    ab = b - a
    den = float(ab @ ab)
    t = 0.0 if den == 0 else min(1.0, max(0.0, float((p - a) @ ab) / den))
    return float(np.hypot(*(a + t * ab - p)))


def _overlap(A, B, eps=1e-9):
    """True when the convex polygons (vertex arrays; points and segments included) intersect or
    touch: no separating axis among the edge normals of either (and the segment directions)."""
    axes = []
    for P in (A, B):
        for a, b in _edges(P):
            e = b - a
            if e @ e > 0:
                axes.append(np.array([-e[1], e[0]]))
                axes.append(e)
    if not axes:
        return bool(np.hypot(*(A[0] - B[0])) <= eps)
    for ax in axes:
        pa, pb = A @ ax, B @ ax
        if pa.max() < pb.min() - eps or pb.max() < pa.min() - eps:
            return False
    return True


def _poly_dist(A, B):
    if _overlap(A, B):
        return 0.0
    d = min(_seg_point_dist(a, b, p) for a, b in _edges(B) for p in A)
    return min(d, min(_seg_point_dist(a, b, p) for a, b in _edges(A) for p in B))


def find_connections(labels, phase, components, conncomp_gap=None, max_neighbors=30, n_neighbors=5):
    """Connections between components (utils_unwrap2d.py:195-444): list of
    (comp_i, comp_j, (row_i, col_i), (row_j, col_j), distance, k_offset, confidence, delta)."""
    n = len(components)
    if n < 2:
        return []
    shape = labels.shape
    slices = ndi.find_objects(labels)
    bpts, hulls, trees = [], [], []
    for lab, _size in components:
        pts = _boundary_points(labels, lab, slices[lab - 1])
        bpts.append(pts)
        hulls.append(_hull(pts[:, ::-1]))       # (x, y) = (col, row), as the reference's geometry
        trees.append(cKDTree(pts))
    # candidate pairs: per component its max_neighbors nearest components by hull distance, then the
    # nearest boundary points of the pair (symmetric, so each pair once)
    cand = {}
    for i in range(n):
        nbr = sorted(((j, _poly_dist(hulls[i], hulls[j])) for j in range(n) if j != i), key=lambda t: t[1])
        for j, _ in nbr[:max_neighbors]:
            key = (min(i, j), max(i, j))
            if key in cand:
                continue
            a, b = key
            d, idx = trees[b].query(bpts[a], k=1)
            q = int(np.argmin(d))
            cand[key] = (float(d[q]), bpts[a][q], bpts[b][idx[q]])
    valid = []
    for (a, b), (dist, pa, pb) in cand.items():
        if conncomp_gap is not None and dist > conncomp_gap:
            continue
        seg = np.stack([pa[::-1], pb[::-1]])    # (x, y)
        if any(_overlap(seg, hulls[k]) for k in range(n) if k != a and k != b):
            continue
        valid.append((a, b, dist, pa, pb))
    valid.sort(key=lambda t: t[2])
    half = n_neighbors // 2
    count = [0] * n
    out = []
    for a, b, dist, pa, pb in valid:
        if count[a] >= max_neighbors or count[b] >= max_neighbors:
            continue
        means = []
        for comp, (r, c) in ((a, pa), (b, pb)):
            r, c = int(r), int(c)
            y0, y1 = max(0, r - half), min(shape[0], r + half + 1)
            x0, x1 = max(0, c - half), min(shape[1], c + half + 1)
            vals = phase[y0:y1, x0:x1][labels[y0:y1, x0:x1] == components[comp][0]]
            means.append(np.nanmean(vals) if vals.size else np.float32(np.nan))
        if np.isnan(means[0]) or np.isnan(means[1]):
            continue
        delta = means[0] - means[1]
        k = int(np.round(delta / (2 * np.pi)))
        frac = (delta / (2 * np.pi)) - k
        conf = max(0, 1.0 - 2 * abs(frac))
        out.append((a, b, (int(pa[0]), int(pa[1])), (int(pb[0]), int(pb[1])), dist, k, conf, delta))
        count[a] += 1
        count[b] += 1
    return out


def solve_offsets(n, edges):
    """Integer k minimising sum W |k_i - k_j - k_e| (k in [-100, 100], k_0 = 0, |k_i - k_j - k_e|
    <= 200) for edges (i, j, k_e, W); None when infeasible (the reference then links nothing)."""
    m = len(edges)
    nv = n + m
    c = np.concatenate([np.zeros(n), np.array([w for *_, w in edges], dtype=np.float64)])
    rows, lo, hi = [], [], []
    for e, (i, j, ke, _w) in enumerate(edges):
        d = np.zeros(nv)
        d[i] += 1.0
        d[j] -= 1.0
        r1 = d.copy(); r1[n + e] = 1.0     # t + (k_i - k_j) >= k_e
        r2 = -d; r2[n + e] = 1.0           # t - (k_i - k_j) >= -k_e
        rows += [r1, r2, d]
        lo += [ke, -ke, ke - 200]
        hi += [np.inf, np.inf, ke + 200]
    lb = np.concatenate([np.full(n, -100.0), np.zeros(m)])
    ub = np.concatenate([np.full(n, 100.0), np.full(m, np.inf)])
    lb[0] = ub[0] = 0.0
    res = milp(c, constraints=LinearConstraint(np.array(rows), lo, hi), integrality=np.ones(nv),
               bounds=Bounds(lb, ub))
    if res.x is None:
        return None
    return [int(round(v)) for v in res.x[:n]]


def link_components(unwrapped, conncomp, conncomp_size=1000, conncomp_gap=None, conncomp_linksize=5,
                    conncomp_linkcount=30, debug=False):
    """One grid of the reference's `_link_2d` (Stack_unwrap2d.py:326-405): small components to
    NaN, then the remaining components shifted by the integer program's 2 pi multiples.
    Returns a float32 copy."""
    phase = np.asarray(unwrapped, dtype=np.float32)
    if not np.any(~np.isnan(phase)):
        return phase
    labels = np.asarray(conncomp).astype(np.int32)
    min_size = max(conncomp_size, 4)
    n_total = int(labels.max())
    sizes = np.bincount(labels.ravel(), minlength=n_total + 1)
    comps = [(i, int(sizes[i])) for i in range(1, n_total + 1) if sizes[i] >= min_size]
    keep = np.isin(labels, np.array([lab for lab, _ in comps], dtype=np.int32)) | (labels == 0)
    phase = phase.copy()
    phase[~keep] = np.nan
    if len(comps) < 2:
        return phase
    conns = find_connections(labels, phase, comps, conncomp_gap, conncomp_linkcount, conncomp_linksize)
    if debug:
        print(f'  link_components: {len(comps)} components, {len(conns)} connections')
    if not conns:
        return phase
    edges = []
    for a, b, _pa, _pb, dist, k, conf, _delta in conns:
        w = conf * np.sqrt(min(comps[a][1], comps[b][1])) / (np.sqrt(dist) + 1.0)
        edges.append((a, b, k, int(w * 1000)))
    ks = solve_offsets(len(comps), edges)
    if ks is None:
        return phase
    adj = {i: set() for i in range(len(comps))}
    for a, b, *_ in edges:
        adj[a].add(b)
        adj[b].add(a)
    seen, todo = {0}, [0]
    while todo:
        for nb in adj[todo.pop()]:
            if nb not in seen:
                seen.add(nb)
                todo.append(nb)
    for i, (lab, _) in enumerate(comps):
        if i in seen and ks[i] != 0:
            phase[labels == lab] += ks[i] * 2 * np.pi
    return phase
