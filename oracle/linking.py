"""Brute-force restatement of the reference's component linking (test infrastructure).

Follows insardev/insardev/Stack_unwrap2d.py:326-405 (`_link_2d`), utils_unwrap2d.py:195-444
(`find_component_connections_fast`) and :886-1040 (`connect_components_ilp_fast`) step by step,
with each geometric or combinatorial step done the slow, obvious way so that it checks the
package's implementation (core_amd/linking.py):
  * boundaries: component pixels with a 4-neighbour outside the component, found by looping over
    the four shifts (the pixel centres OpenCV's contours run through);
  * nearest boundary points: the full pairwise distance matrix;
  * convex hulls: Andrew's monotone chain; hull distance and segment-hull contact by orientation
    tests and point-segment distances;
  * the integer program min sum_e W_e |k_i - k_j - k_e| (k_0 = 0): exhaustive enumeration over a
    window of k values around 0 (small graphs only).
OpenCV's contour polylines (and shapely's nearest points between them) may put a nearest point
between pixel centres, and OR-tools may pick another among equally good offset sets; neither is
reproduced here (parity unpinned for those two choices).
"""
from __future__ import annotations

import itertools

import numpy as np


def boundary_pixels(labels, lab):
    m = labels == lab
    H, W = m.shape
    inner = m.copy()
    for dy, dx in ((-1, 0), (1, 0), (0, -1), (0, 1)):
        sh = np.zeros_like(m)
        ys = slice(max(0, -dy), H - max(0, dy))
        xs = slice(max(0, -dx), W - max(0, dx))
        yd = slice(max(0, dy), H - max(0, -dy))
        xd = slice(max(0, dx), W - max(0, -dx))
        sh[ys, xs] = m[yd, xd]
        inner &= sh
    r, c = np.nonzero(m & ~inner)
    return np.stack([r, c], axis=1).astype(np.float64)


def _cross(o, a, b):
    return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])


def hull(points):
    """Andrew's monotone chain on (x, y) points; collinear input gives its two end points."""
    pts = sorted(set(map(tuple, points)))
    if len(pts) <= 2:
        return np.array(pts, dtype=np.float64)
    lower, upper = [], []
    for p in pts:
        while len(lower) >= 2 and _cross(lower[-2], lower[-1], p) <= 0:
            lower.pop()
        lower.append(p)
    for p in reversed(pts):
        while len(upper) >= 2 and _cross(upper[-2], upper[-1], p) <= 0:
            upper.pop()
        upper.append(p)
    h = lower[:-1] + upper[:-1]
    if len(h) < 3:
        return np.array([pts[0], pts[-1]], dtype=np.float64)
    return np.array(h, dtype=np.float64)


def _segs(P):
    if len(P) == 1:
        return [(P[0], P[0])]
    if len(P) == 2:
        return [(P[0], P[1])]
    return [(P[i], P[(i + 1) % len(P)]) for i in range(len(P))]


def _pt_seg(p, a, b):
    ab = b - a
    den = ab @ ab
    t = 0.0 if den == 0 else np.clip((p - a) @ ab / den, 0.0, 1.0)
    return float(np.linalg.norm(a + t * ab - p))


def _inside(p, P):
    """Point in (or on) a convex polygon with >= 3 counter-clockwise vertices."""
    return all(_cross(a, b, p) >= -1e-9 for a, b in _segs(P))


def _seg_cross(a, b, c, d):
    """Closed segments ab and cd intersect."""
    def orient(p, q, r):
        v = _cross(p, q, r)
        return 0 if abs(v) <= 1e-9 else (1 if v > 0 else -1)

    def on(p, q, r):
        return min(p[0], r[0]) - 1e-9 <= q[0] <= max(p[0], r[0]) + 1e-9 and \
            min(p[1], r[1]) - 1e-9 <= q[1] <= max(p[1], r[1]) + 1e-9
    o1, o2, o3, o4 = orient(a, b, c), orient(a, b, d), orient(c, d, a), orient(c, d, b)
    if o1 != o2 and o3 != o4:
        return True
    return (o1 == 0 and on(a, c, b)) or (o2 == 0 and on(a, d, b)) or (o3 == 0 and on(c, a, d)) or \
        (o4 == 0 and on(c, b, d))


def touches(A, B):
    """Convex polygons (or points / segments) A and B intersect or touch."""
    for a, b in _segs(A):
        for c, d in _segs(B):
            if _seg_cross(a, b, c, d):
                return True
    if len(B) >= 3 and any(_inside(p, B) for p in A):
        return True
    if len(A) >= 3 and any(_inside(p, A) for p in B):
        return True
    return False


def hull_distance(A, B):
    if touches(A, B):
        return 0.0
    return min(min(_pt_seg(p, a, b) for a, b in _segs(B) for p in A),
               min(_pt_seg(p, a, b) for a, b in _segs(A) for p in B))


def connections(labels, phase, comps, gap=None, max_neighbors=30, n_neighbors=5):
    """utils_unwrap2d.py:195-444 with brute-force geometry; comps = [(label, size)]."""
    n = len(comps)
    if n < 2:
        return []
    B = [boundary_pixels(labels, lab) for lab, _ in comps]
    Hs = [hull(b[:, ::-1]) for b in B]
    cand = {}
    for i in range(n):
        order = sorted(((j, hull_distance(Hs[i], Hs[j])) for j in range(n) if j != i), key=lambda t: t[1])
        for j, _ in order[:max_neighbors]:
            a, b = min(i, j), max(i, j)
            if (a, b) in cand:
                continue
            D = np.sqrt(((B[a][:, None, :] - B[b][None, :, :]) ** 2).sum(-1))
            q = np.unravel_index(int(np.argmin(D)), D.shape)
            cand[(a, b)] = (float(D[q]), B[a][q[0]], B[b][q[1]])
    valid = []
    for (a, b), (d, pa, pb) in cand.items():
        if gap is not None and d > gap:
            continue
        seg = np.array([pa[::-1], pb[::-1]])
        if any(touches(seg, Hs[k]) for k in range(n) if k not in (a, b)):
            continue
        valid.append((a, b, d, pa, pb))
    valid.sort(key=lambda t: t[2])
    half = n_neighbors // 2
    cnt = [0] * n
    out = []
    for a, b, d, pa, pb in valid:
        if cnt[a] >= max_neighbors or cnt[b] >= max_neighbors:
            continue
        m = []
        for comp, p in ((a, pa), (b, pb)):
            r, c = int(p[0]), int(p[1])
            win = (slice(max(0, r - half), min(labels.shape[0], r + half + 1)),
                   slice(max(0, c - half), min(labels.shape[1], c + half + 1)))
            vals = phase[win][labels[win] == comps[comp][0]]
            m.append(np.nanmean(vals) if vals.size else np.float32(np.nan))
        if np.isnan(m[0]) or np.isnan(m[1]):
            continue
        delta = m[0] - m[1]
        k = int(np.round(delta / (2 * np.pi)))
        conf = max(0, 1.0 - 2 * abs((delta / (2 * np.pi)) - k))
        out.append((a, b, (int(pa[0]), int(pa[1])), (int(pb[0]), int(pb[1])), d, k, conf, delta))
        cnt[a] += 1
        cnt[b] += 1
    return out


def offsets_bruteforce(n, edges, span=6):
    """min sum W |k_i - k_j - k_e| with k_0 = 0 by enumerating k in [-span, span]^(n-1); returns
    (the lexicographically first optimum, its cost, the number of optima).  edges = [(i, j, k_e, W)].
    Small graphs only: (2 span + 1)^(n - 1) offset sets are scored."""
    if n > 6:
        raise ValueError(f"offsets_bruteforce: {n} components is too many to enumerate")
    ks = np.array(list(itertools.product(range(-span, span + 1), repeat=n - 1)), dtype=np.int64).reshape(-1, n - 1)
    k = np.concatenate([np.zeros((len(ks), 1), np.int64), ks], axis=1)
    cost = np.zeros(len(k), np.int64)
    for i, j, ke, w in edges:
        cost += w * np.abs(k[:, i] - k[:, j] - ke)
    best = int(cost.min())
    first = int(np.argmax(cost == best))
    return [int(v) for v in k[first]], best, int((cost == best).sum())


def link(unwrapped, conncomp, conncomp_size=1000, gap=None, linksize=5, linkcount=30, span=8, info=None):
    """Stack_unwrap2d.py:326-405 for one grid (small graphs: the offsets are enumerated over
    [-span, span]).  `info`, a dict, receives the number of optimal offset sets ('n_opt')."""
    phase = np.asarray(unwrapped, dtype=np.float32).copy()
    if not np.any(~np.isnan(phase)):
        return phase
    labels = np.asarray(conncomp).astype(np.int32)
    min_size = max(conncomp_size, 4)
    n_total = int(labels.max())
    sizes = np.bincount(labels.ravel(), minlength=n_total + 1)
    comps = [(i, int(sizes[i])) for i in range(1, n_total + 1) if sizes[i] >= min_size]
    keep = np.isin(labels, [lab for lab, _ in comps]) | (labels == 0)
    phase[~keep] = np.nan
    if len(comps) < 2:
        return phase
    conns = connections(labels, phase, comps, gap, linkcount, linksize)
    if not conns:
        return phase
    edges = [(a, b, k, int(conf * np.sqrt(min(comps[a][1], comps[b][1])) / (np.sqrt(d) + 1.0) * 1000))
             for a, b, _, _, d, k, conf, _ in conns]
    ks, _, n_opt = offsets_bruteforce(len(comps), edges, span)
    if info is not None:
        info['n_opt'] = n_opt
        info['k'] = ks
    reach, todo = {0}, [0]
    while todo:
        x = todo.pop()
        for a, b, *_ in edges:
            for u, v in ((a, b), (b, a)):
                if u == x and v not in reach:
                    reach.add(v)
                    todo.append(v)
    for i, (lab, _) in enumerate(comps):
        if i in reach and ks[i]:
            phase[labels == lab] += ks[i] * 2 * np.pi
    return phase
