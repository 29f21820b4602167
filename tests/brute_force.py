"""Hierarchy-free point-in-spherical-polygon check (numpy).

Independent of the oracle: the label of p is the crossing parity of the
great-circle path r -> m -> p against every polygon edge, where r is a point
known to be inside (just left of the midpoint of a long outer-ring edge,
the side the reference treats as inside after orientation normalisation) and
m is the normalised midpoint of r and p.  Arc intersection uses the exact
geometric test (both straddle conditions and the crossing point on both
arcs), so it does not share the reference's epsilon rules.  Also returns the
angular distance from p to the nearest edge arc so callers can skip points in
the epsilon band where the reference's tolerance decides the label.
"""
import numpy as np


def _unit(x):
    return x / np.linalg.norm(x, axis=-1, keepdims=True)


def _on_arc(y, a, b, n):
    # y on the minor arc a->b (n = a x b): a x y and y x b both along n
    return (np.einsum("...k,...k->...", np.cross(a, y), n) >= 0) & \
           (np.einsum("...k,...k->...", np.cross(y, b), n) >= 0)


def arc_cross_count(s, t, a, b):
    """Crossings of arcs s[q]->t[q] with every edge a[e]->b[e]: int[q]."""
    n1 = np.cross(s, t)                      # q,3
    n2 = np.cross(a, b)                      # e,3
    sa = a @ n1.T                            # e,q
    sb = b @ n1.T
    ss = s @ n2.T                            # q,e
    st = t @ n2.T
    straddle = (np.sign(sa.T) * np.sign(sb.T) < 0) & (np.sign(ss) * np.sign(st) < 0)
    x = np.cross(n1[:, None, :], n2[None, :, :])  # q,e,3
    x = _unit(x)
    on1 = _on_arc(x, s[:, None, :], t[:, None, :], n1[:, None, :])
    x = np.where(on1[..., None], x, -x)
    ok = _on_arc(x, a[None], b[None], n2[None]) & _on_arc(x, s[:, None, :], t[:, None, :], n1[:, None, :])
    return (straddle & ok).sum(axis=1)


def arc_distance(p, a, b):
    """Angular distance from points p[q] to the nearest arc a[e]->b[e]."""
    n = _unit(np.cross(a, b))
    d = p @ n.T
    proj = _unit(p[:, None, :] - d[..., None] * n[None])
    inside = _on_arc(proj, a[None], b[None], n[None])
    dplane = np.arcsin(np.clip(np.abs(d), 0, 1))
    da = np.arccos(np.clip(p @ a.T, -1, 1))
    db = np.arccos(np.clip(p @ b.T, -1, 1))
    return np.where(inside, dplane, np.minimum(da, db)).min(axis=1)


def labels(pv, pe, pts, chunk=1000):
    """Returns (labels 2=IN/1=OUT, distance to nearest edge) for pts[q,3]."""
    a = pv[pe[:, 0]]
    b = pv[pe[:, 1]]
    # reference point: just left of the midpoint of a long edge of ring 0
    # (edge 0 starts ring 0; ring 0 ends where an edge returns to its first
    # vertex), offset by a fraction of that edge's length
    ring0_end = int(np.nonzero(pe[:, 1] == pe[0, 0])[0][0]) + 1
    lens = np.linalg.norm(b[:ring0_end] - a[:ring0_end], axis=1)
    r = None
    for e in np.argsort(-lens)[:32]:
        n0 = _unit(np.cross(a[e], b[e]))
        off = 1e-3 * lens[e]
        cand = _unit(_unit(a[e] + b[e]) + off * n0)
        if arc_distance(cand[None], a, b)[0] > 0.25 * off:
            r = cand
            break
    if r is None:
        raise ValueError("no clear reference point")
    out = np.zeros(len(pts), dtype=np.int32)
    dist = np.zeros(len(pts))
    for s in range(0, len(pts), chunk):
        p = pts[s:s + chunk]
        m = r[None] + p
        bad = np.linalg.norm(m, axis=1) < 1e-6
        m[bad] = np.cross(r, [0.0, 0.0, 1.0]) if abs(r[2]) < 0.9 else np.cross(r, [1.0, 0.0, 0.0])
        m = _unit(m)
        rr = np.repeat(r[None], len(p), axis=0)
        cnt = arc_cross_count(rr, m, a, b) + arc_cross_count(m, p, a, b)
        out[s:s + chunk] = np.where(cnt % 2 == 0, 2, 1)
        dist[s:s + chunk] = arc_distance(p, a, b)
    return out, dist
