"""Which reference rule makes the oracle's label differ from the geometry.

The reference does not label points by geometry alone: a boundary leaf's
centre is typed by the nearest-ref rule of setBottomNECell / getCPProp
(source/GPISPAHHG.cu:1046-1071, :61-171), empty leaves take types by
propagation (setBottomDECell :1074-1124, setDECell :1292-1363, setIECell
:1366-1392, setTopECell :1395-1419), and a boundary leaf's point gets its
centre's type flipped by the parity of isEEInter crossings along the arc from
the centre (kernel_PISP :1422-1504, isEEInter :351-378).  Given the oracle's
hierarchy and the provenance it records for every cell's type, this module
explains each oracle-vs-geometry disagreement by a root cause, checking every
step of the explanation against exact arc geometry (brute_force.py):

  NE_CENTRE   the leaf (or the cell a propagation chain starts from) is a
              boundary leaf whose centre setBottomNECell typed against the
              geometry;
  ARC_EPS     a boundary leaf with a correctly typed centre whose isEEInter
              parity along centre -> point differs from the exact crossing
              parity (the epsilon rules of isEEInter decide);
  PARITY_EPS  a propagation step across a border (setBottomDECell /
              setDECell parity) whose isInterSeg parity differs from the
              exact crossing parity from the source centre to the border
              midpoint;
  CENTRE_EPS  a centre within the epsilon band of an edge (its geometric
              label is not defined at the reference's tolerance);
  UNEXPLAINED a step that its recorded rule does not produce (a copy whose
              type differs from its source's, an exact parity step that
              does not give the recorded type, an empty leaf crossed by an
              edge): what a wrong oracle would show.
"""
import numpy as np

import brute_force as bf

RULES = {0: "none", 1: "setBottomNECell", 2: "setBottomDECell", 3: "setIECell", 4: "setDECell copy",
         5: "setDECell parity", 6: "setTopECell"}
BAND = 1e-6


class Hierarchy:
    """The oracle's grid, provenance and polygon, with exact-geometry helpers."""

    def __init__(self, o):
        self.o = o
        self.g = o.grid()
        self.cells = self.g["cells"]
        self.bord = self.g["borders"]
        self.v = self.g["v"]
        self.refs = self.g["refs"]
        self.rule, self.src, self.border = o.provenance()
        self.pv, self.pe = o.polygon()
        self._truth = {}

    def centre(self, cid):
        bs, bn = self.cells[cid, 0], self.cells[cid, 1]
        s = self.v[self.bord[bs:bs + bn, 2]].sum(axis=0)
        return s / np.linalg.norm(s)

    def midpoint(self, b):
        m = self.v[self.bord[b, 2]] + self.v[self.bord[b, 3]]
        return m / np.linalg.norm(m)

    def cell_refs(self, cid):
        rs, rn = self.cells[cid, 2], self.cells[cid, 3]
        return self.refs[rs:rs + rn] if rn > 0 else np.zeros(0, np.int32)

    def exact_parity(self, s, t, edges):
        if len(edges) == 0:
            return 0
        a = self.pv[self.pe[edges, 0]]
        b = self.pv[self.pe[edges, 1]]
        return int(bf.arc_cross_count(s[None], t[None], a, b)[0]) % 2

    def truths(self, cids):
        """Geometric label and edge distance of the centres of cells cids."""
        need = [c for c in cids if c not in self._truth]
        if need:
            pts = np.stack([self.centre(c) for c in need])
            lab, dist = bf.labels(self.pv, self.pe, pts, chunk=256)
            for c, l, d in zip(need, lab, dist):
                self._truth[c] = (int(l), float(d))
        return [self._truth[c] for c in cids]

    def root(self, cid, depth=0):
        """Root cause of cell cid's type being wrong (its centre's geometric
        label differs from its type): a category and the cell it lies in."""
        (t, d), = self.truths([cid])
        ty = int(self.cells[cid, 6])
        if d <= BAND:
            return "CENTRE_EPS", cid
        if ty == t:
            return None, cid  # the type is right here
        r = int(self.rule[cid])
        if r == 1:
            return "NE_CENTRE", cid
        if r in (3, 4, 6):  # copies of the source's type
            s = int(self.src[cid])
            if int(self.cells[s, 6]) != ty:
                return "UNEXPLAINED", cid
            if depth > 64:
                return "UNEXPLAINED", cid
            cat, at = self.root(s, depth + 1)
            return (cat, at) if cat else ("UNEXPLAINED", cid)
        if r in (2, 5):  # source type flipped by the crossing parity to a border midpoint
            s, b = int(self.src[cid]), int(self.border[cid])
            ex = self.exact_parity(self.centre(s), self.midpoint(b), self.cell_refs(s))
            st = int(self.cells[s, 6])
            if (st if ex == 0 else 3 - st) != ty:
                return "PARITY_EPS", cid  # the epsilon parity gave the recorded type, the exact one not
            if depth > 64:
                return "UNEXPLAINED", cid
            cat, at = self.root(s, depth + 1)
            return (cat, at) if cat else ("UNEXPLAINED", cid)
        return "UNEXPLAINED", cid

    def explain(self, p, leaf, o_label, g_label):
        """Category of one disagreeing point p (leaf = its oracle leaf)."""
        if leaf < 0:
            return "UNEXPLAINED"
        if self.cells[leaf, 3] > 0:  # boundary leaf: centre type xor isEEInter parity
            (t, d), = self.truths([leaf])
            if d <= BAND:
                return "CENTRE_EPS"
            ty = int(self.cells[leaf, 6])
            if ty != t:
                return "NE_CENTRE"
            ex = self.exact_parity(self.centre(leaf), p, self.cell_refs(leaf))
            if (t if ex == 0 else 3 - t) == g_label:
                return "ARC_EPS"  # exact parity reproduces the geometry; the epsilon rules did not
            return "UNEXPLAINED"
        cat, _ = self.root(leaf)
        return cat or "UNEXPLAINED"  # an empty leaf whose type is right cannot hold a wrong label


def attribute(o, pts, g_lab, g_dist):
    """Categories of the points where the oracle's label differs from the
    geometric one outside the epsilon band: dict category -> count, and the
    per-point list (index, category)."""
    h = Hierarchy(o)
    lab = o.query_xyz(pts)
    leaf = o.query_leaf(pts)
    bad = np.nonzero((lab != g_lab) & (g_dist > BAND))[0]
    out = []
    for i in bad:
        out.append((int(i), h.explain(pts[i], int(leaf[i]), int(lab[i]), int(g_lab[i]))))
    counts = {}
    for _, c in out:
        counts[c] = counts.get(c, 0) + 1
    return counts, out, h
