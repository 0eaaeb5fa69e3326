"""Small periodic systems built in numpy for oracle checks (test helper).

Independent of the product's harness: ghosts are periodic images x + s L
with one rounding per coordinate (what LAMMPS's border exchange produces),
and the full neighbour list is a brute-force O(N^2) search.
"""
import itertools

import numpy as np


class System:
    """Arrays in LAMMPS layout: x, v, f are [nall][3]; ghosts follow locals."""

    def __init__(self, x, v, types, box_lo, box_hi, cut, masses, tags=None, mask=None):
        self.box_lo = np.asarray(box_lo, float)
        self.box_hi = np.asarray(box_hi, float)
        L = self.box_hi - self.box_lo
        n = len(x)
        gx, owner, shift = [], [], []
        for s in itertools.product((-1, 0, 1), repeat=3):
            if s == (0, 0, 0):
                continue
            sv = np.array(s, float) * L
            for i in range(n):
                p = x[i] + sv
                if np.all(p >= self.box_lo - cut) and np.all(p < self.box_hi + cut):
                    gx.append(p); owner.append(i); shift.append(sv)
        self.nlocal = n
        self.nghost = len(gx)
        self.ghost_owner = np.array(owner, dtype=np.int32)
        self.ghost_shift = np.array(shift, float).reshape(-1, 3)
        self.x = np.ascontiguousarray(np.vstack([x, np.array(gx).reshape(-1, 3)]), dtype=float)
        self.v = np.ascontiguousarray(np.vstack([v, np.asarray(v)[self.ghost_owner]]), dtype=float)
        self.f = np.zeros_like(self.x)
        t = np.asarray(types, dtype=np.int32)
        self.type = np.ascontiguousarray(np.concatenate([t, t[self.ghost_owner]]), dtype=np.int32)
        m = np.ones(n, np.int32) if mask is None else np.asarray(mask, np.int32)
        self.mask = np.ascontiguousarray(np.concatenate([m, m[self.ghost_owner]]), dtype=np.int32)
        tg = np.arange(1, n + 1, dtype=np.int32) if tags is None else np.asarray(tags, np.int32)
        self.tag = np.ascontiguousarray(np.concatenate([tg, tg[self.ghost_owner]]), dtype=np.int32)
        self.masses = masses
        self.cut = cut
        self.build_list()

    def build_list(self):
        n = self.nlocal
        c2 = self.cut * self.cut
        off = [0]
        idx = []
        for i in range(n):
            d = self.x - self.x[i]
            r2 = (d * d).sum(1)
            js = np.nonzero(r2 < c2)[0]
            js = js[js != i]
            idx.extend(js.tolist())
            off.append(len(idx))
        self.nbr_off = np.array(off, dtype=np.int32)
        self.nbr_idx = np.array(idx, dtype=np.int32)

    def update_ghosts(self):
        self.x[self.nlocal:] = self.x[self.ghost_owner] + self.ghost_shift
        self.v[self.nlocal:] = self.v[self.ghost_owner]


def fcc(cells, a):
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    pts = []
    for k in range(cells):
        for j in range(cells):
            for i in range(cells):
                for b in base:
                    pts.append((np.array([i, j, k]) + b) * a)
    return np.array(pts)
