"""The reference's main loop with penalty contact (src/mfree_iwf.cpp:84-97)
run on the CPU oracle: the checker of the contact tests."""
import numpy as np

import _oracle as O
from mfree_iwf_amd import METHOD_TL_COROTATIONAL, METHOD_TL_POTENTIAL, METHOD_UL_GODUNOV, METHOD_UL_MONAGHAN

# scene kind -> (library method, oracle method / TL variant)
KIND = {"monaghan": (METHOD_UL_MONAGHAN, O.METHOD_MONAGHAN), "godunov": (METHOD_UL_GODUNOV, O.METHOD_GODUNOV),
        "corot": (METHOD_TL_COROTATIONAL, O.TL_COROTATIONAL), "potential": (METHOD_TL_POTENTIAL, O.TL_POTENTIAL)}


class OracleBodies:
    """The main loop on the oracle: each body's RK4 step with its contact
    forces, then outlines, then contact_apply_body_to_boundary(i, j) for
    every ordered pair."""

    def __init__(self, bodies, kind):
        self.kind = kind
        self.ps = [b.particles.copy() for b in bodies]
        self.prm = [b.params for b in bodies]
        self.hulls = [O.convex_hull(p.x, p.y) for p in self.ps]
        self.fc = [(np.zeros(p.n), np.zeros(p.n)) for p in self.ps]
        self.time = [0.0 for _ in self.ps]
        self.tl = None
        if kind in ("corot", "potential"):
            self.tl = [O.TL(p, float(p.h[0]), search=1, contrib_only=1, variant=KIND[kind][1]) for p in self.ps]

    def step(self):
        for k in range(len(self.ps)):
            if self.tl is None:
                rc, self.ps[k] = O.ul_step(self.ps[k], self.prm[k], 1, method=KIND[self.kind][1], fc=self.fc[k],
                                           threads=16)
            else:
                rc, self.ps[k], self.time[k] = self.tl[k].step(self.ps[k], self.prm[k], 1, time=self.time[k],
                                                               threads=16, fc=self.fc[k])
            assert rc == O.OK
        pts = [O.boundary_update(h, p.x, p.y) for h, p in zip(self.hulls, self.ps)]
        for i in range(len(self.ps)):
            for j in range(len(self.ps)):
                if i != j:
                    p = self.ps[j]
                    self.fc[j] = O.contact(pts[i], p.x, p.y, p.m, self.prm[j].dt)
        return pts
