"""Loader for the reference fixtures in tests/golden/ (see gen_golden.py)."""
import glob
import json
import os

import numpy as np

from mhd3d_mpi_fortran_amd import ics

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STATE = ("rho", "vx", "vy", "vz", "p", "bi", "bj", "bk")
FACE = {"bi": 1, "bj": 2, "bk": 3}


def cases():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


class Case:
    def __init__(self, name):
        self.name = name
        self.z = np.load(os.path.join(GOLDEN, name + ".npz"))
        self.meta = json.loads(str(self.z["meta"]))
        m = self.meta
        self.grid = ics.Grid(*m["grid"], ng=m["ng"], lo=m["lo"], hi=m["hi"])

    def full_state(self, prefix="ic"):
        """ghost-inclusive arrays with the owned values set and ghosts zero."""
        g = self.grid
        shapes = {0: (g.nx, g.ny, g.nz), 1: (g.nx + 1, g.ny, g.nz), 2: (g.nx, g.ny + 1, g.nz),
                  3: (g.nx, g.ny, g.nz + 1)}
        st = {}
        for v in STATE:
            f = FACE.get(v, 0)
            a = np.zeros(shapes[f], order="F")
            g.interior(a, f)[...] = self.z["%s_%s" % (prefix, v)]
            st[v] = a
        return st

    def owned(self, st, v):
        return self.grid.interior(st[v], FACE.get(v, 0))

    def frame(self, n, v):
        return self.z["f%d_%s" % (n, v)]

    def yee(self):
        """get_eijk_Yee on the initial state (the reference's alternative EMF)"""
        z = self.z
        return dict(ei=z["yee_ei"], ej=z["yee_ej"], ek=z["yee_ek"])

    def ops(self):
        z = self.z
        return dict(tcfl=float(z["ops_tcfl"]), ei=z["ops_ei"], ej=z["ops_ej"], ek=z["ops_ek"],
                    flux=[z["ops_flux%d" % q] for q in range(24)])
