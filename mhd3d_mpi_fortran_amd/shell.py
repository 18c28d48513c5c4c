"""Grids given by their node coordinates (the general metric of
``mhd_set_metric_nodes``) and the spherical-shell set-up built on them.

The reference's grid is the tensor product of init_geom (mhd3d3.f90:3379-3515)
only; these grids have no counterpart there.  A ``NodeGrid`` carries the
physical coordinates of its (nx+1)(ny+1)(nz+1) nodes, ghosts included, node
(i,j,k) the lower corner of cell (i,j,k).  ``gpu.Solver`` hands them to the
library, which builds the face normals and areas, cell volumes, cell-B
coefficients and edge factors itself; the numpy geometry below (same
definitions) serves initial states and diagnostics:

  * face area vector of a quadrilateral face = half the cross product of its
    diagonals, oriented along increasing index (exact vector area of the
    face's boundary loop, so the six faces of a cell close exactly);
  * cell volume from the divergence theorem, V = (1/3) sum_f S_f . x_f with
    x_f the mean of the face's corners;
  * div B of a cell = (sum over its faces of B.n A, outward) / V.

Spherical shell: computational (i, j, k) = (r, theta, phi), r log-stretched,
theta kept away from the axis (no polar boundary treatment), phi periodic.
The dipole-plus-inflow state takes its face B from a vector potential's edge
circulations, so the discrete divergence starts at round-off.
"""
import math

import numpy as np

from . import ics


class NodeGrid(ics.Grid):
    """index space of ``ics.Grid`` plus node coordinates X, Y, Z, each of
    shape (nx+1, ny+1, nz+1).  dx, dy, dz are unit spacings: the general
    metric does not use them."""

    def __init__(self, nx0, ny0, nz0, ng, X, Y, Z):
        super().__init__(nx0, ny0, nz0, ng=ng, lo=(0., 0., 0.), hi=(float(nx0), float(ny0), float(nz0)))
        shp = (self.nx + 1, self.ny + 1, self.nz + 1)
        self.nodes = tuple(np.asfortranarray(np.asarray(a, dtype=np.float64)) for a in (X, Y, Z))
        for a in self.nodes:
            if a.shape != shp:
                raise ValueError("node array shape %s, expected %s" % (a.shape, shp))
        self._geo = None

    def physical(self, x, y, z):
        raise NotImplementedError("a node grid has no closed-form map; use cell_centres_phys()")

    def face_normals(self):
        raise NotImplementedError("a node grid's normals vary per face; use geometry()")

    def geometry(self):
        if self._geo is None:
            self._geo = node_geometry(*self.nodes)
        return self._geo

    def cell_centres_phys(self):
        X, Y, Z = self.nodes
        return tuple(_corner_mean(a) for a in (X, Y, Z))


def _corner_mean(a):
    return .125 * (a[:-1, :-1, :-1] + a[1:, :-1, :-1] + a[:-1, 1:, :-1] + a[1:, 1:, :-1] + a[:-1, :-1, 1:] +
                   a[1:, :-1, 1:] + a[:-1, 1:, 1:] + a[1:, 1:, 1:])


def _faces(P, f):
    """corners p00, p10, p01, p11 of every face of family f: P is (3, n1+1, n2+1, n3+1)
    node coordinates; the face's (a, b) directions are (y, z), (z, x), (x, y)"""
    if f == 0:
        return (P[:, :, :-1, :-1], P[:, :, 1:, :-1], P[:, :, :-1, 1:], P[:, :, 1:, 1:])
    if f == 1:
        return (P[:, :-1, :, :-1], P[:, :-1, :, 1:], P[:, 1:, :, :-1], P[:, 1:, :, 1:])
    return (P[:, :-1, :-1, :], P[:, 1:, :-1, :], P[:, :-1, 1:, :], P[:, 1:, 1:, :])


def face_vectors(X, Y, Z):
    """[(S, xc)] per face family: area vectors S (3, ...) and corner means xc"""
    P = np.stack([X, Y, Z])
    out = []
    for f in range(3):
        p00, p10, p01, p11 = _faces(P, f)
        S = .5 * np.cross(p11 - p00, p01 - p10, axis=0)
        xc = .25 * (p00 + p10 + p01 + p11)
        out.append((S, xc))
    return out


def node_geometry(X, Y, Z):
    """dict: S[f] (3, face shape), area[f], n[f] unit normals, vol (cell shape)"""
    fv = face_vectors(X, Y, Z)
    S = [s for s, _ in fv]
    area = [np.sqrt((s * s).sum(axis=0)) for s in S]
    n = [s / a for s, a in zip(S, area)]
    flux = 0.
    for f, (s, xc) in enumerate(fv):
        sx = (s * xc).sum(axis=0)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[f] = slice(None, -1)
        hi[f] = slice(1, None)
        flux = flux + sx[tuple(hi)] - sx[tuple(lo)]
    vol = flux / 3.
    return dict(S=S, area=area, n=n, vol=vol)


def div_b(grid, st):
    """discrete divergence of the face field (bi, bj, bk = B.n) per cell,
    flux balance over the cell divided by its volume"""
    geo = grid.geometry()
    tot = 0.
    for f, key in enumerate(("bi", "bj", "bk")):
        fl = st[key] * geo["area"][f]
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[f] = slice(None, -1)
        hi[f] = slice(1, None)
        tot = tot + fl[tuple(hi)] - fl[tuple(lo)]
    return tot / geo["vol"]


def face_field_from_potential(grid, Afun):
    """face B (B.n per face family) whose flux is the circulation of the
    vector potential Afun(x, y, z) -> (Ax, Ay, Az) around each face; edge
    integrals by Simpson's rule along the straight edge.  Every edge
    integral is shared by the faces that meet there, so div B is zero up to
    rounding."""
    X, Y, Z = grid.nodes
    P = np.stack([X, Y, Z])

    def edge_int(e):
        sl0 = [slice(None)] * 3
        sl1 = [slice(None)] * 3
        sl0[e] = slice(None, -1)
        sl1[e] = slice(1, None)
        p0 = P[(slice(None),) + tuple(sl0)]
        p1 = P[(slice(None),) + tuple(sl1)]
        pm = .5 * (p0 + p1)
        a0 = np.stack(Afun(*p0))
        a1 = np.stack(Afun(*p1))
        am = np.stack(Afun(*pm))
        return (((a0 + 4. * am + a1) / 6.) * (p1 - p0)).sum(axis=0)

    L = [edge_int(e) for e in range(3)]  # L[e]: edges along axis e, node shape with axis e one shorter
    geo = grid.geometry()
    out = {}
    for f, key in enumerate(("bi", "bj", "bk")):
        a, b = (f + 1) % 3, (f + 2) % 3
        # loop p00 -> (+a) -> (+a+b) -> (+b) -> p00, right-handed about axis f
        def sl(arr, axis, part):
            idx = [slice(None)] * 3
            idx[axis] = slice(None, -1) if part == 0 else slice(1, None)
            return arr[tuple(idx)]

        phi = sl(L[a], b, 0) + sl(L[b], a, 1) - sl(L[a], b, 1) - sl(L[b], a, 0)
        out[key] = np.asfortranarray(phi / geo["area"][f])
    return out


def spherical_shell(nr, nth, nph, ng=4, r0=3., r1=30., theta0=.3):
    """node grid of the shell r0 < r < r1 (log-stretched), theta0 < theta <
    pi - theta0, full periodic phi; ghost nodes continue the same maps"""
    dl = math.log(r1 / r0) / nr
    dth = (math.pi - 2. * theta0) / nth
    dph = 2. * math.pi / nph
    if theta0 - ng * dth <= 0.:
        raise ValueError("theta ghost cells reach the axis: raise theta0 or nth")
    I = np.arange(nr + 2 * ng + 1, dtype=np.float64) - ng
    J = np.arange(nth + 2 * ng + 1, dtype=np.float64) - ng
    K = np.arange(nph + 2 * ng + 1, dtype=np.float64) - ng
    r = r0 * np.exp(I * dl)
    th = theta0 + J * dth
    ph = K * dph
    R, TH, PH = np.meshgrid(r, th, ph, indexing="ij")
    X = R * np.sin(TH) * np.cos(PH)
    Y = R * np.sin(TH) * np.sin(PH)
    Z = R * np.cos(TH)
    g = NodeGrid(nr, nth, nph, ng, X, Y, Z)
    g.shell = dict(r0=r0, r1=r1, theta0=theta0)
    return g


def tensor_nodes(grid):
    """nodes of an ``ics.Grid``'s tensor-product faces (its own xf, yf, zf),
    optionally mapped by its affine A: the same cells given to the general metric"""
    X, Y, Z = np.meshgrid(grid.xf, grid.yf, grid.zf, indexing="ij")
    if getattr(grid, "A", None) is not None:
        A = grid.A
        X, Y, Z = (A[0, 0] * X + A[0, 1] * Y + A[0, 2] * Z, A[1, 0] * X + A[1, 1] * Y + A[1, 2] * Z,
                   A[2, 0] * X + A[2, 1] * Y + A[2, 2] * Z)
    return X, Y, Z


def dipole_inflow(grid, seed, moment=30., rho=1., v=4., p=.1, bz=-.1, noise=.01):
    """config-1 style state on a node grid: dipole of the given moment along
    +z at the origin plus a uniform Bz, uniform flow -v along x, pressure p,
    density rho with seeded uniform noise of relative amplitude `noise`.
    Cell fields are Cartesian; face B from the vector potential
    A = m (-y, x, 0)/r^3 + bz (0, x, 0)."""
    def Afun(x, y, z):
        r3 = (x * x + y * y + z * z) ** 1.5
        return (-moment * y / r3, moment * x / r3 + bz * x, 0. * z)

    cs = (grid.nx, grid.ny, grid.nz)
    n = int(np.prod(cs))
    u = ics.uniform01(seed, n).reshape(cs, order="F")
    st = dict(rho=np.asfortranarray(rho * (1. + noise * (2. * u - 1.))), vx=np.full(cs, -v, order="F"),
              vy=np.zeros(cs, order="F"), vz=np.zeros(cs, order="F"), p=np.full(cs, p, order="F"))
    st.update(face_field_from_potential(grid, Afun))
    return st


def uniform_state(grid, rho, v, p, B):
    """a uniform state on a node grid (face B = B.n of each face)"""
    geo = grid.geometry()
    cs = (grid.nx, grid.ny, grid.nz)
    st = dict(rho=np.full(cs, rho, order="F"), vx=np.full(cs, v[0], order="F"), vy=np.full(cs, v[1], order="F"),
              vz=np.full(cs, v[2], order="F"), p=np.full(cs, p, order="F"))
    for f, key in enumerate(("bi", "bj", "bk")):
        nf = geo["n"][f]
        st[key] = np.asfortranarray(nf[0] * B[0] + nf[1] * B[1] + nf[2] * B[2])
    return st


def node_metric(grid):
    """the per-cell and per-edge factors of the general metric (numpy, same
    definitions as the library's host builder):
      cm (18, cells): cell B = sum over the 6 faces (i lo, i hi, j lo, ...) of
         beta_f * cm[3 f + r]: the area-weighted least-squares fit
         B = N^-1 sum_f |S_f| n_f beta_f, N = sum_f |S_f| n_f n_f^T, exact for
         a uniform field on any grid;
      nc (9, cells): unit vector of S_lo + S_hi per family (cell normals for
         the face-normal velocities);
      ih (3, cells): (|S_lo| + |S_hi|) / (2 V), the inverse spacing of the CFL;
      ef[e] (owned edges): |S1||S2| / V_e of the edge's basis dl (the edge),
         a1, a2 (transverse node differences averaged over the two end
         nodes, a1 along axis e+1, a2 along e+2), S1 = a2 x dl, S2 = dl x a1,
         V_e = a1 . (a2 x dl)."""
    X, Y, Z = grid.nodes
    P = np.stack([X, Y, Z])
    fv = face_vectors(X, Y, Z)
    geo = grid.geometry()
    vol = geo["vol"]
    nc = []
    ih = []
    wn = []    # |S_f| n_f of the 6 faces (i lo, i hi, j lo, ...)
    for f, (S, xf) in enumerate(fv):
        lo = [slice(None)] * 4
        hi = [slice(None)] * 4
        lo[f + 1] = slice(None, -1)
        hi[f + 1] = slice(1, None)
        S0, S1 = S[tuple(lo)], S[tuple(hi)]
        a0 = np.sqrt((S0 * S0).sum(axis=0))
        a1 = np.sqrt((S1 * S1).sum(axis=0))
        wn.extend([S0, S1])
        sn = S0 + S1
        nc.extend(list(sn / np.sqrt((sn * sn).sum(axis=0))))
        ih.append(.5 * (a0 + a1) / vol)
    # N = sum_f |S_f| n_f n_f^T = sum_f S_f S_f^T / |S_f|
    N = sum(np.einsum("r...,q...->rq...", w, w) / np.sqrt((w * w).sum(axis=0)) for w in wn)
    Ni = np.moveaxis(np.linalg.inv(np.moveaxis(N, (0, 1), (-2, -1))), (-2, -1), (0, 1))
    cm = []
    for w in wn:
        cm.extend(list(np.einsum("rq...,q...->r...", Ni, w)))
    ng = grid.ng
    ef = []
    for e in range(3):
        a, b = (e + 1) % 3, (e + 2) % 3
        n0 = [grid.nx0 + (e != 0), grid.ny0 + (e != 1), grid.nz0 + (e != 2)]

        def nodes(shift):
            # node coordinates at owned edge starts + shift (per-axis offsets)
            sl = tuple(slice(ng + shift[q], ng + shift[q] + n0[q]) for q in range(3))
            return P[(slice(None),) + sl]

        def unit(q, s):
            o = [0, 0, 0]
            o[q] = s
            return o

        def add(u, v):
            return [u[q] + v[q] for q in range(3)]

        p0 = nodes([0, 0, 0])
        dl = nodes(unit(e, 1)) - p0
        va = 0.
        vb = 0.
        for s2 in (0, 1):
            base = unit(e, s2)
            va = va + .25 * (nodes(add(base, unit(a, 1))) - nodes(add(base, unit(a, -1))))
            vb = vb + .25 * (nodes(add(base, unit(b, 1))) - nodes(add(base, unit(b, -1))))
        S1 = np.cross(vb, dl, axis=0)
        S2 = np.cross(dl, va, axis=0)
        Ve = (va * S1).sum(axis=0)
        ef.append(np.sqrt((S1 * S1).sum(axis=0)) * np.sqrt((S2 * S2).sum(axis=0)) / Ve)
    return dict(cm=np.stack(cm), nc=np.stack(nc), ih=np.stack(ih), ef=ef)
