"""Host-side set-up: tensor-product geometry and initial states.

Geometry follows ``init_geom`` (mhd3d3.f90:3379-3515): face coordinates
``x0(i)=(i-1)*dx0+xmin``, ghost faces mirrored about the domain ends, metric
``dxblock(i)=x(i+1)-x(i)`` per index (so not exactly ``dx0``), cell and face
centres as the reference's 8- and 4-corner means in the reference's summation
order.  Arrays are returned in the reference layout (Fortran order, x fastest),
ghost-inclusive, with ghosts zero: the ghost fill runs on the device.

Initial states (all seeded with SplitMix64 so that every consumer -- tests,
bench, the Fortran driver's IC file -- builds the same bits):
  * ``shocktube``  the reference's own init_ic (mhd3d3.f90:3517-3569);
  * ``orszag_tang`` the 3-D Orszag-Tang vortex of mhd3d2.f90:3420-3438 with
    ``vz=0.01 sin 2 pi z`` and 1e-3 uniform density noise;
  * ``blast``      seeded blast wave, 4 centres, p=100 inside r<radius;
  * ``blobs``      rho = 1 + three seeded Gaussian blobs, p=1, uniform |B|=0.5;
  * ``rough``      uniform random fields (exercises every limiter branch).

A grid may carry an affine metric ``A`` (physical x = A x_comp, e.g. a
sheared box): positions are then physical, vectors are Cartesian, and the
face arrays hold the normal components B.n of each face family.
"""
import math

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n, offset=0):
    """n consecutive SplitMix64 outputs of stream `seed` starting at `offset`."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, n, offset=0):
    """doubles in [0,1) from the top 53 bits."""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def axis_faces(xmin, xmax, n0, ng):
    """ghost-inclusive face coordinates xx0(1..n0+2ng+1) (mhd3d3.f90:3391-3424)."""
    dx0 = (xmax - xmin) / n0
    x0 = np.array([float(i) * dx0 + xmin for i in range(n0 + 1)])
    xx = np.zeros(n0 + 2 * ng + 1)
    xx[ng:ng + n0 + 1] = x0
    for i in range(1, ng + 1):
        xx[i - 1] = 2. * x0[0] - x0[ng + 1 - i]
        xx[n0 + 2 * ng + 1 - i] = 2. * x0[n0] - x0[n0 - ng + i - 1]
    return xx


def metric(xx):
    """dxblock along one axis: x(i+1)-x(i), i=1..n (mhd3d3.f90:3483)."""
    return xx[1:] - xx[:-1]


def _c8(a, b, pattern):
    # 0.125 * 8-corner sum in the reference's order; pattern picks a/b
    s = None
    for c in pattern:
        t = a if c == 0 else b
        s = t if s is None else s + t
    return .125 * s


def _c4(a, b, pattern):
    s = None
    for c in pattern:
        t = a if c == 0 else b
        s = t if s is None else s + t
    return .25 * s


class Grid:
    """Block geometry of one domain (whole domain on one device)."""

    def __init__(self, nx0, ny0, nz0, ng=4, lo=(0., 0., 0.), hi=(1., 1., 1.), A=None):
        self.A = None if A is None else np.asarray(A, dtype=np.float64).reshape(3, 3)
        self.nx0, self.ny0, self.nz0, self.ng = nx0, ny0, nz0, ng
        self.nx, self.ny, self.nz = nx0 + 2 * ng, ny0 + 2 * ng, nz0 + 2 * ng
        self.lo, self.hi = tuple(lo), tuple(hi)
        self.xf = axis_faces(lo[0], hi[0], nx0, ng)
        self.yf = axis_faces(lo[1], hi[1], ny0, ng)
        self.zf = axis_faces(lo[2], hi[2], nz0, ng)
        self.dx, self.dy, self.dz = metric(self.xf), metric(self.yf), metric(self.zf)

    def face_normals(self):
        """unit normals n[d] of the faces normal to computational axis d"""
        A = np.eye(3) if self.A is None else self.A
        a = [A[:, d] for d in range(3)]
        n = np.array([np.cross(a[(d + 1) % 3], a[(d + 2) % 3]) for d in range(3)])
        return n / np.linalg.norm(n, axis=1)[:, None]

    def physical(self, x, y, z):
        """map computational coordinates (broadcastable arrays) to physical"""
        if self.A is None:
            return x, y, z
        A = self.A
        return (A[0, 0] * x + A[0, 1] * y + A[0, 2] * z, A[1, 0] * x + A[1, 1] * y + A[1, 2] * z,
                A[2, 0] * x + A[2, 1] * y + A[2, 2] * z)

    # broadcastable 3-D coordinate arrays (Fortran order is applied at the end)
    def cell_centres(self):
        xa, xb = self.xf[:-1, None, None], self.xf[1:, None, None]
        ya, yb = self.yf[None, :-1, None], self.yf[None, 1:, None]
        za, zb = self.zf[None, None, :-1], self.zf[None, None, 1:]
        # xblock(i,j,k)+xblock(i,j,k+1)+xblock(i+1,j,k+1)+xblock(i+1,j,k)+
        # xblock(i,j+1,k)+xblock(i,j+1,k+1)+xblock(i+1,j+1,k+1)+xblock(i+1,j+1,k)
        xc = _c8(xa, xb, (0, 0, 1, 1, 0, 0, 1, 1))
        yc = _c8(ya, yb, (0, 0, 0, 0, 1, 1, 1, 1))
        zc = _c8(za, zb, (0, 1, 1, 0, 0, 1, 1, 0))
        return xc, yc, zc

    def i_face_centres(self):
        x = self.xf[:, None, None]
        yi = _c4(self.yf[None, :-1, None], self.yf[None, 1:, None], (0, 1, 0, 1))
        zi = _c4(self.zf[None, None, :-1], self.zf[None, None, 1:], (0, 0, 1, 1))
        return x, yi, zi

    def j_face_centres(self):
        xj = _c4(self.xf[:-1, None, None], self.xf[1:, None, None], (0, 1, 0, 1))
        zj = _c4(self.zf[None, None, :-1], self.zf[None, None, 1:], (0, 0, 1, 1))
        return xj, self.yf[None, :, None], zj

    def k_face_centres(self):
        xk = _c4(self.xf[:-1, None, None], self.xf[1:, None, None], (0, 1, 0, 1))
        yk = _c4(self.yf[None, :-1, None], self.yf[None, 1:, None], (0, 0, 1, 1))
        return xk, yk, self.zf[None, None, :]

    def zeros_cell(self):
        return np.zeros((self.nx, self.ny, self.nz), order="F")

    def interior(self, a, face=0):
        g = self.ng
        sl = [slice(g, g + self.nx0), slice(g, g + self.ny0), slice(g, g + self.nz0)]
        if face:
            sl[face - 1] = slice(g, g + (self.nx0, self.ny0, self.nz0)[face - 1] + 1)
        return a[tuple(sl)]


def _full(shape, val=0.):
    return np.full(shape, val, order="F")


def _bcast(a, shape):
    return np.asfortranarray(np.broadcast_to(a, shape).copy())


def _interior_mask(g, shape):
    m = np.zeros(shape, dtype=bool)
    m[g.ng:g.ng + g.nx0, g.ng:g.ng + g.ny0, g.ng:g.ng + g.nz0] = True
    return m


def initial_state(g, kind, seed=0, **kw):
    """dict rho,vx,vy,vz,p,bi,bj,bk of full ghost-inclusive arrays (ghosts 0)."""
    cs = (g.nx, g.ny, g.nz)
    fi, fj, fk = (g.nx + 1, g.ny, g.nz), (g.nx, g.ny + 1, g.nz), (g.nx, g.ny, g.nz + 1)
    pi = 4. * math.atan(1.)
    xc, yc, zc = g.cell_centres()
    st = {}
    if g.A is not None and kind != "blast":
        raise ValueError("initial state %r is defined on the tensor-product grid only" % kind)
    if kind == "shocktube":
        rho = _full(cs, 1.)
        p = _full(cs, 1.)
        zcb = _bcast(zc, cs)
        rho[zcb >= 0.] = .125
        p[zcb >= 0.] = .1
        st = dict(rho=rho, vx=_full(cs), vy=_full(cs), vz=_full(cs), p=p)
        bi = _full(fi, -1.)
        _, _, zi = g.i_face_centres()
        bi[_bcast(zi, fi) < 0.] = 1.
        st.update(bi=bi, bj=_full(fj, 0.), bk=_full(fk, .75))
    elif kind == "orszag_tang":
        amp = kw.get("noise", 1e-3)
        rho = _full(cs, 25. / (36. * pi))
        n = g.nx0 * g.ny0 * g.nz0
        u = uniform01(seed, n).reshape((g.nx0, g.ny0, g.nz0), order="F")
        g.interior(rho)[...] = (25. / (36. * pi)) * (1. + amp * (2. * u - 1.))
        st = dict(rho=rho, p=_full(cs, 5. / (12. * pi)),
                  vx=_bcast(-np.sin(2. * pi * yc), cs), vy=_bcast(np.sin(2. * pi * xc), cs),
                  vz=_bcast(.01 * np.sin(2. * pi * zc), cs))
        b0 = 1. / math.sqrt(4. * pi)
        _, yi, _ = g.i_face_centres()
        xj, _, _ = g.j_face_centres()
        st.update(bi=_bcast(-b0 * np.sin(2. * pi * yi), fi), bj=_bcast(b0 * np.sin(4. * pi * xj), fj),
                  bk=_full(fk, 0.))
    elif kind == "blast":
        ncent = kw.get("ncentres", 4)
        r0 = kw.get("radius", .1)
        pin, pout = kw.get("p_in", 100.), kw.get("p_out", .1)
        u = uniform01(seed, 3 * ncent)
        lo, hi = np.array(g.lo), np.array(g.hi)
        L = hi - lo
        cent = lo + u.reshape(ncent, 3) * L
        xb, yb, zb = (_bcast(c, cs) for c in (xc, yc, zc))
        inside = np.zeros(cs, dtype=bool)
        for c in cent:
            # nearest periodic image in computational coordinates, distance in
            # physical ones (the same thing without a metric)
            if g.A is None:
                d = [np.abs(xb - c[0]), np.abs(yb - c[1]), np.abs(zb - c[2])]
                d = [np.minimum(dd, Lq - dd) for dd, Lq in zip(d, L)]
            else:
                d = [xb - c[0], yb - c[1], zb - c[2]]
                d = g.physical(*[dd - Lq * np.round(dd / Lq) for dd, Lq in zip(d, L)])
            inside |= (d[0] ** 2 + d[1] ** 2 + d[2] ** 2) < r0 * r0
        p = _full(cs, pout)
        p[inside] = pin
        b = 1. / math.sqrt(2.)
        bn = g.face_normals() @ np.array([b, b, 0.])   # normal components of the uniform B
        st = dict(rho=_full(cs, 1.), vx=_full(cs), vy=_full(cs), vz=_full(cs), p=p,
                  bi=_full(fi, bn[0]), bj=_full(fj, bn[1]), bk=_full(fk, bn[2]))
    elif kind == "blobs":
        u = uniform01(seed, 3 * 5 + 2)
        lo, hi = np.array(g.lo), np.array(g.hi)
        L = hi - lo
        xb, yb, zb = (_bcast(c, cs) for c in (xc, yc, zc))
        rho = _full(cs, 1.)
        for m in range(3):
            c = lo + u[5 * m:5 * m + 3] * L
            amp = .5 + 1.5 * u[5 * m + 3]
            w = .05 + .1 * u[5 * m + 4]
            d = [np.abs(xb - c[0]), np.abs(yb - c[1]), np.abs(zb - c[2])]
            d = [np.minimum(dd, Lq - dd) for dd, Lq in zip(d, L)]
            rho = rho + amp * np.exp(-(d[0] ** 2 + d[1] ** 2 + d[2] ** 2) / (w * w))
        th = math.acos(2. * u[15] - 1.)
        ph = 2. * pi * u[16]
        bvec = .5 * np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
        st = dict(rho=np.asfortranarray(rho), vx=_full(cs), vy=_full(cs), vz=_full(cs), p=_full(cs, 1.),
                  bi=_full(fi, bvec[0]), bj=_full(fj, bvec[1]), bk=_full(fk, bvec[2]))
    elif kind == "rough":
        def rnd(shape, idx, lo_, hi_):
            n = shape[0] * shape[1] * shape[2]
            u = uniform01(seed * 1000003 + idx, n)
            return np.asfortranarray((lo_ + (hi_ - lo_) * u).reshape(shape, order="F"))
        st = dict(rho=rnd(cs, 0, .5, 1.5), vx=rnd(cs, 1, -.5, .5), vy=rnd(cs, 2, -.5, .5),
                  vz=rnd(cs, 3, -.5, .5), p=rnd(cs, 4, .5, 1.5),
                  bi=rnd(fi, 5, -.5, .5), bj=rnd(fj, 6, -.5, .5), bk=rnd(fk, 7, -.5, .5))
    else:
        raise ValueError("unknown initial state %r" % kind)
    # ghosts zero: the fill owns them
    out = {}
    for key, a in st.items():
        a = np.asfortranarray(a, dtype=np.float64)
        face = {"bi": 1, "bj": 2, "bk": 3}.get(key, 0)
        z = np.zeros(a.shape, order="F")
        g.interior(z, face)[...] = g.interior(a, face)
        out[key] = z
    return out


def cfl_dt_estimate(g, st, gam=5. / 3., cfl=.3):
    """rough CFL dt of an interior state (host planning aid only)."""
    rho, p = g.interior(st["rho"]), g.interior(st["p"])
    v = np.sqrt(g.interior(st["vx"]) ** 2 + g.interior(st["vy"]) ** 2 + g.interior(st["vz"]) ** 2)
    bi, bj, bk = g.interior(st["bi"], 1), g.interior(st["bj"], 2), g.interior(st["bk"], 3)
    bx = .5 * (bi[1:] + bi[:-1])
    by = .5 * (bj[:, 1:] + bj[:, :-1])
    bz = .5 * (bk[:, :, 1:] + bk[:, :, :-1])
    vc = v + np.sqrt((bx ** 2 + by ** 2 + bz ** 2 + gam * p) / rho)
    dx = g.dx[g.ng:g.ng + g.nx0].min()
    dy = g.dy[g.ng:g.ng + g.ny0].min()
    dz = g.dz[g.ng:g.ng + g.nz0].min()
    return float(cfl / (vc.max() * (1. / dx + 1. / dy + 1. / dz)))
