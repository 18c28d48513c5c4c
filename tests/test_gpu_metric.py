"""Affine metric (sheared box): the path the reference does not have.

What pins it:
  * the Cartesian limit -- the metric kernels with A = I against the
    reference-parity tensor-product kernels, binary64: the same numbers (the
    metric forms reduce term for term; only signs of exact zeros may differ);
  * free-stream preservation -- a uniform state stays uniform on a sheared
    grid (face areas close the cell exactly, the CT update is a curl);
  * a grid rotated by 90 degrees (A a signed permutation): the whole 3-D MHD
    step -- flux along the normals, face-normal velocities in the EMF, cell B
    from the faces -- must reproduce the tensor-product run with permuted
    vector components (the reconstructions are odd, so they commute with the
    permutation exactly);
  * a planar hydrodynamic Riemann problem (uniform normal B) varying along
    the normal of one face family of a sheared grid is the Cartesian 1-D
    problem with spacing h J/|C e_1|: every Cartesian component is an affine
    function of one profile there, so the reconstructions commute with the
    rotation and the runs agree to rounding.  With a tangential field the
    sheared run mixes two profiles in a component (v.n_2 holds both normal
    and tangential velocity), which limiters do not commute with: that case
    is checked at truncation level only;
  * a stretched grid (A diagonal): every metric factor non-trivial, normals
    the axes -- equal to the tensor-product run on the stretched spacing;
  * div B stays at round-off on a sheared blast wave (CT on the metric).
"""
import math

import numpy as np
import pytest

from mhd3d_mpi_fortran_amd import gpu, ics

pytestmark = pytest.mark.gpu

S = .5
SHEAR = np.array([[1., S, 0.], [0., 1., 0.], [0., 0., 1.]])


def _rel(a, b, scale=None):
    return float(np.abs(a - b).max() / max(np.abs(b).max() if scale is None else scale, 1e-300))


def _scales(st):
    """vector components are compared against the vector's magnitude"""
    v = max(np.abs(st[k]).max() for k in ("vx", "vy", "vz"))
    b = max(np.abs(st[k]).max() for k in ("bi", "bj", "bk", "bx", "by", "bz"))
    return {"vx": v, "vy": v, "vz": v, "bi": b, "bj": b, "bk": b, "bx": b, "by": b, "bz": b}


@pytest.mark.parametrize("stepop", [1, 2])
@pytest.mark.parametrize("reconop", [1, 3])
def test_identity_metric_equals_tensor_path(stepop, reconop):
    ng = 4 if reconop == 1 else 3
    g = ics.Grid(20, 16, 12, ng=ng)
    st = ics.initial_state(g, "orszag_tang", 41)
    out = []
    for met in (False, True):
        s = gpu.Solver(gpu.make_params(20, 16, 12, ng=ng, reconop=reconop, stepop=stepop, bc="periodic"), g)
        if met:
            s.set_metric(np.eye(3))
        s.upload(st)
        tau = s.initialize(1e-3)
        s.step(6, tau)
        out.append(s.download())
        s.close()
    sc = _scales(out[0])
    for k in ("rho", "vx", "vy", "vz", "p", "bi", "bj", "bk", "bx", "by", "bz"):
        assert _rel(out[1][k], out[0][k], sc.get(k)) <= 1e-14, k


def _uniform(g, rho, v, p, B):
    cs = (g.nx, g.ny, g.nz)
    n = g.face_normals()
    bn = n @ np.asarray(B)
    st = dict(rho=np.full(cs, rho, order="F"), vx=np.full(cs, v[0], order="F"), vy=np.full(cs, v[1], order="F"),
              vz=np.full(cs, v[2], order="F"), p=np.full(cs, p, order="F"),
              bi=np.full((g.nx + 1, g.ny, g.nz), bn[0], order="F"),
              bj=np.full((g.nx, g.ny + 1, g.nz), bn[1], order="F"),
              bk=np.full((g.nx, g.ny, g.nz + 1), bn[2], order="F"))
    return st


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("mixed", 2e-6)])
def test_free_stream_on_sheared_grid(precision, tol):
    g = ics.Grid(16, 12, 10, ng=4, A=SHEAR)
    v, B = (.3, -.2, .1), (.4, .2, -.3)
    st = _uniform(g, 1.3, v, .7, B)
    s = gpu.Solver(gpu.make_params(16, 12, 10, reconop=1, stepop=1, bc="periodic", precision=precision), g)
    s.upload(st)
    tau = s.initialize(1e-3)
    s.step(10, tau)
    o = s.download()
    s.close()
    for k, ref in (("rho", 1.3), ("p", .7), ("vx", v[0]), ("vy", v[1]), ("vz", v[2]), ("bx", B[0]), ("by", B[1]),
                   ("bz", B[2])):
        a = g.interior(o[k])
        assert np.abs(a - ref).max() <= tol * max(1., abs(ref)), (k, np.abs(a - ref).max())


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-12)])
def test_scaled_grid_equals_stretched_tensor_path(precision, tol):
    """A = diag(a, b, c): normals are the axes, but every metric factor is
    non-trivial (g = 1/a.., CT 1/(bc).., EMF c.., cell B unchanged) -- the run
    must equal the tensor-product run on the stretched spacing"""
    a = (1.5, .8, 1.2)
    g0 = ics.Grid(16, 14, 12, ng=4, hi=a)
    st = ics.initial_state(g0, "orszag_tang", 43)
    g1 = ics.Grid(16, 14, 12, ng=4, A=np.diag(a))
    out = []
    for grid in (g0, g1):
        s = gpu.Solver(gpu.make_params(16, 14, 12, reconop=1, stepop=2, bc="periodic", precision=precision), grid)
        s.upload(st)
        s.initialize(1e-3)
        s.step(8, 1e-3)
        out.append(s.download())
        s.close()
    # stretched spacing differs from a * dx in the last bit: compare to 1e-12
    # (fp32 stencils: to 1e-5)
    sc = _scales(out[0])
    for k in ("rho", "vx", "vy", "vz", "p", "bi", "bj", "bk", "bx", "by", "bz"):
        e = _rel(out[1][k], out[0][k], sc.get(k))
        assert e <= (tol if precision == "fp64" else 1e-5), (k, e)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("mixed", 1e-5)])
def test_rotated_grid_equals_tensor_path(precision, tol):
    """A = 90-degree rotation about z: x_phys = (-y, x, z)"""
    R = np.array([[0., -1., 0.], [1., 0., 0.], [0., 0., 1.]])
    g = ics.Grid(18, 14, 12, ng=4)
    st = ics.initial_state(g, "orszag_tang", 42)
    rot = dict(st, vx=np.asfortranarray(-st["vy"]), vy=np.asfortranarray(st["vx"]))
    gr = ics.Grid(18, 14, 12, ng=4, A=R)
    out = []
    for grid, state in ((g, st), (gr, rot)):
        s = gpu.Solver(gpu.make_params(18, 14, 12, reconop=1, stepop=1, bc="periodic", precision=precision), grid)
        s.upload(state)
        tau = s.initialize(1e-3)
        s.step(8, tau)
        out.append(s.download())
        s.close()
    a, b = out
    pairs = [("rho", b["rho"], a["rho"]), ("p", b["p"], a["p"]), ("vx", b["vx"], -a["vy"]), ("vy", b["vy"], a["vx"]),
             ("vz", b["vz"], a["vz"]), ("bi", b["bi"], a["bi"]), ("bj", b["bj"], a["bj"]), ("bk", b["bk"], a["bk"]),
             ("bx", b["bx"], -a["by"]), ("by", b["by"], a["bx"]), ("bz", b["bz"], a["bz"])]
    sc = _scales(a)
    for name, x, y in pairs:
        assert _rel(x, y, sc.get(name)) <= tol, (name, _rel(x, y, sc.get(name)))


def _planar(n, s, sheared, mhd):
    """planar Riemann problem along the i-face normal, periodic along i (two
    interfaces); hydro with a uniform normal field, or Brio-Wu-like MHD"""
    m = 4
    lx = 1. if sheared else 1. / math.sqrt(1. + s * s)
    A = np.array([[1., s, 0.], [0., 1., 0.], [0., 0., 1.]]) if sheared else None
    g = ics.Grid(n, m, m, ng=4, hi=(lx, 1., 1.), A=A)
    cs = (g.nx, g.ny, g.nz)
    i = np.arange(g.nx)[:, None, None] - g.ng
    left = (i >= n // 4) & (i < 3 * n // 4)
    rho = np.where(left, 1., .125) * np.ones(cs)
    p = np.where(left, 1., .1) * np.ones(cs)
    un = np.where(left, .05, -.05) * np.ones(cs)       # normal velocity
    ut = (np.where(left, .1, -.2) if mhd else np.zeros(1)) * np.ones(cs)
    bn = .75
    bt = (np.where(left, 1., -1.) if mhd else np.zeros(1)) * np.ones(cs)
    if sheared:
        n1 = np.array([1., -s, 0.]) / math.sqrt(1. + s * s)
        t = np.array([s, 1., 0.]) / math.sqrt(1. + s * s)
        n2 = np.array([0., 1., 0.])
        vel = [un * n1[q] + ut * t[q] for q in range(3)]
        bj = bn * (n1 @ n2) + bt * (t @ n2)
        st = dict(rho=rho, p=p, vx=vel[0], vy=vel[1], vz=vel[2],
                  bi=np.full((g.nx + 1, g.ny, g.nz), bn), bj=np.concatenate([bj, bj[:, :1]], axis=1),
                  bk=np.zeros((g.nx, g.ny, g.nz + 1)))
    else:
        st = dict(rho=rho, p=p, vx=un, vy=ut, vz=np.zeros(cs), bi=np.full((g.nx + 1, g.ny, g.nz), bn),
                  bj=np.concatenate([bt, bt[:, :1]], axis=1), bk=np.zeros((g.nx, g.ny, g.nz + 1)))
    st = {k: np.asfortranarray(v, dtype=np.float64) for k, v in st.items()}
    return g, st


def _planar_pair(precision, mhd, n=64, s=S, steps=30):
    runs = {}
    for sheared in (False, True):
        g, st = _planar(n, s, sheared, mhd)
        sv = gpu.Solver(gpu.make_params(n, 4, 4, reconop=1, stepop=2, bc="periodic", precision=precision), g)
        sv.upload(st)
        sv.initialize(2e-3)
        sv.step(steps, 2e-3)
        runs[sheared] = (g, sv.download())
        sv.close()
    gc, c = runs[False]
    gs, h = runs[True]
    n1 = np.array([1., -s, 0.]) / math.sqrt(1. + s * s)
    t = np.array([s, 1., 0.]) / math.sqrt(1. + s * s)
    I = lambda g_, a: g_.interior(a)[:, 1, 1]
    pairs = [("rho", I(gs, h["rho"]), I(gc, c["rho"])), ("p", I(gs, h["p"]), I(gc, c["p"])),
             ("un", n1[0] * I(gs, h["vx"]) + n1[1] * I(gs, h["vy"]), I(gc, c["vx"])),
             ("ut", t[0] * I(gs, h["vx"]) + t[1] * I(gs, h["vy"]), I(gc, c["vy"])),
             ("bt", t[0] * I(gs, h["bx"]) + t[1] * I(gs, h["by"]), I(gc, c["by"])),
             ("bn", n1[0] * I(gs, h["bx"]) + n1[1] * I(gs, h["by"]), I(gc, c["bx"]))]
    moved = np.abs(I(gc, c["rho"]) - np.where((np.arange(n) >= n // 4) & (np.arange(n) < 3 * n // 4), 1., .125))
    assert moved.max() > 1e-2   # the waves did move
    return gs, h, pairs


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("mixed", 2e-5)])
def test_planar_hydro_is_rotated_cartesian(precision, tol):
    gs, h, pairs = _planar_pair(precision, mhd=False)
    for name, a, b in pairs:
        assert np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1.), (name, np.abs(a - b).max())
    for k in ("rho", "p", "vx"):
        a = gs.interior(h[k])
        assert np.abs(a - a[:, :1, :1]).max() <= tol * max(np.abs(a).max(), 1.), k


def test_planar_mhd_close_to_rotated_cartesian():
    gs, h, pairs = _planar_pair("fp64", mhd=True)
    for name, a, b in pairs:
        l1 = np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300)
        print("planar MHD, sheared vs rotated Cartesian: %s L1 %.3e" % (name, l1))
        assert l1 <= .15, (name, l1)
    for k in ("rho", "p", "vx", "by"):
        a = gs.interior(h[k])
        assert np.abs(a - a[:, :1, :1]).max() <= 1e-12 * max(np.abs(a).max(), 1.), k


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sheared_blast_keeps_div_b(precision):
    n = 32
    g = ics.Grid(n, n, n, ng=4, A=SHEAR)
    st = ics.initial_state(g, "blast", 4, radius=.15)
    s = gpu.Solver(gpu.make_params(n, n, n, reconop=1, stepop=1, bc="periodic", precision=precision), g)
    s.upload(st)
    tau = s.initialize(1e-4)
    s.step(20, tau)
    o = s.download()
    assert s.check_state() == 0
    s.close()
    nrm = g.face_normals()
    A = SHEAR
    C = np.array([np.cross(A[:, (d + 1) % 3], A[:, (d + 2) % 3]) for d in range(3)])
    gd = np.linalg.norm(C, axis=1) / np.linalg.det(A)
    bi, bj, bk = g.interior(o["bi"], 1), g.interior(o["bj"], 2), g.interior(o["bk"], 3)
    h = 1. / n
    div = gd[0] * (bi[1:] - bi[:-1]) / h + gd[1] * (bj[:, 1:] - bj[:, :-1]) / h + gd[2] * (bk[:, :, 1:] - bk[:, :, :-1]) / h
    bmax = max(np.abs(bi).max(), np.abs(bj).max(), np.abs(bk).max())
    assert np.abs(div).max() <= 1e-12 * bmax / h
    # the blast did expand
    assert g.interior(o["rho"]).max() > 1.5
    del nrm
