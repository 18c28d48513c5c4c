"""General curvilinear metric (mhd_set_metric_nodes): grids given by node
coordinates -- the spherical shell of the dipole-plus-inflow set-up -- which
the reference's tensor-product grid cannot represent.

What pins it:
  * the Cartesian limit -- a stretched tensor-product grid given by its
    nodes runs through the general kernels (per-face areas and normals,
    per-cell volumes and cell-B coefficients, per-edge EMF factors) and must
    reproduce the reference-parity tensor-product kernels on the same
    spacing to rounding (the formulas differ in rounding only);
  * the affine limit -- a sheared grid given by its nodes equals the affine
    metric path (uniform factors there, per-element factors here), which the
    affine tests pin against rotations and the Cartesian limit in turn;
  * free-stream preservation on the shell -- the six area vectors of every
    cell close exactly, so a uniform unmagnetised flow stays uniform to
    rounding on a curved, log-stretched grid;
  * div B stays at rounding on the shell under CT, from a dipole field whose
    face fluxes are edge circulations of a vector potential;
  * the mixed-precision run of the shell set-up stays close to binary64.
"""
import numpy as np
import pytest

from mhd3d_mpi_fortran_amd import gpu, ics, shell

pytestmark = pytest.mark.gpu

STATE = ("rho", "vx", "vy", "vz", "p", "bi", "bj", "bk", "bx", "by", "bz")


def _scales(st):
    v = max(np.abs(st[k]).max() for k in ("vx", "vy", "vz"))
    b = max(np.abs(st[k]).max() for k in ("bi", "bj", "bk", "bx", "by", "bz"))
    return {"vx": v, "vy": v, "vz": v, "bi": b, "bj": b, "bk": b, "bx": b, "by": b, "bz": b}


def _rel(a, b, scale):
    return float(np.abs(a - b).max() / max(scale, 1e-300))


def _stretched(nx0, ny0, nz0, ng, A=None):
    """tensor grid with smoothly non-uniform spacing on every axis"""
    g = ics.Grid(nx0, ny0, nz0, ng=ng, A=A)
    for ax, amp in (("x", .08), ("y", -.06), ("z", .05)):
        f = getattr(g, ax + "f")
        f = f + amp * np.sin(2. * np.pi * f) / (2. * np.pi)
        setattr(g, ax + "f", f)
        setattr(g, "d" + ax, ics.metric(f))
    return g


def _blast(g, seed, kind="blast"):
    """initial state with seeded 1e-3 velocity noise: the exact symmetry of
    the plain blast leaves velocities at exact zeros and exact ties in the
    EMF's sign-selected velocity columns, where rounding-level differences
    between two correct metric forms pick different columns"""
    st = ics.initial_state(g, kind, seed, radius=.2) if kind == "blast" else ics.initial_state(g, kind, seed)
    rng = np.random.default_rng(seed)
    for k in ("vx", "vy", "vz"):
        st[k] = np.asfortranarray(st[k] + 1e-3 * rng.standard_normal(st[k].shape))
    return st


def _run(grid, params, st, nsteps, tau0=1e-4, nodes=None, A=None):
    s = gpu.Solver(params, grid)
    try:
        if nodes is not None:
            s.set_metric_nodes(*nodes)
        elif A is not None:
            s.set_metric(A)
        s.upload(st)
        tau = s.initialize(tau0)
        s.step(nsteps, tau)
        return s.download()
    finally:
        s.close()


# WENO5 is not positivity-preserving: the 1000:1 pressure jump of the blast
# wave breaks it on either path, so the WENO cases run the Orszag-Tang vortex
@pytest.mark.parametrize("stepop,reconop,kind", [(2, 1, "blast"), (1, 1, "blast"), (2, 2, "blast"),
                                                 (2, 3, "orszag_tang"), (1, 4, "orszag_tang")])
def test_nodes_tensor_grid_equals_tensor_path(stepop, reconop, kind):
    ng = 4 if reconop == 1 else 3
    g = _stretched(20, 16, 12, ng)
    st = _blast(g, 7, kind)
    p = gpu.make_params(20, 16, 12, ng=ng, reconop=reconop, stepop=stepop, bc="even")
    ref = _run(g, p, st, 6)
    got = _run(g, p, st, 6, nodes=shell.tensor_nodes(g))
    sc = _scales(ref)
    for k in STATE:
        assert np.isfinite(ref[k]).all() and np.isfinite(got[k]).all(), k
        err = _rel(got[k], ref[k], sc.get(k, np.abs(ref[k]).max()))
        assert err <= 1e-11, (k, err)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("mixed", 1e-5)])
def test_nodes_affine_grid_equals_affine_path(precision, tol):
    A = np.array([[1., .5, 0.], [0., 1., .2], [0., 0., 1.]])
    g = ics.Grid(20, 16, 12, ng=4, A=A)
    st = _blast(g, 11)
    p = gpu.make_params(20, 16, 12, ng=4, reconop=1, stepop=1, bc="periodic", precision=precision)
    ref = _run(g, p, st, 6)    # Solver applies grid.A (affine path)
    g2 = ics.Grid(20, 16, 12, ng=4)
    got = _run(g2, p, st, 6, nodes=shell.tensor_nodes(g))
    sc = _scales(ref)
    for k in STATE:
        err = _rel(got[k], ref[k], sc.get(k, np.abs(ref[k]).max()))
        assert err <= tol, (k, err)


@pytest.mark.parametrize("stepop,precision,tol", [(2, "fp64", 1e-12), (1, "fp64", 1e-12), (2, "mixed", 2e-6)])
def test_shell_free_stream(stepop, precision, tol):
    g = shell.spherical_shell(16, 16, 24, ng=3, theta0=.6)
    v = (-4., 1., .5)
    st = shell.uniform_state(g, 1., v, .1, (0., 0., 0.))
    p = gpu.make_params(16, 16, 24, ng=3, reconop=3, stepop=stepop, bc=("even", "even", "periodic"),
                        precision=precision)
    out = _run(g, p, st, 10, tau0=2e-3)
    for k, ref in (("rho", 1.), ("vx", v[0]), ("vy", v[1]), ("vz", v[2]), ("p", .1)):
        err = float(np.abs(out[k] - ref).max()) / max(abs(ref), 4.)
        assert err <= tol, (k, err)


@pytest.mark.parametrize("stepop", [2, 1])
def test_shell_dipole_keeps_div_b(stepop):
    g = shell.spherical_shell(24, 16, 32, ng=3, theta0=.6)
    st = shell.dipole_inflow(g, 21)
    d0 = shell.div_b(g, st)
    bmax = max(np.abs(st[k]).max() for k in ("bi", "bj", "bk"))
    hmin = float(np.cbrt(g.geometry()["vol"].min()))
    assert np.abs(d0).max() * hmin / bmax <= 1e-13
    p = gpu.make_params(24, 16, 32, ng=3, reconop=3, stepop=stepop, bc=("even", "even", "periodic"))
    out = _run(g, p, st, 20, tau0=1e-3)
    for k in STATE:
        assert np.isfinite(out[k]).all(), k
    ng = g.ng
    inner = (slice(ng, -ng),) * 3
    assert out["rho"][inner].min() > 0. and out["p"][inner].min() > 0.
    d = shell.div_b(g, out)[inner]
    assert np.abs(d).max() * hmin / bmax <= 1e-12, np.abs(d).max() * hmin / bmax
    # the flow has moved: the state is not the initial one
    assert np.abs(out["bi"] - st["bi"]).max() > 1e-6 * bmax


def test_shell_mixed_close_to_fp64():
    g = shell.spherical_shell(24, 16, 32, ng=3, theta0=.6)
    st = shell.dipole_inflow(g, 22)
    res = {}
    for prec in ("fp64", "mixed"):
        p = gpu.make_params(24, 16, 32, ng=3, reconop=3, stepop=2, bc=("even", "even", "periodic"),
                            precision=prec)
        res[prec] = _run(g, p, st, 20, tau0=1e-3)
    ng = g.ng
    inner = (slice(ng, -ng),) * 3
    a, b = res["mixed"]["rho"][inner], res["fp64"]["rho"][inner]
    l1 = float(np.abs(a - b).mean() / np.abs(b).mean())
    assert l1 <= 1e-5, l1
    d = shell.div_b(g, res["mixed"])[inner]
    bmax = max(np.abs(st[k]).max() for k in ("bi", "bj", "bk"))
    hmin = float(np.cbrt(g.geometry()["vol"].min()))
    assert np.abs(d).max() * hmin / bmax <= 1e-12
