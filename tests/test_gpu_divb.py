"""Divergence monitor (mhd_divb_max) against the same divergence formed in
numpy: on face fields with a non-zero divergence (random faces), so that the
monitor's value is a real number to compare, on the reference's tensor grid,
an affine (sheared) grid and the spherical shell's node grid.  Then after CT
steps of a divergence-free field it reads round-off."""
import numpy as np
import pytest

from mhd3d_mpi_fortran_amd import gpu, ics, shell

pytestmark = pytest.mark.gpu


def _random_faces(g, seed):
    rng = np.random.default_rng(seed)
    return {k: np.asfortranarray(rng.standard_normal(shp)) for k, shp in
            (("bi", (g.nx + 1, g.ny, g.nz)), ("bj", (g.nx, g.ny + 1, g.nz)), ("bk", (g.nx, g.ny, g.nz + 1)))}


def _numpy_divb_tensor(g, st, A=None):
    ng = g.ng
    bi, bj, bk = g.interior(st["bi"], 1), g.interior(st["bj"], 2), g.interior(st["bk"], 3)
    dx = g.dx[ng:ng + g.nx0][:, None, None]
    dy = g.dy[ng:ng + g.ny0][None, :, None]
    dz = g.dz[ng:ng + g.nz0][None, None, :]
    fac = [1., 1., 1.]
    if A is not None:
        A = np.asarray(A, dtype=np.float64)
        C = np.array([np.cross(A[:, (d + 1) % 3], A[:, (d + 2) % 3]) for d in range(3)])
        fac = list(np.linalg.norm(C, axis=1) / np.linalg.det(A))
    return (fac[0] * (bi[1:] - bi[:-1]) / dx + fac[1] * (bj[:, 1:] - bj[:, :-1]) / dy +
            fac[2] * (bk[:, :, 1:] - bk[:, :, :-1]) / dz)


@pytest.mark.parametrize("A", [None, [[1., .5, 0.], [0., 1., .2], [0., 0., 1.]]])
def test_divb_monitor_tensor_and_affine(A):
    g = ics.Grid(14, 12, 10, ng=4, A=A)
    st = ics.initial_state(g, "blast", 3)
    st.update(_random_faces(g, 4))
    with gpu.Solver(gpu.make_params(14, 12, 10, ng=4, bc="periodic"), g) as s:
        s.upload(st)
        got = s.divb_max()
    ref = float(np.abs(_numpy_divb_tensor(g, st, A)).max())
    assert abs(got - ref) <= 1e-13 * ref, (got, ref)


def test_divb_monitor_shell():
    g = shell.spherical_shell(12, 16, 24, ng=3, theta0=.6)
    st = shell.dipole_inflow(g, 5)
    st.update(_random_faces(g, 6))
    with gpu.Solver(gpu.make_params(12, 16, 24, ng=3, reconop=3, stepop=2, bc=("even", "even", "periodic")),
                    g) as s:
        s.upload(st)
        got = s.divb_max()
    ng = g.ng
    ref = float(np.abs(shell.div_b(g, st)[ng:-ng, ng:-ng, ng:-ng]).max())
    assert abs(got - ref) <= 1e-12 * ref, (got, ref)


def test_divb_monitor_reads_round_off_after_ct():
    n = 24
    g = ics.Grid(n, n, n, ng=4)
    st = ics.initial_state(g, "orszag_tang", 8)
    with gpu.Solver(gpu.make_params(n, n, n, ng=4, reconop=1, stepop=2, bc="periodic", precision="mixed"), g) as s:
        s.upload(st)
        tau = s.initialize(2e-3)
        d0 = s.divb_max()
        s.step(10, tau)
        d1 = s.divb_max()
        out = s.download()
    bmax = max(float(np.abs(out[k]).max()) for k in ("bi", "bj", "bk"))
    h = 1. / n
    assert d0 * h / bmax <= 1e-13 and d1 * h / bmax <= 1e-12, (d0, d1)
    assert abs(d1 - float(np.abs(_numpy_divb_tensor(g, out)).max())) <= 1e-12 * bmax / h
