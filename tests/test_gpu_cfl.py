"""CFL step of the mixed-precision Adams-Bashforth path (get_dt,
mhd3d3.f90:1608-1639: the minimum over the ghost-inclusive block).

Where every ghost cell repeats an owned cell (periodic axes whose ghost
spacings repeat the owned ones) the update leaves the CFL partials of the new
state itself, over the owned cells; elsewhere a get_bxyz pass over the whole
block leaves them.  Either way the next get_dt must equal, bit for bit, get_dt
of the same state computed afresh over the ghost-inclusive block (after a
download/upload round trip, which drops the partials)."""
import numpy as np
import pytest

from mhd3d_mpi_fortran_amd import gpu, ics

pytestmark = pytest.mark.gpu

SHEAR = [[1., .5, 0.], [0., 1., 0.], [0., 0., 1.]]


@pytest.mark.parametrize("bc,sheared,stretch", [("periodic", False, False), ("periodic", True, False),
                                                ("even", False, False), ("periodic", False, True)])
def test_step_cfl_partials_equal_fresh_get_dt(bc, sheared, stretch):
    n = 24
    g = ics.Grid(n, n, n, ng=4, A=SHEAR if sheared else None)
    if stretch:  # wider low ghost cells along x: no longer twins of owned cells
        g.dx = g.dx.copy()
        g.dx[:g.ng] *= 1.5
    st = ics.initial_state(g, "blast", seed=4, radius=.15, p_in=20., p_out=.1)
    with gpu.Solver(gpu.make_params(n, n, n, ng=4, reconop=1, stepop=1, bc=bc, precision="mixed"), g) as s:
        s.upload(st)
        tau = s.initialize(1e-4)
        for k in (1, 4):
            tau, _ = s.step(k, tau)
            d_step = s.get_dt()
            s.upload(s.download())
            d_fresh = s.get_dt()
            assert d_step == d_fresh, (k, d_step, d_fresh)
            assert np.isfinite(d_step) and d_step > 0.
