"""The reference's dead alternatives on the GPU (needs a GPU).

  * get_eijk_Yee (mhd3d3.f90:1654-1752; its own calls at 3354 and 3621 are
    commented out): mhd_eijk_yee against the reference's output on the
    initial state of every fixture (tests/golden: the harness of the
    reference build calls it once there) and against the oracle on the
    Adams-Bashforth extrapolation.  The split field carries get_v_center's
    erfc and exp, which the GPU takes from ROCm's device library and the
    reference from glibc, so the bar is not bit equality; and the split
    terms B_b v_b0 v_c1 - B_c v_b1 v_c0 cancel to an E far smaller than
    either term, which turns a last-place difference of a moment into a
    few hundred ulp of E (measured: up to 3.2e-13 of each array's scale),
    so the bar is 1e-12 of the scale where the flux's is 1e-13.  The eight
    reconstructions per edge value are k_recon, pinned bit for bit.
  * get_flux_stress_old (2131-2383): get_flux_stress's operations on the
    extrapolated arrays rhoh, vxh, ... -- mhd_flux_stress_old against the
    oracle's get_flux_stress of the extrapolated state.
  * binary64 on the tensor-product grid only: mixed precision and metrics
    are refused.
"""
import numpy as np
import pytest

import pyoracle as po
from goldens import Case, cases
from mhd3d_mpi_fortran_amd import gpu, ics

pytestmark = pytest.mark.gpu

FLUX_TOL = 1e-13
YEE_TOL = 1e-12


def rel(a, b):
    d = float(np.max(np.abs(a - b)))
    s = float(np.max(np.abs(b)))
    return d / s if s > 0 else d


@pytest.mark.parametrize("name", cases())
def test_yee_matches_reference(name):
    c = Case(name)
    m = c.meta
    s = gpu.Solver(gpu.make_params(*m["grid"], ng=m["ng"], reconop=m["reconop"], stepop=m["stepop"],
                                   bc=m["bc"]), c.grid)
    try:
        s.upload(c.full_state("ic"))
        s.initialize(m["tau"])
        e = s.eijk(yee=True)
        lfm = s.eijk()
    finally:
        s.close()
    ref = c.yee()
    for k in ("ei", "ej", "ek"):
        print(name, k, "%.2e" % rel(e[k], ref[k]))
        assert rel(e[k], ref[k]) <= YEE_TOL, (k, rel(e[k], ref[k]))
        # the LFM EMF stays bit-exact after a Yee call on the same context
        assert np.array_equal(lfm[k], c.ops()[k]), k


@pytest.mark.parametrize("reconop", [1, 2, 3, 4])
def test_yee_and_flux_old_on_extrapolated_state(reconop):
    n, ng = (12, 10, 8), 4
    g = ics.Grid(*n, ng=ng)
    # every component O(1) (the Orszag-Tang vortex's B_z and its stress
    # would be rounding noise after two steps)
    st = ics.initial_state(g, "rough", 40 + reconop)
    s = gpu.Solver(gpu.make_params(*n, ng=ng, reconop=reconop, stepop=1, bc="periodic"), g)
    o = po.Oracle(po.make_params(*n, ng=ng, reconop=reconop, stepop=1, bc="periodic"), g)
    try:
        s.upload(st)
        o.set_state(st)
        tau = s.initialize(1e-3)
        o.initialize(1e-3)
        s.step(2, tau)
        o.step(2, tau)
        s.get_dt()
        o.get_dt()
        s.predict_ab()
        o.predict_ab()
        e = s.eijk(gpu.STATE_PREDICTED, yee=True)
        gpu._ck(gpu.lib().mhd_flux_stress_old(s.h), "mhd_flux_stress_old")
        f = s.flux_stress(compute=False)
    finally:
        s.close()
    ref = o.ops(which=1, yee=True)
    for k in ("ei", "ej", "ek"):
        assert rel(e[k], ref[k]) <= YEE_TOL, (k, rel(e[k], ref[k]))
    for q in range(24):
        assert rel(f[q], ref["flux"][q]) <= FLUX_TOL, (q, rel(f[q], ref["flux"][q]))


def test_yee_refused_where_not_defined():
    n = (8, 8, 8)
    g = ics.Grid(*n, ng=4)
    st = ics.initial_state(g, "orszag_tang", 3)
    s = gpu.Solver(gpu.make_params(*n, ng=4, precision="mixed"), g)
    try:
        s.upload(st)
        s.initialize(1e-3)
        with pytest.raises(gpu.MhdError):
            s.eijk(yee=True)
    finally:
        s.close()
    A = [[1., .5, 0.], [0., 1., 0.], [0., 0., 1.]]
    g2 = ics.Grid(*n, ng=4, A=A)
    s = gpu.Solver(gpu.make_params(*n, ng=4), g2)
    try:
        s.upload(ics.initial_state(g2, "blast", 3))
        s.initialize(1e-3)
        with pytest.raises(gpu.MhdError):
            s.eijk(yee=True)
    finally:
        s.close()
