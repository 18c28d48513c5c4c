"""The CPU oracle against the reference's own outputs (tests/golden/*.npz).

Each fixture was produced by the reference Fortran (mhd3d3.f90 built with the
ifort -r8 semantics, MPI on 2x2x2 ranks) on the same initial state; the oracle
runs the whole domain as one block.  The bar is bit-exact equality: same
operations in the same order on the same binary64 inputs.
"""
import numpy as np
import pytest

import pyoracle as po
from goldens import STATE, Case, cases

ALL = cases()


def _oracle(c):
    m = c.meta
    p = po.make_params(*m["grid"], ng=m["ng"], reconop=m["reconop"], stepop=m["stepop"], bc=m["bc"],
                       gam=m["gam"], cfl=m["cfl"], pdmb=m["pdmb"], eps=m["eps"], alpha=m["alpha"],
                       beta=m["beta"])
    o = po.Oracle(p, c.grid)
    o.set_state(c.full_state("ic"))
    tau = o.initialize(m["tau"])
    return o, tau


def test_fixture_inventory():
    names = set(ALL)
    # every reconstruction, every stepper, both boundary kinds are pinned
    assert {"ot_pdm_rk2_per", "ot_tvd_rk2_per", "ot_weno_rk2_per", "ot_mpweno_rk2_per"} <= names
    assert {"ot_pdm_ab2_per", "ot_pdm_rk3_per", "st_pdm_rk2_even", "st_mpweno_rk3_even"} <= names
    assert len(names) >= 16


@pytest.mark.parametrize("name", ALL)
def test_initial_ops_bit_exact(name):
    """get_dt, get_eijk_LFM and get_flux_stress on the initial state."""
    c = Case(name)
    o, _ = _oracle(c)
    ref = c.ops()
    assert o.get_dt() == ref["tcfl"]
    ops = o.ops()
    for k in ("ei", "ej", "ek"):
        assert np.array_equal(ops[k], ref[k]), k
    for q in range(24):
        assert np.array_equal(ops["flux"][q], ref["flux"][q]), "flux %d" % q


@pytest.mark.parametrize("name", ALL)
def test_initial_yee_bit_exact(name):
    """get_eijk_Yee (mhd3d3.f90:1654-1752), the reference's alternative EMF
    whose own calls are commented out, on the initial state: the fixtures
    hold it from the reference build (its harness calls it once there), all
    four reconstructions (each Yee edge value takes eight reconstruct_3d
    calls in directions 4-6 on top of the Gaussian split of get_v_center)."""
    c = Case(name)
    o, _ = _oracle(c)
    ref = c.yee()
    ops = o.ops(yee=True)
    for k in ("ei", "ej", "ek"):
        assert np.array_equal(ops[k], ref[k]), k


def test_yee_differs_from_lfm():
    """the two EMFs are different routines: on a magnetised flow they must
    not coincide (a Yee entry that fell back to get_eijk_LFM would pass the
    comparison above only where both vanish)"""
    c = Case("ot_pdm_rk2_per")
    lfm, yee = c.ops(), c.yee()
    for k in ("ei", "ej", "ek"):
        d = np.abs(lfm[k] - yee[k]).max()
        assert d > 1e-6 * np.abs(lfm[k]).max(), k


@pytest.mark.parametrize("name", ALL)
def test_trajectory_bit_exact(name):
    """the owned state after each stored frame (nts=1: one step per frame)."""
    c = Case(name)
    m = c.meta
    o, tau = _oracle(c)
    done = 0
    for f in m["frames"]:
        tau, _ = o.step(f - done, m["tau"] if m["stepop"] != 1 else tau)
        done = f
        st = o.get_state()
        assert tau == float(c.z["f%d_tau" % f])
        for v in STATE:
            assert np.array_equal(c.owned(st, v), c.frame(f, v)), (f, v)


def test_planted_mistake_is_caught():
    """a one-ulp change to the oracle's input shows up in the comparison."""
    c = Case("ot_pdm_rk2_per")
    m = c.meta
    st = c.full_state("ic")
    g = c.grid
    g.interior(st["p"])[3, 4, 2] = np.nextafter(g.interior(st["p"])[3, 4, 2], 1.)
    p = po.make_params(*m["grid"], ng=m["ng"], reconop=m["reconop"], stepop=m["stepop"], bc=m["bc"])
    o = po.Oracle(p, g)
    o.set_state(st)
    o.initialize(m["tau"])
    o.step(1, m["tau"])
    assert not np.array_equal(c.owned(o.get_state(), "p"), c.frame(1, "p"))
