"""Per-routine drop-ins of the time stepper against the oracle (needs a GPU).

mhd_advance = get_conserve + advance_corrector1/2 + stage combine +
advance_bijk (mhd3d3.f90:1641-1650, 3842-3933, the stage bodies of
step_rk2/step_rk3 3665-3707, 3733-3785 and step_ab 3622-3627),
mhd_predict_ab = advance_predictor_ab with its advance_previous and tau0=tau
(3616, 3821-3840), mhd_set_tau/mhd_get_tau = the module variables tau, tau0,
and the which_state argument of get_flux_stress / get_eijk_LFM (the state
the reference passes them: current in RK, extrapolated in AB).

Bars:
  * the update routines see no transcendental when both sides start from the
    same state and the same face fluxes and edge EMFs (the oracle's, written
    into the library with mhd_store): bit-exact, every stage of RK2, RK3 and
    AB2, even and periodic boundaries, through several steps;
  * the EMF of the extrapolated state is bit-exact, its flux within 1e-13
    (erfc/exp: ROCm ocml vs glibc);
  * the routine-by-routine chain reproduces mhd_step bit for bit, binary64
    and mixed precision;
  * the CFL step of the next Adams-Bashforth step does not depend on a div B
    monitor call in between, and get_dt right after an upload forms cell B
    from the uploaded faces.
"""
import numpy as np
import pytest

import pyoracle as po
from mhd3d_mpi_fortran_amd import gpu, ics

pytestmark = pytest.mark.gpu

FLUX_TOL = 1e-13
PRIMARY = ("rho", "vx", "vy", "vz", "p", "bi", "bj", "bk")


def rel(a, b):
    d = float(np.max(np.abs(a - b)))
    s = float(np.max(np.abs(b)))
    return d / s if s > 0 else d


def pair(reconop, stepop, bc, kind="rough", seed=5, n=(12, 10, 8), ng=4, precision="fp64"):
    g = ics.Grid(*n, ng=ng)
    st = ics.initial_state(g, kind, seed)
    s = gpu.Solver(gpu.make_params(*n, ng=ng, reconop=reconop, stepop=stepop, bc=bc, precision=precision), g)
    o = po.Oracle(po.make_params(*n, ng=ng, reconop=reconop, stepop=stepop, bc=bc), g)
    s.upload(st)
    o.set_state(st)
    return g, st, s, o


def same(a, b, keys):
    for k in keys:
        assert np.array_equal(a[k], b[k]), k


def same_owned(g, a, b, keys):
    """interior cells and owned faces: between an update and the ghost fill
    the ghost layers hold different leftovers (the reference's stage combine
    averages whole arrays, ghosts included; the library combines the owned
    values), which the fill then overwrites on both sides"""
    for k in keys:
        f = {"bi": 1, "bj": 2, "bk": 3}.get(k, 0)
        assert np.array_equal(g.interior(a[k], f), g.interior(b[k], f)), k


CASES = [(1, 2, "even"), (1, 2, "periodic"), (2, 3, "periodic"), (4, 3, "even"), (1, 1, "periodic"),
         (3, 1, "even")]


@pytest.mark.parametrize("reconop,stepop,bc", CASES)
def test_advance_stages_bit_exact(reconop, stepop, bc):
    print("advance", reconop, stepop, bc, flush=True)
    g, st, s, o = pair(reconop, stepop, bc)
    tau = s.initialize(2e-3)
    assert tau == o.initialize(2e-3)
    for step in range(3):
        tcfl = s.get_dt()
        assert tcfl == o.get_dt()
        if stepop == 1:
            # get_dt sets the module's tau; the predictor then extrapolates
            assert s.get_tau()[0] == tcfl == o.get_tau()[0]
            s.predict_ab()
            o.predict_ab()
            assert s.get_tau() == o.get_tau()
            stages, which = [0], gpu.STATE_PREDICTED
        else:
            s.set_tau(tau, tau)
            o.set_tau(tau, tau)
            stages, which = list(range(1, stepop + 1)), gpu.STATE_CURRENT
        for stage in stages:
            ref = o.ops(which)
            e = s.eijk(which)
            for k in ("ei", "ej", "ek"):
                assert np.array_equal(e[k], ref[k]), (step, stage, k)
            f = s.flux_stress(which)
            for q in range(24):
                assert rel(f[q], ref["flux"][q]) <= FLUX_TOL, (step, stage, q)
            # the update itself from identical products: bit for bit
            s.store_ops(ref)
            assert np.array_equal(s.flux_stress(compute=False)[5], ref["flux"][5])
            s.advance(stage)
            o.advance(stage)
            same_owned(g, s.download(), o.get_state(), PRIMARY)
            s.fill()
            o.fill()
            s.bxyz()
            o.bxyz()
            same(s.download(), o.get_state(), PRIMARY + ("bx", "by", "bz"))


def test_predicted_state_needs_predictor():
    g, st, s, o = pair(1, 1, "periodic")
    s.initialize(1e-3)
    with pytest.raises(gpu.MhdError):
        s.flux_stress(gpu.STATE_PREDICTED)   # no mhd_predict_ab yet
    s.predict_ab()
    s.eijk(gpu.STATE_PREDICTED)
    s.advance(0)
    with pytest.raises(gpu.MhdError):
        s.eijk(gpu.STATE_PREDICTED)           # the state has moved on
    with pytest.raises(gpu.MhdError):
        s.advance(1)                          # not a stage of AB2
    g2, st2, r, o2 = pair(1, 2, "periodic")
    r.initialize(1e-3)
    with pytest.raises(gpu.MhdError):
        r.predict_ab()                        # Runge-Kutta context
    with pytest.raises(gpu.MhdError):
        r.advance(3)
    with pytest.raises(gpu.MhdError):
        r.set_tau(-1., 1.)


def chain(s, stepop, tau):
    """one step routine by routine, as step_rk2/rk3/ab call them"""
    s.get_dt()
    if stepop == 1:
        s.predict_ab()
        stages, which = [0], gpu.STATE_PREDICTED
    else:
        s.set_tau(tau, tau)
        stages, which = list(range(1, stepop + 1)), gpu.STATE_CURRENT
    for stage in stages:
        s.eijk(which)
        s.flux_stress(which)
        s.advance(stage)
        s.fill()
        s.bxyz()


@pytest.mark.parametrize("precision,reconop,stepop,bc,sheared", [
    ("fp64", 1, 2, "periodic", False), ("fp64", 3, 1, "even", False), ("fp64", 2, 3, "even", False),
    ("mixed", 1, 1, "periodic", False), ("mixed", 1, 1, "periodic", True), ("mixed", 2, 3, "even", False),
    ("mixed", 1, 2, "periodic", True), ("fp64", 1, 1, "periodic", True), ("mixed", 1, 1, "even", False),
    ("mixed", 1, 1, "even", True)])
def test_chain_equals_step(precision, reconop, stepop, bc, sheared):
    print("chain", precision, reconop, stepop, bc, sheared, flush=True)
    n = (16, 12, 10)
    A = [[1., .5, 0.], [0., 1., 0.], [0., 0., 1.]] if sheared else None
    g = ics.Grid(*n, ng=4, A=A)
    # (WENO5 is not positivity-preserving on the blast's pressure jump; the
    # Orszag-Tang vortex is set up on the reference's grid only)
    st = (ics.initial_state(g, "blast", 3, radius=.2, p_in=10., p_out=.1) if sheared else
          ics.initial_state(g, "orszag_tang", 3))
    out = []
    for routine in (False, True):
        s = gpu.Solver(gpu.make_params(*n, ng=4, reconop=reconop, stepop=stepop, bc=bc, precision=precision), g)
        # the per-routine entry points run the unfused flux and update
        # kernels (mhd_step's optional fused flux + update agrees with them
        # to fp32 rounding only, tests/test_gpu_fused.py)
        s.set_fusion(False)
        s.upload(st)
        tau = s.initialize(1e-3)
        for _ in range(3):
            if routine:
                chain(s, stepop, tau)
            else:
                s.step(1, tau)
        out.append((s.download(), s.get_tau()))
        s.close()
    (a, ta), (b, tb) = out
    assert ta == tb
    assert np.isfinite(a["rho"]).all()
    same(a, b, PRIMARY + ("bx", "by", "bz"))


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("bc", ["periodic", "even"])
def test_divb_monitor_leaves_cfl_alone(precision, bc):
    """a div B monitor between Adams-Bashforth steps changes nothing"""
    n = (16, 12, 10)
    g = ics.Grid(*n, ng=4)
    st = ics.initial_state(g, "blast", 2, radius=.2, p_in=10., p_out=.1)
    res = []
    for monitor in (False, True):
        s = gpu.Solver(gpu.make_params(*n, ng=4, reconop=1, stepop=1, bc=bc, precision=precision), g)
        s.upload(st)
        s.initialize(1e-3)
        taus = []
        for _ in range(3):
            t, tc = s.step(1)
            taus.append((t, tc))
            if monitor:
                assert s.divb_max() < 1e-10
        res.append((taus, s.download()))
        s.close()
    assert res[0][0] == res[1][0]
    same(res[0][1], res[1][1], PRIMARY)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_get_dt_after_upload(precision):
    """get_dt right after an upload uses cell B of the uploaded faces"""
    g, st, s, o = pair(1, 2, "periodic", kind="orszag_tang", precision=precision)
    s.initialize(1e-3)
    full = s.download()             # ghost-filled state
    s.step(2, 1e-3)                 # cell B of another state in the arrays
    s.upload(full)
    o.set_state(full)
    o.bxyz()
    assert s.get_dt() == o.get_dt()


@pytest.mark.parametrize("bc,sheared,n", [("periodic", True, (16, 12, 10)), ("periodic", False, (13, 9, 7)),
                                          ("even", False, (16, 12, 10)), ("even", True, (9, 11, 8))])
def test_mixed_ab_step_leaves_ghosts_filled(bc, sheared, n):
    """after mixed Adams-Bashforth steps the ghost layers are filled (a
    fill changes nothing), and at a periodic x boundary the faces' fluxes,
    whose stencils reach into the ghosts of the fp32 stencil copies, equal
    their twins' on the opposite boundary bit for bit, as do the EMFs at
    periodic boundaries across their march axes.  (Along a march axis --
    z for ei and ej, y for ek -- the first edge of a line starts the march
    afresh while its twin ends a run, and fp32 contraction may round the two
    differently in the last place; the y and z boundary faces that CT
    updates from those edges then differ in the last place too, and so do
    the y and z face fluxes next to them.  The x boundary faces are updated
    from ej and ek at edges across their march axes.)"""
    print("ghosts after steps", bc, sheared, n, flush=True)
    A = [[1., .5, 0.], [0., 1., 0.], [0., 0., 1.]] if sheared else None
    g = ics.Grid(*n, ng=4, A=A)
    st = ics.initial_state(g, "blast", 7, radius=.2, p_in=10., p_out=.1)
    s = gpu.Solver(gpu.make_params(*n, ng=4, reconop=1, stepop=1, bc=bc, precision="mixed"), g)
    try:
        s.upload(st)
        s.initialize(1e-3)
        s.step(4)
        a = s.download()
        s.fill()
        b = s.download()
        same(a, b, PRIMARY)
        if bc == "periodic":
            f = s.flux_stress(compute=False)
            e = s.eijk(gpu.STATE_PREDICTED, compute=False)
            for q in range(8):
                x = f[q]
                assert np.array_equal(x[0], x[-1]), q
            for key, axes in (("ei", (1,)), ("ej", (0,)), ("ek", (0,))):
                for d in axes:
                    lo = [slice(None)] * 3
                    hi = [slice(None)] * 3
                    lo[d], hi[d] = 0, -1
                    assert np.array_equal(e[key][tuple(lo)], e[key][tuple(hi)]), (key, d)
    finally:
        s.close()
