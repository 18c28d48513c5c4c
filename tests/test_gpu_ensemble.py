"""Batched ensemble (mhd_ensemble_*): independent blocks of uneven sizes and
mixed precisions stepped together by one HIP graph launch per step.

The graph replays the members' own kernel sequences, so each member's state
after k ensemble steps must equal, bit for bit, the same member stepped k
times on its own (mhd_step), which the parity tests pin against the oracle.
The graph is replayed several times with eager work and synchronisation in
between (downloads, a member's own mhd_step), and every replay is checked.
"""
import numpy as np
import pytest

from mhd3d_mpi_fortran_amd import gpu, ics

pytestmark = pytest.mark.gpu

STATE = ("rho", "vx", "vy", "vz", "p", "bi", "bj", "bk", "bx", "by", "bz")

# (n, reconop, stepop, precision): uneven cubes, TVD RK2 in both precisions
# (the ensemble configuration), plus PDM RK3 and WENO5 members
MEMBERS = [(16, 2, 2, "fp64"), (24, 2, 2, "mixed"), (12, 2, 2, "mixed"), (20, 1, 3, "fp64"), (16, 3, 2, "mixed")]


def _member(n, reconop, stepop, precision, seed):
    ng = 4 if reconop == 1 else 3
    g = ics.Grid(n, n, n, ng=ng)
    st = ics.initial_state(g, "blobs", seed)
    s = gpu.Solver(gpu.make_params(n, n, n, ng=ng, reconop=reconop, stepop=stepop, bc="periodic",
                                   precision=precision), g)
    s.upload(st)
    s.initialize(0.)
    tau = .5 * s.get_dt()
    return s, tau


def _equal(a, b):
    for k in STATE:
        assert np.array_equal(a[k], b[k]), k


def test_ensemble_matches_members_stepped_alone():
    ens_members, ref_members, taus = [], [], []
    for q, (n, r, stp, prec) in enumerate(MEMBERS):
        s, tau = _member(n, r, stp, prec, 40 + q)
        o, tau_o = _member(n, r, stp, prec, 40 + q)
        assert tau == tau_o
        ens_members.append(s)
        ref_members.append(o)
        taus.append(tau)
    with gpu.Ensemble(ens_members, taus) as ens:
        for rep, k in enumerate((3, 1, 2)):
            ens.step(k)
            for s, o, tau in zip(ens_members, ref_members, taus):
                o.step(k, tau)
                _equal(s.download(), o.download())
            if rep == 1:
                # a member stepped on its own between replays
                ens_members[1].step(1, taus[1])
                ref_members[1].step(1, taus[1])
                _equal(ens_members[1].download(), ref_members[1].download())
    for s in ens_members + ref_members:
        s.close()


def test_ensemble_rejects_unsupported_members():
    g = ics.Grid(12, 12, 12, ng=4)
    st = ics.initial_state(g, "blobs", 3)
    ab = gpu.Solver(gpu.make_params(12, 12, 12, ng=4, reconop=1, stepop=1, bc="periodic"), g)
    ab.upload(st)
    ab.initialize(1e-4)
    with pytest.raises(gpu.MhdError):
        gpu.Ensemble([ab], [1e-4])
    rk, tau = _member(12, 2, 2, "fp64", 5)
    with pytest.raises(gpu.MhdError):
        gpu.Ensemble([rk, rk], [tau, tau])   # a member listed twice
    ab.close()
    rk.close()


@pytest.mark.parametrize("stepop,precision,reconop", [(2, "fp64", 1), (1, "fp64", 1), (1, "mixed", 1), (3, "mixed", 3)])
def test_stage_overlap_is_bit_identical(stepop, precision, reconop):
    # EMF and CT on the second stream beside the flux and update, against
    # every kernel on one stream in the reference's order
    n = 24
    ng = 4 if reconop == 1 else 3
    g = ics.Grid(n, n, n, ng=ng)
    st = ics.initial_state(g, "orszag_tang", 61)
    out = []
    for overlap in (True, False):
        s = gpu.Solver(gpu.make_params(n, n, n, ng=ng, reconop=reconop, stepop=stepop, bc="periodic",
                                       precision=precision), g)
        s.set_overlap(overlap)
        s.upload(st)
        tau = s.initialize(2e-3)
        s.step(8, tau)
        out.append(s.download())
        s.close()
    _equal(out[0], out[1])
