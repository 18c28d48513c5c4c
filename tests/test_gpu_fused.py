"""Fused face flux + update of the mixed-precision Adams-Bashforth PDM step
against the unfused kernels (needs a GPU).

The fused kernel forms the same fp32 face products with the same functions
and applies the same binary64 corrector, but its floating-point contraction
may round a face value differently in the last place, so the two paths agree
to fp32 rounding rather than bit for bit.  A last-place difference can flip a
limiter branch or an upwind choice (sign of a near-zero corner velocity) at a
few faces of the noisy blast wave, which moves those faces' values by far more
than a rounding step; so after a 6-step trajectory the fields must agree
  * on average to 1e-6 of each field's scale (the bulk: rounding only), and
  * everywhere to 1e-3 (a face or cell missed or misplaced by the tiling
    would differ at O(1)),
and the fused path must be as close to the binary64 solver as the unfused one
(its mean deviation at most 1.5x the unfused one's, plus 1e-8 of the scale;
a largest deviation would compare single flipped faces).  Cases: the reference's grid and the affine sheared grid, with the
fused CFL partials (periodic) and without (even boundaries), on tile-aligned
and unaligned block shapes.
"""
import numpy as np
import pytest

from mhd3d_mpi_fortran_amd import gpu, ics

pytestmark = pytest.mark.gpu

STATE = ("rho", "vx", "vy", "vz", "p", "bi", "bj", "bk", "bx", "by", "bz")


@pytest.mark.parametrize("shape", [(70, 13, 40), (64, 16, 64)])
@pytest.mark.parametrize("sheared,bc", [(False, "periodic"), (True, "periodic"), (False, "even"), (True, "even")])
def test_fused_equals_unfused(shape, sheared, bc):
    print("fused", shape, sheared, bc, flush=True)
    A = [[1., .5, 0.], [0., 1., .2], [0., 0., 1.]] if sheared else None
    g = ics.Grid(*shape, ng=4, A=A)
    st = ics.initial_state(g, "blast", 21, radius=.2, p_in=10., p_out=.1)
    rng = np.random.default_rng(5)
    for k in ("vx", "vy", "vz"):
        st[k] = np.asfortranarray(st[k] + 1e-2 * rng.standard_normal(st[k].shape))
    out = []
    for precision, fused in (("fp64", False), ("mixed", False), ("mixed", True)):
        s = gpu.Solver(gpu.make_params(*shape, ng=4, reconop=1, stepop=1, bc=bc, precision=precision), g)
        if precision == "mixed":
            s.set_fusion(fused)
        s.upload(st)
        tau = s.initialize(1e-4)
        taus = [s.step(1, tau)[0] for _ in range(6)]
        out.append((s.download(), taus))
        s.close()
    (r, _), (a, ta), (b, tb) = out
    assert np.allclose(ta, tb, rtol=1e-6, atol=0), (ta, tb)
    vmax = max(np.abs(a[k]).max() for k in ("vx", "vy", "vz"))
    bmax = max(np.abs(a[k]).max() for k in ("bi", "bj", "bk"))
    for k in STATE:
        scale = vmax if k[0] == "v" else (bmax if k[0] == "b" else np.abs(a[k]).max())
        assert np.isfinite(b[k]).all(), k
        d = np.abs(a[k] - b[k]) / scale
        e_unf = float(np.abs(a[k] - r[k]).mean()) / scale
        e_fus = float(np.abs(b[k] - r[k]).mean()) / scale
        print("%s: fused-unfused mean %.1e max %.1e; mean vs binary64 unfused %.1e fused %.1e"
              % (k, d.mean(), d.max(), e_unf, e_fus))
        assert d.mean() <= 1e-6 and d.max() <= 1e-3, (k, d.mean(), d.max())
        assert e_fus <= 1.5 * e_unf + 1e-8, (k, e_fus, e_unf)
