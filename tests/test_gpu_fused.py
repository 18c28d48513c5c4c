"""Fused face flux + update of the mixed-precision Adams-Bashforth PDM step
against the unfused kernels (needs a GPU).

The fused kernel forms the same fp32 face products with the same functions
and applies the same binary64 corrector, but its floating-point contraction
may round a face value differently in the last place, so the two paths agree
to fp32 rounding rather than bit for bit: a 6-step trajectory to 5e-5 of each
field's scale (the blast wave's steep fronts amplify a last-place difference
a little), on the reference's grid and the affine sheared grid, with the
fused CFL partials (periodic) and without (even boundaries), on tile-aligned
and unaligned block shapes (a face or cell missed by the tiling would differ
at O(1)).
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
    for fused in (False, True):
        s = gpu.Solver(gpu.make_params(*shape, ng=4, reconop=1, stepop=1, bc=bc, precision="mixed"), g)
        s.set_fusion(fused)
        s.upload(st)
        tau = s.initialize(1e-4)
        taus = [s.step(1, tau)[0] for _ in range(6)]
        out.append((s.download(), taus))
        s.close()
    (a, ta), (b, tb) = out
    assert np.allclose(ta, tb, rtol=1e-6, atol=0), (ta, tb)
    vmax = max(np.abs(a[k]).max() for k in ("vx", "vy", "vz"))
    bmax = max(np.abs(a[k]).max() for k in ("bi", "bj", "bk"))
    for k in STATE:
        scale = vmax if k[0] == "v" else (bmax if k[0] == "b" else np.abs(a[k]).max())
        err = float(np.abs(a[k] - b[k]).max()) / scale
        assert np.isfinite(b[k]).all() and err <= 5e-5, (k, err)
