"""Learned face reconstruction: gradients, MFMA inference, training, solver use.

Checks, and what each pins:
  * the oracle's net forward against a numpy restatement, and the closed
    form on linear data: every gate is zero, so the face states are the
    exact midpoint values, which is also what PDM gives there;
  * the oracle's learned flux: with theta < 0 it is plain PDM bit for bit,
    and on a state linear along each axis it equals the PDM flux on every
    face whose windows stay inside the linear region;
  * CPU gradient (oracle, double) against central finite differences;
  * GPU training gradient against the CPU gradient (fp32 vs double);
  * GPU bf16-MFMA forward against the CPU forward with the same bf16
    rounding of weights, inputs and hidden activations (pins the MFMA lane
    layouts and the epilogues);
  * collected samples: normalised inputs in [-1,1], roughness <= theta, PDM
    targets inside [min, max] of the two cells next to the face (a property
    of PDM2's clamp and limiter, mhd3d3.f90:1142-1166);
  * a trained net beats both the untrained net and the 4th-order linear
    interpolation on held-out faces (what the network gains);
  * the GPU's fused learned flux against the oracle's learned flux with the
    same bf16 rounding;
  * the solver with the net stays within the density L1 bound of the
    binary64 solver, next to the same solver with the linear part only.
"""
import numpy as np
import pytest

import pyoracle as po
from mhd3d_mpi_fortran_amd import gpu, ics
from mhd3d_mpi_fortran_amd import surrogate as sg


def _bf16(a):
    a = np.asarray(a, dtype=np.float32).copy()
    u = a.view(np.uint32)
    u += (0x7fff + ((u >> 16) & 1)).astype(np.uint32)
    u &= np.uint32(0xffff0000)
    return u.view(np.float32)


def _samples(rng, n):
    """random samples: inputs and targets in [-1, 1]"""
    return rng.uniform(-1, 1, (n, sg.NSAMPLE)).astype(np.float32)


def _weights(seed, scale=1.):
    rng = np.random.default_rng(seed)
    w = np.zeros(sg.NPARAM, dtype=np.float32)
    w[:1280] = rng.uniform(-1, 1, 1280) / np.sqrt(40) * scale
    w[1280:1312] = rng.uniform(-.1, .1, 32)
    w[1312:2336] = rng.uniform(-1, 1, 1024) / np.sqrt(32) * scale
    w[2336:2368] = rng.uniform(-.1, .1, 32)
    w[2368:2880] = rng.uniform(-1, 1, 512) / np.sqrt(32) * scale
    w[2880:] = rng.uniform(-.1, .1, 16)
    return w


def test_net_forward_restated():
    rng = np.random.default_rng(11)
    w = _weights(5, 2.)
    x = rng.uniform(-1, 1, (300, 40)).astype(np.float32)
    y = po.mlp_forward(w, x)
    ref = sg.net_forward(w, x)
    assert np.abs(y - ref).max() <= 1e-5 * np.abs(ref).max()


def test_net_exact_on_linear_windows():
    """normalised windows of linear data have zero gates: the output is the
    fixed interpolation, which is the face midpoint -- and PDM gives the
    same there (no clamp or limiter acts on monotone linear data)"""
    rng = np.random.default_rng(12)
    w = _weights(6, 3.)
    up = np.array([-.6, -.2, .2, .6, 1.], dtype=np.float32)
    x = np.stack([np.concatenate([up if rng.random() < .5 else -up for _ in range(8)]) for _ in range(50)])
    y = po.mlp_forward(w, x)
    assert np.abs(y).max() <= 1e-6
    f = np.arange(16, dtype=np.float64) * .37 + 2.
    p = po.make_params(8, 1, 1, ng=4, reconop=1)
    fl, fr = po.reconstruct(p, f.reshape(16, 1, 1, order="F"), (9, 1, 1), 4)
    mid = .5 * (f[3:12] + f[4:13])
    assert np.allclose(fl.ravel(), mid, rtol=0, atol=1e-13) and np.allclose(fr.ravel(), mid, rtol=0, atol=1e-13)


def _oracle_pair(g, st, w, theta, bf16_round=False):
    o = po.Oracle(po.make_params(g.nx0, g.ny0, g.nz0, ng=4, reconop=1, stepop=1, bc="even"), g)
    o.set_state(st)
    o.fill()
    o.bxyz()
    plain = o.ops()["flux"]
    o.set_surrogate(w, theta, bf16_round)
    return plain, o.ops()["flux"]


def test_oracle_learned_flux_fallback_is_pdm():
    g = ics.Grid(10, 9, 8, ng=4)
    st = ics.initial_state(g, "orszag_tang", 31)
    plain, learned = _oracle_pair(g, st, _weights(7, 2.), theta=-1.)
    for q in range(24):
        assert np.array_equal(plain[q], learned[q]), q


def test_oracle_learned_flux_on_linear_state():
    g = ics.Grid(14, 13, 12, ng=4)
    st = ics.initial_state(g, "orszag_tang", 32)
    X, Y, Z = np.meshgrid(np.arange(g.nx), np.arange(g.ny), np.arange(g.nz), indexing="ij")
    lin = lambda a, b, c, d: a + b * X + c * Y + d * Z
    st = dict(st)
    st.update(rho=np.asfortranarray(lin(1., .01, .02, -.015)), p=np.asfortranarray(lin(.5, -.004, .003, .002)),
              vx=np.asfortranarray(lin(.1, .01, -.02, .005)), vy=np.asfortranarray(lin(-.2, .003, .01, -.01)),
              vz=np.asfortranarray(lin(.05, -.02, .004, .01)))
    for k, shp in (("bi", (g.nx + 1, g.ny, g.nz)), ("bj", (g.nx, g.ny + 1, g.nz)), ("bk", (g.nx, g.ny, g.nz + 1))):
        a, b, c = np.meshgrid(np.arange(shp[0]), np.arange(shp[1]), np.arange(shp[2]), indexing="ij")
        st[k] = np.asfortranarray({"bi": .3, "bj": -.2, "bk": .25}[k] + .01 * a - .02 * b + .015 * c)
    o = po.Oracle(po.make_params(g.nx0, g.ny0, g.nz0, ng=4, reconop=1, stepop=1, bc="periodic"), g)
    o.set_state(st)
    o.bxyz()          # no ghost fill: the linear data runs through the ghost cells
    plain = o.ops()["flux"]
    o.set_surrogate(_weights(8, 3.), 1.0)
    learned = o.ops()["flux"]
    for d in range(3):
        for q in range(8):
            a, b = plain[8 * d + q], learned[8 * d + q]
            assert np.abs(a - b).max() <= 1e-6 * np.abs(a).max(), (d, q)


def test_cpu_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    w = _weights(1, 2.)
    s = _samples(rng, 64)
    loss, g = po.mlp_loss_grad(w, s)
    for i in (0, 77, 1279, 1290, 1400, 2350, 2400, 2879, 2890):
        h = 2e-3
        wp, wm = w.copy(), w.copy()
        wp[i] += np.float32(h)
        wm[i] -= np.float32(h)
        fd = (po.mlp_loss_grad(wp, s)[0] - po.mlp_loss_grad(wm, s)[0]) / (float(wp[i]) - float(wm[i]))
        assert abs(fd - g[i]) <= 1e-4 * max(abs(g[i]), 1e-3), (i, fd, g[i])


def test_planted_gradient_error_is_caught():
    """a wrong activation derivative in a CPU copy of the backward pass would
    move the gradient well past the comparison tolerance used on the GPU"""
    rng = np.random.default_rng(3)
    w = _weights(2)
    s = _samples(rng, 32)
    _, g = po.mlp_loss_grad(w, s)
    g_bad = g.copy()
    g_bad[:1312] *= 1.01
    assert np.abs(g_bad - g).max() / np.abs(g).max() > 1e-4


@pytest.fixture(scope="module")
def ctx():
    g = ics.Grid(8, 8, 8, ng=4)
    s = gpu.Solver(gpu.make_params(8, 8, 8, ng=4, reconop=1, stepop=2, bc="periodic", precision="mixed",
                                   surrogate=1), g)
    yield s
    s.close()


@pytest.mark.gpu
def test_gpu_gradient_matches_oracle(ctx):
    rng = np.random.default_rng(5)
    w = _weights(3, 2.)
    s = _samples(rng, 777)
    lg, gg = sg.grad(ctx, w, s)
    lc, gc = po.mlp_loss_grad(w, s)
    assert abs(lg - lc) <= 1e-5 * lc
    for a, b in ((0, 1280), (1280, 1312), (1312, 2336), (2336, 2368), (2368, 2880), (2880, 2896)):
        err = np.abs(gg[a:b] - gc[a:b]).max() / np.abs(gc[a:b]).max()
        assert err <= 1e-4, (a, b, err)


@pytest.mark.gpu
def test_mfma_inference_matches_cpu(ctx):
    rng = np.random.default_rng(7)
    w = _weights(4, 2.)
    sg.set_weights(ctx, w)
    x = rng.uniform(-1, 1, (1000, 40)).astype(np.float32)   # not a multiple of 64: exercises the tail wave
    y = sg.infer(ctx, x)
    ref = po.mlp_forward(_bf16(w), _bf16(x), h_round=True)
    exact = po.mlp_forward(w, x)
    scale = np.abs(exact).max()
    assert np.abs(y - ref).max() / scale <= 5e-3
    # and the bf16 path is what it claims: close to fp32 but not identical
    assert np.abs(y - exact).max() / scale <= 3e-2


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["orszag_tang", "blobs"])
def test_gpu_learned_flux_matches_oracle(kind):
    """the configuration-3 set-ups (the Orszag-Tang vortex of configuration 0
    and the Gaussian blobs of configuration 2) at a size the oracle finishes
    in seconds"""
    g = ics.Grid(20, 18, 16, ng=4)
    st = ics.initial_state(g, kind, 33)
    w = _weights(9, 1.)
    s = gpu.Solver(gpu.make_params(20, 18, 16, ng=4, reconop=1, stepop=1, bc="even", precision="mixed",
                                   surrogate=1), g)
    sg.set_weights(s, w)
    sg.configure(s, .5, count=True)
    s.upload(st)
    s.fill()
    f = s.flux_stress()
    faces, fb = sg.stats(s)
    s.close()
    plain, ref = _oracle_pair(g, st, w, .5, bf16_round=True)
    moved = 0
    for q in range(24):
        scale = np.abs(ref[q]).max() + 1e-300
        err = np.abs(f[q] - ref[q]) / scale
        # fp32 stencils against binary64; a face whose roughness sits at theta
        # may take the other branch, so a few faces may differ by more
        assert np.quantile(err, .999) <= 2e-3, (q, np.quantile(err, .999))
        assert err.max() <= 5e-2, (q, err.max())
        moved += int((np.abs(ref[q] - plain[q]) > 1e-3 * scale).sum())
    # and the net did take part: its fluxes differ from PDM's on some faces
    # (the resting blobs have no rough face, so there nothing falls back)
    print("%s: %d of %d faces on the PDM fallback, %d flux values moved by the net" % (kind, fb, faces, moved))
    assert moved > 0 and fb < faces
    if kind == "orszag_tang":
        assert fb > 0


@pytest.mark.gpu
def test_gpu_linear_part_matches_oracle_across_ring_runs():
    """zero weights and no threshold: every face takes the clamped linear part
    (the MLP adds gate * 0), so the learned flux must equal the oracle's
    learned flux to fp32 rounding on every face.  The y and z sweeps read
    their windows from an LDS ring of rows that a workgroup marches through
    16 planes; 37 y faces and 35 z faces give each sweep three runs (the last
    one partial), so a ring slot or row off by one would move a face's
    states by a cell's difference, O(1e-2) of the flux here."""
    g = ics.Grid(20, 36, 34, ng=4)
    st = ics.initial_state(g, "orszag_tang", 34)
    w = np.zeros(sg.NPARAM, np.float32)
    s = gpu.Solver(gpu.make_params(20, 36, 34, ng=4, reconop=1, stepop=1, bc="even", precision="mixed",
                                   surrogate=1), g)
    sg.set_weights(s, w)
    sg.configure(s, 1e30, count=True)
    s.upload(st)
    s.fill()
    f = s.flux_stress()
    faces, fb = sg.stats(s)
    s.close()
    _, ref = _oracle_pair(g, st, w, 1e30)
    assert faces > 0 and fb == 0, (faces, fb)
    worst = 0.
    for q in range(24):
        scale = np.abs(ref[q]).max() + 1e-300
        err = float(np.abs(f[q] - ref[q]).max()) / scale
        worst = max(worst, err)
        assert err <= 1e-5, (q, err)
    print("linear part vs oracle: largest flux error %.2e of the array's scale" % worst)


TRAIN_THETA = .5


@pytest.fixture(scope="module")
def trained():
    """a net trained on 64^3 snapshots (small version of the bench's set).
    Faces up to roughness 0.5 are collected -- twice the solver's threshold --
    so the set holds the windows where PDM departs from the linear part."""
    s, n = sg.build_training_set(n=64, capacity=1 << 20, theta=TRAIN_THETA, stride=16, snapshots=3,
                                 steps_between=6)
    held = sg.build_training_set(n=64, capacity=1 << 16, theta=TRAIN_THETA, stride=7, snapshots=1, steps_between=4,
                                 setups=(("orszag_tang", 901, {}), ("blast", 902, {"radius": .15})))
    hs, hn = held
    held_samples = sg.fetch_samples(hs, hn)
    hs.close()
    samples = sg.fetch_samples(s, min(n, 4096))
    # the first call initialises the net and takes a short epoch-long run
    loss1 = sg.train(s, epochs=1, batch=4096, lr=2e-3, seed=1)
    w1 = sg.get_weights(s)
    loss = sg.train(s, epochs=5, batch=1024, lr=5e-3, seed=2)
    w = sg.get_weights(s)
    s.close()
    print("training loss: first epoch %.3e, final %.3e" % (loss1, loss))
    return dict(w=w, w1=w1, loss=loss, n=n, samples=samples, held=held_samples)


@pytest.mark.gpu
def test_collected_samples_are_consistent(trained):
    smp = trained["samples"]
    x = smp[:, :40].reshape(-1, 8, 5)
    t = smp[:, 40:]
    assert trained["n"] > 500000
    assert np.all(np.abs(x) <= 1. + 1e-6)
    rough = np.abs(x[:, :, :-2] - 2 * x[:, :, 1:-1] + x[:, :, 2:]).max(axis=(1, 2))
    assert np.all(rough <= TRAIN_THETA + 1e-6)
    lo = np.minimum(x[:, :, 1], x[:, :, 2]) - 1e-5
    hi = np.maximum(x[:, :, 1], x[:, :, 2]) + 1e-5
    for side in (0, 8):
        y = t[:, side:side + 8]
        assert np.all((y >= lo) & (y <= hi))


@pytest.mark.gpu
def test_training_is_deterministic():
    """collection writes the samples in face order and the gradient adds its
    partial sums in a fixed order, so two runs from the same snapshots train
    the same net bit for bit (also when the set fills up and samples are
    dropped)"""
    nets = []
    for _ in range(2):
        s, n = sg.build_training_set(n=32, capacity=20000, theta=TRAIN_THETA, stride=3, snapshots=2,
                                     steps_between=3)
        smp = sg.fetch_samples(s, min(n, 2048))
        sg.train(s, epochs=1, batch=1024, lr=5e-3, seed=3)
        nets.append((n, smp, sg.get_weights(s)))
        s.close()
    (n1, s1, w1), (n2, s2, w2) = nets
    assert n1 == n2 == 20000
    assert np.array_equal(s1, s2)
    assert np.array_equal(w1, w2)


@pytest.mark.gpu
def test_trained_net_gains_over_baselines(trained):
    held = trained["held"]
    x, t = held[:, :40], held[:, 40:]
    # outputs clamped to the two cells' range, as the solver uses them
    net = sg.clamp_to_cells(x, po.mlp_forward(trained["w"], x))
    lin = sg.clamp_to_cells(x, sg.linear_baseline(x))
    early = sg.clamp_to_cells(x, po.mlp_forward(trained["w1"], x))
    e_net = float(np.mean((net - t) ** 2))
    e_lin = float(np.mean((lin - t) ** 2))
    e_early = float(np.mean((early - t) ** 2))
    print("held-out MSE (normalised): net %.3e, after 1 epoch %.3e, 4th-order linear %.3e" % (e_net, e_early, e_lin))
    # the trained correction removes most of the linear interpolation's error
    assert e_net < .5 * e_lin
    assert e_net < e_early


@pytest.mark.gpu
def test_solver_with_learned_reconstruction(trained):
    """48^3 Orszag-Tang, AB2 (CFL dt): the mixed solver with the trained net
    against the binary64 solver, next to the same solver with the linear part
    only (zero MLP) and the mixed PDM solver.

    What this measures, stated plainly: the trained correction does not beat
    its own linear part inside the solver.  At theta = 0.25 the seeded density
    noise of this set-up sends ~99 % of the faces to the PDM fallback, the net
    acts on the rest, and its density L1 sits at or above the linear part's
    (1.34x trained here on 2^20 samples -- training is deterministic, so
    the figure is this build's; 1.00x after 8 epochs on 2^24); both
    are two orders of magnitude above fp32 PDM.  The bounds below
    are regression guards on that state, not a claimed gain: the net within
    1.5x of the linear part and below 1e-4, PDM below both."""
    n = 48
    g = ics.Grid(n, n, n, ng=4)
    st = ics.initial_state(g, "orszag_tang", 903)
    ref = gpu.Solver(gpu.make_params(n, n, n, reconop=1, stepop=1, bc="periodic"), g)
    runs = {}
    zero = np.zeros(sg.NPARAM, np.float32)
    for name, w in (("net", trained["w"]), ("linear", zero), ("pdm", None)):
        sur = gpu.Solver(gpu.make_params(n, n, n, reconop=1, stepop=1, bc="periodic", precision="mixed",
                                         surrogate=0 if w is None else 1), g)
        if w is not None:
            sg.set_weights(sur, w)
            sg.configure(sur, sg.THETA, count=True)
        sur.upload(st)
        tau = sur.initialize(1e-4)
        sur.step(20, tau)
        runs[name] = (sur.download(), sg.stats(sur) if w is not None else (1, 0), sur.check_state())
        sur.close()
    ref.upload(st)
    tau = ref.initialize(1e-4)
    ref.step(20, tau)
    a = ref.download()
    ref.close()
    l1 = {}
    for name, (b, (faces, fb), flags) in runs.items():
        l1[name] = np.abs(g.interior(a["rho"]) - g.interior(b["rho"])).sum() / np.abs(g.interior(a["rho"])).sum()
        print("%s: L1(rho) %.3e after 20 steps, fallback %.1f%% of %d faces" % (name, l1[name],
                                                                               100. * fb / max(faces, 1), faces))
        assert faces > 0 and fb < faces
        assert flags == 0
    assert l1["net"] <= 1e-4
    assert l1["net"] <= 1.5 * l1["linear"]
    assert l1["pdm"] < min(l1["net"], l1["linear"])


@pytest.mark.gpu
@pytest.mark.parametrize("grid", ["cartesian", "sheared", "shell"])
def test_all_fallback_equals_mixed_pdm(grid):
    """theta < 0 sends every face to the PDM fallback list: the learned path
    (flux kernel, fix-up of the fallback lists) must then reproduce the
    mixed-precision PDM path, on the reference's grid, the affine sheared grid
    and the spherical shell's node grid.  The two are different kernels (the
    plain flux reconstructs the strided sweeps from rows staged in LDS, the
    fix-up walks its lists through global memory), and fp32 contraction may
    round them differently in the last place (and a component that cancels,
    such as the z energy flux of the Orszag-Tang vortex, keeps that rounding
    at the scale of its terms), so the fluxes of one state must agree to 1e-5
    of the largest flux of their direction and a 6-step trajectory to 1e-5; a
    face the fix-up missed or misplaced would differ at O(1)."""
    from mhd3d_mpi_fortran_amd import shell
    if grid == "shell":
        g = shell.spherical_shell(16, 16, 24, ng=4, theta0=.6)
        states = [shell.dipole_inflow(g, 9)]
        shp, bc = (16, 16, 24), ("even", "even", "periodic")
    else:
        A = [[1., .5, 0.], [0., 1., .2], [0., 0., 1.]] if grid == "sheared" else None
        g = ics.Grid(20, 16, 12, ng=4, A=A)
        states = [ics.initial_state(g, "blast", 13, radius=.2)]
        if A is None:   # (the Orszag-Tang vortex is set up on the reference's grid only)
            states.append(ics.initial_state(g, "orszag_tang", 14))
        shp, bc = (20, 16, 12), "periodic"
    for n_state, st in enumerate(states):
        flux, out = [], []
        for sur in (0, 1):
            s = gpu.Solver(gpu.make_params(*shp, ng=4, reconop=1, stepop=1, bc=bc, precision="mixed", surrogate=sur),
                           g)
            if sur:
                sg.set_weights(s, _weights(11, 1.))
                sg.configure(s, -1., count=True)
            s.upload(st)
            s.fill()
            flux.append(s.flux_stress())
            if n_state == 0:
                s.upload(st)
                tau = s.initialize(1e-4)
                s.step(6, tau)
                out.append(s.download())
            if sur:
                faces, fb = sg.stats(s)
                assert faces > 0 and fb == faces
            s.close()
        for q in range(24):
            scale = max(float(np.abs(flux[0][8 * (q // 8) + c]).max()) for c in range(8))
            assert np.isfinite(flux[1][q]).all()
            assert np.abs(flux[0][q] - flux[1][q]).max() <= 1e-5 * scale, (n_state, q)
        for k in (out[0] if out else ()):
            scale = np.abs(out[0][k]).max() + 1e-300
            assert np.abs(out[0][k] - out[1][k]).max() <= 1e-5 * scale, k
