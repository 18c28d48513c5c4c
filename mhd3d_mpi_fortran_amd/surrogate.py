"""Learned face reconstruction: sample collection, GPU training, weights.

Drives the surrogate entry points of libmhd_gpu (include/mhd_gpu.h):
face samples come from the solver's own binary64 states (normalised 5-point
stencils of the 8 primitives + PDM's left/right states as targets), the
network is trained on the GPU by the library's forward/backward/Adam
kernels, and a MHD_MIXED context created with ``surrogate=1`` evaluates it
with bf16 MFMA inside the fp32 face flux, falling back to PDM where the
smoothness or positivity check fails.
"""
import ctypes as C

import numpy as np

from . import gpu, ics

NPARAM = 2896
NSAMPLE = 56
THETA = 0.25         # smoothness threshold used for training and inference


def _pf(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def set_weights(solver, w):
    w = np.ascontiguousarray(w, dtype=np.float32)
    gpu._ck(gpu.lib().mhd_surrogate_set_weights(solver.h, _pf(w), NPARAM), "mhd_surrogate_set_weights")


def get_weights(solver):
    w = np.zeros(NPARAM, dtype=np.float32)
    gpu._ck(gpu.lib().mhd_surrogate_get_weights(solver.h, _pf(w), NPARAM), "mhd_surrogate_get_weights")
    return w


def configure(solver, theta=THETA, count=False):
    gpu._ck(gpu.lib().mhd_surrogate_configure(solver.h, float(theta), 1 if count else 0), "mhd_surrogate_configure")


def stats(solver):
    f, b = C.c_long(0), C.c_long(0)
    gpu._ck(gpu.lib().mhd_surrogate_stats(solver.h, C.byref(f), C.byref(b)), "mhd_surrogate_stats")
    return f.value, b.value


def collect(solver, theta=THETA, stride=4, capacity=1 << 24):
    n = C.c_long(0)
    gpu._ck(gpu.lib().mhd_surrogate_collect(solver.h, float(theta), int(stride), int(capacity), C.byref(n)),
            "mhd_surrogate_collect")
    return n.value


def fetch_samples(solver, n):
    out = np.zeros((n, NSAMPLE), dtype=np.float32)
    gpu._ck(gpu.lib().mhd_surrogate_fetch_samples(solver.h, _pf(out), n), "mhd_surrogate_fetch_samples")
    return out


def load_samples(solver, samples):
    s = np.ascontiguousarray(samples, dtype=np.float32)
    gpu._ck(gpu.lib().mhd_surrogate_load_samples(solver.h, _pf(s), s.shape[0]), "mhd_surrogate_load_samples")


def save_weights(w, path):
    """raw float32 weights (the Fortran driver's &gpu weights= file)"""
    np.ascontiguousarray(w, dtype="<f4").tofile(path)


def load_weights(path):
    w = np.fromfile(path, dtype="<f4")
    if w.size != NPARAM:
        raise ValueError("%s holds %d values, expected %d" % (path, w.size, NPARAM))
    return w


def train(solver, epochs=4, batch=1024, lr=5e-3, seed=0):
    loss = C.c_double(0.)
    gpu._ck(gpu.lib().mhd_surrogate_train(solver.h, int(epochs), int(batch), float(lr), int(seed), C.byref(loss)),
            "mhd_surrogate_train")
    return loss.value


def grad(solver, w, samples):
    w = np.ascontiguousarray(w, dtype=np.float32)
    s = np.ascontiguousarray(samples, dtype=np.float32)
    g = np.zeros(NPARAM, dtype=np.float32)
    loss = C.c_double(0.)
    gpu._ck(gpu.lib().mhd_surrogate_grad(solver.h, _pf(w), _pf(s), s.shape[0], C.byref(loss), _pf(g)),
            "mhd_surrogate_grad")
    return loss.value, g


def infer(solver, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros((x.shape[0], 16), dtype=np.float32)
    gpu._ck(gpu.lib().mhd_surrogate_infer(solver.h, _pf(x), x.shape[0], _pf(y)), "mhd_surrogate_infer")
    return y


# training set-ups: the Orszag-Tang and Gaussian-blob boxes and blast waves,
# seeds distinct from every benchmark and test state
SETUPS = (("orszag_tang", 101, {}), ("blobs", 102, {}), ("blast", 103, {"radius": .15}),
          ("orszag_tang", 104, {}), ("blobs", 105, {}), ("blast", 106, {"radius": .1}))


def build_training_set(n=128, capacity=1 << 24, theta=THETA, stride=4, snapshots=6, steps_between=8,
                       setups=SETUPS, log=None):
    """run the binary64 solver (PDM, RK2, periodic) on the training set-ups
    and collect face samples from several snapshots of each; returns the
    solver holding the samples (the training context) and the sample count."""
    g = ics.Grid(n, n, n, ng=4)
    s = gpu.Solver(gpu.make_params(n, n, n, ng=4, reconop=1, stepop=2, bc="periodic"), g)
    total = 0
    for kind, seed, kw in setups:
        st = ics.initial_state(g, kind, seed, **kw)
        s.upload(st)
        tau = s.initialize(1e-4)
        tau = ics.cfl_dt_estimate(g, st) * .8
        for snap in range(snapshots):
            total = collect(s, theta, stride, capacity)
            if log:
                log("collect %s seed %d snapshot %d: %d samples" % (kind, seed, snap, total))
            if total >= capacity:
                return s, capacity
            _, tcfl = s.step(steps_between, tau)
            tau = min(tau, .8 * tcfl) if tcfl > 0 else tau
    return s, total


def train_surrogate(n=128, capacity=1 << 24, epochs=2, batch=1024, lr=5e-3, seed=0, stride=8, log=None):
    """collect samples and train; returns (weights, final mean loss, samples)."""
    s, total = build_training_set(n=n, capacity=capacity, stride=stride, log=log)
    loss = train(s, epochs=epochs, batch=batch, lr=lr, seed=seed)
    w = get_weights(s)
    s.close()
    return w, loss, total


def net_forward(w, x):
    """the net in numpy (float64): lin + gate * MLP, as the library defines it"""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 40)
    w = np.asarray(w, dtype=np.float64)
    act = lambda z: z / (1. + np.exp(-1.702 * z))
    a1 = act(x @ w[:1280].reshape(32, 40).T + w[1280:1312])
    a2 = act(a1 @ w[1312:2336].reshape(32, 32).T + w[2336:2368])
    m = a2 @ w[2368:2880].reshape(16, 32).T + w[2880:2896]
    x5 = x.reshape(-1, 8, 5)
    gate = np.abs(x5[:, :, :-2] - 2 * x5[:, :, 1:-1] + x5[:, :, 2:]).max(axis=2)
    return linear_baseline(x) + np.concatenate([gate, gate], axis=1) * m


def linear_baseline(x):
    """normalised left/right states of the 4th-order centred interpolation
    (-f(j-2) + 7 f(j-1) + 7 f(j) - f(j+1))/12 on the same 5-point window --
    the reconstruction available without the net at this stencil width."""
    x = x.reshape(-1, 8, 5)
    v = (-x[:, :, 0] + 7. * x[:, :, 1] + 7. * x[:, :, 2] - x[:, :, 3]) / 12.
    return np.concatenate([v, v], axis=1)


def clamp_to_cells(x, y):
    """clamp normalised left/right states to [min, max] of the two cells next
    to the face (inputs 1 and 2 of each window), as the solver does"""
    x = x.reshape(-1, 8, 5)
    lo = np.minimum(x[:, :, 1], x[:, :, 2])
    hi = np.maximum(x[:, :, 1], x[:, :, 2])
    return np.clip(y, np.concatenate([lo, lo], axis=1), np.concatenate([hi, hi], axis=1))
