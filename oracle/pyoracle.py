"""ctypes binding of the CPU oracle (oracle/mhd_oracle.c).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build(force=False):
    """compile the oracle: no FMA contraction, no -march, like the reference build."""
    src = os.path.join(HERE, "mhd_oracle.c")
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(src):
        subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
                               "-shared", src, "-o", LIB, "-lm"])
    return LIB


class Params(C.Structure):
    _fields_ = [("nx0", C.c_int), ("ny0", C.c_int), ("nz0", C.c_int), ("ng", C.c_int),
                ("gam", C.c_double), ("cfl", C.c_double), ("pdmb", C.c_double), ("eps", C.c_double),
                ("alpha", C.c_double), ("beta", C.c_double), ("reconop", C.c_int), ("stepop", C.c_int),
                ("bc", C.c_int * 3)]


_lib = None
_D = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(Params)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_geometry.argtypes = [C.c_void_p, _D, _D, _D]
        L.orc_set_state.argtypes = [C.c_void_p] + [_D] * 8
        L.orc_get_state.argtypes = [C.c_void_p] + [_D] * 11
        L.orc_initialize.argtypes = [C.c_void_p, _D]
        L.orc_fill.argtypes = [C.c_void_p]
        L.orc_bxyz.argtypes = [C.c_void_p]
        L.orc_get_dt.argtypes = [C.c_void_p]
        L.orc_get_dt.restype = C.c_double
        L.orc_flux_stress.argtypes = [C.c_void_p]
        L.orc_eijk.argtypes = [C.c_void_p]
        L.orc_get_ops.argtypes = [C.c_void_p, _D, _D, _D, _D]
        L.orc_set_tau.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_get_tau.argtypes = [C.c_void_p, _D, _D]
        L.orc_predict_ab.argtypes = [C.c_void_p]
        L.orc_predict_ab.restype = C.c_int
        L.orc_flux_stress_state.argtypes = [C.c_void_p, C.c_int]
        L.orc_eijk_state.argtypes = [C.c_void_p, C.c_int]
        L.orc_eijk_yee_state.argtypes = [C.c_void_p, C.c_int]
        L.orc_advance.argtypes = [C.c_void_p, C.c_int]
        L.orc_advance.restype = C.c_int
        L.orc_set_ops.argtypes = [C.c_void_p, _D, _D, _D, _D]
        L.orc_step.argtypes = [C.c_void_p, C.c_int, _D, _D]
        L.orc_step.restype = C.c_int
        L.orc_reconstruct.argtypes = [C.POINTER(Params), _D, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_int, _D, _D]
        L.orc_v_face.argtypes = [C.c_int, _D, _D, _D, _D, _D, _D]
        L.orc_set_threads.argtypes = [C.c_int]
        _F = C.POINTER(C.c_float)
        L.orc_mlp_forward.argtypes = [_F, _F, C.c_long, C.c_int, _F]
        L.orc_mlp_loss_grad.argtypes = [_F, _F, C.c_long, _D]
        L.orc_set_surrogate.argtypes = [C.c_void_p, _F, C.c_double, C.c_int]
        L.orc_mlp_loss_grad.restype = C.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_D) if a is not None else None


BC = {"even": 0, "periodic": 2}


def make_params(nx0, ny0, nz0, ng=4, reconop=1, stepop=2, bc="periodic", gam=5. / 3., cfl=.3, pdmb=4.,
                eps=1e-8, alpha=2., beta=4.):
    p = Params()
    p.nx0, p.ny0, p.nz0, p.ng = nx0, ny0, nz0, ng
    p.gam, p.cfl, p.pdmb, p.eps, p.alpha, p.beta = gam, cfl, pdmb, eps, alpha, beta
    p.reconop, p.stepop = reconop, stepop
    b = BC[bc] if isinstance(bc, str) else bc
    bcs = b if isinstance(b, (tuple, list)) else (b, b, b)
    for q in range(3):
        p.bc[q] = bcs[q]
    return p


class Oracle:
    """one block of the reference step on the CPU."""

    def __init__(self, params, grid):
        self.p = params
        self.g = grid
        self.h = lib().orc_create(C.byref(params))
        self._keep = [np.ascontiguousarray(grid.dx), np.ascontiguousarray(grid.dy), np.ascontiguousarray(grid.dz)]
        lib().orc_set_geometry(self.h, *[_p(a) for a in self._keep])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_state(self, st):
        arrs = [np.asfortranarray(st[k], dtype=np.float64) for k in ("rho", "vx", "vy", "vz", "p", "bi", "bj", "bk")]
        lib().orc_set_state(self.h, *[_p(a) for a in arrs])

    def get_state(self):
        g = self.g
        cs = (g.nx, g.ny, g.nz)
        out = {k: np.zeros(cs, order="F") for k in ("rho", "vx", "vy", "vz", "p", "bx", "by", "bz")}
        out["bi"] = np.zeros((g.nx + 1, g.ny, g.nz), order="F")
        out["bj"] = np.zeros((g.nx, g.ny + 1, g.nz), order="F")
        out["bk"] = np.zeros((g.nx, g.ny, g.nz + 1), order="F")
        order = ("rho", "vx", "vy", "vz", "p", "bi", "bj", "bk", "bx", "by", "bz")
        lib().orc_get_state(self.h, *[_p(out[k]) for k in order])
        return out

    def initialize(self, tau):
        t = C.c_double(tau)
        lib().orc_initialize(self.h, C.byref(t))
        return t.value

    def fill(self):
        lib().orc_fill(self.h)

    def bxyz(self):
        lib().orc_bxyz(self.h)

    def get_dt(self):
        return lib().orc_get_dt(self.h)

    def set_tau(self, tau, tau0):
        lib().orc_set_tau(self.h, float(tau), float(tau0))

    def get_tau(self):
        t, t0 = C.c_double(0.), C.c_double(0.)
        lib().orc_get_tau(self.h, C.byref(t), C.byref(t0))
        return t.value, t0.value

    def predict_ab(self):
        if lib().orc_predict_ab(self.h):
            raise RuntimeError("orc_predict_ab: not an Adams-Bashforth oracle")

    def advance(self, stage):
        if lib().orc_advance(self.h, int(stage)):
            raise RuntimeError("orc_advance: stage %d not in this stepper" % stage)

    def set_ops(self, ops):
        """store face fluxes and edge EMFs (the layout of ops())"""
        flat = np.concatenate([np.asarray(f, dtype=np.float64).ravel(order="F") for f in ops["flux"]])
        arrs = [np.asfortranarray(ops[k], dtype=np.float64) for k in ("ei", "ej", "ek")]
        lib().orc_set_ops(self.h, *[_p(a) for a in arrs], _p(flat))

    def ops(self, which=0, compute=True, yee=False):
        """run get_eijk_LFM (yee=True: get_eijk_Yee) + get_flux_stress on the
        current state (which=0) or the Adams-Bashforth extrapolation
        (which=1) and return the products; compute=False returns the held
        products as they are."""
        if compute:
            if yee:
                lib().orc_eijk_yee_state(self.h, int(which))
            else:
                lib().orc_eijk_state(self.h, int(which))
            lib().orc_flux_stress_state(self.h, int(which))
        g = self.g
        n0 = (g.nx0, g.ny0, g.nz0)
        ei = np.zeros((n0[0], n0[1] + 1, n0[2] + 1), order="F")
        ej = np.zeros((n0[0] + 1, n0[1], n0[2] + 1), order="F")
        ek = np.zeros((n0[0] + 1, n0[1] + 1, n0[2]), order="F")
        shapes = [(n0[0] + (d == 0), n0[1] + (d == 1), n0[2] + (d == 2)) for d in range(3) for _ in range(8)]
        tot = sum(int(np.prod(s)) for s in shapes)
        flat = np.zeros(tot)
        lib().orc_get_ops(self.h, _p(ei), _p(ej), _p(ek), _p(flat))
        flux, off = [], 0
        for s in shapes:
            n = int(np.prod(s))
            flux.append(flat[off:off + n].reshape(s, order="F"))
            off += n
        return dict(ei=ei, ej=ej, ek=ek, flux=flux)

    def set_surrogate(self, w, theta=1.0, bf16_round=False):
        """learned face states in the flux (PDM fallback); w=None: plain PDM"""
        if w is None:
            lib().orc_set_surrogate(self.h, None, 0., 0)
            return
        self._sur_w = np.ascontiguousarray(w, dtype=np.float32)
        lib().orc_set_surrogate(self.h, _pf(self._sur_w), float(theta), 1 if bf16_round else 0)

    def step(self, nsteps, tau):
        t = C.c_double(tau)
        tc = C.c_double(0.)
        rc = lib().orc_step(self.h, nsteps, C.byref(t), C.byref(tc))
        if rc:
            raise RuntimeError("orc_step failed (%d)" % rc)
        return t.value, tc.value


def reconstruct(params, f, lshape, direction):
    f = np.asfortranarray(f, dtype=np.float64)
    fl = np.zeros(lshape, order="F")
    fr = np.zeros(lshape, order="F")
    lib().orc_reconstruct(C.byref(params), _p(f), f.shape[0], f.shape[1], f.shape[2], lshape[0], lshape[1],
                          lshape[2], direction, _p(fl), _p(fr))
    return fl, fr


def v_face(lam, v):
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    outs = [np.zeros_like(lam) for _ in range(4)]
    lib().orc_v_face(lam.size, _p(lam), _p(v), *[_p(o) for o in outs])
    return outs


NPARAM = 2896


def _pf(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def mlp_forward(w, x, h_round=False):
    """MLP forward of rows of 40 inputs (float32 in, float32 out)."""
    w = np.ascontiguousarray(w, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = x.shape[0]
    y = np.zeros((n, 16), dtype=np.float32)
    lib().orc_mlp_forward(_pf(w), _pf(x), n, 1 if h_round else 0, _pf(y))
    return y


def mlp_loss_grad(w, samples):
    """mean squared error and its gradient (double accumulation)."""
    w = np.ascontiguousarray(w, dtype=np.float32)
    s = np.ascontiguousarray(samples, dtype=np.float32)
    g = np.zeros(NPARAM)
    loss = lib().orc_mlp_loss_grad(_pf(w), _pf(s), s.shape[0], _p(g))
    return loss, g
