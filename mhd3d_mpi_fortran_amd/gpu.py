"""ctypes binding of libmhd_gpu.so (C-ABI in include/mhd_gpu.h).

The Python mirror of the Fortran driver's view of the library: a ``Solver``
owns one context (one block on one GPU) and exposes the reference's per-step
routines by their reference names (get_dt, fill_hydro_mpi+fill_mag_mpi,
get_bxyz, get_flux_stress, get_eijk_LFM, step).  Errors raise ``MhdError``
carrying the library's status code; the library is never replaced by a
fallback: if libmhd_gpu.so is missing this module raises on import of the
binding.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(PKG, "libmhd_gpu.so")

MHD_OK, MHD_ERR_HIP, MHD_ERR_PARAM, MHD_ERR_STATE, MHD_ERR_UNSUPPORTED = 0, 1, 2, 3, 4
BC = {"even": 0, "periodic": 2}
PRECISION = {"fp64": 0, "mixed": 1}
KCLASS = {"flux": 0, "emf": 1, "update": 2, "ct": 3, "fill": 4, "bxyz": 5, "dt": 6, "prev": 7, "surrogate": 8}
FETCH_EI, FETCH_EJ, FETCH_EK, FETCH_VF = 24, 25, 26, 27
STATE_CURRENT, STATE_PREDICTED = 0, 1


class MhdError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        super().__init__("%s failed: %s (code %d)" % (where, _errstr(code), code))


class Params(C.Structure):
    _fields_ = [("nx0", C.c_int), ("ny0", C.c_int), ("nz0", C.c_int), ("ng", C.c_int),
                ("gam", C.c_double), ("cfl", C.c_double), ("pdmb", C.c_double), ("eps", C.c_double),
                ("alpha", C.c_double), ("beta", C.c_double), ("reconop", C.c_int), ("stepop", C.c_int),
                ("bc", C.c_int * 3), ("precision", C.c_int), ("surrogate", C.c_int), ("device", C.c_int)]


_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_lib = None

# every entry point of include/mhd_gpu.h with its ctypes signature
SIGNATURES = {
    "mhd_default_params": (None, [C.POINTER(Params)]),
    "mhd_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(Params)]),
    "mhd_destroy": (None, [C.c_void_p]),
    "mhd_error_string": (C.c_char_p, [C.c_int]),
    "mhd_set_geometry": (C.c_int, [C.c_void_p, _D, _D, _D]),
    "mhd_set_metric": (C.c_int, [C.c_void_p, _D]),
    "mhd_set_metric_nodes": (C.c_int, [C.c_void_p, _D, _D, _D]),
    "mhd_metric_nodes_host": (C.c_long, [C.c_int, C.c_int, C.c_int, C.c_int, _D, _D, _D, _D]),
    "mhd_upload_state": (C.c_int, [C.c_void_p] + [_D] * 8),
    "mhd_download_state": (C.c_int, [C.c_void_p] + [_D] * 11),
    "mhd_initialize": (C.c_int, [C.c_void_p, _D]),
    "mhd_step": (C.c_int, [C.c_void_p, C.c_int, _D, _D]),
    "mhd_get_dt": (C.c_int, [C.c_void_p, _D]),
    "mhd_fill_ghosts": (C.c_int, [C.c_void_p]),
    "mhd_get_bxyz": (C.c_int, [C.c_void_p]),
    "mhd_flux_stress": (C.c_int, [C.c_void_p, C.c_int]),
    "mhd_eijk_lfm": (C.c_int, [C.c_void_p, C.c_int]),
    "mhd_eijk_yee": (C.c_int, [C.c_void_p, C.c_int]),
    "mhd_flux_stress_old": (C.c_int, [C.c_void_p]),
    "mhd_set_tau": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "mhd_get_tau": (C.c_int, [C.c_void_p, _D, _D]),
    "mhd_predict_ab": (C.c_int, [C.c_void_p]),
    "mhd_advance": (C.c_int, [C.c_void_p, C.c_int]),
    "mhd_fetch": (C.c_int, [C.c_void_p, C.c_int, _D]),
    "mhd_store": (C.c_int, [C.c_void_p, C.c_int, _D]),
    "mhd_reconstruct_noalloc": (C.c_int, [C.c_void_p, C.c_int, _D, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_int, _D, _D]),
    "mhd_v_face": (C.c_int, [C.c_void_p, C.c_long, _D, _D, _D, _D, _D, _D]),
    "mhd_reconstruct": (C.c_int, [C.c_void_p, C.c_int, _D, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  _D, _D]),
    "mhd_check_state": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "mhd_synchronize": (C.c_int, [C.c_void_p]),
    "mhd_device_pointer": (C.c_void_p, [C.c_void_p, C.c_int]),
    "mhd_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "mhd_profile_read": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), _D]),
    "mhd_profile_classes": (C.c_int, [C.c_void_p, C.c_uint]),
    "mhd_set_stage_overlap": (C.c_int, [C.c_void_p, C.c_int]),
    "mhd_set_fusion": (C.c_int, [C.c_void_p, C.c_int]),
    "mhd_divb_max": (C.c_int, [C.c_void_p, _D]),
    "mhd_ensemble_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int, _D]),
    "mhd_ensemble_step": (C.c_int, [C.c_void_p, C.c_int]),
    "mhd_ensemble_destroy": (None, [C.c_void_p]),
    "mhd_surrogate_set_weights": (C.c_int, [C.c_void_p, _F, C.c_int]),
    "mhd_surrogate_get_weights": (C.c_int, [C.c_void_p, _F, C.c_int]),
    "mhd_surrogate_configure": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    "mhd_surrogate_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    "mhd_surrogate_collect": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_long, C.POINTER(C.c_long)]),
    "mhd_surrogate_load_samples": (C.c_int, [C.c_void_p, _F, C.c_long]),
    "mhd_surrogate_fetch_samples": (C.c_int, [C.c_void_p, _F, C.c_long]),
    "mhd_surrogate_train": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_uint, _D]),
    "mhd_surrogate_grad": (C.c_int, [C.c_void_p, _F, _F, C.c_int, _D, _F]),
    "mhd_surrogate_infer": (C.c_int, [C.c_void_p, _F, C.c_long, _F]),
}


def lib():
    """load libmhd_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise OSError("libmhd_gpu.so not built: run __graft_entry__.build() or "
                          "python mhd3d_mpi_fortran_amd/build.py")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _errstr(code):
    try:
        return lib().mhd_error_string(code).decode()
    except OSError:
        return "?"


def _ck(rc, where):
    if rc != MHD_OK:
        raise MhdError(rc, where)


def _p(a):
    return None if a is None else a.ctypes.data_as(_D)


def make_params(nx0, ny0, nz0, ng=4, reconop=1, stepop=2, bc="periodic", precision="fp64", surrogate=0,
                gam=5. / 3., cfl=.3, pdmb=4., eps=1e-8, alpha=2., beta=4., device=0):
    p = Params()
    lib().mhd_default_params(C.byref(p))
    p.nx0, p.ny0, p.nz0, p.ng = nx0, ny0, nz0, ng
    p.gam, p.cfl, p.pdmb, p.eps, p.alpha, p.beta = gam, cfl, pdmb, eps, alpha, beta
    p.reconop, p.stepop = reconop, stepop
    b = BC[bc] if isinstance(bc, str) else bc
    bcs = b if isinstance(b, (tuple, list)) else (b, b, b)
    for q in range(3):
        p.bc[q] = BC[bcs[q]] if isinstance(bcs[q], str) else bcs[q]
    p.precision = PRECISION[precision] if isinstance(precision, str) else precision
    p.surrogate = int(surrogate)
    p.device = device
    return p


class Solver:
    """one block of the MHD step on one GPU (a libmhd_gpu context)."""

    STATE = ("rho", "vx", "vy", "vz", "p", "bi", "bj", "bk")

    def __init__(self, params, grid):
        self.params = params
        self.grid = grid
        h = C.c_void_p()
        _ck(lib().mhd_create(C.byref(h), C.byref(params)), "mhd_create")
        self.h = h
        dx = np.ascontiguousarray(grid.dx)
        dy = np.ascontiguousarray(grid.dy)
        dz = np.ascontiguousarray(grid.dz)
        _ck(lib().mhd_set_geometry(self.h, _p(dx), _p(dy), _p(dz)), "mhd_set_geometry")
        if getattr(grid, "nodes", None) is not None:
            self.set_metric_nodes(*grid.nodes)
        elif getattr(grid, "A", None) is not None:
            self.set_metric(grid.A)

    def set_metric(self, A):
        """affine metric x_phys = A x_comp (row-major 3x3)"""
        a = np.ascontiguousarray(np.asarray(A, dtype=np.float64).reshape(3, 3))
        _ck(lib().mhd_set_metric(self.h, _p(a)), "mhd_set_metric")

    def set_metric_nodes(self, X, Y, Z):
        """general metric from the node coordinates, each (nx+1, ny+1, nz+1)
        ghost-inclusive, node (i,j,k) the lower corner of cell (i,j,k)"""
        g = self.grid
        shp = (g.nx + 1, g.ny + 1, g.nz + 1)
        xs = []
        for a in (X, Y, Z):
            a = np.asarray(a, dtype=np.float64)
            if a.shape != shp:
                raise ValueError("node array shape %s, expected %s" % (a.shape, shp))
            xs.append(np.asfortranarray(a))
        _ck(lib().mhd_set_metric_nodes(self.h, *[_p(a.ravel(order="F")) for a in xs]), "mhd_set_metric_nodes")

    def close(self):
        if getattr(self, "h", None):
            lib().mhd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---------------------------------------------------------------- state
    def _shapes(self):
        g = self.grid
        cs = (g.nx, g.ny, g.nz)
        return dict(rho=cs, vx=cs, vy=cs, vz=cs, p=cs, bx=cs, by=cs, bz=cs, bi=(g.nx + 1, g.ny, g.nz),
                    bj=(g.nx, g.ny + 1, g.nz), bk=(g.nx, g.ny, g.nz + 1))

    def upload(self, st):
        arrs = []
        for k, shp in zip(self.STATE, [self._shapes()[k] for k in self.STATE]):
            a = np.asfortranarray(st[k], dtype=np.float64)
            if a.shape != shp:
                raise ValueError("%s has shape %s, expected %s" % (k, a.shape, shp))
            arrs.append(a)
        _ck(lib().mhd_upload_state(self.h, *[_p(a) for a in arrs]), "mhd_upload_state")

    def download(self):
        shp = self._shapes()
        out = {k: np.zeros(s, order="F") for k, s in shp.items()}
        order = ("rho", "vx", "vy", "vz", "p", "bi", "bj", "bk", "bx", "by", "bz")
        _ck(lib().mhd_download_state(self.h, *[_p(out[k]) for k in order]), "mhd_download_state")
        return out

    # ------------------------------------------------------- reference routines
    def initialize(self, tau):
        t = C.c_double(tau)
        _ck(lib().mhd_initialize(self.h, C.byref(t)), "mhd_initialize")
        return t.value

    def step(self, nsteps=1, tau=0.):
        t = C.c_double(tau)
        tc = C.c_double(0.)
        _ck(lib().mhd_step(self.h, int(nsteps), C.byref(t), C.byref(tc)), "mhd_step")
        return t.value, tc.value

    def get_dt(self):
        t = C.c_double(0.)
        _ck(lib().mhd_get_dt(self.h, C.byref(t)), "mhd_get_dt")
        return t.value

    def fill(self):
        _ck(lib().mhd_fill_ghosts(self.h), "mhd_fill_ghosts")

    def bxyz(self):
        _ck(lib().mhd_get_bxyz(self.h), "mhd_get_bxyz")

    def flux_stress(self, which=STATE_CURRENT, compute=True):
        """get_flux_stress of the current state or (Adams-Bashforth, after
        predict_ab) the extrapolated one; the 24 face arrays.  compute=False
        fetches the held arrays."""
        if compute:
            _ck(lib().mhd_flux_stress(self.h, int(which)), "mhd_flux_stress")
        g = self.grid
        out = []
        for d in range(3):
            shp = (g.nx0 + (d == 0), g.ny0 + (d == 1), g.nz0 + (d == 2))
            for q in range(8):
                a = np.zeros(shp, order="F")
                _ck(lib().mhd_fetch(self.h, 8 * d + q, _p(a)), "mhd_fetch")
                out.append(a)
        return out

    def eijk(self, which=STATE_CURRENT, compute=True, yee=False):
        """get_eijk_LFM (or, yee=True, the reference's alternative
        get_eijk_Yee) of the current or the extrapolated state"""
        if compute:
            if yee:
                _ck(lib().mhd_eijk_yee(self.h, int(which)), "mhd_eijk_yee")
            else:
                _ck(lib().mhd_eijk_lfm(self.h, int(which)), "mhd_eijk_lfm")
        g = self.grid
        shp = {FETCH_EI: (g.nx0, g.ny0 + 1, g.nz0 + 1), FETCH_EJ: (g.nx0 + 1, g.ny0, g.nz0 + 1),
               FETCH_EK: (g.nx0 + 1, g.ny0 + 1, g.nz0)}
        out = {}
        for w, name in ((FETCH_EI, "ei"), (FETCH_EJ, "ej"), (FETCH_EK, "ek")):
            a = np.zeros(shp[w], order="F")
            _ck(lib().mhd_fetch(self.h, w, _p(a)), "mhd_fetch")
            out[name] = a
        return out

    def set_tau(self, tau, tau0):
        """the module variables tau and tau0 the update routines use"""
        _ck(lib().mhd_set_tau(self.h, float(tau), float(tau0)), "mhd_set_tau")

    def get_tau(self):
        t, t0 = C.c_double(0.), C.c_double(0.)
        _ck(lib().mhd_get_tau(self.h, C.byref(t), C.byref(t0)), "mhd_get_tau")
        return t.value, t0.value

    def predict_ab(self):
        """advance_predictor_ab (+ advance_previous, tau0 = tau)"""
        _ck(lib().mhd_predict_ab(self.h), "mhd_predict_ab")

    def advance(self, stage):
        """one stage's update from the held fluxes and EMFs: stage 0 the
        Adams-Bashforth corrector, 1..stepop a Runge-Kutta stage"""
        _ck(lib().mhd_advance(self.h, int(stage)), "mhd_advance")

    def store_ops(self, ops):
        """write face fluxes (ops['flux'], 24 arrays) and edge EMFs
        (ops['ei'/'ej'/'ek']) for the next advance"""
        for q, f in enumerate(ops["flux"]):
            a = np.asfortranarray(f, dtype=np.float64)
            _ck(lib().mhd_store(self.h, q, _p(a)), "mhd_store")
        for w, name in ((FETCH_EI, "ei"), (FETCH_EJ, "ej"), (FETCH_EK, "ek")):
            a = np.asfortranarray(ops[name], dtype=np.float64)
            _ck(lib().mhd_store(self.h, w, _p(a)), "mhd_store")

    def divb_max(self):
        """max |div B| over the owned cells (mhd_divb_max)"""
        v = C.c_double(0.)
        _ck(lib().mhd_divb_max(self.h, C.byref(v)), "mhd_divb_max")
        return v.value

    def check_state(self):
        f = C.c_int(0)
        rc = lib().mhd_check_state(self.h, C.byref(f))
        if rc not in (MHD_OK, MHD_ERR_STATE):
            raise MhdError(rc, "mhd_check_state")
        return f.value

    def synchronize(self):
        _ck(lib().mhd_synchronize(self.h), "mhd_synchronize")

    def profile(self, on=True):
        _ck(lib().mhd_profile(self.h, 1 if on else 0), "mhd_profile")

    def set_overlap(self, on=True):
        """EMF and CT on a second stream beside the flux and update (default)"""
        _ck(lib().mhd_set_stage_overlap(self.h, 1 if on else 0), "mhd_set_stage_overlap")

    def set_fusion(self, on=True):
        """fused face flux + update of mixed Adams-Bashforth PDM steps (off by default)"""
        _ck(lib().mhd_set_fusion(self.h, 1 if on else 0), "mhd_set_fusion")

    def profile_classes(self, classes):
        """time only the given kernel classes (names or MHD_K_* codes)"""
        m = 0
        for k in classes:
            m |= 1 << (KCLASS[k] if isinstance(k, str) else k)
        _ck(lib().mhd_profile_classes(self.h, m), "mhd_profile_classes")

    def profile_read(self, kclass):
        n = C.c_int(0)
        ms = C.c_double(0.)
        _ck(lib().mhd_profile_read(self.h, KCLASS[kclass] if isinstance(kclass, str) else kclass, C.byref(n),
                                   C.byref(ms)), "mhd_profile_read")
        return n.value, ms.value


class Ensemble:
    """independent Runge-Kutta solvers of any sizes and precisions stepped
    together, one HIP graph launch per step (mhd_ensemble_*); taus[m] is
    solver m's fixed step"""

    def __init__(self, solvers, taus):
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(self.solvers))(*[s.h for s in self.solvers])
        t = np.ascontiguousarray(taus, dtype=np.float64)
        if t.shape != (len(self.solvers),):
            raise ValueError("one tau per solver")
        h = C.c_void_p()
        _ck(lib().mhd_ensemble_create(C.byref(h), arr, len(self.solvers), _p(t)), "mhd_ensemble_create")
        self.h = h

    def step(self, nsteps=1):
        _ck(lib().mhd_ensemble_step(self.h, int(nsteps)), "mhd_ensemble_step")

    def close(self):
        if getattr(self, "h", None):
            lib().mhd_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def reconstruct(solver, f, lshape, direction, noalloc=False):
    """reconstruct_3d of a host array through the library (noalloc: the
    entry point that stages through the GPU's fixed scratch array)"""
    f = np.asfortranarray(f, dtype=np.float64)
    fl = np.zeros(lshape, order="F")
    fr = np.zeros(lshape, order="F")
    fn = lib().mhd_reconstruct_noalloc if noalloc else lib().mhd_reconstruct
    _ck(fn(solver.h, direction, _p(f), f.shape[0], f.shape[1], f.shape[2], lshape[0], lshape[1], lshape[2], _p(fl),
           _p(fr)), "mhd_reconstruct_noalloc" if noalloc else "mhd_reconstruct")
    return fl, fr


def v_face(solver, lam, v):
    """get_v_face through the library in the solver's flux arithmetic:
    (v0p, v0n, v1p, v1n)"""
    lam = np.ascontiguousarray(lam, dtype=np.float64).ravel()
    v = np.ascontiguousarray(v, dtype=np.float64).ravel()
    if lam.shape != v.shape:
        raise ValueError("lam and v differ in size")
    outs = [np.zeros_like(lam) for _ in range(4)]
    _ck(lib().mhd_v_face(solver.h, lam.size, _p(lam), _p(v), *[_p(o) for o in outs]), "mhd_v_face")
    return outs


def metric_nodes_host(grid):
    """the general metric's arrays of a node grid as the library builds
    them (host code, no device needed): dict with nf[f] (3, face shape),
    af[f], iaf[f], ef[e] (owned edge shape), ivol, cm (18, cell shape),
    nc (9, cell shape), ih (3, cell shape)"""
    nx0, ny0, nz0, ng = grid.nx0, grid.ny0, grid.nz0, grid.ng
    X, Y, Z = (np.asfortranarray(a, dtype=np.float64).ravel(order="F") for a in grid.nodes)
    n = lib().mhd_metric_nodes_host(nx0, ny0, nz0, ng, _p(X), _p(Y), _p(Z), None)
    if n < 0:
        raise MhdError("mhd_metric_nodes_host: %s" % _errstr(-n))
    out = np.zeros(n)
    r = lib().mhd_metric_nodes_host(nx0, ny0, nz0, ng, _p(X), _p(Y), _p(Z), _p(out))
    if r < 0:
        raise MhdError("mhd_metric_nodes_host: %s" % _errstr(-r))
    nx, ny, nz = grid.nx, grid.ny, grid.nz
    fshape = [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]
    eshape = [(nx0, ny0 + 1, nz0 + 1), (nx0 + 1, ny0, nz0 + 1), (nx0 + 1, ny0 + 1, nz0)]
    cs = (nx, ny, nz)
    pos = [0]

    def take(shape, comps=None):
        cnt = int(np.prod(shape)) * (comps or 1)
        a = out[pos[0]:pos[0] + cnt]
        pos[0] += cnt
        if comps:
            return a.reshape((comps,) + tuple(shape[::-1])).transpose((0, 3, 2, 1))
        return a.reshape(shape, order="F")

    res = dict(nf=[], af=[], iaf=[], ef=[])
    for f in range(3):
        res["nf"].append(take(fshape[f], 3))
        res["af"].append(take(fshape[f]))
        res["iaf"].append(take(fshape[f]))
    for e in range(3):
        res["ef"].append(take(eshape[e]))
    res["ivol"] = take(cs)
    res["cm"] = take(cs, 18)
    res["nc"] = take(cs, 9)
    res["ih"] = take(cs, 3)
    assert pos[0] == n
    return res
