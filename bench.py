#!/usr/bin/env python3
"""Benchmark: zone-updates/s of the ideal-MHD step on one MI355X.

Workload (largest single-GPU configuration of the benchmark set): a 256^3
periodic sheared box -- physical x = A x_comp with A = [[1, .5, 0], [0, 1, 0],
[0, 0, 1]] over the unit cube, so the i-faces lean at atan(.5) -- gamma=5/3,
Adams-Bashforth + constrained transport, PDM reconstruction, seeded blast
wave (4 centres, p=100 inside r<0.1, p=0.1 outside, rho=1,
B=(1,1,0)/sqrt(2)).

Two runs from the same initial state:
  * plain: binary64 everywhere (parity-checked against the CPU oracle);
  * fast : fp32 reconstruction/flux/EMF with binary64 conserved
           accumulation and CT face update, plus the learned face
           reconstruction (bf16 MFMA, PDM fallback).  The net is trained
           first, outside the timed region, on face samples of the binary64
           solver at 128^3 (Orszag-Tang, Gaussian blobs and blast set-ups with
           seeds distinct from this workload's), by the library's own GPU
           training kernels.
`value` is the fast run's whole-job zone-updates/s; the plain run's figure,
the speed-up and the density L1 difference between the two after the timed
steps are reported beside it.  A zone-update is one interior cell advanced
one full step.

cpu_baseline: the reference solver (oracle/_ref, built from mhd3d3.f90) on
the host cores with MPI, on a bounded sample of the same workload, timed by
its own step-loop timer; the reference has no sheared metric, so its sample
is the same 256^3 blast wave on its Cartesian grid.  Falls back to the C
oracle when the reference build or MPI is unavailable.
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from mhd3d_mpi_fortran_amd import gpu, ics  # noqa: E402
from mhd3d_mpi_fortran_amd import surrogate as sg  # noqa: E402

HBM_PEAK_GBS = 8000.0
# algorithmic (compulsory) bytes per output point of each kernel class:
# inputs read once + outputs written once (fp64 mode: binary64 inputs and
# outputs; mixed mode: fp32 stencil inputs and fp32 products)
#   emf  : v1, v2, vf (cells) + b1, b2 (faces) read, one edge EMF written
#   flux : 8 primitives read, 8 face products written
ALG_BYTES = {"emf": {"fp64": 5 * 8 + 8, "mixed": 5 * 4 + 4},
             "flux": {"fp64": 8 * 8 + 8 * 8, "mixed": 8 * 4 + 8 * 4}}
# compulsory state traffic of one Adams-Bashforth step per zone: read U^n and
# U^(n-1), write U^(n+1), 8 binary64 words each (rho, rho v, p, three faces)
AB2_STATE_BYTES = 3 * 8 * 8
MLP_FLOPS = 2 * (40 * 32 + 32 * 32 + 32 * 16)


# HBM bytes per launch of the stencil kernels, from the rocprofv3 FETCH_SIZE /
# WRITE_SIZE passes committed under profiles/ (same workload)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "blast256_sheared_traffic.json")
TRAFFIC_KERNELS = {("flux", "mixed", True): ["sur::k_flux_sur<%d, 1>" % d for d in (1, 2, 3)],
                   # the fp32 y and z sweeps march 8 face tiles per workgroup at 256^3
                   ("flux", "mixed", False): ["mhd::k_flux<float, 1, 1, float, float, 1, 1>"] +
                                             ["mhd::k_flux<float, 1, %d, float, float, 1, 8>" % d for d in (2, 3)],
                   ("flux", "fp64", False): ["mhd::k_flux<double, 1, %d, double, double, 1, 1>" % d for d in (1, 2, 3)],
                   # PDM's ei march takes runs of 16 edges, ej and ek runs of 8
                   ("emf", "mixed", True): ["mhd::k_emf_march<float, 1, %d, %d, float, float>" % dt
                                            for dt in ((1, 16), (2, 8), (3, 8))],
                   ("emf", "mixed", False): ["mhd::k_emf_march<float, 1, %d, %d, float, float>" % dt
                                             for dt in ((1, 16), (2, 8), (3, 8))],
                   ("emf", "fp64", False): ["mhd::k_emf_march<double, 1, %d, %d, double, double>" % dt
                                            for dt in ((1, 16), (2, 8), (3, 8))]}


def measured_traffic(kcls, prec, sur, n):
    """mean HBM bytes per launch of the kernel class, or None"""
    if n != 256 or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as fh:
        ks = json.load(fh)["kernels"]
    names = TRAFFIC_KERNELS.get((kcls, prec, bool(sur)))
    if not names or not all(k in ks for k in names):
        return None
    return sum(ks[k]["fetch_bytes"] + ks[k]["write_bytes"] for k in names) / len(names)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


SHEAR = [[1., .5, 0.], [0., 1., 0.], [0., 0., 1.]]


def workload(n, sheared=True):
    g = ics.Grid(n, n, n, ng=4, A=SHEAR if sheared else None)
    st = ics.initial_state(g, "blast", seed=4, radius=.1, p_in=100., p_out=.1)
    return g, st


def run_gpu(g, st, n, precision, weights, warmup, steps, breakdown=True):
    """warm-up, then the timed steps with HIP events around the flux launches
    only (the roofline's kernel: events around every launch of a step cost
    the stream time and distort the kernels between them).  In the timed
    steps a stage's EMF and CT run on a second stream beside its flux and
    update, so the flux kernels share the chip.  Then, for the per-class
    breakdown, the same steps replayed from the same initial state once per
    kernel class, each replay timing that class alone with the stages run
    serially (the kernels' own durations, which add up to a serial step)"""
    s = gpu.Solver(gpu.make_params(n, n, n, ng=4, reconop=1, stepop=1, bc="periodic", precision=precision,
                                   surrogate=1 if weights is not None else 0), g)
    s.upload(st)
    tau0 = s.initialize(1e-4)
    sur = None
    if weights is not None:
        sg.set_weights(s, weights)
        sg.configure(s, sg.THETA, count=True)   # count fallbacks over the warm-up steps only
    tau = tau0
    if warmup:
        tau, _ = s.step(warmup, tau)
    if weights is not None:
        faces, fb = sg.stats(s)
        sur = dict(faces=faces, fallbacks=fb)
        sg.configure(s, sg.THETA, count=False)
    s.synchronize()
    s.profile_classes(["flux"])
    t0 = time.perf_counter()
    tau, tcfl = s.step(steps, tau)
    s.synchronize()
    wall = time.perf_counter() - t0
    flux_timed = s.profile_read("flux")
    s.profile(False)
    flags = s.check_state()
    st_out = s.download()
    prof = {}
    if breakdown:
        s.set_overlap(False)
        for k in gpu.KCLASS:
            s.upload(st)
            tau = s.initialize(1e-4)
            if warmup:
                tau, _ = s.step(warmup, tau)
            s.synchronize()
            before = s.profile_read(k)
            s.profile_classes([k])
            s.step(steps, tau)
            s.synchronize()
            after = s.profile_read(k)
            s.profile(False)
            prof[k] = (after[0] - before[0], after[1] - before[1])
    s.close()
    units = {"emf": [n * (n + 1) * (n + 1), (n + 1) * n * (n + 1), (n + 1) * (n + 1) * n],
             "flux": [(n + 1) * n * n, n * (n + 1) * n, n * n * (n + 1)]}
    return dict(wall=wall, tau=tau, prof=prof, flux_timed=flux_timed, flags=flags, state=st_out, units=units,
                sur=sur)


def l1_rel(a, b, g):
    ra, rb = g.interior(a["rho"]), g.interior(b["rho"])
    return float(np.abs(ra - rb).sum() / np.abs(rb).sum())


def max_divb(st, g):
    """max |div B| = |sum over faces of (B.n) area| / volume"""
    bi, bj, bk = g.interior(st["bi"], 1), g.interior(st["bj"], 2), g.interior(st["bk"], 3)
    ng = g.ng
    dx = g.dx[ng:ng + g.nx0][:, None, None]
    dy = g.dy[ng:ng + g.ny0][None, :, None]
    dz = g.dz[ng:ng + g.nz0][None, None, :]
    fac = [1., 1., 1.]
    if g.A is not None:
        A = g.A
        C = np.array([np.cross(A[:, (d + 1) % 3], A[:, (d + 2) % 3]) for d in range(3)])
        fac = list(np.linalg.norm(C, axis=1) / np.linalg.det(A))
    div = (fac[0] * (bi[1:] - bi[:-1]) / dx + fac[1] * (bj[:, 1:] - bj[:, :-1]) / dy +
           fac[2] * (bk[:, :, 1:] - bk[:, :, :-1]) / dz)
    return float(np.abs(div).max())


FIELDS = ("rho", "vx", "vy", "vz", "p", "bi", "bj", "bk", "bx", "by", "bz")
FACE = {"bi": 1, "bj": 2, "bk": 3}


def field_errors(a, b, g):
    """every output field of state a against state b over the owned cells
    and faces: L1 = sum |a - b| / sum |b| and the largest difference over the
    largest value of the field's group (velocity, magnetic field, or itself)"""
    own = {k: (g.interior(a[k], FACE.get(k, 0)), g.interior(b[k], FACE.get(k, 0))) for k in FIELDS}
    mx = {k: float(np.abs(v[1]).max()) for k, v in own.items()}
    grp = {"v": max(mx[k] for k in ("vx", "vy", "vz")), "b": max(mx[k] for k in ("bi", "bj", "bk", "bx", "by", "bz"))}
    out = {}
    for k, (x, y) in own.items():
        d = np.abs(x - y)
        scale = grp["v"] if k[0] == "v" else (grp["b"] if k[0] == "b" else mx[k])
        out[k] = {"L1": float(d.sum() / max(np.abs(y).sum(), 1e-300)), "max_rel": float(d.max() / max(scale, 1e-300))}
    return out


def accuracy(g, st, n, weights, steps):
    """the headline's accuracy half over the configuration's full step count:
    binary64, fp32 + learned reconstruction and fp32 + PDM from the same
    initial state, each `steps` Adams-Bashforth steps (CFL dt), then every
    output field of the two fp32 runs against binary64, max |div B| and the
    state flags"""
    runs = [("fp64", "fp64", None), ("fp32_pdm", "mixed", None)]
    if weights is not None:
        runs.insert(1, ("learned", "mixed", weights))
    states, rec = {}, {"steps": steps, "state_flags": {}, "max_divB": {}}
    for name, prec, w in runs:
        s = gpu.Solver(gpu.make_params(n, n, n, ng=4, reconop=1, stepop=1, bc="periodic", precision=prec,
                                       surrogate=0 if w is None else 1), g)
        try:
            if w is not None:
                sg.set_weights(s, w)
                sg.configure(s, sg.THETA, count=True)
            s.upload(st)
            tau = s.initialize(1e-4)
            t0 = time.perf_counter()
            s.step(steps, tau)
            s.synchronize()
            log("accuracy run %s: %d steps in %.2f s" % (name, steps, time.perf_counter() - t0))
            rec["state_flags"][name] = s.check_state()
            if w is not None:
                faces, fb = sg.stats(s)
                rec["learned_fallback_fraction"] = fb / max(faces, 1)
            states[name] = s.download()
        finally:
            s.close()
        rec["max_divB"][name] = max_divb(states[name], g)
    rec["fields_vs_fp64"] = {name: field_errors(states[name], states["fp64"], g) for name in states if name != "fp64"}
    rec["density_L1_vs_fp64"] = {name: e["rho"]["L1"] for name, e in rec["fields_vs_fp64"].items()}
    return rec


SHELL_SHAPE = (96, 64, 128)


def shell_config(steps, warmup=2):
    """spherical-shell dipole-plus-inflow set-up: 96 (r, log-stretched 3-30)
    x 64 (theta) x 128 (phi), WENO5, CT, RK2, binary64 against fp32 stencils
    with binary64 accumulation; the general metric built from the node
    coordinates.  Fixed dt = 0.5 x the CFL-0.3 dt of the initial state (the
    inflow compresses the dayside, so the initial dt alone is not stable for
    the whole run).  Returns the record of both runs."""
    from mhd3d_mpi_fortran_amd import shell
    nr, nth, nph = SHELL_SHAPE
    g = shell.spherical_shell(nr, nth, nph, ng=3)
    st = shell.dipole_inflow(g, seed=1)
    geo = g.geometry()
    bmax = max(float(np.abs(st[k]).max()) for k in ("bi", "bj", "bk"))
    hmin = float(np.cbrt(geo["vol"].min()))
    out, states, tau_used = {}, {}, None
    for prec in ("fp64", "mixed"):
        p = gpu.make_params(nr, nth, nph, ng=3, reconop=3, stepop=2, bc=("even", "even", "periodic"),
                            precision=prec)
        s = gpu.Solver(p, g)
        try:
            s.upload(st)
            s.initialize(0.)
            if tau_used is None:
                tau_used = .5 * s.get_dt()
            s.step(warmup, tau_used)
            s.synchronize()
            t0 = time.perf_counter()
            s.step(steps, tau_used)
            s.synchronize()
            wall = time.perf_counter() - t0
            flags = s.check_state()
            states[prec] = s.download()
        finally:
            s.close()
        d = shell.div_b(g, states[prec])[g.ng:-g.ng, g.ng:-g.ng, g.ng:-g.ng]
        out[prec] = {"value": nr * nth * nph * steps / wall, "ms_per_step": wall / steps * 1e3,
                     "state_flags": flags, "max_divB_normalised": float(np.abs(d).max()) * hmin / bmax}
    ra, rb = g.interior(states["mixed"]["rho"]), g.interior(states["fp64"]["rho"])
    return {"workload": "spherical shell %dx%dx%d (r log 3-30, theta .3-pi-.3, phi periodic), dipole + "
                        "inflow v=4 along -x, Bz=-0.1, 1%% density noise; WENO5, CT, RK2, %d steps (+%d warm-up), "
                        "dt %.4e fixed" % (nr, nth, nph, steps, warmup, tau_used),
            "fp64": out["fp64"], "mixed": out["mixed"],
            "speedup_mixed_vs_fp64": out["mixed"]["value"] / out["fp64"]["value"],
            "density_L1_mixed_vs_fp64": float(np.abs(ra - rb).sum() / np.abs(rb).sum()),
            "initial_max_divB_normalised": float(np.abs(shell.div_b(g, st)).max()) * hmin / bmax}


ENSEMBLE_SIZES = (32, 48, 64, 96)


def ensemble_config(steps, nbox=16, warmup=2):
    """batched ensemble: nbox independent periodic boxes with edges drawn
    (SplitMix64, seed 2) from 32, 48, 64, 96; Gaussian blobs (rho = 1 + 3
    seeded blobs, p = 1, uniform |B| = 0.5, v = 0), TVD, RK2, fixed dt =
    0.5 x each box's initial CFL-0.3 dt.  The boxes step together, one HIP
    graph launch per step (mhd_ensemble_*), in fp32 stencils + binary64
    accumulation and in binary64; the mixed ensemble is also timed box by
    box (mhd_step on each box in turn) for comparison."""
    u = ics.uniform01(2, nbox)
    sizes = [ENSEMBLE_SIZES[min(int(x * 4), 3)] for x in u]
    zones = sum(n ** 3 for n in sizes)
    rec = {"workload": "%d periodic boxes, edges %s; Gaussian blobs, TVD, RK2, %d steps (+%d warm-up), dt = 0.5 x "
                       "initial CFL dt per box" % (nbox, sizes, steps, warmup, ), "zones": zones}
    states = {}
    for prec in ("mixed", "fp64"):
        solvers, taus, grids = [], [], []
        for b, n in enumerate(sizes):
            g = ics.Grid(n, n, n, ng=3)
            st = ics.initial_state(g, "blobs", 100 + b)
            s = gpu.Solver(gpu.make_params(n, n, n, ng=3, reconop=2, stepop=2, bc="periodic", precision=prec), g)
            s.upload(st)
            s.initialize(0.)
            taus.append(.5 * s.get_dt())
            solvers.append(s)
            grids.append(g)
        try:
            with gpu.Ensemble(solvers, taus) as ens:
                ens.step(warmup)
                t0 = time.perf_counter()
                ens.step(steps)
                wall = time.perf_counter() - t0
            r = {"ensemble": {"value": zones * steps / wall, "ms_per_step": wall / steps * 1e3}}
            if prec == "mixed":
                # the same boxes, one after the other with their own launches
                for s, tau in zip(solvers, taus):
                    s.step(warmup, tau)
                for s in solvers:
                    s.synchronize()
                t0 = time.perf_counter()
                for s, tau in zip(solvers, taus):
                    s.step(steps, tau)
                wall = time.perf_counter() - t0
                r["box_by_box"] = {"value": zones * steps / wall, "ms_per_step": wall / steps * 1e3}
                r["ensemble_speedup_vs_box_by_box"] = r["ensemble"]["value"] / r["box_by_box"]["value"]
                # the comparison with binary64 is taken at the ensemble's
                # step count: re-run the mixed ensemble from the start
                for s, g, b in zip(solvers, grids, range(nbox)):
                    s.upload(ics.initial_state(g, "blobs", 100 + b))
                    s.initialize(0.)
                with gpu.Ensemble(solvers, taus) as ens:
                    ens.step(warmup + steps)
            states[prec] = [(g, s.download()) for g, s in zip(grids, solvers)]
            rec[prec] = r
        finally:
            for s in solvers:
                s.close()
    num = sum(float(np.abs(g.interior(a["rho"]) - g.interior(b["rho"])).sum())
              for (g, a), (_, b) in zip(states["mixed"], states["fp64"]))
    den = sum(float(np.abs(g.interior(b["rho"])).sum()) for g, b in states["fp64"])
    rec["density_L1_mixed_vs_fp64"] = num / den
    rec["speedup_mixed_vs_fp64"] = rec["mixed"]["ensemble"]["value"] / rec["fp64"]["ensemble"]["value"]
    return rec


def host_cpus():
    """the CPUs this process may run on, their physical cores (hyperthread
    siblings counted once) and the CPU model"""
    try:
        cpus = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cpus = list(range(os.cpu_count() or 1))
    cores = set()
    for c in cpus:
        try:
            base = "/sys/devices/system/cpu/cpu%d/topology/" % c
            with open(base + "physical_package_id") as fh:
                pkg = fh.read().strip()
            with open(base + "core_id") as fh:
                core = fh.read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("?", str(c)))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return len(cpus), len(cores), model


def rank_grid(p, n):
    """xprocs, yprocs, zprocs with every factor >= 2 (the reference leaves a
    direction's right-side ghosts unfilled with one rank, mhd3d3.f90:375-388)
    dividing n, for the largest usable rank count <= p: powers of two, split
    as evenly as possible"""
    best = None
    for a in range(1, 12):
        for b in range(1, 12):
            for c in range(1, 12):
                q = (2 ** a) * (2 ** b) * (2 ** c)
                if q <= p and all(n % (2 ** e) == 0 for e in (a, b, c)):
                    key = (q, -max(a, b, c))
                    if best is None or key > best[0]:
                        best = (key, (2 ** a, 2 ** b, 2 ** c))
    return best[1] if best else None


def cpu_reference(g, st, n, nsteps, px, model):
    """time the reference Fortran (oracle/_ref) with MPI on px[0]*px[1]*px[2]
    ranks; None if unavailable."""
    exe = os.path.join(ROOT, "oracle", "_ref", "mhd3d3_ref_periodic.x")
    mpiexec = "/opt/conda/bin/mpiexec"
    if not (os.path.exists(exe) and os.path.exists(mpiexec)):
        return None
    ranks = px[0] * px[1] * px[2]
    run = tempfile.mkdtemp(prefix="mhdref_bench_")
    try:
        with open(os.path.join(run, "ic.bin"), "wb") as fh:
            for v in ("rho", "vx", "vy", "vz", "p", "bi", "bj", "bk"):
                face = {"bi": 1, "bj": 2, "bk": 3}.get(v, 0)
                fh.write(np.asfortranarray(g.interior(st[v], face)).tobytes(order="F"))
        deck = (" &dom\n xmin=0., xmax=1., ymin=0., ymax=1., zmin=0., zmax=1.\n/\n"
                " &grid\n nx0G=%d, ny0G=%d, nz0G=%d, xprocs=%d, yprocs=%d, zprocs=%d\n/\n"
                " &comset\n nord=8, reconop=1, stepop=1, outop=0\n/\n"
                " &list\n nts=1, tau=1.e-4, nframes=%d, nrst=0, nfdump=10\n/\n"
                " &free\n cfl=.3, pdmb=4., eps=1.e-8, alpha=2., beta=4.\n/\n") % (n, n, n, px[0], px[1], px[2],
                                                                                  nsteps)
        with open(os.path.join(run, "mhd3d3.in"), "w") as fh:
            fh.write(deck)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run([mpiexec, "-n", str(ranks), exe], cwd=run, capture_output=True, text=True,
                           timeout=900, env=env)
        if r.returncode != 0:
            log("reference run failed:", r.stdout[-500:], r.stderr[-500:])
            return None
        for line in r.stdout.splitlines():
            if "STEPLOOP_SECONDS" in line:
                t = float(line.split()[-1])
                return dict(value=n ** 3 * nsteps / t, unit="zone-updates/s", cores=ranks, kind="reference",
                            cpu_model=model, step_loop_s=t,
                            sample="%d^3 blast wave on the reference's Cartesian grid, PDM AB2, %d steps, mpiexec -n %d "
                                   "(%dx%dx%d, one rank per physical core), step loop timed by MPI_Wtime; flang -O3 "
                                   "-fdefault-real-8, no fast-math" % (n, nsteps, ranks, px[0], px[1], px[2]))
        return None
    except (subprocess.TimeoutExpired, OSError) as e:
        log("reference run failed:", e)
        return None
    finally:
        shutil.rmtree(run, ignore_errors=True)


def cpu_port(g, st, n, nsteps, threads, model):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    po.lib().orc_set_threads(threads)
    o = po.Oracle(po.make_params(n, n, n, ng=4, reconop=1, stepop=1, bc="periodic"), g)
    o.set_state(st)
    tau = o.initialize(1e-4)
    t0 = time.perf_counter()
    o.step(nsteps, tau)
    t = time.perf_counter() - t0
    return dict(value=n ** 3 * nsteps / t, unit="zone-updates/s", cores=threads, kind="port", cpu_model=model,
                sample="%d^3 blast wave on the Cartesian grid, PDM AB2, %d steps, C oracle with %d OpenMP threads" % (n, nsteps, threads))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--accuracy-steps", type=int, default=500,
                    help="steps of the accuracy comparison (binary64 vs learned vs fp32 + PDM), apart from the "
                         "timed steps; 0: skip it")
    ap.add_argument("--only", choices=["plain", "fast", "both"], default="both")
    ap.add_argument("--no-surrogate", action="store_true", help="fast run without the learned reconstruction")
    ap.add_argument("--weights", help="float32 file of trained weights (skips training)")
    ap.add_argument("--train-n", type=int, default=128)
    ap.add_argument("--train-epochs", type=int, default=2)
    ap.add_argument("--theta", type=float, default=None, help="smoothness threshold of the learned reconstruction")
    ap.add_argument("--shell-steps", type=int, default=200,
                    help="steps of the spherical-shell side record (0: skip it)")
    ap.add_argument("--shell-only", action="store_true", help="print only the spherical-shell record")
    ap.add_argument("--ensemble-steps", type=int, default=100,
                    help="steps of the batched-ensemble side record (0: skip it)")
    ap.add_argument("--ensemble-only", action="store_true", help="print only the batched-ensemble record")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the MPI decomposition of the reference is replaced by a single "
              "whole-domain block; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.shell_only:
        print(json.dumps(shell_config(args.shell_steps)), flush=True)
        return
    if args.ensemble_only:
        print(json.dumps(ensemble_config(args.ensemble_steps)), flush=True)
        return
    n = args.n
    if args.theta is not None:
        sg.THETA = args.theta
    g, st = workload(n)
    weights, train_info = None, None
    if args.only in ("fast", "both") and not args.no_surrogate:
        if args.weights:
            weights = np.fromfile(args.weights, dtype=np.float32)
            train_info = {"weights": os.path.basename(args.weights)}
        else:
            log("training the learned reconstruction at %d^3 ..." % args.train_n)
            t0 = time.perf_counter()
            weights, loss, nsamp = sg.train_surrogate(n=args.train_n, capacity=1 << 24, epochs=args.train_epochs,
                                                      log=log)
            train_info = {"samples": nsamp, "epochs": args.train_epochs, "batch": 1024,
                          "final_mean_loss": loss, "collect_and_train_s": time.perf_counter() - t0,
                          "sample_grid": "%d^3" % args.train_n}
            log("trained on %d samples in %.1f s, loss %.3e" % (nsamp, train_info["collect_and_train_s"], loss))
    sur = weights is not None
    res = {}
    if args.only in ("plain", "both"):
        log("plain fp64 run ...")
        res["plain"] = run_gpu(g, st, n, "fp64", None, args.warmup, args.steps, breakdown=args.only == "plain")
    if args.only in ("fast", "both"):
        log("fast run (mixed precision%s) ..." % (" + learned reconstruction" if sur else ""))
        res["fast"] = run_gpu(g, st, n, "mixed", weights, args.warmup, args.steps)
        if sur and args.only == "both":
            log("mixed precision with PDM (no learned reconstruction) ...")
            res["mixed_pdm"] = run_gpu(g, st, n, "mixed", None, args.warmup, args.steps)
    head = res.get("fast") or res["plain"]
    head_name = "fast" if "fast" in res else "plain"
    prec = "mixed" if head_name == "fast" else "fp64"
    K = args.steps
    value = n ** 3 * K / head["wall"]

    # dominant kernel of the headline run (the face flux), timed by HIP
    # events on the library's stream over the timed region
    kcls = "flux"
    launches, ms = head["flux_timed"]
    avg_s = ms / launches / 1e3
    units = sum(head["units"][kcls]) / 3.
    alg = ALG_BYTES[kcls][prec] * units
    achieved = alg / avg_s / 1e9

    out = {
        "metric": "zone-updates/s on 1 MI355X: learned+FP32 vs plain FP64 at equal L1 error",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": head["wall"] / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 stencils + f64 accumulate" if prec == "mixed" else "f64",
        "data": "synthetic (seeded blast wave)",
        "config": {"workload": "%d^3 periodic sheared-box blast wave, AB2 + CT, PDM, gamma=5/3" % n,
                   "mode": "mixed precision" + (" + learned reconstruction" if (sur and prec == "mixed") else ""),
                   "grid": "sheared Cartesian, x = A x_comp, A = [[1,.5,0],[0,1,0],[0,0,1]]",
                   "parallelism": "single GPU, whole-domain block"},
        "roofline": {"bound": "hbm", "kernel": kcls, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(kcls, prec, sur, n),
                     "avg_launch_ms": avg_s * 1e3, "alg_bytes_per_launch": alg,
                     "note": "timed region: the flux shares the chip with the EMF and CT on the second stream"},
    }
    if head["prof"].get(kcls, (0, 0))[0]:
        # the same kernels alone (serial replay of the same steps)
        li, msi = head["prof"][kcls]
        ach_i = alg / (msi / li / 1e3) / 1e9
        out["roofline"]["isolated"] = {"avg_launch_ms": msi / li, "achieved": ach_i, "frac": ach_i / HBM_PEAK_GBS}
    if "plain" in res and "fast" in res:
        pv = n ** 3 * K / res["plain"]["wall"]
        out["plain_fp64"] = {"value": pv, "ms_per_step": res["plain"]["wall"] / K * 1e3,
                             "max_divB": max_divb(res["plain"]["state"], g)}
        out["speedup_vs_plain_fp64"] = value / pv
        out["density_L1_after_timed_steps"] = l1_rel(res["fast"]["state"], res["plain"]["state"], g)
        out["max_divB"] = max_divb(res["fast"]["state"], g)
        if args.accuracy_steps > 0:
            log("accuracy comparison over %d steps ..." % args.accuracy_steps)
            acc = accuracy(g, st, n, weights, args.accuracy_steps)
            out["accuracy"] = acc
            key = "learned" if sur else "fp32_pdm"
            out["density_L1_vs_plain_fp64"] = acc["density_L1_vs_fp64"][key]
            out["density_L1_steps"] = args.accuracy_steps
            out["max_divB"] = acc["max_divB"][key]
    if "mixed_pdm" in res and "plain" in res:
        r = res["mixed_pdm"]
        out["mixed_fp32_pdm"] = {"value": n ** 3 * K / r["wall"], "ms_per_step": r["wall"] / K * 1e3,
                                 "speedup_vs_plain_fp64": res["plain"]["wall"] / r["wall"],
                                 "density_L1_after_timed_steps": l1_rel(r["state"], res["plain"]["state"], g),
                                 "max_divB": max_divb(r["state"], g),
                                 "kernel_ms_per_step": {k: v[1] / K for k, v in r["prof"].items() if v[0]}}
    out["kernel_ms_per_step"] = {k: v[1] / K for k, v in head["prof"].items() if v[0]}
    out["step_roofline"] = {"alg_bytes_per_zone_update": AB2_STATE_BYTES,
                            "achieved_GBs": value * AB2_STATE_BYTES / 1e9,
                            "frac": value * AB2_STATE_BYTES / 1e9 / HBM_PEAK_GBS}
    if sur and head_name == "fast":
        s = head["sur"]
        # the MLP runs only on the listed faces that use the net (k_flux_mlp,
        # in the "surrogate" kernel class with the fallback fix-up)
        out["surrogate"] = {"theta": sg.THETA, "fallback_fraction": s["fallbacks"] / max(s["faces"], 1),
                            "faces_counted": s["faces"], "training": train_info,
                            "mlp_flops_per_net_face": MLP_FLOPS}
    if args.shell_steps > 0:
        log("spherical-shell record (%d steps) ..." % args.shell_steps)
        out["shell_config"] = shell_config(args.shell_steps)
    if args.ensemble_steps > 0:
        log("batched-ensemble record (%d steps) ..." % args.ensemble_steps)
        out["ensemble_config"] = ensemble_config(args.ensemble_steps)
    if not args.no_cpu_baseline:
        ncpu, ncore, model = host_cpus()
        px = rank_grid(ncore, n)
        log("cpu baseline: reference on %s MPI ranks (%d CPUs, %d physical cores: %s) ..." % (px, ncpu, ncore, model))
        gc, stc = workload(n, sheared=False)   # the reference's own (Cartesian) geometry
        cb = cpu_reference(gc, stc, n, args.cpu_steps, px, model) if px else None
        if cb is None:
            m = 64
            g2, st2 = workload(m, sheared=False)
            cb = cpu_port(g2, st2, m, 10, ncore, model)
        cb["host_cpus"] = ncpu
        cb["host_physical_cores"] = ncore
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
