"""The Fortran driver (ISO_C_BINDING over libmhd_gpu.so) against the
reference's frames: same namelist deck format, same initial states."""
import os
import subprocess

import numpy as np
import pytest

from goldens import STATE, Case
from mhd3d_mpi_fortran_amd import build

pytestmark = pytest.mark.gpu

DECK = """ &dom
 xmin={lo[0]!r}, xmax={hi[0]!r}, ymin={lo[1]!r}, ymax={hi[1]!r}, zmin={lo[2]!r}, zmax={hi[2]!r}
/
 &grid
 nx0G={n[0]}, ny0G={n[1]}, nz0G={n[2]}, xprocs=2, yprocs=2, zprocs=2
/
 &comset
 nord={nord}, reconop={reconop}, stepop={stepop}, outop=2
/
 &list
 nts=1, tau={tau!r}, nframes={nframes}, nrst=0, nfdump=10
/
 &free
 cfl=.3, pdmb=4., eps=1.e-8, alpha=2., beta=4.
/
 &gpu
 bc_x='{bc}', bc_y='{bc}', bc_z='{bc}', outdir='{out}', icfile='{ic}', frames=.true.
/
"""


def read_frame(path, n):
    raw = open(path, "rb").read()
    off = 24 + 16
    out = {}
    shapes = dict(rho=(0, 0, 0), vx=(0, 0, 0), vy=(0, 0, 0), vz=(0, 0, 0), p=(0, 0, 0), bi=(1, 0, 0),
                  bj=(0, 1, 0), bk=(0, 0, 1))
    for v in STATE:
        shp = tuple(a + b for a, b in zip(n, shapes[v]))
        cnt = int(np.prod(shp))
        out[v] = np.frombuffer(raw, "<f8", cnt, off).reshape(shp, order="F")
        off += 8 * cnt
    return out


@pytest.mark.parametrize("name,use_icfile", [("st_pdm_rk2_even", False), ("st_mpweno_rk3_even", False),
                                              ("ot_pdm_ab2_per", True), ("ot_weno_rk2_per", True)])
def test_driver_matches_reference(tmp_path, name, use_icfile):
    assert os.path.exists(build.DRIVER), "Fortran driver not built"
    c = Case(name)
    m = c.meta
    ic = ""
    if use_icfile:
        ic = str(tmp_path / "ic.bin")
        with open(ic, "wb") as fh:
            for v in STATE:
                fh.write(np.asfortranarray(c.z["ic_" + v]).tobytes(order="F"))
    deck = tmp_path / "mhd3d3.in"
    deck.write_text(DECK.format(lo=m["lo"], hi=m["hi"], n=m["grid"], nord=2 * m["ng"], reconop=m["reconop"],
                                stepop=m["stepop"], tau=m["tau"], nframes=max(m["frames"]),
                                bc=m["bc"], out=str(tmp_path), ic=ic))
    r = subprocess.run([build.DRIVER, str(deck)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "STEPLOOP_SECONDS" in r.stdout
    f0 = read_frame(tmp_path / "frame_0000.bin", m["grid"])
    for v in STATE:
        assert np.array_equal(f0[v], c.z["ic_" + v]), v
    for fr in m["frames"]:
        got = read_frame(tmp_path / ("frame_%04d.bin" % fr), m["grid"])
        ref = {v: c.frame(fr, v) for v in STATE}
        vmax = max(np.abs(ref[k]).max() for k in ("vx", "vy", "vz"))
        bmax = max(np.abs(ref[k]).max() for k in ("bi", "bj", "bk"))
        for v in STATE:
            scale = vmax if v in ("vx", "vy", "vz") else bmax if v in ("bi", "bj", "bk") else np.abs(ref[v]).max()
            err = np.abs(got[v] - ref[v]).max() / scale
            assert err <= 1e-12, (fr, v, err)


DECK_GPU = """ &dom
 xmin=0., xmax=1., ymin=0., ymax=1., zmin=0., zmax=1.
/
 &grid
 nx0G={n}, ny0G={n}, nz0G={n}, xprocs=1, yprocs=1, zprocs=1
/
 &comset
 nord=8, reconop=1, stepop=1, outop=2
/
 &list
 nts=4, tau=1.e-4, nframes=2, nrst=0, nfdump=10
/
 &free
 cfl=.3, pdmb=4., eps=1.e-8, alpha=2., beta=4.
/
 &gpu
 bc_x='periodic', bc_y='periodic', bc_z='periodic', outdir='{out}', icfile='{ic}', frames=.true.,
 precision='mixed', surrogate=1, weights='{w}', theta=0.25,
 metric=1., 0.5, 0., 0., 1., 0., 0., 0., 1.
/
"""


def test_driver_metric_and_learned_reconstruction(tmp_path):
    """the driver's &gpu metric= and weights= reach the library: its frames
    equal the Python binding's run of the same sheared blast wave"""
    from mhd3d_mpi_fortran_amd import gpu, ics
    from mhd3d_mpi_fortran_amd import surrogate as sg
    n = 24
    A = [[1., .5, 0.], [0., 1., 0.], [0., 0., 1.]]
    g = ics.Grid(n, n, n, ng=4, A=A)
    st = ics.initial_state(g, "blast", 11, radius=.2)
    rng = np.random.default_rng(3)
    w = (rng.standard_normal(sg.NPARAM) * .05).astype(np.float32)
    sg.save_weights(w, tmp_path / "w.bin")
    assert np.array_equal(sg.load_weights(tmp_path / "w.bin"), w)
    ic = tmp_path / "ic.bin"
    with open(ic, "wb") as fh:
        for v in STATE:
            face = {"bi": 1, "bj": 2, "bk": 3}.get(v, 0)
            fh.write(np.asfortranarray(g.interior(st[v], face)).tobytes(order="F"))
    deck = tmp_path / "mhd3d3.in"
    deck.write_text(DECK_GPU.format(n=n, out=str(tmp_path), ic=str(ic), w=str(tmp_path / "w.bin")))
    r = subprocess.run([build.DRIVER, str(deck)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    s = gpu.Solver(gpu.make_params(n, n, n, reconop=1, stepop=1, bc="periodic", precision="mixed", surrogate=1), g)
    sg.set_weights(s, w)
    sg.configure(s, .25)
    s.upload(st)
    tau = s.initialize(1e-4)
    for fr in (1, 2):
        for _ in range(4):
            tau, _ = s.step(1, tau)
        ref = s.download()
        got = read_frame(tmp_path / ("frame_%04d.bin" % fr), (n, n, n))
        for v in STATE:
            face = {"bi": 1, "bj": 2, "bk": 3}.get(v, 0)
            assert np.array_equal(got[v], g.interior(ref[v], face)), (fr, v)
    s.close()


def read_xyz(path, n, nframes):
    """one dataset of outputxyz: (nx0G, ny0G, nz0G, nframes+1) float64"""
    a = np.fromfile(path, "<f8")
    assert a.size == n[0] * n[1] * n[2] * (nframes + 1), (path, a.size)
    return a.reshape(tuple(n) + (nframes + 1,), order="F")


def read_records(path):
    """Fortran sequential unformatted records: 4-byte length, data, length"""
    raw = open(path, "rb").read()
    out, off = [], 0
    while off < len(raw):
        ln = int(np.frombuffer(raw, "<i4", 1, off)[0])
        assert int(np.frombuffer(raw, "<i4", 1, off + 4 + ln)[0]) == ln
        out.append(raw[off + 4:off + 4 + ln])
        off += 8 + ln
    return out


def test_driver_outputxyz_layout(tmp_path):
    """the reference's outputxyz datasets (rho vx vy vz p bx by bz, shape
    (nx0G, ny0G, nz0G, nframes+1), mhd3d3.f90:2966-3109) and output_simt's
    stime.out records (3112-3127): frame n of each dataset equals the raw
    frame of the same run (cell B from its faces, as get_bxyz forms it), and
    stime.out holds the frame times"""
    c = Case("ot_pdm_ab2_per")
    m = c.meta
    ic = str(tmp_path / "ic.bin")
    with open(ic, "wb") as fh:
        for v in STATE:
            fh.write(np.asfortranarray(c.z["ic_" + v]).tobytes(order="F"))
    nframes = max(m["frames"])
    deck = tmp_path / "mhd3d3.in"
    deck.write_text(DECK.format(lo=m["lo"], hi=m["hi"], n=m["grid"], nord=2 * m["ng"], reconop=m["reconop"],
                                stepop=m["stepop"], tau=m["tau"], nframes=nframes, bc=m["bc"], out=str(tmp_path),
                                ic=ic))
    r = subprocess.run([build.DRIVER, str(deck)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    n = m["grid"]
    data = {v: read_xyz(tmp_path / (v + ".xyz"), n, nframes) for v in ("rho", "vx", "vy", "vz", "p", "bx", "by", "bz")}
    for fr in range(nframes + 1):
        raw = read_frame(tmp_path / ("frame_%04d.bin" % fr), n)
        for v in ("rho", "vx", "vy", "vz", "p"):
            assert np.array_equal(data[v][..., fr], raw[v]), (fr, v)
        assert np.array_equal(data["bx"][..., fr], .5 * (raw["bi"][1:] + raw["bi"][:-1])), fr
        assert np.array_equal(data["by"][..., fr], .5 * (raw["bj"][:, 1:] + raw["bj"][:, :-1])), fr
        assert np.array_equal(data["bz"][..., fr], .5 * (raw["bk"][:, :, 1:] + raw["bk"][:, :, :-1])), fr
    recs = read_records(tmp_path / "stime.out")
    times = [float(np.frombuffer(b, "<f8")[0]) for b in recs]
    assert len(times) == nframes + 1 and times[0] == 0.
    assert all(b > a for a, b in zip(times, times[1:]))
    # the driver prints the same frame times
    printed = [float(line.split("time=")[1]) for line in r.stdout.splitlines() if "output frame number" in line]
    assert np.allclose(times[1:], printed, rtol=1e-6)
