"""The C-ABI library loads and exports every entry point include/mhd_gpu.h
declares; the Python and Fortran bindings agree with the C struct layout.
No compute calls: these run without a GPU."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from mhd3d_mpi_fortran_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mhd_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mhd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("mhd_create", "mhd_step", "mhd_get_dt", "mhd_flux_stress", "mhd_eijk_lfm",
                 "mhd_fill_ghosts", "mhd_reconstruct", "mhd_initialize"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = C.CDLL(gpu.LIBPATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_ctypes_signatures_cover_header():
    assert set(declared_functions()) == set(gpu.SIGNATURES)


def test_params_layout_matches_c():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "mhd_gpu.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(mhd_params), offsetof(mhd_params, gam),
         offsetof(mhd_params, reconop), offsetof(mhd_params, bc), offsetof(mhd_params, precision),
         offsetof(mhd_params, surrogate), offsetof(mhd_params, device));
  return 0; }
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.check_call(["gcc", "-I" + os.path.join(ROOT, "include"), c, "-o", exe])
        vals = [int(v) for v in subprocess.check_output([exe]).split()]
    P = gpu.Params
    assert vals == [C.sizeof(P), P.gam.offset, P.reconop.offset, P.bc.offset, P.precision.offset,
                    P.surrogate.offset, P.device.offset]


def test_default_params_are_the_reference_deck():
    p = gpu.Params()
    gpu.lib().mhd_default_params(C.byref(p))
    assert (p.cfl, p.pdmb, p.eps, p.alpha, p.beta) == (.3, 4., 1e-8, 2., 4.)
    assert p.gam == 5. / 3. and p.ng == 4


@pytest.mark.parametrize("bad", [dict(reconop=5), dict(stepop=0), dict(ng=3, reconop=1), dict(nx0=2),
                                 dict(bc=(0, 1, 0)), dict(precision=7)])
def test_create_rejects_bad_parameters(bad):
    kw = dict(nx0=8, ny0=8, nz0=8, ng=4, reconop=1, stepop=2)
    bc = bad.pop("bc", "periodic")
    kw.update(bad)
    p = gpu.make_params(**kw, bc=bc)
    h = C.c_void_p()
    rc = gpu.lib().mhd_create(C.byref(h), C.byref(p))
    assert rc == gpu.MHD_ERR_PARAM and not h.value


def test_error_strings():
    assert gpu.lib().mhd_error_string(0) == b"ok"
    assert b"parameter" in gpu.lib().mhd_error_string(2)
