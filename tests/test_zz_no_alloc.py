"""mhd_reconstruct_noalloc -- reconstruct_3d staged through the GPU's fixed
scratch array, allocating nothing -- against mhd_reconstruct, which stages
through per-call device buffers (needs a GPU).

Both run the same kernel on the same values: bit-exact, for two calls in a
row on each context (a large input, then a smaller one, so the second call
runs over the first call's leftovers in the shared array), for all four
methods and all six directions; plus the size limit, which is checked on
the host before anything reaches the device.
"""
import numpy as np
import pytest

from mhd3d_mpi_fortran_amd import gpu, ics

pytestmark = pytest.mark.gpu


def _lshape(g, direction):
    n = [g.nx, g.ny, g.nz]
    n0 = [g.nx0, g.ny0, g.nz0]
    if direction <= 3:
        lsh = list(n0)
        lsh[direction - 1] += 1
    else:
        lsh = list(n)
        lsh[direction - 4] = n0[direction - 4] + 1
    return tuple(lsh)


@pytest.mark.parametrize("reconop", [1, 2, 3, 4])
def test_noalloc_equals_reconstruct(reconop):
    ng = 4
    for direction in range(1, 7):
        for call, n0 in enumerate(((40, 36, 30), (10, 9, 7))):
            print("noalloc reconop %d dir %d call %d shape %s" % (reconop, direction, call, n0), flush=True)
            g = ics.Grid(*n0, ng=ng)
            rng = np.random.default_rng(1000 * reconop + 10 * direction + call)
            f = np.asfortranarray(rng.standard_normal((g.nx, g.ny, g.nz)))
            s = gpu.Solver(gpu.make_params(*n0, ng=ng, reconop=reconop), g)
            lsh = _lshape(g, direction)
            a = gpu.reconstruct(s, f, lsh, direction)
            b = gpu.reconstruct(s, f, lsh, direction, noalloc=True)
            s.close()
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (direction, call)


def test_noalloc_size_limit():
    g = ics.Grid(8, 8, 8, ng=4)
    s = gpu.Solver(gpu.make_params(8, 8, 8, ng=4, reconop=1), g)
    # f alone needs more than the 2^22 doubles of the scratch array
    big = np.zeros((172, 172, 142), order="F")
    assert big.size > 1 << 22
    with pytest.raises(gpu.MhdError) as e:
        gpu.reconstruct(s, big, (10, 10, 10), 1, noalloc=True)
    assert e.value.code == gpu.MHD_ERR_PARAM
    # a valid call on the same context still works afterwards
    f = np.asfortranarray(np.random.default_rng(1).standard_normal((g.nx, g.ny, g.nz)))
    a = gpu.reconstruct(s, f, (9, 8, 8), 1)
    b = gpu.reconstruct(s, f, (9, 8, 8), 1, noalloc=True)
    s.close()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
