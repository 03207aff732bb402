"""GPU test of sc_get_neighbor_lists_noalloc, the list export that allocates
nothing (its rows are assembled in the library's static device scratch):
called through the C-ABI, it gives exactly the lists of
sc_get_neighbor_lists, on two calls in a row, the second on a smaller
system; sizes beyond the scratch are refused."""
import ctypes

import numpy as np
import pytest

from simulationcode_amd import _native as nat
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _lists(ctx, fn, M):
    N = ctx.N
    out = [np.zeros(N, np.int64), np.zeros((N, M), np.int64), np.zeros(N, np.int64), np.zeros((N, M), np.int64)]
    rc = getattr(ctx._lib, fn)(ctx._h, M, *(a.ctypes.data for a in out))
    return rc, out


@pytest.mark.parametrize("precision", [nat.PREC_FP64, nat.PREC_MIXED])
def test_noalloc_lists_equal_allocating_lists(precision):
    for kind, N, dims, seed, steps in (("berthier", 6000, 3, 81, 40), ("lj", 900, 2, 82, 15)):
        op, r, D, v, rc = H.make_system(kind, N, dims, seed=seed)
        ctx = nat.Context(H.sc_params_from(op, rc, precision=precision, cell_divisor=2), D)
        ctx.set_state(r=r, v=v)
        ctx.update_neighbor_lists()
        ctx.do_time_steps(steps)  # lists from a rebuild at other positions than the current ones
        M = int(op.maxneighbours)
        rc1, a = _lists(ctx, "sc_get_neighbor_lists", M)
        rc2, b = _lists(ctx, "sc_get_neighbor_lists_noalloc", M)
        assert rc1 == 0 and rc2 == 0, ctx._lib.sc_last_error(ctx._h)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        assert a[0].sum() > 0 and a[2].sum() > 0
        # rows past their count are zero
        k = np.arange(M)[None, :]
        assert np.all(b[1][k >= b[0][:, None]] == 0) and np.all(b[3][k >= b[2][:, None]] == 0)
        # only the half list, then only counts
        h = [np.zeros(N, np.int64), np.zeros((N, M), np.int64)]
        assert ctx._lib.sc_get_neighbor_lists_noalloc(ctx._h, M, None, None, h[0].ctypes.data, h[1].ctypes.data) == 0
        assert np.array_equal(h[0], a[2]) and np.array_equal(h[1], a[3])
        c = np.zeros(N, np.int64)
        assert ctx._lib.sc_get_neighbor_lists_noalloc(ctx._h, M, c.ctypes.data, None, None, None) == 0
        assert np.array_equal(c, a[0])
        ctx.close()


def test_noalloc_refuses_sizes_beyond_the_scratch_and_reports_overflow():
    op, r, D, v, rc = H.make_system("berthier", 30000, 3, seed=83)
    ctx = nat.Context(H.sc_params_from(op, rc, cell_divisor=2), D)
    ctx.set_state(r=r, v=v)
    ctx.update_neighbor_lists()
    N, M = ctx.N, 200  # 2 lists x 30000 x 201 words > 8388576
    buf = [np.zeros(N, np.int64), np.zeros((N, M), np.int64), np.zeros(N, np.int64), np.zeros((N, M), np.int64)]
    assert ctx._lib.sc_get_neighbor_lists_noalloc(ctx._h, M, *(x.ctypes.data for x in buf)) == nat.SC_ERR_ARG
    assert b"scratch" in ctx._lib.sc_last_error(ctx._h)
    # one list fits
    assert ctx._lib.sc_get_neighbor_lists_noalloc(ctx._h, M, buf[0].ctypes.data, buf[1].ctypes.data, None, None) == 0
    # rows longer than maxneigh: overflow, as sc_get_neighbor_lists reports it
    small = int(buf[0].max()) - 1
    b2 = [np.zeros(N, np.int64), np.zeros((N, small), np.int64)]
    assert ctx._lib.sc_get_neighbor_lists_noalloc(ctx._h, small, b2[0].ctypes.data, b2[1].ctypes.data,
                                                  None, None) == nat.SC_ERR_OVERFLOW
    ctx.close()
