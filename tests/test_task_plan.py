"""Host logic of the gradient engine's plane-task schedule (phx_plane_task_plan, a host
function of the C library; no GPU call): the forward and adjoint sweeps of P = bmax *
n_modes planes in launches of at most one workgroup per CU, an adjoint plane only after
the launch that completes the forward sweeps of every mode of its position."""
import ctypes
import math
import os

import pytest


def _plan(planes, modes, cus, max_launches=8):
    from phaser_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    buf = (ctypes.c_int * (4 * max_launches))()
    n = lib.phx_plane_task_plan(planes, modes, cus, buf, max_launches)
    return n, [tuple(buf[4 * k:4 * k + 4]) for k in range(max(n, 0))]


def test_known_schedule_384_planes_on_256_cus():
    n, plan = _plan(64 * 6, 6, 256)
    assert n == 3
    assert plan == [(0, 256, 0, 0), (256, 128, 0, 128), (384, 0, 128, 256)]


@pytest.mark.parametrize('positions', [8, 40, 43, 48, 64, 100, 128])
@pytest.mark.parametrize('modes', [1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize('cus', [80, 240, 256, 304])
def test_schedule_invariants(positions, modes, cus):
    P = positions * modes
    n, plan = _plan(P, modes, cus)
    if P <= cus or P > 2 * cus:
        assert n == 0
        return
    assert n >= 0
    if n == 0:
        return   # no launch saved: the two sweeps run as separate launches
    assert n < 2 * math.ceil(P / cus)
    fwd_done = set()
    fwd_seen, adj_seen = [], []
    for f0, nf, a0, na in plan:
        assert 0 < nf + na <= cus and nf >= 0 and na >= 0
        for q in range(a0, a0 + na):
            b = q // modes
            # every mode of position b finished its forward sweep in an earlier launch
            assert all(b * modes + m in fwd_done for m in range(modes))
        fwd_seen += range(f0, f0 + nf)
        adj_seen += range(a0, a0 + na)
        fwd_done |= set(range(f0, f0 + nf))
    assert fwd_seen == list(range(P)) and adj_seen == list(range(P))


def test_rejects_planes_not_a_multiple_of_modes():
    n, _ = _plan(385, 6, 256)
    assert n == -1000   # PHX_EINVAL
