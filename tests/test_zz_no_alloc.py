"""ntt_sr_step_static: the SRPIC step without a workspace argument (the
fixed-point current accumulator lives in a static device array of the
library).  Called through the C ABI, it gives exactly the results of
ntt_sr_step on an identical box -- fields bit for bit, particles as a
multiset (a re-sort may order a tile differently) -- on two calls in a row,
the second on a smaller box (the array keeps the first call's contents,
which the step clears).  A box beyond the array's
size is rejected with NTT_ERR_CAPACITY, and a warm call changes the device's
free memory by nothing."""
import ctypes as C

import numpy as np
import pytest
import torch

from test_gpu_engine import make_engine, snapshot
from test_gpu_ensemble import rows

pytestmark = pytest.mark.gpu

BORIS, VAY = 2, 4


def _static_step(eng):
    """One engine step with ntt_sr_step_static in place of ntt_sr_step."""
    from entity_amd import _abi
    eng.maybe_sort()
    (em, cur, buff), sp, prm, metric = eng._step_args()
    prm.step = eng.step
    prm.em_half_out = None
    st = eng.lib.ntt_sr_step_static(C.byref(em), C.byref(cur), C.byref(buff), sp,
                                    len(eng.species), C.byref(prm), C.byref(metric),
                                    torch.cuda.current_stream().cuda_stream)
    _abi.check(st, "ntt_sr_step_static")
    eng.step += 1
    eng.time += eng.params.dt


def _same(a, b):
    pa, fa = snapshot(a)
    pb, fb = snapshot(b)
    for k in ("em", "cur"):
        np.testing.assert_array_equal(fa[k], fb[k], err_msg=k)
    for ha, hb in zip(pa, pb):       # as multisets: the order inside a tile may differ
        np.testing.assert_array_equal(rows(ha), rows(hb))


CALLS = [  # (dim, n, order, pusher, filters, steps): a larger box, then a smaller one
    (3, (32, 16, 16), 2, VAY, 2, 5),
    (2, (32, 24), 1, BORIS, 1, 3),
]


def test_static_step_equals_step(hip):
    from entity_amd.kernels import device_errors
    for dim, n, order, alg, nf, nsteps in CALLS:
        ref = make_engine(dim, n, order, alg, nf, sort_interval=2)
        got = make_engine(dim, n, order, alg, nf, sort_interval=2)
        for _ in range(nsteps):
            ref.step_forward()
            _static_step(got)
        torch.cuda.synchronize()
        assert device_errors() == 0
        _same(got, ref)


def test_static_step_allocates_nothing_and_rejects_large_boxes(hip):
    from entity_amd import _abi
    eng = make_engine(2, (32, 24), 1, BORIS, 1)
    _static_step(eng)                 # warm: kernels and the array's address looked up
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(3):
        _static_step(eng)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0
    # 1200 x 1200 cells need (1204^2 * 24) B > 32 MiB: refused before any launch
    (em, cur, buff), sp, prm, metric = eng._step_args()
    big = [_abi.NttField.from_buffer_copy(f) for f in (em, cur, buff)]
    for f in big:
        f.n[0], f.n[1] = 1200, 1200
    need = C.c_uint64(0)
    assert eng.lib.ntt_sr_step_workspace_size(C.byref(big[1]), C.byref(need)) == _abi.NTT_OK
    assert need.value > (32 << 20)
    st = eng.lib.ntt_sr_step_static(C.byref(big[0]), C.byref(big[1]), C.byref(big[2]), sp,
                                    len(eng.species), C.byref(prm), C.byref(metric), None)
    assert st == _abi.NTT_ERR_CAPACITY
