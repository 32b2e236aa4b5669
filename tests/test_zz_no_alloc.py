"""mck_results_add_noalloc (include/mck.h): the result assembly of mck_results_add without
any allocation, its work arrays in the library's fixed device array.  Through the C-ABI it
must give the same accumulators as mck_results_add, bit for bit, on two calls in a row
(the second with fewer elements, so stale contents of the fixed array would show), with
device arrays and with the device-resident results of the last slave_work; and reject
host arrays and tasks beyond its limit."""
import ctypes as C

import numpy as np
import pytest

from mc_kernel_amd import abi, synth

pytestmark = pytest.mark.gpu

MAX_PAIRS = 131072


def _accumulate(k, res_cls, p, calls, noalloc):
    """Fresh accumulators, then every (elems, arrays) call; returns the accumulators."""
    res = res_cls(k, 1, p.ntet, p.ntet, 1, 0)
    fn = k.lib.mck_results_add_noalloc if noalloc else k.lib.mck_results_add
    for elems, arrays in calls:
        elems = np.ascontiguousarray(elems, np.int32)
        ptrs = [None if a is None else (a.ctypes.data if isinstance(a, np.ndarray) else a.data_ptr())
                for a in arrays]
        rc = fn(len(elems), abi.ptr(elems, C.c_int32), *ptrs, None, None, 1)
        assert rc == 0, k.lib.mck_last_error().decode()
    return res.get()


def test_noalloc_equals_results_add():
    import torch
    import mc_kernel_amd as mk
    from mc_kernel_amd.results import Results
    p = synth.make_problem("tiny", ntet=12)
    k = mk.MCKernel(p.setup)
    try:
        a = k.slave_work(p.verts, p.conn)
        small = np.arange(3, 8)
        k.halton_step = 0
        b = k.slave_work(p.verts, p.conn[small])
        elems_a = np.arange(1, p.ntet + 1, dtype=np.int32)
        elems_b = (small + 1).astype(np.int32)

        def host(r):
            return [np.ascontiguousarray(r["values"]), np.ascontiguousarray(r["variance"]),
                    np.ascontiguousarray(r["niterations"]), np.ascontiguousarray(r["computation_time"])]

        def dev(r):
            return [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in host(r)]

        da, db = dev(a), dev(b)
        torch.cuda.synchronize()
        ref = _accumulate(k, Results, p, [(elems_a, host(a)), (elems_b, host(b))], False)
        got = _accumulate(k, Results, p, [(elems_a, da), (elems_b, db)], True)
        assert np.abs(ref["kernel"]).max() > 0
        for key in ("kernel", "variance", "niterations", "computation_time", "element_proc"):
            np.testing.assert_array_equal(got[key], ref[key], err_msg=key)

        # the device-resident results of the last slave_work (NULL arrays), twice
        def resident(noalloc):
            res = Results(k, 1, p.ntet, p.ntet, 1, 0)
            fn = k.lib.mck_results_add_noalloc if noalloc else k.lib.mck_results_add
            for conn, elems in ((p.conn, elems_a), (p.conn[small], elems_b)):
                k.halton_step = 0
                k.slave_work(p.verts, conn)
                assert fn(len(elems), abi.ptr(elems, C.c_int32), None, None, None, None, None, None, 1) == 0, \
                    k.lib.mck_last_error().decode()
            return res.get()
        r1, r2 = resident(False), resident(True)
        for key in ("kernel", "variance", "niterations", "element_proc"):
            np.testing.assert_array_equal(r2[key], r1[key], err_msg=key)

        # host arrays and oversized tasks are rejected before any device work
        Results(k, 1, p.ntet, p.ntet, 1, 0)
        arrs = host(a)
        rc = k.lib.mck_results_add_noalloc(len(elems_a), abi.ptr(elems_a, C.c_int32),
                                           *[x.ctypes.data for x in arrs], None, None, 1)
        assert rc == abi.MCK_EINVAL and b"device memory" in k.lib.mck_last_error()
        big = np.full(MAX_PAIRS + 1, -1, np.int32)
        Results(k, 1, MAX_PAIRS + 1, MAX_PAIRS + 1, 1, 0)
        rc = k.lib.mck_results_add_noalloc(len(big), abi.ptr(big, C.c_int32), *[t.data_ptr() for t in da],
                                           None, None, 1)
        assert rc == abi.MCK_EINVAL and b"pairs per task" in k.lib.mck_last_error()
    finally:
        k.close()
