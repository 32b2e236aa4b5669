"""The master's result assembly with its accumulators in device memory (mck_results_*).

Replaces extract_receive_buffer (src/master_queue.f90:253-323) and the conversions of
finalize (src/master_queue.f90:327-581): every task's kernel values, variances and model
integrals are added to K_x, Var and Bg_Model of their inversion-mesh points on the GPU,
and the single-precision output variables are produced there.
"""
import ctypes as C

import numpy as np

from . import abi
from .abi import (MCK_RES_ERROR, MCK_RES_HETERO, MCK_RES_KERNEL,  # noqa: F401 (re-exported)
                  MCK_RES_MODEL, MCK_RES_NITER, MCK_RES_VARIANCE)

ONVERTICES, VOLUMETRIC = 0, 1


def _addr(a):
    """Address of a host numpy array or of a device tensor (None stays NULL)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


class Results:
    """Accumulators of one inversion mesh (int_type ONVERTICES: npoints = vertices and
    connectivity [nelements][nbasis] 1-based; VOLUMETRIC: npoints = nelements)."""

    def __init__(self, kernel, int_type, npoints, nelements, nbasis=1, nmodel=0, connectivity=None, nhet=0):
        from . import _check
        self.k = kernel
        self.int_type, self.npoints, self.nelements = int_type, npoints, nelements
        self.nbasis, self.nmodel, self.nhet, self.nk = nbasis, nmodel, nhet, kernel.nk
        conn = None if connectivity is None else np.ascontiguousarray(connectivity, np.int32)
        self._check = _check
        _check(kernel.lib.mck_results_init(int_type, npoints, nelements, nbasis, nmodel, nhet,
                                           abi.ptr(conn, C.c_int32)))

    def add(self, elems, values=None, variance=None, niterations=None, computation_time=None,
            model=None, rank=1, hetero=None):
        """One task (extract_receive_buffer).  values/variance/niterations/computation_time
        all None: the last slave_work's results, still in device memory.  Host arrays or
        device tensors otherwise, laid out as slave_work returns them."""
        elems = np.ascontiguousarray(elems, np.int32)
        keep = []

        def arr(a, dt):
            if a is None or not isinstance(a, np.ndarray):
                return a
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return a

        args = [arr(values, np.float64), arr(variance, np.float64), arr(niterations, np.int32),
                arr(computation_time, np.float64), arr(model, np.float64), arr(hetero, np.float64)]
        self._check(self.k.lib.mck_results_add(len(elems), abi.ptr(elems, C.c_int32),
                                               *[_addr(a) for a in args], rank))

    def get(self):
        """The FP64 accumulators as the device holds them (point-major)."""
        npt, ne, nk = self.npoints, self.nelements, self.nk
        out = dict(kernel=np.zeros((npt, nk)), variance=np.zeros((npt, nk)),
                   model=np.zeros((npt, self.nmodel)) if self.nmodel else None,
                   hetero=np.zeros((npt, self.nhet)) if self.nhet else None,
                   niterations=np.zeros((ne, nk), np.int32), computation_time=np.zeros(ne),
                   element_proc=np.zeros(ne, np.int32))
        self._check(self.k.lib.mck_results_get(
            abi.ptr(out["kernel"], C.c_double), abi.ptr(out["variance"], C.c_double),
            abi.ptr(out["model"], C.c_double), abi.ptr(out["hetero"], C.c_double),
            abi.ptr(out["niterations"], C.c_int32),
            abi.ptr(out["computation_time"], C.c_double), abi.ptr(out["element_proc"], C.c_int32)))
        return out

    def load(self, kernel=None, variance=None, model=None, niterations=None, computation_time=None,
             element_proc=None, hetero=None):
        """Replaces accumulators (resume from intermediate results, or merge partial sums)."""
        keep = [None if a is None else np.ascontiguousarray(a, dt) for a, dt in
                ((kernel, np.float64), (variance, np.float64), (model, np.float64), (hetero, np.float64),
                 (niterations, np.int32), (computation_time, np.float64), (element_proc, np.int32))]
        self._check(self.k.lib.mck_results_load(*[_addr(a) for a in keep]))

    def export(self, which):
        """One single-precision output variable of finalize, (nentries, npoints)."""
        if which == MCK_RES_NITER:
            shape = (self.nk, self.nelements)
        elif which == MCK_RES_MODEL:
            shape = (self.nmodel, self.npoints)
        elif which == MCK_RES_HETERO:
            shape = (self.nhet, self.npoints)
        else:
            shape = (self.nk, self.npoints)
        out = np.zeros(shape, np.float32)
        self._check(self.k.lib.mck_results_export(which, out.ctypes.data))
        return out
