"""MI355X implementation of MC Kernel's per-point kernel evaluation and Monte Carlo
integration path (the work unit `slave_work` of src/slave.f90).

The compute path is the HIP library libmck.so behind the C-ABI of include/mck.h; this
package is its Python binding.  There is no CPU fallback: every call goes to the GPU and
fails loudly when the library or a GPU is missing.
"""
import ctypes as C
import os

import numpy as np

from . import abi

__all__ = ["MCKernel", "MckError", "load_library", "bind_hip_runtime", "abi"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class MckError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("mck error %d: %s" % (code, msg))
        self.code = code


def library_path():
    """The in-tree libmck.so, or the build that MCK_LIB names (an A/B build of the same
    sources with other compile-time settings)."""
    return os.environ.get("MCK_LIB") or os.path.join(_HERE, "libmck.so")


def bind_hip_runtime():
    """Makes the HIP libraries of this package share PyTorch's HIP runtime.

    PyTorch-ROCm ships its own libamdhip64 / libhsa-runtime64 with the same sonames as
    ROCm's.  A process that holds device tensors of torch and device memory of libmck.so
    must run one runtime: once torch is imported, the dynamic loader resolves libmck.so's
    libamdhip64.so.7 to the copy torch already loaded.  Loading libmck.so first would
    map ROCm's copy beside torch's, and torch's device initialisation then fails.
    Call before loading any HIP library of the package (load_library does)."""
    try:
        import torch  # noqa: F401  (loads torch's HIP runtime)
    except ImportError:
        pass


def load_library():
    """Loads libmck.so (built in-tree by __graft_entry__.build())."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError("libmck.so is not built (run __graft_entry__.build()); "
                               "there is no CPU fallback")
        bind_hip_runtime()
        _LIB = abi.declare(C.CDLL(path))
    return _LIB


def _check(rc):
    if rc != 0:
        raise MckError(rc, load_library().mck_last_error().decode())


class MCKernel:
    """One process-wide instance of the GPU path (mck_init / mck_finalize)."""

    def __init__(self, setup):
        self.lib = load_library()
        self.setup = setup          # the caller keeps the arrays behind it alive
        self.nk = setup.nkernel
        self.nt = setup.nt
        self.R = setup.nrec
        _check(self.lib.mck_init(C.byref(setup)))
        self.halton_step = 0
        full = any(p not in (0, 1, 6, 7) for p in
                   np.ctypeslib.as_array(setup.kernel_param, shape=(setup.nkernel,)))
        self.nd = 6 if full else 1

    def close(self):
        if self.lib is not None:
            self.lib.mck_finalize()
            self.lib = None

    def slave_work(self, verts, conn, halton_offsets=None, plain=False, nbasis=1, sequential=False):
        """mck_slave_work for a task of tetrahedra (conn (ne, 4)) or voxels (conn (ne, 8));
        returns a dict of numpy arrays.  values/variance are (ne, nk) for volumetric
        elements (nbasis 1) and (ne, nbasis, nk) for basis functions on the vertices.
        plain: mck_slave_work_plain; sequential: mck_slave_work_sequential (the reference's
        own stream from self.halton_step, each element's start returned as "offsets")."""
        verts = np.ascontiguousarray(verts, np.float64)
        conn = np.ascontiguousarray(conn, np.int32)
        ne, nvpe = conn.shape
        etype = 4 if nvpe == 4 else 5
        shape = (ne, self.nk) if nbasis == 1 else (ne, nbasis, self.nk)
        vals = np.zeros(shape)
        var = np.zeros(shape)
        nit = np.zeros((ne, self.nk), np.int32)
        tm = np.zeros(ne)
        step = np.array([self.halton_step], np.int64)
        offs = None if halton_offsets is None else np.ascontiguousarray(halton_offsets, np.int64)
        if sequential:
            if offs is not None:
                raise ValueError("the sequential stream takes no halton_offsets")
            offs = np.zeros(ne, np.int64)
            fn = self.lib.mck_slave_work_sequential
        else:
            fn = self.lib.mck_slave_work_plain if plain else self.lib.mck_slave_work
        _check(fn(etype, nvpe, ne, nbasis, verts.shape[0], abi.ptr(verts, C.c_double),
                  abi.ptr(conn, C.c_int32), abi.ptr(vals, C.c_double), abi.ptr(var, C.c_double),
                  abi.ptr(nit, C.c_int32), abi.ptr(tm, C.c_double), abi.ptr(step, C.c_int64),
                  abi.ptr(offs, C.c_int64)))
        self.halton_step = int(step[0])
        out = dict(values=vals, variance=var, niterations=nit, computation_time=tm,
                   halton_step=int(step[0]), stats=self.stats())
        if sequential:
            out["offsets"] = offs
        return out

    def tune(self, xyz, reps=2):
        """mck_tune on the points xyz [n][3]: the fastest backward evaluation, applied to
        the later calls; returns its description (dict of mck_tuning_t)."""
        xyz = np.ascontiguousarray(xyz, np.float64)
        t = abi.MckTuning()
        _check(self.lib.mck_tune(xyz.shape[0], abi.ptr(xyz, C.c_double), reps, C.byref(t)))
        out = {f: getattr(t, f) for f, _ in abi.MckTuning._fields_}
        out["candidate_ms"] = list(t.candidate_ms)
        return out

    def set_tuning(self, tuning):
        t = abi.MckTuning()
        for f in ("bwd_path", "mfma_waves", "mfma_xcd"):
            setattr(t, f, tuning[f])
        _check(self.lib.mck_set_tuning(C.byref(t)))

    def stats(self):
        s = abi.MckStats()
        _check(self.lib.mck_get_stats(C.byref(s)))
        return {f: getattr(s, f) for f, _ in abi.MckStats._fields_}

    def locate(self, frame, xyz):
        xyz = np.ascontiguousarray(xyz, np.float64)
        n = xyz.shape[0]
        elem = np.zeros(n, np.int32)
        xi = np.zeros(n)
        eta = np.zeros(n)
        ties = np.zeros(n, np.int32)
        _check(self.lib.mck_locate_points(frame, n, abi.ptr(xyz, C.c_double),
                                          abi.ptr(elem, C.c_int32), abi.ptr(xi, C.c_double),
                                          abi.ptr(eta, C.c_double), abi.ptr(ties, C.c_int32)))
        return elem, xi, eta, ties

    def load_strain(self, frame, xyz):
        xyz = np.ascontiguousarray(xyz, np.float64)
        out = np.zeros((xyz.shape[0], self.nd, self.nt), np.float32)
        _check(self.lib.mck_load_strain_points(frame, xyz.shape[0], abi.ptr(xyz, C.c_double),
                                               abi.ptr(out, C.c_float)))
        return out

    def integrate_model(self, verts, conn, nbasis=1, halton_offsets=None):
        """mck_integrate_model: (model [ne][nbasis][12], nsteps [ne]); advances halton_step."""
        verts = np.ascontiguousarray(verts, np.float64).reshape(-1, 3)
        conn = np.ascontiguousarray(conn, np.int32)
        ne, nvpe = conn.shape
        model = np.zeros((ne, nbasis, 12))
        nsteps = np.zeros(ne, np.int32)
        step = np.array([self.halton_step], np.int64)
        offs = None if halton_offsets is None else np.ascontiguousarray(halton_offsets, np.int64)
        _check(self.lib.mck_integrate_model(4 if nvpe == 4 else 5, nvpe, ne, nbasis, verts.shape[0],
                                            abi.ptr(verts, C.c_double), abi.ptr(conn, C.c_int32),
                                            abi.ptr(model, C.c_double), abi.ptr(nsteps, C.c_int32),
                                            abi.ptr(step, C.c_int64), abi.ptr(offs, C.c_int64)))
        self.halton_step = int(step[0])
        return model, nsteps

    def set_hetero(self, rtp, dlnhet):
        """mck_set_hetero: the heterogeneity model, rtp (n, 3) = radius km, latitude,
        longitude (degrees), dlnhet (n, 3) = dlnvp, dlnvs, dlnrho."""
        rtp = np.ascontiguousarray(rtp, np.float64).reshape(-1, 3)
        dln = np.ascontiguousarray(dlnhet, np.float32).reshape(-1, 3)
        if rtp.shape[0] != dln.shape[0]:
            raise ValueError("rtp and dlnhet differ in length")
        _check(self.lib.mck_set_hetero(rtp.shape[0], abi.ptr(rtp, C.c_double), abi.ptr(dln, C.c_float)))

    def integrate_hetero(self, verts, conn, nbasis=1, halton_offsets=None):
        """mck_integrate_hetero: (hetero [ne][nbasis][3], nsteps [ne]); advances halton_step."""
        verts = np.ascontiguousarray(verts, np.float64).reshape(-1, 3)
        conn = np.ascontiguousarray(conn, np.int32)
        ne, nvpe = conn.shape
        out = np.zeros((ne, nbasis, 3))
        nsteps = np.zeros(ne, np.int32)
        step = np.array([self.halton_step], np.int64)
        offs = None if halton_offsets is None else np.ascontiguousarray(halton_offsets, np.int64)
        _check(self.lib.mck_integrate_hetero(4 if nvpe == 4 else 5, nvpe, ne, nbasis, verts.shape[0],
                                             abi.ptr(verts, C.c_double), abi.ptr(conn, C.c_int32),
                                             abi.ptr(out, C.c_double), abi.ptr(nsteps, C.c_int32),
                                             abi.ptr(step, C.c_int64), abi.ptr(offs, C.c_int64)))
        self.halton_step = int(step[0])
        return out, nsteps

    def results(self, int_type, npoints, nelements, nbasis=1, nmodel=0, connectivity=None, nhet=0):
        """mck_results_init: the master's accumulators in device memory (results.Results)."""
        from .results import Results
        return Results(self, int_type, npoints, nelements, nbasis, nmodel, connectivity, nhet)

    def seismograms(self):
        """[nrec][nt] seismograms the kernels use (the set-up's, or by reciprocity)."""
        out = np.zeros((self.R, self.nt))
        _check(self.lib.mck_seismograms(abi.ptr(out, C.c_double)))
        return out

    def base_kernels(self, xyz):
        xyz = np.ascontiguousarray(xyz, np.float64)
        out = np.zeros((xyz.shape[0], self.nk, 6))
        _check(self.lib.mck_base_kernels(xyz.shape[0], abi.ptr(xyz, C.c_double),
                                         abi.ptr(out, C.c_double)))
        return out


def load_het_rtpv(path):
    """The heterogeneity file read by load_het_rtpv (heterogeneities.f90:63-110): a first
    line with the number of points, then one line per point with radius (km), latitude
    and longitude (degrees) and dlnvp, dlnvs, dlnrho.  Returns (rtp (n, 3) float64,
    dlnhet (n, 3) float32)."""
    with open(path) as f:
        n = int(f.readline().split()[0])
        rows = np.loadtxt(f, ndmin=2, max_rows=n)
    if rows.shape[0] != n or rows.shape[1] < 6:
        raise ValueError("heterogeneity file %s: expected %d rows of 6 columns" % (path, n))
    if np.any(np.abs(rows[:, 1]) > 90.0):
        raise ValueError("heterogeneity file %s: latitude outside [-90, 90]" % path)
    return rows[:, :3].astype(np.float64), rows[:, 3:6].astype(np.float32)
