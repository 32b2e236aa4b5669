"""Merged database files (include/mck_dbio.h, libmck_dbio.so): opening the file the
reference reads as <dir>/merged_instaseis_db.nc4 (readfields.f90:291-467, 499-658,
2925-3078) into an mck_db_t for mck_setup_t, and writing generated databases in that
layout (repack_database.py:151-187)."""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class MckDbFile(C.Structure):
    _fields_ = [
        ("db", abi.MckDb), ("nt", C.c_int32), ("dt", C.c_double),
        ("planet_radius", C.c_double), ("source_shift", C.c_double),
        ("source_depth_km", C.c_double), ("file_version", C.c_int32),
        ("stf_dump", C.POINTER(C.c_double)), ("stf_d_dump", C.POINTER(C.c_double)),
        ("sem_mesh", C.POINTER(C.c_int32)),
        ("mesh_vp", C.POINTER(C.c_float)), ("mesh_vs", C.POINTER(C.c_float)),
        ("mesh_rho", C.POINTER(C.c_float)), ("mesh_phi", C.POINTER(C.c_float)),
        ("mesh_xi", C.POINTER(C.c_float)), ("mesh_eta", C.POINTER(C.c_float)),
        ("mapped", C.c_int32), ("priv", C.c_void_p),
    ]


EXPORTS = ["mck_db_open", "mck_db_close", "mck_db_write", "mck_dbio_last_error"]


def library_path():
    # MCK_DBIO_LIB: another build of the same sources (the sanitizer build of the tests)
    return os.environ.get("MCK_DBIO_LIB") or os.path.join(HERE, "libmck_dbio.so")


def load_library():
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError("libmck_dbio.so is not built (run __graft_entry__.build(); needs HDF5)")
        lib = C.CDLL(path)
        lib.mck_db_open.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.POINTER(MckDbFile)]
        lib.mck_db_open.restype = C.c_int
        lib.mck_db_close.argtypes = [C.POINTER(MckDbFile)]
        lib.mck_db_close.restype = None
        lib.mck_db_write.argtypes = [C.c_char_p, C.POINTER(MckDbFile)]
        lib.mck_db_write.restype = C.c_int
        lib.mck_dbio_last_error.restype = C.c_char_p
        _LIB = lib
    return _LIB


class DbIoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("mck_dbio error %d: %s" % (code, msg))
        self.code = code


def _check(rc):
    if rc != 0:
        raise DbIoError(rc, load_library().mck_dbio_last_error().decode())


class DbFile:
    """An opened merged database.  .db is the mck_db_t to place in mck_setup_t.fwd/.bwd
    (valid until close()); the scalars and arrays mirror mck_dbfile_t."""

    def __init__(self, path, nvar_expected=0, map_snapshots=True):
        self.lib = load_library()
        self.f = MckDbFile()
        _check(self.lib.mck_db_open(os.fsencode(path), nvar_expected, 1 if map_snapshots else 0,
                                    C.byref(self.f)))

    @property
    def db(self):
        return self.f.db

    def __getattr__(self, name):
        if name in ("nt", "dt", "planet_radius", "source_shift", "source_depth_km", "file_version",
                    "mapped"):
            return getattr(self.f, name)
        raise AttributeError(name)

    def array(self, name):
        """numpy view of one array (valid until close())."""
        f, d = self.f, self.f.db
        ne, npnt, nt = d.nelem, d.npoints, f.nt
        shapes = {"mesh_s": (d.mesh_s, (npnt,)), "mesh_z": (d.mesh_z, (npnt,)),
                  "mp_s": (d.mp_s, (ne,)), "mp_z": (d.mp_z, (ne,)),
                  "fem_mesh": (d.fem_mesh, (ne, 4)), "eltype": (d.eltype, (ne,)),
                  "axis": (d.axis, (ne,)), "snapshots": (d.snapshots, (ne, d.nvar, 5, 5, nt)),
                  "stf_dump": (f.stf_dump, (nt,)), "stf_d_dump": (f.stf_d_dump, (nt,)),
                  "sem_mesh": (f.sem_mesh, (ne, 5, 5))}
        for m in ("vp", "vs", "rho", "phi", "xi", "eta"):
            shapes["mesh_" + m] = (getattr(f, "mesh_" + m), (npnt,))
        ptr, shape = shapes[name]
        if not ptr:
            return None
        return np.ctypeslib.as_array(ptr, shape=shape)

    def close(self):
        if self.lib is not None:
            self.lib.mck_db_close(C.byref(self.f))
            self.lib = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def write_problem_db(path, problem, which):
    """Writes the fwd (which = 'fwd') or bwd database of a synth.Problem as a merged
    database file (snapshots, mesh, STF of the set-up, time shift and radius)."""
    s = problem.setup
    f = MckDbFile()
    f.db = s.fwd if which == "fwd" else s.bwd
    f.nt, f.dt = problem.nt, problem.dt
    f.planet_radius = s.planet_radius
    f.source_shift = s.timeshift_fwd if which == "fwd" else s.timeshift_bwd
    f.source_depth_km = problem.src_depth
    f.file_version = 3
    stf = np.ascontiguousarray(problem.stf, np.float64)
    stf_d = np.ascontiguousarray(np.gradient(stf, problem.dt), np.float64)
    sem = np.ascontiguousarray(problem.mesh.sem, np.int32)
    f.stf_dump = abi.ptr(stf, C.c_double)
    f.stf_d_dump = abi.ptr(stf_d, C.c_double)
    f.sem_mesh = abi.ptr(sem, C.c_int32)
    _check(load_library().mck_db_write(os.fsencode(path), C.byref(f)))
