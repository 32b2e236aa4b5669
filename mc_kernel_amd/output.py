"""Output files of a kernel run (include/mck_output.h, in libmck_dbio.so): the XDMF/NetCDF,
CSR and ASCII dumps of finalize (master_queue.f90:327-581, inversion_mesh.f90:1886-2265)
and the master's intermediate results (master_queue.f90:586-705), written from the
single-precision variables that results.Results.export produces on the GPU."""
import ctypes as C
import os

import numpy as np

from . import abi, dbfile
from .results import (MCK_RES_ERROR, MCK_RES_HETERO, MCK_RES_KERNEL, MCK_RES_MODEL,
                      MCK_RES_NITER, MCK_RES_VARIANCE, ONVERTICES)

EXPORTS = ["mck_write_xdmf", "mck_write_csr", "mck_write_intermediate", "mck_read_intermediate"]

# background-model parameter names of the model variable (background_model.f90:34-37)
MODEL_NAMES = ["vp", "vs", "rho", "vph", "vpv", "vsh", "vsv", "eta", "phi", "xi", "lam", "mu"]
# heterogeneity parameters as the output stores them (heterogeneities.f90:19-22)
HETERO_NAMES = ["dlnvp", "dlnvs", "dlnrho"]


class MckOutVar(C.Structure):
    _fields_ = [("name", C.c_char_p), ("on_cells", C.c_int32), ("nentries", C.c_int32),
                ("entry_names", C.POINTER(C.c_char_p)), ("values", C.POINTER(C.c_float))]


def _lib():
    lib = dbfile.load_library()
    if not getattr(lib, "_mck_output_declared", False):
        lib.mck_write_xdmf.argtypes = [C.c_char_p, C.c_char_p, C.c_int32, abi.f64p, C.c_int32, C.c_int32,
                                       abi.i32p, abi.i32p, C.c_int32, C.POINTER(MckOutVar)]
        lib.mck_write_xdmf.restype = C.c_int
        lib.mck_write_csr.argtypes = [C.c_char_p, C.c_int32, C.c_int32, abi.f32p, C.c_double, C.c_int32]
        lib.mck_write_csr.restype = C.c_int
        lib.mck_write_intermediate.argtypes = [C.c_char_p] + [C.c_int32] * 5 + [abi.f32p] * 4 + [abi.i32p] * 2
        lib.mck_write_intermediate.restype = C.c_int
        lib.mck_read_intermediate.argtypes = [C.c_char_p] + [C.c_int32] * 5 + [abi.f32p] * 4 + [abi.i32p] * 2
        lib.mck_read_intermediate.restype = C.c_int
        lib._mck_output_declared = True
    return lib


def write_xdmf(filename, element_type, vertices, connectivity, variables, block_id=None):
    """dump_data_xdmf: variables = [(name, on_cells, values (nentries, n) f32, entry
    names or None)]; writes <filename>.nc and <filename>.xdmf."""
    lib = _lib()
    vertices = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
    connectivity = np.ascontiguousarray(connectivity, np.int32)
    block = None if block_id is None else np.ascontiguousarray(block_id, np.int32)
    keep = []
    arr = (MckOutVar * max(len(variables), 1))()
    for i, (name, on_cells, values, names) in enumerate(variables):
        values = np.ascontiguousarray(values, np.float32)
        keep.append(values)
        enames = None
        if names is not None:
            enames = (C.c_char_p * len(names))(*[n.encode() for n in names])
            keep.append(enames)
        arr[i] = MckOutVar(name.encode(), 1 if on_cells else 0, values.shape[0],
                           C.cast(enames, C.POINTER(C.c_char_p)) if enames is not None else None,
                           abi.ptr(values, C.c_float))
    dbfile._check(lib.mck_write_xdmf(os.fsencode(filename), element_type.encode(), vertices.shape[0],
                                     abi.ptr(vertices, C.c_double), connectivity.shape[0],
                                     connectivity.shape[1], abi.ptr(connectivity, C.c_int32),
                                     abi.ptr(block, C.c_int32), len(variables), arr))


def write_csr(filename, values, threshold, ascii=False):
    """dump_*_data_csr / dump_*_data_ascii of values (nkernel, npoints) f32."""
    values = np.ascontiguousarray(values, np.float32)
    dbfile._check(_lib().mck_write_csr(os.fsencode(filename), values.shape[1], values.shape[0],
                                       abi.ptr(values, C.c_float), float(threshold), 1 if ascii else 0))


def write_intermediate(path, kernel, variance, niterations, element_proc, bg_model=None, het_model=None):
    """create_intermediate / dump_intermediate: kernel, variance (nkernel, npoints) f32,
    niterations (nkernel, nelements) i32, element_proc (nelements,) i32."""
    k = np.ascontiguousarray(kernel, np.float32)
    v = np.ascontiguousarray(variance, np.float32)
    ni = np.ascontiguousarray(niterations, np.int32)
    pr = np.ascontiguousarray(element_proc, np.int32)
    bg = None if bg_model is None else np.ascontiguousarray(bg_model, np.float32)
    het = None if het_model is None else np.ascontiguousarray(het_model, np.float32)
    dbfile._check(_lib().mck_write_intermediate(
        os.fsencode(path), k.shape[1], pr.shape[0], k.shape[0], 0 if bg is None else bg.shape[0],
        0 if het is None else het.shape[0], abi.ptr(k, C.c_float), abi.ptr(v, C.c_float),
        abi.ptr(bg, C.c_float), abi.ptr(het, C.c_float), abi.ptr(ni, C.c_int32), abi.ptr(pr, C.c_int32)))


def read_intermediate(path, npoints, nelements, nkernel, nmodel=0, nhet=0):
    """read_intermediate: the arrays of write_intermediate, and completed = element_proc >= 1
    (master_queue.f90:652-653)."""
    out = dict(kernel=np.zeros((nkernel, npoints), np.float32), variance=np.zeros((nkernel, npoints), np.float32),
               bg_model=np.zeros((nmodel, npoints), np.float32) if nmodel else None,
               het_model=np.zeros((nhet, npoints), np.float32) if nhet else None,
               niterations=np.zeros((nkernel, nelements), np.int32), element_proc=np.zeros(nelements, np.int32))
    dbfile._check(_lib().mck_read_intermediate(
        os.fsencode(path), npoints, nelements, nkernel, nmodel, nhet, abi.ptr(out["kernel"], C.c_float),
        abi.ptr(out["variance"], C.c_float), abi.ptr(out["bg_model"], C.c_float),
        abi.ptr(out["het_model"], C.c_float), abi.ptr(out["niterations"], C.c_int32),
        abi.ptr(out["element_proc"], C.c_int32)))
    out["completed"] = out["element_proc"] >= 1
    return out


def finalize(results, output_file, kernel_names, element_type, vertices, connectivity,
             dump_type="xdmf", allowed_error=0.0, volumes=None, block_id=None):
    """finalize (master_queue.f90:327-581) from device-resident accumulators.  dump_type
    'xdmf': the kernel, error and model variables (node data in 'onvertices' mode, cell
    data in 'volumetric' mode) go to <output_file>_kernel.xdmf/.nc; 'csr' or 'ascii':
    the kernels go to <output_file>_kernel.val/.ind/.pnt instead (|K| > allowed_error).
    In every case niterations, computation_time, element_proc and volume follow as cell
    data of <output_file>_kernel.xdmf/.nc (master_queue.f90:510-534).  volumes: per
    element; None computes tetrahedron volumes and leaves the variable out otherwise."""
    cells = results.int_type != ONVERTICES
    vertices = np.asarray(vertices, np.float64).reshape(-1, 3)
    connectivity = np.asarray(connectivity, np.int32)
    var = []
    if dump_type == "xdmf":
        var.append(("kernel", cells, results.export(MCK_RES_KERNEL), ["K_x_" + n for n in kernel_names]))
        var.append(("error", cells, results.export(MCK_RES_ERROR), ["err_" + n for n in kernel_names]))
        if results.nmodel:
            var.append(("model", cells, results.export(MCK_RES_MODEL), MODEL_NAMES[:results.nmodel]))
        if results.nhet:
            var.append(("hetero", cells, results.export(MCK_RES_HETERO), HETERO_NAMES[:results.nhet]))
    elif dump_type in ("csr", "ascii"):
        write_csr(output_file + "_kernel", results.export(MCK_RES_KERNEL), allowed_error, ascii=dump_type == "ascii")
    else:
        raise ValueError("dump_type must be 'xdmf', 'csr' or 'ascii'")
    raw = results.get()
    ne = results.nelements
    if volumes is None and element_type == "tet":
        v = vertices[connectivity - 1]
        volumes = np.abs(np.einsum("ij,ij->i", v[:, 1] - v[:, 0],
                                   np.cross(v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]))) / 6.0
    var.append(("niterations", True, results.export(MCK_RES_NITER), ["nit_" + n for n in kernel_names]))
    var.append(("computation_time", True, raw["computation_time"].astype(np.float32).reshape(1, ne), None))
    var.append(("element_proc", True, raw["element_proc"].astype(np.float32).reshape(1, ne), None))
    if volumes is not None:
        var.append(("volume", True, np.asarray(volumes, np.float32).reshape(1, ne), None))
    write_xdmf(output_file + "_kernel", element_type, vertices, connectivity, var, block_id)
    return var


__all__ = ["write_xdmf", "write_csr", "write_intermediate", "read_intermediate", "finalize",
           "MCK_RES_VARIANCE", "EXPORTS"]
