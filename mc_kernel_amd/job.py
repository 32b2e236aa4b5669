"""A whole MC Kernel run on one GPU: the master's task loop (master_queue.f90, get_next_task
:223-249 and extract_receive_buffer :253-323) and the slave's work units (do_slave,
slave.f90:205-219: the model integrals, then slave_work) in one process, with the
results accumulated in device memory and finalize's files written at the end.

Tasks are consecutive runs of `nelems_per_task` inversion elements.  Each element
reserves its own range of the Halton stream (the MCKernel's halton_step advances by
whole tasks), so without model integrals splitting the mesh into tasks does not change
any element's result; with them, each task's model integrals reserve their range of the
stream before its kernels do, as the reference draws them first."""
import numpy as np

from . import output
from .results import ONVERTICES, VOLUMETRIC


def run_job(kernel, vertices, connectivity, nbasis=1, nelems_per_task=1000,
            int_over_background=False, hetero=None, output_file=None, kernel_names=None,
            dump_type="xdmf", allowed_error=0.0, rank=1, progress=None):
    """Runs every element of the inversion mesh through `kernel` (an MCKernel).

    vertices      (nvertices, 3) metres; connectivity (nelements, 4) tetrahedra or
                  (nelements, 8) voxels, 1-based
    nbasis        1: volumetric results (one value per element); 4 (tetrahedra) or 8
                  (voxels): 'onvertices' results (one value per vertex)
    hetero        None, or (rtp, dlnhet) of a heterogeneity model to integrate as well
    output_file   None, or the prefix of finalize's files (<prefix>_kernel.xdmf/.nc ...)
    Returns the results.Results accumulators (device memory, valid until kernel.close())."""
    vertices = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
    conn = np.ascontiguousarray(connectivity, np.int32)
    ne, nvpe = conn.shape
    if nvpe not in (4, 8):
        raise ValueError("inversion elements must be tetrahedra (4 vertices) or voxels (8)")
    onvertices = nbasis > 1
    if hetero is not None:
        kernel.set_hetero(*hetero)
    res = kernel.results(ONVERTICES if onvertices else VOLUMETRIC,
                         vertices.shape[0] if onvertices else ne, ne, nbasis,
                         12 if int_over_background else 0, conn if onvertices else None,
                         nhet=3 if hetero is not None else 0)
    for t0 in range(0, ne, nelems_per_task):
        c = conn[t0:t0 + nelems_per_task]
        # the order of slave.f90:450-461: background model, heterogeneities, kernels; each
        # reserves its own stream range for the task's elements
        if int_over_background:
            kernel.integrate_model(vertices, c, nbasis=nbasis)
        if hetero is not None:
            kernel.integrate_hetero(vertices, c, nbasis=nbasis)
        kernel.slave_work(vertices, c, nbasis=nbasis)
        res.add(np.arange(t0 + 1, t0 + len(c) + 1, dtype=np.int32), rank=rank)
        if progress:
            progress(t0 + len(c), ne)
    if output_file:
        names = kernel_names or ["k%d" % i for i in range(kernel.nk)]
        output.finalize(res, output_file, names, "tet" if nvpe == 4 else "vox", vertices, conn,
                        dump_type=dump_type, allowed_error=allowed_error)
    return res
