#!/usr/bin/env python3
"""Throughput of the MI355X MC Kernel path: MC point-kernel evaluations per second.

Workload c1 (one GPU): merged fwd/bwd databases of a 20,000-quad axisymmetric mesh with
nt = 2048 (nfft 4096), 10 receivers at 10..100 deg, Gabor filters 20/30/40/60 s, CC and AM
misfits for vp and vs (160 kernels), 50,000 tetrahedral voxels with Halton points,
16 points per MC step, relative MC tolerance 1e-3.  The databases are generated directly
in HBM.  A step is one task of --task voxels evaluated by mck_slave_work (the reference's
work unit, slave.f90:273-810); the value is whole-job evaluations per second over the
timed steps.  One evaluation is one (voxel, MC point, kernel) value actually computed
(converged and travel-time-masked kernels excluded).

Also reported: the roofline of the dominant kernel (algorithmic bytes per launch over its
hipEvent time), the FP64 CPU oracle timed on a bounded sample on the host cores, and the
plain one-point-at-a-time GPU form on a small sample.
"""
import argparse
import ctypes as C
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MC point-kernel evals/sec at <=1e-5 rel err vs FP64; % of HBM peak"
HBM_PEAK_GBS = 8000.0


def workload(name):
    from mc_kernel_amd import synth
    return synth.make_problem(name, fill=False)


def unmasked_voxels(p):
    """Voxels with at least one receiver that is not travel-time masked."""
    from mc_kernel_amd import synth
    v = p.verts.reshape(p.ntet, 4, 3)
    c = v.mean(axis=1)
    src = synth.geo_xyz(p.src_lat, p.src_lon, synth.R_EARTH - p.src_depth * 1e3)
    vp = float(p.model[0].max()); vs = float(p.model[1].max())
    keep = np.zeros(p.ntet, bool)
    for r in range(p.nrec):
        rec = synth.geo_xyz(p.rec_lat[r], p.rec_lon[r], synth.R_EARTH)
        d = np.linalg.norm(c - src, axis=1) + np.linalg.norm(c - rec, axis=1)
        ks = p.kernel_receiver == r
        params = p.kernel_param[ks]
        tw = p.kernel_window[ks][:, 1]
        tlp = tw[np.isin(params, [0, 1, 6, 7])].max(initial=0.0)
        tls = tw[~np.isin(params, [0, 1, 6, 7])].max(initial=0.0)
        keep |= ~((d / vp > tlp) & (d / vs > tls))
    return np.nonzero(keep)[0]


def _cpu_worker(args):
    """One host core: the FP64 oracle on a bounded sample of voxels (lazy database);
    returns the evaluations, the timed seconds and the oracle's results of the voxels."""
    name, voxels, max_iter = args
    sys.path.insert(0, ROOT)
    from mc_kernel_amd import synth
    from mc_kernel_amd import abi
    from oracle import oracle
    p = synth.make_problem(name, fill=False, max_iter=max_iter)
    m = p.mesh
    fwd = np.empty((m.nelem, 10, 5, 5, p.nt), np.float32)
    bwd = np.empty((m.nelem, 5, 5, 5, p.nt), np.float32)
    p.setup.fwd.snapshots = abi.ptr(fwd, C.c_float)
    p.setup.bwd.snapshots = abi.ptr(bwd, C.c_float)
    lib = synth.synth_lib()
    lib.synth_lazy_create.argtypes = [abi.i32p, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                      C.c_uint64, C.c_uint64]
    lib.synth_lazy_create.restype = C.c_void_p
    user = lib.synth_lazy_create(abi.ptr(m.sem, C.c_int32), 10, 5, p.nt, p.dt, 1001, 1002)
    o = oracle.Oracle(p.setup)
    o.set_lazy_fill(C.cast(lib.synth_fill_element, C.c_void_p), user)
    conn = p.conn[voxels]
    stride = (p.max_iter + 1) * p.npoints_per_step
    offs = np.asarray(voxels, np.int64) * stride
    o.slave_work(p.verts, conn, halton_offsets=offs)          # generates the touched elements
    t0 = time.perf_counter()
    r = o.slave_work(p.verts, conn, halton_offsets=offs)      # timed
    return r["evaluations"], time.perf_counter() - t0, voxels, r["values"], r["niterations"]


def c4_task(layer, nvox):
    """The c4 problem and its task: the nvox voxels of the 1.25 deg grid's layer `layer`
    nearest to the epicentre (corners [8 nvox][3], connectivity [nvox][8])."""
    from mc_kernel_amd import synth
    p = synth.make_problem("c4")
    v, c, L, _ = synth.voxel_grid()
    sel = synth.voxel_slice(v, c, L, layer, p.src_lat, p.src_lon, nvox)
    verts = np.ascontiguousarray(v.reshape(-1, 8, 3)[sel].reshape(-1, 3))
    conn = np.arange(1, 8 * len(sel) + 1, dtype=np.int32).reshape(-1, 8)
    return p, verts, conn


def _cpu_worker_c4(args):
    """One host core: the FP64 oracle on voxels of the c4 slice, its databases generated
    element by element as they are touched into sparse host mappings; the same return
    as _cpu_worker."""
    layer, nvox, voxels, max_iter = args
    sys.path.insert(0, ROOT)
    from mc_kernel_amd import synth
    from oracle import oracle
    p, verts, conn = c4_task(layer, nvox)
    p.max_iter = max_iter
    p.setup.max_iter = max_iter
    m = p.mesh
    maps = []
    for db, nvar in ((p.setup.fwd, 10), (p.setup.bwd, 5)):
        n = m.nelem * nvar * 25 * p.nt * 4
        a = synth.sparse_mapping(n)
        maps.append((a, n))
        db.snapshots = C.cast(C.c_void_p(a), C.POINTER(C.c_float))
    lib = synth.synth_lib()
    from mc_kernel_amd import abi
    lib.synth_lazy_create.argtypes = [abi.i32p, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                      C.c_uint64, C.c_uint64]
    lib.synth_lazy_create.restype = C.c_void_p
    user = lib.synth_lazy_create(abi.ptr(m.sem, C.c_int32), 10, 5, p.nt, p.dt, 1001, 1002)
    o = oracle.Oracle(p.setup)
    o.set_lazy_fill(C.cast(lib.synth_fill_element, C.c_void_p), user)
    sub = conn[voxels]
    offs = np.asarray(voxels, np.int64) * ((max_iter + 1) * p.npoints_per_step)
    o.slave_work(verts, sub, halton_offsets=offs)                # generates the touched elements
    t0 = time.perf_counter()
    r = o.slave_work(verts, sub, halton_offsets=offs)            # timed
    dt = time.perf_counter() - t0
    return r["evaluations"], dt, voxels, r["values"], r["niterations"]


def cpu_baseline_c4(workers, layer, nvox, max_iter):
    voxels = np.linspace(0, nvox - 1, workers).astype(int)
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(workers) as pool:
        res = pool.map(_cpu_worker_c4, [(layer, nvox, np.array([v]), max_iter) for v in voxels])
    wall = time.perf_counter() - t0
    evals = sum(r[0] for r in res)
    tmax = max(r[1] for r in res)
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    line = {"value": evals / tmax, "unit": "evals/s", "cores": workers, "kind": "port",
            "sample": "%d voxels of the c4 slice (one per core), %d MC step(s) each (max_iter=%d), 16 points/step; "
                      "FP64 oracle, one process per core, timed region = slave_work only" % (workers, max_iter, max_iter),
            "seconds": tmax, "evaluations": int(evals), "total_wall_s": wall,
            "nproc": os.cpu_count(), "affinity_cpus": affinity, "cpu_model": cpu_model()}
    vox = np.concatenate([r[2] for r in res])
    vals = np.concatenate([r[3] for r in res])
    nit = np.concatenate([r[4] for r in res])
    return line, (vox, vals, nit, max_iter)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(name, workers, per_worker, max_iter):
    p = workload(name)
    cand = unmasked_voxels(p)
    need = workers * per_worker
    sel = cand[np.linspace(0, len(cand) - 1, need).astype(int)] if len(cand) >= need else cand
    chunks = [sel[i::workers] for i in range(workers)]
    chunks = [c for c in chunks if len(c)]
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(len(chunks)) as pool:
        res = pool.map(_cpu_worker, [(name, c, max_iter) for c in chunks])
    wall = time.perf_counter() - t0
    evals = sum(r[0] for r in res)
    tmax = max(r[1] for r in res)
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    line = {"value": evals / tmax, "unit": "evals/s", "cores": len(chunks), "kind": "port",
            "sample": "%d unmasked voxels of %s, %d MC steps each (max_iter=%d), %d points/step; "
                      "FP64 oracle, one process per core, timed region = slave_work only"
                      % (len(sel), name, max_iter, max_iter, p.npoints_per_step),
            "seconds": tmax, "evaluations": int(evals), "total_wall_s": wall,
            "nproc": os.cpu_count(), "affinity_cpus": affinity, "cpu_model": cpu_model()}
    # the oracle's results of the sample, for the GPU's self-check on the same voxels
    vox = np.concatenate([r[2] for r in res])
    vals = np.concatenate([r[3] for r in res])
    nit = np.concatenate([r[4] for r in res])
    return line, (vox, vals, nit, max_iter)


def self_check(p, sample):
    """Runs the CPU baseline's voxels on the GPU with the oracle's stream positions and
    MAXIMUM_ITERATIONS (the databases are bit-identical: the device generator reproduces
    the host generator's samples) and compares: niterations bit for bit, kernel values as
    the largest relative L2 error over kernels with a nonzero oracle norm."""
    import mc_kernel_amd as mk
    vox, ref_vals, ref_nit, max_iter = sample
    offs = np.asarray(vox, np.int64) * ((max_iter + 1) * p.npoints_per_step)
    old = p.setup.max_iter
    p.setup.max_iter = max_iter
    try:
        k = mk.MCKernel(p.setup)
        try:
            r = k.slave_work(p.verts, p.conn[vox], halton_offsets=offs)
        finally:
            k.close()
    finally:
        p.setup.max_iter = old
    errs = []
    for j in range(ref_vals.shape[1]):
        n = np.linalg.norm(ref_vals[:, j])
        if n > 0:
            errs.append(float(np.linalg.norm(r["values"][:, j] - ref_vals[:, j]) / n))
    return {"voxels": int(len(vox)), "max_iter": int(max_iter), "max_rel_err": max(errs) if errs else None,
            "kernels_compared": len(errs), "niterations_equal": bool(np.array_equal(r["niterations"], ref_nit)),
            "tolerance": 1e-5}


def page_db_to_host(p, lib, dbs, pinned):
    """Moves the generated databases to host memory (page-locked or pageable), for the
    library to page them (mck_setup_t.db_page_elems); returns the host buffers to keep."""
    m = p.mesh
    keep = []
    lib.synth_gpu_download.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    lib.synth_gpu_free.argtypes = [C.c_void_p]
    lib.synth_gpu_host_alloc.argtypes = [C.c_size_t, C.POINTER(C.c_void_p)]
    for i, (nvar, db) in enumerate(((10, p.setup.fwd), (5, p.setup.bwd))):
        n = m.nelem * nvar * 25 * p.nt
        if pinned:
            h = C.c_void_p()
            if lib.synth_gpu_host_alloc(n * 4, C.byref(h)):
                raise RuntimeError("page-locked host allocation failed")
            ptr = h.value
            keep.append(h)
        else:
            arr = np.empty(n, np.float32)
            ptr = arr.ctypes.data
            keep.append(arr)
        if lib.synth_gpu_download(dbs[i], ptr, n):
            raise RuntimeError("database download failed")
        lib.synth_gpu_free(dbs[i])
        db.snapshots = C.cast(C.c_void_p(ptr), C.POINTER(C.c_float))
    return keep


def measured_traffic(path, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary: lines
    'traffic <kernel>: <bytes per launch>' written from the FETCH_SIZE and WRITE_SIZE
    runs, each scaled by the same counter's reading for a plain copy kernel."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for ln in f:
            if ln.startswith("traffic %s:" % kernel):
                return float(ln.split(":", 1)[1].split()[0])
    return None


def measured_counters(path, kernel):
    """Issue rates of `kernel` from the committed counter summary: the row
    '| <kernel> | VALU busy % | MFMA busy % | VALU instructions per wave | MFMA instructions
    per wave |' of its derived table."""
    if not os.path.exists(path):
        return None
    keys = ("valu_busy", "mfma_busy", "valu_insts_per_wave", "mfma_insts_per_wave")
    with open(path) as f:
        for ln in f:
            cells = [c.strip() for c in ln.strip().strip("|").split("|")]
            if len(cells) >= 5 and cells[0] == kernel:
                try:
                    v = [float(c) for c in cells[1:5]]
                except ValueError:
                    continue
                return {keys[0]: round(v[0] / 100, 3), keys[1]: round(v[1] / 100, 3), keys[2]: v[2], keys[3]: v[3]}
    return None


def bench_autograd(args):
    """c3: the stage of the c1 workload (nt 2048, 10 receivers, 4 Gabor filters, 160
    kernels) at --autograd-points points with their strain traces; forward and the
    hand-written adjoint (FFTs recomputed), timed with events on torch's stream."""
    import torch
    import mc_kernel_amd as mk
    from mc_kernel_amd import autograd as ad
    from mc_kernel_amd import synth
    p = synth.make_problem("c1", nelem_target=2000, ntet=200)
    k = mk.MCKernel(p.setup)
    nd, nt, nw, nf, nr, nk = ad.dims()
    P = args.autograd_points
    x = p.sample_points(P, seed=31)
    fw = torch.from_numpy(k.load_strain(0, x)).cuda()
    bw = torch.empty((P, nr, nd, nt), dtype=torch.float32, device="cuda")
    for r in range(nr):
        bw[:, r] = torch.from_numpy(k.load_strain(1 + r, x)).cuda()
    tf = ad.transfer_functions()
    fw.requires_grad_(True); bw.requires_grad_(True); tf.requires_grad_(True)
    w = torch.rand((P, nk, 6), dtype=torch.float64, device="cuda")
    fwd_ms, bwd_ms = [], []
    for i in range(args.warmup + args.steps):
        fw.grad = bw.grad = tf.grad = None
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record()
        base = ad.filter_convolve_window(tf, fw, bw)
        e1.record()
        (base * w).sum().backward()
        e2.record()
        torch.cuda.synchronize()
        if i >= args.warmup:
            fwd_ms.append(e0.elapsed_time(e1))
            bwd_ms.append(e1.elapsed_time(e2))
        print("autograd step %d: forward %.2f ms, backward %.2f ms" % (i, e0.elapsed_time(e1), e1.elapsed_time(e2)),
              file=sys.stderr, flush=True)
    k.close()
    tf_b, fw_b, bw_b = 8 * nf * nw, 4 * P * nd * nt, 4 * P * nr * nd * nt
    base_b = 8 * P * nk * 6
    f_ms, b_ms = float(np.mean(fwd_ms)), float(np.mean(bwd_ms))
    line = {
        "metric": "c3 filter-convolve-window-integrate stage: forward and adjoint time",
        "value": f_ms + b_ms, "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": f_ms + b_ms, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (f64 accumulation)", "data": "strain traces of a synthetic DB at random points",
        "config": {"workload": "c3: c1 stage (nt 2048, 10 receivers, 4 Gabor, 160 kernels), "
                               "%d points, gradients w.r.t. G, fwd and bwd traces" % P},
        "forward_ms": f_ms, "backward_ms": b_ms,
        "forward_algorithmic_GBs": (tf_b + fw_b + bw_b + base_b) / (f_ms * 1e-3) / 1e9,
        "backward_algorithmic_GBs": (2 * (tf_b + fw_b + bw_b) + 2 * base_b) / (b_ms * 1e-3) / 1e9,
    }
    print(json.dumps(line), flush=True)


def bench_assembly(args):
    """Result assembly at the c4 scale (0.8M voxels, 1600 kernels; tasks of
    --assembly-task voxels, volumetric): each step adds one task's kernel values and
    variances, already in device memory, to the device-resident K_x and Var
    (mck_results_add, extract_receive_buffer master_queue.f90:253-323)."""
    import torch
    import mc_kernel_amd as mk
    from mc_kernel_amd import synth
    from mc_kernel_amd.results import VOLUMETRIC
    npoints, nk, task = args.assembly_points, 1600, args.assembly_task
    p = synth.make_problem("tiny", receivers_deg=[20.0 + i for i in range(50)],
                           filters=[("Gabor", (30.0 + 5.0 * i, 0.5, 0.0, 0.0)) for i in range(8)],
                           kernels=synth._kernels(50, 8, ["CC", "AM"], ["vp", "vs"]))
    assert p.nkernel == nk
    cpu = None
    if not args.no_cpu:
        # the serial accumulation of the reference's master loop, on a bounded sample
        from oracle import oracle
        rng = np.random.default_rng(0)
        ncpu = min(task, 20000)
        ref = oracle.Results(VOLUMETRIC, ncpu, ncpu, 1, nk)
        vals = rng.standard_normal((ncpu, 1, nk))
        nit = np.ones((ncpu, nk), np.int32)
        elems = np.arange(1, ncpu + 1, dtype=np.int32)
        t0 = time.perf_counter()
        for _ in range(3):
            ref.add(elems, vals, vals, nit)
        dt = (time.perf_counter() - t0) / 3
        cpu = {"value": ncpu * nk / dt, "unit": "kernel values/s", "cores": 1, "kind": "port",
               "sample": "3 tasks of %d voxels x %d kernels into %d-voxel accumulators (FP64 oracle, serial)"
                         % (ncpu, nk, ncpu)}
    k = mk.MCKernel(p.setup)
    res = k.results(VOLUMETRIC, npoints, npoints, 1, 0)
    g = torch.Generator(device="cuda").manual_seed(1)
    vals = torch.randn((task, 1, nk), dtype=torch.float64, device="cuda", generator=g)
    var = vals.abs()
    nit = torch.ones((task, nk), dtype=torch.int32, device="cuda")
    ct = torch.zeros(task, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ntask = npoints // task
    times = []
    for i in range(args.warmup + args.steps):
        elems = (np.arange(task, dtype=np.int32) + (i % ntask) * task + 1).astype(np.int32)
        t0 = time.perf_counter()
        res.add(elems, vals, var, nit, ct, None, 1)
        t1 = time.perf_counter()
        if i >= args.warmup:
            times.append(t1 - t0)
        print("assembly step %d: %.3f ms" % (i, (t1 - t0) * 1e3), file=sys.stderr, flush=True)
    k.close()
    ms = float(np.mean(times)) * 1e3
    # algorithmic bytes of one k_results_add launch (K_x and Var together): the task's
    # values and variances read once, both accumulators' rows read and written once
    launch_bytes = 2 * task * nk * 8 * 3
    line = {
        "metric": "result assembly: kernel values accumulated per second (K_x and Var)",
        "value": task * nk / (ms * 1e-3), "unit": "kernel values/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic task results in device memory",
        "config": {"workload": "c4 accumulators: %d voxels x %d kernels, volumetric; tasks of %d voxels"
                               % (npoints, nk, task)},
        "k_results_add_algorithmic_bytes": launch_bytes, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_integrals(args):
    """The model integrals of a c1-sized task (50,000 tetrahedra, one basis function):
    integrate_1d_model (mck_integrate_model, 100 points per step) and integrate_3d_model
    (mck_integrate_hetero, 10 points per step) over a global heterogeneity model on a
    1-degree grid at 10 radii (648,000 points); the FP64 oracle on a bounded sample of
    the same elements as the CPU baseline."""
    import mc_kernel_amd as mk
    from mc_kernel_amd import synth
    p = synth.make_problem("tiny", ntet=args.integrals_tets, tet_edge=300e3)
    verts, conn = np.asarray(p.verts, float).reshape(-1, 3), np.asarray(p.conn, np.int32)
    r = np.linspace(3500.0, 6371.0, 10)
    lat = np.arange(-89.5, 90.0, 1.0)
    lon = np.arange(-179.5, 180.0, 1.0)
    rtp = np.stack(np.meshgrid(r, lat, lon, indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(2)
    dln = (rng.standard_normal((len(rtp), 3)) * 0.05).astype(np.float32)
    cpu = None
    if not args.no_cpu:
        from oracle import oracle
        ns = min(len(conn), 200)
        o = oracle.Oracle(p.setup)
        t0 = time.perf_counter()
        o.integrate_model(verts, conn[:ns])
        t1 = time.perf_counter()
        oracle.integrate_hetero(rtp, dln, verts, conn[:20])
        t2 = time.perf_counter()
        cpu = {"model": {"value": ns / (t1 - t0), "unit": "elements/s", "cores": 1, "kind": "port",
                         "sample": "%d elements, FP64 oracle" % ns},
               "hetero": {"value": 20 / (t2 - t1), "unit": "elements/s", "cores": 1, "kind": "port",
                          "sample": "20 elements, FP64 oracle with brute-force nearest points"}}
    k = mk.MCKernel(p.setup)
    k.set_hetero(rtp, dln)
    tm, th, steps_m, steps_h = [], [], 0, 0
    for i in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        _, nsm = k.integrate_model(verts, conn)
        t1 = time.perf_counter()
        _, nsh = k.integrate_hetero(verts, conn)
        t2 = time.perf_counter()
        if i >= args.warmup:
            tm.append(t1 - t0)
            th.append(t2 - t1)
            steps_m, steps_h = int(nsm.sum()), int(nsh.sum())
        print("integrals step %d: model %.1f ms, hetero %.1f ms" % (i, (t1 - t0) * 1e3, (t2 - t1) * 1e3),
              file=sys.stderr, flush=True)
    k.close()
    m, h = float(np.mean(tm)), float(np.mean(th))
    line = {
        "metric": "model integrals: elements integrated per second (1-D background, 3-D heterogeneities)",
        "value": len(conn) / m, "unit": "elements/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": (m + h) * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (fp32 model values)", "data": "synthetic tetrahedra and heterogeneity grid",
        "config": {"workload": "%d tetrahedra, 1 basis function; heterogeneity model %d points (1 deg x 10 radii)"
                               % (len(conn), len(rtp))},
        "model_ms": m * 1e3, "hetero_ms": h * 1e3, "hetero_elements_per_s": len(conn) / h,
        "model_mc_steps": steps_m, "hetero_mc_steps": steps_h, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def roofline(steps, config, traffic_path):
    """The roofline object of the dominant stage over the timed steps' statistics."""
    launches = sum(s[1]["strain_launches"] for s in steps) // 2
    fwd_b = sum(s[1]["fwd_bytes"] for s in steps)
    bwd_b = sum(s[1]["bwd_bytes"] for s in steps)
    fwd_ms = sum(s[1]["strain_fwd_ms"] for s in steps)
    bwd_ms = sum(s[1]["strain_bwd_ms"] for s in steps)
    sp_ms = sum(s[1]["spectra_ms"] for s in steps)
    mf_ms = sum(s[1]["mfma_ms"] for s in steps)
    mf_launches = sum(s[1]["mfma_launches"] for s in steps)
    mfma_path = all(s[1]["bwd_path"] == 1 for s in steps)
    # the backward stage dominates: with the matrix-core path it is two kernels, the
    # element spectra (which read the merged blocks) and the matrix-core evaluation; its
    # algorithmic bytes are the backward blocks of every receiver's element set
    if mfma_path:
        bwd_kern, bwd_kms = "k_elem_spectra+k_bwd_mfma", sp_ms + mf_ms
    else:
        bwd_kern, bwd_kms = "k_bwd_eval", bwd_ms
    if bwd_kms >= fwd_ms:
        kern, byt, ms = bwd_kern, bwd_b, bwd_kms
    else:
        kern, byt, ms = "k_fwd_traces+k_fwd_spectra", fwd_b, fwd_ms
    achieved = (byt / launches) / (ms / launches * 1e-3) / 1e9 if ms > 0 else 0.0
    # the committed counter summary is of the c1 workload
    traffic = measured_traffic(traffic_path, kern) if config == "c1" else None
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kern,
            "bytes_per_launch": byt / max(launches, 1), "avg_launch_ms": ms / max(launches, 1),
            "launches": launches,
            "algorithmic_bytes": "merged blocks once per launch: fwd 4 nt 25 x 10 variables per distinct "
                                 "forward element (+ 48 B per point), bwd 4 nt 25 x 5 variables per distinct "
                                 "(receiver, backward element)",
            "alg_bytes_whole_job": sum(s[1]["alg_bytes"] for s in steps) / len(steps),
            "timing": "HIP events on the library's stream around each kernel of every launch"}
    if mfma_path:
        roof["kernels"] = {
            "k_elem_spectra": {"avg_launch_ms": sp_ms / max(mf_launches, 1), "launches": mf_launches},
            "k_bwd_mfma": {"avg_launch_ms": mf_ms / max(mf_launches, 1), "launches": mf_launches,
                           "waves_per_launch": sum(s[1]["mfma_groups"] for s in steps) / max(mf_launches, 1)},
            "k_fwd_traces+k_fwd_spectra": {"avg_launch_ms": fwd_ms / max(launches, 1), "launches": launches},
            "backward_stage_ms_per_launch": bwd_ms / max(launches, 1)}
    if traffic:
        # the counter-measured HBM bytes of the same kernels over their launch time
        roof["measured_frac"] = round(traffic / (ms / launches * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    if config == "c1":
        # what does bound the matrix-core kernel: its issue rates, from the committed counters
        busy = measured_counters(os.path.join(ROOT, "profiles", "c1_bwd_mfma_counters.md"), "k_bwd_mfma")
        if busy:
            roof["k_bwd_mfma_counters"] = busy

    return roof


def bench_c4(args):
    """c4: 200 receivers x 4 Gabor x CC x vp/vs (1600 kernels), 100,000-quad databases at
    nt 4096 (614 GB, more than HBM and than the host memory a run may take).  Both
    databases are reserved as sparse host mappings; the element pages that the voxel
    slice can touch in any frame (synth.touched_elements) are generated there and the
    library pages them (db_page_elems).  A step is the slice as one task."""
    import mc_kernel_amd as mk
    from mc_kernel_amd import synth
    cpu, sample = None, None
    if not args.no_cpu:
        cpu, sample = cpu_baseline_c4(args.cpu_workers, args.c4_layer, args.c4_voxels, args.c4_cpu_steps)
        print("cpu baseline:", json.dumps(cpu), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    p, verts, conn = c4_task(args.c4_layer, args.c4_voxels)
    src = synth.geo_xyz(p.src_lat, p.src_lon, 1.0)
    recs = np.array([synth.geo_xyz(a, b, 1.0) for a, b in zip(p.rec_lat, p.rec_lon)])
    masks = (synth.touched_elements(p.mesh, src[None], verts, conn),
             synth.touched_elements(p.mesh, recs, verts, conn))
    page = args.db_page_elems if args.db_page_elems > 0 else 256
    dbs = synth.PagedSyntheticDatabases(p, masks, page)
    npages = [len(x) for x in dbs.pages]
    p.setup.db_device_pages = args.db_device_pages if args.db_device_pages > 0 else max(npages) + 2
    setup_s = time.perf_counter() - t0
    print("c4: %d voxels of layer %d, %d + %d pages of %d elements generated (%.1f GB) in %.1f s"
          % (len(conn), args.c4_layer, npages[0], npages[1], page, dbs.generated_bytes / 1e9, setup_s),
          file=sys.stderr, flush=True)
    k = mk.MCKernel(p.setup)
    steps, warm_paged = [], 0
    for i in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        r = k.slave_work(verts, conn)
        dt = time.perf_counter() - t0
        if i >= args.warmup:
            steps.append((dt, r["stats"]))
        else:
            warm_paged += r["stats"]["paged_bytes"]
        print("step %d: %.3f s, %d evals, %d MC steps, %.1f GB paged" % (
              i, dt, r["stats"]["evaluations"], r["stats"]["mc_steps"], r["stats"]["paged_bytes"] / 1e9),
              file=sys.stderr, flush=True)
    k.close()
    tot_t = sum(s[0] for s in steps)
    tot_e = sum(s[1]["evaluations"] for s in steps)
    roof = roofline(steps, "c4", None)
    check = None
    if sample is not None:
        check = self_check_task(p, verts, conn, sample)
        print("self-check:", json.dumps(check), file=sys.stderr, flush=True)
    dbs.free()
    line = {
        "metric": METRIC, "value": tot_e / tot_t, "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": tot_t / len(steps) * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 accumulation)",
        "data": "synthetic (seeded AxiSEM-style merged DB; the pages the slice touches generated into sparse host "
                "mappings and paged to HBM)",
        "config": {"workload": "c4: %d-quad DB, nt=%d (nfft %d), 200 receivers at the Gibraltar benchmark "
                               "stations, Gibraltar CMT, 4 Gabor filters, CC x vp/vs (1600 kernels), slice of "
                               "the 30-layer 1.25 deg voxel grid: the %d voxels of layer %d nearest the "
                               "epicentre, 16 pts/step, max_iter 64, rel tol 0.001, STF deconvolution on"
                               % (p.mesh.nelem, p.nt, 2 * p.nt, len(conn), args.c4_layer),
                   "task_voxels": int(len(conn)), "parallelism": "none (1 GPU)", "evaluations": int(tot_e)},
        "paging": {"page_elems": page, "pages_generated": {"fwd": npages[0], "bwd": npages[1]},
                   "generated_gb": dbs.generated_bytes / 1e9,
                   "db_bytes_total": float(p.mesh.nelem * 15 * 25 * p.nt * 4),
                   "device_pages_per_db": int(p.setup.db_device_pages),
                   "paged_gb_warmup": warm_paged / 1e9,
                   "paged_gb_timed": sum(s[1]["paged_bytes"] for s in steps) / 1e9},
        "roofline": roof, "cpu_baseline": cpu, "self_check": check,
    }
    print(json.dumps(line), flush=True)


def self_check_task(p, verts, conn, sample):
    """The CPU sample's voxels of a task on the GPU with the oracle's stream positions and
    MAXIMUM_ITERATIONS; niterations bit for bit and the largest relative L2 error."""
    import mc_kernel_amd as mk
    vox, ref_vals, ref_nit, max_iter = sample
    offs = np.asarray(vox, np.int64) * ((max_iter + 1) * p.npoints_per_step)
    old = p.setup.max_iter
    p.setup.max_iter = max_iter
    try:
        k = mk.MCKernel(p.setup)
        try:
            r = k.slave_work(verts, conn[vox], halton_offsets=offs)
        finally:
            k.close()
    finally:
        p.setup.max_iter = old
    errs = []
    for j in range(ref_vals.shape[1]):
        n = np.linalg.norm(ref_vals[:, j])
        if n > 0:
            errs.append(float(np.linalg.norm(r["values"][:, j] - ref_vals[:, j]) / n))
    return {"voxels": int(len(vox)), "max_iter": int(max_iter), "max_rel_err": max(errs) if errs else None,
            "kernels_compared": len(errs), "niterations_equal": bool(np.array_equal(r["niterations"], ref_nit)),
            "tolerance": 1e-5}


CANDIDATES = ["per-pair FFTs (k_bwd_eval)", "matrix cores, 1 wave per 16 pairs",
              "matrix cores, 1 wave per 16 pairs, XCD-contiguous groups",
              "matrix cores, 2 waves per element (spectra in LDS)", "matrix cores, 4 waves per element (spectra in LDS)"]
CAND_ORDER = [4, 0, 1, 2, 3]          # candidate_ms index of each row above


def tune_report(args):
    """mck_tune on each workload's kernel set (2048 points of its first tetrahedra; c4 on a
    3,000-quad database): the time of every candidate, the choice and its gain over the
    default (one JSON line)."""
    import mc_kernel_amd as mk
    from mc_kernel_amd import synth
    rows = []
    for name in ("c0", "c1", "c2", "c4"):
        kw = {"nelem_target": 3000, "ntet": 64} if name == "c4" else {}
        p = synth.make_problem(name, fill=False, **kw)
        dev = synth.DeviceDatabases(p)
        k = mk.MCKernel(p.setup)
        try:
            x = p.sample_points(2048, seed=17)
            t = k.tune(x, reps=3)
        finally:
            k.close()
            dev.free()
        cms = t["candidate_ms"]
        row = {"config": name, "nt": p.nt, "kernels": p.nkernel, "receivers": p.nrec,
               "candidates_ms": {CANDIDATES[i]: (cms[j] or None) for i, j in enumerate(CAND_ORDER)},
               "per_launch_choice_ms": t["auto_ms"] or None,
               "costs_us": {"group": t["cost_group_us"], "element_run": t["cost_run_us"], "pair": t["cost_pair_us"]},
               "chosen": {"bwd_path": t["bwd_path"], "mfma_waves": t["mfma_waves"], "mfma_xcd": t["mfma_xcd"]},
               "best_ms": t["best_ms"], "default_ms": t["default_ms"],
               "gain_over_default": t["default_ms"] / t["best_ms"] if t["best_ms"] else None,
               "gain_over_per_pair": (cms[4] / t["best_ms"]) if cms[4] and t["best_ms"] else None}
        rows.append(row)
        print("tune %s: %s" % (name, json.dumps(row)), file=sys.stderr, flush=True)
    print(json.dumps({"tune_report": rows, "points": 2048, "reps": 3,
                      "timed": "mck_base_kernels of the points (all receivers), best of 3 after one untimed run"}),
          flush=True)


def task_curve(k, p, sizes, v0, nvox):
    """evals/s of the same voxels [v0, v0 + nvox) run as tasks of each size (the master
    hands out tasks of ELEMENTS_PER_TASK voxels, master_queue.f90:223-249); one warm-up
    task of each size first."""
    out = []
    for T in sizes:
        T = min(T, nvox)
        k.slave_work(p.verts, p.conn[v0:v0 + T])
        evals, dev_ms, steps, t0 = 0, 0.0, 0, time.perf_counter()
        for a in range(v0, v0 + nvox, T):
            r = k.slave_work(p.verts, p.conn[a:min(a + T, v0 + nvox)])
            evals += r["stats"]["evaluations"]
            dev_ms += r["stats"]["device_ms"]
            steps += r["stats"]["mc_steps"]
        dt = time.perf_counter() - t0
        out.append({"task_voxels": T, "tasks": (nvox + T - 1) // T, "seconds": dt, "evals_per_s": evals / dt,
                    "device_seconds": dev_ms * 1e-3, "mc_steps": steps,
                    "host_ms_per_mc_step": (dt - dev_ms * 1e-3) * 1e3 / max(steps, 1)})
        print("task curve: %d-voxel tasks: %.1f M evals/s (%.2f s, %.2f s on the device, %d MC steps)"
              % (T, evals / dt / 1e6, dt, dev_ms * 1e-3, steps), file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c1")
    ap.add_argument("--task", type=int, default=0, help="voxels per step (0 = whole job)")
    ap.add_argument("--first-voxel", type=int, default=0,
                    help="tasks start at this voxel (the synthetic meshes are spatially clustered)")
    # the GPU box's CPU share is 16 cores (its process guard counts workers beyond them)
    ap.add_argument("--cpu-workers", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--cpu-voxels-per-worker", type=int, default=4)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-voxels", type=int, default=16)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run a plain copy kernel of this size (HBM counter calibration)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "c1_hbm_traffic.md"),
                    help="summary of the FETCH_SIZE/WRITE_SIZE counter runs for the roofline")
    ap.add_argument("--autograd", action="store_true",
                    help="time the differentiable filter-convolve-window-integrate stage at "
                         "the c3 size instead (forward and adjoint, one JSON line)")
    ap.add_argument("--autograd-points", type=int, default=10000)
    ap.add_argument("--assembly", action="store_true",
                    help="time the result assembly at the c4 scale instead (one JSON line)")
    ap.add_argument("--assembly-points", type=int, default=800000)
    ap.add_argument("--assembly-task", type=int, default=50000)
    ap.add_argument("--integrals", action="store_true",
                    help="time the model and heterogeneity integrals of a task instead (one JSON line)")
    ap.add_argument("--integrals-tets", type=int, default=50000)
    ap.add_argument("--task-curve", default="",
                    help="comma-separated task sizes: evals/s of the voxels [--first-voxel, "
                         "+--curve-voxels) run as tasks of each size (one JSON field)")
    ap.add_argument("--curve-voxels", type=int, default=10000)
    ap.add_argument("--db-page-elems", type=int, default=0,
                    help="keep the databases in host memory and page them, N elements per "
                         "page (0: resident in HBM, the default workload)")
    ap.add_argument("--db-host-memory", choices=["pinned", "pageable"], default="pinned")
    ap.add_argument("--db-device-pages", type=int, default=0,
                    help="device page buffers per database when paging (0: 2; c4: every page generated)")
    ap.add_argument("--tune", action="store_true",
                    help="mck_tune on 2048 points of the task before the timed steps (choice in the line)")
    ap.add_argument("--tune-report", action="store_true",
                    help="the tuner's candidate times for c0, c1, c2 and c4 instead (one JSON line)")
    ap.add_argument("--c4-layer", type=int, default=15, help="c4: layer of the voxel slice (0 = top)")
    ap.add_argument("--c4-voxels", type=int, default=1024, help="c4: voxels of the slice")
    ap.add_argument("--c4-cpu-steps", type=int, default=1, help="c4: MC steps of the CPU sample's voxels")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(the path would shard by voxels, one process per GPU; not built here)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    if args.autograd:
        bench_autograd(args)
        return
    if args.assembly:
        bench_assembly(args)
        return
    if args.integrals:
        bench_integrals(args)
        return
    if args.tune_report:
        tune_report(args)
        return
    if args.config == "c4":
        bench_c4(args)
        return

    # CPU baseline first, in fresh processes, before this process touches the GPU
    cpu, sample = None, None
    if not args.no_cpu:
        cpu, sample = cpu_baseline(args.config, args.cpu_workers, args.cpu_voxels_per_worker, args.cpu_steps)
        print("cpu baseline:", json.dumps(cpu), file=sys.stderr, flush=True)

    import mc_kernel_amd as mk
    from mc_kernel_amd import synth
    p = workload(args.config)
    devdb = synth.DeviceDatabases(p)
    lib, dbs = devdb.lib, devdb.ptrs
    if args.copy_calibration_mib > 0:
        ms = C.c_float()
        nbytes = args.copy_calibration_mib << 20
        if lib.synth_gpu_copy_calibration(nbytes, 5, C.byref(ms)):
            raise RuntimeError("copy calibration failed")
        print("copy calibration: %d bytes read + %d written per launch, %.3f ms, %.1f GB/s"
              % (nbytes, nbytes, ms.value, 2 * nbytes / ms.value / 1e6), file=sys.stderr, flush=True)
    host_dbs = None
    if args.db_page_elems > 0:
        host_dbs = page_db_to_host(p, lib, dbs, args.db_host_memory == "pinned")
        p.setup.db_page_elems = args.db_page_elems
        p.setup.db_device_pages = args.db_device_pages
    k = mk.MCKernel(p.setup)
    tuning = k.tune(p.sample_points(2048, seed=17), reps=2) if args.tune else None
    v0 = min(max(args.first_voxel, 0), p.ntet - 1)
    T = args.task if args.task > 0 else p.ntet - v0
    ntask = (p.ntet - v0 + T - 1) // T
    steps = []
    for i in range(args.warmup + args.steps):
        t = i % ntask
        conn = p.conn[v0 + t * T:v0 + (t + 1) * T]
        t0 = time.perf_counter()
        r = k.slave_work(p.verts, conn)
        dt = time.perf_counter() - t0
        if i >= args.warmup:
            steps.append((dt, r["stats"]))
        print("step %d: %.3f s, %d evals, %d MC steps, %.1f GB paged" % (
              i, dt, r["stats"]["evaluations"], r["stats"]["mc_steps"], r["stats"]["paged_bytes"] / 1e9),
              file=sys.stderr, flush=True)
    tot_t = sum(s[0] for s in steps)
    tot_e = sum(s[1]["evaluations"] for s in steps)
    dev_ms = sum(s[1]["device_ms"] for s in steps)
    roof = roofline(steps, args.config, args.traffic)

    curve = None
    if args.task_curve:
        sizes = [int(x) for x in args.task_curve.split(",") if x]
        curve = task_curve(k, p, sizes, v0, min(args.curve_voxels, p.ntet - v0))

    # plain one-point-at-a-time form on a small sample of the same workload
    plain = None
    if args.plain_voxels > 0:
        sel = unmasked_voxels(p)[: args.plain_voxels]
        conn = p.conn[sel]
        k.halton_step = 0
        t0 = time.perf_counter()
        rp = k.slave_work(p.verts, conn, plain=True)
        tp = time.perf_counter() - t0
        k.halton_step = 0
        t0 = time.perf_counter()
        rb = k.slave_work(p.verts, conn)
        tb = time.perf_counter() - t0
        plain = {"definition": "plain = mck_slave_work_plain, the reference's loop order: one voxel at a "
                               "time, its MC steps one after another (the points of one step evaluated "
                               "together); batched = mck_slave_work on the same voxels",
                 "voxels": int(len(sel)), "plain_evals_per_s": rp["stats"]["evaluations"] / tp,
                 "batched_same_sample_evals_per_s": rb["stats"]["evaluations"] / tb,
                 "batched_over_plain_same_sample": (rb["stats"]["evaluations"] / tb) / (rp["stats"]["evaluations"] / tp),
                 "identical_results": bool(np.array_equal(rp["values"], rb["values"]))}
        print("plain sample:", json.dumps(plain), file=sys.stderr, flush=True)
    k.close()

    value = tot_e / tot_t
    line = {
        "metric": METRIC, "value": value, "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * tot_t / len(steps), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 accumulation)",
        "data": "synthetic (seeded AxiSEM-style merged DB generated in HBM)",
        "config": {"workload": "%s: %d-quad DB, nt=%d, %d receivers, %d filters, %d kernels "
                               "(CC/AM x vp/vs), %d tets, %d pts/step, max_iter %d, rel tol %g, "
                               "STF deconvolution %s"
                               % (args.config, p.mesh.nelem, p.nt, p.nrec, len(p.filters), p.nkernel,
                                  p.ntet, p.npoints_per_step, p.max_iter, p.allowed_relative_error,
                                  "on" if p.deconv_stf else "off"),
                   "task_voxels": T, "first_voxel": v0, "parallelism": "none (1 GPU)",
                   "evaluations": int(tot_e), "device_ms": dev_ms},
        "roofline": roof, "cpu_baseline": cpu, "plain_form": plain,
    }
    if curve is not None:
        line["task_curve"] = curve
    if tuning is not None:
        line["tuning"] = tuning
    if sample is not None and host_dbs is None:
        line["self_check"] = self_check(p, sample)
        print("self-check:", json.dumps(line["self_check"]), file=sys.stderr, flush=True)
    if host_dbs is not None:
        paged = sum(s[1]["paged_bytes"] for s in steps)
        line["config"]["databases"] = ("paged from %s host memory, %d elements per page, "
                                       "%d device page buffers per database"
                                       % (args.db_host_memory, args.db_page_elems,
                                          max(2, args.db_device_pages)))
        line["paging"] = {"bytes": paged, "host_to_device_GBs": paged / tot_t / 1e9}
        line["data"] = "synthetic (seeded AxiSEM-style merged DB generated in HBM, moved to host memory)"
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
