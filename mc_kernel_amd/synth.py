"""Seeded synthetic workloads: AxiSEM-style merged databases, receivers, kernels and
tetrahedral inversion meshes, assembled into an `MckSetup` (include/mck.h).

The workloads follow the benchmark configurations of this path:
  c0    2,000-quad DB, nt=1024, dt=1 s, 1 explosive source, 1 receiver at 30 deg,
        2 CC vp kernels (Gabor 40 s / 30 s, 60 s windows), 500 mantle tetrahedra,
        Halton points, 32 points per MC step, up to 8 steps.
  c1    20,000-quad DB, nt=2048, 10 receivers at 10..100 deg, Gabor 20/30/40/60 s,
        CC and AM for vp and vs (160 kernels), 50,000 tetrahedra, relative MC
        tolerance 1e-3, 16 points per step, MAXIMUM_ITERATIONS 64 (<= 1,024 points).
  c2    c1's database shape with 50 receivers at 10..108 deg, Gabor 10/20/30/40/60 s and a
        Butterworth band pass 80-8 s, CC and AM for vp and vs (1200 kernels), 200,000
        tetrahedra, 16 points per step, up to 256 steps.
  tiny  a few-hundred-quad DB with nt=256, dt=4 s for CPU tests.
"""
import ctypes as C
import math
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_SYNTH = None


def synth_lib():
    global _SYNTH
    if _SYNTH is None:
        path = os.path.join(_HERE, "libmck_synth.so")
        if not os.path.exists(path):
            raise RuntimeError("libmck_synth.so is not built; run __graft_entry__.build()")
        lib = C.CDLL(path)
        lib.synth_mesh_create.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]
        lib.synth_mesh_create.restype = C.c_void_p
        lib.synth_mesh_sizes.argtypes = [C.c_void_p, abi.i32p, abi.i32p]
        lib.synth_mesh_get.argtypes = [C.c_void_p, abi.f32p, abi.f32p, abi.f32p, abi.f32p,
                                       abi.i32p, abi.i32p, abi.i32p, abi.i32p]
        lib.synth_mesh_free.argtypes = [C.c_void_p]
        lib.synth_db_fill.argtypes = [C.c_int32, abi.i32p, C.c_int32, C.c_int32, C.c_double,
                                      C.c_uint64, abi.f32p, C.c_int32]
        lib.synth_model_table.argtypes = [C.c_int32, abi.f32p, abi.f32p, C.c_double, C.c_float,
                                          abi.f32p]
        lib.synth_model_table.restype = C.c_int32
        _SYNTH = lib
    return _SYNTH


R_EARTH = 6371.0e3
R_CMB = 3480.0e3
R_CENTRAL = 1200.0e3

# approximate P travel times (s) against epicentral distance (deg)
_TP_DEG = np.array([0, 10, 20, 30, 40, 50, 60, 70, 80, 90, 100, 120, 140, 160, 180.0])
_TP_SEC = np.array([0, 145, 280, 370, 460, 540, 610, 675, 735, 780, 820, 900, 980, 1060, 1140.0])


def p_time(delta_deg):
    return float(np.interp(delta_deg, _TP_DEG, _TP_SEC))


def mesh_params(nelem_target):
    """Quarter divisions n, ring layers and shell layers giving ~nelem_target quads with
    near-unit aspect ratio."""
    n = max(3, int(round(math.sqrt(nelem_target / 10.5))))
    mring = 2
    nshell = max(4, int(round((nelem_target - 2 * n * n - 4 * n * mring) / (4 * n))))
    return n, mring, nshell


class Mesh:
    def __init__(self, nelem_target, R=R_EARTH, rc=R_CENTRAL):
        lib = synth_lib()
        n, mring, nshell = mesh_params(nelem_target)
        h = lib.synth_mesh_create(R, rc, n, mring, nshell)
        ne, npnt = C.c_int32(), C.c_int32()
        lib.synth_mesh_sizes(h, C.byref(ne), C.byref(npnt))
        self.nelem, self.npoints = ne.value, npnt.value
        self.mesh_s = np.zeros(self.npoints, np.float32)
        self.mesh_z = np.zeros(self.npoints, np.float32)
        self.mp_s = np.zeros(self.nelem, np.float32)
        self.mp_z = np.zeros(self.nelem, np.float32)
        self.fem = np.zeros((self.nelem, 4), np.int32)
        self.sem = np.zeros((self.nelem, 5, 5), np.int32)
        self.eltype = np.zeros(self.nelem, np.int32)
        self.axis = np.zeros(self.nelem, np.int32)
        P = abi.ptr
        lib.synth_mesh_get(h, P(self.mesh_s, C.c_float), P(self.mesh_z, C.c_float),
                           P(self.mp_s, C.c_float), P(self.mp_z, C.c_float),
                           P(self.fem, C.c_int32), P(self.sem, C.c_int32),
                           P(self.eltype, C.c_int32), P(self.axis, C.c_int32))
        lib.synth_mesh_free(h)
        self.R = R


def fill_db(mesh, nvar, nt, dt, seed, nthreads=None):
    """merged_snapshots [nelem][nvar][5][5][nt] float32 (CPU generator)."""
    lib = synth_lib()
    out = np.empty((mesh.nelem, nvar, 5, 5, nt), np.float32)
    nth = nthreads or min(16, os.cpu_count() or 1)
    lib.synth_db_fill(mesh.nelem, abi.ptr(mesh.sem, C.c_int32), nvar, nt, dt, seed,
                      abi.ptr(out, C.c_float), nth)
    return out


def synth_gpu_lib():
    """The device-side database generator (libmck_synth_gpu.so): benchmark and test input,
    bit-identical to fill_db."""
    from . import bind_hip_runtime
    path = os.path.join(_HERE, "libmck_synth_gpu.so")
    if not os.path.exists(path):
        raise RuntimeError("libmck_synth_gpu.so is not built; run __graft_entry__.build()")
    bind_hip_runtime()
    lib = C.CDLL(path)
    lib.synth_gpu_db.argtypes = [C.c_int, abi.i32p, C.c_int, C.c_int, C.c_double, C.c_uint64,
                                 C.POINTER(C.c_void_p)]
    lib.synth_gpu_download.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    lib.synth_gpu_free.argtypes = [C.c_void_p]
    lib.synth_gpu_host_alloc.argtypes = [C.c_size_t, C.POINTER(C.c_void_p)]
    lib.synth_gpu_copy_calibration.argtypes = [C.c_size_t, C.c_int, C.POINTER(C.c_float)]
    return lib


class DeviceDatabases:
    """Both merged databases of a Problem generated in device memory (fwd 10 variables,
    bwd 5, the seeds of fill_db); points the Problem's setup at them.  free() releases
    them; download(which) copies one to a host array."""

    def __init__(self, p):
        self.lib = synth_gpu_lib()
        self.p = p
        m = p.mesh
        self.ptrs = []
        for nvar, seed in ((10, p.seed * 1000 + 1), (5, p.seed * 1000 + 2)):
            d = C.c_void_p()
            rc = self.lib.synth_gpu_db(m.nelem, abi.ptr(m.sem, C.c_int32), nvar, p.nt, p.dt, seed, C.byref(d))
            if rc:
                raise RuntimeError("device database generation failed (%d)" % rc)
            self.ptrs.append(d)
        self.attach()

    def attach(self):
        self.p.setup.fwd.snapshots = C.cast(self.ptrs[0], C.POINTER(C.c_float))
        self.p.setup.bwd.snapshots = C.cast(self.ptrs[1], C.POINTER(C.c_float))

    def download(self, which):
        m = self.p.mesh
        nvar = 10 if which == 0 else 5
        out = np.empty((m.nelem, nvar, 5, 5, self.p.nt), np.float32)
        if self.lib.synth_gpu_download(self.ptrs[which], out.ctypes.data, out.size):
            raise RuntimeError("database download failed")
        return out

    def free(self):
        for d in self.ptrs:
            if d:
                self.lib.synth_gpu_free(d)
        self.ptrs = []


def model_table(mesh, dz=100.0):
    lib = synth_lib()
    nd = lib.synth_model_table(mesh.npoints, abi.ptr(mesh.mesh_s, C.c_float),
                               abi.ptr(mesh.mesh_z, C.c_float), mesh.R, dz, None)
    out = np.zeros((6, nd), np.float32)
    lib.synth_model_table(mesh.npoints, abi.ptr(mesh.mesh_s, C.c_float),
                          abi.ptr(mesh.mesh_z, C.c_float), mesh.R, dz,
                          abi.ptr(out, C.c_float))
    return out


def destination(lat, lon, delta, az):
    """Point at epicentral distance delta (deg) and azimuth az (deg) from (lat, lon)."""
    la, lo, d, a = map(math.radians, (lat, lon, delta, az))
    la2 = math.asin(math.sin(la) * math.cos(d) + math.cos(la) * math.sin(d) * math.cos(a))
    lo2 = lo + math.atan2(math.sin(a) * math.sin(d) * math.cos(la),
                          math.cos(d) - math.sin(la) * math.sin(la2))
    return math.degrees(la2), (math.degrees(lo2) + 540.0) % 360.0 - 180.0


def epicentral_distance(lat1, lon1, lat2, lon2):
    """Great-circle distance (deg) between two points given in degrees."""
    a1, o1, a2, o2 = map(math.radians, (lat1, lon1, lat2, lon2))
    c = math.sin(a1) * math.sin(a2) + math.cos(a1) * math.cos(a2) * math.cos(o2 - o1)
    return math.degrees(math.acos(max(-1.0, min(1.0, c))))


def morton3(q):
    """Morton codes of integer coordinates q[n][3] (21 bits each)."""
    def spread(v):
        v = v.astype(np.uint64) & np.uint64(0x1fffff)
        v = (v | (v << np.uint64(32))) & np.uint64(0x1f00000000ffff)
        v = (v | (v << np.uint64(16))) & np.uint64(0x1f0000ff0000ff)
        v = (v | (v << np.uint64(8))) & np.uint64(0x100f00f00f00f00f)
        v = (v | (v << np.uint64(4))) & np.uint64(0x10c30c30c30c30c3)
        v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
        return v
    return spread(q[:, 0]) | (spread(q[:, 1]) << np.uint64(1)) | (spread(q[:, 2]) << np.uint64(2))


def random_tets(ntet, edge, seed, rmin=R_CMB, rmax=R_EARTH, accept=None, accept_frac=0.0):
    """Regular tetrahedra of the given edge length (m) with random orientation, centres
    uniform in the shell [rmin + edge, rmax - edge], in Morton order of their centres.
    A fraction accept_frac of the centres is drawn from the region where accept(c) holds.
    Returns vertices [4*ntet][3] and 1-based connectivity [ntet][4]."""
    rng = np.random.default_rng(seed)
    lo, hi = (rmin + edge) ** 3, (rmax - edge) ** 3

    def draw(n):
        r = np.cbrt(rng.uniform(lo, hi, n))
        v = rng.normal(size=(n, 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        return v * r[:, None]

    nacc = int(round(ntet * accept_frac)) if accept is not None else 0
    parts = [draw(ntet - nacc)]
    got = 0
    while got < nacc:
        c = draw(max(64, 8 * nacc))
        c = c[accept(c)][: nacc - got]
        parts.append(c)
        got += len(c)
    centres = np.concatenate(parts)
    base = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float)
    base *= edge / (2.0 * math.sqrt(2.0))
    # random rotations from normalised quaternions
    q = rng.normal(size=(ntet, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    w, x, y, z = q.T
    rot = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)
    verts = centres[:, None, :] + np.einsum("nij,kj->nki", rot, base)
    qc = ((centres / rmax + 1.0) * 0.5 * (2 ** 21 - 1)).astype(np.int64)
    order = np.argsort(morton3(qc), kind="stable")
    verts = np.ascontiguousarray(verts[order].reshape(-1, 3))
    conn = np.arange(1, 4 * ntet + 1, dtype=np.int32).reshape(ntet, 4)
    return verts, conn


def geo_xyz(lat, lon, r):
    la, lo = math.radians(lat), math.radians(lon)
    return np.array([math.cos(la) * math.cos(lo), math.cos(la) * math.sin(lo), math.sin(la)]) * r


def wavelet(t, t0, period):
    x = t - t0
    return np.exp(-(x / period) ** 2) * np.cos(2 * np.pi * x / period)


class Problem:
    """All inputs of one workload plus the MckSetup that points into them."""

    def __init__(self, name, nelem_target, nt, dt, receivers_deg, filters, kernels, ntet,
                 tet_edge, npoints_per_step, max_iter, allowed_error=1e-4,
                 allowed_relative_error=1e-2, component="Z", mij=None, seed=1,
                 mesh_seed=2, src=(20.0, 30.0, 100.0), azimuth=135.0, deconv_stf=1,
                 int_scheme="parseval", db_threads=None, fill=True, unmasked_frac=0.0,
                 receivers_latlon=None, s_windows=True, window_min_start=0.0):
        self.name = name
        self.seed = seed
        self.mesh = Mesh(nelem_target)
        self.nt, self.dt = nt, dt
        m = self.mesh
        self.fwd_u = fill_db(m, 10, nt, dt, seed * 1000 + 1, db_threads) if fill else None
        self.bwd_u = fill_db(m, 5, nt, dt, seed * 1000 + 2, db_threads) if fill else None
        self.model = model_table(m)
        self.src_lat, self.src_lon, self.src_depth = src
        if mij is None:
            mij = [1.0e20, 1.0e20, 1.0e20, 0.0, 0.0, 0.0]   # explosion, N m
        self.mij = np.array(mij, np.float64)
        lats, lons = [], []
        if receivers_latlon is not None:
            # receivers given by their coordinates: distances from the source
            for la, lo in receivers_latlon:
                lats.append(float(la))
                lons.append(float(lo))
            receivers_deg = [epicentral_distance(self.src_lat, self.src_lon, la, lo) for la, lo in receivers_latlon]
        else:
            for d in receivers_deg:
                la, lo = destination(self.src_lat, self.src_lon, d, azimuth)
                lats.append(la)
                lons.append(lo)
        self.receivers_deg = list(receivers_deg)
        self.rec_lat = np.array(lats, np.float64)
        self.rec_lon = np.array(lons, np.float64)
        self.nrec = len(receivers_deg)
        self.component = component
        t = np.arange(nt) * dt
        seis = np.zeros((self.nrec, nt))
        for r, d in enumerate(receivers_deg):
            tp = p_time(d)
            for per in (20.0, 30.0, 40.0, 60.0):
                seis[r] += wavelet(t, tp + 20.0, per) + 0.7 * wavelet(t, 1.8 * tp + 20.0, per)
        self.seis = np.ascontiguousarray(seis * 1e-6)
        # the database's source time function (stf_dump): a Gaussian one sample wide, so
        # that its deconvolution (add_stfs) keeps the filters vanishing at the Nyquist
        # frequency, as the reference requires
        self.stf = np.exp(-((t - 10.0) / dt) ** 2)
        self.stf /= self.stf.sum() * dt
        # filters: list of (class, (f1, f2, f3, f4)); kernels: list of dicts
        self.filters = (abi.MckFilter * len(filters))()
        for i, (cls, fr) in enumerate(filters):
            self.filters[i].filterclass = abi.FILTERS[cls]
            for j in range(4):
                self.filters[i].frequencies[j] = fr[j]
        nk = len(kernels)
        self.kernel_receiver = np.array([k["rec"] for k in kernels], np.int32)
        self.kernel_filter = np.array([k["filter"] for k in kernels], np.int32)
        self.kernel_misfit = np.array([abi.MISFITS[k["misfit"]] for k in kernels], np.int32)
        self.kernel_param = np.array([abi.PARAMS[k["param"]] for k in kernels], np.int32)
        # 60 s windows 10 s after the P time (vp-like parameters) or after 1.8 x the P time
        # (the others; s_windows False: the P window for every parameter), starting no
        # earlier than window_min_start
        win = []
        for k in kernels:
            tp = p_time(receivers_deg[k["rec"]])
            ts = tp if (k["param"] in ("vp", "lam", "vph", "vpv") or not s_windows) else 1.8 * tp
            t0 = max(ts, window_min_start - 10.0) + 10.0
            win.append((t0, t0 + 60.0))
        self.kernel_window = np.array(win, np.float64)
        self.nkernel = nk
        src_xyz = geo_xyz(self.src_lat, self.src_lon, R_EARTH - self.src_depth * 1e3)
        rec_xyz = geo_xyz(self.rec_lat[0], self.rec_lon[0], R_EARTH)
        vmax = float(self.model[0].max())
        tlast = float(self.kernel_window[self.kernel_receiver == 0][:, 1].max())

        def inside(c):   # not travel-time masked for the first receiver (slave.f90:928-950)
            d = np.linalg.norm(c - src_xyz, axis=1) + np.linalg.norm(c - rec_xyz, axis=1)
            return d / vmax < 0.9 * tlast

        self.verts, self.conn = random_tets(ntet, tet_edge, mesh_seed, accept=inside,
                                            accept_frac=unmasked_frac)
        self.ntet = ntet
        self.npoints_per_step = npoints_per_step
        self.max_iter = max_iter
        self.allowed_error = allowed_error
        self.allowed_relative_error = allowed_relative_error
        self.deconv_stf = deconv_stf
        self.int_scheme = abi.INT_SCHEMES[int_scheme]
        self.setup = self._make_setup()

    def _db(self, u, nvar):
        m = self.mesh
        d = abi.MckDb()
        d.npol, d.nelem, d.npoints = 4, m.nelem, m.npoints
        d.mesh_s = abi.ptr(m.mesh_s, C.c_float)
        d.mesh_z = abi.ptr(m.mesh_z, C.c_float)
        d.mp_s = abi.ptr(m.mp_s, C.c_float)
        d.mp_z = abi.ptr(m.mp_z, C.c_float)
        d.fem_mesh = abi.ptr(m.fem, C.c_int32)
        d.eltype = abi.ptr(m.eltype, C.c_int32)
        d.axis = abi.ptr(m.axis, C.c_int32)
        d.nvar = nvar
        d.snapshots = abi.ptr(u, C.c_float)
        d.amplitude = 1.0e20
        return d

    def _make_setup(self):
        s = abi.MckSetup()
        s.fwd = self._db(self.fwd_u, 10)
        s.bwd = self._db(self.bwd_u, 5)
        s.nt, s.dt, s.planet_radius = self.nt, self.dt, self.mesh.R
        s.timeshift_fwd, s.timeshift_bwd = 15.0, 15.0
        s.stf_fwd = abi.ptr(self.stf, C.c_double)
        s.src_lat, s.src_lon, s.src_depth_km = self.src_lat, self.src_lon, self.src_depth
        for i in range(6):
            s.src_mij[i] = self.mij[i]
        s.stf_src = None
        s.stf_src_shift = 0.0
        s.nrec = self.nrec
        s.rec_lat = abi.ptr(self.rec_lat, C.c_double)
        s.rec_lon = abi.ptr(self.rec_lon, C.c_double)
        s.component = self.component.encode()
        s.seis = abi.ptr(self.seis, C.c_double)
        s.nfilter = len(self.filters)
        s.filters = self.filters
        s.nkernel = self.nkernel
        s.kernel_receiver = abi.ptr(self.kernel_receiver, C.c_int32)
        s.kernel_filter = abi.ptr(self.kernel_filter, C.c_int32)
        s.kernel_misfit = abi.ptr(self.kernel_misfit, C.c_int32)
        s.kernel_param = abi.ptr(self.kernel_param, C.c_int32)
        s.kernel_window = abi.ptr(self.kernel_window, C.c_double)
        s.model_ndepth = self.model.shape[1]
        s.model_dr = 100.0
        names = ["model_vp", "model_vs", "model_rho", "model_phi", "model_xi", "model_eta"]
        for i, nm in enumerate(names):
            setattr(s, nm, abi.ptr(self.model[i], C.c_float))
        s.vp_max = float(self.model[0].max())
        s.vs_max = float(self.model[1].max())
        s.npoints_per_step = self.npoints_per_step
        s.max_iter = self.max_iter
        s.allowed_error = self.allowed_error
        s.allowed_relative_error = self.allowed_relative_error
        s.int_over_volume = 1
        s.relative_kernel = 1
        s.quasirandom = 1
        s.int_scheme = self.int_scheme
        s.deconv_stf = self.deconv_stf
        s.damp_radius = -1.0
        s.mask_src_rec = 0
        return s

    def voxels(self, nvox, seed=7, dlat=4.0, dlon=4.0, dr=300e3):
        """Voxel inversion elements (blocks in radius, colatitude and longitude given by
        their 8 corners, random_voxel.f90): nvox blocks of dlat x dlon degrees and dr
        metres scattered between the source and the first receiver; (verts, conn) with
        conn (nvox, 8) 1-based."""
        rng = np.random.default_rng(seed)
        verts, conn = [], []
        for i in range(nvox):
            f = rng.uniform(0.0, 1.0)
            lat = self.src_lat + f * (self.rec_lat[0] - self.src_lat) + rng.uniform(-5, 5)
            lon = self.src_lon + f * (self.rec_lon[0] - self.src_lon) + rng.uniform(-5, 5)
            r0 = rng.uniform(R_EARTH - 2800e3, R_EARTH - 100e3 - dr)
            for (a, b, c) in np.ndindex(2, 2, 2):
                verts.append(geo_xyz(lat + a * dlat, lon + b * dlon, r0 + c * dr))
            conn.append(np.arange(8 * i + 1, 8 * i + 9))
        return np.ascontiguousarray(np.array(verts, np.float64)), np.ascontiguousarray(np.array(conn, np.int32))

    def sample_points(self, npts, seed=3):
        """Uniform points inside the first tetrahedra (for point-level parity tests)."""
        rng = np.random.default_rng(seed)
        tet = rng.integers(0, self.ntet, npts)
        b = rng.dirichlet(np.ones(4), npts)
        v = self.verts.reshape(self.ntet, 4, 3)[tet]
        return np.ascontiguousarray(np.einsum("nk,nkj->nj", b, v))


def _kernels(nrec, nfilter, misfits, params):
    out = []
    for r in range(nrec):
        for f in range(nfilter):
            for mf in misfits:
                for p in params:
                    out.append({"rec": r, "filter": f, "misfit": mf, "param": p})
    return out


def make_problem(name, **over):
    """Named workloads (see module docstring); keyword arguments override fields."""
    if name == "c0":
        kw = dict(nelem_target=2000, nt=1024, dt=1.0, receivers_deg=[30.0],
                  filters=[("Gabor", (40.0, 0.5, 0.0, 0.0)), ("Gabor", (30.0, 0.5, 0.0, 0.0))],
                  kernels=_kernels(1, 2, ["CC"], ["vp"]), ntet=500, tet_edge=400e3,
                  npoints_per_step=32, max_iter=8)
    elif name == "c1":
        kw = dict(nelem_target=20000, nt=2048, dt=1.0,
                  receivers_deg=[10.0 * (i + 1) for i in range(10)],
                  filters=[("Gabor", (p, 0.5, 0.0, 0.0)) for p in (20.0, 30.0, 40.0, 60.0)],
                  kernels=_kernels(10, 4, ["CC", "AM"], ["vp", "vs"]), ntet=50000,
                  tet_edge=300e3, npoints_per_step=16, max_iter=64,
                  allowed_error=1e-30, allowed_relative_error=1e-3)
    elif name == "c2":
        # 50 receivers x 6 filters x CC/AM x vp/vs = 1200 kernels (the reference's rho
        # kernels of this configuration are rejected, as the reference aborts on them).
        # Receivers from 10 deg on, as in c1: nearer ones would put their windows before
        # the synthetic wavefields' first arrivals (100 s), where a kernel is a small
        # remainder of the fwd x bwd energy and single-precision spectra lose digits
        kw = dict(nelem_target=20000, nt=2048, dt=1.0,
                  receivers_deg=[10.0 + 2.0 * i for i in range(50)],
                  filters=[("Gabor", (p, 0.5, 0.0, 0.0)) for p in (10.0, 20.0, 30.0, 40.0, 60.0)]
                          + [("Butterw_BP", (80.0, 8.0, 4.0, 0.0))],
                  kernels=_kernels(50, 6, ["CC", "AM"], ["vp", "vs"]), ntet=200000,
                  tet_edge=300e3, npoints_per_step=16, max_iter=256,
                  allowed_error=1e-30, allowed_relative_error=1e-3)
    elif name == "c4":
        # the reference's Gibraltar benchmark event (CMTSOLUTION: 35.74 N, 3.72 W, 12 km,
        # moment tensor in dyne cm -> N m) and its 200 stations' coordinates (3.8..174.3
        # deg); 4 Gabor filters x CC x vp/vs = 1600 kernels, both parameters on the P
        # window as in the reference's benchmark kernel file, windows starting no earlier
        # than 220 s (where the synthetic wavefields' convolution has energy: arrivals
        # from 100 s in both); 100,000-quad databases at nt 4096 (614 GB: paged, see
        # PagedSyntheticDatabases); the inversion mesh is the 30-layer 1.25 deg voxel
        # grid (voxel_grid), of which a task takes a slice
        kw = dict(nelem_target=100000, nt=4096, dt=1.0, receivers_deg=None,
                  receivers_latlon=gibraltar_receivers(), src=(35.74, -3.72, 12.0),
                  mij=[x * 1e-7 for x in (-3.85e24, -2.76e25, 3.145e25, 1.188e25, 1.345e25, -1.649e25)],
                  filters=[("Gabor", (p, 0.5, 0.0, 0.0)) for p in (20.0, 30.0, 40.0, 60.0)],
                  kernels=_kernels(200, 4, ["CC"], ["vp", "vs"]), ntet=1, tet_edge=300e3,
                  npoints_per_step=16, max_iter=64, allowed_error=1e-30, allowed_relative_error=1e-3,
                  fill=False, s_windows=False, window_min_start=220.0)
    elif name == "tiny":
        kw = dict(nelem_target=450, nt=256, dt=4.0, receivers_deg=[30.0, 60.0],
                  filters=[("Gabor", (40.0, 0.5, 0.0, 0.0)), ("Gabor", (60.0, 0.5, 0.0, 0.0))],
                  kernels=_kernels(2, 2, ["CC", "AM"], ["vp", "vs"]), ntet=24,
                  tet_edge=500e3, npoints_per_step=8, max_iter=6, unmasked_frac=0.75)
    else:
        raise ValueError(name)
    kw.update(over)
    return Problem(name, **kw)


# --------------------------------------------------------------------------------------
# c4: receivers, voxel grid, databases paged from a sparse host mapping
# --------------------------------------------------------------------------------------

def gibraltar_receivers():
    """(lat, lon) in degrees of the reference's 200-station Gibraltar benchmark
    (data/gibraltar_receivers_200.txt, extracted by data/extract_receivers.py)."""
    d = np.loadtxt(os.path.join(_HERE, "data", "gibraltar_receivers_200.txt"))
    return [tuple(r) for r in d]


def voxel_grid(nlayers=30, dlat=1.25, rmin=R_CMB, rmax=R_EARTH):
    """The global equal-area voxel grid: nlayers shells of equal thickness between rmin
    and rmax, colatitude bands of dlat degrees, band i cut into round(360 / dlat x
    sin(mid colatitude)) equal longitude blocks from -180 deg.  Voxels are blocks in
    radius, colatitude and longitude given by their 8 corners (the inversion mesh's TYPE=VOX
    elements, random_voxel.f90): bottom face (colat0, lon0), (colat0, lon1), (colat1, lon1),
    (colat1, lon0), then the top face.  Order: layer (top down), band (north to south),
    longitude.  Returns corners [8 nvox][3] in m, 1-based connectivity [nvox][8] and the
    (layer, band) of every voxel."""
    radii = np.linspace(rmax, rmin, nlayers + 1)
    nb = int(round(180.0 / dlat))
    th_edges = np.radians(np.arange(nb + 1) * dlat)
    blocks = []
    for i in range(nb):
        nl = max(1, int(round(360.0 / dlat * math.sin(0.5 * (th_edges[i] + th_edges[i + 1])))))
        ph = np.radians(np.linspace(-180.0, 180.0, nl + 1))
        blocks.append((th_edges[i], th_edges[i + 1], ph))
    corners, lay, band = [], [], []
    for k in range(nlayers):
        r_top, r_bot = radii[k], radii[k + 1]
        for i, (t0, t1, ph) in enumerate(blocks):
            n = len(ph) - 1
            p0, p1 = ph[:-1], ph[1:]
            th = np.array([t0, t0, t1, t1])
            c = np.empty((n, 8, 3))
            for j, r in enumerate((r_bot, r_top)):
                for q, (tq, pq) in enumerate(zip(th, (p0, p1, p1, p0))):
                    c[:, 4 * j + q, 0] = r * np.sin(tq) * np.cos(pq)
                    c[:, 4 * j + q, 1] = r * np.sin(tq) * np.sin(pq)
                    c[:, 4 * j + q, 2] = r * np.cos(tq)
            corners.append(c)
            lay.append(np.full(n, k, np.int32))
            band.append(np.full(n, i, np.int32))
    c = np.concatenate(corners)
    nv = len(c)
    verts = np.ascontiguousarray(c.reshape(-1, 3))
    conn = np.arange(1, 8 * nv + 1, dtype=np.int32).reshape(nv, 8)
    return verts, conn, np.concatenate(lay), np.concatenate(band)


def write_vox_mesh(path, verts, conn):
    """The inversion mesh in the abaqus text format of the reference's voxel meshes
    (Meshes/vox_15l_5deg.dat: *NODE in km, *ELEMENT, TYPE=VOX with 8 node ids)."""
    with open(path, "w") as f:
        f.write("*HEADING\nvoxel grid\n**\n*PART, NAME=Part-Default\n**\n*NODE, NSET=ALLNODES\n")
        for i, (x, y, z) in enumerate(np.asarray(verts) / 1e3):
            f.write("%d, %.9e, %.9e, %.9e\n" % (i + 1, x, y, z))
        f.write("**\n*ELEMENT, TYPE=VOX, ELSET=EB2\n")
        for i, row in enumerate(np.asarray(conn)):
            f.write("%d, %s\n" % (i + 1, ", ".join(str(int(v)) for v in row)))
        f.write("**\n*END PART\n")


def read_vox_mesh(path):
    """(verts [n][3] in m, conn [nvox][8] 1-based) of an abaqus TYPE=VOX mesh file."""
    import gzip
    op = gzip.open if path.endswith(".gz") else open
    verts, conn, mode = [], [], None
    with op(path, "rt") as f:
        for ln in f:
            if ln.startswith("*"):
                u = ln.upper()
                mode = "n" if u.startswith("*NODE") else ("e" if u.startswith("*ELEMENT") and "VOX" in u else None)
                continue
            parts = [x for x in ln.replace(",", " ").split()]
            if mode == "n" and len(parts) == 4:
                verts.append([float(x) for x in parts[1:]])
            elif mode == "e" and len(parts) == 9:
                conn.append([int(x) for x in parts[1:]])
    return np.array(verts) * 1e3, np.array(conn, np.int32)


def voxel_slice(verts, conn, layer_of, layer, lat, lon, n):
    """The n voxels of one layer whose corner centroids lie nearest to (lat, lon), in
    grid order."""
    sel = np.nonzero(layer_of == layer)[0]
    c = np.asarray(verts).reshape(-1, 8, 3)[sel].mean(axis=1)
    c /= np.linalg.norm(c, axis=1)[:, None]
    p = geo_xyz(lat, lon, 1.0)
    order = np.argsort(-(c @ p), kind="stable")[:n]
    return np.sort(sel[order])


def _frame_angles(pole, pts):
    """angular distance (rad) of unit vectors pts [n][3] from the unit vector pole"""
    return np.arccos(np.clip(pts @ pole, -1.0, 1.0))


def touched_elements(mesh, poles, verts, conn, margin_deg=0.5):
    """Elements of the (s, z) mesh that points of the voxels (verts, conn) can reach in
    any of the frames whose poles (unit vectors: source, receivers) are given: elements
    whose radius range meets the voxels' and whose angle from the pole range meets the
    voxels' angular range from that pole (each voxel's range sampled on a 9 x 9 grid over
    its colatitude-longitude face, widened by the sampling step and margin_deg).
    Returns a bool mask [nelem]."""
    c = np.asarray(verts).reshape(-1, 8, 3)
    r = np.linalg.norm(c, axis=2)
    rmin, rmax = r.min(), r.max()
    th = np.arccos(np.clip(c[..., 2] / r, -1, 1))
    ph = np.arctan2(c[..., 1], c[..., 0])
    t0, t1, p0, p1 = th.min(1), th.max(1), ph.min(1), ph.max(1)
    g = np.linspace(0.0, 1.0, 9)
    tt = t0[:, None, None] + (t1 - t0)[:, None, None] * g[None, :, None]
    pp = p0[:, None, None] + (p1 - p0)[:, None, None] * g[None, None, :]
    pts = np.stack([np.sin(tt) * np.cos(pp), np.sin(tt) * np.sin(pp), np.cos(tt) * np.ones_like(pp)], -1).reshape(-1, 3)
    step = max(float((t1 - t0).max()), float(((p1 - p0) * np.sin(0.5 * (t0 + t1))).max())) / 8.0
    widen = step + math.radians(margin_deg)
    # element extents in the frame's polar coordinates, from its GLL points
    s_ = mesh.mesh_s.astype(np.float64)[mesh.sem.reshape(mesh.nelem, 25)]
    z_ = mesh.mesh_z.astype(np.float64)[mesh.sem.reshape(mesh.nelem, 25)]
    er = np.hypot(s_, z_)
    ea = np.arctan2(s_, z_)
    er0, er1, ea0, ea1 = er.min(1), er.max(1), ea.min(1), ea.max(1)
    pad_r = 0.25 * (er1 - er0)
    pad_a = 0.25 * (ea1 - ea0)
    in_r = (er1 + pad_r >= rmin - 50e3) & (er0 - pad_r <= rmax + 50e3)
    mask = np.zeros(mesh.nelem, bool)
    for pole in np.atleast_2d(poles):
        pole = pole / np.linalg.norm(pole)
        a = _frame_angles(pole, pts)
        a0, a1 = a.min() - widen, a.max() + widen
        mask |= in_r & (ea1 + pad_a >= a0) & (ea0 - pad_a <= a1)
    return mask


def _libc():
    libc = C.CDLL(None, use_errno=True)
    libc.mmap.restype = C.c_void_p
    libc.mmap.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_long]
    libc.munmap.argtypes = [C.c_void_p, C.c_size_t]
    return libc


def sparse_mapping(nbytes):
    """Address of a private anonymous mapping of nbytes that reserves nothing: memory is
    taken only where written (pages never written read as zeros)."""
    addr = _libc().mmap(None, nbytes, 0x3, 0x02 | 0x20 | 0x4000, -1, 0)
    if addr is None or addr == C.c_void_p(-1).value:
        raise OSError(C.get_errno(), "sparse mapping of %d bytes failed" % nbytes)
    return addr


def release_mapping(addr, nbytes):
    _libc().munmap(addr, nbytes)


class PagedSyntheticDatabases:
    """Both merged databases of a Problem too large for device or host memory: the whole
    address range of each is reserved as a sparse host mapping (memory is taken only where
    written), and only the element pages of page_elems elements that hold an element of
    masks[which] are generated (the device generator, bit-identical to fill_db) and
    written into it.  The setup's snapshots point at the mappings and page from them
    (db_page_elems = page_elems).  generated_bytes: the bytes written."""

    def __init__(self, p, masks, page_elems, chunk_elems=2048):
        self.lib = synth_gpu_lib()
        self.p = p
        m = p.mesh
        self.maps, self.generated_bytes, self.pages = [], 0, []
        for which, (nvar, seed) in enumerate(((10, p.seed * 1000 + 1), (5, p.seed * 1000 + 2))):
            elem_floats = nvar * 25 * p.nt
            size = m.nelem * elem_floats * 4
            addr = sparse_mapping(size)
            self.maps.append((addr, size))
            pages = np.unique(np.nonzero(masks[which])[0] // page_elems)
            self.pages.append(pages)
            # runs of consecutive pages, generated chunk_elems elements at a time
            for pg in pages:
                e0, e1 = int(pg) * page_elems, min(m.nelem, (int(pg) + 1) * page_elems)
                for a in range(e0, e1, chunk_elems):
                    b = min(e1, a + chunk_elems)
                    d = C.c_void_p()
                    sem = np.ascontiguousarray(m.sem[a:b])
                    rc = self.lib.synth_gpu_db(b - a, abi.ptr(sem, C.c_int32), nvar, p.nt, p.dt, seed, C.byref(d))
                    if rc:
                        raise RuntimeError("device database generation failed (%d)" % rc)
                    try:
                        if self.lib.synth_gpu_download(d, addr + a * elem_floats * 4, (b - a) * elem_floats):
                            raise RuntimeError("database download failed")
                    finally:
                        self.lib.synth_gpu_free(d)
                    self.generated_bytes += (b - a) * elem_floats * 4
        p.setup.fwd.snapshots = C.cast(C.c_void_p(self.maps[0][0]), C.POINTER(C.c_float))
        p.setup.bwd.snapshots = C.cast(C.c_void_p(self.maps[1][0]), C.POINTER(C.c_float))
        p.setup.db_page_elems = page_elems

    def free(self):
        for addr, size in self.maps:
            release_mapping(addr, size)
        self.maps = []
