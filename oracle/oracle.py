"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the mc_kernel_amd package.
"""
import ctypes as C
import os

import numpy as np

from mc_kernel_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        from mc_kernel_amd import build
        # MCK_ORACLE_LIB: another build of the same sources (the sanitizer build)
        path = os.environ.get("MCK_ORACLE_LIB") or build.build_oracle(force=False)
        L = C.CDLL(path)
        f64p, f32p, i32p, i64p = abi.f64p, abi.f32p, abi.i32p, abi.i64p
        L.ora_gll_points.argtypes = [C.c_int, f64p]
        L.ora_glj_points.argtypes = [C.c_int, f64p]
        L.ora_lagrange_derivs.argtypes = [C.c_int, f64p, f64p]
        L.ora_halton.argtypes = [C.c_int, i64p, i32p, i64p, C.c_int64, C.c_int, f64p]
        L.ora_tet_points.argtypes = [f64p, C.c_int, f64p, f64p]
        L.ora_tet_volume.argtypes = [f64p]
        L.ora_tet_volume.restype = C.c_double
        L.ora_point_in_tet.argtypes = [f64p, f64p]
        L.ora_norm2.argtypes = [C.c_int, f64p]
        L.ora_norm2.restype = C.c_double
        L.ora_mapping.argtypes = [C.c_double, C.c_double, f64p, C.c_int, f64p]
        L.ora_inv_jacobian.argtypes = [C.c_double, C.c_double, f64p, C.c_int, f64p]
        L.ora_inv_mapping.argtypes = [C.c_double, C.c_double, f64p, C.c_int, f64p]
        L.ora_inside_element.argtypes = [C.c_double, C.c_double, f64p, C.c_int, C.c_double,
                                         f64p, f64p]
        L.ora_strain_pole.argtypes = [C.c_int, f64p, C.c_int, C.c_int, f64p, C.c_int, C.c_int, f64p]
        L.ora_straintrace_pole.argtypes = [C.c_int, f64p, C.c_int, C.c_int, f64p, C.c_int, C.c_int, f64p]
        L.ora_rotate_frame.argtypes = [C.c_int, f64p, C.c_double, C.c_double, f64p, f64p, f64p]
        L.ora_nextpow2.argtypes = [C.c_int]
        L.ora_rfft.argtypes = [C.c_int, C.c_double, f64p, f64p]
        L.ora_irfft.argtypes = [C.c_int, C.c_double, f64p, f64p]
        L.ora_taperandzeropad.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, f64p, f64p]
        L.ora_cumsum_trapezoidal.argtypes = [C.c_int, C.c_double, f64p, f64p]
        L.ora_filter_tf.argtypes = [C.POINTER(abi.MckFilter), C.c_int, C.c_double, f64p]
        L.ora_nearest.argtypes = [C.c_int, f32p, f32p, C.c_double, C.c_double, C.c_int, i32p, f32p]
        L.ora_create.argtypes = [C.POINTER(abi.MckSetup), C.c_char_p, C.c_int]
        L.ora_create.restype = C.c_void_p
        L.ora_destroy.argtypes = [C.c_void_p]
        L.ora_locate.argtypes = [C.c_void_p, C.c_int, C.c_int, f64p, i32p, f64p, f64p, i32p]
        L.ora_load_strain.argtypes = [C.c_void_p, C.c_int, C.c_int, f64p, f64p]
        L.ora_base_kernels.argtypes = [C.c_void_p, C.c_int, f64p, f64p]
        L.ora_slave_work.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, f64p, i32p, f64p,
                                     f64p, i32p, i64p, i64p, i64p, i64p]
        L.ora_splitmix64.argtypes = [C.c_uint64]
        L.ora_splitmix64.restype = C.c_uint64
        L.ora_prng.argtypes = [C.c_int64, C.c_int]
        L.ora_prng.restype = C.c_double
        L.ora_prng_points.argtypes = [C.c_int64, C.c_int, f64p]
        L.ora_vox_range.argtypes = [f64p, f64p]
        L.ora_vox_points.argtypes = [f64p, C.c_int, f64p, f64p]
        L.ora_vox_volume.argtypes = [f64p]
        L.ora_vox_volume.restype = C.c_double
        L.ora_vox_center.argtypes = [f64p, f64p]
        L.ora_point_in_vox.argtypes = [f64p, f64p]
        L.ora_tet_weights.argtypes = [f64p, C.c_int, f64p, f64p]
        L.ora_kernel_info.argtypes = [C.c_void_p, C.c_int, f64p, C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]
        L.ora_strain_nd.argtypes = [C.c_void_p]
        L.ora_set_lazy_fill.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.ora_integrate_parseval.argtypes = [C.c_int, C.c_double, f64p, f64p]
        L.ora_basekernel.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, f64p, f64p, f64p]
        L.ora_dampen_field.argtypes = [f64p, C.c_int, C.c_int, f64p, f64p, C.c_double]
        L.ora_add_stfs.argtypes = [C.c_int, C.c_double, C.c_int, f64p, C.c_double, f64p, C.c_double,
                                   f64p, f64p, C.c_char_p, C.c_int]
        L.ora_integrate_parseval.restype = C.c_double
        L.ora_physical.argtypes = [C.c_int, f64p] + [C.c_float] * 6 + [C.c_int]
        L.ora_physical.restype = C.c_double
        L.ora_param_needs.argtypes = [C.c_int, i32p]
        L.ora_azim_factor.argtypes = [C.c_double, f64p, C.c_int, C.c_int]
        L.ora_azim_factor.restype = C.c_double
        L.ora_cut_window.argtypes = [C.c_int, f64p, C.c_double, C.c_double, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]
        L.ora_rotate_voigt.argtypes = [C.c_int, f64p, C.c_int, C.c_double, C.c_double]
        L.ora_rotate_voigt_inv.argtypes = [C.c_int, f64p, C.c_double, C.c_double]
        L.ora_seismograms.argtypes = [C.c_void_p, f64p]
        L.ora_integrate_model.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, f64p, i32p,
                                          f64p, i32p, i64p, i64p]
        L.ora_het_xyz.argtypes = [C.c_int, f64p, f32p]
        L.ora_het_nearest.argtypes = [C.c_int, f32p, C.c_float, C.c_float, C.c_float]
        L.ora_integrate_hetero.argtypes = [C.c_int, f64p, f32p, C.c_int, C.c_int, C.c_int, C.c_int, f64p, i32p,
                                           f64p, i32p, i64p, i64p]
        L.ora_results_add.argtypes = [C.c_int] * 6 + [i32p, C.c_int, i32p, f64p, f64p, i32p, f64p, f64p,
                                                       C.c_int, f64p, f64p, f64p, i32p, f64p, i32p]
        L.ora_results_export.argtypes = [C.c_int, C.c_int, C.c_int, f64p, f64p, i32p, f32p]
        L.ora_results_export.restype = None
        L.ora_ad_transfer.argtypes = [C.c_void_p, f64p]
        L.ora_ad_forward.argtypes = [C.c_void_p, C.c_int, f64p, f64p, f64p, f64p]
        L.ora_ad_backward.argtypes = [C.c_void_p, C.c_int, f64p, f64p, f64p, f64p, f64p, f64p, f64p]
        L.ora_mc_run.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                 f64p, f64p, f64p, i32p]
        _LIB = L
    return _LIB


def P(a, t=C.c_double):
    return abi.ptr(a, t)


def gll_points(n):
    x = np.zeros(n + 1)
    lib().ora_gll_points(n, P(x))
    return x


def glj_points(n):
    x = np.zeros(n + 1)
    lib().ora_glj_points(n, P(x))
    return x


def lagrange_derivs(n, x):
    G = np.zeros((n + 1) * (n + 1))
    lib().ora_lagrange_derivs(n, P(np.ascontiguousarray(x, float)), P(G))
    return G.reshape(n + 1, n + 1).T   # G[i, j] = l_i'(x_j)


def halton(npoints, step=0, seed=(20, 20, 20), base=(2, 3, 5), leap=(2, 2, 2)):
    nd = len(base)
    r = np.zeros((npoints, nd))
    s = np.array(seed, np.int64)
    b = np.array(base, np.int32)
    l = np.array(leap, np.int64)
    lib().ora_halton(nd, P(s, C.c_int64), P(b, C.c_int32), P(l, C.c_int64), step, npoints, P(r))
    return r


def tet_points(v, coords):
    v = np.ascontiguousarray(v, float)
    coords = np.ascontiguousarray(coords, float)
    out = np.zeros_like(coords)
    lib().ora_tet_points(P(v), coords.shape[0], P(coords), P(out))
    return out


def tet_volume(v):
    return lib().ora_tet_volume(P(np.ascontiguousarray(v, float)))


def tet_weights(v, pts):
    """The four vertex basis functions of tetrahedron v[4][3] at pts[n][3]: (w[n][4], bad)"""
    v = np.ascontiguousarray(v, float)
    pts = np.ascontiguousarray(pts, float)
    w = np.zeros((pts.shape[0], 4))
    bad = lib().ora_tet_weights(P(v), pts.shape[0], P(pts), P(w))
    return w, bool(bad)


def prng_points(step, n):
    r = np.zeros((n, 3))
    lib().ora_prng_points(step, n, P(r))
    return r


def vox_range(v):
    g = np.zeros(6)
    lib().ora_vox_range(P(np.ascontiguousarray(v, float)), P(g))
    return g


def vox_points(v, coords):
    coords = np.ascontiguousarray(coords, float)
    out = np.zeros_like(coords)
    lib().ora_vox_points(P(np.ascontiguousarray(v, float)), coords.shape[0], P(coords), P(out))
    return out


def vox_volume(v):
    return lib().ora_vox_volume(P(np.ascontiguousarray(v, float)))


def vox_center(v):
    c = np.zeros(3)
    lib().ora_vox_center(P(np.ascontiguousarray(v, float)), P(c))
    return c


def point_in_vox(v, q):
    return bool(lib().ora_point_in_vox(P(np.ascontiguousarray(v, float)), P(np.ascontiguousarray(q, float))))


def point_in_tet(v, q):
    v = np.ascontiguousarray(v, float)
    q = np.ascontiguousarray(q, float)
    return bool(lib().ora_point_in_tet(P(v), P(q)))


def rfft(x, dt=1.0):
    x = np.ascontiguousarray(x, float)
    X = np.zeros((len(x) // 2 + 1) * 2)
    lib().ora_rfft(len(x), dt, P(x), P(X))
    return X[0::2] + 1j * X[1::2]


def irfft(X, n, df):
    Xa = np.zeros(2 * (n // 2 + 1))
    Xa[0::2] = X.real
    Xa[1::2] = X.imag
    x = np.zeros(n)
    lib().ora_irfft(n, df, P(Xa), P(x))
    return x


def taperandzeropad(x, nout, ntaper, end_only=False):
    x = np.ascontiguousarray(x, float)
    out = np.zeros(nout)
    lib().ora_taperandzeropad(len(x), nout, ntaper, int(end_only), P(x), P(out))
    return out


def cumsum_trapezoidal(x, dt):
    x = np.ascontiguousarray(x, float)
    out = np.zeros_like(x)
    lib().ora_cumsum_trapezoidal(len(x), dt, P(x), P(out))
    return out


def filter_tf(filterclass, freqs, nfreq, df):
    f = abi.MckFilter()
    f.filterclass = abi.FILTERS[filterclass]
    for i in range(4):
        f.frequencies[i] = freqs[i]
    tf = np.zeros(2 * nfreq)
    lib().ora_filter_tf(C.byref(f), nfreq, df, P(tf))
    return tf[0::2] + 1j * tf[1::2]


def mapping(xi, eta, nodes, eltype):
    sz = np.zeros(2)
    lib().ora_mapping(xi, eta, P(np.asfortranarray(nodes, float).ravel(order="F")), eltype, P(sz))
    return sz


def inv_mapping(s, z, nodes, eltype):
    out = np.zeros(2)
    lib().ora_inv_mapping(s, z, P(np.asfortranarray(nodes, float).ravel(order="F")), eltype, P(out))
    return out


def strain_pole(pole, u, nodes, eltype, axial, trace=False):
    """u [nvar][jpol][ipol][nsamp] -> strain [6][jpol][ipol][nsamp] (or trace)."""
    u = np.ascontiguousarray(u, float)
    ns = u.shape[-1]
    nd = np.asfortranarray(nodes, float).ravel(order="F")
    if trace:
        out = np.zeros((5, 5, ns))
        lib().ora_straintrace_pole(pole, P(u), ns, 4, P(nd), eltype, int(axial), P(out))
    else:
        out = np.zeros((6, 5, 5, ns))
        lib().ora_strain_pole(pole, P(u), ns, 4, P(nd), eltype, int(axial), P(out))
    return out


def nearest(mp_s, mp_z, s, z, nn=12):
    idx = np.zeros(nn, np.int32)
    dist = np.zeros(nn, np.float32)
    tie = lib().ora_nearest(len(mp_s), P(mp_s, C.c_float), P(mp_z, C.c_float), s, z, nn,
                            P(idx, C.c_int32), P(dist, C.c_float))
    return idx, dist, tie


def integrate_parseval(f, g, dt):
    f = np.ascontiguousarray(f, float)
    g = np.ascontiguousarray(g, float)
    return lib().ora_integrate_parseval(len(f), dt, P(f), P(g))


def basekernel(ibk, full, F, B):
    """calc_basekernel: F, B complex [nd][nw] -> complex [nw]"""
    F = np.ascontiguousarray(F, np.complex128)
    B = np.ascontiguousarray(B, np.complex128)
    nd, nw = F.shape
    out = np.zeros(nw, np.complex128)
    lib().ora_basekernel(ibk, int(full), nd, nw, P(F.view(float)), P(B.view(float)), P(out.view(float)))
    return out


def dampen_field(base, pts, c, rmax):
    """dampen_field on base [npts][nk][6] (a damped copy)"""
    b = np.array(base, float, order="C")
    pts = np.ascontiguousarray(pts, float)
    c = np.ascontiguousarray(c, float)
    lib().ora_dampen_field(P(b), b.shape[0], b.shape[1], P(pts), P(c), rmax)
    return b


def add_stfs(tf, stf_sem, dt, amplitude=1.0, stf_src=None, stf_shift=0.0):
    """add_stfs on a filter response tf [nw] (N = 2 (nw - 1)); (rc, tf_fwd, tf_bwd, message)"""
    stf_sem = np.ascontiguousarray(stf_sem, float)
    nt = len(stf_sem)
    tff = np.array(tf, np.complex128)
    tfb = np.zeros_like(tff)
    N = 2 * (len(tff) - 1)
    src = None if stf_src is None else np.ascontiguousarray(stf_src, float)
    msg = C.create_string_buffer(256)
    rc = lib().ora_add_stfs(nt, dt, N, P(stf_sem), amplitude, P(src), stf_shift, P(tff.view(float)),
                            P(tfb.view(float)), msg, 256)
    return rc, tff, tfb, msg.value.decode()


def physical(param, base, vp, vs, rho, phi=1.0, xi=1.0, eta=1.0, relative=True):
    b = np.ascontiguousarray(base, float)
    return lib().ora_physical(abi.PARAMS[param], P(b), vp, vs, rho, phi, xi, eta, int(relative))


def param_needs(param):
    nb = np.zeros(6, np.int32)
    rc = lib().ora_param_needs(abi.PARAMS[param], P(nb, C.c_int32))
    return rc, nb.astype(bool)


def azim_factor(phi, mij, isim, ikind):
    m = np.ascontiguousarray(mij, float)
    return lib().ora_azim_factor(phi, P(m), isim, ikind)


def rotate_frame(xyz, lon, colat):
    """Points [n][3] into the frame whose pole is at (lon, colat) (radians): s, phi, z."""
    xyz = np.ascontiguousarray(xyz, float)
    n = xyz.shape[0]
    s, ph, z = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().ora_rotate_frame(n, P(xyz), lon, colat, P(s), P(ph), P(z))
    return s, ph, z


def cut_window(t, w0, w1):
    """Indices of the samples of time axis t inside [w0, w1] (first, count)."""
    t = np.ascontiguousarray(t, float)
    i0, n = C.c_int(), C.c_int()
    lib().ora_cut_window(len(t), P(t), w0, w1, C.byref(i0), C.byref(n))
    return i0.value, n.value


def rotate_voigt(t, phi, theta=0.0, to_earth=False):
    """t [6] or [6][nt] Voigt (xx, yy, zz, yz, xz, xy) -> rotated copy."""
    a = np.array(t, float, order="C")
    nt = 1 if a.ndim == 1 else a.shape[1]
    lib().ora_rotate_voigt(int(to_earth), P(a), nt, phi, theta)
    return a


def rotate_voigt_inv(t, phi, theta=0.0, earth_to_xyz_src=True):
    """t [6] Voigt: xyz_earth -> xyz_src (phi, theta) or xyz -> src (phi); rotated copy."""
    a = np.array(t, float, order="C")
    lib().ora_rotate_voigt_inv(2 if earth_to_xyz_src else 3, P(a), phi, theta)
    return a


def mc_run(vals, volume, allowed_error=0.0, allowed_rel=0.0):
    """Integrator alone: vals [nsteps][npts][n] -> (integral, variance, converged)."""
    vals = np.ascontiguousarray(vals, float)
    ns, npts, n = vals.shape
    I = np.zeros(n)
    V = np.zeros(n)
    cv = np.zeros(n, np.int32)
    rc = lib().ora_mc_run(n, volume, allowed_error, allowed_rel, ns, npts, P(vals), P(I), P(V),
                          P(cv, C.c_int32))
    if rc:
        raise RuntimeError("oracle mc_run failed (%d)" % rc)
    return I, V, cv.astype(bool)


def ora_nomega(nt):
    return lib().ora_nextpow2(nt) + 1


class Oracle:
    """Whole-path oracle for one MckSetup (the Problem must stay alive)."""

    def __init__(self, setup):
        err = C.create_string_buffer(512)
        self.h = lib().ora_create(C.byref(setup), err, 512)
        if not self.h:
            raise RuntimeError("oracle setup failed: " + err.value.decode())
        self.nk = setup.nkernel
        self.nt = setup.nt
        self.nrec = setup.nrec
        self.nfilter = setup.nfilter
        self.nomega = ora_nomega(setup.nt)
        self.nd = lib().ora_strain_nd(self.h)
        self.halton_step = 0

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_destroy(self.h)
            self.h = None

    def set_lazy_fill(self, fn_ptr, user_ptr):
        """Generate database elements on first use (fn_ptr: C function address)."""
        lib().ora_set_lazy_fill(self.h, fn_ptr, user_ptr)

    def locate(self, frame, xyz):
        xyz = np.ascontiguousarray(xyz, float)
        n = xyz.shape[0]
        elem = np.zeros(n, np.int32)
        xi = np.zeros(n)
        eta = np.zeros(n)
        ties = np.zeros(n, np.int32)
        rc = lib().ora_locate(self.h, frame, n, P(xyz), P(elem, C.c_int32), P(xi), P(eta),
                              P(ties, C.c_int32))
        if rc:
            raise RuntimeError("oracle locate failed (%d)" % rc)
        return elem, xi, eta, ties

    def load_strain(self, frame, xyz):
        xyz = np.ascontiguousarray(xyz, float)
        out = np.zeros((xyz.shape[0], self.nd, self.nt))
        rc = lib().ora_load_strain(self.h, frame, xyz.shape[0], P(xyz), P(out))
        if rc:
            raise RuntimeError("oracle load_strain failed (%d)" % rc)
        return out

    def integrate_model(self, verts, conn, nbasis=1, halton_step=0, halton_offsets=None):
        """integrate_1d_model for a task: (model [ne][nbasis][12], nsteps [ne], new step)."""
        verts = np.ascontiguousarray(verts, float).reshape(-1, 3)
        conn = np.ascontiguousarray(conn, np.int32)
        ne, nvpe = conn.shape
        model = np.zeros((ne, nbasis, 12))
        nsteps = np.zeros(ne, np.int32)
        step = np.array([halton_step], np.int64)
        offs = None if halton_offsets is None else np.ascontiguousarray(halton_offsets, np.int64)
        rc = lib().ora_integrate_model(self.h, 4 if nvpe == 4 else 5, nvpe, ne, nbasis, P(verts),
                                       P(conn, C.c_int32), P(model), P(nsteps, C.c_int32),
                                       P(step, C.c_int64), None if offs is None else P(offs, C.c_int64))
        if rc:
            raise RuntimeError("oracle integrate_model failed (%d)" % rc)
        return model, nsteps, int(step[0])

    def seismograms(self):
        """[nrec][nt] seismograms by reciprocity from the bwd database."""
        out = np.zeros((self.nrec, self.nt))
        rc = lib().ora_seismograms(self.h, P(out))
        if rc:
            raise RuntimeError("oracle seismograms failed (%d)" % rc)
        return out

    def base_kernels(self, xyz):
        xyz = np.ascontiguousarray(xyz, float)
        out = np.zeros((xyz.shape[0], self.nk, 6))
        rc = lib().ora_base_kernels(self.h, xyz.shape[0], P(xyz), P(out))
        if rc:
            raise RuntimeError("oracle base_kernels failed (%d)" % rc)
        return out

    def kernel_info(self, k):
        norm = C.c_double()
        ncut, ifirst = C.c_int(), C.c_int()
        lib().ora_kernel_info(self.h, k, C.byref(norm), C.byref(ncut), C.byref(ifirst))
        return norm.value, ncut.value, ifirst.value

    def slave_work(self, verts, conn, halton_offsets=None, nbasis=1):
        """values/variance (ne, nk) for volumetric elements, (ne, nbasis, nk) otherwise;
        conn (ne, 4) tetrahedra or (ne, 8) voxels"""
        verts = np.ascontiguousarray(verts, float)
        conn = np.ascontiguousarray(conn, np.int32)
        ne, nvpe = conn.shape
        etype = 4 if nvpe == 4 else 5
        shape = (ne, self.nk) if nbasis == 1 else (ne, nbasis, self.nk)
        vals = np.zeros(shape)
        var = np.zeros(shape)
        nit = np.zeros((ne, self.nk), np.int32)
        step = np.array([self.halton_step], np.int64)
        used = np.zeros(ne, np.int64)
        ev = np.zeros(1, np.int64)
        offs = None if halton_offsets is None else np.ascontiguousarray(halton_offsets, np.int64)
        rc = lib().ora_slave_work(self.h, etype, nvpe, ne, nbasis, P(verts), P(conn, C.c_int32), P(vals), P(var),
                                  P(nit, C.c_int32), P(step, C.c_int64), P(offs, C.c_int64),
                                  P(used, C.c_int64), P(ev, C.c_int64))
        if rc:
            raise RuntimeError("oracle slave_work failed (%d)" % rc)
        self.halton_step = int(step[0])
        return dict(values=vals, variance=var, niterations=nit, offsets=used,
                    evaluations=int(ev[0]), halton_step=int(step[0]))

    # the filter-convolve-window-integrate stage (G complex [nfilter][nomega], fw
    # [npts][nd][nt], bw [npts][nrec][nd][nt]) and its FP64 adjoint
    def ad_transfer(self):
        G = np.zeros((self.nfilter, self.nomega, 2))
        lib().ora_ad_transfer(self.h, P(G))
        return G[..., 0] + 1j * G[..., 1]

    def _ad_inputs(self, G, fw, bw):
        G = np.ascontiguousarray(np.stack([np.real(G), np.imag(G)], -1), float)
        fw = np.ascontiguousarray(fw, float)
        bw = np.ascontiguousarray(bw, float)
        assert G.shape == (self.nfilter, self.nomega, 2)
        assert fw.shape[1:] == (self.nd, self.nt) and bw.shape[1:] == (self.nrec, self.nd, self.nt)
        assert bw.shape[0] == fw.shape[0]
        return G, fw, bw

    def ad_forward(self, G, fw, bw):
        G, fw, bw = self._ad_inputs(G, fw, bw)
        out = np.zeros((fw.shape[0], self.nk, 6))
        rc = lib().ora_ad_forward(self.h, fw.shape[0], P(G), P(fw), P(bw), P(out))
        if rc:
            raise RuntimeError("oracle ad_forward failed (%d)" % rc)
        return out

    def ad_backward(self, G, fw, bw, gbase):
        G, fw, bw = self._ad_inputs(G, fw, bw)
        gbase = np.ascontiguousarray(gbase, float)
        assert gbase.shape == (fw.shape[0], self.nk, 6)
        gG = np.zeros_like(G)
        gfw = np.zeros_like(fw)
        gbw = np.zeros_like(bw)
        rc = lib().ora_ad_backward(self.h, fw.shape[0], P(G), P(fw), P(bw), P(gbase), P(gG),
                                   P(gfw), P(gbw))
        if rc:
            raise RuntimeError("oracle ad_backward failed (%d)" % rc)
        return gG[..., 0] + 1j * gG[..., 1], gfw, gbw


class Results:
    """The master's accumulators on the CPU (ora_results_add / ora_results_export),
    the checker of mck_results_*: K, V [npoints][nk], M [npoints][nmodel], niter
    [nelements][nk], ctime, proc [nelements]."""

    def __init__(self, int_type, npoints, nelements, nbasis, nk, nmodel=0, connectivity=None, nhet=0):
        self.int_type, self.np, self.ne, self.nb, self.nk, self.nm = int_type, npoints, nelements, nbasis, nk, nmodel
        self.nh = nhet
        self.H = np.zeros((npoints, max(nhet, 1)))
        self.conn = None if connectivity is None else np.ascontiguousarray(connectivity, np.int32)
        self.K = np.zeros((npoints, nk))
        self.V = np.zeros((npoints, nk))
        self.M = np.zeros((npoints, max(nmodel, 1)))
        self.niter = np.zeros((nelements, nk), np.int32)
        self.ctime = np.zeros(nelements)
        self.proc = np.zeros(nelements, np.int32)

    def add(self, elems, values, variance, niterations, computation_time=None, model=None, rank=1, hetero=None):
        elems = np.ascontiguousarray(elems, np.int32)
        values = np.ascontiguousarray(values, float)
        variance = np.ascontiguousarray(variance, float)
        niterations = np.ascontiguousarray(niterations, np.int32)
        ct = None if computation_time is None else np.ascontiguousarray(computation_time, float)
        md = None if model is None else np.ascontiguousarray(model, float)
        rc = lib().ora_results_add(self.int_type, self.nb, self.nk, self.nm, self.np, self.ne,
                                   None if self.conn is None else P(self.conn, C.c_int32), len(elems),
                                   P(elems, C.c_int32), P(values), P(variance), P(niterations, C.c_int32),
                                   None if ct is None else P(ct), None if md is None else P(md), rank,
                                   P(self.K), P(self.V), P(self.M), P(self.niter, C.c_int32), P(self.ctime),
                                   P(self.proc, C.c_int32))
        if rc:
            raise RuntimeError("oracle results_add: id out of range")
        if self.nh and hetero is not None:
            # Het_Model by the same loop (master_queue.f90:301-304): no kernel columns
            hd = np.ascontiguousarray(hetero, float)
            z = np.zeros(1)
            lib().ora_results_add(self.int_type, self.nb, 0, self.nh, self.np, self.ne,
                                  None if self.conn is None else P(self.conn, C.c_int32), len(elems),
                                  P(elems, C.c_int32), P(z), P(z), P(niterations, C.c_int32),
                                  None if ct is None else P(ct), P(hd), rank, P(z), P(z), P(self.H),
                                  P(self.niter, C.c_int32), P(self.ctime), P(self.proc, C.c_int32))

    def export(self, which):
        from mc_kernel_amd import abi as A
        if which == A.MCK_RES_NITER:
            a, ncol, n = None, self.nk, self.ne
        elif which == A.MCK_RES_MODEL:
            a, ncol, n = self.M, self.nm, self.np
        elif which == A.MCK_RES_HETERO:
            a, ncol, n = self.H, self.nh, self.np
        else:
            a, ncol, n = (self.K if which == A.MCK_RES_KERNEL else self.V), self.nk, self.np
        out = np.zeros((ncol, n), np.float32)
        lib().ora_results_export(which, n, ncol, None if a is None else P(np.ascontiguousarray(a)),
                                 P(self.K), P(self.niter, C.c_int32), P(out, C.c_float))
        return out


def het_xyz(rtp):
    """load_het_rtpv's single-precision Cartesian coordinates of (r km, lat, lon) rows."""
    rtp = np.ascontiguousarray(rtp, float).reshape(-1, 3)
    out = np.zeros((rtp.shape[0], 3), np.float32)
    if lib().ora_het_xyz(rtp.shape[0], P(rtp), P(out, C.c_float)):
        raise ValueError("latitude outside [-90, 90]")
    return out


def het_nearest(xyz, q):
    """0-based nearest model point of each query row (brute force, fp32 distances)."""
    xyz = np.ascontiguousarray(xyz, np.float32)
    return np.array([lib().ora_het_nearest(xyz.shape[0], P(xyz, C.c_float), float(a), float(b), float(c))
                     for a, b, c in np.asarray(q, np.float32)])


def integrate_hetero(rtp, dln, verts, conn, nbasis=1, halton_step=0, halton_offsets=None):
    """integrate_3d_model for a task: (hetero [ne][nbasis][3], nsteps [ne], new step)."""
    rtp = np.ascontiguousarray(rtp, float).reshape(-1, 3)
    dln = np.ascontiguousarray(dln, np.float32).reshape(-1, 3)
    verts = np.ascontiguousarray(verts, float).reshape(-1, 3)
    conn = np.ascontiguousarray(conn, np.int32)
    ne, nvpe = conn.shape
    out = np.zeros((ne, nbasis, 3))
    nsteps = np.zeros(ne, np.int32)
    step = np.array([halton_step], np.int64)
    offs = None if halton_offsets is None else np.ascontiguousarray(halton_offsets, np.int64)
    rc = lib().ora_integrate_hetero(rtp.shape[0], P(rtp), P(dln, C.c_float), 4 if nvpe == 4 else 5, nvpe, ne,
                                    nbasis, P(verts), P(conn, C.c_int32), P(out), P(nsteps, C.c_int32),
                                    P(step, C.c_int64), None if offs is None else P(offs, C.c_int64))
    if rc:
        raise RuntimeError("oracle integrate_hetero failed (%d)" % rc)
    return out, nsteps, int(step[0])
