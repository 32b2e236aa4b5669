"""ctypes mirror of include/mck.h (the C-ABI of the MI355X MC Kernel path).

The structures below must match include/mck.h field by field; tests/test_abi.py checks
their sizes against the compiled library.
"""
import ctypes as C

MCK_OK = 0
MCK_EINVAL = 1
MCK_ENOTINIT = 2
MCK_ENOMEM = 3
MCK_EHIP = 4
MCK_ENOTFOUND = 5
MCK_ENAN = 6
MCK_EUNSUPPORTED = 7

# result variables of mck_results_export
MCK_RES_KERNEL = 0
MCK_RES_VARIANCE = 1
MCK_RES_ERROR = 2
MCK_RES_MODEL = 3
MCK_RES_NITER = 4
MCK_RES_HETERO = 5

PARAMS = {"lam": 0, "vp": 1, "vs": 2, "mu": 3, "vsh": 4, "vsv": 5, "vph": 6, "vpv": 7,
          "eta": 8, "rho": 9}
MISFITS = {"CC": 0, "AM": 1}
INT_SCHEMES = {"parseval": 0, "trapezoidal": 1}
FILTERS = {"ident": 0, "Gabor": 1, "Butterw_LP": 2, "Butterw_HP": 3, "Butterw_BP": 4,
           "Butterw_LP_ZP": 5, "Butterw_HP_ZP": 6, "Butterw_BP_ZP": 7}

f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)


class MckDb(C.Structure):
    _fields_ = [
        ("npol", C.c_int32), ("nelem", C.c_int32), ("npoints", C.c_int32),
        ("mesh_s", f32p), ("mesh_z", f32p), ("mp_s", f32p), ("mp_z", f32p),
        ("fem_mesh", i32p), ("eltype", i32p), ("axis", i32p),
        ("nvar", C.c_int32), ("snapshots", f32p), ("amplitude", C.c_double),
    ]


class MckFilter(C.Structure):
    _fields_ = [("filterclass", C.c_int32), ("frequencies", C.c_double * 4)]


class MckSetup(C.Structure):
    _fields_ = [
        ("fwd", MckDb), ("bwd", MckDb),
        ("nt", C.c_int32), ("dt", C.c_double), ("planet_radius", C.c_double),
        ("timeshift_fwd", C.c_double), ("timeshift_bwd", C.c_double), ("stf_fwd", f64p),
        ("src_lat", C.c_double), ("src_lon", C.c_double), ("src_depth_km", C.c_double),
        ("src_mij", C.c_double * 6), ("stf_src", f64p), ("stf_src_shift", C.c_double),
        ("nrec", C.c_int32), ("rec_lat", f64p), ("rec_lon", f64p), ("component", C.c_char),
        ("seis", f64p),
        ("nfilter", C.c_int32), ("filters", C.POINTER(MckFilter)),
        ("nkernel", C.c_int32), ("kernel_receiver", i32p), ("kernel_filter", i32p),
        ("kernel_misfit", i32p), ("kernel_param", i32p), ("kernel_window", f64p),
        ("model_ndepth", C.c_int32), ("model_dr", C.c_float),
        ("model_vp", f32p), ("model_vs", f32p), ("model_rho", f32p),
        ("model_phi", f32p), ("model_xi", f32p), ("model_eta", f32p),
        ("vp_max", C.c_float), ("vs_max", C.c_float),
        ("npoints_per_step", C.c_int32), ("max_iter", C.c_int32),
        ("allowed_error", C.c_double), ("allowed_relative_error", C.c_double),
        ("int_over_volume", C.c_int32), ("relative_kernel", C.c_int32),
        ("quasirandom", C.c_int32), ("int_scheme", C.c_int32), ("deconv_stf", C.c_int32),
        ("damp_radius", C.c_double), ("mask_src_rec", C.c_int32),
        ("db_page_elems", C.c_int32), ("db_device_pages", C.c_int32),
    ]


class MckStats(C.Structure):
    _fields_ = [
        ("evaluations", C.c_int64), ("points", C.c_int64),
        ("fwd_locations", C.c_int64), ("bwd_locations", C.c_int64),
        ("fwd_bytes", C.c_int64), ("bwd_bytes", C.c_int64),
        ("mc_steps", C.c_int32), ("strain_launches", C.c_int32),
        ("device_ms", C.c_double), ("strain_fwd_ms", C.c_double),
        ("strain_bwd_ms", C.c_double), ("wall_ms", C.c_double),
        ("paged_bytes", C.c_int64),
        ("spt", C.c_int32), ("nd_f", C.c_int32), ("nd_b", C.c_int32), ("bwd_path", C.c_int32),
        ("fwd_elements", C.c_int64), ("bwd_elem_receivers", C.c_int64), ("alg_bytes", C.c_int64),
        ("spectra_ms", C.c_double), ("mfma_ms", C.c_double),
        ("mfma_launches", C.c_int32), ("mfma_groups", C.c_int32),
        ("mfma_runs", C.c_int64), ("pair_ms", C.c_double), ("pair_pairs", C.c_int64),
        ("pair_launches", C.c_int32), ("reserved", C.c_int32),
    ]


class MckTuning(C.Structure):
    _fields_ = [("bwd_path", C.c_int32), ("mfma_waves", C.c_int32), ("mfma_xcd", C.c_int32),
                ("ncandidates", C.c_int32), ("best_ms", C.c_double), ("default_ms", C.c_double),
                ("candidate_ms", C.c_double * 5),
                ("cost_group_us", C.c_double), ("cost_run_us", C.c_double), ("cost_pair_us", C.c_double),
                ("auto_ms", C.c_double)]


def ptr(arr, ctype):
    """Pointer to the data of a C-contiguous numpy array (None stays NULL)."""
    if arr is None:
        return None
    return arr.ctypes.data_as(C.POINTER(ctype))


# exported symbols of libmck.so, in the order of include/mck.h
EXPORTS = [
    "mck_init", "mck_slave_work", "mck_slave_work_plain", "mck_slave_work_sequential", "mck_locate_points",
    "mck_load_strain_points", "mck_base_kernels", "mck_seismograms", "mck_integrate_model", "mck_set_hetero", "mck_integrate_hetero", "mck_ad_dims", "mck_ad_transfer",
    "mck_ad_forward", "mck_ad_backward", "mck_results_init", "mck_results_add", "mck_results_add_noalloc",
    "mck_results_get",
    "mck_results_load", "mck_results_export", "mck_tune", "mck_set_tuning", "mck_get_stats", "mck_finalize",
    "mck_last_error",
]


def declare(lib):
    """Attach argument/return types of every C-ABI entry point to a loaded library."""
    lib.mck_init.argtypes = [C.POINTER(MckSetup)]
    lib.mck_init.restype = C.c_int
    lib.mck_seismograms.argtypes = [f64p]
    lib.mck_seismograms.restype = C.c_int
    lib.mck_integrate_model.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, f64p, i32p,
                                        f64p, i32p, i64p, i64p]
    lib.mck_integrate_model.restype = C.c_int
    lib.mck_set_hetero.argtypes = [C.c_int32, f64p, f32p]
    lib.mck_set_hetero.restype = C.c_int
    lib.mck_integrate_hetero.argtypes = lib.mck_integrate_model.argtypes
    lib.mck_integrate_hetero.restype = C.c_int
    work = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, f64p, i32p,
            f64p, f64p, i32p, f64p, i64p, i64p]
    lib.mck_slave_work.argtypes = work
    lib.mck_slave_work.restype = C.c_int
    lib.mck_slave_work_plain.argtypes = work
    lib.mck_slave_work_plain.restype = C.c_int
    lib.mck_tune.argtypes = [C.c_int32, f64p, C.c_int32, C.POINTER(MckTuning)]
    lib.mck_tune.restype = C.c_int
    lib.mck_set_tuning.argtypes = [C.POINTER(MckTuning)]
    lib.mck_set_tuning.restype = C.c_int
    lib.mck_slave_work_sequential.argtypes = work
    lib.mck_slave_work_sequential.restype = C.c_int
    lib.mck_locate_points.argtypes = [C.c_int32, C.c_int32, f64p, i32p, f64p, f64p, i32p]
    lib.mck_locate_points.restype = C.c_int
    lib.mck_load_strain_points.argtypes = [C.c_int32, C.c_int32, f64p, f32p]
    lib.mck_load_strain_points.restype = C.c_int
    lib.mck_base_kernels.argtypes = [C.c_int32, f64p, f64p]
    lib.mck_base_kernels.restype = C.c_int
    lib.mck_ad_dims.argtypes = [i32p] * 6
    lib.mck_ad_dims.restype = C.c_int
    lib.mck_ad_transfer.argtypes = [f32p]
    lib.mck_ad_transfer.restype = C.c_int
    # device pointers and the stream travel as addresses
    lib.mck_ad_forward.argtypes = [C.c_int32] + [C.c_void_p] * 5
    lib.mck_ad_forward.restype = C.c_int
    lib.mck_ad_backward.argtypes = [C.c_int32] + [C.c_void_p] * 8
    lib.mck_ad_backward.restype = C.c_int
    lib.mck_results_init.argtypes = [C.c_int32] * 6 + [i32p]
    lib.mck_results_init.restype = C.c_int
    # result arrays may be host or device memory: passed as addresses
    lib.mck_results_add.argtypes = [C.c_int32, i32p] + [C.c_void_p] * 6 + [C.c_int32]
    lib.mck_results_add.restype = C.c_int
    lib.mck_results_add_noalloc.argtypes = [C.c_int32, i32p] + [C.c_void_p] * 6 + [C.c_int32]
    lib.mck_results_add_noalloc.restype = C.c_int
    lib.mck_results_get.argtypes = [f64p, f64p, f64p, f64p, i32p, f64p, i32p]
    lib.mck_results_get.restype = C.c_int
    lib.mck_results_load.argtypes = [C.c_void_p] * 7
    lib.mck_results_load.restype = C.c_int
    lib.mck_results_export.argtypes = [C.c_int32, C.c_void_p]
    lib.mck_results_export.restype = C.c_int
    lib.mck_get_stats.argtypes = [C.POINTER(MckStats)]
    lib.mck_get_stats.restype = C.c_int
    lib.mck_finalize.argtypes = []
    lib.mck_finalize.restype = None
    lib.mck_last_error.argtypes = []
    lib.mck_last_error.restype = C.c_char_p
    return lib
