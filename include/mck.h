/*
 * mck.h - C-ABI of the MI355X implementation of MC Kernel's per-point kernel
 * evaluation and Monte Carlo integration path.
 *
 * The boundary is the reference's work unit: `slave_work` in src/slave.f90:273-810,
 * called once per task by `do_slave` (src/slave.f90:205-219), plus the once-per-process
 * setup that `do_slave` performs before its task loop (src/slave.f90:94-173).
 * Every entry point takes plain pointers and sizes; arrays are caller-owned and use
 * Fortran (column-major) order where the reference passes Fortran arrays, so a Fortran
 * host binds them with iso_c_binding without copies.
 *
 * Error convention: every int-returning function returns 0 on success and a nonzero
 * MCK_E* code on failure; mck_last_error() then describes the failure.  The reference
 * aborts the whole MPI job instead (pabort, src/commpi.F90:515-556); a Fortran host keeps
 * that behaviour by calling pabort on a nonzero return.
 *
 * Threading: one call in flight per process; one GPU per process (the device that is
 * current when mck_init is called).
 */
#ifndef MCK_H
#define MCK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------------------ */
#define MCK_OK              0
#define MCK_EINVAL          1   /* invalid argument / inconsistent setup               */
#define MCK_ENOTINIT        2   /* mck_init has not been called                        */
#define MCK_ENOMEM          3   /* device or host allocation failed                    */
#define MCK_EHIP            4   /* a HIP runtime call failed                           */
#define MCK_ENOTFOUND       5   /* point not inside any of the 12 nearest elements     */
                                /* (reference: 'element not found', readfields.f90:1100) */
#define MCK_ENAN            6   /* NaN in an MC integral (mc_integration.f90:113-130)  */
#define MCK_EUNSUPPORTED    7   /* feature the reference aborts on (e.g. rho kernels)  */

/* ---- model parameter codes (kernel.f90:665-832, type_parameter.f90:487-491) ---------- */
#define MCK_PARAM_LAM  0
#define MCK_PARAM_VP   1
#define MCK_PARAM_VS   2
#define MCK_PARAM_MU   3
#define MCK_PARAM_VSH  4
#define MCK_PARAM_VSV  5
#define MCK_PARAM_VPH  6
#define MCK_PARAM_VPV  7
#define MCK_PARAM_ETA  8
#define MCK_PARAM_RHO  9   /* rejected with MCK_EUNSUPPORTED: kernel.f90:554-556,824-826 */

/* ---- misfit types and integration schemes (kernel.f90:410-510) --------------------- */
#define MCK_MISFIT_CC  0
#define MCK_MISFIT_AM  1
#define MCK_INT_PARSEVAL     0
#define MCK_INT_TRAPEZOIDAL  1

/* ---- filter classes (filtering.f90:78-230) ---------------------------------------- */
#define MCK_FILTER_IDENT       0
#define MCK_FILTER_GABOR       1   /* freq = [centre period, sigma/period, tshift, -]   */
#define MCK_FILTER_BUTTER_LP   2   /* freq = [period, -, order, -]                       */
#define MCK_FILTER_BUTTER_HP   3
#define MCK_FILTER_BUTTER_BP   4   /* freq = [long period, short period, order, -]       */
#define MCK_FILTER_BUTTER_LP_ZP 5
#define MCK_FILTER_BUTTER_HP_ZP 6
#define MCK_FILTER_BUTTER_BP_ZP 7

/*
 * One merged AxiSEM wavefield database (one of fwd / bwd), already in host memory.
 * Mirrors the merged NetCDF layout read by readfields.f90 (group `Mesh` and variable
 * `merged_snapshots`, readfields.f90:2925-2998, 2673-2768; repack_database.py:151-187).
 */
typedef struct {
    int32_t npol;              /* polynomial order, 4                                    */
    int32_t nelem;             /* number of spectral elements                            */
    int32_t npoints;           /* number of global GLL points                            */
    const float   *mesh_s;     /* [npoints] s coordinate of GLL points (m)  Mesh/mesh_S  */
    const float   *mesh_z;     /* [npoints]                                  Mesh/mesh_Z */
    const float   *mp_s;       /* [nelem] element midpoints (m)           Mesh/mp_mesh_S */
    const float   *mp_z;       /* [nelem]                                 Mesh/mp_mesh_Z */
    const int32_t *fem_mesh;   /* [nelem][4] 0-based corner point ids      Mesh/fem_mesh */
    const int32_t *eltype;     /* [nelem] 0 spheroid, 1 subpar, 2 semino, 3 semiso       */
    const int32_t *axis;       /* [nelem] 1 if the element touches the axis              */
    int32_t nvar;              /* 10 (fwd: MZZ s,z; MXX s,z; MXZ s,p,z; MXY s,p,z)       */
                               /*  5 (bwd: PZ s,z; PX s,p,z)                             */
    const float   *snapshots;  /* merged_snapshots, C dims [nelem][nvar][npol+1][npol+1][nt] */
                               /* (= Fortran u(nt, ipol, jpol, nvar, nelem))             */
    double amplitude;          /* global attribute "scalar source magnitude"             */
} mck_db_t;

/* One seismic filter as read from filters.dat (type_parameter.f90:695-766). */
typedef struct {
    int32_t filterclass;       /* MCK_FILTER_*                                           */
    double  frequencies[4];    /* class-dependent parameters, see MCK_FILTER_*           */
} mck_filter_t;

/*
 * Once-per-process setup: the products of do_slave's setup phase
 * (src/slave.f90:94-173: read_parameters, read_source, read_receiver, read_filter,
 * read_kernel, sem_data%open_files/read_meshes/load_seismogram).
 */
typedef struct {
    mck_db_t fwd;              /* forward (source) database                              */
    mck_db_t bwd;              /* backward (receiver, reciprocity) database              */
    int32_t nt;                /* number of strain dumps (ndumps)                        */
    double  dt;                /* strain dump sampling rate in s                         */
    double  planet_radius;     /* m                                                      */
    double  timeshift_fwd;     /* "source shift factor in sec" of fwd / bwd databases,   */
    double  timeshift_bwd;     /*  applied when deconv_stf == 0 (slave.f90:531,596)      */
    const double *stf_fwd;     /* [nt] stf_dump of the fwd DB (used when deconv_stf)     */

    /* source (source.f90:67-110): lat/lon in degrees, depth in km, Mij in N m,          */
    /* order rr, tt, pp, rt, rp, tp                                                      */
    double  src_lat, src_lon, src_depth_km;
    double  src_mij[6];
    const double *stf_src;     /* [nt] earthquake STF sampled at dt (deconv_stf only),   */
                               /* NULL = unit impulse                                    */
    double  stf_src_shift;     /* time shift of the earthquake STF in s                  */

    /* receivers (type_parameter.f90:396-517)                                            */
    int32_t nrec;
    const double *rec_lat;     /* [nrec] degrees                                         */
    const double *rec_lon;     /* [nrec] degrees                                         */
    char    component;         /* 'Z', 'R' or 'T'                                        */
    const double *seis;        /* [nrec][nt] velocity seismogram of each receiver        */
                               /* (sem_data%seis, readfields.f90:1004-1048); NULL: mck_init */
                               /* computes them by reciprocity from the bwd database     */
                               /* (load_fw_points_rdbm, readfields.f90:1469-1604)         */

    /* filters and kernels (type_parameter.f90:521-691)                                  */
    int32_t nfilter;
    const mck_filter_t *filters;
    int32_t nkernel;           /* kernels are grouped by receiver in receiver order      */
    const int32_t *kernel_receiver; /* [nkernel] 0-based receiver of each kernel         */
    const int32_t *kernel_filter;   /* [nkernel] 0-based filter index                    */
    const int32_t *kernel_misfit;   /* [nkernel] MCK_MISFIT_*                            */
    const int32_t *kernel_param;    /* [nkernel] MCK_PARAM_*                             */
    const double  *kernel_window;   /* [nkernel][2] time window (s)                      */

    /* 1-D background model on a radial grid (interpolate_mesh.f90:45-109)               */
    int32_t model_ndepth;
    float   model_dr;          /* 100 m in the reference                                 */
    const float *model_vp, *model_vs, *model_rho, *model_phi, *model_xi, *model_eta;
    float   vp_max, vs_max;    /* maximum velocities of the model (travel-time mask)     */

    /* Monte Carlo and kernel options (inparam_basic, type_parameter.f90:139-274)        */
    int32_t npoints_per_step;  /* POINTS_PER_MC_STEP                                     */
    int32_t max_iter;          /* MAXIMUM_ITERATIONS                                     */
    double  allowed_error;     /* ALLOWED_ERROR                                          */
    double  allowed_relative_error; /* ALLOWED_RELATIVE_ERROR                            */
    int32_t int_over_volume;   /* !NO_INT_OVER_VOLUME                                    */
    int32_t relative_kernel;   /* !KERNEL_FOR_ABSOLUTE_PERTURBATIONS                     */
    int32_t quasirandom;       /* !USE_PSEUDORANDOM_NUMBERS: 1 Halton points in         */
                               /* tetrahedra; 0 the library's pseudo-random stream (the  */
                               /* reference's random_number is clock-seeded); voxels     */
                               /* always take pseudo-random points                       */
    int32_t int_scheme;        /* MCK_INT_*                                              */
    int32_t deconv_stf;        /* !NO_DECONVOLVE_STF                                     */
    double  damp_radius;       /* DAMP_RADIUS_SOURCE_RECEIVER (m); <= 0 disables         */
    int32_t mask_src_rec;      /* MASK_SOURCE_RECEIVER                                   */

    /* Databases larger than device memory (no reference counterpart: readfields.f90
     * reads hyperslabs from NetCDF on demand).  0: each database's snapshots are copied
     * to device memory once (used in place when `snapshots` already is device memory),
     * unless the two together exceed 3/4 of the free device memory, in which case they
     * are paged with pages of 1/64 of it and buffers filling half of it.  > 0: snapshots
     * in host memory are paged: they stay where they are (the caller keeps them valid
     * and unchanged until mck_finalize) and every launch copies the element pages its
     * points touch, db_page_elems elements each, into device page buffers.  Results
     * are identical to the resident mode.                                               */
    int32_t db_page_elems;
    /* Device page buffers per database when paging (0 or < 2: 2; at most the number
     * of pages + 2 are made).  The first db_device_pages - 2 distinct pages used keep a
     * buffer of their own once loaded (a working set of at most that many pages is
     * copied once, wherever its pages lie); the others rotate through the last two
     * buffers (copies overlap the kernels).  A launch whose pages all keep a buffer of
     * their own runs once over all its points, through a table of the page buffers;
     * otherwise once per page.                                                          */
    int32_t db_device_pages;
} mck_setup_t;

/* Work-unit statistics of the last mck_slave_work call (device timings via hipEvents). */
typedef struct {
    int64_t evaluations;       /* computed (voxel, MC point, kernel) values              */
    int64_t points;            /* MC points drawn                                        */
    int64_t fwd_locations;     /* points located in the fwd mesh with an active receiver */
                               /* (their forward spectra are computed)                   */
    int64_t bwd_locations;     /* point locations in the bwd mesh (sum over receivers)   */
    int64_t fwd_bytes;         /* algorithmic DB bytes of the strain gathers (fwd)       */
    int64_t bwd_bytes;         /* algorithmic DB bytes of the strain gathers (bwd)       */
    int32_t mc_steps;          /* step-synchronous MC iterations performed               */
    int32_t strain_launches;   /* launches of the strain-gather/FFT kernel               */
    double  device_ms;         /* device time of the whole call (hipEvent)               */
    double  strain_fwd_ms;     /* summed time of the fwd strain-gather and FFT kernels   */
    double  strain_bwd_ms;     /* summed time of the bwd strain-gather/FFT kernel        */
    double  wall_ms;           /* host wall time of the whole call                       */
    int64_t paged_bytes;       /* snapshot bytes copied host -> device by paging         */
    int32_t spt;               /* samples per thread of the strain/FFT kernels           */
    int32_t nd_f, nd_b;        /* carried forward / backward strain components           */
    int32_t bwd_path;          /* backward evaluation: 0 per-pair FFTs, 1 element        */
                               /* spectra and a matrix-core contraction                  */
    int64_t fwd_elements;      /* sum over launches of the distinct fwd elements |F|     */
    int64_t bwd_elem_receivers;/* sum over launches of sum_r |G_r|, the distinct bwd     */
                               /* elements located by the points in receiver r's frame   */
    int64_t alg_bytes;         /* algorithmic bytes of the whole call, summed over its   */
                               /* launches: 4 nt 25 (10 |F| + 5 sum_r |G_r|) + 48 P      */
                               /* + 24 V K (merged blocks once per launch and receiver,  */
                               /* point coordinates, MC state of the evaluated voxels)   */
    double  spectra_ms;        /* summed time of the element-spectra kernels (path 1)    */
    double  mfma_ms;           /* summed time of the matrix-core evaluation kernels      */
                               /* (path 1); both timed with events on the library stream */
    int32_t mfma_launches;     /* launches of the matrix-core evaluation kernel          */
    int32_t mfma_groups;       /* waves (groups of 16 pairs) over those launches         */
    int64_t mfma_runs;         /* element runs (element spectra) over those launches     */
    double  pair_ms;           /* summed time of the per-pair kernel (k_bwd_eval)        */
    int64_t pair_pairs;        /* pairs it evaluated                                     */
    int32_t pair_launches;     /* its launches                                           */
    int32_t reserved;
} mck_stats_t;

/*
 * Replaces the setup half of do_slave (src/slave.f90:94-173): uploads both databases to
 * HBM, builds the point-location grids, the derivative matrices (calc_gradient_terms,
 * readfields.f90:3084-3150), the filters (filtering.f90:78-470) and the per-kernel
 * window/seismogram integration weights (kernel.f90:181-385).
 */
int mck_init(const mck_setup_t *setup);

/*
 * Replaces slave_work (src/slave.f90:273-810) for one task of inversion elements.
 *  elem_type       4 = tetrahedron, 5 = voxel (block in radius, colatitude and longitude
 *                  given by its 8 corners, random_voxel.f90); others rejected
 *  nvpe            vertices per element (4 or 8)
 *  nelems          elements in this task
 *  nbasis          basis functions per element: 1 = volumetric, 4 = linear functions on
 *                  the tetrahedron's vertices (inversion_mesh.f90:530-622, one Monte Carlo
 *                  integrator per basis function, slave.f90:421, 489-497, 723-744); voxels
 *                  take 1 or 8 (every weight 1, as weights() gives them)
 *  vertices        [nvertices][3] metres (Fortran vertices(3, nvertices))
 *  conn            [nelems][nvpe] 1-based vertex ids (Fortran connectivity(nvpe, nelems));
 *                  elements whose first id is <= 0 are skipped (empty slots of a last task)
 *  kernel_values   out [nelems][nbasis][nkernel]   (Fortran (nkernel, nbasis, nelems))
 *  kernel_variance out [nelems][nbasis][nkernel]
 *  niterations     out [nelems][nkernel]           (Fortran (nkernel, nelems))
 *  computation_time out [nelems] seconds (device time of the task / live elements)
 *  halton_step     in/out position of the Halton stream (halton_sequence.f90:109-153)
 *  halton_offsets  optional [nelems] Halton stream position at which each element starts;
 *                  NULL = each element reserves (max_iter+1)*npoints_per_step draws of the
 *                  stream starting at *halton_step (voxels are then independent)
 */
int mck_slave_work(int32_t elem_type, int32_t nvpe, int32_t nelems, int32_t nbasis,
                   int32_t nvertices, const double *vertices, const int32_t *conn,
                   double *kernel_values, double *kernel_variance, int32_t *niterations,
                   double *computation_time, int64_t *halton_step,
                   const int64_t *halton_offsets);

/*
 * The same work unit evaluated the way the reference loop runs it: one element and one
 * MC step at a time, each step a separate sequence of launches over npoints_per_step
 * points.  Same arguments and results as mck_slave_work; used as the one-point-at-a-time
 * comparator for the batched form.
 */
int mck_slave_work_plain(int32_t elem_type, int32_t nvpe, int32_t nelems, int32_t nbasis,
                         int32_t nvertices, const double *vertices, const int32_t *conn,
                         double *kernel_values, double *kernel_variance,
                         int32_t *niterations, double *computation_time,
                         int64_t *halton_step, const int64_t *halton_offsets);

/*
 * The work unit on the reference's own Halton stream (slave_work with the module-level
 * position of halton_sequence.f90:109-153, advanced by every draw, including the draw
 * before the max_iter exit, slave.f90:501-511): elements one after another in the
 * plain form's order, element e starting where element e-1's draws ended, so that a task
 * gives the reference's points without positions computed elsewhere.
 *  halton_step     in: the position at which the first element starts;
 *                  out: the position after the task's last draw
 *  offsets_used    optional out [nelems]: the position at which each element started
 *                  (passed as mck_slave_work's halton_offsets, the batched form
 *                  evaluates the same points)
 * Other arguments and results as mck_slave_work.
 */
int mck_slave_work_sequential(int32_t elem_type, int32_t nvpe, int32_t nelems, int32_t nbasis,
                              int32_t nvertices, const double *vertices, const int32_t *conn,
                              double *kernel_values, double *kernel_variance,
                              int32_t *niterations, double *computation_time,
                              int64_t *halton_step, int64_t *offsets_used);

/*
 * Point location, replacing the front half of load_fw_points / load_bw_points
 * (readfields.f90:1192-1226, 1359-1418) and find_element (readfields.f90:1052-1134).
 *  frame      0 = fwd (source frame), 1+r = bwd frame of receiver r
 *  xyz        [npts][3] metres
 *  elem       out [npts] 1-based element id (the reference's kd-tree index)
 *  xi, eta    out [npts] reference coordinates
 *  ties       out [npts] 1 where two of the 13 nearest midpoints have equal fp32
 *             distance (the reference's heap order then decides; NULL allowed)
 */
int mck_locate_points(int32_t frame, int32_t npts, const double *xyz,
                      int32_t *elem, double *xi, double *eta, int32_t *ties);

/*
 * Strain traces at points, replacing load_fw_points (readfields.f90:1138-1308) for
 * frame 0 and load_bw_points (readfields.f90:1312-1465) for frame 1+r.
 *  traces     out [npts][nd][nt] float (Fortran (nt, nd, npts)); nd = 6 for full strain,
 *             1 for the trace.  The fwd result is before cumsum_trapezoidal.
 */
int mck_load_strain_points(int32_t frame, int32_t npts, const double *xyz, float *traces);

/*
 * Base-kernel values for given points, replacing the receiver loop body of slave_work
 * (slave.f90:588-640: load_bw_points, rfft, calc_basekernel, apply_filter, irfft,
 * calc_misfit_kernel) for all receivers, one MC batch, without masks or damping.
 *  base       out [npts][nkernel][6] double (Fortran kernelvalue_basekers(npts, nkernel, 6)
 *             transposed): base kernels lambda, mu, rho, a, b, c of each kernel; slots the
 *             kernel's model parameter does not need are 0
 */
int mck_base_kernels(int32_t npts, const double *xyz, double *base);

/*
 * The seismograms the kernels' windows use, seis[nrec][nt]: a copy of mck_setup_t.seis,
 * or, when that was NULL, the ones mck_init computed by reciprocity from the backward
 * database (sem_data%load_seismogram, readfields.f90:1004-1048, 1469-1604: the source
 * point located in each receiver frame, the moment tensor rotated into that frame and
 * contracted with the backward strain there).
 */
int mck_seismograms(double *seis);

/*
 * The model integrals of a task, integrate_1d_model (slave.f90:815-869, computed by the
 * reference for every element when INT_OVER_BACKGROUND_MODEL is set, slave.f90:450-455,
 * and returned as slave_result%model, slave.f90:475, 781): per element and basis
 * function the Monte Carlo integral (volume 1, allowed error 1e-2, relative 1e-2) of the
 * 12 background-model parameters vp vs rho vph vpv vsh vsv eta phi xi lam mu
 * (background_model.f90:34-37) times the basis function, 100 points per step (Halton in
 * tetrahedra whatever quasirandom says, the library's pseudo-random stream in voxels),
 * at most 9999 steps.  Elements and vertices as mck_slave_work.
 *  model           out [nelems][nbasis][12]   (Fortran model(12, nbasis, nelems))
 *  nsteps          out [nelems] MC steps taken (nullable)
 *  halton_step     in/out; with halton_offsets == NULL element e's points start at
 *                  *halton_step + e * 999900 (9999 steps of 100) and *halton_step
 *                  advances past the task
 *  halton_offsets  optional [nelems] stream position of each element's first point
 */
int mck_integrate_model(int32_t elem_type, int32_t nvpe, int32_t nelems, int32_t nbasis,
                        int32_t nvertices, const double *vertices, const int32_t *conn,
                        double *model, int32_t *nsteps, int64_t *halton_step,
                        const int64_t *halton_offsets);

/*
 * The heterogeneity model for mck_integrate_hetero, replacing load_het_rtpv and
 * build_hetero_kdtree (heterogeneities.f90:63-136, called by do_slave, slave.f90:126-134
 * when INT_OVER_3D_HETEROGENEITIES is set): nhet points, rtp [nhet][3] = radius in km,
 * latitude and longitude in degrees (the columns of the heterogeneity file; a latitude
 * outside [-90, 90] is rejected as the reference stops on it), dlnhet [nhet][3] = dlnvp,
 * dlnvs, dlnrho.  The points are converted to single-precision Cartesian coordinates as
 * the reference does and held in device memory.
 */
int mck_set_hetero(int32_t nhet, const double *rtp, const float *dlnhet);

/*
 * The heterogeneity integrals of a task, integrate_3d_model (slave.f90:874-924, computed
 * for every element when INT_OVER_3D_HETEROGENEITIES is set, slave.f90:459-461, and
 * returned as slave_result%hetero_model): per element and basis function the Monte Carlo
 * integral (volume 1, allowed error 1e-2, relative 1e-1) of dlnvp, dlnvs, dlnrho of the
 * nearest model point (load_model_coeffs, heterogeneities.f90:141-188; equal fp32
 * distances take the lower model index) times the basis function, 10 points per step,
 * at most 9999 steps.  Arguments as mck_integrate_model:
 *  hetero          out [nelems][nbasis][3]   (Fortran hetero_model(3, nbasis, nelems))
 *  halton_step     with halton_offsets == NULL element e's points start at
 *                  *halton_step + e * 99990 (9999 steps of 10)
 */
int mck_integrate_hetero(int32_t elem_type, int32_t nvpe, int32_t nelems, int32_t nbasis,
                         int32_t nvertices, const double *vertices, const int32_t *conn,
                         double *hetero, int32_t *nsteps, int64_t *halton_step,
                         const int64_t *halton_offsets);

/*
 * The filter-convolve-window-integrate stage of the receiver loop (slave.f90:515-534,
 * 586-640: rfft of the tapered traces, calc_basekernel kernel.f90:514-570,
 * apply_filter_2d kernel.f90:946-956, irfft and calc_misfit_kernel kernel.f90:410-510) as a
 * differentiable function, for a gradient check by automatic differentiation.  The
 * reference has no such interface: it is the entry of the PyTorch binding.  Inputs:
 *  tf         [nfilter][nomega][2] float: the filter responses G (filtering.f90:78-230);
 *             the stage applies TF_fwd TF_bwd = G^2 x (STF deconvolution factors, time
 *             shifts) as the set-up defines them
 *  fw         [npts][nd][nt] float: earth-frame fwd strain traces before
 *             cumsum_trapezoidal (as mck_load_strain_points frame 0 returns them)
 *  bw         [npts][nrec][nd][nt] float: bwd strain traces of every receiver
 * The filtered seismogram windows and their normalisations (kernel.f90:181-385) are the
 * set-up's and enter as constants.  All arrays are DEVICE memory; work is queued on
 * `stream` (a hipStream_t, NULL = default stream) and complete on return.  Scratch
 * buffers are the library's: calls must not overlap.
 */
/* nd (1 trace, 6 full strain), nt, nomega = N/2 + 1, nfilter, nrec, nkernel */
int mck_ad_dims(int32_t *nd, int32_t *nt, int32_t *nomega, int32_t *nfilter, int32_t *nrec,
                int32_t *nkernel);
/* the set-up's filter responses G, host [nfilter][nomega][2] */
int mck_ad_transfer(float *tf);
/* base [npts][nkernel][6] double, as mck_base_kernels */
int mck_ad_forward(int32_t npts, const float *tf, const float *fw, const float *bw, double *base,
                   void *stream);
/* Adjoint for the weights gbase [npts][nkernel][6] of the base kernels: gtf (d/dRe G +
 * i d/dIm G of sum gbase base), gfw, gbw, shaped as tf, fw, bw.  The FFTs are
 * recomputed from the traces. */
int mck_ad_backward(int32_t npts, const float *tf, const float *fw, const float *bw,
                    const double *gbase, float *gtf, float *gfw, float *gbw, void *stream);

/*
 * Result assembly: the master's accumulators K_x, Var, Bg_Model, niterations,
 * computation_time and element_proc (master_queue.f90:253-323 extract_receive_buffer,
 * the arrays allocated by init_queue, master_queue.f90:60-219), held in device memory.
 * They belong to the set-up of mck_init (nkernel columns) and are freed by mck_init and
 * mck_finalize.
 *  int_type     0 = 'onvertices' (a basis function's result goes to its vertex,
 *               connectivity(ibasis, ielement)), 1 = 'volumetric' (to the element)
 *  npoints      inversion-mesh vertices (onvertices) or elements (volumetric)
 *  nelements    inversion-mesh elements
 *  nbasis       basis functions per element (the tasks' nbasis)
 *  nmodel       0, or 12 to accumulate the background-model integrals
 *               (INT_OVER_BACKGROUND_MODEL, mck_integrate_model)
 *  nhet         0, or 3 to accumulate the heterogeneity integrals (Het_Model,
 *               INT_OVER_3D_HETEROGENEITIES, mck_integrate_hetero)
 *  connectivity [nelements][nbasis] 1-based vertex of each basis function (Fortran
 *               connectivity(nbasis, nelements)); NULL in volumetric mode
 */
int mck_results_init(int32_t int_type, int32_t npoints, int32_t nelements, int32_t nbasis,
                     int32_t nmodel, int32_t nhet, const int32_t *connectivity);

/*
 * Adds one task's results, replacing extract_receive_buffer (master_queue.f90:253-323):
 * K_x(ipoint,:) += kernel_values(:, ibasis, iel), Var likewise, Bg_Model with the model
 * integrals; niterations(ielement,:), computation_time(ielement) and element_proc(ielement)
 * = rank are assigned.  Contributions to one point are added in task order (element,
 * then basis function), as the reference's loop adds them.
 *  elems           [nelems] 1-based inversion-mesh element of each task slot; -1 = empty
 *  kernel_values,  as mck_slave_work returns them ([nelems][nbasis][nkernel],
 *  kernel_variance [nelems][nbasis][nkernel], [nelems][nkernel], [nelems]), host or
 *  niterations,    device memory.  All four NULL: the results of the last mck_slave_work
 *  computation_time  call, still in device memory (no host round trip); nelems and
 *                  nbasis must be that call's.  computation_time alone may be NULL (0 s).
 *  model           [nelems][nbasis][12] model integrals when nmodel = 12 (host or device;
 *                  NULL = those of the last mck_integrate_model call, kept in device
 *                  memory); ignored when nmodel = 0
 *  hetero          [nelems][nbasis][3] heterogeneity integrals when nhet = 3, likewise
 *                  (NULL = the last mck_integrate_hetero call's); ignored when nhet = 0
 */
int mck_results_add(int32_t nelems, const int32_t *elems, const double *kernel_values,
                    const double *kernel_variance, const int32_t *niterations,
                    const double *computation_time, const double *model, const double *hetero,
                    int32_t rank);

/*
 * mck_results_add without any allocation (no counterpart in the reference): for callers
 * that cannot allocate device memory - code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc.  Same arguments and result as mck_results_add, with
 * two limits: kernel_values, kernel_variance, niterations, model and hetero are device
 * arrays or NULL (the device-resident results of the last mck_slave_work /
 * mck_integrate_model / mck_integrate_hetero), and a task holds at most 131072 (element,
 * basis function) pairs (MCK_EINVAL beyond).  Its work arrays live in one fixed 16 MiB
 * device array per GPU, kept until the process ends and shared by every call on that
 * GPU: calls that overlap on one GPU race.
 */
int mck_results_add_noalloc(int32_t nelems, const int32_t *elems, const double *kernel_values,
                            const double *kernel_variance, const int32_t *niterations,
                            const double *computation_time, const double *model, const double *hetero,
                            int32_t rank);

/* The accumulators themselves (each nullable), FP64 point-major as the device holds them:
 * kernel, variance [npoints][nkernel], model [npoints][nmodel], hetero [npoints][nhet],
 * niterations [nelements][nkernel], computation_time, element_proc [nelements]. */
int mck_results_get(double *kernel, double *variance, double *model, double *hetero,
                    int32_t *niterations, double *computation_time, int32_t *element_proc);

/* Replaces the accumulators (each nullable = unchanged), laid out as mck_results_get
 * returns them: resuming from intermediate results (read_intermediate,
 * master_queue.f90:615-659) or merging the partial sums of several devices. */
int mck_results_load(const double *kernel, const double *variance, const double *model,
                     const double *hetero, const int32_t *niterations,
                     const double *computation_time, const int32_t *element_proc);

#define MCK_RES_KERNEL    0   /* real(K_x, sp)                        [nkernel][npoints]   */
#define MCK_RES_VARIANCE  1   /* real(Var, sp)                        [nkernel][npoints]   */
#define MCK_RES_ERROR     2   /* sqrt(Var/|K_x|) where |K_x| > 1e-30, else 1000 (finalize,
                                 master_queue.f90:383-387)            [nkernel][npoints]   */
#define MCK_RES_MODEL     3   /* real(Bg_Model, sp)                   [nmodel][npoints]    */
#define MCK_RES_NITER     4   /* real(niterations, sp)                [nkernel][nelements] */
#define MCK_RES_HETERO    5   /* real(Het_Model, sp)                  [nhet][npoints]      */
/* One output variable of finalize (master_queue.f90:327-581) in single precision, laid
 * out as the output files hold it: entry-major, i.e. Fortran (npoints, nentries). */
int mck_results_export(int32_t which, float *out);

/*
 * Measured choice of the backward evaluation (the counterpart of FFTW's planner, which the
 * reference selects with FFTW_PLAN, type_parameter.f90:237, fft.f90:262-270): the
 * candidates are timed on the caller's points with the databases and kernel set of
 * mck_init, and the fastest is applied to every later call until mck_finalize.
 */
typedef struct {
    int32_t bwd_path;          /* 0 per-pair FFTs (k_bwd_eval), 1 element spectra and the */
                               /* matrix cores (k_elem_spectra + k_bwd_mfma), -1 chosen   */
                               /* per launch by the costs below (the default)             */
    int32_t mfma_waves;        /* matrix cores: 1 wave per group of 16 pairs, or 2 / 4     */
                               /* waves per element sharing its spectra through LDS        */
    int32_t mfma_xcd;          /* 1: groups in XCD-contiguous ranges                       */
    int32_t ncandidates;       /* candidates timed                                         */
    double  best_ms;           /* best time of the chosen candidate on the points          */
    double  default_ms;        /* best time of the default (1, 1, 0; 0 when the kernel set */
                               /* has no matrix-core form)                                 */
    double  candidate_ms[5];   /* best time of each candidate (bwd_path, mfma_waves,       */
                               /* mfma_xcd) = (1,1,0), (1,1,1), (1,2,0), (1,4,0), (0,1,0); */
                               /* 0 when not timed                                         */
    /* per-launch choice (bwd_path -1): the matrix cores when cost_group_us x groups of   */
    /* 16 pairs + cost_run_us x element runs < cost_pair_us x pairs, else the per-pair    */
    /* kernel; measured by mck_tune (the matrix-core and per-pair kernels' event times),  */
    /* defaults scaled from c1's at mck_init                                              */
    double  cost_group_us, cost_run_us, cost_pair_us;
    double  auto_ms;           /* best time of the per-launch choice on the points        */
} mck_tuning_t;

/*
 * mck_tune: times every candidate reps times (after one untimed run) evaluating the base
 * kernels of the npts points xyz[npts][3] (all receivers, as mck_base_kernels), measures
 * the per-launch costs from the kernels' event times, and applies the per-launch choice
 * with the fastest matrix-core variant (or a fixed candidate that beats it on the points);
 * *out describes it.  Statistics (mck_get_stats) after mck_tune are those of its runs.  Setting MCK_TUNE=1 in the environment makes mck_init
 * tune itself on 2048 points spread through the mantle.  mck_set_tuning applies a choice
 * saved from an earlier mck_tune (no timing).  Results are the same for every choice to
 * the tolerance of the path; the choice only moves time.
 */
int mck_tune(int32_t npts, const double *xyz, int32_t reps, mck_tuning_t *out);
int mck_set_tuning(const mck_tuning_t *tuning);

/* Statistics of the last mck_slave_work / mck_slave_work_plain call. */
int mck_get_stats(mck_stats_t *stats);

/* Frees device memory and resets the library (reference: freeme routines). */
void mck_finalize(void);

/* Human-readable description of the last error (thread-local, never NULL). */
const char *mck_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MCK_H */
