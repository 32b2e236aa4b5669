/*
 * mck_oracle.h - FP64 CPU restatement of MC Kernel's per-point kernel evaluation and
 * Monte Carlo integration path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP implementation:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (mc_kernel_amd) never links, loads or calls it.
 *
 * Every function follows one reference routine, cited file:line in mck_oracle.c,
 * restated operation by operation (same evaluation order, no FMA contraction:
 * built with -ffp-contract=off) so that element indices and Halton points are
 * bit-identical to the reference's.
 *
 * Parity pins (tests/test_oracle_golden.py): Halton known answers (test_halton.f90) and the
 * equality of the quad-precision and double-division forms; GLL/GLJ nodes and derivative
 * matrices (test_spectral_basis.f90); mapping / inverse mapping / inverse Jacobian /
 * inside-element known answers (test_finite_elem_mapping.f90); polynomial strain closed
 * forms (test_sem_derivatives.f90); Voigt rotation known answers and R T R^T
 * (test_rotations.f90); frame rotation geometry; azimuthal factors under rotation of the
 * source; Gabor and Butterworth golden responses (tests/gaborresponse_*.txt,
 * butterresponse*.txt), FFT identities, taper, cumsum fixtures; cut_timewindow known
 * answers and the Parseval integral (test_kernel.f90); tabulate_kernels truth table;
 * isotropic physical kernels by the chain rule; MC known answers (test_mc_integration.f90);
 * tet volume of tests/vertices.TEST; the 12-nearest order against the reference's own
 * kdtree2.f90 (oracle/_ref); the differentiable filter-convolve-window-integrate stage:
 * its forward equals ora_base_kernels bit for bit, its adjoint equals FP64 central
 * differences of the forward (exact for a function linear/quadratic in every input);
 * the heterogeneity model's nearest point against the reference's own 3-D kd-tree
 * (oracle/_ref), index and fp32 distance, and the heterogeneity integrals' closed forms
 * (tests/test_oracle_hetero.py); the result assembly's hand sums, scatter formulation and
 * error transform (tests/test_oracle_results.py, mck_oracle_results.c); in
 * tests/test_oracle_pins.py: dampen_field by the reference's known answer
 * (test_readfields.f90:1189-1217), calc_basekernel's lambda and mu by branch equality,
 * the isotropic-strain ratio K_mu = 2/3 K_lambda, symmetry and unit strains, add_stfs by
 * time-domain convolutions with the database and source STFs and a circular shift, and
 * its stop at the highest frequency, and the slave_work loop by iteration counts, draws
 * and evaluations counted by hand (max_iter exits, travel-time masks, a receiver inside
 * the element).
 *
 * Parity unpinned (restated from the reference text, no fixture or closed form checks
 * them): calc_basekernel's k_a, k_b, k_c (kernel.f90:553-567, k_c's product as written),
 * which no kernel of the benchmark sets uses; the anisotropic physical kernels vsh, vsv,
 * vph, vpv, eta (kernel.f90:745-822).
 */
#ifndef MCK_ORACLE_H
#define MCK_ORACLE_H

#include <stdint.h>
#include "../include/mck.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORA_PI 3.1415926535898   /* global_parameters.f90:40 (truncated pi) */

/* spectral basis */
void ora_gll_points(int n, double *x);
void ora_glj_points(int n, double *x);
void ora_lagrange_derivs(int n, const double *x, double *G);   /* G[i + j*(n+1)] = l_i'(x_j) */
void ora_lagrange_basis(int n, const double *pts, double x, double *l);

/* halton / geometry */
void ora_halton(int ndim, const int64_t *seed, const int32_t *base, const int64_t *leap,
                int64_t step, int npoints, double *r);       /* r[npoints][ndim] */
void ora_tet_points(const double *v, int npoints, const double *coords, double *out);
double ora_tet_volume(const double *v);                      /* v[4][3] */
int    ora_tet_weights(const double *v, int npts, const double *pts, double *w); /* w[npts][4] */
/* pseudo-random stream (stand-in for random_number) and voxels v[8][3] */
uint64_t ora_splitmix64(uint64_t z);
double ora_prng(int64_t idx, int d);
void   ora_prng_points(int64_t step, int npoints, double *r);   /* r[npoints][3] */
void   ora_vox_range(const double *v, double *range);           /* r, colat, lon min/max */
void   ora_vox_points(const double *v, int npoints, const double *coords, double *out);
double ora_vox_volume(const double *v);
void   ora_vox_center(const double *v, double *c);
int    ora_point_in_vox(const double *v, const double *p);
int    ora_point_in_tet(const double *v, const double *p);       /* v[4][3], p[3] */
double ora_norm2(int n, const double *x);                    /* gfortran NORM2 */

/* mapping */
void ora_mapping(double xi, double eta, const double *nodes, int eltype, double *sz);
void ora_inv_jacobian(double xi, double eta, const double *nodes, int eltype, double *ij);
void ora_inv_mapping(double s, double z, const double *nodes, int eltype, double *xieta);
int  ora_inside_element(double s, double z, const double *nodes, int eltype, double tol,
                        double *xi, double *eta);

/* strain at the 25 nodes of one element: u[var][jpol][ipol][nsamp] (merged block, dp) */
void ora_strain_pole(int pole, const double *u, int nsamp, int npol, const double *nodes,
                     int eltype, int axial, double *strain);  /* strain[6][jpol][ipol][nsamp] */
void ora_straintrace_pole(int pole, const double *u, int nsamp, int npol,
                          const double *nodes, int eltype, int axial, double *tr);

/* rotations */
void ora_rotate_frame(int n, const double *xyz, double lon, double colat,
                      double *s, double *phi, double *z);

/* FFT / filtering */
int  ora_nextpow2(int n);
void ora_rfft(int n, double dt, const double *x, double *X);     /* X[(n/2+1)][2]   */
void ora_irfft(int n, double df, const double *X, double *x);
void ora_taperandzeropad(int nin, int nout, int ntaper, int end_only, const double *x,
                         double *out);
void ora_cumsum_trapezoidal(int n, double dt, const double *x, double *out);
void ora_filter_tf(const mck_filter_t *f, int nfreq, double df, double *tf); /* [nfreq][2] */

/* kNN over element midpoints (fp32 squared distances, kdtree2.f90:1541-1552) */
int  ora_nearest(int nelem, const float *mp_s, const float *mp_z, double s, double z,
                 int nn, int32_t *idx, float *dist);   /* returns 1 if a tie affects order */

/* whole-path context */
typedef struct ora_ctx ora_ctx;
ora_ctx *ora_create(const mck_setup_t *setup, char *err, int errlen);
void     ora_destroy(ora_ctx *ctx);
int      ora_locate(ora_ctx *ctx, int frame, int npts, const double *xyz, int32_t *elem,
                    double *xi, double *eta, int32_t *ties);
int      ora_load_strain(ora_ctx *ctx, int frame, int npts, const double *xyz,
                         double *traces);      /* [npts][nd][nt] */
int      ora_base_kernels(ora_ctx *ctx, int npts, const double *xyz, double *base);
/* Seismogram of every receiver by reciprocity from the bwd database, seis[nrec][nt]
 * (load_seismogram / load_fw_points_rdbm, readfields.f90:1004-1048, 1469-1604); used by
 * ora_create when the set-up's seis is NULL. */
int      ora_seismograms(ora_ctx *ctx, double *seis);
/* integrate_1d_model (slave.f90:815-869) for a task: model[nelems][nbasis][12] MC
 * integrals of the background-model parameters, nsteps[nelems] (nullable) */
int      ora_integrate_model(ora_ctx *ctx, int elem_type, int nvpe, int nelems, int nbasis,
                             const double *vertices, const int32_t *conn, double *model,
                             int32_t *nsteps, int64_t *halton_step, const int64_t *halton_offsets);
/* extract_receive_buffer (master_queue.f90:253-323) for one task, serial sums into
 * K, V [npoints][nk], M [npoints][nmodel], niter_g [nelements][nk], ctime_g, proc
 * [nelements]; returns nonzero for an element or point id out of range */
int      ora_results_add(int int_type, int nbasis, int nk, int nmodel, int npoints, int nelements,
                         const int32_t *conn, int nelems, const int32_t *elems, const double *values,
                         const double *variance, const int32_t *niter, const double *ctime,
                         const double *model, int rank, double *K, double *V, double *M,
                         int32_t *niter_g, double *ctime_g, int32_t *proc);
/* finalize's sp conversions (MCK_RES_*): a [np][ncol] -> out [ncol][np]; kx for the error,
 * ia for niterations */
void     ora_results_export(int which, int np, int ncol, const double *a, const double *kx,
                            const int32_t *ia, float *out);
/* load_het_rtpv's coordinates (heterogeneities.f90:63-110), the brute-force nearest
 * model point of load_model_coeffs (heterogeneities.f90:141-188), and
 * integrate_3d_model (slave.f90:874-924) for a task: hetero[nelems][nbasis][3] */
int      ora_het_xyz(int nhet, const double *rtp, float *xyz);
int      ora_het_nearest(int nhet, const float *xyz, float qx, float qy, float qz);
int      ora_integrate_hetero(int nhet, const double *rtp, const float *dlnhet, int elem_type, int nvpe,
                              int nelems, int nbasis, const double *vertices, const int32_t *conn,
                              double *hetero, int32_t *nsteps, int64_t *halton_step,
                              const int64_t *halton_offsets);
/* Test hook: which 2 = xyz_earth -> xyz_src (rotations.f90:615-671), 3 = xyz -> src
 * (rotations.f90:423-455); t[6] Voigt, in place. */
void     ora_rotate_voigt_inv(int which, double *t, double phi, double theta);
int      ora_slave_work(ora_ctx *ctx, int elem_type, int nvpe, int nelems, int nbasis,
                        const double *vertices,
                        const int32_t *conn, double *values, double *variance,
                        int32_t *niterations, int64_t *halton_step,
                        const int64_t *halton_offsets, int64_t *offsets_used,
                        int64_t *evaluations);
/* seismogram/kernel setup products (kernel.f90:181-385) for inspection */
int      ora_kernel_info(ora_ctx *ctx, int k, double *normalization, int *ncut, int *ifirst);
int      ora_strain_nd(ora_ctx *ctx);
/* Voigt tensor rotations alone (test hook): t[6][nt] in place */
void     ora_rotate_voigt(int which, double *t, int nt, double phi, double theta);
/* physical kernels alone (test hook) */
double   ora_physical(int param, const double *b, float vp, float vs, float rho, float phi,
                      float xi, float eta, int rel);
/* tabulate_kernels alone (test hook): nb = lambda, mu, rho, a, b, c */
int      ora_param_needs(int param, int32_t *nb);
/* azim_factor alone (test hook): isim 1..4, ikind 1/2 */
double   ora_azim_factor(double phi, const double *mij, int isim, int ikind);
/* cut_timewindow on a time axis t[n] (test hook and kernel set-up) */
void     ora_cut_window(int n, const double *t, double w0, double w1, int *ifirst, int *ncut);
/* calc_basekernel alone (test hook): F, B [nd][nw][2] -> out [nw][2]; ibk 0..5 */
void     ora_basekernel(int ibk, int full, int nd, int nw, const double *F, const double *B, double *out);
/* dampen_field alone (test hook): base[npts][nk][6] scaled by |x - c| / rmax inside rmax */
void     ora_dampen_field(double *base, int npts, int nk, const double *pts, const double *c, double rmax);
/* add_stfs alone (test hook): tff in = filter response, out = forward; tfb out = backward;
 * stf_src NULL = unit impulse; -1 where the reference stops (response at the highest
 * frequency above 1e-2 of its maximum) */
int      ora_add_stfs(int nt, double dt, int N, const double *stf_sem, double amplitude, const double *stf_src,
                      double stf_shift, double *tff, double *tfb, char *err, int errlen);
/* integrate_parseval alone (test hook) */
double   ora_integrate_parseval(int n, double dt, const double *f, const double *g);
/* the Monte Carlo integrator alone (test hook): vals[nsteps][npts][n] */
int      ora_mc_run(int n, double volume, double allowed_error, double allowed_rel, int nsteps,
                    int npts, const double *vals, double *integral, double *variance,
                    int32_t *conv);
/* The filter-convolve-window-integrate stage as a function of the filter responses
 * G [nfilter][nomega][2] and the strain traces fw [npts][nd][nt] (before
 * cumsum_trapezoidal), bw [npts][nrec][nd][nt]: base [npts][nkernel][6] as
 * ora_base_kernels, and its FP64 adjoint for the weights gbase [npts][nkernel][6]:
 * gG as d/dRe + i d/dIm, gfw, gbw.  ora_ad_transfer returns the set-up's own G. */
void     ora_ad_transfer(ora_ctx *ctx, double *G);
int      ora_ad_forward(ora_ctx *ctx, int npts, const double *G, const double *fw,
                        const double *bw, double *base);
int      ora_ad_backward(ora_ctx *ctx, int npts, const double *G, const double *fw,
                         const double *bw, const double *gbase, double *gG, double *gfw,
                         double *gbw);
/* Optional on-demand generation of database elements (which: 0 fwd, 1 bwd): the
 * snapshots arrays must then be writable; used for bounded CPU-baseline samples. */
typedef void (*ora_fill_fn)(int which, int elem, float *dst, void *user);
void     ora_set_lazy_fill(ora_ctx *ctx, ora_fill_fn fn, void *user);

#ifdef __cplusplus
}
#endif
#endif
