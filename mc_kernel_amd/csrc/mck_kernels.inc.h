// mck_kernels.hip - HIP kernels of the MC Kernel per-point kernel-evaluation and Monte
// Carlo integration path for gfx950.
//
//  k_points       Halton points in tetrahedra (halton_sequence.f90:109-153,
//                 polytopes.f90:47-115) and the radial model index of every point
//  k_locate       point location in one frame: rotation into the source/receiver frame
//                 (rotations.f90:737-782), clamp to the surface, exact 12 nearest element
//                 midpoints with single-precision distances (kdtree2.f90:1541-1552) and
//                 the Newton inverse mappings (finite_elem_mapping.f90) - bit-identical
//                 element indices
//  k_fwd_traces   gather + strain + interpolation + azimuthal factors of the forward
//                 field (sem_derivatives.f90, spectral_basis.f90:181-314, readfields.f90:
//                 1138-1465) on the matrix cores, the points of one element together
//  k_fwd_spectra  cumsum (simple_routines.f90:314-341), taper + zero pad (fft.f90:
//                 546-601) and LDS Stockham real FFT (replacing FFTW in fft.f90:375-413)
//                 of those traces, points in element order
//  k_bwd_eval     the same for the backward field of every active (point, receiver) pair,
//                 fused with the fwd x bwd base kernels (kernel.f90:514-570), the filtered
//                 Parseval window integral (kernel.f90:410-510, 946-1020) in the frequency
//                 domain with FP64 accumulation and the physical kernels (kernel.f90:
//                 665-832)
//  k_strain_traces  earth-frame strain time series of one frame (point-level API)
//  k_mc           per-voxel MC sums, variance and convergence (mc_integration.f90:69-143)
//                 and the iteration bookkeeping of slave_work (slave.f90:501-757)
// Included once by mck_api.hip (single translation unit).
#pragma once
#include "mck_device.h"

namespace mck {

__constant__ Basis c_basis;

// ------------------------------------------------------------------------------------
// element mappings (finite_elem_mapping.f90); E = element constants:
//  E[0..3] corner s, E[4..7] corner z (fp32 values), spheroid: E[8..11] r, E[12..15]
//  theta; semino/semiso: E[8] a, E[9] b, E[10] theta_bar, E[11] dtheta.
// ------------------------------------------------------------------------------------
// The snapshots [nvar][5][5][nt] of element e in the view db, or null when the view does
// not hold it (a point not located, an element of another page)
__device__ inline const float *elem_u(const DevDb &db, int e)
{
    const size_t stride = (size_t)db.nvar * 25 * db.nt;
    if (db.page_u) {
        if (e < 0 || e >= db.nelem) return nullptr;
        const float *b = db.page_u[e / db.page_elems];
        return b ? b + (size_t)(e % db.page_elems) * stride : nullptr;
    }
    if (e < db.u_first || e >= db.u_last) return nullptr;
    return db.u + (size_t)(e - db.u_first) * stride;
}

__device__ inline void d_mapping(double xi, double eta, const double *E, int type, double *sz)
{
    if (type == 0) {
        const double *r = E + 8, *th = E + 12;
        sz[0] = (1 + eta) * r[3] / 2 * sin(((1 - xi) * th[3] + (1 + xi) * th[2]) / 2)
              + (1 - eta) * r[0] / 2 * sin(((1 - xi) * th[0] + (1 + xi) * th[1]) / 2);
        sz[1] = (1 + eta) * r[3] / 2 * cos(((1 - xi) * th[3] + (1 + xi) * th[2]) / 2)
              + (1 - eta) * r[0] / 2 * cos(((1 - xi) * th[0] + (1 + xi) * th[1]) / 2);
    } else if (type == 1) {
        double xip = 1 + xi, xim = 1 - xi, etap = 1 + eta, etam = 1 - eta;
        double shp0 = xim * etam / 4, shp1 = xip * etam / 4, shp2 = xip * etap / 4, shp3 = xim * etap / 4;
        double s = 0.0, z = 0.0;
        s = s + shp0 * E[0]; z = z + shp0 * E[4];
        s = s + shp1 * E[1]; z = z + shp1 * E[5];
        s = s + shp2 * E[2]; z = z + shp2 * E[6];
        s = s + shp3 * E[3]; z = z + shp3 * E[7];
        sz[0] = s; sz[1] = z;
    } else if (type == 2) {
        double a = E[8], b = E[9], tb = E[10], dt = E[11];
        double sb = ((1 + xi) * E[1] + (1 - xi) * E[0]) / 2;
        double zb = ((1 + xi) * E[5] + (1 - xi) * E[4]) / 2;
        double st = a * cos(tb + xi * dt / 2);
        double zt = b * sin(tb + xi * dt / 2);
        sz[0] = (sb + st) / 2 + eta * (st - sb) / 2;
        sz[1] = (zb + zt) / 2 + eta * (zt - zb) / 2;
    } else {
        double a = E[8], b = E[9], tb = E[10], dt = E[11];
        double st = ((1 + xi) * E[2] + (1 - xi) * E[3]) / 2;
        double zt = ((1 + xi) * E[6] + (1 - xi) * E[7]) / 2;
        double sb = a * cos(tb + xi * dt / 2);
        double zb = b * sin(tb + xi * dt / 2);
        sz[0] = (sb + st) / 2 + eta * (st - sb) / 2;
        sz[1] = (zb + zt) / 2 + eta * (zt - zb) / 2;
    }
}

// inverse Jacobian, column-major: ij[0]=(1,1) ij[1]=(2,1) ij[2]=(1,2) ij[3]=(2,2)
__device__ inline void d_inv_jacobian(double xi, double eta, const double *E, int type, double *ij)
{
    double J0, J1, J2, J3;
    if (type == 0) {
        const double *r = E + 8, *th = E + 12;
        J0 = (1 + eta) * r[3] * (th[2] - th[3]) / 4 * cos(((1 - xi) * th[3] + (1 + xi) * th[2]) / 2)
           + (1 - eta) * r[0] * (th[1] - th[0]) / 4 * cos(((1 - xi) * th[0] + (1 + xi) * th[1]) / 2);
        J2 = 0.5 * (r[3] * sin(((1 - xi) * th[3] + (1 + xi) * th[2]) / 2)
                  - r[0] * sin(((1 - xi) * th[0] + (1 + xi) * th[1]) / 2));
        J1 = -(1 + eta) * r[3] * (th[2] - th[3]) / 4 * sin(((1 - xi) * th[3] + (1 + xi) * th[2]) / 2)
           - (1 - eta) * r[0] * (th[1] - th[0]) / 4 * sin(((1 - xi) * th[0] + (1 + xi) * th[1]) / 2);
        J3 = 0.5 * (r[3] * cos(((1 - xi) * th[3] + (1 + xi) * th[2]) / 2)
                  - r[0] * cos(((1 - xi) * th[0] + (1 + xi) * th[1]) / 2));
    } else if (type == 1) {
        double xip = 1 + xi, xim = 1 - xi, etap = 1 + eta, etam = 1 - eta;
        double d0[4] = {-etam / 4, etam / 4, etap / 4, -etap / 4};
        double d1[4] = {-xim / 4, -xip / 4, xip / 4, xim / 4};
        J0 = J1 = J2 = J3 = 0.0;
        for (int i = 0; i < 4; ++i) {
            J0 = J0 + E[i] * d0[i];
            J2 = J2 + E[i] * d1[i];
            J1 = J1 + E[4 + i] * d0[i];
            J3 = J3 + E[4 + i] * d1[i];
        }
    } else {
        double a = E[8], b = E[9], tb = E[10], dt = E[11];
        double sc = a * cos(tb + xi * dt / 2), zc = b * sin(tb + xi * dt / 2);
        double dsc = -a * dt * sin(tb + xi * dt / 2) / 2;
        double dzc = b * dt * cos(tb + xi * dt / 2) / 2;
        double sb, zb, st, zt, dsb, dzb, dst, dzt;
        if (type == 2) {
            sb = ((1 + xi) * E[1] + (1 - xi) * E[0]) / 2;
            zb = ((1 + xi) * E[5] + (1 - xi) * E[4]) / 2;
            st = sc; zt = zc;
            dsb = (E[1] - E[0]) / 2;
            dzb = (E[5] - E[4]) / 2;
            dst = dsc; dzt = dzc;
        } else {
            st = ((1 + xi) * E[2] + (1 - xi) * E[3]) / 2;
            zt = ((1 + xi) * E[6] + (1 - xi) * E[7]) / 2;
            sb = sc; zb = zc;
            dsb = dsc; dzb = dzc;
            dst = (E[2] - E[3]) / 2;
            dzt = (E[6] - E[7]) / 2;
        }
        J0 = (dsb + dst) / 2 + eta * (dst - dsb) / 2;
        J2 = (st - sb) / 2;
        J1 = (dzb + dzt) / 2 + eta * (dzt - dzb) / 2;
        J3 = (zt - zb) / 2;
    }
    double det = J0 * J3 - J1 * J2;
    ij[0] = J3 / det;
    ij[1] = -J1 / det;
    ij[2] = -J2 / det;
    ij[3] = J0 / det;
}

// inv_mapping_* + inside_element (finite_elem_mapping.f90:101-141, 137-200, 403-433,
// 505-535): 10 Newton steps from the centre; single-precision threshold 1e-7**2 except
// for spheroids (1d-7**2).
__device__ inline bool d_inside(double s, double z, const double *E, int type, double *xi_out,
                                double *eta_out)
{
    const double thr = (type == 0) ? (1e-7 * 1e-7) : (double)(1e-7f * 1e-7f);
    double xi = 0.0, eta = 0.0, sz[2], ij[4];
    for (int i = 1; i <= 10; ++i) {
        d_mapping(xi, eta, E, type, sz);
        double ds = s - sz[0], dz = z - sz[1];
        if ((ds * ds + dz * dz) / (s * s + z * z) < thr) break;
        d_inv_jacobian(xi, eta, E, type, ij);
        xi = xi + ij[0] * ds + ij[2] * dz;
        eta = eta + ij[1] * ds + ij[3] * dz;
    }
    *xi_out = xi;
    *eta_out = eta;
    const double tol = 1e-3;
    return (xi >= -1 - tol && xi <= 1 + tol && eta >= -1 - tol && eta <= 1 + tol);
}

// gfortran NORM2 of a 3-vector (scaled sum of squares)
__device__ inline double d_norm2(double a, double b, double c)
{
    double v[3] = {a, b, c}, scale = 0.0, res = 1.0;
    for (int i = 0; i < 3; ++i) {
        if (v[i] != 0.0) {
            double ax = fabs(v[i]);
            if (scale < ax) { double q = scale / ax; res = 1.0 + res * q * q; scale = ax; }
            else { double q = ax / scale; res = res + q * q; }
        }
    }
    return sqrt(res) * scale;
}

// ------------------------------------------------------------------------------------
// k_points: one thread per MC point of the working voxels of this chunk
// ------------------------------------------------------------------------------------
// Pseudo-random stream standing in for Fortran's random_number (which the reference
// seeds from the clock): value d of draw idx is the SplitMix64 output of the counter
// 0x5EED0000 + (3 idx + d + 1) * 0x9E3779B97F4A7C15 (Steele, Lea and Flood 2014), top 53
// bits scaled to [0, 1).  Draw positions are assigned as the Halton positions are.
__device__ inline double d_prng(int64_t idx, int d)
{
    uint64_t z = 0x5EED0000ULL + ((uint64_t)idx * 3u + (uint64_t)d + 1u) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}
// Point `idx` of an element's stream mapped into its geometry V (a tetrahedron's 4
// vertices, or a voxel's range): Halton (quasi, tetrahedra) or the pseudo-random
// stream in [0, 1)^3, then generate_random_points_vox (random_voxel.f90:45-86) or the
// Rocchini-Cignoni fold (polytopes.f90:88-110).
__device__ inline void draw_point(int64_t idx, const double *V, int vox, int quasi, double *x)
{
    double c[4];
    if (quasi && !vox) {
        const int base[3] = {2, 3, 5};
        for (int d = 0; d < 3; ++d) {
            int64_t s2 = 20 + idx * 2;
            double bk = (double)base[d];
            double acc = 0.0;
            while (s2 != 0) {
                int64_t digit = s2 % base[d];
                acc = acc + (double)digit / bk;   // = dp(digit * qp(1/b^k)), see halton tests
                bk = bk * (double)base[d];
                s2 = s2 / base[d];
            }
            c[d] = acc;
        }
    } else {
        for (int d = 0; d < 3; ++d) c[d] = d_prng(idx, d);
    }
    if (vox) {
        // generate_random_points_vox (random_voxel.f90:45-86)
        const double r = pow(c[0] * (V[1] * V[1] * V[1] - V[0] * V[0] * V[0]) + V[0] * V[0] * V[0], 1.0 / 3.0);
        const double th = acos(cos(V[2]) + (cos(V[3]) - cos(V[2])) * c[1]);
        const double ph = V[4] + fabs(V[5] - V[4]) * c[2];
        x[0] = r * cos(ph) * sin(th);
        x[1] = r * sin(ph) * sin(th);
        x[2] = r * cos(th);
    } else {
        // Rocchini-Cignoni fold (polytopes.f90:88-110)
        if (c[0] + c[1] > 1.0) { c[0] = 1.0 - c[0]; c[1] = 1.0 - c[1]; }
        if (c[1] + c[2] > 1.0) {
            double t = c[2];
            c[2] = 1.0 - c[0] - c[1];
            c[1] = 1.0 - t;
        } else if (c[0] + c[1] + c[2] > 1.0) {
            double t = c[2];
            c[2] = c[0] + c[1] + c[2] - 1.0;
            c[0] = 1.0 - c[1] - t;
        }
        c[3] = 1.0 - (c[0] + c[1] + c[2]);
        for (int i = 0; i < 3; ++i) {
            double a = 0.0;
            a = a + V[0 * 3 + i] * c[0];
            a = a + V[1 * 3 + i] * c[1];
            a = a + V[2 * 3 + i] * c[2];
            a = a + V[3 * 3 + i] * c[3];
            x[i] = a;
        }
    }
}

// linear basis function b of tetrahedron V at x from its barycentric inverse A
// (inversion_mesh.f90:554-563)
__device__ inline double tet_bary(const double *A, const double *V, const double *x, int b)
{
    const double dx = x[0] - V[0], dy = x[1] - V[1], dz = x[2] - V[2];
    return A[4 * b] * dx + A[4 * b + 1] * dy + A[4 * b + 2] * dz + A[4 * b + 3] * 1.0;
}


// With nbasis 4 it also evaluates the vertex basis functions at the point (weights,
// inversion_mesh.f90:530-622): bweights[point][4] from abinv [nvox][4][4].
// geom: tetrahedra [nvox][4][3] vertices; voxels [nvox][12] whose first 6 values are the
// spherical range r_min, r_max, colat_min, colat_max, lon_min, lon_max.
// Tetrahedra take Halton (quasi) or pseudo-random numbers, voxels pseudo-random ones
// (generate_random_points, inversion_mesh.f90:626-667).
__global__ void k_points(int nwork, int P, const int *work_voxel, const int64_t *halton_base,
                         const int *voxel_draws, const double *geom, int vox, int quasi,
                         int ndepth, float dr, double *xyz, int *model_idx, int nbasis,
                         const double *abinv, double *bweights, int *err)
{
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nwork * P) return;
    int w = gid / P, p = gid - w * P;
    int v = work_voxel[w];
    // the current draw of voxel v is its (voxel_draws[v] - 1)-th
    int64_t idx = halton_base[v] + (int64_t)(voxel_draws[v] - 1) * P + p;
    const double *V = geom + (size_t)v * 12;
    double x[3];
    draw_point(idx, V, vox, quasi, x);
    xyz[3 * (size_t)gid + 0] = x[0];
    xyz[3 * (size_t)gid + 1] = x[1];
    xyz[3 * (size_t)gid + 2] = x[2];
    double r = d_norm2(x[0], x[1], x[2]);
    int id = (int)(r / (double)dr + 1);
    id = id < 1 ? 1 : (id > ndepth ? ndepth : id);
    model_idx[gid] = id - 1;
    if (nbasis > 1 && !vox) {
        const double *A = abinv + (size_t)v * 16;
        bool bad = false;
        for (int b = 0; b < 4; ++b) {
            const double w = tet_bary(A, V, x, b);
            bweights[(size_t)gid * 4 + b] = w;
            bad |= (w < -1e-9 || w > 1.0 + 1e-9);
        }
        if (bad) atomicOr(err, 4);
    }
}

// ------------------------------------------------------------------------------------
// k_model_integral: integrate_1d_model (slave.f90:815-869), one workgroup per element
// running its whole Monte Carlo loop: 100 points per step (Halton in tetrahedra whatever
// quasirandom says, the pseudo-random stream in voxels), the 12 background-model
// parameters of the radial table (combine(), background_model.f90:122-198, fp32; order
// vp vs rho vph vpv vsh vsv eta phi xi lam mu) times the basis function's weight, one
// integrator per (basis function, parameter) with volume 1 and errors 1e-2
// (check_montecarlo_integral, mc_integration.f90:69-143: the 100-point sums in point
// order, as the reference's sum()), until every integrator has converged or after 9999
// steps.  base[e]: stream position of the element's first point.  out[e][nbasis][12],
// nsteps[e]; err bit 4: a basis weight outside [0, 1], bit 8: NaN.
// ------------------------------------------------------------------------------------
constexpr int kModelPts = 100, kModelPar = 12, kModelMaxStep = 9999;

__global__ __launch_bounds__(128) void k_model_integral(int nelem, const double *geom, int vox, int nbasis,
                                                        const double *abinv, const int64_t *base,
                                                        const uint8_t *live, const float *model, int ndepth,
                                                        float dr, double *out, int *nsteps, int *err)
{
    __shared__ double val[kModelPts][kModelPar];
    __shared__ double wt[kModelPts][8];
    __shared__ int nconv;
    const int e = blockIdx.x, tid = threadIdx.x;
    if (e >= nelem) return;
    const int nint = nbasis * kModelPar;
    const bool owner = tid < nint;
    const int b = tid / kModelPar, k = tid - (tid / kModelPar) * kModelPar;
    if (!live[e]) {
        if (owner) out[((size_t)e * nbasis + b) * kModelPar + k] = 0.0;
        if (tid == 0) nsteps[e] = 0;
        return;
    }
    const double *V = geom + (size_t)e * 12;
    double fsum = 0.0, f2sum = 0.0, integ = 0.0;
    int64_t nmodels = 0;
    bool conv = false;
    int istep = 0;
    for (;;) {
        // allallconverged (mc_integration.f90:322-340)
        if (tid == 0) nconv = 0;
        __syncthreads();
        if (owner && conv) atomicAdd(&nconv, 1);
        __syncthreads();
        if (nconv == nint) break;
        if (tid < kModelPts) {
            double x[3];
            draw_point(base[e] + (int64_t)istep * kModelPts + tid, V, vox, 1, x);
            const double r = d_norm2(x[0], x[1], x[2]);
            int id = (int)(r / (double)dr + 1);
            id = (id < 1 ? 1 : (id > ndepth ? ndepth : id)) - 1;
            const float vp = model[0 * ndepth + id], vs = model[1 * ndepth + id], rho = model[2 * ndepth + id];
            const float phi = model[3 * ndepth + id], xi = model[4 * ndepth + id], eta = model[5 * ndepth + id];
            const float lam = rho * (vp * vp - 2 * (vs * vs)), mu = rho * (vs * vs);
            const float vsh = vs, vph = vp, vsv = vsh / sqrtf(xi), vpv = vph * sqrtf(phi);
            const float c[kModelPar] = {vp, vs, rho, vph, vpv, vsh, vsv, eta, phi, xi, lam, mu};
#pragma unroll
            for (int j = 0; j < kModelPar; ++j) val[tid][j] = (double)c[j];
            bool bad = false;
            for (int bb = 0; bb < nbasis; ++bb) {
                double w = 1.0;
                if (nbasis > 1 && !vox) {
                    w = tet_bary(abinv + (size_t)e * 16, V, x, bb);
                    bad |= (w < -1e-9 || w > 1.0 + 1e-9);
                }
                wt[tid][bb] = w;
            }
            if (bad) atomicOr(err, 4);
        }
        __syncthreads();
        if (owner) {
            nmodels += kModelPts;
            if (!conv) {
                double s1 = 0.0, s2 = 0.0;
#pragma unroll 4
                for (int i = 0; i < kModelPts; ++i) s1 += val[i][k] * wt[i][b];
#pragma unroll 4
                for (int i = 0; i < kModelPts; ++i) {
                    const double f = val[i][k] * wt[i][b];
                    s2 += f * f;
                }
                fsum = fsum + s1;
                f2sum = f2sum + s2;
                integ = fsum * 1.0 / (double)nmodels;
                double var;
                if (nmodels > 2) {
                    var = 1.0 * 1.0 / (double)(nmodels - 1)
                        * ((f2sum / (double)nmodels) - (fsum / (double)nmodels) * (fsum / (double)nmodels));
                    if (var < 1e-2 * 1e-2) conv = true;
                    else if (sqrt(var) < 1e-2 * fabs(integ)) conv = true;
                } else {
                    var = integ * 1e10;
                }
                if (integ != integ || var != var) atomicOr(err, 8);
            }
        }
        __syncthreads();
        ++istep;
        if (istep >= kModelMaxStep) break;
    }
    if (owner) out[((size_t)e * nbasis + b) * kModelPar + k] = integ;
    if (tid == 0) nsteps[e] = istep;
}

// ------------------------------------------------------------------------------------
// k_window_spectra: the frequency-domain window integral of every kernel's filtered
// seismogram window (cut_and_add_seismogram and the misfit normalisation,
// kernel.f90:181-385).  One thread per (distinct window h, bin q):
//   S = sum_n w_h[n] e^{2 pi i q (ifirst_h + n) / N}   (FP64, samples in order, angles
//       from an exact table of 2 pi m / N)
//   H = -norm df c_q S TF_fwd(q) TF_bwd(q) ts(q),  W = -norm df c_q S rest(q) ts(q)
// with c_q = 1 at DC and Nyquist, else 2; both stored in single precision, and H once
// more as (re, -im) in Hm (rows fs apart; nullable).
// ------------------------------------------------------------------------------------
__device__ inline double2 cmul_d(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__global__ __launch_bounds__(256) void k_window_spectra(int nH, int nw, int N, double df, const double *w_all,
                                                         const int *w_off, const int *ifirst, const int *ncut,
                                                         const double *norm, const int *filt,
                                                         const double *cosT, const double *sinT,
                                                         const double2 *tff, const double2 *tfb,
                                                         const double2 *ts, const double2 *rest, float2 *H,
                                                         float2 *W, float2 *Hm, int fs)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nH * nw) return;
    const int h = (int)(i / nw), q = (int)(i - (long long)h * nw);
    const double *w = w_all + w_off[h];
    int m = (int)(((long long)q * ifirst[h]) % N);
    double sr = 0.0, si = 0.0;
    for (int n = 0; n < ncut[h]; ++n) {
        sr += w[n] * cosT[m];
        si += w[n] * sinT[m];
        m += q;
        if (m >= N) m -= N;
    }
    const double ck = (q == 0 || q == nw - 1) ? 1.0 : 2.0;
    const double sc = -norm[h] * df * ck;
    const double2 S = make_double2(sc * sr, sc * si);
    const int f = filt[h];
    const double2 hv = cmul_d(cmul_d(cmul_d(S, tff[(size_t)f * nw + q]), tfb[(size_t)f * nw + q]), ts[q]);
    const double2 wv = cmul_d(cmul_d(S, rest[q]), ts[q]);
    H[i] = make_float2((float)hv.x, (float)hv.y);
    W[i] = make_float2((float)wv.x, (float)wv.y);
    if (Hm) Hm[(size_t)h * fs + q] = make_float2((float)hv.x, -(float)hv.y);
}

// ------------------------------------------------------------------------------------
// k_hetero_integral: integrate_3d_model (slave.f90:874-924), one wave per element.
// The heterogeneity model is a point cloud (load_het_rtpv, heterogeneities.f90:63-110)
// held in a uniform grid of cells; each random point takes the parameters of its nearest
// model point (load_model_coeffs, heterogeneities.f90:141-188: kdtree2_n_nearest with
// nn = 1 on single-precision coordinates, distance ((dx^2 + dy^2) + dz^2) in fp32 as
// kdtree2.f90:1541-1543 sums it).  Equal distances take the lower model index.  10
// points per step, 3 parameters (dlnvp, dlnvs, dlnrho) times the basis function, one
// integrator per (basis, parameter) with volume 1, allowed error 1e-2 and relative error
// 1e-1, at most 9999 steps.
// ------------------------------------------------------------------------------------
constexpr int kHetPts = 10, kHetPar = 3, kHetMaxStep = 9999;

struct HetGrid {
    int n, nx, ny, nz;
    double x0, y0, z0, h;            // grid origin and cell size (m)
    const float *xyz;                // [n][3] model points sorted by cell (fp32)
    const int *idx;                  // [n] model index of each sorted point
    const int *cell_start;           // [nx * ny * nz + 1]
    const float *dln;                // [n][3] dlnvp, dlnvs, dlnrho by model index
};

__device__ inline int het_cell(double q, double o, double h, int n) {
    int c = (int)floor((q - o) / h);
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

// Exact nearest model point: shells of cells of growing Chebyshev radius around the
// query's cell, until every cell not yet searched lies farther than the best distance.
__device__ inline int het_nearest(const HetGrid &g, float qx, float qy, float qz) {
    const int cx = het_cell(qx, g.x0, g.h, g.nx), cy = het_cell(qy, g.y0, g.h, g.ny),
              cz = het_cell(qz, g.z0, g.h, g.nz);
    float best = 3.0e38f;
    int bi = 0x7fffffff;
    for (int k = 0;; ++k) {
        const int zlo = max(cz - k, 0), zhi = min(cz + k, g.nz - 1);
        const int ylo = max(cy - k, 0), yhi = min(cy + k, g.ny - 1);
        const int xlo = max(cx - k, 0), xhi = min(cx + k, g.nx - 1);
        for (int z = zlo; z <= zhi; ++z)
            for (int y = ylo; y <= yhi; ++y) {
                const bool face = abs(z - cz) == k || abs(y - cy) == k;
                for (int x = xlo; x <= xhi; x += (face || k == 0) ? 1 : max(1, xhi - xlo)) {
                    if (!face && abs(x - cx) != k) continue;
                    const int c = (z * g.ny + y) * g.nx + x;
                    for (int j = g.cell_start[c]; j < g.cell_start[c + 1]; ++j) {
                        const float dx = g.xyz[3 * j] - qx, dy = g.xyz[3 * j + 1] - qy, dz = g.xyz[3 * j + 2] - qz;
                        const float d = (dx * dx + dy * dy) + dz * dz;
                        const int id = g.idx[j];
                        if (d < best || (d == best && id < bi)) { best = d; bi = id; }
                    }
                }
            }
        // distance from the query to the faces of the searched block that are inside the
        // grid; beyond them lie the cells not searched yet
        double m = 1e300;
        if (cx - k > 0) m = fmin(m, (double)qx - (g.x0 + (cx - k) * g.h));
        if (cx + k < g.nx - 1) m = fmin(m, g.x0 + (cx + k + 1) * g.h - (double)qx);
        if (cy - k > 0) m = fmin(m, (double)qy - (g.y0 + (cy - k) * g.h));
        if (cy + k < g.ny - 1) m = fmin(m, g.y0 + (cy + k + 1) * g.h - (double)qy);
        if (cz - k > 0) m = fmin(m, (double)qz - (g.z0 + (cz - k) * g.h));
        if (cz + k < g.nz - 1) m = fmin(m, g.z0 + (cz + k + 1) * g.h - (double)qz);
        if (m >= 1e300) break;                               // the block covers the grid
        if (bi != 0x7fffffff && m > 0.0 && m * m * (1.0 - 1e-5) > (double)best) break;
    }
    return bi;
}

__global__ __launch_bounds__(64) void k_hetero_integral(int nelem, const double *geom, int vox, int nbasis,
                                                         const double *abinv, const int64_t *base,
                                                         const uint8_t *live, HetGrid g, double *out,
                                                         int *nsteps, int *err)
{
    __shared__ double val[kHetPts][kHetPar];
    __shared__ double wt[kHetPts][8];
    __shared__ int nconv;
    const int e = blockIdx.x, tid = threadIdx.x;
    if (e >= nelem) return;
    const int nint = nbasis * kHetPar;
    const bool owner = tid < nint;
    const int b = tid / kHetPar, k = tid - (tid / kHetPar) * kHetPar;
    if (!live[e]) {
        if (owner) out[((size_t)e * nbasis + b) * kHetPar + k] = 0.0;
        if (tid == 0) nsteps[e] = 0;
        return;
    }
    const double *V = geom + (size_t)e * 12;
    double fsum = 0.0, f2sum = 0.0, integ = 0.0;
    int64_t nmodels = 0;
    bool conv = false;
    int istep = 0;
    for (;;) {
        // allallconverged (mc_integration.f90:322-340)
        if (tid == 0) nconv = 0;
        __syncthreads();
        if (owner && conv) atomicAdd(&nconv, 1);
        __syncthreads();
        if (nconv == nint) break;
        if (tid < kHetPts) {
            double x[3];
            draw_point(base[e] + (int64_t)istep * kHetPts + tid, V, vox, 1, x);
            const int j = het_nearest(g, (float)x[0], (float)x[1], (float)x[2]);
#pragma unroll
            for (int q = 0; q < kHetPar; ++q) val[tid][q] = (double)g.dln[3 * (size_t)j + q];
            bool bad = false;
            for (int bb = 0; bb < nbasis; ++bb) {
                double w = 1.0;
                if (nbasis > 1 && !vox) {
                    w = tet_bary(abinv + (size_t)e * 16, V, x, bb);
                    bad |= (w < -1e-9 || w > 1.0 + 1e-9);
                }
                wt[tid][bb] = w;
            }
            if (bad) atomicOr(err, 4);
        }
        __syncthreads();
        if (owner) {
            nmodels += kHetPts;
            if (!conv) {
                double s1 = 0.0, s2 = 0.0;
                for (int i = 0; i < kHetPts; ++i) s1 += val[i][k] * wt[i][b];
                for (int i = 0; i < kHetPts; ++i) {
                    const double f = val[i][k] * wt[i][b];
                    s2 += f * f;
                }
                fsum = fsum + s1;
                f2sum = f2sum + s2;
                integ = fsum * 1.0 / (double)nmodels;
                double var;
                if (nmodels > 2) {
                    var = 1.0 * 1.0 / (double)(nmodels - 1)
                        * ((f2sum / (double)nmodels) - (fsum / (double)nmodels) * (fsum / (double)nmodels));
                    if (var < 1e-2 * 1e-2) conv = true;
                    else if (sqrt(var) < 1e-1 * fabs(integ)) conv = true;
                } else {
                    var = integ * 1e10;
                }
                if (integ != integ || var != var) atomicOr(err, 8);
            }
        }
        __syncthreads();
        ++istep;
        if (istep >= kHetMaxStep) break;
    }
    if (owner) out[((size_t)e * nbasis + b) * kHetPar + k] = integ;
    if (tid == 0) nsteps[e] = istep;
}

// ------------------------------------------------------------------------------------
// k_locate: one thread per (point, frame).  Frame 0 = source, 1 + r = receiver r.
// active (nullable) gates bwd frames: active[slot_of_point * R + r].
// ------------------------------------------------------------------------------------
__device__ inline void knn_insert(float d, int e, float *bd, int *bi, int &cnt)
{
    const int K = kNN + 1;
    if (cnt == K && !(d < bd[K - 1] || (d == bd[K - 1] && e < bi[K - 1]))) return;
    int p = (cnt < K) ? cnt++ : K - 1;
    while (p > 0 && (bd[p - 1] > d || (bd[p - 1] == d && bi[p - 1] > e))) {
        bd[p] = bd[p - 1];
        bi[p] = bi[p - 1];
        --p;
    }
    bd[p] = d;
    bi[p] = e;
}

__global__ void k_locate(DevDb db, const DevFrame *frames, int frame0, int nframes, int npts,
                         const double *xyz, double planet_radius, const uint8_t *active,
                         int P, int R, Located *out, int *err)
{
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= npts * nframes) return;
    int p = gid / nframes, f = frame0 + (gid - (gid / nframes) * nframes);
    Located res;
    res.elem = -1; res.tie = 0; res.xi = 0; res.eta = 0; res.phi = 0;
    if (f > 0 && active && !active[(p / P) * R + (f - 1)]) {
        out[gid] = res;
        return;
    }
    const DevFrame fr = frames[f];
    double x = xyz[3 * (size_t)p], y = xyz[3 * (size_t)p + 1], zz = xyz[3 * (size_t)p + 2];
    // rotate_frame_rd
    double xc = x * fr.cos_lon + y * fr.sin_lon;
    double yc = -x * fr.sin_lon + y * fr.cos_lon;
    double zc = zz;
    double xp = xc * fr.cos_colat - zc * fr.sin_colat;
    double yp = yc;
    double zp = xc * fr.sin_colat + zc * fr.cos_colat;
    double s = sqrt(xp * xp + yp * yp), z = zp;
    double pc = atan2(yp, xp);
    res.phi = (pc < 0.0) ? 2.0 * 3.1415926535898 + pc : pc;
    double d2 = s * s + z * z, Rp = planet_radius;
    if (d2 > Rp * Rp) { s = s / sqrt(d2) * Rp; z = z / sqrt(d2) * Rp; }
    // exact 13 nearest midpoints (12 tested, the 13th detects ties at the boundary)
    float qs = (float)s, qz = (float)z;
    float bd[kNN + 1];
    int bi[kNN + 1], cnt = 0;
    int cx = (int)floor((s - db.gs0) / db.gh), cz = (int)floor((z - db.gz0) / db.gh);
    cx = cx < 0 ? 0 : (cx >= db.gnx ? db.gnx - 1 : cx);
    cz = cz < 0 ? 0 : (cz >= db.gnz ? db.gnz - 1 : cz);
    for (int ring = 0;; ++ring) {
        int x0 = cx - ring, x1 = cx + ring, z0 = cz - ring, z1 = cz + ring;
        for (int iz = z0; iz <= z1; ++iz) {
            if (iz < 0 || iz >= db.gnz) continue;
            bool edge_z = (iz == z0 || iz == z1);
            for (int ix = x0; ix <= x1; ++ix) {
                if (ix < 0 || ix >= db.gnx) continue;
                if (!edge_z && ix != x0 && ix != x1) continue;
                int c = iz * db.gnx + ix;
                for (int k = db.cell_start[c]; k < db.cell_start[c + 1]; ++k) {
                    int e = db.cell_elems[k];
                    float dx = db.mp_s[e] - qs, dz = db.mp_z[e] - qz;
                    float sd = 0.0f;
                    sd = sd + dx * dx;
                    sd = sd + dz * dz;
                    knn_insert(sd, e, bd, bi, cnt);
                }
            }
        }
        bool all = (x0 <= 0 && z0 <= 0 && x1 >= db.gnx - 1 && z1 >= db.gnz - 1);
        if (all) break;
        if (cnt == kNN + 1) {
            double inf = 1e300;
            double lo_s = x0 <= 0 ? -inf : db.gs0 + x0 * db.gh, hi_s = x1 >= db.gnx - 1 ? inf : db.gs0 + (x1 + 1) * db.gh;
            double lo_z = z0 <= 0 ? -inf : db.gz0 + z0 * db.gh, hi_z = z1 >= db.gnz - 1 ? inf : db.gz0 + (z1 + 1) * db.gh;
            double dm = fmin(fmin((double)qs - lo_s, hi_s - (double)qs), fmin((double)qz - lo_z, hi_z - (double)qz));
            if (dm > 0 && (double)bd[kNN] < dm * dm * (1.0 - 1e-6)) break;
        }
    }
    for (int j = 1; j < cnt; ++j) if (bd[j] == bd[j - 1]) res.tie = 1;
    for (int c = 0; c < kNN && c < cnt; ++c) {
        int e = bi[c];
        double xi, eta;
        if (d_inside(s, z, db.econ + (size_t)e * 16, db.eltype[e], &xi, &eta)) {
            res.elem = e; res.xi = xi; res.eta = eta;
            break;
        }
    }
    if (res.elem < 0 && err) atomicOr(err, 1);
    out[gid] = res;
}

// ------------------------------------------------------------------------------------
// Strain operator tables.  Strain components (Voigt ss, pp, zz, zp, sz, sp) of one
// simulation as linear combinations of d/ds, d/dz and 1/s of its variables
// (sem_derivatives.f90:281-649).  op: 0 = d/ds, 1 = d/dz, 2 = 1/s.
// ------------------------------------------------------------------------------------
struct Term { signed char comp, var, op; float coef; };
// monopole (a = u_s, b = u_z)
__constant__ Term kMono[] = {{0, 0, 0, 1.f}, {1, 0, 2, 1.f}, {2, 1, 1, 1.f},
                             {4, 0, 1, .5f}, {4, 1, 0, .5f}};
__constant__ Term kDi[] = {{0, 0, 0, 1.f}, {1, 0, 2, 1.f}, {1, 1, 2, -1.f}, {2, 2, 1, 1.f},
                           {3, 2, 2, -.5f}, {3, 1, 1, -.5f}, {4, 0, 1, .5f}, {4, 2, 0, .5f},
                           {5, 0, 2, -.5f}, {5, 1, 2, .5f}, {5, 1, 0, -.5f}};
__constant__ Term kQuad[] = {{0, 0, 0, 1.f}, {1, 0, 2, 1.f}, {1, 1, 2, -2.f}, {2, 2, 1, 1.f},
                             {3, 2, 2, -1.f}, {3, 1, 1, -.5f}, {4, 0, 1, .5f}, {4, 2, 0, .5f},
                             {5, 1, 2, .5f}, {5, 0, 2, -1.f}, {5, 1, 0, -.5f}};
// traces (straintrace_*_td): comp field unused
__constant__ Term kTrMono[] = {{0, 0, 0, 1.f}, {0, 0, 2, 1.f}, {0, 1, 1, 1.f}};
__constant__ Term kTrDi[] = {{0, 0, 0, 1.f}, {0, 0, 2, 1.f}, {0, 1, 2, -1.f}, {0, 2, 1, 1.f}};
__constant__ Term kTrQuad[] = {{0, 0, 0, 1.f}, {0, 0, 2, 1.f}, {0, 1, 2, -2.f}, {0, 2, 1, 1.f}};

// Voigt rotations (rotations.f90:382-419 and 555-612) applied to a 6-vector
__device__ inline void rot_src_to_xyz(double *t, double sp, double cp)
{
    double a1 = t[0], a2 = t[1], a3 = t[2], a4 = t[3], a5 = t[4], a6 = t[5];
    t[0] = a1 * (cp * cp) + sp * (-2 * a6 * cp + a2 * sp);
    t[1] = a2 * (cp * cp) + sp * (2 * a6 * cp + a1 * sp);
    t[2] = a3;
    t[3] = a4 * cp + a5 * sp;
    t[4] = a5 * cp - a4 * sp;
    t[5] = (a1 - a2) * cp * sp + a6 * (cp * cp - sp * sp);
}

__device__ inline void rot_xyz_to_earth(double *t, double sp, double cp, double st, double ct)
{
    double a1 = t[0], a2 = t[1], a3 = t[2], a4 = t[3], a5 = t[4], a6 = t[5];
    t[0] = a1 * (cp * cp) * (ct * ct) - 2 * a6 * cp * ct * sp + a2 * (sp * sp)
         + 2 * a5 * (cp * cp) * ct * st - 2 * a4 * cp * sp * st + a3 * (cp * cp) * (st * st);
    t[1] = a2 * (cp * cp) + sp * (2 * a6 * cp * ct + a1 * (ct * ct) * sp
                                  + st * (2 * a4 * cp + 2 * a5 * ct * sp + a3 * sp * st));
    t[2] = a3 * (ct * ct) + st * (-2 * a5 * ct + a1 * st);
    t[3] = ct * (a4 * cp + a5 * ct * sp + a3 * sp * st) - st * (a6 * cp + a1 * ct * sp + a5 * sp * st);
    t[4] = ct * (a5 * cp * ct - a4 * sp + a3 * cp * st) - st * (a1 * cp * ct - a6 * sp + a5 * cp * st);
    t[5] = cp * st * (a4 * cp + a5 * ct * sp + a3 * sp * st) - sp * (a2 * cp + a6 * ct * sp + a4 * sp * st)
         + cp * ct * (a6 * cp + a1 * ct * sp + a5 * sp * st);
}

__device__ inline double azim_factor(double phi, const double *mij, int isim, int ikind)
{
    switch (isim) {
    case 1: return ikind == 1 ? mij[0] : 0.0;
    case 2: return ikind == 1 ? 0.5 * (mij[1] + mij[2]) : 0.0;
    case 3: return ikind == 1 ? mij[3] * cos(phi) + mij[4] * sin(phi) : -mij[3] * sin(phi) + mij[4] * cos(phi);
    default: return ikind == 1 ? 0.5 * (mij[1] - mij[2]) * cos(2.0 * phi) + mij[5] * sin(2.0 * phi)
                               : -0.5 * (mij[1] - mij[2]) * sin(2.0 * phi) + mij[5] * cos(2.0 * phi);
    }
}


// Per-element table of the inverse Jacobian (4) and 1/s at the 25 collocation points
// (finite_elem_mapping.f90), [nelem][25][5]: computed once per database.
__global__ void k_elem_tables(DevDb db, double *out)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= db.nelem * 25) return;
    const int e = g / 25, node = g - (g / 25) * 25;
    const int i = node % 5, j = node / 5;
    const double *E = db.econ + (size_t)e * 16;
    const double *xp = db.axis[e] ? c_basis.glj : c_basis.gll;
    double ij[4], sz[2];
    d_inv_jacobian(xp[i], c_basis.gll[j], E, db.eltype[e], ij);
    d_mapping(xp[i], c_basis.gll[j], E, db.eltype[e], sz);
    double *o = out + (size_t)g * 5;
    o[0] = ij[0]; o[1] = ij[1]; o[2] = ij[2]; o[3] = ij[3];
    o[4] = 1.0 / sz[0];
}

// Builds the per-point weights W[c][v][jpol*5+ipol] such that carried component c at
// this point is sum_v sum_node W * u_v(node, t).
//  earth mode (src_mode false): c are the earth-frame Voigt components (or the trace);
//  source mode: c are the source-frame components comps[c] of the frame, with the
//  azimuthal factors, the sign flip of components 4 and 6 and 1/amplitude applied;
//  S.Q[i][c] is then the rotation of component c into earth-frame component i.
// The set-up runs in two parts separated by a barrier: op_setup_parts (element tables,
// Lagrange bases, output transform) and op_weights.  nthr threads cooperate (tid < nthr);
// threads role_l .. role_l + 9 form the ten Lagrange basis values, role_xf .. role_xf +
// nd - 1 the rotation columns and role_m .. role_m + nsim - 1 the output transform of one
// simulation each (each value is computed exactly as by a single thread).
__device__ void op_setup_parts(OpShared &S, const DevDb &db, const PathConf &pc, const DevFrame &fr,
                               const Located &L, bool fwd, int nd, const int *comps, bool src_mode,
                               int tid, int nthr, int role_l, int role_xf, int role_m)
{
    const int e = L.elem;
    if (tid == 0) { S.elem = e; S.axial = db.axis[e]; S.type = db.eltype[e]; }
    for (int i = tid; i < 125; i += nthr) {
        // inverse Jacobian and 1/s at the 25 nodes (element table, see k_elem_tables)
        const double v = db.ejac[(size_t)e * 125 + i];
        const int n = i / 5, c = i - (i / 5) * 5;
        if (c < 4) S.invj[n][c] = v; else S.sinv[n] = v;
    }
    if (tid >= role_l && tid < role_l + 10) {
        // Lagrange basis at xi (GLJ on the axis) or eta (spectral_basis.f90:203-217)
        const int which = tid - role_l;
        const bool is_xi = which < 5;
        const int i = is_xi ? which : which - 5;
        const double *pts = is_xi ? (db.axis[e] ? c_basis.glj : c_basis.gll) : c_basis.gll;
        const double x = is_xi ? L.xi : L.eta;
        double v = 1.0;
        for (int m = 0; m < 5; ++m) {
            if (m == i) continue;
            v = v * (x - pts[m]) / (pts[i] - pts[m]);
        }
        if (is_xi) S.li[i] = v; else S.lj[i] = v;
    }
    if (pc.full && src_mode && tid >= role_xf && tid < role_xf + nd) {
        // rotation of carried component c: src -> xyz -> earth (full strain only)
        const int c = tid - role_xf;
        const double sp = sin(L.phi), cp = cos(L.phi);
        double t[6] = {0, 0, 0, 0, 0, 0};
        t[comps[c]] = 1.0;
        rot_src_to_xyz(t, sp, cp);
        rot_xyz_to_earth(t, fr.sin_lon, fr.cos_lon, fr.sin_colat, fr.cos_colat);
        for (int i = 0; i < 6; ++i) S.Q[i][c] = (float)t[i];
    }
    const int nsim = fwd ? pc.nsim_fwd : 1;
    if (tid >= role_m && tid < role_m + nsim) {
        // output transform of simulation si: sign flip of comps 4, 6 (readfields.f90:
        // 2180-2184), azimuthal factors, 1/amplitude (and, in earth mode, the rotations)
        const int si = tid - role_m;
        const double amp = db.amplitude;
        double az[6];
        if (fwd) {
            int isim = pc.fwd_sim[si];
            double a1 = azim_factor(L.phi, pc.mij, isim, 1), a2 = azim_factor(L.phi, pc.mij, isim, 2);
            az[0] = a1; az[1] = a1; az[2] = a1; az[3] = a2; az[4] = a1; az[5] = a2;
        } else if (pc.bwd_isim == 1) {
            for (int m = 0; m < 6; ++m) az[m] = 1.0;
        } else {
            double a1 = pc.bwd_fi[1] * cos(L.phi) + pc.bwd_fi[2] * sin(L.phi);
            double a2 = -pc.bwd_fi[1] * sin(L.phi) + pc.bwd_fi[2] * cos(L.phi);
            az[0] = a1; az[1] = a1; az[2] = a1; az[3] = a2; az[4] = a1 * 2; az[5] = a2 * 2;
        }
        if (!pc.full) {
            for (int m = 0; m < 6; ++m) S.M[0][si][m] = (m == 0) ? az[0] / amp : 0.0;
        } else if (src_mode) {
            for (int c = 0; c < 6; ++c)
                for (int m = 0; m < 6; ++m)
                    S.M[c][si][m] = (c < nd && m == comps[c])
                                  ? ((m == 3 || m == 5) ? -1.0 : 1.0) * az[m] / amp : 0.0;
        } else {
            const double sp = sin(L.phi), cp = cos(L.phi);
            for (int m = 0; m < 6; ++m) {
                double t[6] = {0, 0, 0, 0, 0, 0};
                t[m] = 1.0;
                rot_src_to_xyz(t, sp, cp);
                rot_xyz_to_earth(t, fr.sin_lon, fr.cos_lon, fr.sin_colat, fr.cos_colat);
                double f = ((m == 3 || m == 5) ? -1.0 : 1.0) * az[m] / amp;
                for (int c = 0; c < 6; ++c) S.M[c][si][m] = t[c] * f;
            }
        }
    }
}

// W[(c * nvar + v) * 25 + node] from the tables of op_setup_parts (W: LDS or global,
// f32 for the spectra kernels, f64 for the seismograms)
template <class T>
__device__ void op_weights(const OpShared &S, const PathConf &pc, bool fwd, int nvar, const int *vars,
                           int nd, int tid, int nthr, T *W)
{
    const bool axial = S.axial != 0;
    const int total = nd * nvar * 25;
    const double *Gx = axial ? c_basis.G1 : c_basis.G2;
    for (int idx = tid; idx < total; idx += nthr) {
        int n = idx % 25, v = (idx / 25) % nvar, c = idx / (25 * nvar);
        int k = n % 5, l = n / 5;
        int var = vars[v];
        // simulation of this variable and its local index
        int si, pole, vl;
        if (fwd) {
            // variables 0-1 MZZ, 2-3 MXX+MYY, 4-6 MXZ, 7-9 MXY
            int isim = var < 2 ? 1 : (var < 4 ? 2 : (var < 7 ? 3 : 4));
            pole = isim <= 2 ? 0 : isim - 2;
            vl = var - (isim == 1 ? 0 : isim == 2 ? 2 : isim == 3 ? 4 : 7);
            si = 0;
            for (int q = 0; q < pc.nsim_fwd; ++q) if (pc.fwd_sim[q] == isim) si = q;
        } else {
            pole = pc.bwd_isim == 1 ? 0 : 1;
            vl = var - (pc.bwd_isim == 1 ? 0 : 2);
            si = 0;
        }
        const Term *tab;
        int nt;
        if (pc.full) {
            tab = pole == 0 ? kMono : (pole == 1 ? kDi : kQuad);
            nt = pole == 0 ? 5 : 11;
        } else {
            tab = pole == 0 ? kTrMono : (pole == 1 ? kTrDi : kTrQuad);
            nt = pole == 0 ? 3 : 4;
        }
        double as = 0.0, azz = 0.0, af = 0.0;
        for (int q = 0; q < nt; ++q) {
            if (tab[q].var != vl) continue;
            double m = S.M[c][si][pc.full ? tab[q].comp : 0] * (double)tab[q].coef;
            if (tab[q].op == 0) as += m; else if (tab[q].op == 1) azz += m; else af += m;
        }
        double w = 0.0;
        // xi-derivative part: sum_i Pxi(i, l) * GT(i, k),  GT(i, k) = Gx[k + 5 i]
        for (int i = 0; i < 5; ++i) {
            const double *ij = S.invj[i + 5 * l];
            double Lil = S.li[i] * S.lj[l];
            double pxi = Lil * (as * ij[0] + azz * ij[2]);
            if (axial && i == 0) pxi += af * Lil * ij[0];
            w += pxi * Gx[k + 5 * i];
        }
        // eta-derivative part: sum_j Peta(k, j) * G(l, j),  G(l, j) = G2[l + 5 j]
        for (int j = 0; j < 5; ++j) {
            const double *ij = S.invj[k + 5 * j];
            double Lkj = S.li[k] * S.lj[j];
            double peta = Lkj * (as * ij[1] + azz * ij[3]);
            if (axial && k == 0) peta += af * Lkj * ij[1];
            w += peta * c_basis.G2[l + 5 * j];
        }
        if (!(axial && k == 0)) w += af * S.li[k] * S.lj[l] * S.sinv[k + 5 * l];
        W[(c * nvar + v) * 25 + n] = (T)w;
    }
}

__device__ void build_weights(OpShared &S, const DevDb &db, const PathConf &pc, const DevFrame &fr,
                              const Located &L, bool fwd, int nvar, const int *vars, int nd,
                              const int *comps, bool src_mode)
{
    op_setup_parts(S, db, pc, fr, L, fwd, nd, comps, src_mode, threadIdx.x, blockDim.x, 160, 192, 224);
    __syncthreads();
    op_weights(S, pc, fwd, nvar, vars, nd, threadIdx.x, blockDim.x, S.W);
    __syncthreads();
}

// ------------------------------------------------------------------------------------
// k_seis: the seismogram of receiver blockIdx.x by reciprocity (load_fw_points_rdbm,
// readfields.f90:1469-1604): the full strain of simulation bwd_isim at the source point
// located in the receiver's frame (comps 4 and 6 negated as load_strain_point_merged_bwd
// does, readfields.f90:2390-2393; no azimuthal factor, no amplitude), from FP64 weights
// and FP64 sums, contracted with the rotated moment tensor mb[r][6] (already divided by
// the amplitude) in the reference's order of operations (readfields.f90:1562-1597).
// comp: 0 Z, 1 R, 2 T; pc.full must be 1 (full strain tables).  Receivers whose element
// lies outside db's page are skipped.
// ------------------------------------------------------------------------------------
__global__ __launch_bounds__(kStrainThreads) void k_seis(DevDb db, PathConf pc, const Located *loc,
                                                       const double *mb, int comp, double *seis)
{
    __shared__ OpShared S;
    __shared__ double Wd[6 * 3 * 25];
    const int r = blockIdx.x, tid = threadIdx.x;
    const Located L = loc[r];
    const float *ue = elem_u(db, L.elem);
    if (!ue) return;
    DevFrame none;
    none.cos_lon = 1.0; none.sin_lon = 0.0; none.cos_colat = 1.0; none.sin_colat = 0.0;
    // element tables and Lagrange bases only (no output-transform roles)
    op_setup_parts(S, db, pc, none, L, false, 6, nullptr, false, tid, kStrainThreads, 160, -64, -64);
    if (tid < 36) {
        const int c = tid / 6, m = tid - (tid / 6) * 6;
        S.M[c][0][m] = c == m ? ((m == 3 || m == 5) ? -1.0 : 1.0) : 0.0;
    }
    __syncthreads();
    const int nvar = pc.nvar_bwd;
    op_weights(S, pc, false, nvar, pc.bwd_vars, 6, tid, kStrainThreads, Wd);
    __syncthreads();
    const int nt = pc.nt;
    const size_t blk = (size_t)25 * nt;
    const double *m = mb + 6 * r;
    for (int t = tid; t < nt; t += kStrainThreads) {
        double e[6] = {0, 0, 0, 0, 0, 0};
        for (int v = 0; v < nvar; ++v) {
            const float *uv = ue + (size_t)pc.bwd_vars[v] * blk + t;
            for (int n = 0; n < 25; ++n) {
                const double x = (double)uv[(size_t)n * nt];
#pragma unroll
                for (int c = 0; c < 6; ++c) e[c] += Wd[(c * nvar + v) * 25 + n] * x;
            }
        }
        double a = 0.0;
        if (comp == 0) {
            a = a + m[0] * e[0];
            a = a + m[1] * e[1];
            a = a + m[2] * e[2];
            a = a + 2 * m[4] * e[4];
        } else if (comp == 1) {
            a = a + m[0] * e[0];
            a = a + m[1] * e[1];
            a = a + m[2] * e[2];
            a = a + m[4] * e[4] * 2;
            a = -a;
        } else {
            a = a + m[3] * e[3] * 2;
            a = a + m[5] * e[5] * 2;
        }
        seis[(size_t)r * nt + t] = a;
    }
}

// ------------------------------------------------------------------------------------
// k_op_setup: the operator set-up of the spectra kernels, one wave per item, so that its
// serial FP64 chains (sine and cosine of the azimuth, frame rotations, Lagrange bases)
// run at full occupancy instead of inside the FFT workgroups.
//  fwd: items are points (frame 0), outputs indexed by point id;
//  bwd: items are (point, receiver) pairs q = p * R + r (frame 1 + r), outputs indexed
//       by the pair's position in the launch order.
// W: [item][nd][nvar][25] gather weights (as build_weights in source mode); Q: [item][6][6]
// rotation of the carried components into the earth frame (columns >= nd and the trace
// mode: zero).
// ------------------------------------------------------------------------------------
constexpr int kSetupWaves = 4;

__global__ __launch_bounds__(64 * kSetupWaves) void k_op_setup(DevDb db, PathConf pc, const DevFrame *frames,
                                                               int n, const int *order, const int *keys, int R,
                                                               const Located *loc, int fwd, float *Wout, float *Qout)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    OpShared &S = *reinterpret_cast<OpShared *>(smem_raw + wave * op_bytes(0));
    const int lb = blockIdx.x * kSetupWaves + wave;
    const int q = lb < n ? (order ? order[lb] : lb) : 0;
    Located L;
    L.elem = -1;
    if (lb < n && !(keys && keys[lb] >= db.nelem)) L = loc[q];
    const bool live = L.elem >= 0;
    const int nvar = fwd ? pc.nvar_fwd : pc.nvar_bwd;
    const int *vars = fwd ? pc.fwd_vars : pc.bwd_vars;
    const int nd = fwd ? pc.nd_f : pc.nd_b;
    const int *comps = fwd ? pc.comps_f : pc.comps_b;
    const int frame = fwd ? 0 : 1 + (q - (q / R) * R);
    if (live) op_setup_parts(S, db, pc, frames[frame], L, fwd != 0, nd, comps, true, lane, 64, 40, 50, 56);
    __syncthreads();
    if (!live) return;
    const size_t item = fwd ? (size_t)q : (size_t)lb;
    op_weights(S, pc, fwd != 0, nvar, vars, nd, lane, 64, Wout + item * (size_t)(nd * nvar * 25));
    if (lane < 36) {
        const int i = lane / 6, c = lane - (lane / 6) * 6;
        Qout[item * 36 + lane] = (pc.full && c < nd) ? S.Q[i][c] : 0.0f;
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs; block b of a grid padded to a
// multiple of 8 takes logical item (b % 8) * (grid / 8) + b / 8, so each XCD walks one
// contiguous range of the element-sorted work and neighbours share its L2.
__device__ inline int xcd_block(int n)
{
    const int b = blockIdx.x, per = gridDim.x >> 3;
    const int lb = (b & 7) * per + (b >> 3);
    return lb < n ? lb : -1;
}

// the same over the first n items of a grid that may be larger: every XCD takes one
// contiguous eighth of those n (blocks past them return -1)
__device__ inline int xcd_block_n(int n)
{
    const int b = blockIdx.x, per = (n + 7) >> 3, i = b >> 3;
    if (i >= per) return -1;
    const int lb = (b & 7) * per + i;
    return lb < n ? lb : -1;
}

// acc[c][q] = sum_v sum_node W[c][v][node] u_v(node, t0 + q): the strain (or trace) of
// carried component c at this thread's SPT samples
// loads rows n0..n0+NB-1 of one variable (this thread's samples)
template <int SPT, int NB>
__device__ inline void load_rows(const float *uv, int n0, int nt, int t0, bool full_vec, float (&x)[NB][SPT])
{
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const float *src = uv + (size_t)(n0 + b) * nt;
        if (full_vec && (SPT % 4) == 0) {
#pragma unroll
            for (int q = 0; q < SPT; q += 4) {
                float4 t = *reinterpret_cast<const float4 *>(src + q);
                x[b][q] = t.x; x[b][q + 1] = t.y; x[b][q + 2] = t.z; x[b][q + 3] = t.w;
            }
        } else {
#pragma unroll
            for (int q = 0; q < SPT; ++q) x[b][q] = (t0 + q < nt) ? src[q] : 0.0f;
        }
    }
}

// FMAs of rows n0..n0+NB-1 of variable v into the carried components.  Monopole fields
// (ND 4: ss, pp, zz, sz) come as (u_s, u_z) variable pairs; u_s feeds ss, pp and sz, u_z
// feeds zz and sz - the other weights are exactly zero.
template <int SPT, int ND, int NB>
__device__ inline void fma_rows(const OpShared &S, int nvar, int v, int n0, const float (&x)[NB][SPT],
                                float (&acc)[ND][SPT])
{
    if (ND == 4) {
        if (v & 1) {
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const float w2 = S.W[(2 * nvar + v) * 25 + n0 + b];
                const float w3 = S.W[(3 * nvar + v) * 25 + n0 + b];
#pragma unroll
                for (int q = 0; q < SPT; ++q) {
                    acc[2 % ND][q] = __builtin_fmaf(w2, x[b][q], acc[2 % ND][q]);
                    acc[3 % ND][q] = __builtin_fmaf(w3, x[b][q], acc[3 % ND][q]);
                }
            }
        } else {
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const float w0 = S.W[(0 * nvar + v) * 25 + n0 + b];
                const float w1 = S.W[(1 * nvar + v) * 25 + n0 + b];
                const float w3 = S.W[(3 * nvar + v) * 25 + n0 + b];
#pragma unroll
                for (int q = 0; q < SPT; ++q) {
                    acc[0][q] = __builtin_fmaf(w0, x[b][q], acc[0][q]);
                    acc[1 % ND][q] = __builtin_fmaf(w1, x[b][q], acc[1 % ND][q]);
                    acc[3 % ND][q] = __builtin_fmaf(w3, x[b][q], acc[3 % ND][q]);
                }
            }
        }
    } else {
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int c = 0; c < ND; ++c) {
                const float w = S.W[(c * nvar + v) * 25 + n0 + b];
#pragma unroll
                for (int q = 0; q < SPT; ++q) acc[c][q] = __builtin_fmaf(w, x[b][q], acc[c][q]);
            }
    }
}

// acc[c][q] = sum_v sum_node W[c][v][node] u_v(node, t0 + q): the strain (or trace) of
// carried component c at this thread's SPT samples.  The 25 node rows of a variable are
// read in 5 batches of 5; for the monopole components the loads of batch b + 1 are in
// flight while batch b is summed.
template <int SPT, int ND>
__device__ inline void gather_strain(const OpShared &S, const float *ue, size_t blk, int nvar,
                                     const int *vars, int t0, int nt, float (&acc)[ND][SPT])
{
#pragma unroll
    for (int c = 0; c < ND; ++c)
#pragma unroll
        for (int q = 0; q < SPT; ++q) acc[c][q] = 0.0f;
    if (t0 >= nt) return;
    const bool full_vec = (t0 + SPT <= nt) && ((nt & 3) == 0);
    constexpr int NB = 5;
    if (ND != 4) {
        // dense components: one batch in flight (register budget)
        for (int v = 0; v < nvar; ++v) {
            const float *uv = ue + (size_t)vars[v] * blk + t0;
            for (int n0 = 0; n0 < 25; n0 += NB) {
                float x[NB][SPT];
                load_rows<SPT, NB>(uv, n0, nt, t0, full_vec, x);
                fma_rows<SPT, ND, NB>(S, nvar, v, n0, x, acc);
            }
        }
        return;
    }
    for (int v = 0; v < nvar; ++v) {
        const float *uv = ue + (size_t)vars[v] * blk + t0;
        float xa[NB][SPT], xb[NB][SPT];
        load_rows<SPT, NB>(uv, 0, nt, t0, full_vec, xa);
        load_rows<SPT, NB>(uv, 5, nt, t0, full_vec, xb);
        fma_rows<SPT, ND, NB>(S, nvar, v, 0, xa, acc);
        load_rows<SPT, NB>(uv, 10, nt, t0, full_vec, xa);
        fma_rows<SPT, ND, NB>(S, nvar, v, 5, xb, acc);
        load_rows<SPT, NB>(uv, 15, nt, t0, full_vec, xb);
        fma_rows<SPT, ND, NB>(S, nvar, v, 10, xa, acc);
        load_rows<SPT, NB>(uv, 20, nt, t0, full_vec, xa);
        fma_rows<SPT, ND, NB>(S, nvar, v, 15, xb, acc);
        fma_rows<SPT, ND, NB>(S, nvar, v, 20, xa, acc);
    }
}

// component c of the accumulators, selected with static indices so that acc stays in
// registers; samples beyond nt are zero
template <int SPT, int ND>
__device__ inline void select_comp(const float (&acc)[ND][SPT], int c, int t0, int nt, float (&x)[SPT])
{
#pragma unroll
    for (int q = 0; q < SPT; ++q) x[q] = 0.0f;
#pragma unroll
    for (int cc = 0; cc < ND; ++cc)
        if (cc == c)
#pragma unroll
            for (int q = 0; q < SPT; ++q) x[q] = acc[cc][q];
#pragma unroll
    for (int q = 0; q < SPT; ++q) if (t0 + q >= nt) x[q] = 0.0f;
}

// cumsum_trapezoidal (simple_routines.f90:314-341): c0 = x0, c1 = c0 + x1,
// ct = c(t-1) + x(t-1) + xt, times dt/2; a block-wide scan of the per-thread totals.  The
// running sums are kept in double precision (the reference's whole cumsum is FP64) and
// rounded once per sample.
template <int SPT>
__device__ inline void cumsum_trapezoidal(float (&x)[SPT], int t0, int nt, float h, float *edge_x,
                                          double *wave_tot)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    edge_x[tid + 1] = x[SPT - 1];
    if (tid == 0) edge_x[0] = 0.0f;
    __syncthreads();
    float prev = edge_x[tid];
    double d[SPT], run = 0.0;
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        int t = t0 + q;
        float xm1 = q == 0 ? prev : x[q - 1];
        d[q] = (t <= 1) ? (double)x[q] : ((double)xm1 + (double)x[q]);
        run += d[q];
    }
    double incl = run;
    for (int o = 1; o < 64; o <<= 1) {
        double y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) wave_tot[wave] = incl;
    __syncthreads();
    double wbase = 0.0;
    for (int w2 = 0; w2 < wave; ++w2) wbase += wave_tot[w2];
    double cs = wbase + (incl - run);
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        cs += d[q];
        x[q] = (t0 + q < nt) ? (float)(cs * (double)h) : 0.0f;
    }
    __syncthreads();
}

// taper factors of this thread's samples: ntaper samples at both ends of the nt-sample
// trace (fft.f90:546-601), 1 elsewhere
template <int SPT>
__device__ inline void taper_factors(int t0, int nt, const PathConf &pc, float (&tap)[SPT])
{
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int t = t0 + q;
        tap[q] = 1.0f;
        if (t < pc.ntaper) tap[q] = pc.taper[t];
        else if (t >= nt - pc.ntaper && t < nt) tap[q] = pc.taper[nt - 1 - t];
    }
}

// The FFTs and the frequency-domain products allow FMA contraction (fp contract fast):
// they are FP32 work with a 1e-5 tolerance.  The translation unit is built with
// -ffp-contract=off so that the point-location and Halton arithmetic match the
// reference operation for operation.

__device__ inline float2 cmulf(float2 a, float2 b)
{
#pragma clang fp contract(fast)
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// w_N^i = exp(-2 pi i i / N) from the quarter table T[r] = w_N^r, r < N/4 (in LDS):
// w_N^(q N/4 + r) = (-i)^q w_N^r
__device__ inline float2 twiddle(const float2 *T, int i, int lq)
{
    const int q = (i >> lq) & 3;
    const float2 w = T[tq_slot(i & ((1 << lq) - 1))];
    return q == 0 ? w : q == 1 ? make_float2(w.y, -w.x) : q == 2 ? make_float2(-w.x, -w.y) : make_float2(-w.y, w.x);
}

// copies the quarter table (N/4 = M/2 entries) into LDS at the padded slots tq_slot(r);
// the caller's barrier publishes it
__device__ inline void load_twiddles(const float2 *twq, int M, float2 *Tq)
{
    for (int r = threadIdx.x; r < (M >> 1); r += kStrainThreads) Tq[tq_slot(r)] = twq[r];
}

// Packs the zero-padded real sequence x (length 2M) as z[n] = x[2n] + i x[2n+1] and runs
// the length-M Stockham autosort FFT in LDS (radix 4, after one radix-2 pass when log2 M
// is odd).  z[n] vanishes for n >= M/2 (nt <= M), so the radix-2 pass reduces to
// out[2n] = out[2n+1] = z[n] and is fused with the packing.  Returns the buffer holding
// Z; ends with a barrier.  From 2 samples per thread on, M = 256 SPT (the host sizes both
// from nt), so the passes' sizes and strides can be compile-time constants: UNROLL takes
// them so and unrolls the passes and each thread's butterflies (more registers: off for
// kernels that hold other spectra in registers meanwhile).
template <int SPT>
constexpr int fft_size_of_spt() { return SPT >= 2 ? 256 * SPT : 0; }

template <int SPT, bool UNROLL = true>
__device__ inline const float2 *lds_fft(const float (&x)[SPT], int t0, int nt, int M_, float2 *bufA,
                                        float2 *bufB, const float2 *Tq)
{
#pragma clang fp contract(fast)
    constexpr int MC = UNROLL ? fft_size_of_spt<SPT>() : 0;
    const int M = MC ? MC : M_;
    const int tid = threadIdx.x;
    const int lM = 31 - __clz(M);
    float2 *in = bufA, *ob = bufB;
    int Ns = 1;
    if (lM & 1) {
        // thread holds samples t0..t0+SPT-1 = z[t0/2 .. (t0+SPT)/2 - 1]
        if (SPT >= 2) {
#pragma unroll
            for (int q = 0; q < SPT; q += 2) {
                const int n = (t0 + q) >> 1;
                if (n < (M >> 1)) {
                    const float2 z = make_float2(t0 + q < nt ? x[q] : 0.0f, t0 + q + 1 < nt ? x[q + 1] : 0.0f);
                    reinterpret_cast<float4 *>(in)[n] = make_float4(z.x, z.y, z.x, z.y);
                }
            }
        } else {
            // one sample per thread: pair up through LDS as floats
            float *fa = reinterpret_cast<float *>(ob);
            for (int n = tid; n < M; n += kStrainThreads) ob[n] = make_float2(0.0f, 0.0f);
            __syncthreads();
            if (t0 < nt) fa[t0] = x[0];
            __syncthreads();
            for (int n = tid; n < (M >> 1); n += kStrainThreads) {
                const float2 z = ob[n];
                in[2 * n] = z;
                in[2 * n + 1] = z;
            }
        }
        Ns = 2;
    } else {
        for (int n = tid; n < M; n += kStrainThreads) in[n] = make_float2(0.0f, 0.0f);
        __syncthreads();
        float *fa = reinterpret_cast<float *>(in);
#pragma unroll
        for (int q = 0; q < SPT; ++q)
            if (t0 + q < nt) fa[t0 + q] = x[q];
    }
    __syncthreads();
    const int Mq = M >> 2;
    // (a compile-time size unrolls the passes and each thread's butterflies)
    constexpr int UP = MC ? 8 : 1, UJ = MC ? 4 : 1;
#pragma unroll UP
    for (int lNs = Ns == 1 ? 0 : 1; Ns < M; Ns <<= 2, lNs += 2) {
        const int lstride = lM - lNs - 2;            // M / (4 Ns)
#pragma unroll UJ
        for (int j = tid; j < Mq; j += kStrainThreads) {
            const int kk = j & (Ns - 1);
            const float2 v0 = in[j], v1 = in[j + Mq], v2 = in[j + 2 * Mq], v3 = in[j + 3 * Mq];
            // w_M^t = w_N^(2t), N = 2M
            const int ti = kk << (lstride + 1);
            // ti < N/4: w1 lies in the table's quadrant; w2 and w3 from the table too
            const float2 w1 = Tq[tq_slot(ti)], w2 = twiddle(Tq, 2 * ti, lM - 1), w3 = twiddle(Tq, 3 * ti, lM - 1);
            const float2 a1 = make_float2(v1.x * w1.x - v1.y * w1.y, v1.x * w1.y + v1.y * w1.x);
            const float2 a2 = make_float2(v2.x * w2.x - v2.y * w2.y, v2.x * w2.y + v2.y * w2.x);
            const float2 a3 = make_float2(v3.x * w3.x - v3.y * w3.y, v3.x * w3.y + v3.y * w3.x);
            const float2 s02 = make_float2(v0.x + a2.x, v0.y + a2.y), d02 = make_float2(v0.x - a2.x, v0.y - a2.y);
            const float2 s13 = make_float2(a1.x + a3.x, a1.y + a3.y), d13 = make_float2(a1.x - a3.x, a1.y - a3.y);
            const int o = ((j >> lNs) << (lNs + 2)) + kk;
            ob[o] = make_float2(s02.x + s13.x, s02.y + s13.y);
            // y1 = d02 - i d13, y3 = d02 + i d13
            ob[o + Ns] = make_float2(d02.x + d13.y, d02.y - d13.x);
            ob[o + 2 * Ns] = make_float2(s02.x - s13.x, s02.y - s13.y);
            ob[o + 3 * Ns] = make_float2(d02.x - d13.y, d02.y + d13.x);
        }
        __syncthreads();
        float2 *tmp = in; in = ob; ob = tmp;
    }
    return in;
}

// real-FFT post-processing: X_k = E_k + e^{-2 pi i k/N} O_k (k = 0..M), times dt (the
// FFTW r2c convention of fft.f90:375-413 scaled as rfft_type%rfft does)
__device__ inline float2 rfft_bin(const float2 *Z, int k, int M, const float2 *Tq, float dt)
{
#pragma clang fp contract(fast)
    float2 zk = Z[k == M ? 0 : k];
    float2 zm = Z[(M - k) & (M - 1)];
    float2 zc = make_float2(zm.x, -zm.y);
    float2 E = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y + zc.y));
    float2 dd = make_float2(zk.x - zc.x, zk.y - zc.y);
    float2 O = make_float2(0.5f * dd.y, -0.5f * dd.x);
    float2 w = twiddle(Tq, k, 30 - __clz(M));     // w_N^k, log2(N/4) = log2(M) - 1
    float2 X = make_float2(E.x + w.x * O.x - w.y * O.y, E.y + w.x * O.y + w.y * O.x);
    return make_float2(X.x * dt, X.y * dt);
}


__device__ inline double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// one DPP step on both halves of a double; rows outside ROWMASK receive +0
template <int CTRL, int ROWMASK>
__device__ inline double dpp_d(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xF, false);
    return __hiloint2double(hi, lo);
}

// wave64 sum through DPP moves (quad permutes, row mirrors, row broadcasts); the total
// is exact in lane 63 only
__device__ inline double wave_sum_lane63(double v)
{
    v += dpp_d<0xB1, 0xF>(v);    // quad_perm [1,0,3,2]
    v += dpp_d<0x4E, 0xF>(v);    // quad_perm [2,3,0,1]
    v += dpp_d<0x141, 0xF>(v);   // row_half_mirror
    v += dpp_d<0x140, 0xF>(v);   // row_mirror: every lane holds its row's sum
    v += dpp_d<0x142, 0xA>(v);   // row_bcast:15 into rows 1 and 3
    v += dpp_d<0x143, 0xC>(v);   // row_bcast:31 into rows 2 and 3
    return v;
}

// ---- double-precision LDS real FFT (the element spectra, computed once per element) ----

// w_N^i = exp(-2 pi i i / N) from the quarter table T[r] = w_N^r, r < N/4 = 2^lq (global
// memory): w_N^(q N/4 + r) = (-i)^q w_N^r
__device__ inline double2 twiddle_d(const double2 *__restrict__ T, int i, int lq)
{
    const int q = (i >> lq) & 3;
    const double2 w = T[i & ((1 << lq) - 1)];
    return q == 0 ? w : q == 1 ? make_double2(w.y, -w.x) : q == 2 ? make_double2(-w.x, -w.y) : make_double2(-w.y, w.x);
}

__device__ inline double2 cmul_dd(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// lds_fft in double precision: packs the zero-padded real sequence x (length 2M) as
// z[n] = x[2n] + i x[2n+1] and runs the length-M Stockham radix-4 FFT (one fused radix-2
// pass first when log2 M is odd); twiddles from the table in global memory.  Returns the
// buffer holding Z; ends with a barrier.
template <int SPT>
__device__ inline const double2 *lds_fft_d(const double (&x)[SPT], int t0, int nt, int M_, double2 *bufA,
                                           double2 *bufB, const double2 *__restrict__ twd)
{
    constexpr int MC = fft_size_of_spt<SPT>();
    const int M = MC ? MC : M_;
    const int tid = threadIdx.x;
    const int lM = 31 - __clz(M);
    double2 *in = bufA, *ob = bufB;
    int Ns = 1;
    if (lM & 1) {
        if (SPT >= 2) {
#pragma unroll
            for (int q = 0; q < SPT; q += 2) {
                const int n = (t0 + q) >> 1;
                if (n < (M >> 1)) {
                    const double2 z = make_double2(t0 + q < nt ? x[q] : 0.0, t0 + q + 1 < nt ? x[q + 1] : 0.0);
                    in[2 * n] = z;
                    in[2 * n + 1] = z;
                }
            }
        } else {
            double *fa = reinterpret_cast<double *>(ob);
            for (int n = tid; n < M; n += kStrainThreads) ob[n] = make_double2(0.0, 0.0);
            __syncthreads();
            if (t0 < nt) fa[t0] = x[0];
            __syncthreads();
            for (int n = tid; n < (M >> 1); n += kStrainThreads) {
                const double2 z = ob[n];
                in[2 * n] = z;
                in[2 * n + 1] = z;
            }
        }
        Ns = 2;
    } else {
        for (int n = tid; n < M; n += kStrainThreads) in[n] = make_double2(0.0, 0.0);
        __syncthreads();
        double *fa = reinterpret_cast<double *>(in);
#pragma unroll
        for (int q = 0; q < SPT; ++q)
            if (t0 + q < nt) fa[t0 + q] = x[q];
    }
    __syncthreads();
    const int Mq = M >> 2;
    // (a compile-time size unrolls the passes and each thread's butterflies)
    constexpr int UP = MC ? 8 : 1, UJ = MC ? 4 : 1;
#pragma unroll UP
    for (int lNs = Ns == 1 ? 0 : 1; Ns < M; Ns <<= 2, lNs += 2) {
        const int lstride = lM - lNs - 2;            // M / (4 Ns)
#pragma unroll UJ
        for (int j = tid; j < Mq; j += kStrainThreads) {
            const int kk = j & (Ns - 1);
            const double2 v0 = in[j], v1 = in[j + Mq], v2 = in[j + 2 * Mq], v3 = in[j + 3 * Mq];
            // w_M^t = w_N^(2t), N = 2M; the index lies in the table's quadrant
            const int ti = kk << (lstride + 1);
            const double2 w1 = twd[ti], w2 = cmul_dd(w1, w1), w3 = cmul_dd(w1, w2);
            const double2 a1 = cmul_dd(v1, w1), a2 = cmul_dd(v2, w2), a3 = cmul_dd(v3, w3);
            const double2 s02 = make_double2(v0.x + a2.x, v0.y + a2.y), d02 = make_double2(v0.x - a2.x, v0.y - a2.y);
            const double2 s13 = make_double2(a1.x + a3.x, a1.y + a3.y), d13 = make_double2(a1.x - a3.x, a1.y - a3.y);
            const int o = ((j >> lNs) << (lNs + 2)) + kk;
            ob[o] = make_double2(s02.x + s13.x, s02.y + s13.y);
            ob[o + Ns] = make_double2(d02.x + d13.y, d02.y - d13.x);
            ob[o + 2 * Ns] = make_double2(s02.x - s13.x, s02.y - s13.y);
            ob[o + 3 * Ns] = make_double2(d02.x - d13.y, d02.y + d13.x);
        }
        __syncthreads();
        double2 *tmp = in; in = ob; ob = tmp;
    }
    return in;
}

// real-FFT post-processing in double precision, times dt, rounded to single precision
__device__ inline float2 rfft_bin_d(const double2 *Z, int k, int M, const double2 *__restrict__ twd, double dt)
{
    const double2 zk = Z[k == M ? 0 : k];
    const double2 zm = Z[(M - k) & (M - 1)];
    const double2 zc = make_double2(zm.x, -zm.y);
    const double2 E = make_double2(0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y));
    const double2 dd = make_double2(zk.x - zc.x, zk.y - zc.y);
    const double2 O = make_double2(0.5 * dd.y, -0.5 * dd.x);
    const double2 w = twiddle_d(twd, k, 30 - __clz(M));     // w_N^k, log2(N/4) = log2(M) - 1
    const double2 X = make_double2(E.x + w.x * O.x - w.y * O.y, E.y + w.x * O.y + w.y * O.x);
    return make_float2((float)(X.x * dt), (float)(X.y * dt));
}

// ------------------------------------------------------------------------------------
// k_strain_traces: one workgroup per point; earth-frame strain (or trace) time series
// [point][nd][nt] of one frame (load_fw_points / load_bw_points, readfields.f90:689-1063).
// ------------------------------------------------------------------------------------
template <int SPT>
__global__ __launch_bounds__(kStrainThreads) void k_strain_traces(
    DevDb db, PathConf pc, const DevFrame *frames, int fwd, int frame, int npts, const Located *loc,
    float *trace_out)
{
    __shared__ OpShared S;
    const int p = blockIdx.x;
    if (p >= npts) return;
    const Located L = loc[p];
    const float *ue = elem_u(db, L.elem);
    if (!ue) return;                                           // not located / not in this page
    const int nvar = fwd ? pc.nvar_fwd : pc.nvar_bwd;
    const int *vars = fwd ? pc.fwd_vars : pc.bwd_vars;
    build_weights(S, db, pc, frames[frame], L, fwd != 0, nvar, vars, pc.nd, nullptr, false);
    const int nt = pc.nt, nd = pc.nd, t0 = threadIdx.x * SPT;
    const size_t blk = (size_t)25 * nt;
    float acc[6][SPT];
    gather_strain<SPT, 6>(S, ue, blk, nvar, vars, t0, nt, acc);
#pragma unroll
    for (int c = 0; c < 6; ++c)
        if (c < nd)
#pragma unroll
            for (int q = 0; q < SPT; ++q)
                if (t0 + q < nt) trace_out[((size_t)p * nd + c) * nt + t0 + q] = acc[c][q];
}

// ------------------------------------------------------------------------------------
// k_fwd_traces: the forward strain gather (sem_derivatives.f90, spectral_basis.f90:181-314,
// readfields.f90:1138-1465: u_c(t) = sum_j W_cj u_j(t) over the nvar x 25 nodal traces j
// of the point's element, the weights from k_op_setup) on the matrix cores.  One wave per
// window of PW = 16 / ND sorted items (windows in XCD-contiguous ranges; blockIdx.y: a share
// of the samples when a launch has few windows): rows r = ND i + c
// (item i0 + i, component c) of a 16-row block, K = the traces in steps of 4,
// v_mfma_f32_16x16x4f32 with A = W of row l % 16 at trace 4 s + l / 16 and B = that trace
// at 16 samples; lane l loads 4 samples
// 64 t + 4 (l % 16) .. + 3 of its trace and feeds them to 4 products (sample 64 t + 4 j + q
// in column j of product q), so that a row's 4 consecutive samples land in one lane.  The
// window's items of one element share every trace load; a window that spans several
// elements takes one pass per element with the other rows' weights zero.
// T: [sorted item][ND][nt]; items not located in db (or keyed past the last element) are
// left unwritten, as k_fwd_spectra skips them.
// ------------------------------------------------------------------------------------
// a float4 of global memory at base + off (base: a generic pointer into device memory;
// global loads take a 32-bit offset from a uniform base)
typedef float vec4f __attribute__((ext_vector_type(4)));
typedef const float __attribute__((address_space(1))) gfloat;
typedef const vec4f __attribute__((address_space(1))) gvec4f;
__device__ __forceinline__ float4 gload4(const float *base, int off)
{
    const vec4f v = *(gvec4f *)((gfloat *)base + (unsigned)off);
    return make_float4(v.x, v.y, v.z, v.w);
}

// One element pass of k_fwd_traces over all super-tiles: A operands from registers (wa,
// KS <= 32) or from LDS (sWa[s][64]); WHOLE: nt a multiple of 64 (no partial tile).  The
// trace loads of half the K steps (8 when the A operands are in LDS) are issued together,
// then multiplied (two batches per super-tile: fewer registers, more waves in flight).
template <int ND, int KS, bool WHOLE>
__device__ __forceinline__ void fwd_traces_pass(const float *ue, const int *toff, const float *wa,
                                                const float *sWa, float mf, int e, const int (&eout)[4],
                                                int i0, int l, int nt, int tb, int te, float *T)
{
    constexpr int PW = 16 / ND;
    constexpr bool WREG = KS <= 32;
    constexpr int CH = WREG ? (KS + 1) / 2 : 8;                // K steps loaded together
    const int r = l & 15, kq = l >> 4;
    for (int t = tb; t < te; ++t) {
        const int s0 = (t << 6) + 4 * r;                       // this lane's 4 samples
        v4f d[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) d[q] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 1
        for (int c0 = 0; c0 < KS; c0 += CH) {
            float4 b[CH];
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const int s = c0 + c;
                b[c] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (s < KS) {
                    const int off = toff[4 * s + kq] + s0;
                    if (WHOLE) {
                        b[c] = gload4(ue, off);
                    } else {
                        if (s0 < nt) b[c].x = ue[off];
                        if (s0 + 1 < nt) b[c].y = ue[off + 1];
                        if (s0 + 2 < nt) b[c].z = ue[off + 2];
                        if (s0 + 3 < nt) b[c].w = ue[off + 3];
                    }
                }
            }
            __builtin_amdgcn_sched_barrier(0);                 // all loads in flight first
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const int s = c0 + c;
                if (s >= KS) break;
                const float a = (WREG ? wa[s < KS ? s : 0] : sWa[s * 64 + l]) * mf;
                d[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b[c].x, d[0], 0, 0, 0);
                d[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b[c].y, d[1], 0, 0, 0);
                d[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b[c].z, d[2], 0, 0, 0);
                d[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b[c].w, d[3], 0, 0, 0);
            }
        }
        // lane l: rows 4 kq + v, samples s0 .. s0 + 3
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int ro = 4 * kq + v, io = ro / ND;
            if (io >= PW || eout[v] != e) continue;
            float *dst = T + ((size_t)(i0 + io) * ND + (ro - io * ND)) * nt + s0;
            if (WHOLE) *reinterpret_cast<float4 *>(dst) = make_float4(d[0][v], d[1][v], d[2][v], d[3][v]);
            else
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (s0 + q < nt) dst[q] = d[q][v];
        }
    }
}

template <int ND, int NV>
__global__ __launch_bounds__(64, 2) void k_fwd_traces(DevDb db, PathConf pc, int npts, const int *nlive,
                                                     const int *order, const int *keys, const Located *loc,
                                                     const float *Wf, float *T)
{
    constexpr int J = NV * 25, KS = (J + 3) / 4, PW = 16 / ND;
    constexpr bool WREG = KS <= 32;
    __shared__ int toff[4 * KS];                               // trace j -> offset in the element
    __shared__ float sWa[WREG ? 1 : KS * 64];                  // A operands when not in registers
    // windows in XCD-contiguous ranges: the windows of an element share one L2
    const int n = nlive ? min(*nlive, npts) : npts;
    const int w = xcd_block_n((n + PW - 1) / PW);
    if (w < 0) return;
    const int i0 = w * PW;
    const int l = threadIdx.x, r = l & 15, kq = l >> 4, nt = pc.nt;
    // (traces past J: offset 0 with weight 0)
    for (int j = l; j < 4 * KS; j += 64) {
        const int v = j / 25;
        toff[j] = j < J ? (pc.fwd_vars[v] * 25 + (j - v * 25)) * nt : 0;
    }
    // this lane's A row: item i0 + r / ND, component r % ND
    const int ir = r / ND, cr = r - ir * ND, item = i0 + ir;
    int erow = -1;
    const float *wrow = nullptr;
    if (ir < PW && item < n && !(keys && keys[item] >= db.nelem)) {
        const int p = order ? order[item] : item;
        erow = loc[p].elem;
        wrow = Wf + ((size_t)p * ND + cr) * J;
    }
    float wa[WREG ? KS : 1];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int j = 4 * s + kq;
        const float w = (wrow && j < J) ? wrow[j] : 0.0f;
        if constexpr (WREG) wa[s] = w;
        else sWa[s * 64 + l] = w;
    }
    // element of the output rows 4 kq + v of this lane
    int eout[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) eout[v] = __shfl(erow, 4 * kq + v, 64);
    __syncthreads();
    const bool whole = (nt & 63) == 0;                         // super-tiles of 64 samples
    // this block's share of the super-tiles (gridDim.y blocks per window: small launches
    // spread a window's samples over several waves)
    const int nst = (nt + 63) >> 6, tb = (int)blockIdx.y * nst / (int)gridDim.y,
              te = ((int)blockIdx.y + 1) * nst / (int)gridDim.y;
    int prev = -1;
    for (;;) {
        // the next element of the window (the wave's smallest above prev)
        int e = erow > prev ? erow : INT_MAX;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) e = min(e, __shfl_xor(e, o, 64));
        e = __builtin_amdgcn_readfirstlane(e);                 // (uniform: a scalar base)
        if (e == INT_MAX) break;
        prev = e;
        const float *ue = elem_u(db, e);
        if (!ue) continue;                                     // not in this page
        const float mf = erow == e ? 1.0f : 0.0f;
        if (whole) fwd_traces_pass<ND, KS, true>(ue, toff, wa, sWa, mf, e, eout, i0, l, nt, tb, te, T);
        else fwd_traces_pass<ND, KS, false>(ue, toff, wa, sWa, mf, e, eout, i0, l, nt, tb, te, T);
    }
}

// ------------------------------------------------------------------------------------
// k_fwd_spectra: one workgroup per point of the forward frame, points in element order.
// The point's strain traces (k_fwd_traces: T[sorted item][ND][nt]), cumsum
// (simple_routines.f90:314-341), taper + zero pad (fft.f90:546-601) and LDS real FFT
// (fft.f90:375-413).
// F: [point][ND][fstride] spectra of the carried source-frame components (bins < nomega).
// ------------------------------------------------------------------------------------
template <int SPT, int ND>
__global__ __launch_bounds__(kStrainThreads, (SPT <= 8 && ND <= 4) ? 3 : (SPT <= 8 ? 2 : 1)) void k_fwd_spectra(
    DevDb db, PathConf pc, int npts, const int *nlive, const int *order, const int *keys, const Located *loc,
    const float *T, const float2 *twq, float2 *F)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2 *bufA = reinterpret_cast<float2 *>(smem_raw);
    float2 *bufB = bufA + pc.M;
    float2 *Tq = bufA + 2 * pc.M;
    __shared__ double wave_tot[kStrainThreads / 64];
    __shared__ float edge_x[kStrainThreads + 1];

    // nlive (device, nullable): the sorted items before the first one keyed past the last
    // element; the XCD ranges cover only those
    const int lb = xcd_block_n(nlive ? min(*nlive, npts) : npts);
    if (lb < 0) return;
    if (keys && keys[lb] >= db.nelem) return;                  // spectra nobody reads
    const int p = order ? order[lb] : lb;
    const Located L = loc[p];
    const float *ue = elem_u(db, L.elem);
    if (!ue) return;                                           // not located / not in this page
    load_twiddles(twq, pc.M, Tq);
    __syncthreads();

    const int tid = threadIdx.x, nt = pc.nt, t0 = tid * SPT;
    const int M = fft_size_of_spt<SPT>() ? fft_size_of_spt<SPT>() : pc.M;      // (= pc.M)
    // the point's strain traces (k_fwd_traces)
    float acc[ND][SPT];
    {
        const float *tr = T + (size_t)lb * ND * nt + t0;
        const bool full_vec = (t0 + SPT <= nt) && ((nt & 3) == 0) && (SPT % 4) == 0;
#pragma unroll
        for (int c = 0; c < ND; ++c) {
            if (full_vec) {
#pragma unroll
                for (int q = 0; q < SPT; q += 4) {
                    const float4 v = *reinterpret_cast<const float4 *>(tr + (size_t)c * nt + q);
                    acc[c][q] = v.x; acc[c][q + 1] = v.y; acc[c][q + 2] = v.z; acc[c][q + 3] = v.w;
                }
            } else {
#pragma unroll
                for (int q = 0; q < SPT; ++q) acc[c][q] = (t0 + q < nt) ? tr[(size_t)c * nt + q] : 0.0f;
            }
        }
    }
    const float fdt = (float)pc.dt, h = (float)(pc.dt * 0.5);
    float tap[SPT];
    taper_factors(t0, nt, pc, tap);
#pragma unroll
    for (int c = 0; c < ND; ++c) {
        float x[SPT];
        select_comp(acc, c, t0, nt, x);
        cumsum_trapezoidal(x, t0, nt, h, edge_x, wave_tot);
#pragma unroll
        for (int q = 0; q < SPT; ++q) x[q] *= tap[q];
        const float2 *Z = lds_fft(x, t0, nt, M, bufA, bufB, Tq);
        float2 *dst = F + ((size_t)p * ND + c) * pc.fstride;
        for (int k = tid; k < pc.fstride; k += kStrainThreads)
            dst[k] = k <= M ? rfft_bin(Z, k, M, Tq, fdt) : make_float2(0.0f, 0.0f);
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------
// k_bwd_eval: one workgroup per active (point, receiver) pair, pairs in element order.
// The backward strain spectra (as k_fwd_spectra, without cumsum) stay in registers and
// are contracted at once with the forward spectra of the point: base kernels
// (kernel.f90:514-570) in the earth frame through the per-pair rotations, and the filtered
// Parseval window integrals (kernel.f90:410-510, 946-1020) as sum_k Re(bk_k H_k) with
// FP64 accumulation; then the physical kernels (kernel.f90:665-832).
//  phys: [point][nk] values of the receiver's kernels; base_out (optional): [point][nk][6]
// ------------------------------------------------------------------------------------
template <int SPT, int NDF, int NDB>
__global__ __launch_bounds__(kStrainThreads, (SPT <= 8 && NDB <= 4) ? 3 : (SPT <= 8 ? 2 : 1)) void k_bwd_eval(
    DevDb db, PathConf pc, int npairs, const int *order, int R,
    const Located *loc_b, const float *Wb, const float *Qb, const float *Qf, const float2 *twq,
    const float2 *F, const RecDesc *recs, const KernDesc *kern, const float2 *H,
    const int *model_idx, const float *model /* [6][ndepth] */, int ndepth, int relative, int nk,
    double *phys, double *base_out, const double *xyz, const uint8_t *active, int P,
    const double *rec_xyz)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    OpShared &S = *reinterpret_cast<OpShared *>(smem_raw);
    float2 *bufA = reinterpret_cast<float2 *>(smem_raw + pc.op_bytes_b);
    float2 *bufB = bufA + pc.M;
    float2 *Tq = bufA + max(2 * pc.M, pc.nbk_lds * (pc.M + 1));
    double *slot_red = reinterpret_cast<double *>(Tq + tq_len(pc.M));   // [4][max_slots]
    double *slotval = slot_red + 4 * pc.max_slots;                      // [max_slots]
    __shared__ EvalShared E;

    const int lb = xcd_block(npairs);
    if (lb < 0) return;
    const int q = order ? order[lb] : lb;
    const int p = q / R, r = q - (q / R) * R;
    const Located L = loc_b[q];
    const float *ue = elem_u(db, L.elem);
    if (!ue) return;                                           // not located / not in this page
    load_twiddles(twq, pc.M, Tq);
    const int tid = threadIdx.x;
    {
        // gather weights and rotation of the pair, rotation of the point's forward
        // components (frame 0) (k_op_setup)
        const int nw = NDB * pc.nvar_bwd * 25;
        const float *src = Wb + (size_t)lb * nw;
        for (int i = tid; i < nw; i += kStrainThreads) S.W[i] = src[i];
        if (tid < 36) (&S.Q[0][0])[tid] = Qb[(size_t)lb * 36 + tid];
        else if (NDF > 1 && tid >= 64 && tid < 100) (&E.Qf[0][0])[tid - 64] = Qf[(size_t)p * 36 + tid - 64];
    }
    __syncthreads();
    if (NDF > 1) {
        // F.B and trace coefficients
        if (tid < NDF * NDB) {
            const int m = tid / NDB, n = tid - (tid / NDB) * NDB;
            float a = 0.0f;
            for (int i = 0; i < 6; ++i) a += E.Qf[i][m] * S.Q[i][n];
            E.Amu[m][n] = a;
        } else if (tid >= 64 && tid < 64 + NDF) {
            const int m = tid - 64;
            E.uf[m] = pc.comps_f[m] < 3 ? 1.0f : 0.0f;     // the trace: rotation invariant
        } else if (tid >= 128 && tid < 128 + NDB) {
            const int n = tid - 128;
            E.ub[n] = pc.comps_b[n] < 3 ? 1.0f : 0.0f;
        }
        // (made visible by the barriers of the FFTs below)
    }

    const int nt = pc.nt, M = pc.M, nw = pc.nomega, t0 = tid * SPT;
    const size_t blk = (size_t)25 * nt;
    float2 spec[NDB][SPT], specM[NDB];
    {
        float acc[NDB][SPT];
        gather_strain<SPT, NDB>(S, ue, blk, pc.nvar_bwd, pc.bwd_vars, t0, nt, acc);
        const float fdt = (float)pc.dt;
        float tap[SPT];
        taper_factors(t0, nt, pc, tap);
        for (int c = 0; c < NDB; ++c) {
            float x[SPT];
            select_comp(acc, c, t0, nt, x);
#pragma unroll
            for (int q = 0; q < SPT; ++q) x[q] *= tap[q];
            const float2 *Z = lds_fft<SPT, false>(x, t0, nt, M, bufA, bufB, Tq);
#pragma unroll
            for (int j = 0; j < SPT; ++j) {
                const int k = tid + kStrainThreads * j;
                const float2 X = k < M ? rfft_bin(Z, k, M, Tq, fdt) : make_float2(0.0f, 0.0f);
#pragma unroll
                for (int cc = 0; cc < NDB; ++cc) if (cc == c) spec[cc][j] = X;
            }
            const float2 XM = rfft_bin(Z, M, M, Tq, fdt);
#pragma unroll
            for (int cc = 0; cc < NDB; ++cc) if (cc == c) specM[cc] = XM;
            __syncthreads();
        }
    }

    // frequency-domain contraction.  Pass A: the base kernels bk_b(k) of every bin from the
    // forward spectra (HBM) and the backward spectra (registers) into LDS (over the FFT
    // buffers); thread tid owns k = tid + 256 j, k = M goes to thread M % 256.
    float2 *bkL = bufA;                                  // [pc.nbk_lds][M + 1]
    const float2 *Fp = F + (size_t)p * NDF * pc.fstride;
    const int need = pc.need_bk;
    const int kM_owner = M & (kStrainThreads - 1);
    constexpr int NJ = SPT + 1;                          // bins per thread (the last: k = M)
    constexpr int JB = NJ < 3 ? NJ : 3;                  // bins whose loads are issued together
#pragma unroll
    for (int j0 = 0; j0 < NJ; j0 += JB) {
#pragma clang fp contract(fast)
        float2 fv[JB][NDF];
#pragma unroll
        for (int jj = 0; jj < JB; ++jj) {
            const int j = j0 + jj;
            const int k = j < SPT ? tid + kStrainThreads * j : M;
            const bool ok = j < SPT ? (k < M) : (j == SPT && tid == kM_owner);
#pragma unroll
            for (int c = 0; c < NDF; ++c)
                fv[jj][c] = (j < NJ && ok) ? Fp[(size_t)c * pc.fstride + k] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int jj = 0; jj < JB; ++jj) {
            const int j = j0 + jj;
            if (j >= NJ) continue;
            const int k = j < SPT ? tid + kStrainThreads * j : M;
            if (j < SPT ? (k >= M) : (tid != kM_owner)) continue;
            const float2 *f = fv[jj];
            float2 b[NDB];
#pragma unroll
            for (int c = 0; c < NDB; ++c) b[c] = j < SPT ? spec[c][j < SPT ? j : 0] : specM[c];
            float2 bk0, bk1 = make_float2(0.f, 0.f), bk3 = bk1, bk4 = bk1, bk5 = bk1;
            if (NDF == 1) {
                const float2 pr = cmulf(f[0], b[0]);
                bk0 = make_float2(-pr.x, -pr.y);
                bk1 = make_float2(-2.0f * pr.x, -2.0f * pr.y);
            } else {
                float2 trf = make_float2(0.f, 0.f), trb = trf;
#pragma unroll
                for (int m = 0; m < NDF; ++m) { trf.x += E.uf[m] * f[m].x; trf.y += E.uf[m] * f[m].y; }
#pragma unroll
                for (int n = 0; n < NDB; ++n) { trb.x += E.ub[n] * b[n].x; trb.y += E.ub[n] * b[n].y; }
                const float2 l = cmulf(trf, trb);
                bk0 = make_float2(-l.x, -l.y);
                if (need & 2) {
                    float2 mu = make_float2(0.f, 0.f);
#pragma unroll
                    for (int m = 0; m < NDF; ++m) {
                        float2 g = make_float2(0.f, 0.f);
#pragma unroll
                        for (int n = 0; n < NDB; ++n) { g.x += E.Amu[m][n] * b[n].x; g.y += E.Amu[m][n] * b[n].y; }
                        const float2 t = cmulf(f[m], g);
                        mu.x += t.x; mu.y += t.y;
                    }
                    bk1 = make_float2(-2.0f * mu.x, -2.0f * mu.y);
                }
                if (need & (8 | 16 | 32)) {
                    float2 fe[6], be[6];
#pragma unroll
                    for (int i = 0; i < 6; ++i) {
                        fe[i] = make_float2(0.f, 0.f);
                        be[i] = make_float2(0.f, 0.f);
#pragma unroll
                        for (int m = 0; m < NDF; ++m) { fe[i].x += E.Qf[i][m] * f[m].x; fe[i].y += E.Qf[i][m] * f[m].y; }
#pragma unroll
                        for (int n = 0; n < NDB; ++n) { be[i].x += S.Q[i][n] * b[n].x; be[i].y += S.Q[i][n] * b[n].y; }
                    }
                    bk3 = cmulf(make_float2(fe[1].x + fe[0].x, fe[1].y + fe[0].y),
                                make_float2(be[1].x + be[0].x, be[1].y + be[0].y));
                    const float2 q5 = cmulf(be[4], fe[4]), q4 = cmulf(be[3], fe[3]);
                    bk4 = make_float2(4.0f * (q5.x + q4.x), 4.0f * (q5.y + q4.y));
                    // k_c exactly as kernel.f90:564-567 writes it
                    const float2 c1 = cmulf(make_float2(fe[0].x + fe[1].x, fe[0].y + fe[1].y), be[2]);
                    const float2 c2 = cmulf(cmulf(be[0], be[1]), fe[2]);
                    bk5 = make_float2(c1.x + c2.x, c1.y + c2.y);
                }
            }
            const int stride = M + 1;
            if (pc.bk_lds[0] >= 0) bkL[pc.bk_lds[0] * stride + k] = bk0;
            if (pc.bk_lds[1] >= 0) bkL[pc.bk_lds[1] * stride + k] = bk1;
            if (pc.bk_lds[3] >= 0) bkL[pc.bk_lds[3] * stride + k] = bk3;
            if (pc.bk_lds[4] >= 0) bkL[pc.bk_lds[4] * stride + k] = bk4;
            if (pc.bk_lds[5] >= 0) bkL[pc.bk_lds[5] * stride + k] = bk5;
            // combinations read by single physical kernels
            if (pc.bk_lds[6] >= 0)
                bkL[pc.bk_lds[6] * stride + k] = make_float2(bk1.x - 2.0f * bk0.x, bk1.y - 2.0f * bk0.y);
            if (pc.bk_lds[7] >= 0)
                bkL[pc.bk_lds[7] * stride + k] = make_float2(bk0.x - bk3.x - bk5.x, bk0.y - bk3.y - bk5.y);
            if (pc.bk_lds[8] >= 0)
                bkL[pc.bk_lds[8] * stride + k] = make_float2(bk1.x - 2.0f * bk0.x - bk4.x, bk1.y - 2.0f * bk0.y - bk4.y);
        }
    }
    // Pass B: sum_k Re(bk_b(k) H_h(k)) in FP64 for every slot (h, b) of the receiver; the
    // slots are ordered by h so that each window spectrum is read once per bin.  A thread
    // reads only the bins it wrote in pass A (the same k = tid + 256 j, and k = M on the
    // same thread), so no barrier separates the passes.
    const RecDesc &rd = recs[r];
    const int lane = tid & 63, wave = tid >> 6;
    // window spectra of the next group are loaded while the current one is summed
    float2 hv[NJ], hn[NJ];
    auto load_h = [&](int hh, float2 (&h)[NJ]) {
        const float2 *Hh = H + (size_t)hh * nw;
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int k = j < SPT ? tid + kStrainThreads * j : M;
            const bool ok = j < SPT ? (k < M) : (tid == kM_owner);
            h[j] = ok ? Hh[k] : make_float2(0.f, 0.f);
        }
    };
    if (rd.ngroup > 0) load_h(rd.grp_h[0], hv);
    for (int g = 0; g < rd.ngroup; ++g) {
        const int s0 = rd.grp_first[g], ng = rd.grp_n[g];
        if (g + 1 < rd.ngroup) load_h(rd.grp_h[g + 1], hn);
        int off[kSlotGroup];
#pragma unroll
        for (int t = 0; t < kSlotGroup; ++t) off[t] = t < ng ? rd.slot_lds[s0 + t] * (M + 1) : 0;
        double a2[kSlotGroup];
#pragma unroll
        for (int t = 0; t < kSlotGroup; ++t) a2[t] = 0.0;
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int k = j < SPT ? tid + kStrainThreads * j : M;
            if (j < SPT ? (k >= M) : (tid != kM_owner)) continue;
#pragma unroll
            for (int t = 0; t < kSlotGroup; ++t)
                if (t < ng) {
                    const float2 v = bkL[off[t] + k];
                    // Re(v h) rounded once in fp32 (the inputs carry fp32 rounding anyway),
                    // accumulated in fp64
                    a2[t] += (double)__builtin_fmaf(v.x, hv[j].x, -(v.y * hv[j].y));
                }
        }
#pragma unroll
        for (int j = 0; j < NJ; ++j) hv[j] = hn[j];
#pragma unroll
        for (int t = 0; t < kSlotGroup; ++t)
            if (t < ng) {
                const double v = wave_sum_lane63(a2[t]);
                if (lane == 63) slot_red[wave * pc.max_slots + s0 + t] = v;
            }
    }
    __syncthreads();
    if (tid < rd.nslot) {
        double t = 0.0;
        for (int w = 0; w < kStrainThreads / 64; ++w) t += slot_red[w * pc.max_slots + tid];
        slotval[tid] = t;
    }
    __syncthreads();
    // physical kernels (kernel.f90:665-832) with the model at the point (fp32 values)
    const int mi = model_idx[p];
    const float vp = model[0 * ndepth + mi], vs = model[1 * ndepth + mi], rho_f = model[2 * ndepth + mi];
    const float phi_m = model[3 * ndepth + mi], xi_m = model[4 * ndepth + mi], eta_m = model[5 * ndepth + mi];
    const double rho = rho_f;
    // dampen_field after every processed receiver (slave.f90:701-706): the base kernels of
    // receiver r are scaled by the receiver and source factors of r and of every later
    // receiver processed in this step (readfields.f90:3184-3204)
    double damp = 1.0;
    if (pc.damp_radius > 0.0 && active) {
        const double x0 = xyz[3 * (size_t)p], x1 = xyz[3 * (size_t)p + 1], x2 = xyz[3 * (size_t)p + 2];
        const double ds = sqrt((x0 - pc.src_r[0]) * (x0 - pc.src_r[0]) + (x1 - pc.src_r[1]) * (x1 - pc.src_r[1])
                               + (x2 - pc.src_r[2]) * (x2 - pc.src_r[2]));
        const double fs = ds < pc.damp_radius ? ds / pc.damp_radius : 1.0;
        for (int r2 = r; r2 < R; ++r2) {
            if (!active[(size_t)(p / P) * R + r2]) continue;
            const double *c = rec_xyz + 3 * r2;
            const double dr = sqrt((x0 - c[0]) * (x0 - c[0]) + (x1 - c[1]) * (x1 - c[1]) + (x2 - c[2]) * (x2 - c[2]));
            damp = damp * (dr < pc.damp_radius ? dr / pc.damp_radius : 1.0) * fs;
        }
    }
    for (int kk = tid; kk < rd.nkernel; kk += kStrainThreads) {
        const int k = rd.first_kernel + kk;
        const KernDesc &kd = kern[k];
        double b[kNumBk];
        for (int qq = 0; qq < kNumBk; ++qq) b[qq] = kd.slot[qq] >= 0 ? slotval[kd.slot[qq]] * damp : 0.0;
        if (base_out) {
            for (int qq = 0; qq < 6; ++qq) base_out[((size_t)p * nk + k) * 6 + qq] = b[qq];
            continue;
        }
        // (the receiver's slots use the combinations: b[6] = mu - 2 lambda, b[7] = lambda
        // - a - c, b[8] = mu - 2 lambda - b)
        double v = 0.0;
        float vsh = vs, vph = vp, vsv = vsh / sqrtf(xi_m), vpv = vph * sqrtf(phi_m);
        switch (kd.param) {
        case 0: v = relative ? b[0] * (double)(rho_f * (vp * vp - 2 * (vs * vs))) : b[0]; break;
        case 3: v = relative ? b[1] * (double)(rho_f * (vs * vs)) : b[1]; break;
        case 1: v = relative ? 2.0 * rho * (double)(vp * vp) * b[0] : 2.0 * rho * (double)vp * b[0]; break;
        case 2: v = (relative ? 2.0 * rho * (double)(vs * vs) : 2.0 * rho * (double)vs) * b[6]; break;
        case 4: v = (relative ? 2.0 * rho * (double)(vsh * vsh) : 2.0 * rho * (double)vsh)
                    * (b[8] + 2 * b[5] * (double)(1 - eta_m)); break;
        case 5: v = (relative ? 2.0 * rho * (double)(vsv * vsv) : 2.0 * rho * (double)vsv) * b[4]; break;
        case 6: v = (relative ? 2.0 * rho * (double)(vph * vph) : 2.0 * rho * (double)vph) * (b[3] + b[5] * (double)eta_m); break;
        case 7: v = (relative ? 2.0 * rho * (double)(vpv * vpv) : 2.0 * rho * (double)vpv) * b[7]; break;
        case 8: v = relative ? 2.0 * rho * (double)(vph * vph - vsh * vsh) * b[5] * (double)eta_m
                             : 2.0 * rho * (double)(vph * vph - vsh * vsh) * b[5]; break;
        default: v = 0.0;
        }
        phys[(size_t)p * nk + k] = v;
    }
}

// sort keys: element of every point (fwd) / element * R + receiver of every active pair
// with a located element (bwd, compacted through an atomic counter); keys past the last
// element sort last
// first item of every database page in element-sorted keys: out[b] = number of keys
// below b * page_elems * scale, b = 0 .. npages (keys are element * scale + minor key;
// keys of unlocated items sort last)
__global__ void k_page_bounds(int n, const int *keys, int page_elems, int npages, int scale, int *out)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > npages) return;
    const long long e = (long long)b * page_elems * scale;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = lo + ((hi - lo) >> 1);
        if (keys[mid] < e) lo = mid + 1; else hi = mid;
    }
    out[b] = lo;
}

// (points of a voxel without an active receiver - all travel-time masked or converged -
// get the key nelem too: their forward spectra would never be read)
__global__ void k_keys_points(int n, const Located *loc, int nelem, const uint8_t *active, int P, int R,
                              int *keys, int *vals)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int e = loc[i].elem;
    bool any = active == nullptr;
    if (active)
        for (int r = 0; r < R; ++r) any |= active[(size_t)(i / P) * R + r] != 0;
    keys[i] = (e < 0 || !any) ? nelem : e;
    vals[i] = i;
}

__global__ void k_keys_pairs(int npairs, int R, int P, const uint8_t *active, const Located *loc,
                             int *keys, int *vals, int *count)
{
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npairs) return;
    int p = q / R, r = q - (q / R) * R;
    if (active && !active[(p / P) * R + r]) return;
    int e = loc[q].elem;
    if (e < 0) return;
    int pos = atomicAdd(count, 1);
    keys[pos] = e * R + r;
    vals[pos] = q;
}

// Algorithmic-byte accounting from the element-sorted keys of a launch: counters[0] +=
// distinct keys below nkeys (elements read), counters[1] += keys below nkeys (items)
__global__ void k_count_runs(int n, const int *keys, int nkeys, unsigned long long *counters)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = i < n ? keys[i] : nkeys;
    const bool valid = k < nkeys;
    const bool start = valid && (i == 0 || keys[i - 1] != k);
    const unsigned long long nv = __popcll(__ballot(valid)), ns = __popcll(__ballot(start));
    if ((threadIdx.x & 63) == 0) {
        if (ns) atomicAdd(counters, ns);
        if (nv) atomicAdd(counters + 1, nv);
    }
}

// points of the launch with at least one evaluated receiver: counter += P per voxel
__global__ void k_count_points(int nwork, int R, int P, const uint8_t *active, unsigned long long *counter)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    bool any = false;
    if (w < nwork)
        for (int r = 0; r < R; ++r) any |= active[(size_t)w * R + r] != 0;
    const unsigned long long n = __popcll(__ballot(any));
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(counter, n * (unsigned long long)P);
}

// ------------------------------------------------------------------------------------
// Monte Carlo bookkeeping
// ------------------------------------------------------------------------------------

// slave.f90:501-511: loop while not all converged; draw points; exit when any kernel
// has niterations == max_iter.  Produces the ordered list of voxels to evaluate.
__global__ void k_step_begin(int nvox, int nk, int max_iter, VoxState st, const uint8_t *live,
                             uint8_t *working)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    working[v] = 0;
    if (st.done[v] || !live[v]) return;
    bool all = true, anymax = false;
    for (int i = 0; i < st.nbasis * nk; ++i)
        if (!st.conv[(size_t)v * st.nbasis * nk + i]) all = false;
    for (int k = 0; k < nk; ++k)
        if (st.niter[(size_t)v * nk + k] == max_iter) anymax = true;
    if (all) { st.done[v] = 1; return; }
    st.draws[v] += 1;
    if (anymax) { st.done[v] = 1; return; }
    working[v] = 1;
}

// single-workgroup ordered compaction of the working flags
__global__ __launch_bounds__(1024) void k_compact(int n, const uint8_t *flag, int *list, int *count)
{
    __shared__ int part[1024];
    const int tid = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b0 = tid * per, b1 = min(n, b0 + per);
    int c = 0;
    for (int i = b0; i < b1; ++i) c += flag[i] ? 1 : 0;
    part[tid] = c;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int y = tid >= o ? part[tid - o] : 0;
        __syncthreads();
        part[tid] += y;
        __syncthreads();
    }
    int pos = part[tid] - c;
    for (int i = b0; i < b1; ++i) if (flag[i]) list[pos++] = i;
    if (tid == 1023) *count = part[1023];
}

// receivers evaluated this step: not masked and not all converged in every basis
// function (slave.f90:539-584, allallconverged mc_integration.f90:322-340)
__global__ void k_receiver_active(int nwork, int R, int nk, const int *work_voxel,
                                  const uint8_t *masked, const int *rec_first, const int *rec_count,
                                  VoxState st, uint8_t *active)
{
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nwork * R) return;
    int w = gid / R, r = gid - (gid / R) * R;
    int v = work_voxel[w];
    uint8_t a = 0;
    if (!masked[(size_t)v * R + r]) {
        for (int b = 0; b < st.nbasis && !a; ++b)
            for (int k = rec_first[r]; k < rec_first[r] + rec_count[r]; ++k)
                if (!st.conv[((size_t)v * st.nbasis + b) * nk + k]) { a = 1; break; }
    }
    active[gid] = a;
}

// check_montecarlo_integral (mc_integration.f90:69-143) of every basis function and the
// niterations bookkeeping.  A kernel is evaluated while it is not converged in every
// basis function (allisconverged, mc_integration.f90:343-359); basis function b sees the
// physical kernel of the weighted base kernels, phys x w_b (slave.f90:723-744).
__global__ void k_mc(int nwork, int P, int nk, int R, const int *work_voxel, const double *phys,
                     const double *bweights /* [nwork * P][4] of tetrahedra, nbasis 4; null = 1 */,
                     const uint8_t *masked, const uint8_t *active, const int *k_rec,
                     const int *k_nbk, const double *volume, const double *allowed_var, double allowed_rel, VoxState st,
                     unsigned long long *evals, int *err)
{
    const int w = blockIdx.x;
    if (w >= nwork) return;
    const int v = work_voxel[w];
    const int nb = st.nbasis;
    const int64_t n_old = st.nmodels[v];
    const int64_t n = n_old + P;
    const double V = volume[v];
    unsigned long long my_evals = 0;
    for (int k = threadIdx.x; k < nk; k += blockDim.x) {
        const size_t vk = (size_t)v * nk + k;
        const int r = k_rec[k];
        const bool msk = masked[(size_t)v * R + r] != 0;
        bool was_conv = true;
        for (int b = 0; b < nb; ++b)
            if (!st.conv[((size_t)v * nb + b) * nk + k]) was_conv = false;
        if (msk) st.niter[vk] += 1;
        else if (active[(size_t)w * R + r] && !was_conv) { st.niter[vk] += k_nbk[k]; my_evals += P; }
        if (was_conv) continue;
        for (int b = 0; b < nb; ++b) {
            const size_t vbk = ((size_t)v * nb + b) * nk + k;
            if (st.conv[vbk]) continue;
            double s = 0.0, s2 = 0.0;
            if (!msk) {
                for (int p = 0; p < P; ++p) {
                    const double wt = bweights ? bweights[((size_t)w * P + p) * 4 + b] : 1.0;
                    s += phys[((size_t)w * P + p) * nk + k] * wt;
                }
                for (int p = 0; p < P; ++p) {
                    const double wt = bweights ? bweights[((size_t)w * P + p) * 4 + b] : 1.0;
                    double x = phys[((size_t)w * P + p) * nk + k] * wt;
                    s2 += x * x;
                }
            }
            double fs = st.fsum[vbk] + s, f2 = st.f2sum[vbk] + s2;
            st.fsum[vbk] = fs;
            st.f2sum[vbk] = f2;
            double integral = fs * V / (double)n, variance;
            bool conv = false;
            if (n > 2) {
                variance = V * V / (double)(n - 1) * ((f2 / (double)n) - (fs / (double)n) * (fs / (double)n));
                if (variance < allowed_var[v]) conv = true;
                else if (sqrt(variance) < allowed_rel * fabs(integral)) conv = true;
            } else {
                variance = integral * 1e10;
            }
            st.integral[vbk] = integral;
            st.variance[vbk] = variance;
            st.conv[vbk] = conv ? 1 : 0;
            if (integral != integral || variance != variance) atomicOr(err, 2);
        }
    }
    // block sum of evaluations
    __shared__ unsigned long long red[16];
    unsigned long long t = my_evals;
    for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int i = 0; i < (int)(blockDim.x + 63) / 64; ++i) tot += red[i];
        atomicAdd(evals, tot);
        st.nmodels[v] = n;
        st.steps[v] += 1;
    }
}


}  // namespace mck
