/*
 * mck_oracle_core.c - FP64 CPU restatement of the element-level routines of MC Kernel's
 * hot path.  TEST INFRASTRUCTURE: the checker for the HIP implementation, never part of
 * the product (see mck_oracle.h).  Each function cites the reference routine it
 * restates (paths relative to the reference repository).
 */
#include "mck_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------
 * Spectral basis (src/spectral_basis.f90)
 * GLL nodes are the zeros of P_{n-1}^{(1,1)} plus +-1 (zelegl, spectral_basis.f90:413-452);
 * GLJ(0,1) nodes are the zeros of P_{n-1}^{(1,2)} plus +-1 (zemngl2, :601-642).  Both are
 * found here by Newton iteration with deflation on the Jacobi recurrence.
 * ---------------------------------------------------------------------------------- */
static void jacobi_pd(int n, double a, double b, double x, double *p, double *dp)
{
    /* value of P_n^{(a,b)}(x) and its derivative (n+a+b+1)/2 P_{n-1}^{(a+1,b+1)}(x) */
    double p0 = 1.0, p1, pn = 1.0;
    int k;
    if (n == 0) { *p = 1.0; *dp = 0.0; return; }
    p1 = 0.5 * ((a + b + 2.0) * x + (a - b));
    pn = p1;
    for (k = 2; k <= n; ++k) {
        double c  = 2.0 * k + a + b;
        double a1 = 2.0 * k * (k + a + b) * (c - 2.0);
        double a2 = (c - 1.0) * (a * a - b * b);
        double a3 = (c - 2.0) * (c - 1.0) * c;
        double a4 = 2.0 * (k + a - 1.0) * (k + b - 1.0) * c;
        pn = ((a2 + a3 * x) * p1 - a4 * p0) / a1;
        p0 = p1; p1 = pn;
    }
    *p = pn;
    if (dp) {
        double q, dq;
        jacobi_pd(n - 1, a + 1.0, b + 1.0, x, &q, &dq);
        *dp = 0.5 * (n + a + b + 1.0) * q;
    }
}

static void jacobi_zeros(int n, double a, double b, double *z)
{
    int k, it, i;
    for (k = 0; k < n; ++k) {
        double r = -cos((2.0 * k + 1.0) * M_PI / (2.0 * n));
        if (k > 0) r = 0.5 * (r + z[k - 1]);
        for (it = 0; it < 100; ++it) {
            double p, dp, s = 0.0, delta;
            for (i = 0; i < k; ++i) s += 1.0 / (r - z[i]);
            jacobi_pd(n, a, b, r, &p, &dp);
            delta = -p / (dp - s * p);
            r += delta;
            if (fabs(delta) < 1e-17) break;
        }
        z[k] = r;
    }
}

void ora_gll_points(int n, double *x)
{
    x[0] = -1.0; x[n] = 1.0;
    if (n > 1) jacobi_zeros(n - 1, 1.0, 1.0, x + 1);
}

void ora_glj_points(int n, double *x)
{
    x[0] = -1.0; x[n] = 1.0;
    if (n > 1) jacobi_zeros(n - 1, 1.0, 2.0, x + 1);
}

/* G(i,j) = l_i'(x_j) (def_lagrange_derivs_gll/glj, spectral_basis.f90:50-100),
 * barycentric form of the Lagrange derivative. */
void ora_lagrange_derivs(int n, const double *x, double *G)
{
    int i, j, m;
    double lam[32];
    for (i = 0; i <= n; ++i) {
        double p = 1.0;
        for (m = 0; m <= n; ++m) if (m != i) p *= (x[i] - x[m]);
        lam[i] = 1.0 / p;
    }
    for (j = 0; j <= n; ++j)
        for (i = 0; i <= n; ++i) {
            double v = 0.0;
            if (i == j) {
                for (m = 0; m <= n; ++m) if (m != i) v += 1.0 / (x[i] - x[m]);
            } else {
                v = lam[i] / (lam[j] * (x[j] - x[i]));
            }
            G[i + j * (n + 1)] = v;
        }
}

/* l_i(x) as in lagrange_interpol_2D_td (spectral_basis.f90:181-314) */
void ora_lagrange_basis(int n, const double *pts, double x, double *l)
{
    int i, m;
    for (i = 0; i <= n; ++i) {
        double v = 1.0;
        for (m = 0; m <= n; ++m) {
            if (m == i) continue;
            v = v * (x - pts[m]) / (pts[i] - pts[m]);
        }
        l[i] = v;
    }
}

/* ------------------------------------------------------------------------------------
 * Halton sequence (src/halton_sequence.f90:109-153).  Index of point j (0-based) in
 * dimension d: seed(d) + (step + j) * leap(d); digits taken least significant first;
 * base_inv kept in quad precision (qp = __float128), each digit*base_inv product rounded
 * to double and accumulated in double.
 * ---------------------------------------------------------------------------------- */
void ora_halton(int ndim, const int64_t *seed, const int32_t *base, const int64_t *leap,
                int64_t step, int npoints, double *r)
{
    int d, j;
    for (j = 0; j < npoints; ++j)
        for (d = 0; d < ndim; ++d) {
            int64_t s2 = seed[d] + (step + j) * leap[d];
            __float128 binv = (__float128)1 / (__float128)base[d];
            double acc = 0.0;
            while (s2 != 0) {
                int64_t digit = s2 % base[d];
                acc = acc + (double)((__float128)digit * binv);
                binv = binv / (__float128)base[d];
                s2 = s2 / base[d];
            }
            r[j * ndim + d] = acc;
        }
}

/* generate_random_points_tet (src/polytopes.f90:47-115): Rocchini-Cignoni fold.
 * v[4][3] vertices, coords[npoints][3] uniform numbers, out[npoints][3]. */
void ora_tet_points(const double *v, int npoints, const double *coords, double *out)
{
    int ip, i;
    for (ip = 0; ip < npoints; ++ip) {
        double c[4], t;
        c[0] = coords[3 * ip + 0];
        c[1] = coords[3 * ip + 1];
        c[2] = coords[3 * ip + 2];
        if (c[0] + c[1] > 1.0) {
            c[0] = 1.0 - c[0];
            c[1] = 1.0 - c[1];
        }
        if (c[1] + c[2] > 1.0) {
            t = c[2];
            c[2] = 1.0 - c[0] - c[1];
            c[1] = 1.0 - t;
        } else if (c[0] + c[1] + c[2] > 1.0) {
            t = c[2];
            c[2] = c[0] + c[1] + c[2] - 1.0;
            c[0] = 1.0 - c[1] - t;
        }
        c[3] = 1.0 - (c[0] + c[1] + c[2]);
        for (i = 0; i < 3; ++i) {
            double acc = 0.0;
            acc = acc + v[0 * 3 + i] * c[0];
            acc = acc + v[1 * 3 + i] * c[1];
            acc = acc + v[2 * 3 + i] * c[2];
            acc = acc + v[3 * 3 + i] * c[3];
            out[3 * ip + i] = acc;
        }
    }
}

/* ------------------------------------------------------------------------------------
 * Pseudo-random numbers.  The reference draws them from Fortran's random_number, seeded
 * by the run time library (clock / OS entropy), so its streams cannot be reproduced.
 * This build uses a counter-based generator instead: value d of draw idx of a stream is
 * the SplitMix64 output (Steele, Lea and Flood 2014) of the counter
 * 0x5EED_0000 + (3 idx + d + 1) * 0x9E3779B97F4A7C15, its top 53 bits scaled to [0, 1).
 * Stream positions are assigned exactly as the Halton positions are.
 * ---------------------------------------------------------------------------------- */
/* the SplitMix64 output function of a state z */
uint64_t ora_splitmix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double ora_prng(int64_t idx, int d)
{
    uint64_t z = ora_splitmix64(0x5EED0000ULL + ((uint64_t)idx * 3u + (uint64_t)d + 1u) * 0x9E3779B97F4A7C15ULL);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

void ora_prng_points(int64_t step, int npoints, double *r)
{
    int j, d;
    for (j = 0; j < npoints; ++j)
        for (d = 0; d < 3; ++d) r[3 * j + d] = ora_prng(step + j, d);
}

/* ------------------------------------------------------------------------------------
 * Voxels: blocks bounded by two radii, two colatitudes and two longitudes, given by their
 * 8 corners (src/random_voxel.f90).  range[6] = r_min, r_max, colat_min, colat_max,
 * lon_min, lon_max of the corners (cartesian_to_spherical_range and cartesian_to_spherical_points, :166-209).
 * ---------------------------------------------------------------------------------- */
void ora_vox_range(const double *v, double *range)
{
    int i;
    for (i = 0; i < 8; ++i) {
        double x = v[3 * i], y = v[3 * i + 1], z = v[3 * i + 2];
        double r = sqrt(x * x + y * y + z * z), c = acos(z / r), l = atan2(y, x);
        if (i == 0 || r < range[0]) range[0] = r;
        if (i == 0 || r > range[1]) range[1] = r;
        if (i == 0 || c < range[2]) range[2] = c;
        if (i == 0 || c > range[3]) range[3] = c;
        if (i == 0 || l < range[4]) range[4] = l;
        if (i == 0 || l > range[5]) range[5] = l;
    }
}

static void sph_to_xyz(double r, double colat, double lon, double *u)
{
    /* spherical_to_cartesian_point (random_voxel.f90:214-227) */
    u[0] = r * cos(lon) * sin(colat);
    u[1] = r * sin(lon) * sin(colat);
    u[2] = r * cos(colat);
}

/* generate_random_points_vox (random_voxel.f90:45-86): r uniform in r^3, cos(colat)
 * uniform, longitude uniform; coords[npoints][3] uniform numbers */
void ora_vox_points(const double *v, int npoints, const double *coords, double *out)
{
    double g[6];
    int i;
    ora_vox_range(v, g);
    for (i = 0; i < npoints; ++i) {
        double r = pow(coords[3 * i] * (g[1] * g[1] * g[1] - g[0] * g[0] * g[0]) + g[0] * g[0] * g[0], 1.0 / 3.0);
        double c = acos(cos(g[2]) + (cos(g[3]) - cos(g[2])) * coords[3 * i + 1]);
        double l = g[4] + fabs(g[5] - g[4]) * coords[3 * i + 2];
        sph_to_xyz(r, c, l, out + 3 * i);
    }
}

/* get_volume_vox (random_voxel.f90:91-120): shell volume times the area fraction of the
 * latitude-longitude quadrilateral */
double ora_vox_volume(const double *v)
{
    double g[6], shell, frac, lat1, lat2;
    ora_vox_range(v, g);
    lat1 = ORA_PI / 2.0 - g[2];
    lat2 = ORA_PI / 2.0 - g[3];
    shell = (4.0 / 3.0) * ORA_PI * (g[1] * g[1] * g[1] - g[0] * g[0] * g[0]);
    frac = fabs(g[4] - g[5]) * fabs(sin(lat1) - sin(lat2)) / (4.0 * ORA_PI);
    return shell * frac;
}

/* get_center_vox (random_voxel.f90:125-145), as written: the third spherical coordinate
 * (longitude) of the centre is taken from the colatitude range */
void ora_vox_center(const double *v, double *c)
{
    double g[6];
    ora_vox_range(v, g);
    sph_to_xyz((g[1] - g[0]) / 2.0 + g[0], (g[3] - g[2]) / 2.0 + g[2], (g[3] - g[2]) / 2.0 + g[2], c);
}

/* point_in_voxel (random_voxel.f90:232-259): inclusive range test */
int ora_point_in_vox(const double *v, const double *p)
{
    double g[6], r, c, l;
    ora_vox_range(v, g);
    r = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    c = acos(p[2] / r);
    l = atan2(p[1], p[0]);
    return r >= g[0] && r <= g[1] && c >= g[2] && c <= g[3] && l >= g[4] && l <= g[5];
}

/* get_volume_tet (polytopes.f90:495-506) with determinant_4 in quad precision
 * (polytopes.f90:320-347). */
double ora_tet_volume(const double *v)
{
    __float128 a[4][4];
    int i, k;
    __float128 d;
    for (i = 0; i < 4; ++i) {
        for (k = 0; k < 3; ++k) a[i][k] = v[i * 3 + k];
        a[i][3] = 1;
    }
#define A(r, c) a[(r) - 1][(c) - 1]
    d =   A(1,1) * ( A(2,2) * ( A(3,3) * A(4,4) - A(3,4) * A(4,3) )
                   - A(2,3) * ( A(3,2) * A(4,4) - A(3,4) * A(4,2) )
                   + A(2,4) * ( A(3,2) * A(4,3) - A(3,3) * A(4,2) ) )
        - A(1,2) * ( A(2,1) * ( A(3,3) * A(4,4) - A(3,4) * A(4,3) )
                   - A(2,3) * ( A(3,1) * A(4,4) - A(3,4) * A(4,1) )
                   + A(2,4) * ( A(3,1) * A(4,3) - A(3,3) * A(4,1) ) )
        + A(1,3) * ( A(2,1) * ( A(3,2) * A(4,4) - A(3,4) * A(4,2) )
                   - A(2,2) * ( A(3,1) * A(4,4) - A(3,4) * A(4,1) )
                   + A(2,4) * ( A(3,1) * A(4,2) - A(3,2) * A(4,1) ) )
        - A(1,4) * ( A(2,1) * ( A(3,2) * A(4,3) - A(3,3) * A(4,2) )
                   - A(2,2) * ( A(3,1) * A(4,3) - A(3,3) * A(4,1) )
                   + A(2,3) * ( A(3,1) * A(4,2) - A(3,2) * A(4,1) ) );
#undef A
    {
        double dd = (double)d;
        return fabs(dd) / 6.0;
    }
}

/* determinant_4_qp (polytopes.f90:351-380): cofactor expansion along the first row, in
 * quad precision; m[r][c] = a(r+1, c+1) */
static __float128 det4q(__float128 m[4][4])
{
#define A(r, c) m[(r) - 1][(c) - 1]
    return A(1,1) * ( A(2,2) * ( A(3,3) * A(4,4) - A(3,4) * A(4,3) )
                    - A(2,3) * ( A(3,2) * A(4,4) - A(3,4) * A(4,2) )
                    + A(2,4) * ( A(3,2) * A(4,3) - A(3,3) * A(4,2) ) )
         - A(1,2) * ( A(2,1) * ( A(3,3) * A(4,4) - A(3,4) * A(4,3) )
                    - A(2,3) * ( A(3,1) * A(4,4) - A(3,4) * A(4,1) )
                    + A(2,4) * ( A(3,1) * A(4,3) - A(3,3) * A(4,1) ) )
         + A(1,3) * ( A(2,1) * ( A(3,2) * A(4,4) - A(3,4) * A(4,2) )
                    - A(2,2) * ( A(3,1) * A(4,4) - A(3,4) * A(4,1) )
                    + A(2,4) * ( A(3,1) * A(4,2) - A(3,2) * A(4,1) ) )
         - A(1,4) * ( A(2,1) * ( A(3,2) * A(4,3) - A(3,3) * A(4,2) )
                    - A(2,2) * ( A(3,1) * A(4,3) - A(3,3) * A(4,1) )
                    + A(2,3) * ( A(3,1) * A(4,2) - A(3,2) * A(4,1) ) );
#undef A
}

/* point_in_tetrahedron (polytopes.f90:660-715): barycentric coordinates as ratios of
 * quad-precision 4x4 determinants; inside unless one is below -1e-9.  v[4][3], p[3]. */
int ora_point_in_tet(const double *v, const double *p)
{
    __float128 m0[4][4], m[4][4], d0;
    int r, c, i;
    for (c = 0; c < 4; ++c) {
        for (r = 0; r < 3; ++r) m0[r][c] = v[3 * c + r];
        m0[3][c] = 1;
    }
    d0 = det4q(m0);
    if ((d0 < 0 ? -d0 : d0) < (__float128)1.92592994438723585305597794258492732e-34Q) return 0;
    for (i = 0; i < 4; ++i) {
        __float128 di;
        for (r = 0; r < 4; ++r) for (c = 0; c < 4; ++c) m[r][c] = m0[r][c];
        for (r = 0; r < 3; ++r) m[r][i] = p[r];
        di = det4q(m) / d0;
        if (di < (__float128)(-1e-9)) return 0;
    }
    return 1;
}

/* gfortran's inline NORM2: scaled sum of squares (BLAS dnrm2 style). */
double ora_norm2(int n, const double *x)
{
    double scale = 0.0, res = 1.0;
    int i;
    for (i = 0; i < n; ++i) {
        if (x[i] != 0.0) {
            double ax = fabs(x[i]);
            if (scale < ax) {
                double val = scale / ax;
                res = 1.0 + res * val * val;
                scale = ax;
            } else {
                double val = ax / scale;
                res = res + val * val;
            }
        }
    }
    return sqrt(res) * scale;
}

/* ------------------------------------------------------------------------------------
 * Element mappings (src/finite_elem_mapping.f90).  nodes is Fortran nodes(4,2):
 * nodes[i] = s of corner i+1, nodes[4+i] = z of corner i+1.
 * ---------------------------------------------------------------------------------- */
#define NS(i) nodes[(i) - 1]
#define NZ(i) nodes[4 + (i) - 1]

static void theta_r(const double *nodes, double *theta, double *r)
{
    int i;
    for (i = 1; i <= 4; ++i) {
        r[i - 1] = sqrt(NS(i) * NS(i) + NZ(i) * NZ(i));
        if (r[i - 1] != 0.0) theta[i - 1] = acos(NZ(i) / r[i - 1]);
        else theta[i - 1] = 0.0;
    }
}

static void compute_ab(double *a, double *b, double s1, double z1, double s2, double z2)
{
    *a = sqrt(fabs((s2 * s2 * (z1 * z1) - z2 * z2 * (s1 * s1)) / (z1 * z1 - z2 * z2)));
    *b = sqrt(fabs((z1 * z1 * (s2 * s2) - z2 * z2 * (s1 * s1)) / (s2 * s2 - s1 * s1)));
}

static double compute_theta(double s, double z, double a, double b)
{
    if (s >= 0.0) return atan2(z * a, s * b);
    if (z > 0) return ORA_PI / 2;
    if (z < 0) return -ORA_PI / 2;
    return 0.0; /* undefined in the reference (uninitialised); not reached for valid meshes */
}

static void params_semino(const double *nodes, double *a, double *b, double *tbar,
                          double *dth)
{
    double t3, t4;
    compute_ab(a, b, NS(4), NZ(4), NS(3), NZ(3));
    t3 = compute_theta(NS(3), NZ(3), *a, *b);
    t4 = compute_theta(NS(4), NZ(4), *a, *b);
    *tbar = (t4 + t3) / 2;
    *dth = t3 - t4;
}

static void params_semiso(const double *nodes, double *a, double *b, double *tbar,
                          double *dth)
{
    double t1, t2;
    compute_ab(a, b, NS(1), NZ(1), NS(2), NZ(2));
    t2 = compute_theta(NS(2), NZ(2), *a, *b);
    t1 = compute_theta(NS(1), NZ(1), *a, *b);
    *tbar = (t1 + t2) / 2;
    *dth = t2 - t1;
}

static void sz_xi(double *s, double *z, double xi, double a, double b, double tbar,
                  double dth)
{
    *s = a * cos(tbar + xi * dth / 2);
    *z = b * sin(tbar + xi * dth / 2);
}

static void dsdxi_dzdxi(double *dsdxi, double *dzdxi, double xi, double a, double b,
                        double tbar, double dth)
{
    *dsdxi = -a * dth * sin(tbar + xi * dth / 2) / 2;
    *dzdxi = b * dth * cos(tbar + xi * dth / 2) / 2;
}

static void map_semino(double xi, double eta, const double *nodes, double *sz)
{
    double a, b, tb, dt, sb, zb, st, zt;
    params_semino(nodes, &a, &b, &tb, &dt);
    sb = ((1 + xi) * NS(2) + (1 - xi) * NS(1)) / 2;
    zb = ((1 + xi) * NZ(2) + (1 - xi) * NZ(1)) / 2;
    sz_xi(&st, &zt, xi, a, b, tb, dt);
    sz[0] = (sb + st) / 2 + eta * (st - sb) / 2;
    sz[1] = (zb + zt) / 2 + eta * (zt - zb) / 2;
}

static void map_semiso(double xi, double eta, const double *nodes, double *sz)
{
    double a, b, tb, dt, sb, zb, st, zt;
    params_semiso(nodes, &a, &b, &tb, &dt);
    st = ((1 + xi) * NS(3) + (1 - xi) * NS(4)) / 2;
    zt = ((1 + xi) * NZ(3) + (1 - xi) * NZ(4)) / 2;
    sz_xi(&sb, &zb, xi, a, b, tb, dt);
    sz[0] = (sb + st) / 2 + eta * (st - sb) / 2;
    sz[1] = (zb + zt) / 2 + eta * (zt - zb) / 2;
}

static void jac_semino(double xi, double eta, const double *nodes, double *J)
{
    double a, b, tb, dt, sb, zb, st, zt, dsb, dzb, dst, dzt;
    params_semino(nodes, &a, &b, &tb, &dt);
    sb = ((1 + xi) * NS(2) + (1 - xi) * NS(1)) / 2;
    zb = ((1 + xi) * NZ(2) + (1 - xi) * NZ(1)) / 2;
    sz_xi(&st, &zt, xi, a, b, tb, dt);
    dsb = (NS(2) - NS(1)) / 2;
    dzb = (NZ(2) - NZ(1)) / 2;
    dsdxi_dzdxi(&dst, &dzt, xi, a, b, tb, dt);
    J[0] = (dsb + dst) / 2 + eta * (dst - dsb) / 2;   /* J(1,1) */
    J[2] = (st - sb) / 2;                             /* J(1,2) */
    J[1] = (dzb + dzt) / 2 + eta * (dzt - dzb) / 2;   /* J(2,1) */
    J[3] = (zt - zb) / 2;                             /* J(2,2) */
}

static void jac_semiso(double xi, double eta, const double *nodes, double *J)
{
    double a, b, tb, dt, sb, zb, st, zt, dsb, dzb, dst, dzt;
    params_semiso(nodes, &a, &b, &tb, &dt);
    st = ((1 + xi) * NS(3) + (1 - xi) * NS(4)) / 2;
    zt = ((1 + xi) * NZ(3) + (1 - xi) * NZ(4)) / 2;
    sz_xi(&sb, &zb, xi, a, b, tb, dt);
    dsdxi_dzdxi(&dsb, &dzb, xi, a, b, tb, dt);
    dst = (NS(3) - NS(4)) / 2;
    dzt = (NZ(3) - NZ(4)) / 2;
    J[0] = (dsb + dst) / 2 + eta * (dst - dsb) / 2;
    J[2] = (st - sb) / 2;
    J[1] = (dzb + dzt) / 2 + eta * (dzt - dzb) / 2;
    J[3] = (zt - zb) / 2;
}

static void map_spheroid(double xi, double eta, const double *nodes, double *sz)
{
    double th[4], r[4];
    theta_r(nodes, th, r);
    sz[0] = (1 + eta) * r[3] / 2 * sin(((1 - xi) * th[3] + (1 + xi) * th[2]) / 2)
          + (1 - eta) * r[0] / 2 * sin(((1 - xi) * th[0] + (1 + xi) * th[1]) / 2);
    sz[1] = (1 + eta) * r[3] / 2 * cos(((1 - xi) * th[3] + (1 + xi) * th[2]) / 2)
          + (1 - eta) * r[0] / 2 * cos(((1 - xi) * th[0] + (1 + xi) * th[1]) / 2);
}

static void jac_spheroid(double xi, double eta, const double *nodes, double *J)
{
    double th[4], r[4];
    theta_r(nodes, th, r);
    J[0] = (1 + eta) * r[3] * (th[2] - th[3]) / 4
               * cos(((1 - xi) * th[3] + (1 + xi) * th[2]) / 2)
         + (1 - eta) * r[0] * (th[1] - th[0]) / 4
               * cos(((1 - xi) * th[0] + (1 + xi) * th[1]) / 2);
    J[2] = 0.5 * (r[3] * sin(((1 - xi) * th[3] + (1 + xi) * th[2]) / 2)
                - r[0] * sin(((1 - xi) * th[0] + (1 + xi) * th[1]) / 2));
    J[1] = -(1 + eta) * r[3] * (th[2] - th[3]) / 4
               * sin(((1 - xi) * th[3] + (1 + xi) * th[2]) / 2)
         - (1 - eta) * r[0] * (th[1] - th[0]) / 4
               * sin(((1 - xi) * th[0] + (1 + xi) * th[1]) / 2);
    J[3] = 0.5 * (r[3] * cos(((1 - xi) * th[3] + (1 + xi) * th[2]) / 2)
                - r[0] * cos(((1 - xi) * th[0] + (1 + xi) * th[1]) / 2));
}

static void map_subpar(double xi, double eta, const double *nodes, double *sz)
{
    double xip = 1 + xi, xim = 1 - xi, etap = 1 + eta, etam = 1 - eta, shp[4];
    int i;
    shp[0] = xim * etam / 4;
    shp[1] = xip * etam / 4;
    shp[2] = xip * etap / 4;
    shp[3] = xim * etap / 4;
    sz[0] = 0.0; sz[1] = 0.0;
    for (i = 1; i <= 4; ++i) {
        sz[0] = sz[0] + shp[i - 1] * NS(i);
        sz[1] = sz[1] + shp[i - 1] * NZ(i);
    }
}

static void jac_subpar(double xi, double eta, const double *nodes, double *J)
{
    double xip = 1 + xi, xim = 1 - xi, etap = 1 + eta, etam = 1 - eta;
    double d[4][2];
    int i;
    d[0][0] = -etam / 4; d[1][0] = etam / 4; d[2][0] = etap / 4; d[3][0] = -etap / 4;
    d[0][1] = -xim / 4;  d[1][1] = -xip / 4; d[2][1] = xip / 4;  d[3][1] = xim / 4;
    J[0] = J[1] = J[2] = J[3] = 0.0;
    for (i = 1; i <= 4; ++i) {
        J[0] = J[0] + NS(i) * d[i - 1][0];   /* (1,1) */
        J[2] = J[2] + NS(i) * d[i - 1][1];   /* (1,2) */
        J[1] = J[1] + NZ(i) * d[i - 1][0];   /* (2,1) */
        J[3] = J[3] + NZ(i) * d[i - 1][1];   /* (2,2) */
    }
}

void ora_mapping(double xi, double eta, const double *nodes, int eltype, double *sz)
{
    switch (eltype) {
    case 0: map_spheroid(xi, eta, nodes, sz); break;
    case 1: map_subpar(xi, eta, nodes, sz); break;
    case 2: map_semino(xi, eta, nodes, sz); break;
    default: map_semiso(xi, eta, nodes, sz); break;
    }
}

/* inv_jacobian (finite_elem_mapping.f90:216-234, 281-304, 476-486, 585-595);
 * ij column-major: ij[0]=(1,1) ij[1]=(2,1) ij[2]=(1,2) ij[3]=(2,2) */
void ora_inv_jacobian(double xi, double eta, const double *nodes, int eltype, double *ij)
{
    double J[4], det;
    switch (eltype) {
    case 0: jac_spheroid(xi, eta, nodes, J); break;
    case 1: jac_subpar(xi, eta, nodes, J); break;
    case 2: jac_semino(xi, eta, nodes, J); break;
    default: jac_semiso(xi, eta, nodes, J); break;
    }
    det = J[0] * J[3] - J[1] * J[2];
    ij[0] = J[3] / det;
    ij[1] = -J[1] / det;
    ij[2] = -J[2] / det;
    ij[3] = J[0] / det;
}

/* inv_mapping_* (finite_elem_mapping.f90:137-200, 403-433, 505-535): Newton from the
 * element centre, 10 iterations; convergence threshold is the single-precision literal
 * 1e-7**2 for semino/semiso/subpar and the double-precision 1d-7**2 for spheroids. */
void ora_inv_mapping(double s, double z, const double *nodes, int eltype, double *xieta)
{
    const double thr = (eltype == 0) ? (1e-7 * 1e-7) : (double)(1e-7f * 1e-7f);
    double xi = 0.0, eta = 0.0, sz[2], ij[4], ds, dz;
    int i;
    for (i = 1; i <= 10; ++i) {
        ora_mapping(xi, eta, nodes, eltype, sz);
        ds = s - sz[0];
        dz = z - sz[1];
        if ((ds * ds + dz * dz) / (s * s + z * z) < thr) break;
        ora_inv_jacobian(xi, eta, nodes, eltype, ij);
        xi = xi + ij[0] * ds + ij[2] * dz;
        eta = eta + ij[1] * ds + ij[3] * dz;
    }
    xieta[0] = xi;
    xieta[1] = eta;
}

/* inside_element (finite_elem_mapping.f90:101-141) */
int ora_inside_element(double s, double z, const double *nodes, int eltype, double tol,
                       double *xi, double *eta)
{
    double xe[2];
    ora_inv_mapping(s, z, nodes, eltype, xe);
    *xi = xe[0];
    *eta = xe[1];
    return (xe[0] >= -1 - tol && xe[0] <= 1 + tol && xe[1] >= -1 - tol && xe[1] <= 1 + tol);
}
#undef NS
#undef NZ

/* ------------------------------------------------------------------------------------
 * SEM strain (src/sem_derivatives.f90).  Arrays f[jpol][ipol][nsamp] (time fastest,
 * matching Fortran f(nsamp, 0:npol, 0:npol)).
 * ---------------------------------------------------------------------------------- */
typedef struct {
    int npol, ns;
    const double *G, *GT;       /* G[i + j*(n+1)] column-major like the Fortran arrays */
    const double *xi, *eta;     /* collocation points */
    const double *nodes;
    int eltype;
    double invj[8][8][4];       /* inverse Jacobians at the nodes [ipol][jpol] */
    double s[8][8];             /* s coordinate of the nodes */
} ora_elem;

#define F3(f, t, i, j) (f)[((j) * (E->npol + 1) + (i)) * E->ns + (t)]

static void elem_prepare(ora_elem *E)
{
    int i, j;
    for (i = 0; i <= E->npol; ++i)
        for (j = 0; j <= E->npol; ++j) {
            double sz[2];
            ora_inv_jacobian(E->xi[i], E->eta[j], E->nodes, E->eltype, E->invj[i][j]);
            ora_mapping(E->xi[i], E->eta[j], E->nodes, E->eltype, sz);
            E->s[i][j] = sz[0];
        }
}

/* axisym_gradient_td (sem_derivatives.f90:934-977) -> grad[2][...] */
static void axisym_gradient(const ora_elem *E, const double *f, double *gs, double *gz)
{
    int n = E->npol, ns = E->ns, i, j, k, t;
    size_t sz = (size_t)(n + 1) * (n + 1) * ns;
    double *m1 = (double *)calloc(sz, sizeof(double));
    double *m2 = (double *)calloc(sz, sizeof(double));
    /* mxm_btd(GT, f): m1(:,i,j) = sum_k GT(i,k) f(:,k,j) */
    for (j = 0; j <= n; ++j)
        for (i = 0; i <= n; ++i)
            for (k = 0; k <= n; ++k) {
                double g = E->GT[i + k * (n + 1)];
                for (t = 0; t < ns; ++t) m1[(j * (n + 1) + i) * ns + t] += g * F3(f, t, k, j);
            }
    /* mxm_atd(f, G): m2(:,i,j) = sum_k f(:,i,k) G(k,j) */
    for (j = 0; j <= n; ++j)
        for (i = 0; i <= n; ++i)
            for (k = 0; k <= n; ++k) {
                double g = E->G[k + j * (n + 1)];
                for (t = 0; t < ns; ++t) m2[(j * (n + 1) + i) * ns + t] += F3(f, t, i, k) * g;
            }
    for (j = 0; j <= n; ++j)
        for (i = 0; i <= n; ++i) {
            const double *ij = E->invj[i][j];
            for (t = 0; t < ns; ++t) {
                size_t o = (j * (n + 1) + i) * ns + t;
                gs[o] = ij[0] * m1[o] + ij[1] * m2[o];
                gz[o] = ij[2] * m1[o] + ij[3] * m2[o];
            }
        }
    free(m1);
    free(m2);
}

/* f_over_s_td (sem_derivatives.f90:775-818) with dsdf_axis_td on the axis (:861-895) */
static void f_over_s(const ora_elem *E, const double *f, int axial, double *out)
{
    int n = E->npol, ns = E->ns, i, j, k, t;
    for (j = 0; j <= n; ++j)
        for (i = (axial ? 1 : 0); i <= n; ++i)
            for (t = 0; t < ns; ++t)
                out[(j * (n + 1) + i) * ns + t] = F3(f, t, i, j) / E->s[i][j];
    if (axial) {
        double *m1 = (double *)calloc((size_t)(n + 1) * ns, sizeof(double));
        double *m2 = (double *)calloc((size_t)(n + 1) * ns, sizeof(double));
        for (j = 0; j <= n; ++j)
            for (k = 0; k <= n; ++k) {
                double g1 = E->GT[0 + k * (n + 1)];
                double g2 = E->G[k + j * (n + 1)];
                for (t = 0; t < ns; ++t) {
                    m1[j * ns + t] += g1 * F3(f, t, k, j);
                    m2[j * ns + t] += F3(f, t, 0, k) * g2;
                }
            }
        for (j = 0; j <= n; ++j) {
            const double *ij = E->invj[0][j];
            for (t = 0; t < ns; ++t)
                out[(j * (n + 1) + 0) * ns + t] = ij[0] * m1[j * ns + t] + ij[1] * m2[j * ns + t];
        }
        free(m1);
        free(m2);
    }
}
#undef F3

static void setup_elem(ora_elem *E, int npol, int nsamp, const double *nodes, int eltype,
                       int axial)
{
    static int init = 0;
    static double gll[8], glj[8], G1[64], G2[64], G1T[64], G2T[64];
    int i, j;
    if (!init || npol != 4) {
        ora_gll_points(npol, gll);
        ora_glj_points(npol, glj);
        ora_lagrange_derivs(npol, glj, G1);
        ora_lagrange_derivs(npol, gll, G2);
        for (i = 0; i <= npol; ++i)
            for (j = 0; j <= npol; ++j) {
                G1T[i + j * (npol + 1)] = G1[j + i * (npol + 1)];
                G2T[i + j * (npol + 1)] = G2[j + i * (npol + 1)];
            }
        init = (npol == 4);
    }
    E->npol = npol;
    E->ns = nsamp;
    E->G = G2;
    E->GT = axial ? G1T : G2T;
    E->xi = axial ? glj : gll;
    E->eta = gll;
    E->nodes = nodes;
    E->eltype = eltype;
    elem_prepare(E);
}

/*
 * strain_monopole_td / strain_dipole_td / strain_quadpole_td
 * (sem_derivatives.f90:281-649).  pole 0 = monopole (2 vars), 1 = dipole, 2 = quadpole
 * (3 vars).  u[var][jpol][ipol][nsamp]; strain[6][jpol][ipol][nsamp] (Voigt
 * ss, pp, zz, zp, sz, sp).
 */
void ora_strain_pole(int pole, const double *u, int nsamp, int npol, const double *nodes,
                     int eltype, int axial, double *strain)
{
    ora_elem E;
    size_t n = (size_t)(npol + 1) * (npol + 1) * nsamp, o;
    double *g1s, *g1z, *g2s, *g2z, *g3s, *g3z, *tmp, *fos;
    const double *u1 = u, *u2 = u + n, *u3 = u + 2 * n;
    setup_elem(&E, npol, nsamp, nodes, eltype, axial);
    g1s = (double *)malloc(8 * n * sizeof(double));
    g1z = g1s + n; g2s = g1z + n; g2z = g2s + n; g3s = g2z + n; g3z = g3s + n;
    tmp = g3z + n; fos = tmp + n;
    axisym_gradient(&E, u1, g1s, g1z);
    axisym_gradient(&E, u2, g2s, g2z);
    if (pole == 0) {
        for (o = 0; o < n; ++o) strain[0 * n + o] = g1s[o];
        f_over_s(&E, u1, axial, fos);
        for (o = 0; o < n; ++o) strain[1 * n + o] = fos[o];
        for (o = 0; o < n; ++o) strain[2 * n + o] = g2z[o];
        for (o = 0; o < n; ++o) strain[3 * n + o] = 0.0;
        for (o = 0; o < n; ++o) strain[4 * n + o] = (g1z[o] + g2s[o]) / 2.0;
        for (o = 0; o < n; ++o) strain[5 * n + o] = 0.0;
    } else {
        axisym_gradient(&E, u3, g3s, g3z);
        for (o = 0; o < n; ++o) strain[0 * n + o] = g1s[o];
        for (o = 0; o < n; ++o)
            tmp[o] = (pole == 1) ? (u1[o] - u2[o]) : (u1[o] - 2 * u2[o]);
        f_over_s(&E, tmp, axial, fos);
        for (o = 0; o < n; ++o) strain[1 * n + o] = fos[o];
        for (o = 0; o < n; ++o) strain[2 * n + o] = g3z[o];
        f_over_s(&E, u3, axial, fos);
        if (pole == 1)
            for (o = 0; o < n; ++o) strain[3 * n + o] = -0.5 * (fos[o] + g2z[o]);
        else
            for (o = 0; o < n; ++o) strain[3 * n + o] = -fos[o] - g2z[o] / 2.0;
        for (o = 0; o < n; ++o) strain[4 * n + o] = (g1z[o] + g3s[o]) / 2.0;
        for (o = 0; o < n; ++o)
            tmp[o] = (pole == 1) ? ((u1[o] - u2[o]) / 2) : (0.5 * u2[o] - u1[o]);
        f_over_s(&E, tmp, axial, fos);
        if (pole == 1)
            for (o = 0; o < n; ++o) strain[5 * n + o] = -fos[o] - g2s[o] / 2.0;
        else
            for (o = 0; o < n; ++o) strain[5 * n + o] = fos[o] - g2s[o] / 2.0;
    }
    free(g1s);
}

/* straintrace_monopole_td / _dipole_td / _quadpole_td (sem_derivatives.f90:340-649) */
void ora_straintrace_pole(int pole, const double *u, int nsamp, int npol,
                          const double *nodes, int eltype, int axial, double *tr)
{
    ora_elem E;
    size_t n = (size_t)(npol + 1) * (npol + 1) * nsamp, o;
    double *g1s, *g1z, *g3s, *g3z, *tmp, *fos;
    const double *u1 = u, *u2 = u + n, *u3 = u + 2 * n;
    setup_elem(&E, npol, nsamp, nodes, eltype, axial);
    g1s = (double *)malloc(6 * n * sizeof(double));
    g1z = g1s + n; g3s = g1z + n; g3z = g3s + n; tmp = g3z + n; fos = tmp + n;
    axisym_gradient(&E, u1, g1s, g1z);
    if (pole == 0) {
        axisym_gradient(&E, u2, g3s, g3z);
        f_over_s(&E, u1, axial, fos);
    } else {
        axisym_gradient(&E, u3, g3s, g3z);
        for (o = 0; o < n; ++o)
            tmp[o] = (pole == 1) ? (u1[o] - u2[o]) : (u1[o] - 2 * u2[o]);
        f_over_s(&E, tmp, axial, fos);
    }
    for (o = 0; o < n; ++o) tr[o] = g1s[o] + fos[o] + g3z[o];
    free(g1s);
}

/* ------------------------------------------------------------------------------------
 * rotate_frame_rd (src/rotations.f90:737-782)
 * ---------------------------------------------------------------------------------- */
void ora_rotate_frame(int n, const double *xyz, double lon, double colat, double *s,
                      double *phi, double *z)
{
    int i;
    for (i = 0; i < n; ++i) {
        double x = xyz[3 * i], y = xyz[3 * i + 1], zz = xyz[3 * i + 2];
        double xc = x * cos(lon) + y * sin(lon);
        double yc = -x * sin(lon) + y * cos(lon);
        double zc = zz;
        double xp = xc * cos(colat) - zc * sin(colat);
        double yp = yc;
        double zp = xc * sin(colat) + zc * cos(colat);
        double pc;
        s[i] = sqrt(xp * xp + yp * yp);
        z[i] = zp;
        pc = atan2(yp, xp);
        phi[i] = (pc < 0.0) ? 2.0 * ORA_PI + pc : pc;
    }
}

/* ------------------------------------------------------------------------------------
 * FFT (replaces FFTW in src/fft.f90:375-457): X = dt * DFT(x), x = df * IDFT(X) with the
 * imaginary parts of DC and Nyquist ignored (FFTW c2r semantics).
 * ---------------------------------------------------------------------------------- */
int ora_nextpow2(int n)
{
    int p = 2;
    while (p < n) p *= 2;
    return p;
}

static void cfft(int n, double *re, double *im, int sign)
{
    int i, j, k, m, len;
    for (i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (len = 2; len <= n; len <<= 1) {
        int half = len >> 1;
        for (k = 0; k < half; ++k) {
            double ang = sign * 2.0 * M_PI * k / len;
            double wr = cos(ang), wi = sin(ang);
            for (m = k; m < n; m += len) {
                double xr = re[m + half] * wr - im[m + half] * wi;
                double xi = re[m + half] * wi + im[m + half] * wr;
                re[m + half] = re[m] - xr;
                im[m + half] = im[m] - xi;
                re[m] += xr;
                im[m] += xi;
            }
        }
    }
}

void ora_rfft(int n, double dt, const double *x, double *X)
{
    double *re = (double *)malloc(2 * n * sizeof(double)), *im = re + n;
    int k;
    for (k = 0; k < n; ++k) { re[k] = x[k]; im[k] = 0.0; }
    cfft(n, re, im, -1);
    for (k = 0; k <= n / 2; ++k) {
        X[2 * k] = re[k] * dt;
        X[2 * k + 1] = im[k] * dt;
    }
    free(re);
}

void ora_irfft(int n, double df, const double *X, double *x)
{
    double *re = (double *)malloc(2 * n * sizeof(double)), *im = re + n;
    int k;
    re[0] = X[0]; im[0] = 0.0;
    re[n / 2] = X[n]; im[n / 2] = 0.0;
    for (k = 1; k < n / 2; ++k) {
        re[k] = X[2 * k]; im[k] = X[2 * k + 1];
        re[n - k] = X[2 * k]; im[n - k] = -X[2 * k + 1];
    }
    cfft(n, re, im, +1);
    for (k = 0; k < n; ++k) x[k] = re[k] * df;
    free(re);
}

/* taperandzeropad_1d (src/fft.f90:512-545); window evaluated in single precision with
 * the real (fp32) decay constant D = 3.3 */
void ora_taperandzeropad(int nin, int nout, int ntaper, int end_only, const double *x,
                         double *out)
{
    int i;
    for (i = 0; i < nout; ++i) out[i] = 0.0;
    if (ntaper > 0) {
        double *win = (double *)malloc(nin * sizeof(double));
        const float D = 3.3f;
        for (i = 0; i < nin; ++i) win[i] = 1.0;
        if (!end_only)
            for (i = 1; i <= ntaper; ++i) {
                float q = D * (float)(ntaper - i + 1) / (float)ntaper;
                win[i - 1] = (double)expf(-(q * q));
            }
        for (i = 0; i <= ntaper - 1; ++i) {
            float q = D * (float)(ntaper - i) / (float)ntaper;
            win[nin - i - 1] = (double)expf(-(q * q));
        }
        for (i = 0; i < nin; ++i) out[i] = x[i] * win[i];
        free(win);
    } else {
        for (i = 0; i < nin; ++i) out[i] = x[i];
    }
}

/* cumsum_trapezoidal (src/simple_routines.f90:314-341) */
void ora_cumsum_trapezoidal(int n, double dt, const double *x, double *out)
{
    int i;
    out[0] = x[0];
    if (n > 1) {
        out[1] = out[0] + x[1];
        for (i = 2; i < n; ++i) out[i] = out[i - 1] + x[i - 1] + x[i];
    }
    for (i = 0; i < n; ++i) out[i] = out[i] * dt * 0.5;
}

/* ------------------------------------------------------------------------------------
 * Filters (src/filtering.f90:78-230, 699-794).  Complex division uses Smith's method,
 * as gfortran does under its default -fcx-fortran-rules.
 * ---------------------------------------------------------------------------------- */
typedef struct { double re, im; } cplx;

static cplx cmul(cplx a, cplx b) { cplx r; r.re = a.re * b.re - a.im * b.im; r.im = a.re * b.im + a.im * b.re; return r; }
static cplx cdiv(cplx a, cplx b)
{
    cplx r;
    if (fabs(b.re) >= fabs(b.im)) {
        double ratio = b.im / b.re, den = b.re + b.im * ratio;
        r.re = (a.re + a.im * ratio) / den;
        r.im = (a.im - a.re * ratio) / den;
    } else {
        double ratio = b.re / b.im, den = b.im + b.re * ratio;
        r.re = (a.re * ratio + a.im) / den;
        r.im = (a.im * ratio - a.re) / den;
    }
    return r;
}
static cplx cexpi(double th) { cplx r; r.re = cos(th); r.im = sin(th); return r; }

static void butter_lp(const double *f, int nf, double pc, int n, cplx *out)
{
    double wc = 2 * ORA_PI / pc;
    cplx s[64];
    int k, i;
    for (k = 1; k <= n; ++k) {
        cplx e = cexpi((double)(2 * k + n - 1) * ORA_PI / (double)(2 * n));
        s[k - 1].re = wc * e.re; s[k - 1].im = wc * e.im;
    }
    for (i = 0; i < nf; ++i) {
        cplx v = {1.0, 0.0};
        for (k = 1; k <= n; ++k) {
            cplx d;
            d.re = (0.0 - s[k - 1].re) / wc;
            d.im = (2 * ORA_PI * f[i] - s[k - 1].im) / wc;
            v = cdiv(v, d);
        }
        out[i] = v;
    }
}

static void butter_hp(const double *f, int nf, double pc, int n, cplx *out)
{
    double wc = 2 * ORA_PI / pc;
    cplx s[64];
    int k, i;
    for (k = 1; k <= n; ++k) {
        cplx e = cexpi((double)(2 * k + n - 1) * ORA_PI / (double)(2 * n));
        s[k - 1].re = wc * e.re; s[k - 1].im = wc * e.im;
    }
    for (i = 0; i < nf; ++i) {
        cplx v = {1.0, 0.0};
        if (i == 0) { out[0].re = 0.0; out[0].im = 0.0; continue; }
        for (k = 1; k <= n; ++k) {
            cplx num = {wc, 0.0}, den = {0.0, 2 * ORA_PI * f[i]}, q, d;
            q = cdiv(num, den);
            d.re = q.re - s[k - 1].re / wc;
            d.im = q.im - s[k - 1].im / wc;
            v = cdiv(v, d);
        }
        out[i] = v;
    }
}

void ora_filter_tf(const mck_filter_t *flt, int nfreq, double df, double *tf)
{
    double *f = (double *)malloc(nfreq * sizeof(double));
    cplx *a = (cplx *)malloc(nfreq * sizeof(cplx)), *b = (cplx *)malloc(nfreq * sizeof(cplx));
    int i;
    const double *fr = flt->frequencies;
    for (i = 0; i < nfreq; ++i) f[i] = df * i;
    switch (flt->filterclass) {
    case MCK_FILTER_IDENT:
        for (i = 0; i < nfreq; ++i) { a[i].re = 1.0; a[i].im = 0.0; }
        break;
    case MCK_FILTER_GABOR: {
        double fc = 1 / fr[0], sig = fr[1], ts = fr[2];
        for (i = 0; i < nfreq; ++i) {
            double g = 0.0;
            cplx ph;
            if (f[i] > 0.0) {
                double l = log(f[i] / fc), ls = log(sig);
                g = exp(-(l * l) / (2 * (ls * ls)));
            }
            ph = cexpi((-2.0 * ORA_PI) * f[i] * ts);
            a[i].re = g * ph.re;
            a[i].im = g * ph.im;
        }
        break;
    }
    case MCK_FILTER_BUTTER_LP: case MCK_FILTER_BUTTER_LP_ZP:
        butter_lp(f, nfreq, fr[0], (int)fr[2], a);
        break;
    case MCK_FILTER_BUTTER_HP: case MCK_FILTER_BUTTER_HP_ZP:
        butter_hp(f, nfreq, fr[0], (int)fr[2], a);
        break;
    default: /* band passes */
        butter_hp(f, nfreq, fr[0], (int)fr[2], a);
        butter_lp(f, nfreq, fr[1], (int)fr[2], b);
        for (i = 0; i < nfreq; ++i) a[i] = cmul(a[i], b[i]);
        break;
    }
    if (flt->filterclass == MCK_FILTER_BUTTER_LP_ZP || flt->filterclass == MCK_FILTER_BUTTER_HP_ZP ||
        flt->filterclass == MCK_FILTER_BUTTER_BP_ZP)
        for (i = 0; i < nfreq; ++i) {
            cplx c = {a[i].re, -a[i].im};
            a[i] = cmul(a[i], c);
        }
    for (i = 0; i < nfreq; ++i) { tf[2 * i] = a[i].re; tf[2 * i + 1] = a[i].im; }
    free(f); free(a); free(b);
}

/* ------------------------------------------------------------------------------------
 * 12 nearest element midpoints (find_element, readfields.f90:1065-1073, with the
 * squared distance of kdtree2.f90:1541-1552 accumulated in single precision).
 * Brute force; sorted by (distance, index).  Returns 1 when two of the nn+1 nearest
 * distances are equal, i.e. when the kd-tree's heap order could differ.
 * ---------------------------------------------------------------------------------- */
int ora_nearest(int nelem, const float *mp_s, const float *mp_z, double s, double z,
                int nn, int32_t *idx, float *dist)
{
    float qs = (float)s, qz = (float)z;
    int cnt = 0, i, j, tie = 0;
    int K = nn + 1;
    int32_t bi[64];
    float bd[64];
    for (i = 0; i < nelem; ++i) {
        float dx = mp_s[i] - qs, dz = mp_z[i] - qz;
        float sd = 0.0f;
        sd = sd + dx * dx;
        sd = sd + dz * dz;
        if (cnt < K || sd < bd[cnt - 1] || (sd == bd[cnt - 1] && i < bi[cnt - 1])) {
            int p = (cnt < K) ? cnt++ : K - 1;
            while (p > 0 && (bd[p - 1] > sd || (bd[p - 1] == sd && bi[p - 1] > i))) {
                bd[p] = bd[p - 1]; bi[p] = bi[p - 1]; --p;
            }
            bd[p] = sd; bi[p] = i;
        }
    }
    for (j = 1; j < cnt; ++j) if (bd[j] == bd[j - 1]) tie = 1;
    for (j = 0; j < nn && j < cnt; ++j) { idx[j] = bi[j]; dist[j] = bd[j]; }
    return tie;
}
