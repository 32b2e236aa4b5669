// synth.cpp - seeded synthetic AxiSEM-style merged wavefield databases and inversion
// meshes for tests and benchmarks.
//
// Mesh: an axisymmetric half-disk (s >= 0, |z| <= R) of npol = 4 quads with the four
// element types of the AxiSEM mapping (finite_elem_mapping.f90): a central grid of
// bilinear (subpar, type 1) quads, straight-edged transition rings (subpar), one ring of
// semi-curved quads (semino, type 2, north; semiso, type 3, south) ending on the circle
// r = rc, and spheroidal shells (type 0) from rc to R.  Axis elements have their
// xi = -1 edge on s = 0 (GLJ collocation in xi, sem_derivatives.f90:803-816).  Elements
// are numbered in Morton order of their midpoints (the repacker sorts them spatially,
// repack_database.py:85-87).  Shared GLL points are merged, so traces are continuous.
//
// Traces: for every global GLL point and variable, a sum of 8 Gaussian-windowed 40 s
// period wavelets with amplitudes N(0,1) and arrival times U(100, 900) s drawn from a
// counter-based hash of (seed, point, variable, wavelet), evaluated in double precision
// (synth_math.h) and rounded to float.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

#include "synth_math.h"

namespace {

const double kPi = 3.1415926535898;  // global_parameters.f90:40

struct Mesh {
    std::vector<float> ps, pz;          // global points
    std::vector<int32_t> fem, sem;      // [e][4], [e][jpol][ipol] (C order)
    std::vector<int32_t> eltype, axis;
    std::vector<float> mps, mpz;
};

// ---- element mappings (same formulas as the reference, used to place GLL points) ----
void compute_ab(double &a, double &b, double s1, double z1, double s2, double z2) {
    a = std::sqrt(std::fabs((s2 * s2 * (z1 * z1) - z2 * z2 * (s1 * s1)) / (z1 * z1 - z2 * z2)));
    b = std::sqrt(std::fabs((z1 * z1 * (s2 * s2) - z2 * z2 * (s1 * s1)) / (s2 * s2 - s1 * s1)));
}
double compute_theta(double s, double z, double a, double b) {
    if (s >= 0.0) return std::atan2(z * a, s * b);
    return z > 0 ? kPi / 2 : -kPi / 2;
}
void mapping(double xi, double eta, const double *n /* s1..s4,z1..z4 */, int type, double *sz) {
    const double *S = n, *Z = n + 4;
    if (type == 1) {
        double sh[4] = {(1 - xi) * (1 - eta) / 4, (1 + xi) * (1 - eta) / 4,
                        (1 + xi) * (1 + eta) / 4, (1 - xi) * (1 + eta) / 4};
        sz[0] = sz[1] = 0;
        for (int i = 0; i < 4; ++i) { sz[0] += sh[i] * S[i]; sz[1] += sh[i] * Z[i]; }
    } else if (type == 0) {
        double th[4], r[4];
        for (int i = 0; i < 4; ++i) {
            r[i] = std::sqrt(S[i] * S[i] + Z[i] * Z[i]);
            th[i] = r[i] != 0 ? std::acos(Z[i] / r[i]) : 0.0;
        }
        double a = ((1 - xi) * th[3] + (1 + xi) * th[2]) / 2, b = ((1 - xi) * th[0] + (1 + xi) * th[1]) / 2;
        sz[0] = (1 + eta) * r[3] / 2 * std::sin(a) + (1 - eta) * r[0] / 2 * std::sin(b);
        sz[1] = (1 + eta) * r[3] / 2 * std::cos(a) + (1 - eta) * r[0] / 2 * std::cos(b);
    } else {
        // semino: straight 1-2, curved 4-3; semiso: curved 1-2, straight 4-3
        int c1 = type == 2 ? 3 : 0, c2 = type == 2 ? 2 : 1;  // curved edge corners
        double a, b;
        compute_ab(a, b, S[c1], Z[c1], S[c2], Z[c2]);
        double t1 = compute_theta(S[c1], Z[c1], a, b), t2 = compute_theta(S[c2], Z[c2], a, b);
        double tb = (t1 + t2) / 2, dt = t2 - t1;
        double sc = a * std::cos(tb + xi * dt / 2), zc = b * std::sin(tb + xi * dt / 2);
        int l1 = type == 2 ? 0 : 3, l2 = type == 2 ? 1 : 2;  // straight edge corners
        double sl = ((1 + xi) * S[l2] + (1 - xi) * S[l1]) / 2, zl = ((1 + xi) * Z[l2] + (1 - xi) * Z[l1]) / 2;
        double sb = type == 2 ? sl : sc, zb = type == 2 ? zl : zc;
        double st = type == 2 ? sc : sl, zt = type == 2 ? zc : zl;
        sz[0] = (sb + st) / 2 + eta * (st - sb) / 2;
        sz[1] = (zb + zt) / 2 + eta * (zt - zb) / 2;
    }
}

void gll4(double *x) { double q = std::sqrt(3.0 / 7.0); x[0] = -1; x[1] = -q; x[2] = 0; x[3] = q; x[4] = 1; }
void glj4(double *x) {
    // zeros of P_3^{(1,2)}: -0.50778762955831502, 0.13230082077732325, 0.70882014211432509
    // refined by Newton on the Jacobi recurrence
    double z[3] = {-0.50778762955831502, 0.13230082077732325, 0.70882014211432509};
    x[0] = -1; x[1] = z[0]; x[2] = z[1]; x[3] = z[2]; x[4] = 1;
}

uint64_t morton2(uint32_t a, uint32_t b) {
    auto spread = [](uint64_t v) {
        v &= 0xffffffffull;
        v = (v | (v << 16)) & 0x0000ffff0000ffffull;
        v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
        v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
        v = (v | (v << 2)) & 0x3333333333333333ull;
        v = (v | (v << 1)) & 0x5555555555555555ull;
        return v;
    };
    return spread(a) | (spread(b) << 1);
}

struct PointSet {
    std::unordered_map<int64_t, std::vector<int>> cells;
    std::vector<double> s, z;
    double h = 1.0;
    int64_t key(int64_t i, int64_t j) const { return (i << 32) ^ (j & 0xffffffffll); }
    int insert(double ps, double pz) {
        int64_t ci = (int64_t)std::floor(ps / h), cj = (int64_t)std::floor(pz / h);
        for (int64_t di = -1; di <= 1; ++di)
            for (int64_t dj = -1; dj <= 1; ++dj) {
                auto it = cells.find(key(ci + di, cj + dj));
                if (it == cells.end()) continue;
                for (int id : it->second)
                    if (std::fabs(s[id] - ps) < 1e-3 && std::fabs(z[id] - pz) < 1e-3) return id;
            }
        int id = (int)s.size();
        s.push_back(ps); z.push_back(pz);
        cells[key(ci, cj)].push_back(id);
        return id;
    }
};

}  // namespace

extern "C" {

// Builds the mesh.  n: quarter divisions of the central grid (4n elements along every
// ring), mring: ring layers (the last one semi-curved), nshell: spheroidal layers.
void *synth_mesh_create(double R, double rc, int n, int mring, int nshell) {
    Mesh *M = new Mesh;
    const int nth = 4 * n;
    // central grid, rounded towards a disk: p' = p * (linf / l2)^0.6
    const double a = 0.62 * rc;
    auto central = [&](int i, int j, double &s, double &z) {
        double s0 = a * i / n, z0 = a * ((double)j / n - 1.0);
        double linf = std::max(std::fabs(s0), std::fabs(z0)), l2 = std::sqrt(s0 * s0 + z0 * z0);
        double f = l2 > 0 ? std::pow(linf / l2, 0.6) : 1.0;
        s = s0 * f; z = z0 * f;
    };
    // boundary walk of the central grid, north pole -> south pole (4n segments)
    std::vector<double> bs(nth + 1), bz(nth + 1), th(nth + 1);
    for (int k = 0; k <= nth; ++k) {
        int i, j;
        if (k <= n) { i = k; j = 2 * n; }
        else if (k <= 3 * n) { i = n; j = 2 * n - (k - n); }
        else { i = 4 * n - k; j = 0; }
        central(i, j, bs[k], bz[k]);
        th[k] = std::atan2(bs[k], bz[k]);  // colatitude
    }
    // radii of ring polygons (between boundary and rc) and shells (rc -> R)
    std::vector<double> shell_r(nshell + 1);
    for (int l = 0; l <= nshell; ++l) shell_r[l] = rc * std::pow(R / rc, (double)l / nshell);

    struct Quad { double n[8]; int type, axis; };
    std::vector<Quad> quads;
    auto fr = [](double v) { return (double)(float)v; };  // corners as stored (fp32)
    auto add = [&](double s1, double z1, double s2, double z2, double s3, double z3, double s4,
                   double z4, int type, int axis) {
        Quad q;
        q.n[0] = fr(s1); q.n[1] = fr(s2); q.n[2] = fr(s3); q.n[3] = fr(s4);
        q.n[4] = fr(z1); q.n[5] = fr(z2); q.n[6] = fr(z3); q.n[7] = fr(z4);
        if (std::fabs(q.n[0]) < 1e-6) q.n[0] = 0; if (std::fabs(q.n[3]) < 1e-6) q.n[3] = 0;
        q.type = type; q.axis = axis;
        quads.push_back(q);
    };
    // central grid
    for (int j = 0; j < 2 * n; ++j)
        for (int i = 0; i < n; ++i) {
            double s1, z1, s2, z2, s3, z3, s4, z4;
            central(i, j, s1, z1); central(i + 1, j, s2, z2);
            central(i + 1, j + 1, s3, z3); central(i, j + 1, s4, z4);
            add(s1, z1, s2, z2, s3, z3, s4, z4, 1, i == 0);
        }
    // ring polygon k on ray th[k] at fraction t between boundary and rc
    auto ringpt = [&](int k, double t, double &s, double &z) {
        double rb = std::sqrt(bs[k] * bs[k] + bz[k] * bz[k]);
        double r = rb * (1 - t) + rc * t;
        if (t <= 0) { s = bs[k]; z = bz[k]; return; }
        s = r * std::sin(th[k]); z = r * std::cos(th[k]);
        if (k == 0 || k == nth) s = 0;
    };
    for (int m = 0; m < mring; ++m) {
        double t0 = (double)m / mring, t1 = (double)(m + 1) / mring;
        for (int k = 0; k < nth; ++k) {
            bool north = k < 2 * n;
            int kl = north ? k : k + 1, kr = north ? k + 1 : k;  // xi = -1 side first
            double s1, z1, s2, z2, s3, z3, s4, z4;
            int ax = (k == 0 || k == nth - 1);
            if (m < mring - 1) {
                if (north) {
                    ringpt(kl, t0, s1, z1); ringpt(kr, t0, s2, z2);
                    ringpt(kr, t1, s3, z3); ringpt(kl, t1, s4, z4);
                } else {
                    ringpt(kl, t1, s1, z1); ringpt(kr, t1, s2, z2);
                    ringpt(kr, t0, s3, z3); ringpt(kl, t0, s4, z4);
                }
                add(s1, z1, s2, z2, s3, z3, s4, z4, 1, ax);
            } else if (north) {  // semino: straight bottom (1-2), arc top (4-3)
                ringpt(kl, t0, s1, z1); ringpt(kr, t0, s2, z2);
                s3 = rc * std::sin(th[kr]); z3 = rc * std::cos(th[kr]);
                s4 = rc * std::sin(th[kl]); z4 = rc * std::cos(th[kl]);
                if (kl == 0) s4 = 0;
                add(s1, z1, s2, z2, s3, z3, s4, z4, 2, ax);
            } else {             // semiso: arc bottom (1-2), straight top (4-3)
                s1 = rc * std::sin(th[kl]); z1 = rc * std::cos(th[kl]);
                s2 = rc * std::sin(th[kr]); z2 = rc * std::cos(th[kr]);
                if (kl == nth) s1 = 0;
                ringpt(kr, t0, s3, z3); ringpt(kl, t0, s4, z4);
                add(s1, z1, s2, z2, s3, z3, s4, z4, 3, ax);
            }
        }
    }
    // spheroidal shells
    for (int l = 0; l < nshell; ++l)
        for (int k = 0; k < nth; ++k) {
            bool north = k < 2 * n;
            int kl = north ? k : k + 1, kr = north ? k + 1 : k;
            double r0 = shell_r[l], r1 = shell_r[l + 1];
            double s1 = r0 * std::sin(th[kl]), z1 = r0 * std::cos(th[kl]);
            double s2 = r0 * std::sin(th[kr]), z2 = r0 * std::cos(th[kr]);
            double s3 = r1 * std::sin(th[kr]), z3 = r1 * std::cos(th[kr]);
            double s4 = r1 * std::sin(th[kl]), z4 = r1 * std::cos(th[kl]);
            if (kl == 0 || kl == nth) { s1 = 0; s4 = 0; }
            add(s1, z1, s2, z2, s3, z3, s4, z4, 0, (k == 0 || k == nth - 1));
        }
    // spatial order of elements: Morton code of the midpoint
    const int ne = (int)quads.size();
    std::vector<double> mids(2 * ne);
    std::vector<std::pair<uint64_t, int>> order(ne);
    for (int e = 0; e < ne; ++e) {
        double sz[2];
        mapping(0.0, 0.0, quads[e].n, quads[e].type, sz);
        mids[2 * e] = sz[0]; mids[2 * e + 1] = sz[1];
        uint32_t qi = (uint32_t)std::min(65535.0, std::max(0.0, sz[0] / R * 65535.0));
        uint32_t qj = (uint32_t)std::min(65535.0, std::max(0.0, (sz[1] + R) / (2 * R) * 65535.0));
        order[e] = {morton2(qi, qj), e};
    }
    std::sort(order.begin(), order.end());
    PointSet P;
    P.h = 1.0;
    M->fem.resize(4 * ne); M->sem.resize(25 * ne);
    M->eltype.resize(ne); M->axis.resize(ne); M->mps.resize(ne); M->mpz.resize(ne);
    double gll[5], glj[5];
    gll4(gll); glj4(glj);
    for (int o = 0; o < ne; ++o) {
        const Quad &q = quads[order[o].second];
        // corners first, exactly as stored
        int cid[4];
        for (int c = 0; c < 4; ++c) cid[c] = P.insert(q.n[c], q.n[4 + c]);
        const double *xs = q.axis ? glj : gll;
        for (int j = 0; j < 5; ++j)
            for (int i = 0; i < 5; ++i) {
                int id;
                if (i == 0 && j == 0) id = cid[0];
                else if (i == 4 && j == 0) id = cid[1];
                else if (i == 4 && j == 4) id = cid[2];
                else if (i == 0 && j == 4) id = cid[3];
                else {
                    double sz[2];
                    mapping(xs[i], gll[j], q.n, q.type, sz);
                    if (q.axis && i == 0) sz[0] = 0.0;
                    id = P.insert(sz[0], sz[1]);
                }
                M->sem[25 * o + 5 * j + i] = id;
            }
        for (int c = 0; c < 4; ++c) M->fem[4 * o + c] = cid[c];
        M->eltype[o] = q.type;
        M->axis[o] = q.axis;
        M->mps[o] = (float)mids[2 * order[o].second];
        M->mpz[o] = (float)mids[2 * order[o].second + 1];
    }
    M->ps.resize(P.s.size()); M->pz.resize(P.s.size());
    for (size_t i = 0; i < P.s.size(); ++i) { M->ps[i] = (float)P.s[i]; M->pz[i] = (float)P.z[i]; }
    return M;
}

void synth_mesh_sizes(void *h, int32_t *nelem, int32_t *npoints) {
    Mesh *M = (Mesh *)h;
    *nelem = (int32_t)M->eltype.size();
    *npoints = (int32_t)M->ps.size();
}

void synth_mesh_get(void *h, float *mesh_s, float *mesh_z, float *mp_s, float *mp_z,
                    int32_t *fem, int32_t *sem, int32_t *eltype, int32_t *axis) {
    Mesh *M = (Mesh *)h;
    std::memcpy(mesh_s, M->ps.data(), M->ps.size() * 4);
    std::memcpy(mesh_z, M->pz.data(), M->pz.size() * 4);
    std::memcpy(mp_s, M->mps.data(), M->mps.size() * 4);
    std::memcpy(mp_z, M->mpz.data(), M->mpz.size() * 4);
    std::memcpy(fem, M->fem.data(), M->fem.size() * 4);
    std::memcpy(sem, M->sem.data(), M->sem.size() * 4);
    std::memcpy(eltype, M->eltype.data(), M->eltype.size() * 4);
    std::memcpy(axis, M->axis.data(), M->axis.size() * 4);
}

void synth_mesh_free(void *h) { delete (Mesh *)h; }

// one trace: sum of 8 wavelets, window |t - tau| <= 240 s (synth_math.h, shared with
// the device generator so that both produce the same floats)
void synth_trace(uint64_t seed, int64_t point, int var, int nt, double dt, float *out) {
    std::vector<double> acc(nt, 0.0);
    for (int w = 0; w < 8; ++w) {
        const synth_math::Wavelet W = synth_math::wavelet(seed, point, var, w, nt, dt);
        for (int i = W.i0; i <= W.i1; ++i) acc[i] += synth_math::wavelet_sample(W, i, dt);
    }
    for (int i = 0; i < nt; ++i) out[i] = (float)acc[i];
}

// merged_snapshots [nelem][nvar][5][5][nt] for a mesh's point ids sem[e][j][i]
void synth_db_fill(int32_t nelem, const int32_t *sem, int32_t nvar, int32_t nt, double dt,
                   uint64_t seed, float *out, int32_t nthreads) {
    if (nthreads < 1) nthreads = 1;
    auto work = [&](int t0) {
        for (int e = t0; e < nelem; e += nthreads)
            for (int v = 0; v < nvar; ++v)
                for (int j = 0; j < 5; ++j)
                    for (int i = 0; i < 5; ++i)
                        synth_trace(seed, sem[25 * e + 5 * j + i], v, nt, dt,
                                    out + ((((size_t)e * nvar + v) * 5 + j) * 5 + i) * nt);
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(work, t);
    for (auto &x : th) x.join();
}

// on-demand filler of single elements (merged layout), for bounded CPU samples
struct LazyDb {
    const int32_t *sem;
    int32_t nvar[2], nt;
    double dt;
    uint64_t seed[2];
};

void *synth_lazy_create(const int32_t *sem, int32_t nvar_fwd, int32_t nvar_bwd, int32_t nt,
                        double dt, uint64_t seed_fwd, uint64_t seed_bwd) {
    LazyDb *L = new LazyDb;
    L->sem = sem; L->nvar[0] = nvar_fwd; L->nvar[1] = nvar_bwd; L->nt = nt; L->dt = dt;
    L->seed[0] = seed_fwd; L->seed[1] = seed_bwd;
    return L;
}

void synth_lazy_free(void *h) { delete (LazyDb *)h; }

void synth_fill_element(int which, int e, float *dst, void *user) {
    LazyDb *L = (LazyDb *)user;
    const int nv = L->nvar[which], nt = L->nt;
    for (int v = 0; v < nv; ++v)
        for (int j = 0; j < 5; ++j)
            for (int i = 0; i < 5; ++i)
                synth_trace(L->seed[which], L->sem[25 * e + 5 * j + i], v, nt, L->dt,
                            dst + (((size_t)v * 5 + j) * 5 + i) * nt);
}

// radial model table (create_interpolator, interpolate_mesh.f90:45-79): value of the
// GLL point nearest (fp32 distance) to (0, (i-1) dz); the model at a point is given by
// the smooth profile below.
static void profile(double r, double R, float *v) {
    double x = 1.0 - r / R;
    v[0] = (float)(8000.0 + 5500.0 * x);            // vp
    v[1] = (float)(4500.0 + 2800.0 * x);            // vs
    v[2] = (float)(3300.0 + 9000.0 * x * x);        // rho
    v[3] = 1.0f; v[4] = 1.0f; v[5] = 1.0f;          // phi, xi, eta
}

int32_t synth_model_table(int32_t npoints, const float *ps, const float *pz, double R, float dz,
                          float *out /* [6][ndepth] or NULL */) {
    int32_t nd = (int32_t)(R / dz) + 1;
    if (!out) return nd;
    std::vector<int> idx(npoints);
    for (int i = 0; i < npoints; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return pz[a] < pz[b]; });
    for (int d = 0; d < nd; ++d) {
        float qz = (float)d * dz;
        // binary search the z-sorted list, then scan outwards
        int lo = 0, hi = npoints;
        while (lo < hi) { int mid = (lo + hi) / 2; if (pz[idx[mid]] < qz) lo = mid + 1; else hi = mid; }
        float best = 3.4e38f; int bi = -1;
        for (int dir = -1; dir <= 1; dir += 2)
            for (int k = (dir < 0 ? lo - 1 : lo); k >= 0 && k < npoints; k += dir) {
                float dzz = pz[idx[k]] - qz;
                if (dzz * dzz > best) break;
                float sd = 0.0f;
                sd = sd + (ps[idx[k]] - 0.0f) * (ps[idx[k]] - 0.0f);
                sd = sd + dzz * dzz;
                if (sd < best || (sd == best && idx[k] < bi)) { best = sd; bi = idx[k]; }
            }
        float v[6];
        profile(std::sqrt((double)ps[bi] * ps[bi] + (double)pz[bi] * pz[bi]), R, v);
        for (int c = 0; c < 6; ++c) out[(size_t)c * nd + d] = v[c];
    }
    return nd;
}

}  // extern "C"
