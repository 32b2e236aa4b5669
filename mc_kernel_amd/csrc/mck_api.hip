// mck_api.hip - host side of the C-ABI (include/mck.h): once-per-process setup (HBM
// residency of the merged databases, point-location grids, derivative matrices,
// filters and per-kernel window spectra) and the step-synchronous batched evaluation
// of slave_work (src/slave.f90:273-810) over all voxels of a task.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <array>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "mck.h"
#include "mck_device.h"

#include "mck_kernels.inc.h"
#include "mck_mfma.inc.h"
#include "mck_ad.inc.h"
#include "mck_results.inc.h"

using namespace mck;
using cd = std::complex<double>;

namespace {

thread_local std::string g_err = "no error";
const double kPi = 3.1415926535898;  // global_parameters.f90:40
const double kDeg2Rad = kPi / 180.0;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHECK(x)                                                                   \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess)                                                         \
            return fail(MCK_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));    \
    } while (0)

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t want) {
        if (want <= n) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc((void **)&p, std::max<size_t>(want, 1) * sizeof(T));
        if (e == hipSuccess) n = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// ---- host FFT (double), used for the per-kernel seismogram set-up only --------------
void fft_inplace(std::vector<cd> &a, int sign) {
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        for (size_t k = 0; k < len / 2; ++k) {
            double ang = sign * 2.0 * M_PI * (double)k / (double)len;
            cd w(std::cos(ang), std::sin(ang));
            for (size_t m = k; m < n; m += len) {
                cd t = a[m + len / 2] * w;
                a[m + len / 2] = a[m] - t;
                a[m] += t;
            }
        }
    }
}
// X = dt * DFT(x) (r2c), x = df * IDFT(X) (c2r, imaginary parts of DC and Nyquist ignored)
std::vector<cd> rfft(const std::vector<double> &x, int N, double dt) {
    std::vector<cd> a(N, 0.0);
    for (int i = 0; i < (int)x.size() && i < N; ++i) a[i] = x[i];
    fft_inplace(a, -1);
    std::vector<cd> X(N / 2 + 1);
    for (int k = 0; k <= N / 2; ++k) X[k] = a[k] * dt;
    return X;
}
std::vector<double> irfft(const std::vector<cd> &X, int N, double df) {
    std::vector<cd> a(N);
    a[0] = X[0].real();
    a[N / 2] = X[N / 2].real();
    for (int k = 1; k < N / 2; ++k) { a[k] = X[k]; a[N - k] = std::conj(X[k]); }
    fft_inplace(a, +1);
    std::vector<double> x(N);
    for (int i = 0; i < N; ++i) x[i] = a[i].real() * df;
    return x;
}

int nextpow2(int n) { int p = 2; while (p < n) p *= 2; return p; }

// ---- filters (filtering.f90:78-230, 699-794) -----------------------------------------
std::vector<cd> butter_lp(const std::vector<double> &f, double pc, int n) {
    double wc = 2 * kPi / pc;
    std::vector<cd> s(n), out(f.size());
    for (int k = 1; k <= n; ++k) {
        double th = (double)(2 * k + n - 1) * kPi / (double)(2 * n);
        s[k - 1] = cd(wc * std::cos(th), wc * std::sin(th));
    }
    for (size_t i = 0; i < f.size(); ++i) {
        cd v = 1.0;
        for (int k = 0; k < n; ++k) v /= (cd(0.0, 2 * kPi * f[i]) - s[k]) / wc;
        out[i] = v;
    }
    return out;
}
std::vector<cd> butter_hp(const std::vector<double> &f, double pc, int n) {
    double wc = 2 * kPi / pc;
    std::vector<cd> s(n), out(f.size());
    for (int k = 1; k <= n; ++k) {
        double th = (double)(2 * k + n - 1) * kPi / (double)(2 * n);
        s[k - 1] = cd(wc * std::cos(th), wc * std::sin(th));
    }
    for (size_t i = 0; i < f.size(); ++i) {
        if (i == 0) { out[0] = 0.0; continue; }
        cd v = 1.0;
        for (int k = 0; k < n; ++k) v /= wc / cd(0.0, 2 * kPi * f[i]) - s[k] / wc;
        out[i] = v;
    }
    return out;
}
std::vector<cd> filter_tf(const mck_filter_t &flt, const std::vector<double> &f) {
    const double *fr = flt.frequencies;
    std::vector<cd> tf(f.size());
    switch (flt.filterclass) {
    case MCK_FILTER_IDENT: std::fill(tf.begin(), tf.end(), cd(1.0, 0.0)); break;
    case MCK_FILTER_GABOR: {
        double fc = 1 / fr[0], ls = std::log(fr[1]);
        for (size_t i = 0; i < f.size(); ++i) {
            double g = 0.0;
            if (f[i] > 0) { double l = std::log(f[i] / fc); g = std::exp(-(l * l) / (2 * (ls * ls))); }
            double th = (-2.0 * kPi) * f[i] * fr[2];
            tf[i] = cd(g * std::cos(th), g * std::sin(th));
        }
        break;
    }
    case MCK_FILTER_BUTTER_LP: case MCK_FILTER_BUTTER_LP_ZP: tf = butter_lp(f, fr[0], (int)fr[2]); break;
    case MCK_FILTER_BUTTER_HP: case MCK_FILTER_BUTTER_HP_ZP: tf = butter_hp(f, fr[0], (int)fr[2]); break;
    default: {
        auto a = butter_hp(f, fr[0], (int)fr[2]), b = butter_lp(f, fr[1], (int)fr[2]);
        for (size_t i = 0; i < f.size(); ++i) tf[i] = a[i] * b[i];
    }
    }
    if (flt.filterclass >= MCK_FILTER_BUTTER_LP_ZP)
        for (auto &v : tf) v = v * std::conj(v);
    return tf;
}

std::vector<double> taper_zeropad(const double *x, int nin, int nout, int ntaper, bool end_only) {
    std::vector<double> out(nout, 0.0), win(nin, 1.0);
    const float D = 3.3f;
    if (ntaper > 0) {
        if (!end_only)
            for (int i = 1; i <= ntaper; ++i) {
                float q = D * (float)(ntaper - i + 1) / (float)ntaper;
                win[i - 1] = (double)expf(-(q * q));
            }
        for (int i = 0; i < ntaper; ++i) {
            float q = D * (float)(ntaper - i) / (float)ntaper;
            win[nin - 1 - i] = (double)expf(-(q * q));
        }
    }
    for (int i = 0; i < nin; ++i) out[i] = x[i] * win[i];
    return out;
}

std::vector<double> cumsum_trap(const double *x, int n, double dt) {
    std::vector<double> c(n);
    c[0] = x[0];
    if (n > 1) c[1] = c[0] + x[1];
    for (int i = 2; i < n; ++i) c[i] = c[i - 1] + x[i - 1] + x[i];
    for (auto &v : c) v = v * dt * 0.5;
    return c;
}

// ---- spectral basis: GLL / GLJ(0,1) nodes and Lagrange derivative matrices ---------
void jacobi(int n, double a, double b, double x, double &p, double &dp) {
    auto val = [](int n_, double a_, double b_, double x_) {
        double p0 = 1.0, p1 = 0.5 * ((a_ + b_ + 2.0) * x_ + (a_ - b_)), pn = n_ == 0 ? 1.0 : p1;
        for (int k = 2; k <= n_; ++k) {
            double c = 2.0 * k + a_ + b_;
            double a1 = 2.0 * k * (k + a_ + b_) * (c - 2.0), a2 = (c - 1.0) * (a_ * a_ - b_ * b_);
            double a3 = (c - 2.0) * (c - 1.0) * c, a4 = 2.0 * (k + a_ - 1.0) * (k + b_ - 1.0) * c;
            pn = ((a2 + a3 * x_) * p1 - a4 * p0) / a1;
            p0 = p1; p1 = pn;
        }
        return pn;
    };
    p = val(n, a, b, x);
    dp = n == 0 ? 0.0 : 0.5 * (n + a + b + 1.0) * val(n - 1, a + 1, b + 1, x);
}
void lobatto(int n, double a, double b, double *x) {
    x[0] = -1.0; x[n] = 1.0;
    int m = n - 1;
    for (int k = 0; k < m; ++k) {
        double r = -std::cos((2.0 * k + 1.0) * M_PI / (2.0 * m));
        if (k > 0) r = 0.5 * (r + x[k]);
        for (int it = 0; it < 100; ++it) {
            double p, dp, s = 0.0;
            for (int i = 0; i < k; ++i) s += 1.0 / (r - x[i + 1]);
            jacobi(m, a + 1, b + 1, r, p, dp);
            double d = -p / (dp - s * p);
            r += d;
            if (std::fabs(d) < 1e-17) break;
        }
        x[k + 1] = r;
    }
}
void deriv_matrix(const double *x, double *G) {
    double lam[5];
    for (int i = 0; i < 5; ++i) {
        double p = 1.0;
        for (int m = 0; m < 5; ++m) if (m != i) p *= (x[i] - x[m]);
        lam[i] = 1.0 / p;
    }
    for (int j = 0; j < 5; ++j)
        for (int i = 0; i < 5; ++i) {
            double v = 0.0;
            if (i == j) { for (int m = 0; m < 5; ++m) if (m != i) v += 1.0 / (x[i] - x[m]); }
            else v = lam[i] / (lam[j] * (x[j] - x[i]));
            G[i + 5 * j] = v;
        }
}

// ---- element constants (same arithmetic as finite_elem_mapping.f90's helpers) ------
void element_constants(const float *ms, const float *mz, const int *fem, int type, double *E) {
    for (int c = 0; c < 4; ++c) { E[c] = (double)ms[fem[c]]; E[4 + c] = (double)mz[fem[c]]; }
    for (int c = 8; c < 16; ++c) E[c] = 0.0;
    auto ab = [](double &a, double &b, double s1, double z1, double s2, double z2) {
        a = std::sqrt(std::fabs((s2 * s2 * (z1 * z1) - z2 * z2 * (s1 * s1)) / (z1 * z1 - z2 * z2)));
        b = std::sqrt(std::fabs((z1 * z1 * (s2 * s2) - z2 * z2 * (s1 * s1)) / (s2 * s2 - s1 * s1)));
    };
    auto theta = [](double s, double z, double a, double b) {
        if (s >= 0.0) return std::atan2(z * a, s * b);
        if (z > 0) return kPi / 2;
        if (z < 0) return -kPi / 2;
        return 0.0;
    };
    if (type == 0) {
        for (int i = 0; i < 4; ++i) {
            double r = std::sqrt(E[i] * E[i] + E[4 + i] * E[4 + i]);
            E[8 + i] = r;
            E[12 + i] = r != 0.0 ? std::acos(E[4 + i] / r) : 0.0;
        }
    } else if (type == 2) {
        double a, b;
        ab(a, b, E[3], E[7], E[2], E[6]);
        double t3 = theta(E[2], E[6], a, b), t4 = theta(E[3], E[7], a, b);
        E[8] = a; E[9] = b; E[10] = (t4 + t3) / 2; E[11] = t3 - t4;
    } else if (type == 3) {
        double a, b;
        ab(a, b, E[0], E[4], E[1], E[5]);
        double t2 = theta(E[1], E[5], a, b), t1 = theta(E[0], E[4], a, b);
        E[8] = a; E[9] = b; E[10] = (t1 + t2) / 2; E[11] = t2 - t1;
    }
}

double norm2_3(double a, double b, double c) {
    double v[3] = {a, b, c}, scale = 0.0, res = 1.0;
    for (double x : v)
        if (x != 0.0) {
            double ax = std::fabs(x);
            if (scale < ax) { double q = scale / ax; res = 1.0 + res * q * q; scale = ax; }
            else { double q = ax / scale; res = res + q * q; }
        }
    return std::sqrt(res) * scale;
}

double tet_volume(const double *v) {
    __float128 a[4][4];
    for (int i = 0; i < 4; ++i) {
        for (int k = 0; k < 3; ++k) a[i][k] = v[i * 3 + k];
        a[i][3] = 1;
    }
#define A(r, c) a[(r) - 1][(c) - 1]
    __float128 d =
          A(1,1) * ( A(2,2) * ( A(3,3) * A(4,4) - A(3,4) * A(4,3) )
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
    return std::fabs((double)d) / 6.0;
}

struct DbDev {
    DevDb d{};
    float *u_owned = nullptr;      // device copy of the snapshots (null if caller-resident)
    // paged snapshots (databases larger than device memory): the caller's host copy is
    // paged through nslot device buffers of page_elems elements.  The first nslot - 2
    // distinct pages used keep a buffer of their own once loaded (a working set that fits
    // stays resident wherever its pages lie); the others rotate through the last two.
    bool paged = false;
    int page_elems = 0, npages = 0, nslot = 0;
    size_t elem_floats = 0;        // floats per element: nvar x 25 x nt
    const float *host_u = nullptr;
    std::vector<float *> slot;
    std::vector<int> slot_page;    // page held by each buffer (-1: none)
    std::vector<int> page_slot;    // own buffer of each page (-1: none yet, or rotating)
    int nfixed_used = 0;           // own buffers handed out
    // page table of the own buffers (device copy and its host image): launches whose
    // pages all have one run once over all their items
    DBuf<const float *> ptable;
    std::vector<const float *> ptable_host;
    std::vector<hipEvent_t> copied, freed;
    int rot = 0;                   // next rotating buffer
    void release_pages() {
        for (size_t i = 0; i < slot.size(); ++i) {
            if (slot[i]) (void)hipFree(slot[i]);
            if (copied[i]) (void)hipEventDestroy(copied[i]);
            if (freed[i]) (void)hipEventDestroy(freed[i]);
        }
        slot.clear(); slot_page.clear(); page_slot.clear(); copied.clear(); freed.clear();
        ptable.release(); ptable_host.clear();
        paged = false; page_elems = npages = nslot = rot = nfixed_used = 0; host_u = nullptr;
    }
    DBuf<float> mp_s, mp_z;
    DBuf<double> econ, ejac;
    DBuf<int> eltype, axis, cell_start, cell_elems;
    std::vector<double> econ_host;
    std::vector<int> eltype_host, axis_host;
    int nvar = 0;
};

struct State {
    bool init = false;
    mck_setup_t S{};
    PathConf pc{};
    DbDev fwd, bwd;
    int R = 0, nk = 0, nframes = 0;
    DBuf<DevFrame> frames;
    std::vector<DevFrame> frames_host;
    DBuf<float2> twq, H;
    DBuf<double2> twq_d;                     // the quarter twiddle table in double precision
    DBuf<RecDesc> recs, recs_base;           // slot layouts: physical kernels / base kernels
    DBuf<KernDesc> kerns, kerns_base;
    PathConf pc_base{};                      // pc with the base-kernel layout
    DBuf<float> model;
    std::vector<int> k_rec, k_nbk, rec_first, rec_count;
    DBuf<int> d_k_rec, d_k_nbk, d_rec_first, d_rec_count;
    DBuf<double> d_rec_r;                    // [R][3] receiver positions (m)
    std::vector<double> t_last_p, t_last_s, rec_r;  // rec_r [R][3]
    double src_r[3];
    // per-task and per-chunk buffers
    DBuf<double> tet, volume, allowed_var;
    int nbasis = 1;                          // basis functions per element of the task
    int vox = 0;                             // elements of the task are voxels
    DBuf<double> abinv, bweights;            // [nvox][4][4], [step points][4] (nbasis 4)
    DBuf<int64_t> hbase;
    DBuf<uint8_t> masked, live, working, active;
    DBuf<int> work_list, nwork;
    DBuf<double> fsum, f2sum, integral, variance;
    DBuf<uint8_t> conv, done;
    DBuf<int> niter, draws, steps;
    DBuf<int64_t> nmodels;
    DBuf<double> xyz;
    DBuf<int> model_idx;
    DBuf<Located> loc;
    DBuf<float2> Fspec;                 // [point][nd_f][fstride] forward spectra
    DBuf<float> Ftr;                    // [sorted point][nd_f][nt] forward strain traces
    // operator set-up (k_op_setup): gather weights and rotations of the points (by point
    // id) and of the pairs (by position in the pair order)
    DBuf<float> opWf, opQf, opWb, opQb;
    DBuf<int> page_bounds;              // [npages + 1] first item of every database page
    DBuf<int> nlive;                    // [2]: 0, points of a launch with an active receiver
    std::vector<double> seis_host;      // [nrec][nt] seismograms by reciprocity (seis NULL)
    hipStream_t copy_stream = nullptr;  // page copies of paged databases
    DBuf<int> keys_in, keys_out, vals_in, vals_out, pair_count;
    // backward evaluation through element spectra (k_elem_spectra + k_bwd_mfma): element
    // runs of a launch and the nodal spectra of a batch of runs
    DBuf<int> worktab;
    DBuf<float2> Uspec;                           // element spectra
    DBuf<float2> Hp;                              // window spectra (re, -im), rows fstride
    bool mfma_ok = false, mfma_ok_base = false;   // kernel sets the matrix-core path serves
    // events around the backward kernels of the current step, read by collect_events:
    // matrix-core launches (start, after the spectra, after the matrix-core kernel) and
    // per-pair launches (start, end, null)
    std::vector<std::array<hipEvent_t, 3>> kev;
    // per-launch choice of the backward evaluation (bwd_path_env -1): microseconds per
    // group of 16 pairs and per element run of the matrix-core path, per pair of the
    // per-pair path (set at init from c1's measured costs, measured by mck_tune)
    double cost_group_us = 0.0, cost_run_us = 0.0, cost_pair_us = 0.0;
    int bwd_path_env = -1;                        // MCK_BWD_PATH: 0 per-pair, 1 matrix cores
    int mfma_xcd = 1;                             // MCK_MFMA_XCD: groups in XCD-contiguous ranges
    int mfma_waves = 1;                           // MCK_MFMA_WAVES: waves per element run (1, 2, 4)
    size_t uspec_budget = 0;                      // bytes of element spectra per batch
    DBuf<unsigned char> sort_tmp;
    DBuf<double> phys;
    DBuf<unsigned long long> evals;
    DBuf<int> err;
    DBuf<unsigned long long> elem_count;
    int chunk_points = 0;
    hipStream_t stream = nullptr;
    mck_stats_t stats{};
    int64_t stats_vk = 0;                    // evaluated (voxel, kernel) states of a task
    // differentiable filter-convolve-window-integrate stage: window spectra without their
    // filter (W_h), the filter of each window spectrum, the filter responses G of the
    // set-up, and scratch of the forward and adjoint launches
    int nH = 0;
    std::vector<float> ad_tf_host;           // [nfilter][nomega][2]
    DBuf<float2> ad_W, ad_Hg, ad_Fs, ad_U;
    DBuf<float2> ad_Bs, ad_dFs, ad_Ups;      // backward stage buffers of a chunk of points
    DBuf<double2> ad_part;
    DBuf<int> ad_hfilt;
    // results of the last mck_slave_work (device: integral, variance, niter) and
    // mck_integrate_model (device: last_model), for mck_results_add without a host copy
    int last_nelems = 0, last_nbasis = 0, last_model_nelems = 0, last_model_nbasis = 0;
    std::vector<double> last_ctime;
    DBuf<double> last_model, last_hetero;
    DBuf<double> int_geom, int_abinv;       // model / heterogeneity integrals of a task
    DBuf<int64_t> int_base;
    DBuf<uint8_t> int_live;
    DBuf<int> int_nsteps;
    int last_hetero_nelems = 0, last_hetero_nbasis = 0;
    // heterogeneity model (mck_set_hetero): point cloud in a uniform grid
    HetGrid hgrid{};
    bool het_ready = false;
    DBuf<float> het_xyz, het_dln;
    DBuf<int> het_idx, het_cells;
    // the master's accumulators (mck_results_*)
    struct Results {
        bool init = false;
        int int_type = 0, npoints = 0, nelements = 0, nbasis = 1, nmodel = 0, nhet = 0;
        std::vector<int> conn;               // [nelements][nbasis] 1-based vertices
        DBuf<double> K, V, M, H, ctime;      // [npoints][nk] x 2, [npoints][nmodel], [npoints][nhet]
        DBuf<int> niter, proc;               // [nelements][nk], [nelements]
        DBuf<int> dconn, elems, pt_in, pt_out, sl_in, sl_out, seg_point, seg_count, seg_start, nseg, elem_idx;
        DBuf<unsigned char> tmp;             // hipcub scratch
        DBuf<double> stage;                  // host inputs staged in device memory
        DBuf<int> stage_i;
        void release() {
            K.release(); V.release(); M.release(); H.release(); ctime.release(); niter.release(); proc.release();
            for (DBuf<int> *b : {&dconn, &elems, &pt_in, &pt_out, &sl_in, &sl_out, &seg_point, &seg_count,
                                 &seg_start, &nseg, &elem_idx})
                b->release();
            tmp.release(); stage.release(); stage_i.release();
            conn.clear();
            init = false;
        }
    } res;
};
State g;

// page_elems > 0: keep the snapshots in host memory and page them through nslot (>= 2)
// device buffers (see run_paged)
int upload_db(DbDev &D, const mck_db_t &db, int nt, const char *name, int page_elems, int nslot,
              hipStream_t st) {
    if (db.npol != 4) return fail(MCK_EINVAL, std::string(name) + ": only npol = 4 is supported");
    if (db.nelem <= 0 || !db.snapshots || !db.mp_s || !db.fem_mesh)
        return fail(MCK_EINVAL, std::string(name) + ": incomplete mesh");
    const int ne = db.nelem;
    D.nvar = db.nvar;
    // element constants and grid
    D.econ_host.assign((size_t)ne * 16, 0.0);
    D.eltype_host.assign(db.eltype, db.eltype + ne);
    D.axis_host.assign(db.axis, db.axis + ne);
    for (int e = 0; e < ne; ++e) {
        int t = db.eltype[e];
        if (t < 0 || t > 3) return fail(MCK_EINVAL, std::string(name) + ": unknown element type");
        element_constants(db.mesh_s, db.mesh_z, db.fem_mesh + 4 * e, t, &D.econ_host[(size_t)e * 16]);
    }
    float smin = 1e30f, smax = -1e30f, zmin = 1e30f, zmax = -1e30f;
    for (int e = 0; e < ne; ++e) {
        smin = std::min(smin, db.mp_s[e]); smax = std::max(smax, db.mp_s[e]);
        zmin = std::min(zmin, db.mp_z[e]); zmax = std::max(zmax, db.mp_z[e]);
    }
    double area = std::max(1.0, (double)(smax - smin)) * std::max(1.0, (double)(zmax - zmin));
    double h = std::sqrt(area * 2.0 / ne);
    int gnx = std::max(1, (int)std::ceil((smax - smin) / h) + 1);
    int gnz = std::max(1, (int)std::ceil((zmax - zmin) / h) + 1);
    std::vector<int> cnt((size_t)gnx * gnz + 1, 0), cell(ne);
    for (int e = 0; e < ne; ++e) {
        int ix = std::min(gnx - 1, std::max(0, (int)std::floor((db.mp_s[e] - smin) / h)));
        int iz = std::min(gnz - 1, std::max(0, (int)std::floor((db.mp_z[e] - zmin) / h)));
        cell[e] = iz * gnx + ix;
        cnt[cell[e] + 1]++;
    }
    for (size_t i = 1; i < cnt.size(); ++i) cnt[i] += cnt[i - 1];
    std::vector<int> elems(ne), fill(cnt.begin(), cnt.end() - 1);
    for (int e = 0; e < ne; ++e) elems[fill[cell[e]]++] = e;
    HIPCHECK(D.mp_s.ensure(ne));
    HIPCHECK(D.mp_z.ensure(ne));
    HIPCHECK(D.econ.ensure((size_t)ne * 16));
    HIPCHECK(D.eltype.ensure(ne));
    HIPCHECK(D.axis.ensure(ne));
    HIPCHECK(D.cell_start.ensure(cnt.size()));
    HIPCHECK(D.cell_elems.ensure(ne));
    HIPCHECK(hipMemcpy(D.mp_s.p, db.mp_s, ne * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(D.mp_z.p, db.mp_z, ne * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(D.econ.p, D.econ_host.data(), (size_t)ne * 16 * 8, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(D.eltype.p, db.eltype, ne * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(D.axis.p, db.axis, ne * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(D.cell_start.p, cnt.data(), cnt.size() * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(D.cell_elems.p, elems.data(), ne * 4, hipMemcpyHostToDevice));
    // snapshots: use in place when they already live in device memory, page them from
    // host memory when asked to, else copy them once
    D.elem_floats = (size_t)db.nvar * 25 * nt;
    const size_t nbytes = (size_t)ne * D.elem_floats * sizeof(float);
    hipPointerAttribute_t attr;
    bool on_device = false;
    if (hipPointerGetAttributes(&attr, db.snapshots) == hipSuccess)
        on_device = (attr.type == hipMemoryTypeDevice);
    (void)hipGetLastError();
    const float *u = db.snapshots;
    if (D.u_owned) { (void)hipFree(D.u_owned); D.u_owned = nullptr; }
    D.release_pages();
    if (!on_device && page_elems > 0 && page_elems < ne) {
        D.paged = true;
        D.page_elems = page_elems;
        D.npages = (ne + page_elems - 1) / page_elems;
        D.nslot = std::max(2, std::min(nslot, D.npages + 2));
        D.host_u = db.snapshots;
        D.slot.assign(D.nslot, nullptr);
        D.slot_page.assign(D.nslot, -1);
        D.page_slot.assign(D.npages, -1);
        D.nfixed_used = 0;
        D.ptable_host.assign(D.npages, nullptr);
        HIPCHECK(D.ptable.ensure(D.npages));
        HIPCHECK(hipMemcpy(D.ptable.p, D.ptable_host.data(), D.npages * sizeof(float *), hipMemcpyHostToDevice));
        D.copied.assign(D.nslot, nullptr);
        D.freed.assign(D.nslot, nullptr);
        for (int i = 0; i < D.nslot; ++i) {
            HIPCHECK(hipMalloc((void **)&D.slot[i], (size_t)page_elems * D.elem_floats * sizeof(float)));
            HIPCHECK(hipEventCreateWithFlags(&D.copied[i], hipEventDisableTiming));
            HIPCHECK(hipEventCreateWithFlags(&D.freed[i], hipEventDisableTiming));
            HIPCHECK(hipEventRecord(D.freed[i], st));
        }
        u = D.slot[0];
    } else if (!on_device) {
        HIPCHECK(hipMalloc((void **)&D.u_owned, nbytes));
        HIPCHECK(hipMemcpy(D.u_owned, db.snapshots, nbytes, hipMemcpyHostToDevice));
        u = D.u_owned;
    }
    D.d.nelem = ne; D.d.nvar = db.nvar; D.d.nt = nt;
    D.d.u = u;
    D.d.u_first = 0;
    D.d.u_last = D.paged ? std::min(ne, D.page_elems) : ne;
    D.d.mp_s = D.mp_s.p; D.d.mp_z = D.mp_z.p;
    D.d.econ = D.econ.p; D.d.eltype = D.eltype.p; D.d.axis = D.axis.p;
    D.d.gnx = gnx; D.d.gnz = gnz; D.d.gs0 = smin; D.d.gz0 = zmin; D.d.gh = h;
    D.d.cell_start = D.cell_start.p; D.d.cell_elems = D.cell_elems.p;
    D.d.amplitude = db.amplitude;
    return MCK_OK;
}

int needs_of(int param, int *nb) {
    std::fill(nb, nb + 6, 0);
    switch (param) {
    case MCK_PARAM_LAM: case MCK_PARAM_VP: nb[0] = 1; break;
    case MCK_PARAM_VS: nb[0] = nb[1] = 1; break;
    case MCK_PARAM_MU: nb[1] = 1; break;
    case MCK_PARAM_VSH: nb[0] = nb[1] = nb[4] = nb[5] = 1; break;
    case MCK_PARAM_VSV: nb[4] = 1; break;
    case MCK_PARAM_VPH: nb[3] = nb[5] = 1; break;
    case MCK_PARAM_VPV: case MCK_PARAM_ETA: nb[0] = nb[3] = nb[5] = 1; break;
    default: return -1;
    }
    return 0;
}

// Per-kernel window spectrum H_k = -norm df c_k S_k TF_fwd_k TF_bwd_k [TS_fwd TS_bwd],
// the frequency-domain form of apply_filter + irfft + calc_misfit_kernel
// (kernel.f90:410-510, 946-1020): value = sum_k Re(basekernel_k H_k).
int build_kernels(State &G) {
    const mck_setup_t &S = G.S;
    const int nt = S.nt, N = 2 * nextpow2(nt), nw = N / 2 + 1;
    const double dt = S.dt, df = 1.0 / (dt * N);
    std::vector<double> f(nw);
    for (int k = 0; k < nw; ++k) f[k] = df * k;
    std::vector<std::vector<cd>> tff(S.nfilter), tfb(S.nfilter);
    G.ad_tf_host.assign((size_t)S.nfilter * nw * 2, 0.0f);
    for (int i = 0; i < S.nfilter; ++i) {
        tff[i] = filter_tf(S.filters[i], f);
        tfb[i] = tff[i];
        for (int k = 0; k < nw; ++k) {
            G.ad_tf_host[((size_t)i * nw + k) * 2] = (float)tff[i][k].real();
            G.ad_tf_host[((size_t)i * nw + k) * 2 + 1] = (float)tff[i][k].imag();
        }
        if (S.deconv_stf) {
            // add_stfs (filtering.f90:241-470)
            if (!S.stf_fwd) return fail(MCK_EINVAL, "deconv_stf needs stf_fwd");
            auto Fs = rfft(taper_zeropad(S.stf_fwd, nt, N, 2, false), N, dt);
            std::vector<double> e(nt, 0.0);
            for (int t = 0; t < nt; ++t) e[t] = S.stf_src ? S.stf_src[t] : (t == 0 ? 1.0 : 0.0);
            auto Fe = rfft(taper_zeropad(e.data(), nt, N, 2, true), N, dt);
            auto lp = butter_lp(f, 1.0 / (f[nw - 1] * 0.75), 12);
            double mf = 0, mb = 0;
            for (int k = 0; k < nw; ++k) {
                cd l = lp[k] * std::conj(lp[k]);
                cd a = tff[i][k] / Fs[k] * S.fwd.amplitude;
                cd b = a;
                double th = 2 * kPi * f[k] * S.stf_src_shift;
                a = a * Fe[k] * cd(std::cos(th), std::sin(th)) * l;
                b = b * l;
                if (std::isnan(std::abs(a))) a = 0.0;
                if (std::isnan(std::abs(b))) b = 0.0;
                tff[i][k] = a; tfb[i][k] = b;
                mf = std::max(mf, std::abs(a)); mb = std::max(mb, std::abs(b));
            }
            if (std::abs(tff[i][nw - 1]) / mf > 1e-2 || std::abs(tfb[i][nw - 1]) / mb > 1e-2)
                return fail(MCK_EINVAL, "filter is not vanishing fast enough for high frequencies");
        }
    }
    // the factors of TF_fwd TF_bwd besides the filter response G (G^2 rest_f rest_b):
    // add_stfs of a unit response, non-finite values (0/0 bins) set to 0
    std::vector<cd> rest(nw, 1.0);
    if (S.deconv_stf) {
        auto Fs = rfft(taper_zeropad(S.stf_fwd, nt, N, 2, false), N, dt);
        std::vector<double> e(nt, 0.0);
        for (int t = 0; t < nt; ++t) e[t] = S.stf_src ? S.stf_src[t] : (t == 0 ? 1.0 : 0.0);
        auto Fe = rfft(taper_zeropad(e.data(), nt, N, 2, true), N, dt);
        auto lp = butter_lp(f, 1.0 / (f[nw - 1] * 0.75), 12);
        for (int k = 0; k < nw; ++k) {
            cd l = lp[k] * std::conj(lp[k]);
            cd a = cd(1.0) / Fs[k] * S.fwd.amplitude;
            double th = 2 * kPi * f[k] * S.stf_src_shift;
            cd rf = a * Fe[k] * cd(std::cos(th), std::sin(th)) * l, rb = a * l;
            if (!std::isfinite(std::abs(rf))) rf = 0.0;
            if (!std::isfinite(std::abs(rb))) rb = 0.0;
            rest[k] = rf * rb;
        }
    }
    std::vector<cd> ts(nw, 1.0);
    if (!S.deconv_stf)
        for (int k = 0; k < nw; ++k) {
            double a = 2 * kPi * f[k] * S.timeshift_fwd, b = 2 * kPi * f[k] * S.timeshift_bwd;
            ts[k] = cd(std::cos(a), std::sin(a)) * cd(std::cos(b), std::sin(b));
        }
    // seismograms per receiver (cut_and_add_seismogram, kernel.f90:181-385)
    std::map<std::tuple<int, int, int, int>, int> hidx;  // (rec, filter, misfit, window) -> H
    std::map<std::tuple<int, int, int>, std::vector<double>> filtered;  // (rec, filter, misfit)
    // per distinct window: the weighted samples of the filtered seismogram in the window
    // (concatenated), first sample, length, normalisation, filter; the window spectra
    // themselves are computed on the device (k_window_spectra)
    std::vector<double> w_all, h_norm;
    std::vector<int> w_off, h_first, h_ncut;
    std::vector<int> hfilt;
    std::vector<RecDesc> recs(G.R);
    std::vector<int> h_of(G.nk, 0);
    for (int r = 0; r < G.R; ++r) { recs[r] = RecDesc{}; recs[r].nslot = 0; recs[r].first_kernel = -1; recs[r].nkernel = 0; }
    G.k_rec.assign(G.nk, 0);
    G.k_nbk.assign(G.nk, 0);
    G.t_last_p.assign(G.R, 0.0);
    G.t_last_s.assign(G.R, 0.0);
    for (int k = 0; k < G.nk; ++k) {
        int r = S.kernel_receiver[k], fi = S.kernel_filter[k], mf = S.kernel_misfit[k], pa = S.kernel_param[k];
        if (r < 0 || r >= G.R || fi < 0 || fi >= S.nfilter) return fail(MCK_EINVAL, "kernel references unknown receiver/filter");
        if (k > 0 && r < S.kernel_receiver[k - 1]) return fail(MCK_EINVAL, "kernels must be grouped by receiver");
        if (pa == MCK_PARAM_RHO) return fail(MCK_EUNSUPPORTED, "density kernels are not implemented (kernel.f90:554)");
        int nb[6];
        if (needs_of(pa, nb)) return fail(MCK_EINVAL, "unknown model parameter");
        if (!G.pc.full && (nb[1] || nb[3] || nb[4] || nb[5]))
            return fail(MCK_EINVAL, "kernel needs the full strain tensor");
        RecDesc &rd = recs[r];
        if (rd.first_kernel < 0) rd.first_kernel = k;
        rd.nkernel++;
        G.k_rec[k] = r;
        for (int b = 0; b < 6; ++b) G.k_nbk[k] += nb[b];
        double tw0 = S.kernel_window[2 * k], tw1 = S.kernel_window[2 * k + 1];
        if (pa == MCK_PARAM_VP || pa == MCK_PARAM_VPH || pa == MCK_PARAM_VPV || pa == MCK_PARAM_LAM)
            G.t_last_p[r] = std::max(G.t_last_p[r], tw1);
        else
            G.t_last_s[r] = std::max(G.t_last_s[r], tw1);
        if (tw0 < 0.0 || tw1 > (N - 1) * dt || tw1 - tw0 <= 0)
            return fail(MCK_EINVAL, "kernel time window outside the seismogram");
        int ifirst = -1, ncut = 0;
        for (int i = 0; i < N; ++i) {
            double t = i * dt;
            if (t <= tw1 && t >= tw0) { if (ifirst < 0) ifirst = i; ++ncut; }
            if (t > tw1) break;
        }
        if (ncut == 0) return fail(MCK_EINVAL, "kernel time window length is zero");
        auto key = std::make_tuple(r, fi, mf, ifirst * 100000 + ncut);
        int h;
        auto it = hidx.find(key);
        if (it != hidx.end()) h = it->second;
        else {
            // the filtered seismogram depends on (receiver, filter, misfit) only: computed
            // once for all the windows cut from it
            auto fkey = std::make_tuple(r, fi, mf);
            auto fit = filtered.find(fkey);
            if (fit == filtered.end()) {
                const double *seis = S.seis + (size_t)r * nt;
                std::vector<double> x = (mf == MCK_MISFIT_AM) ? cumsum_trap(seis, nt, dt)
                                                              : std::vector<double>(seis, seis + nt);
                auto X = rfft(x, N, dt);
                for (int q = 0; q < nw; ++q) {
                    if (!S.deconv_stf) {
                        double a = 2 * kPi * f[q] * S.timeshift_fwd;
                        X[q] *= cd(std::cos(a), std::sin(a));
                    }
                    X[q] *= tff[fi][q];
                }
                fit = filtered.emplace(fkey, irfft(X, N, df)).first;
            }
            const std::vector<double> &xs = fit->second;
            std::vector<double> w(ncut);
            double norm_term = 0.0;
            for (int n = 0; n < ncut; ++n) norm_term += xs[ifirst + n] * xs[ifirst + n];
            norm_term *= dt;
            double norm = norm_term > 1e-100 ? 1.0 / norm_term : 0.0;
            for (int n = 0; n < ncut; ++n) {
                double om = 1.0;
                if (S.int_scheme == MCK_INT_TRAPEZOIDAL) {
                    om = 0.0;
                    if (n == 0) om += 0.5;
                    if (n == ncut - 1) om += 0.5;
                    if (n > 0 && n < ncut - 1) om = 1.0;
                }
                w[n] = dt * om * xs[ifirst + n];
            }
            h = (int)hfilt.size();
            w_off.push_back((int)w_all.size());
            w_all.insert(w_all.end(), w.begin(), w.end());
            h_first.push_back(ifirst);
            h_ncut.push_back(ncut);
            h_norm.push_back(norm);
            hfilt.push_back(fi);
            hidx[key] = h;
        }
        h_of[k] = h;
    }
    for (int r = 0; r < G.R; ++r)
        if (recs[r].first_kernel < 0) return fail(MCK_EINVAL, "receiver without kernels");
    // Two slot layouts: the base kernels themselves (mck_base_kernels and the
    // differentiable stage) and, for the physical kernels, the model-independent
    // combinations each parameter reads (one slot where the reference combines two)
    auto layout = [&](bool combo, std::vector<RecDesc> &lr, std::vector<KernDesc> &lk, PathConf &pc) -> int {
        lr = recs;
        lk.assign(G.nk, KernDesc{});
        int stored = 0;
        for (int k = 0; k < G.nk; ++k) {
            int nb[kNumBk] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            if (needs_of(S.kernel_param[k], nb)) return fail(MCK_EINVAL, "unknown model parameter");
            if (combo) {
                switch (S.kernel_param[k]) {
                case MCK_PARAM_VS: nb[0] = nb[1] = 0; nb[6] = 1; break;                 // mu - 2 lambda
                case MCK_PARAM_VPV: nb[0] = nb[3] = nb[5] = 0; nb[7] = 1; break;       // lambda - a - c
                case MCK_PARAM_VSH: nb[0] = nb[1] = nb[4] = 0; nb[8] = 1; break;       // + 2 (1 - eta) c
                default: break;
                }
            }
            RecDesc &rd = lr[S.kernel_receiver[k]];
            KernDesc &kd = lk[k];
            kd.param = S.kernel_param[k];
            for (int b = 0; b < kNumBk; ++b) {
                kd.slot[b] = -1;
                if (!nb[b]) continue;
                stored |= 1 << b;
                int s = -1;
                for (int q = 0; q < rd.nslot; ++q)
                    if (rd.slot_h[q] == h_of[k] && rd.slot_bk[q] == b) s = q;
                if (s < 0) {
                    if (rd.nslot >= kMaxSlots) return fail(MCK_EUNSUPPORTED, "too many distinct windows per receiver");
                    s = rd.nslot++;
                    rd.slot_h[s] = h_of[k];
                    rd.slot_bk[s] = b;
                }
                kd.slot[b] = s;
            }
        }
        // base kernels formed per bin: the stored ones and the parts of the combinations
        int need = stored & 0x3f;
        if (stored & (1 << 6)) need |= 1 | 2;
        if (stored & (1 << 7)) need |= 1 | 8 | 32;
        if (stored & (1 << 8)) need |= 1 | 2 | 16;
        pc.need_bk = need;
        pc.nbk_lds = 0;
        for (int b = 0; b < kNumBk; ++b) pc.bk_lds[b] = (stored >> b) & 1 ? pc.nbk_lds++ : -1;
        // order each receiver's slots by window spectrum (one H read per bin and group)
        for (int r = 0; r < G.R; ++r) {
            RecDesc &rd = lr[r];
            std::vector<int> perm(rd.nslot), inv(rd.nslot);
            for (int q = 0; q < rd.nslot; ++q) perm[q] = q;
            std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return rd.slot_h[a] < rd.slot_h[b]; });
            RecDesc o = rd;
            for (int q = 0; q < rd.nslot; ++q) {
                rd.slot_h[q] = o.slot_h[perm[q]];
                rd.slot_bk[q] = o.slot_bk[perm[q]];
                inv[perm[q]] = q;
            }
            for (int k = rd.first_kernel; k < rd.first_kernel + rd.nkernel; ++k)
                for (int b = 0; b < kNumBk; ++b)
                    if (lk[k].slot[b] >= 0) lk[k].slot[b] = inv[lk[k].slot[b]];
            for (int q = 0; q < rd.nslot; ++q) rd.slot_lds[q] = pc.bk_lds[rd.slot_bk[q]];
            rd.ngroup = 0;
            for (int q = 0; q < rd.nslot;) {
                int e = q + 1;
                while (e < rd.nslot && e - q < kSlotGroup && rd.slot_h[e] == rd.slot_h[q]) ++e;
                rd.grp_h[rd.ngroup] = rd.slot_h[q];
                rd.grp_first[rd.ngroup] = q;
                rd.grp_n[rd.ngroup] = e - q;
                ++rd.ngroup;
                q = e;
            }
            int n = 0;
            for (int c = 0; c < 3; ++c) {
                const int b = n;
                for (int q = 0; q < rd.nslot; ++q)
                    if ((rd.slot_bk[q] == 0 ? 0 : rd.slot_bk[q] == 1 ? 1 : 2) == c) {
                        rd.cls_pos[q] = n;
                        rd.cls_h[n++] = rd.slot_h[q];
                    }
                while ((n - b) & 15) { rd.cls_h[n] = rd.cls_h[b]; ++n; }
                rd.cls_end[c] = n;
            }
        }
        pc.max_slots = 1;
        pc.max_cls = 16;
        for (int r = 0; r < G.R; ++r) {
            pc.max_slots = std::max(pc.max_slots, lr[r].nslot);
            pc.max_cls = std::max(pc.max_cls, lr[r].cls_end[2]);
        }
        return MCK_OK;
    };
    std::vector<RecDesc> recs_base;
    std::vector<KernDesc> kerns, kerns_base;
    G.pc_base = G.pc;
    int lrc;
    if ((lrc = layout(false, recs_base, kerns_base, G.pc_base))) return lrc;
    {
        std::vector<RecDesc> rc2;
        if ((lrc = layout(true, rc2, kerns, G.pc))) return lrc;
        recs = rc2;
    }
    {
        // window spectra on the device: exact angle table 2 pi m / N, the filters' transfer
        // functions, time shifts and STF factors in FP64
        const int nH = (int)hfilt.size();
        std::vector<double> cosT(N), sinT(N);
        for (int m = 0; m < N; ++m) {
            const double ang = 2 * kPi * (double)m / (double)N;
            cosT[m] = std::cos(ang);
            sinT[m] = std::sin(ang);
        }
        std::vector<double2> tf2((size_t)S.nfilter * nw * 2), tr2((size_t)nw * 2);
        for (int i = 0; i < S.nfilter; ++i)
            for (int q = 0; q < nw; ++q) {
                tf2[(size_t)i * nw + q] = make_double2(tff[i][q].real(), tff[i][q].imag());
                tf2[((size_t)S.nfilter + i) * nw + q] = make_double2(tfb[i][q].real(), tfb[i][q].imag());
            }
        for (int q = 0; q < nw; ++q) {
            tr2[q] = make_double2(ts[q].real(), ts[q].imag());
            tr2[(size_t)nw + q] = make_double2(rest[q].real(), rest[q].imag());
        }
        DBuf<double> d_w, d_tab, d_norm;
        DBuf<double2> d_tf, d_tr;
        DBuf<int> d_idx;
        HIPCHECK(d_w.ensure(std::max<size_t>(w_all.size(), 1)));
        HIPCHECK(d_tab.ensure(2 * (size_t)N));
        HIPCHECK(d_norm.ensure(nH));
        HIPCHECK(d_tf.ensure(tf2.size()));
        HIPCHECK(d_tr.ensure(tr2.size()));
        HIPCHECK(d_idx.ensure(4 * (size_t)nH));
        std::vector<int> idx(4 * (size_t)nH);
        std::copy(w_off.begin(), w_off.end(), idx.begin());
        std::copy(h_first.begin(), h_first.end(), idx.begin() + nH);
        std::copy(h_ncut.begin(), h_ncut.end(), idx.begin() + 2 * nH);
        std::copy(hfilt.begin(), hfilt.end(), idx.begin() + 3 * nH);
        if (!w_all.empty()) HIPCHECK(hipMemcpy(d_w.p, w_all.data(), w_all.size() * 8, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(d_tab.p, cosT.data(), (size_t)N * 8, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(d_tab.p + N, sinT.data(), (size_t)N * 8, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(d_norm.p, h_norm.data(), (size_t)nH * 8, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(d_tf.p, tf2.data(), tf2.size() * sizeof(double2), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(d_tr.p, tr2.data(), tr2.size() * sizeof(double2), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(d_idx.p, idx.data(), idx.size() * 4, hipMemcpyHostToDevice));
        HIPCHECK(G.H.ensure((size_t)nH * nw));
        HIPCHECK(G.ad_W.ensure((size_t)nH * nw));
        // and for the matrix-core backward path as (re, -im), rows fstride apart, zero
        // beyond nomega
        HIPCHECK(G.Hp.ensure((size_t)nH * G.pc.fstride));
        HIPCHECK(hipMemsetAsync(G.Hp.p, 0, (size_t)nH * G.pc.fstride * sizeof(float2), G.stream));
        const long long nth = (long long)nH * nw;
        hipLaunchKernelGGL(k_window_spectra, dim3((unsigned)((nth + 255) / 256)), dim3(256), 0, G.stream, nH, nw, N,
                           df, d_w.p, d_idx.p, d_idx.p + nH, d_idx.p + 2 * nH, d_norm.p, d_idx.p + 3 * nH, d_tab.p,
                           d_tab.p + N, d_tf.p, d_tf.p + (size_t)S.nfilter * nw, d_tr.p, d_tr.p + nw, G.H.p,
                           G.ad_W.p, G.Hp.p, G.pc.fstride);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(G.stream));
        d_w.release(); d_tab.release(); d_norm.release(); d_tf.release(); d_tr.release(); d_idx.release();
        G.nH = nH;
    }
    HIPCHECK(G.ad_hfilt.ensure(hfilt.size()));
    HIPCHECK(hipMemcpy(G.ad_hfilt.p, hfilt.data(), hfilt.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(G.recs.ensure(G.R));
    HIPCHECK(hipMemcpy(G.recs.p, recs.data(), G.R * sizeof(RecDesc), hipMemcpyHostToDevice));
    HIPCHECK(G.kerns.ensure(G.nk));
    HIPCHECK(hipMemcpy(G.kerns.p, kerns.data(), G.nk * sizeof(KernDesc), hipMemcpyHostToDevice));
    HIPCHECK(G.recs_base.ensure(G.R));
    HIPCHECK(hipMemcpy(G.recs_base.p, recs_base.data(), G.R * sizeof(RecDesc), hipMemcpyHostToDevice));
    HIPCHECK(G.kerns_base.ensure(G.nk));
    HIPCHECK(hipMemcpy(G.kerns_base.p, kerns_base.data(), G.nk * sizeof(KernDesc), hipMemcpyHostToDevice));
    // the matrix-core backward path forms the lambda and mu base kernels only
    {
        bool iso = true;
        for (int k = 0; k < G.nk; ++k) {
            const int p = S.kernel_param[k];
            if (!(p == MCK_PARAM_LAM || p == MCK_PARAM_VP || p == MCK_PARAM_VS || p == MCK_PARAM_MU)) iso = false;
        }
        G.mfma_ok = iso && (G.pc.need_bk & ~3) == 0;
        G.mfma_ok_base = iso && (G.pc_base.need_bk & ~3) == 0;
    }
    G.rec_first.resize(G.R);
    G.rec_count.resize(G.R);
    for (int r = 0; r < G.R; ++r) { G.rec_first[r] = recs[r].first_kernel; G.rec_count[r] = recs[r].nkernel; }
    HIPCHECK(G.d_k_rec.ensure(G.nk));
    HIPCHECK(G.d_k_nbk.ensure(G.nk));
    HIPCHECK(G.d_rec_first.ensure(G.R));
    HIPCHECK(G.d_rec_count.ensure(G.R));
    HIPCHECK(hipMemcpy(G.d_k_rec.p, G.k_rec.data(), G.nk * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(G.d_k_nbk.p, G.k_nbk.data(), G.nk * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(G.d_rec_first.p, G.rec_first.data(), G.R * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(G.d_rec_count.p, G.rec_count.data(), G.R * 4, hipMemcpyHostToDevice));
    return MCK_OK;
}


// LDS of the forward kernel: two FFT buffers and the quarter twiddle table
size_t strain_smem(const PathConf &pc) {
    return (2 * (size_t)pc.M + tq_len(pc.M)) * sizeof(float2);
}

// samples per thread of the strain kernels for a trace length
int spt_for(int nt) {
    int spt = (nt + kStrainThreads - 1) / kStrainThreads;
    int p = 1;
    while (p < spt) p *= 2;
    return p;
}

unsigned grid_x8(int n) { return 8u * (unsigned)((n + 7) / 8); }

// operator set-up of n points (fwd) or pairs (bwd) in launch order (null = identity)
hipError_t launch_setup(State &G, bool fwd, int n, const int *order, const int *keys, const Located *loc) {
    if (n <= 0) return hipSuccess;
    const DbDev &D = fwd ? G.fwd : G.bwd;
    const size_t smem = (size_t)kSetupWaves * op_bytes(0);
    hipLaunchKernelGGL(k_op_setup, dim3((unsigned)((n + kSetupWaves - 1) / kSetupWaves)), dim3(64 * kSetupWaves),
                       smem, G.stream, D.d, G.pc, G.frames.p, n, order, keys, G.R, loc, fwd ? 1 : 0,
                       fwd ? G.opWf.p : G.opWb.p, fwd ? G.opQf.p : G.opQb.p);
    return hipGetLastError();
}

// forward spectra of npts points (order: element-sorted point ids, or null = identity)
// whose elements db holds; the operator set-up of the points (launch_setup) precedes it
hipError_t launch_fwd(State &G, const DevDb &db, int npts, const int *nlive, const int *order, const int *keys,
                      const Located *loc, float2 *F) {
    if (npts <= 0) return hipSuccess;
    const size_t smem = strain_smem(G.pc);
    const dim3 grid(grid_x8(npts)), block(kStrainThreads);
    const int nd = G.pc.nd_f, nv = G.pc.nvar_fwd;
    // the strain traces of the points (16 / nd points per wave; a launch of fewer than
    // 8,192 windows splits each window's samples over up to nt / 64 waves), then their
    // spectra
    bool traced = false;
    const int nwin = (npts + 16 / nd - 1) / (16 / nd);
    const int split = std::max(1, std::min((G.pc.nt + 63) / 64, 8192 / std::max(nwin, 1)));
#define TRC(ND, NV)                                                                               \
    if (nd == ND && nv == NV) {                                                                   \
        hipLaunchKernelGGL((k_fwd_traces<ND, NV>), dim3(grid_x8(nwin), split), dim3(64), 0, G.stream, \
                           db, G.pc, npts, nlive, order, keys, loc, G.opWf.p, G.Ftr.p);           \
        traced = true;                                                                            \
    }
#define TRC_NV(ND) TRC(ND, 2) TRC(ND, 3) TRC(ND, 4) TRC(ND, 5) TRC(ND, 6) TRC(ND, 7) TRC(ND, 8) TRC(ND, 10)
    TRC_NV(1) TRC(4, 2) TRC(4, 4) TRC_NV(6)                    // (monopole fields: 2 or 4 variables)
#undef TRC_NV
#undef TRC
    if (!traced) return hipErrorInvalidValue;
    {
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
#define FWD(SP, ND)                                                                               \
    if (spt_for(G.pc.nt) == SP && nd == ND) {                                                     \
        if (smem > 65536)                                                                         \
            (void)hipFuncSetAttribute((const void *)&k_fwd_spectra<SP, ND>,                       \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);     \
        hipLaunchKernelGGL((k_fwd_spectra<SP, ND>), grid, block, smem, G.stream, db, G.pc,        \
                           npts, nlive, order, keys, loc, G.Ftr.p, G.twq.p, F);                  \
        return hipGetLastError();                                                                 \
    }
#define FWD_ND(SP) FWD(SP, 1) FWD(SP, 4) FWD(SP, 6)
    FWD_ND(1) FWD_ND(2) FWD_ND(4) FWD_ND(8) FWD_ND(16)
#undef FWD_ND
#undef FWD
    return hipErrorInvalidValue;
}

// LDS of the backward kernel: FFT buffers, reused for the base-kernel spectra, the
// quarter twiddle table and the per-slot sums (4 waves + total) of the largest receiver
size_t bwd_smem(const PathConf &pc) {
    const int M = pc.M, nbk = pc.nbk_lds, max_slots = pc.max_slots;
    return (size_t)pc.op_bytes_b
         + (std::max<size_t>(2 * (size_t)M, (size_t)nbk * (M + 1)) + tq_len(M)) * sizeof(float2)
         + (size_t)5 * max_slots * sizeof(double);
}

// backward strain + evaluation of npairs (point, receiver) pairs q = p * R + r whose
// elements db holds; the operator set-up of the pairs (launch_setup) precedes it and
// Wb, Qb point at the set-up of the first pair of this launch
hipError_t launch_bwd(State &G, const DevDb &db, int npairs, const int *order, const Located *loc_b,
                      const float *Wb, const float *Qb, double *phys, double *base_out,
                      const uint8_t *active, int P) {
    if (npairs <= 0) return hipSuccess;
    // base kernels come from the base-kernel slot layout, physical kernels from the
    // combined one
    const PathConf &pc = base_out ? G.pc_base : G.pc;
    const RecDesc *recs = base_out ? G.recs_base.p : G.recs.p;
    const KernDesc *kerns = base_out ? G.kerns_base.p : G.kerns.p;
    const size_t smem = bwd_smem(pc);
    const dim3 grid(grid_x8(npairs)), block(kStrainThreads);
    const int nf = pc.nd_f, nb = pc.nd_b;
    G.stats.spt = spt_for(pc.nt);
    G.stats.nd_f = nf;
    G.stats.nd_b = nb;
    G.stats.bwd_path = 0;
    std::array<hipEvent_t, 3> ev{};
    hipError_t e0 = hipEventCreate(&ev[0]);
    if (e0 == hipSuccess) e0 = hipEventCreate(&ev[1]);
    if (e0 != hipSuccess) return e0;
    G.kev.push_back(ev);
    if ((e0 = hipEventRecord(ev[0], G.stream)) != hipSuccess) return e0;
    G.stats.pair_launches += 1;
    G.stats.pair_pairs += npairs;
#define BWD(SP, NF, NB)                                                                           \
    if (spt_for(pc.nt) == SP && nf == NF && nb == NB) {                                           \
        if (smem > 65536)                                                                         \
            (void)hipFuncSetAttribute((const void *)&k_bwd_eval<SP, NF, NB>,                      \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);     \
        hipLaunchKernelGGL((k_bwd_eval<SP, NF, NB>), grid, block, smem, G.stream, db, pc,         \
                           npairs, order, G.R, loc_b, Wb, Qb, G.opQf.p, G.twq.p,                 \
                           G.Fspec.p, recs, kerns, G.H.p, G.model_idx.p, G.model.p,              \
                           G.S.model_ndepth, G.S.relative_kernel, G.nk, phys, base_out,         \
                           G.xyz.p, active, P, G.d_rec_r.p);                                     \
        const hipError_t el = hipGetLastError();                                                  \
        return el != hipSuccess ? el : hipEventRecord(ev[1], G.stream);                           \
    }
#define BWD_ND(SP) BWD(SP, 1, 1) BWD(SP, 4, 4) BWD(SP, 4, 6) BWD(SP, 6, 4) BWD(SP, 6, 6)
    BWD_ND(1) BWD_ND(2) BWD_ND(4) BWD_ND(8) BWD_ND(16)
#undef BWD_ND
#undef BWD
    return hipErrorInvalidValue;
}

// Backward evaluation of npairs pairs sorted by (element, receiver) (keys, order) whose
// elements db holds, through element spectra: the element runs and the groups of at most
// 16 pairs of one (element, receiver), then per batch of element runs k_elem_spectra and
// k_bwd_mfma.  Wb, Qb: the operator set-up of the first pair.
// (a launch returns a hipError_t, or an MCK_* code whose message is already set)
inline int launch_status(hipError_t e) {
    return e == hipSuccess ? MCK_OK : fail(MCK_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
}
inline int launch_status(int rc) { return rc; }

int launch_bwd_mfma(State &G, const DevDb &db, int npairs, const int *keys, const int *order, const Located *loc_b,
                    const float *Wb, const float *Qb, double *phys, double *base_out, const uint8_t *active, int P) {
    if (npairs <= 0) return MCK_OK;
    const PathConf &pc = base_out ? G.pc_base : G.pc;
    const RecDesc *recs = base_out ? G.recs_base.p : G.recs.p;
    const KernDesc *kerns = base_out ? G.kerns_base.p : G.kerns.p;
    const int R = G.R, n = npairs;
    hipStream_t st = G.stream;
    // work tables: flags and their inclusive scans, element runs, key runs, groups
    HIPCHECK(G.worktab.ensure((size_t)13 * n + 9));
    int *eflag = G.worktab.p, *kflag = eflag + n, *epos = kflag + n, *kpos = epos + n, *gflag = kpos + n,
        *gpos = gflag + n, *erun_first = gpos + n, *erun_elem = erun_first + n + 1, *krun_first = erun_elem + n,
        *group_first = krun_first + n, *group_erun = group_first + n + 1, *erun_group_first = group_erun + n,
        *counts = erun_group_first + n + 1;
    const dim3 gr((n + 255) / 256), bl(256);
    hipLaunchKernelGGL(k_run_flags, gr, bl, 0, st, n, keys, R, eflag, kflag);
    HIPCHECK(hipGetLastError());
    size_t need = 0;
    HIPCHECK(hipcub::DeviceScan::InclusiveSum(nullptr, need, eflag, epos, n, st));
    HIPCHECK(G.sort_tmp.ensure(need));
    HIPCHECK(hipcub::DeviceScan::InclusiveSum(G.sort_tmp.p, need, eflag, epos, n, st));
    HIPCHECK(hipcub::DeviceScan::InclusiveSum(G.sort_tmp.p, need, kflag, kpos, n, st));
    hipLaunchKernelGGL(k_run_table, gr, bl, 0, st, n, keys, R, eflag, epos, kflag, kpos, erun_first, erun_elem,
                       counts, krun_first);
    HIPCHECK(hipGetLastError());
    hipLaunchKernelGGL(k_group_flags, gr, bl, 0, st, n, kpos, krun_first, gflag);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipcub::DeviceScan::InclusiveSum(G.sort_tmp.p, need, gflag, gpos, n, st));
    hipLaunchKernelGGL(k_group_table, gr, bl, 0, st, n, gflag, gpos, eflag, epos, group_first, group_erun,
                       erun_group_first, counts + 1);
    HIPCHECK(hipGetLastError());
    int cnt[2] = {0, 0};
    HIPCHECK(hipMemcpyAsync(cnt, counts, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    const int nerun = cnt[0], ngroups = cnt[1];
    // the per-launch choice: the per-pair kernel when it costs less for these pairs
    if (G.bwd_path_env < 0 && G.cost_pair_us * npairs < G.cost_group_us * ngroups + G.cost_run_us * nerun)
        return launch_status(launch_bwd(G, db, npairs, order, loc_b, Wb, Qb, phys, base_out, active, P));
    G.stats.mfma_runs += nerun;
    const int J = pc.nvar_bwd * 25, KS = mfma_ks(J);
    const size_t per_elem = (size_t)J * pc.ustride * sizeof(float2);
    const int batch = (int)std::max<size_t>(1, std::min<size_t>((size_t)nerun, G.uspec_budget / per_elem));
    HIPCHECK(G.Uspec.ensure((size_t)batch * J * pc.ustride));
    const size_t smem_s = 2 * (size_t)pc.M * sizeof(double2);
    const size_t smem_m = mfma_wave_lds(KS, pc.nd_b, pc.max_cls);
    const int spt = spt_for(pc.nt);
    G.stats.spt = spt;
    G.stats.nd_f = pc.nd_f;
    G.stats.nd_b = pc.nd_b;
    G.stats.bwd_path = 1;
    for (int u0 = 0; u0 < nerun; u0 += batch) {
        const int nb = std::min(batch, nerun - u0);
        bool launched = false;
#define SPEC(SP)                                                                                  \
        if (spt == SP) {                                                                          \
            if (smem_s > 65536)                                                                   \
                (void)hipFuncSetAttribute((const void *)&k_elem_spectra<SP>,                      \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem_s); \
            hipLaunchKernelGGL((k_elem_spectra<SP>), dim3(nb), dim3(kStrainThreads), smem_s, st, db, pc, u0, \
                               nerun, erun_elem, G.twq_d.p, G.Uspec.p);                           \
            launched = true;                                                                      \
        }
        std::array<hipEvent_t, 3> ev{};
        for (auto &e : ev) HIPCHECK(hipEventCreate(&e));
        G.kev.push_back(ev);
        HIPCHECK(hipEventRecord(ev[0], st));
        SPEC(1) SPEC(2) SPEC(4) SPEC(8) SPEC(16)
#undef SPEC
        if (!launched) return fail(MCK_EINVAL, "no element-spectra kernel for this trace length");
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventRecord(ev[1], st));
        launched = false;
        // W waves per element run sharing its spectra through LDS, or one wave per group
        int W = G.mfma_waves;
        if (W > 1 && mfma_wg_lds(KS, pc.nd_b, pc.max_cls, W) > 160 * 1024) W = W > 2 ? 2 : 1;
        if (W > 1 && mfma_wg_lds(KS, pc.nd_b, pc.max_cls, W) > 160 * 1024) W = 1;
        const size_t smem_w = mfma_wg_lds(KS, pc.nd_b, pc.max_cls, W);
#define EVAL1(NF, NB, NV)                                                                         \
        if (pc.nd_f == NF && pc.nd_b == NB && pc.nvar_bwd == NV) {                                \
            if (smem_m > 65536)                                                                   \
                (void)hipFuncSetAttribute((const void *)&k_bwd_mfma<NF, NB, NV>,                  \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem_m); \
            hipLaunchKernelGGL((k_bwd_mfma<NF, NB, NV>), dim3(grid_x8(ngroups)), dim3(64), smem_m, st, pc, u0, nb, \
                               ngroups, group_first, group_erun, G.Uspec.p, order, R, Wb, Qb, G.opQf.p, \
                               G.Fspec.p, recs, kerns, G.Hp.p, pc.max_cls, G.model_idx.p, G.model.p, \
                               G.S.model_ndepth, G.S.relative_kernel, G.nk, phys, base_out, G.xyz.p, \
                               active, P, G.d_rec_r.p, G.mfma_xcd);                               \
            launched = true;                                                                      \
        }
#define EVALW(NF, NB, NV, WW)                                                                     \
        if (W == WW && pc.nd_f == NF && pc.nd_b == NB && pc.nvar_bwd == NV) {                     \
            if (smem_w > 65536)                                                                   \
                (void)hipFuncSetAttribute((const void *)&k_bwd_mfma_wg<NF, NB, NV, WW>,           \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem_w); \
            hipLaunchKernelGGL((k_bwd_mfma_wg<NF, NB, NV, WW>), dim3(nb), dim3(64 * WW), smem_w, st, pc, u0, nb, \
                               erun_group_first, group_first, G.Uspec.p, order, R, Wb, Qb, G.opQf.p,   \
                               G.Fspec.p, recs, kerns, G.Hp.p, pc.max_cls, G.model_idx.p, G.model.p, \
                               G.S.model_ndepth, G.S.relative_kernel, G.nk, phys, base_out, G.xyz.p, \
                               active, P, G.d_rec_r.p);                                           \
            launched = true;                                                                      \
        }
#define EVAL(NF, NB, NV) \
        if (W == 1) { EVAL1(NF, NB, NV) } else { EVALW(NF, NB, NV, 2) EVALW(NF, NB, NV, 4) }
        EVAL(1, 1, 2) EVAL(1, 1, 3) EVAL(4, 4, 2) EVAL(6, 4, 2) EVAL(4, 6, 3) EVAL(6, 6, 3)
#undef EVAL
#undef EVALW
#undef EVAL1
        if (!launched) return fail(MCK_EINVAL, "no matrix-core kernel for these strain components");
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventRecord(ev[2], st));
        G.stats.mfma_launches += 1;
        G.stats.mfma_groups += ngroups;
    }
    return MCK_OK;
}

// the backward path of a launch: element spectra and matrix cores when the kernel set
// allows it (MCK_BWD_PATH=0 forces the per-pair kernel, 1 the matrix-core one)
bool use_mfma(const State &G, bool base) {
    const bool ok = base ? G.mfma_ok_base : G.mfma_ok;
    if (G.bwd_path_env == 0) return false;
    return ok;
}

hipError_t launch_traces(State &G, const DevDb &db, int fwd, int frame, int npts, const Located *loc,
                         float *out) {
    if (npts <= 0) return hipSuccess;
    switch (spt_for(G.pc.nt)) {
#define CASE(SP)                                                                                  \
    case SP:                                                                                      \
        hipLaunchKernelGGL((k_strain_traces<SP>), dim3(npts), dim3(kStrainThreads), 0, G.stream,  \
                           db, G.pc, G.frames.p, fwd, frame, npts, loc, out);                    \
        break;
        CASE(1) CASE(2) CASE(4) CASE(8) CASE(16)
#undef CASE
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Runs launch(db, i0, i1) over the items [i0, i1) of a launch whose elements the device
// view db holds: once over [0, n) when database D is resident; otherwise page by page.
// The first nslot - 2 distinct pages used are copied into buffers of their own on first
// use and stay; any other page is copied into one of the two rotating buffers (on the copy stream: the copy of
// the next page overlaps the kernels of this one, and a buffer is refilled only after
// the kernels that read it).  sorted_keys: element id * key_scale + minor key of every
// item in launch order, ascending (pages then cover disjoint item ranges and untouched
// pages are not copied); null: every page covers all items and the kernels skip the items
// outside it.

template <class Launch>
int run_paged(State &G, DbDev &D, int n, const int *sorted_keys, Launch &&launch, int key_scale = 1) {
    if (n <= 0) return MCK_OK;
    int lrc;
    if (!D.paged) {
        if ((lrc = launch_status(launch(D.d, 0, n)))) return lrc;
        return MCK_OK;
    }
    std::vector<int> bounds(D.npages + 1, n);
    bounds[0] = 0;
    if (sorted_keys) {
        HIPCHECK(G.page_bounds.ensure(D.npages + 1));
        hipLaunchKernelGGL(k_page_bounds, dim3((D.npages + 256) / 256), dim3(256), 0, G.stream, n, sorted_keys,
                           D.page_elems, D.npages, key_scale, G.page_bounds.p);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(bounds.data(), G.page_bounds.p, (D.npages + 1) * sizeof(int),
                                hipMemcpyDeviceToHost, G.stream));
        HIPCHECK(hipStreamSynchronize(G.stream));
    }
    const int nfixed = D.nslot - 2;
    if (sorted_keys) {
        // every page of the launch in a buffer of its own (already, or one still free):
        // the missing ones are copied, the page table updated, and one launch runs over
        // all the items
        int need = 0;
        for (int b = 0; b < D.npages; ++b)
            if (bounds[b + 1] > bounds[b] && D.page_slot[b] < 0) ++need;
        if (need <= nfixed - D.nfixed_used) {
            bool table = false;
            for (int b = 0; b < D.npages; ++b) {
                if (bounds[b + 1] <= bounds[b]) continue;
                int k = D.page_slot[b];
                if (k < 0) {
                    k = D.nfixed_used++;
                    D.page_slot[b] = k;
                }
                if (D.slot_page[k] != b) {
                    const int e0 = b * D.page_elems, e1 = std::min(D.d.nelem, e0 + D.page_elems);
                    HIPCHECK(hipStreamWaitEvent(G.copy_stream, D.freed[k], 0));
                    HIPCHECK(hipMemcpyAsync(D.slot[k], D.host_u + (size_t)e0 * D.elem_floats,
                                            (size_t)(e1 - e0) * D.elem_floats * sizeof(float), hipMemcpyHostToDevice,
                                            G.copy_stream));
                    HIPCHECK(hipEventRecord(D.copied[k], G.copy_stream));
                    D.slot_page[k] = b;
                    G.stats.paged_bytes += (int64_t)((size_t)(e1 - e0) * D.elem_floats * sizeof(float));
                }
                // (the host image changes only here, after the stream synchronisation
                // above: an earlier copy of it has completed)
                if (D.ptable_host[b] != D.slot[k]) {
                    D.ptable_host[b] = D.slot[k];
                    table = true;
                }
                HIPCHECK(hipStreamWaitEvent(G.stream, D.copied[k], 0));
            }
            if (table)
                HIPCHECK(hipMemcpyAsync(D.ptable.p, D.ptable_host.data(), D.npages * sizeof(float *),
                                        hipMemcpyHostToDevice, G.stream));
            DevDb d = D.d;
            d.u = nullptr;
            d.u_first = d.u_last = 0;
            d.page_u = D.ptable.p;
            d.page_elems = D.page_elems;
            return launch_status(launch(d, 0, n));
        }
    }
    for (int b = 0; b < D.npages; ++b) {
        const int i0 = sorted_keys ? bounds[b] : 0, i1 = sorted_keys ? bounds[b + 1] : n;
        if (i1 <= i0) continue;
        int k = D.page_slot[b];
        if (k < 0 && D.nfixed_used < nfixed) {
            k = D.nfixed_used++;
            D.page_slot[b] = k;
        }
        if (k < 0) {
            // a rotating buffer: the one that already holds the page, else the next one
            for (int j = nfixed; j < D.nslot; ++j) if (D.slot_page[j] == b) k = j;
            if (k < 0) { k = nfixed + D.rot; D.rot ^= 1; }
        }
        const int e0 = b * D.page_elems, e1 = std::min(D.d.nelem, e0 + D.page_elems);
        if (D.slot_page[k] != b) {
            HIPCHECK(hipStreamWaitEvent(G.copy_stream, D.freed[k], 0));
            HIPCHECK(hipMemcpyAsync(D.slot[k], D.host_u + (size_t)e0 * D.elem_floats,
                                    (size_t)(e1 - e0) * D.elem_floats * sizeof(float), hipMemcpyHostToDevice,
                                    G.copy_stream));
            HIPCHECK(hipEventRecord(D.copied[k], G.copy_stream));
            D.slot_page[k] = b;
            G.stats.paged_bytes += (int64_t)((size_t)(e1 - e0) * D.elem_floats * sizeof(float));
        }
        HIPCHECK(hipStreamWaitEvent(G.stream, D.copied[k], 0));
        DevDb d = D.d;
        d.u = D.slot[k];
        d.u_first = e0;
        d.u_last = e1;
        if ((lrc = launch_status(launch(d, i0, i1)))) return lrc;
        HIPCHECK(hipEventRecord(D.freed[k], G.stream));
    }
    return MCK_OK;
}

// Voigt rotations of the moment tensor for the reciprocal seismograms, host FP64, the
// same expressions as the reference:
// rotate_symm_tensor_voigt_xyz_src_to_xyz_earth_1d (rotations.f90:555-612)
void h_rot_xyz_src_to_earth(double *t, double phi, double theta) {
    const double sp = std::sin(phi), cp = std::cos(phi), st = std::sin(theta), ct = std::cos(theta);
    const double a1 = t[0], a2 = t[1], a3 = t[2], a4 = t[3], a5 = t[4], a6 = t[5];
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

// rotate_symm_tensor_voigt_xyz_earth_to_xyz_src_1d (rotations.f90:615-671)
void h_rot_earth_to_xyz_src(double *t, double phi, double theta) {
    const double sp = std::sin(phi), cp = std::cos(phi), st = std::sin(theta), ct = std::cos(theta);
    const double a1 = t[0], a2 = t[1], a3 = t[2], a4 = t[3], a5 = t[4], a6 = t[5];
    t[0] = a1 * (cp * cp) * (ct * ct) + 2 * a6 * cp * (ct * ct) * sp + a2 * (ct * ct) * (sp * sp)
         - 2 * a5 * cp * ct * st - 2 * a4 * ct * sp * st + a3 * (st * st);
    t[1] = a2 * (cp * cp) + sp * (-2 * a6 * cp + a1 * sp);
    t[2] = a3 * (ct * ct) + st * (2 * a5 * cp * ct + 2 * a4 * ct * sp + a1 * (cp * cp) * st
                                  + 2 * a6 * cp * sp * st + a2 * (sp * sp) * st);
    t[3] = cp * (a4 * ct + a6 * cp * st + a2 * sp * st) - sp * (a5 * ct + a1 * cp * st + a6 * sp * st);
    t[4] = ct * sp * (a4 * ct + a6 * cp * st + a2 * sp * st) - st * (a3 * ct + a5 * cp * st + a4 * sp * st)
         + cp * ct * (a5 * ct + a1 * cp * st + a6 * sp * st);
    t[5] = cp * (a6 * cp * ct + a2 * ct * sp - a4 * st) - sp * (a1 * cp * ct + a6 * ct * sp - a5 * st);
}

// rotate_symm_tensor_voigt_xyz_to_src_1d (rotations.f90:423-455)
void h_rot_xyz_to_src(double *t, double phi) {
    const double sp = std::sin(phi), cp = std::cos(phi);
    const double a1 = t[0], a2 = t[1], a3 = t[2], a4 = t[3], a5 = t[4], a6 = t[5];
    t[0] = a1 * (cp * cp) + sp * (2 * a6 * cp + a2 * sp);
    t[1] = a2 * (cp * cp) + sp * (-2 * a6 * cp + a1 * sp);
    t[2] = a3;
    t[3] = a4 * cp - a5 * sp;
    t[4] = a5 * cp + a4 * sp;
    t[5] = cp * (a6 * cp + a2 * sp) - sp * (a1 * cp + a6 * sp);
}

// load_seismogram (readfields.f90:1004-1048) / load_fw_points_rdbm (readfields.f90:
// 1469-1604): the seismogram of every receiver by reciprocity into G.seis_host.  The
// source point is located in each receiver frame in the fwd mesh (k_locate, as
// find_element on fwdmesh); the moment tensor is rotated source xyz -> earth ->
// receiver xyz -> receiver (s, phi, z) and divided by the bwd amplitude on the host;
// k_seis contracts it with the bwd strain at that element (FP64).
int reciprocity_seismograms(State &G) {
    const mck_setup_t &S = G.S;
    const int R = G.R, nt = S.nt;
    if (S.fwd.nelem != S.bwd.nelem)
        return fail(MCK_EINVAL, "seismograms by reciprocity need the same mesh in both databases");
    double *d_xyz = nullptr, *d_mb = nullptr, *d_seis = nullptr;
    Located *d_loc = nullptr;
    std::vector<Located> loc(R);
    std::vector<double> mb((size_t)R * 6);
    G.seis_host.assign((size_t)R * nt, 0.0);
    int rc = MCK_OK;
    hipError_t e = hipMalloc((void **)&d_xyz, 3 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&d_loc, (size_t)R * sizeof(Located));
    if (e == hipSuccess) e = hipMalloc((void **)&d_mb, (size_t)R * 6 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&d_seis, (size_t)R * nt * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(d_xyz, G.src_r, 3 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemsetAsync(G.err.p, 0, 4, G.stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_locate, dim3((R + 127) / 128), dim3(128), 0, G.stream, G.fwd.d, G.frames.p, 1, R, 1,
                           d_xyz, S.planet_radius, nullptr, 1, R, d_loc, G.err.p);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(loc.data(), d_loc, (size_t)R * sizeof(Located), hipMemcpyDeviceToHost, G.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(G.stream);
    int err = 0;
    if (e == hipSuccess) e = hipMemcpy(&err, G.err.p, 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && (err & 1)) rc = fail(MCK_ENOTFOUND, "source point not found in the mesh for the reciprocal seismograms");
    if (e == hipSuccess && rc == MCK_OK) {
        const double *m = S.src_mij;
        const double slon = S.src_lon * kDeg2Rad, scolat = (90 - S.src_lat) * kDeg2Rad;
        for (int r = 0; r < R; ++r) {
            // Voigt (tt, pp, rr, rp, rt, tp) of the CMT order (rr, tt, pp, rt, rp, tp), source.f90:90-95
            double *t = &mb[(size_t)r * 6];
            t[0] = m[1]; t[1] = m[2]; t[2] = m[0]; t[3] = m[4]; t[4] = m[3]; t[5] = m[5];
            h_rot_xyz_src_to_earth(t, slon, scolat);
            h_rot_earth_to_xyz_src(t, S.rec_lon[r] * kDeg2Rad, (90 - S.rec_lat[r]) * kDeg2Rad);
            h_rot_xyz_to_src(t, loc[r].phi);
            for (int i = 0; i < 6; ++i) t[i] = t[i] / S.bwd.amplitude;
        }
        e = hipMemcpy(d_mb, mb.data(), mb.size() * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemsetAsync(d_seis, 0, (size_t)R * nt * sizeof(double), G.stream);
    }
    if (e == hipSuccess && rc == MCK_OK) {
        PathConf pcs = G.pc;
        pcs.full = 1;
        const int comp = S.component == 'Z' ? 0 : (S.component == 'R' ? 1 : 2);
        rc = run_paged(G, G.bwd, R, nullptr, [&](const DevDb &d, int, int) {
            hipLaunchKernelGGL(k_seis, dim3(R), dim3(kStrainThreads), 0, G.stream, d, pcs, d_loc, d_mb, comp, d_seis);
            return hipGetLastError();
        });
    }
    if (e == hipSuccess && rc == MCK_OK)
        e = hipMemcpyAsync(G.seis_host.data(), d_seis, (size_t)R * nt * sizeof(double), hipMemcpyDeviceToHost, G.stream);
    if (e == hipSuccess && rc == MCK_OK) e = hipStreamSynchronize(G.stream);
    (void)hipFree(d_xyz); (void)hipFree(d_loc); (void)hipFree(d_mb); (void)hipFree(d_seis);
    if (rc) return rc;
    if (e != hipSuccess) return fail(MCK_EHIP, std::string("reciprocal seismograms: ") + hipGetErrorString(e));
    return MCK_OK;
}

// stable sort of n (element, id) pairs from keys_in/vals_in into keys_out/vals_out
int sort_by_element(State &G, int n, int nelem) {
    if (n <= 1) {
        if (n == 1) {
            HIPCHECK(hipMemcpyAsync(G.vals_out.p, G.vals_in.p, 4, hipMemcpyDeviceToDevice, G.stream));
            HIPCHECK(hipMemcpyAsync(G.keys_out.p, G.keys_in.p, 4, hipMemcpyDeviceToDevice, G.stream));
        }
        return MCK_OK;
    }
    int bits = 1;
    while ((1 << bits) <= nelem) ++bits;
    size_t need = 0;
    HIPCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, G.keys_in.p, G.keys_out.p, G.vals_in.p,
                                                G.vals_out.p, n, 0, bits, G.stream));
    HIPCHECK(G.sort_tmp.ensure(need));
    HIPCHECK(hipcub::DeviceRadixSort::SortPairs(G.sort_tmp.p, need, G.keys_in.p, G.keys_out.p, G.vals_in.p,
                                                G.vals_out.p, n, 0, bits, G.stream));
    return MCK_OK;
}

int setup_tables(State &G) {
    const mck_setup_t &S = G.S;
    // basis
    Basis B;
    lobatto(4, 0.0, 0.0, B.gll);
    lobatto(4, 0.0, 1.0, B.glj);
    deriv_matrix(B.gll, B.G2);
    deriv_matrix(B.glj, B.G1);
    HIPCHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_basis), &B, sizeof(Basis)));
    // frames: cos/sin of lon and colat, evaluated on the host
    G.frames_host.resize(1 + G.R);
    auto mk = [](double latd, double lond) {
        DevFrame f;
        double lon = lond * kDeg2Rad, colat = (90 - latd) * kDeg2Rad;
        f.cos_lon = std::cos(lon); f.sin_lon = std::sin(lon);
        f.cos_colat = std::cos(colat); f.sin_colat = std::sin(colat);
        return f;
    };
    G.frames_host[0] = mk(S.src_lat, S.src_lon);
    for (int r = 0; r < G.R; ++r) G.frames_host[1 + r] = mk(S.rec_lat[r], S.rec_lon[r]);
    HIPCHECK(G.frames.ensure(1 + G.R));
    HIPCHECK(hipMemcpy(G.frames.p, G.frames_host.data(), (1 + G.R) * sizeof(DevFrame), hipMemcpyHostToDevice));
    // source / receiver positions (source.f90:115-131, receiver.f90 get_r)
    {
        double lat = S.src_lat * kDeg2Rad, lon = S.src_lon * kDeg2Rad;
        double rad = S.planet_radius - S.src_depth_km * 1e3;
        G.src_r[0] = std::cos(lat) * std::cos(lon) * rad;
        G.src_r[1] = std::cos(lat) * std::sin(lon) * rad;
        G.src_r[2] = std::sin(lat) * rad;
        G.rec_r.resize(3 * G.R);
        for (int r = 0; r < G.R; ++r) {
            double la = S.rec_lat[r] * kDeg2Rad, lo = S.rec_lon[r] * kDeg2Rad;
            G.rec_r[3 * r] = std::cos(la) * std::cos(lo) * S.planet_radius;
            G.rec_r[3 * r + 1] = std::cos(la) * std::sin(lo) * S.planet_radius;
            G.rec_r[3 * r + 2] = std::sin(la) * S.planet_radius;
        }
        HIPCHECK(G.d_rec_r.ensure(3 * G.R));
        HIPCHECK(hipMemcpy(G.d_rec_r.p, G.rec_r.data(), 3 * G.R * sizeof(double), hipMemcpyHostToDevice));
        for (int i = 0; i < 3; ++i) G.pc.src_r[i] = G.src_r[i];
        G.pc.damp_radius = S.damp_radius;
    }
    // quarter twiddle table w_N^r = exp(-2 pi i r / N), r < N/4, N = 2M (the kernels derive
    // the other quadrants and the length-M twiddles w_M^t = w_N^(2t) from it)
    const int M = G.pc.M, N = 2 * M;
    std::vector<float2> twq(N / 4);
    for (int r = 0; r < N / 4; ++r) {
        double a = -2.0 * M_PI * r / N;
        twq[r] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    HIPCHECK(G.twq.ensure(N / 4));
    HIPCHECK(hipMemcpy(G.twq.p, twq.data(), (N / 4) * sizeof(float2), hipMemcpyHostToDevice));
    {
        std::vector<double2> twd(N / 4);
        for (int r = 0; r < N / 4; ++r) {
            const double a = -2.0 * M_PI * r / N;
            twd[r] = make_double2(std::cos(a), std::sin(a));
        }
        HIPCHECK(G.twq_d.ensure(N / 4));
        HIPCHECK(hipMemcpy(G.twq_d.p, twd.data(), (N / 4) * sizeof(double2), hipMemcpyHostToDevice));
    }
    // radial model table [6][ndepth]
    const int nd = S.model_ndepth;
    std::vector<float> mt((size_t)6 * nd);
    const float *src[6] = {S.model_vp, S.model_vs, S.model_rho, S.model_phi, S.model_xi, S.model_eta};
    for (int c = 0; c < 6; ++c) {
        if (!src[c]) return fail(MCK_EINVAL, "model table missing");
        std::memcpy(&mt[(size_t)c * nd], src[c], nd * sizeof(float));
    }
    HIPCHECK(G.model.ensure(mt.size()));
    HIPCHECK(hipMemcpy(G.model.p, mt.data(), mt.size() * 4, hipMemcpyHostToDevice));
    return MCK_OK;
}

int ensure_chunk(State &G, int points) {
    const int R = G.R;
    const size_t npairs = (size_t)points * std::max(R, 1);
    HIPCHECK(G.xyz.ensure((size_t)points * 3));
    HIPCHECK(G.model_idx.ensure(points));
    HIPCHECK(G.loc.ensure((size_t)points * (1 + R)));
    HIPCHECK(G.Fspec.ensure((size_t)points * G.pc.nd_f * G.pc.fstride));
    HIPCHECK(G.Ftr.ensure((size_t)points * G.pc.nd_f * G.pc.nt));
    HIPCHECK(G.opWf.ensure((size_t)points * G.pc.nd_f * G.pc.nvar_fwd * 25));
    HIPCHECK(G.opQf.ensure((size_t)points * 36));
    HIPCHECK(G.opWb.ensure(npairs * G.pc.nd_b * G.pc.nvar_bwd * 25));
    HIPCHECK(G.opQb.ensure(npairs * 36));
    HIPCHECK(G.keys_in.ensure(npairs));
    HIPCHECK(G.keys_out.ensure(npairs));
    HIPCHECK(G.vals_in.ensure(npairs));
    HIPCHECK(G.vals_out.ensure(npairs));
    HIPCHECK(G.pair_count.ensure(1));
    return MCK_OK;
}

int scratch_address(unsigned char **out);

int init_impl(const mck_setup_t *s) {
    if (!s) return fail(MCK_EINVAL, "null setup");
    State &G = g;
    G.S = *s;
    const mck_setup_t &S = G.S;
    if (S.nt < 16 || S.nt > 4096) return fail(MCK_EUNSUPPORTED, "nt must be in [16, 4096]");
    if (S.nrec <= 0 || S.nkernel <= 0) return fail(MCK_EINVAL, "no receivers or kernels");
    if (S.fwd.nvar != 10 || S.bwd.nvar != 5) return fail(MCK_EINVAL, "merged databases need 10 (fwd) / 5 (bwd) variables");
    if (S.component != 'Z' && S.component != 'R' && S.component != 'T')
        return fail(MCK_EINVAL, "component must be Z, R or T");
    if (!G.stream) HIPCHECK(hipStreamCreateWithFlags(&G.stream, hipStreamNonBlocking));
    G.R = S.nrec;
    G.nk = S.nkernel;
    G.nframes = 1 + G.R;
    PathConf &pc = G.pc;
    pc = PathConf{};
    pc.nt = S.nt;
    pc.M = nextpow2(S.nt);
    // the LDS FFTs take M = 256 samples-per-thread as a compile-time size from 2 samples
    // per thread on
    if (spt_for(pc.nt) >= 2 && pc.M != 256 * spt_for(pc.nt)) return fail(MCK_EUNSUPPORTED, "FFT size");
    pc.nomega = pc.M + 1;
    pc.ustride = (pc.nomega + 15) / 16 * 16;
    pc.fstride = pc.ustride;
    pc.dt = S.dt;
    // strain type (type_parameter.f90:477-486)
    pc.full = 0;
    for (int k = 0; k < S.nkernel; ++k) {
        int p = S.kernel_param[k];
        if (!(p == MCK_PARAM_LAM || p == MCK_PARAM_VP || p == MCK_PARAM_VPH || p == MCK_PARAM_VPV)) pc.full = 1;
    }
    pc.nd = pc.full ? 6 : 1;
    pc.planet_radius = S.planet_radius;
    for (int i = 0; i < 6; ++i) pc.mij[i] = S.src_mij[i];
    // simulations with a non-zero azimuthal factor and the variables they read
    const double *m = S.src_mij;
    bool act[4] = {m[0] != 0.0, (m[1] + m[2]) != 0.0, m[3] != 0.0 || m[4] != 0.0,
                   (m[1] - m[2]) != 0.0 || m[5] != 0.0};
    const int v0[4] = {0, 2, 4, 7}, nv[4] = {2, 2, 3, 3};
    pc.nsim_fwd = 0;
    pc.nvar_fwd = 0;
    for (int i = 0; i < 4; ++i) {
        if (!act[i]) continue;
        pc.fwd_sim[pc.nsim_fwd++] = i + 1;
        for (int q = 0; q < nv[i]; ++q) pc.fwd_vars[pc.nvar_fwd++] = v0[i] + q;
    }
    if (pc.nsim_fwd == 0) return fail(MCK_EINVAL, "moment tensor is zero");
    if (S.component == 'Z') {
        pc.bwd_isim = 1; pc.nvar_bwd = 2; pc.bwd_vars[0] = 0; pc.bwd_vars[1] = 1;
    } else {
        pc.bwd_isim = 2; pc.nvar_bwd = 3; pc.bwd_vars[0] = 2; pc.bwd_vars[1] = 3; pc.bwd_vars[2] = 4;
        pc.bwd_fi[0] = 0; pc.bwd_fi[1] = S.component == 'R' ? 1.0 : 0.0; pc.bwd_fi[2] = S.component == 'T' ? 1.0 : 0.0;
    }
    // carried components: the trace, or the source-frame Voigt components with a non-zero
    // azimuthal factor (monopole simulations have no zp and sp strain)
    const bool fwd_mono = !act[2] && !act[3];
    const bool bwd_mono = pc.bwd_isim == 1;
    auto carry = [&](bool mono, int *nd, int *comps) {
        if (!pc.full) { *nd = 1; comps[0] = 0; return; }
        const int mc[4] = {0, 1, 2, 4};
        *nd = mono ? 4 : 6;
        for (int i = 0; i < *nd; ++i) comps[i] = mono ? mc[i] : i;
    };
    carry(fwd_mono, &pc.nd_f, pc.comps_f);
    carry(bwd_mono, &pc.nd_b, pc.comps_b);
    pc.op_bytes_b = (int)op_bytes(pc.nd_b * pc.nvar_bwd * 25);
    if (pc.nd_f == 1 && pc.nd_b != 1) return fail(MCK_EINVAL, "inconsistent strain types");
    pc.ntaper = 10;
    for (int j = 0; j < 10; ++j) {
        float q = 3.3f * (float)(10 - j) / 10.0f;
        pc.taper[j] = expf(-(q * q));
    }
    int rc;
    if (S.db_page_elems < 0 || S.db_device_pages < 0)
        return fail(MCK_EINVAL, "db_page_elems and db_device_pages must be >= 0");
    if (!G.copy_stream) HIPCHECK(hipStreamCreateWithFlags(&G.copy_stream, hipStreamNonBlocking));
    {
        // paging: as asked, or by itself when the two databases' snapshots exceed 3/4 of
        // the free device memory (two pages of each then take 1/16 of it)
        size_t free_b = 0, total_b = 0;
        HIPCHECK(hipMemGetInfo(&free_b, &total_b));
        const size_t ef = (size_t)S.fwd.nvar * 25 * S.nt * 4, eb = (size_t)S.bwd.nvar * 25 * S.nt * 4;
        const size_t need = (size_t)std::max(S.fwd.nelem, 0) * ef + (size_t)std::max(S.bwd.nelem, 0) * eb;
        int pf = S.db_page_elems, pb = S.db_page_elems, nsf = S.db_device_pages, nsb = S.db_device_pages;
        if (pf == 0 && need > free_b / 4 * 3) {
            // automatic paging: pages of 1/64 of the free memory, buffers filling half of it
            // in proportion to the two databases' sizes
            pf = (int)std::max<size_t>(1, std::min<size_t>(free_b / 64 / ef, INT32_MAX));
            pb = (int)std::max<size_t>(1, std::min<size_t>(free_b / 64 / eb, INT32_MAX));
            const double share = (double)S.fwd.nelem * ef / (double)need;
            nsf = (int)std::max(2.0, share * 32);
            nsb = (int)std::max(2.0, (1.0 - share) * 32);
        }
        if ((rc = upload_db(G.fwd, S.fwd, S.nt, "fwd", pf, nsf, G.stream))) return rc;
        if ((rc = upload_db(G.bwd, S.bwd, S.nt, "bwd", pb, nsb, G.stream))) return rc;
    }
    if ((rc = setup_tables(G))) return rc;
    for (DbDev *D : {&G.fwd, &G.bwd}) {
        // inverse Jacobians and 1/s at the nodes of every element (needs the basis)
        HIPCHECK(D->ejac.ensure((size_t)D->d.nelem * 125));
        hipLaunchKernelGGL(k_elem_tables, dim3((D->d.nelem * 25 + 255) / 256), dim3(256), 0, G.stream, D->d,
                           D->ejac.p);
        HIPCHECK(hipGetLastError());
        D->d.ejac = D->ejac.p;
    }
    HIPCHECK(hipStreamSynchronize(G.stream));
    // seismograms: a copy of the caller's, or by reciprocity from the bwd database
    if (S.seis) {
        G.seis_host.assign(S.seis, S.seis + (size_t)G.R * S.nt);
    } else {
        HIPCHECK(G.err.ensure(1));
        if ((rc = reciprocity_seismograms(G))) return rc;
    }
    G.S.seis = G.seis_host.data();
    if ((rc = build_kernels(G))) return rc;
    HIPCHECK(G.evals.ensure(1));
    HIPCHECK(G.err.ensure(1));
    HIPCHECK(G.elem_count.ensure(8));
    HIPCHECK(G.nwork.ensure(1));
    // points per chunk: forward spectra, locations and sort buffers within a memory budget
    size_t free_b = 0, total_b = 0;
    HIPCHECK(hipMemGetInfo(&free_b, &total_b));
    size_t per_point = (size_t)pc.nd_f * pc.fstride * sizeof(float2) + (size_t)pc.nd_f * pc.nt * sizeof(float)
                     + (size_t)(1 + G.R) * sizeof(Located)
                     + (size_t)G.R * 16 + 64
                     + ((size_t)pc.nd_f * pc.nvar_fwd * 25 + 36) * sizeof(float)
                     + (size_t)G.R * ((size_t)pc.nd_b * pc.nvar_bwd * 25 + 36) * sizeof(float);
    size_t budget = std::min<size_t>(free_b / 3, (size_t)64 << 30);
    G.chunk_points = (int)std::max<size_t>(1024, std::min<size_t>(budget / per_point, 1 << 21));
    // element spectra of the matrix-core path: batches of element runs within an eighth of
    // the free memory (at most 16 GiB)
    G.uspec_budget = std::min<size_t>(free_b / 8, (size_t)16 << 30);
    G.bwd_path_env = -1;
    if (const char *e = std::getenv("MCK_BWD_PATH")) G.bwd_path_env = std::atoi(e);
    // groups in XCD-contiguous ranges (the groups of an element share one L2) unless
    // MCK_MFMA_XCD=0
    G.mfma_xcd = 1;
    if (const char *e = std::getenv("MCK_MFMA_XCD")) G.mfma_xcd = std::atoi(e) != 0;
    G.mfma_waves = 1;
    if (const char *e = std::getenv("MCK_MFMA_WAVES")) {
        const int w = std::atoi(e);
        G.mfma_waves = (w == 2 || w == 4) ? w : 1;
    }
    {
        // c1 (nt 2048, 4 + 4 components, 2 backward variables): 0.355 us per group, 0.93 us
        // per element run, 0.0985 us per pair (k_bwd_mfma 22.6 ms / 63,650 groups,
        // k_elem_spectra 3.0 ms / 3,239 runs, k_bwd_eval 86.2 ms / 875k pairs), scaled
        // with the trace length and the carried components / traces
        const double sc = G.pc.nt / 2048.0, J = G.pc.nvar_bwd * 25.0, KS = std::ceil(J / 4.0);
        G.cost_group_us = 0.355 * sc * (KS * G.pc.nd_b) / (13.0 * 4.0) * (0.5 + 0.5 * G.pc.nd_f / 4.0);
        G.cost_run_us = 0.93 * sc * J / 50.0;
        G.cost_pair_us = 0.0985 * sc * (G.pc.nd_f + G.pc.nd_b) / 8.0;
    }
    // tuning knob: points per evaluation launch (smaller launches keep the forward
    // spectra of a launch in the last-level cache)
    if (const char *e = std::getenv("MCK_CHUNK_POINTS")) {
        const long v = std::atol(e);
        if (v > 0) G.chunk_points = (int)std::min<long>(v, G.chunk_points);
    }
    // the address of the no-allocation entry point's fixed device array on this device
    {
        unsigned char *sp = nullptr;
        if ((rc = scratch_address(&sp))) return rc;
    }
    G.init = true;
    return MCK_OK;
}

// point_in_tetrahedron (polytopes.f90:660-715): barycentric coordinates as ratios of
// quad-precision 4x4 determinants (cofactor expansion along the first row,
// determinant_4_qp polytopes.f90:351-380); inside unless one is below -1e-9.
__float128 det4q(__float128 m[4][4]) {
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

bool point_in_tet(const double *v, const double *p) {
    __float128 m0[4][4], m[4][4];
    for (int c = 0; c < 4; ++c) {
        for (int r = 0; r < 3; ++r) m0[r][c] = v[3 * c + r];
        m0[3][c] = 1;
    }
    const __float128 d0 = det4q(m0);
    // epsilon of the quad-precision kind
    if ((d0 < 0 ? -d0 : d0) < (__float128)1.92592994438723585305597794258492732e-34Q) return false;
    for (int i = 0; i < 4; ++i) {
        for (int r = 0; r < 4; ++r) for (int c = 0; c < 4; ++c) m[r][c] = m0[r][c];
        for (int r = 0; r < 3; ++r) m[r][i] = p[r];
        if (det4q(m) / d0 < (__float128)(-1e-9)) return false;
    }
    return true;
}

// init_weight_tet_mesh (inversion_mesh.f90:1086-1116): inverse of the matrix with columns
// (v_j - v_1, 1), j = 1..4, by cofactors; A[4 i + j] = inv(i, j), so that the barycentric
// coordinate i of x is sum_j inv(i, j) (x - v_1, 1)_j
void tet_barycentric_inverse(const double *v, double *A) {
    double m[4][4];
    for (int c = 0; c < 4; ++c) {
        for (int r = 0; r < 3; ++r) m[r][c] = v[3 * c + r] - v[r];
        m[3][c] = 1.0;
    }
    auto minor3 = [&](int r0, int c0) {
        double t[9];
        int n = 0;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c)
                if (r != r0 && c != c0) t[n++] = m[r][c];
        return t[0] * (t[4] * t[8] - t[5] * t[7]) - t[1] * (t[3] * t[8] - t[5] * t[6])
             + t[2] * (t[3] * t[7] - t[4] * t[6]);
    };
    double cof[4][4], det = 0.0;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) cof[r][c] = (((r + c) & 1) ? -1.0 : 1.0) * minor3(r, c);
    for (int c = 0; c < 4; ++c) det += m[0][c] * cof[0][c];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) A[4 * i + j] = cof[j][i] / det;
}

// Voxels (random_voxel.f90): spherical range r_min, r_max, colat_min, colat_max,
// lon_min, lon_max of the 8 corners (cartesian_to_spherical_range, :166-209)
void vox_range(const double *v, double *g) {
    for (int i = 0; i < 8; ++i) {
        const double x = v[3 * i], y = v[3 * i + 1], z = v[3 * i + 2];
        const double r = std::sqrt(x * x + y * y + z * z), c = std::acos(z / r), l = std::atan2(y, x);
        if (i == 0 || r < g[0]) g[0] = r;
        if (i == 0 || r > g[1]) g[1] = r;
        if (i == 0 || c < g[2]) g[2] = c;
        if (i == 0 || c > g[3]) g[3] = c;
        if (i == 0 || l < g[4]) g[4] = l;
        if (i == 0 || l > g[5]) g[5] = l;
    }
}

// get_volume_vox (random_voxel.f90:91-120)
double vox_volume(const double *g) {
    const double lat1 = kPi / 2.0 - g[2], lat2 = kPi / 2.0 - g[3];
    const double shell = (4.0 / 3.0) * kPi * (g[1] * g[1] * g[1] - g[0] * g[0] * g[0]);
    return shell * (std::fabs(g[4] - g[5]) * std::fabs(std::sin(lat1) - std::sin(lat2)) / (4.0 * kPi));
}

// get_center_vox (random_voxel.f90:125-145) as written: the longitude of the centre is
// the mid colatitude
void vox_center(const double *g, double *c) {
    const double r = (g[1] - g[0]) / 2.0 + g[0], th = (g[3] - g[2]) / 2.0 + g[2], ph = th;
    c[0] = r * std::cos(ph) * std::sin(th);
    c[1] = r * std::sin(ph) * std::sin(th);
    c[2] = r * std::cos(th);
}

// point_in_voxel (random_voxel.f90:232-259)
bool point_in_vox(const double *g, const double *p) {
    const double r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    const double c = std::acos(p[2] / r), l = std::atan2(p[1], p[0]);
    return r >= g[0] && r <= g[1] && c >= g[2] && c <= g[3] && l >= g[4] && l <= g[5];
}

struct Task {
    int nelems = 0, P = 0, nbasis = 1, vox = 0;
    std::vector<double> abinv;   // [ne][4][4] (nbasis 4)
    std::vector<double> tet;     // [ne][12]: tetrahedron vertices, or a voxel's range
    std::vector<double> vol, allowed_var;
    std::vector<uint8_t> masked, live;
    std::vector<int64_t> hbase;
    // the reference's sequential stream (mck_slave_work_sequential): each element starts
    // where the previous one's draws ended; seq_end is the position after the task
    bool sequential = false;
    int64_t seq_end = 0;
};

int prepare_task(State &G, Task &T, int elem_type, int nvpe, int nelems, int nbasis, int nvert,
                 const double *vertices, const int32_t *conn, const int64_t *halton_step,
                 const int64_t *halton_offsets) {
    const mck_setup_t &S = G.S;
    const bool vox = elem_type == 5;
    if (!((elem_type == 4 && nvpe == 4) || (vox && nvpe == 8)))
        return fail(MCK_EUNSUPPORTED, "inversion elements must be tetrahedra (type 4) or voxels (type 5)");
    if (!vox && nbasis != 1 && nbasis != 4)
        return fail(MCK_EINVAL, "tetrahedra have 1 (volumetric) or 4 (on-vertex) basis functions");
    if (vox && nbasis != 1 && nbasis != 8)
        return fail(MCK_EINVAL, "voxels have 1 (volumetric) or 8 (on-vertex) basis functions");
    if (nelems <= 0) return fail(MCK_EINVAL, "empty task");
    T.nelems = nelems;
    T.nbasis = nbasis;
    T.vox = vox ? 1 : 0;
    if (nbasis > 1 && !vox) T.abinv.assign((size_t)nelems * 16, 0.0);
    T.P = S.npoints_per_step;
    T.tet.assign((size_t)nelems * 12, 0.0);
    T.vol.assign(nelems, 1.0);
    T.allowed_var.assign(nelems, 0.0);
    T.masked.assign((size_t)nelems * G.R, 0);
    T.live.assign(nelems, 1);
    T.hbase.assign(nelems, 0);
    const int64_t stride = (int64_t)(S.max_iter + 1) * T.P;
    for (int e = 0; e < nelems; ++e) {
        if (conn[nvpe * e] <= 0) { T.live[e] = 0; continue; }
        double ev[24];
        for (int i = 0; i < nvpe; ++i) {
            int vid = conn[nvpe * e + i] - 1;
            if (vid < 0 || vid >= nvert) return fail(MCK_EINVAL, "connectivity out of range");
            for (int c = 0; c < 3; ++c) ev[3 * i + c] = vertices[3 * vid + c];
        }
        double *v = &T.tet[(size_t)e * 12];
        double c[3], vol;
        if (vox) {
            vox_range(ev, v);
            if (!(v[0] > 0.0) || v[1] <= v[0] || v[3] <= v[2] || v[5] <= v[4])
                return fail(MCK_EINVAL, "degenerate voxel");
            // element containing the source: no integration, zero results (slave.f90:466-482)
            if (S.mask_src_rec && point_in_vox(v, G.src_r)) { T.live[e] = 0; continue; }
            vol = S.int_over_volume ? vox_volume(v) : 1.0;
            vox_center(v, c);
        } else {
            std::copy(ev, ev + 12, v);
            if (nbasis > 1) tet_barycentric_inverse(v, &T.abinv[(size_t)e * 16]);
            if (S.mask_src_rec && point_in_tet(v, G.src_r)) { T.live[e] = 0; continue; }
            vol = S.int_over_volume ? tet_volume(v) : 1.0;
            for (int i = 0; i < 3; ++i) c[i] = (v[i] + v[3 + i] + v[6 + i] + v[9 + i]) / 4.0;
        }
        T.vol[e] = vol;
        double ae = S.allowed_error * vol;
        T.allowed_var[e] = ae * ae;
        // travel_time_larger_Tmax (slave.f90:928-950)
        for (int r = 0; r < G.R; ++r) {
            double d = norm2_3(G.src_r[0] - c[0], G.src_r[1] - c[1], G.src_r[2] - c[2])
                     + norm2_3(G.rec_r[3 * r] - c[0], G.rec_r[3 * r + 1] - c[1], G.rec_r[3 * r + 2] - c[2]);
            double tp = d / (double)S.vp_max, ts = d / (double)S.vs_max;
            T.masked[(size_t)e * G.R + r] = (tp > G.t_last_p[r] && ts > G.t_last_s[r]) ? 1 : 0;
            // receiver inside the element (slave.f90:539-545): same treatment
            if (S.mask_src_rec && (vox ? point_in_vox(v, &G.rec_r[3 * r]) : point_in_tet(v, &G.rec_r[3 * r])))
                T.masked[(size_t)e * G.R + r] = 1;
        }
        T.hbase[e] = halton_offsets ? halton_offsets[e] : (*halton_step + (int64_t)e * stride);
    }
    const int ne = nelems, nk = G.nk, nbk = nelems * nbasis;
    G.nbasis = nbasis;
    G.vox = T.vox;
    HIPCHECK(G.tet.ensure((size_t)ne * 12));
    if (nbasis > 1 && !vox) {
        HIPCHECK(G.abinv.ensure((size_t)ne * 16));
        HIPCHECK(hipMemcpyAsync(G.abinv.p, T.abinv.data(), (size_t)ne * 16 * 8, hipMemcpyHostToDevice, G.stream));
    }
    HIPCHECK(G.volume.ensure(ne));
    HIPCHECK(G.allowed_var.ensure(ne));
    HIPCHECK(G.hbase.ensure(ne));
    HIPCHECK(G.masked.ensure((size_t)ne * G.R));
    HIPCHECK(G.live.ensure(ne));
    HIPCHECK(G.working.ensure(ne));
    HIPCHECK(G.work_list.ensure(ne));
    HIPCHECK(G.active.ensure((size_t)ne * G.R));
    HIPCHECK(G.fsum.ensure((size_t)nbk * nk));
    HIPCHECK(G.f2sum.ensure((size_t)nbk * nk));
    HIPCHECK(G.integral.ensure((size_t)nbk * nk));
    HIPCHECK(G.variance.ensure((size_t)nbk * nk));
    HIPCHECK(G.conv.ensure((size_t)nbk * nk));
    HIPCHECK(G.niter.ensure((size_t)ne * nk));
    HIPCHECK(G.nmodels.ensure(ne));
    HIPCHECK(G.draws.ensure(ne));
    HIPCHECK(G.done.ensure(ne));
    HIPCHECK(G.steps.ensure(ne));
    hipStream_t st = G.stream;
    HIPCHECK(hipMemcpyAsync(G.tet.p, T.tet.data(), T.tet.size() * 8, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(G.volume.p, T.vol.data(), ne * 8, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(G.allowed_var.p, T.allowed_var.data(), ne * 8, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(G.hbase.p, T.hbase.data(), ne * 8, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(G.masked.p, T.masked.data(), T.masked.size(), hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(G.live.p, T.live.data(), ne, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemsetAsync(G.fsum.p, 0, (size_t)nbk * nk * 8, st));
    HIPCHECK(hipMemsetAsync(G.f2sum.p, 0, (size_t)nbk * nk * 8, st));
    HIPCHECK(hipMemsetAsync(G.integral.p, 0, (size_t)nbk * nk * 8, st));
    HIPCHECK(hipMemsetAsync(G.variance.p, 0, (size_t)nbk * nk * 8, st));
    HIPCHECK(hipMemsetAsync(G.conv.p, 0, (size_t)nbk * nk, st));
    HIPCHECK(hipMemsetAsync(G.niter.p, 0, (size_t)ne * nk * 4, st));
    HIPCHECK(hipMemsetAsync(G.nmodels.p, 0, ne * 8, st));
    HIPCHECK(hipMemsetAsync(G.draws.p, 0, ne * 4, st));
    HIPCHECK(hipMemsetAsync(G.done.p, 0, ne, st));
    HIPCHECK(hipMemsetAsync(G.steps.p, 0, ne * 4, st));
    HIPCHECK(hipMemsetAsync(G.evals.p, 0, 8, st));
    HIPCHECK(hipMemsetAsync(G.err.p, 0, 4, st));
    HIPCHECK(hipMemsetAsync(G.elem_count.p, 0, 8 * 8, st));
    return MCK_OK;
}

VoxState vox_state(State &G) {
    VoxState v;
    v.nbasis = G.nbasis;
    v.fsum = G.fsum.p; v.f2sum = G.f2sum.p; v.integral = G.integral.p; v.variance = G.variance.p;
    v.conv = G.conv.p; v.niter = G.niter.p; v.nmodels = G.nmodels.p; v.draws = G.draws.p;
    v.done = G.done.p; v.steps = G.steps.p;
    return v;
}

// Evaluates the MC points of working voxels [w0, w0 + nw) of the current step; writes
// physical kernel values phys[(w0 + w) * P + p][k].
int eval_chunk(State &G, int P, int w0, int nw, std::vector<std::pair<hipEvent_t, hipEvent_t>> &evs) {
    const int np = nw * P, R = G.R;
    const int *wl = G.work_list.p + w0;
    const uint8_t *act = G.active.p + (size_t)w0 * R;
    hipStream_t st = G.stream;
    int rc;
    if ((rc = ensure_chunk(G, np))) return rc;
    Located *loc_f = G.loc.p, *loc_b = G.loc.p + np;
    const bool tet_basis = G.nbasis > 1 && !G.vox;
    hipLaunchKernelGGL(k_points, dim3((np + 255) / 256), dim3(256), 0, st, nw, P, wl, G.hbase.p,
                       G.draws.p, G.tet.p, G.vox, G.S.quasirandom, G.S.model_ndepth, G.S.model_dr, G.xyz.p,
                       G.model_idx.p, G.nbasis, G.abinv.p,
                       tet_basis ? G.bweights.p + (size_t)w0 * P * 4 : nullptr, G.err.p);
    HIPCHECK(hipGetLastError());
    // fwd and bwd meshes may differ: locate frame 0 in the fwd mesh, frames 1.. in the bwd mesh
    hipLaunchKernelGGL(k_locate, dim3((unsigned)((np + 127) / 128)), dim3(128), 0, st, G.fwd.d,
                       G.frames.p, 0, 1, np, G.xyz.p, G.S.planet_radius, nullptr, P, R, loc_f, G.err.p);
    HIPCHECK(hipGetLastError());
    hipLaunchKernelGGL(k_locate, dim3((unsigned)(((long long)np * R + 127) / 128)), dim3(128), 0, st,
                       G.bwd.d, G.frames.p, 1, R, np, G.xyz.p, G.S.planet_radius, act, P, R, loc_b, G.err.p);
    HIPCHECK(hipGetLastError());
    hipLaunchKernelGGL(k_count_points, dim3((nw + 255) / 256), dim3(256), 0, st, nw, R, P, act,
                       G.elem_count.p + 4);
    HIPCHECK(hipGetLastError());
    // forward: points in element order
    hipLaunchKernelGGL(k_keys_points, dim3((np + 255) / 256), dim3(256), 0, st, np, loc_f, G.fwd.d.nelem,
                       act, P, R, G.keys_in.p, G.vals_in.p);
    HIPCHECK(hipGetLastError());
    if ((rc = sort_by_element(G, np, G.fwd.d.nelem))) return rc;
    hipLaunchKernelGGL(k_count_runs, dim3((np + 255) / 256), dim3(256), 0, st, np, G.keys_out.p,
                       G.fwd.d.nelem, G.elem_count.p + 0);
    HIPCHECK(hipGetLastError());
    HIPCHECK(launch_setup(G, true, np, G.vals_out.p, G.keys_out.p, loc_f));
    // points with an active receiver come first in the sorted order: their count
    HIPCHECK(G.nlive.ensure(2));
    hipLaunchKernelGGL(k_page_bounds, dim3(1), dim3(64), 0, st, np, G.keys_out.p, G.fwd.d.nelem, 1, 1, G.nlive.p);
    HIPCHECK(hipGetLastError());
    hipEvent_t a, b;
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    HIPCHECK(hipEventRecord(a, st));
    if ((rc = run_paged(G, G.fwd, np, G.keys_out.p, [&](const DevDb &d, int i0, int i1) {
             // (a paged launch covers one page's range, which holds no inactive point)
             return launch_fwd(G, d, i1 - i0, G.fwd.paged ? nullptr : G.nlive.p + 1, G.vals_out.p + i0,
                               G.keys_out.p + i0, loc_f, G.Fspec.p);
         })))
        return rc;
    HIPCHECK(hipEventRecord(b, st));
    evs.push_back({a, b});
    // backward: active located pairs in element order
    HIPCHECK(hipMemsetAsync(G.pair_count.p, 0, 4, st));
    hipLaunchKernelGGL(k_keys_pairs, dim3((unsigned)(((long long)np * R + 255) / 256)), dim3(256), 0, st,
                       np * R, R, P, act, loc_b, G.keys_in.p, G.vals_in.p, G.pair_count.p);
    HIPCHECK(hipGetLastError());
    int npairs = 0;
    HIPCHECK(hipMemcpyAsync(&npairs, G.pair_count.p, 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    // (element, receiver) order: the runs of equal keys count sum_r |G_r|
    if ((rc = sort_by_element(G, npairs, G.bwd.d.nelem * R))) return rc;
    if (npairs > 0) {
        hipLaunchKernelGGL(k_count_runs, dim3((npairs + 255) / 256), dim3(256), 0, st, npairs, G.keys_out.p,
                           G.bwd.d.nelem * R, G.elem_count.p + 2);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(launch_setup(G, false, npairs, G.vals_out.p, nullptr, loc_b));
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    HIPCHECK(hipEventRecord(a, st));
    {
        const size_t nwb = (size_t)G.pc.nd_b * G.pc.nvar_bwd * 25;
        double *phys = G.phys.p + (size_t)w0 * P * G.nk;
        if (use_mfma(G, false)) {
            if ((rc = run_paged(G, G.bwd, npairs, G.keys_out.p, [&](const DevDb &d, int i0, int i1) {
                     return launch_bwd_mfma(G, d, i1 - i0, G.keys_out.p + i0, G.vals_out.p + i0, loc_b,
                                            G.opWb.p + (size_t)i0 * nwb, G.opQb.p + (size_t)i0 * 36, phys,
                                            nullptr, act, P);
                 }, R)))
                return rc;
        } else if ((rc = run_paged(G, G.bwd, npairs, G.keys_out.p, [&](const DevDb &d, int i0, int i1) {
                        return launch_bwd(G, d, i1 - i0, G.vals_out.p + i0, loc_b, G.opWb.p + (size_t)i0 * nwb,
                                          G.opQb.p + (size_t)i0 * 36, phys, nullptr, act, P);
                    }, R)))
            return rc;
    }
    HIPCHECK(hipEventRecord(b, st));
    evs.push_back({a, b});
    G.stats.strain_launches += 2;
    return MCK_OK;
}

// the element-spectra and matrix-core times of the launches so far (their events are
// then released); keep = false releases them unread, after an error
int fold_kernel_events(State &G, bool keep = true) {
    int rc = MCK_OK;
    for (auto &ev : G.kev) {
        float a = 0.f, b = 0.f;
        if (keep && rc == MCK_OK) {
            const bool pair = ev[2] == nullptr;
            hipError_t e = hipEventSynchronize(pair ? ev[1] : ev[2]);
            if (e == hipSuccess) e = hipEventElapsedTime(&a, ev[0], ev[1]);
            if (e == hipSuccess && !pair) e = hipEventElapsedTime(&b, ev[1], ev[2]);
            if (e != hipSuccess) rc = fail(MCK_EHIP, hipGetErrorString(e));
            if (pair) {
                G.stats.pair_ms += a;
            } else {
                G.stats.spectra_ms += a;
                G.stats.mfma_ms += b;
            }
        }
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    G.kev.clear();
    return rc;
}

int collect_events(State &G, std::vector<std::pair<hipEvent_t, hipEvent_t>> &evs) {
    for (size_t i = 0; i < evs.size(); ++i) {
        float ms = 0.f;
        HIPCHECK(hipEventSynchronize(evs[i].second));
        HIPCHECK(hipEventElapsedTime(&ms, evs[i].first, evs[i].second));
        if (i % 2 == 0) G.stats.strain_fwd_ms += ms; else G.stats.strain_bwd_ms += ms;
        (void)hipEventDestroy(evs[i].first);
        (void)hipEventDestroy(evs[i].second);
    }
    evs.clear();
    return fold_kernel_events(G);
}

// One MC step of every voxel that is still iterating.  Returns the number of voxels
// evaluated in *nwork_out.
int mc_step(State &G, const Task &T, int *nwork_out, std::vector<std::pair<hipEvent_t, hipEvent_t>> &evs) {
    const int ne = T.nelems, P = T.P, R = G.R, nk = G.nk;
    hipStream_t st = G.stream;
    VoxState vs = vox_state(G);
    hipLaunchKernelGGL(k_step_begin, dim3((ne + 255) / 256), dim3(256), 0, st, ne, nk, G.S.max_iter, vs,
                       G.live.p, G.working.p);
    HIPCHECK(hipGetLastError());
    hipLaunchKernelGGL(k_compact, dim3(1), dim3(1024), 0, st, ne, G.working.p, G.work_list.p, G.nwork.p);
    HIPCHECK(hipGetLastError());
    int nwork = 0;
    HIPCHECK(hipMemcpyAsync(&nwork, G.nwork.p, 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    *nwork_out = nwork;
    if (nwork == 0) return MCK_OK;
    hipLaunchKernelGGL(k_receiver_active, dim3((nwork * R + 255) / 256), dim3(256), 0, st, nwork, R, nk,
                       G.work_list.p, G.masked.p, G.d_rec_first.p, G.d_rec_count.p, vs, G.active.p);
    HIPCHECK(hipGetLastError());
    HIPCHECK(G.phys.ensure((size_t)nwork * P * nk));
    if (G.nbasis > 1 && !G.vox) HIPCHECK(G.bweights.ensure((size_t)nwork * P * 4));
    const int wc = std::max(1, G.chunk_points / P);
    int rc;
    for (int w0 = 0; w0 < nwork; w0 += wc)
        if ((rc = eval_chunk(G, P, w0, std::min(wc, nwork - w0), evs))) return rc;
    hipLaunchKernelGGL(k_mc, dim3(nwork), dim3(256), 0, st, nwork, P, nk, R, G.work_list.p, G.phys.p,
                       G.nbasis > 1 && !G.vox ? G.bweights.p : nullptr, G.masked.p, G.active.p, G.d_k_rec.p, G.d_k_nbk.p, G.volume.p, G.allowed_var.p,
                       G.S.allowed_relative_error, vs, G.evals.p, G.err.p);
    HIPCHECK(hipGetLastError());
    G.stats.points += (int64_t)nwork * P;
    G.stats.mc_steps += 1;
    G.stats_vk += (int64_t)nwork * nk;
    return MCK_OK;
}

int finish_task(State &G, const Task &T, double *values, double *variance, int32_t *niterations,
                double *ctime, int64_t *halton_step, const int64_t *halton_offsets, double dev_ms) {
    const int ne = T.nelems, nk = G.nk, nb = T.nbasis;
    hipStream_t st = G.stream;
    std::vector<int> draws(ne), steps(ne);
    unsigned long long ev = 0, cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int err = 0;
    HIPCHECK(hipMemcpyAsync(cnt, G.elem_count.p, 8 * 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(values, G.integral.p, (size_t)ne * nb * nk * 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(variance, G.variance.p, (size_t)ne * nb * nk * 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(niterations, G.niter.p, (size_t)ne * nk * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(draws.data(), G.draws.p, ne * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(steps.data(), G.steps.p, ne * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&ev, G.evals.p, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&err, G.err.p, 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    G.stats.evaluations = (int64_t)ev;
    {
        // algorithmic bytes of the path: every merged block the reference reads, once per
        // launch - forward blocks (10 variables) of the distinct forward elements |F|,
        // backward blocks (5 variables) of the distinct elements of every receiver's frame
        // (sum_r |G_r|) - the coordinates of the points (48 B) and the MC state (Sum f,
        // Sum f^2, integral of every evaluated voxel and kernel, 24 B)
        const int64_t blk = (int64_t)25 * G.pc.nt * sizeof(float);
        G.stats.fwd_locations = (int64_t)cnt[1];
        G.stats.bwd_locations = (int64_t)cnt[3];
        G.stats.fwd_elements = (int64_t)cnt[0];
        G.stats.bwd_elem_receivers = (int64_t)cnt[2];
        G.stats.fwd_bytes = (int64_t)cnt[0] * 10 * blk + 48 * (int64_t)cnt[1];
        G.stats.bwd_bytes = (int64_t)cnt[2] * 5 * blk;
        G.stats.alg_bytes = G.stats.fwd_bytes + G.stats.bwd_bytes + 24 * G.stats_vk;
    }
    if (err & 1) return fail(MCK_ENOTFOUND, "element not found: a point is outside the 12 nearest elements");
    if (err & 2) return fail(MCK_ENAN, "NaN in a Monte Carlo integral");
    if (err & 4) return fail(MCK_EINVAL, "basis-function weight outside [0, 1]: point outside its element");
    long long tot_steps = 0;
    for (int e = 0; e < ne; ++e) tot_steps += steps[e];
    // skipped slots (padding, the source's element) are zeroed with the whole task at its
    // start and never written: the device arrays and their host copies hold zeros there
    G.last_ctime.assign(ne, 0.0);
    for (int e = 0; e < ne; ++e)
        if (T.live[e]) G.last_ctime[e] = tot_steps ? dev_ms * 1e-3 * steps[e] / (double)tot_steps : 0.0;
    if (ctime) std::copy(G.last_ctime.begin(), G.last_ctime.end(), ctime);
    G.last_nelems = ne;
    G.last_nbasis = nb;
    // the Halton stream position after the task
    if (T.sequential) {
        *halton_step = T.seq_end;
    } else if (halton_offsets) {
        int last = -1;
        for (int e = 0; e < ne; ++e) if (T.live[e]) last = e;
        if (last >= 0) *halton_step = T.hbase[last] + (int64_t)draws[last] * T.P;
    } else {
        *halton_step += (int64_t)ne * (G.S.max_iter + 1) * T.P;
    }
    return MCK_OK;
}

// One task in the batched form, or (plain) in the reference's loop order: one element at
// a time, its MC steps one after another.  sequential (plain only): the reference's own
// Halton stream, halton_sequence.f90:109-153 with one global position that every draw
// advances (slave.f90:506-511 draws before the max_iter exit): element e starts where
// element e-1's draws ended; offsets_out (nullable) receives every element's start.
int slave_work_body(Task &T, bool plain, bool sequential, int32_t elem_type, int32_t nvpe, int32_t nelems,
                    int32_t nbasis, int32_t nvertices, const double *vertices, const int32_t *conn,
                    double *values, double *variance, int32_t *niterations, double *ctime,
                    int64_t *halton_step, const int64_t *halton_offsets, int64_t *offsets_out) {
    State &G = g;
    if (!G.init) return fail(MCK_ENOTINIT, "mck_init has not been called");
    if (!vertices || !conn || !values || !variance || !niterations || !halton_step)
        return fail(MCK_EINVAL, "null argument");
    auto t0 = std::chrono::steady_clock::now();
    G.stats = mck_stats_t{};
    (void)fold_kernel_events(G, false);   // events an earlier failed call left
    G.stats_vk = 0;
    G.last_nelems = G.last_nbasis = 0;
    int rc;
    if ((rc = prepare_task(G, T, elem_type, nvpe, nelems, nbasis, nvertices, vertices, conn, halton_step,
                           halton_offsets)))
        return rc;
    T.sequential = sequential;
    T.seq_end = *halton_step;
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    HIPCHECK(hipEventRecord(e0, G.stream));
    std::vector<std::pair<hipEvent_t, hipEvent_t>> evs;
    if (!plain) {
        for (;;) {
            int nwork = 0;
            if ((rc = mc_step(G, T, &nwork, evs))) return rc;
            if ((rc = collect_events(G, evs))) return rc;
            if (nwork == 0) break;
        }
    } else {
        // reference loop order: one element at a time, one MC step at a time
        std::vector<uint8_t> only(nelems, 0);
        for (int e = 0; e < nelems; ++e) {
            if (sequential) {
                T.hbase[e] = T.seq_end;
                if (offsets_out) offsets_out[e] = T.seq_end;
            }
            if (!T.live[e]) continue;
            only.assign(nelems, 0);
            only[e] = 1;
            HIPCHECK(hipMemcpyAsync(G.live.p, only.data(), nelems, hipMemcpyHostToDevice, G.stream));
            if (sequential)
                HIPCHECK(hipMemcpyAsync(G.hbase.p + e, &T.hbase[e], 8, hipMemcpyHostToDevice, G.stream));
            for (;;) {
                int nwork = 0;
                if ((rc = mc_step(G, T, &nwork, evs))) return rc;
                if ((rc = collect_events(G, evs))) return rc;
                if (nwork == 0) break;
            }
            if (sequential) {
                int d = 0;
                HIPCHECK(hipMemcpyAsync(&d, G.draws.p + e, 4, hipMemcpyDeviceToHost, G.stream));
                HIPCHECK(hipStreamSynchronize(G.stream));
                T.seq_end += (int64_t)d * T.P;
            }
        }
    }
    HIPCHECK(hipEventRecord(e1, G.stream));
    HIPCHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    G.stats.device_ms = ms;
    if ((rc = finish_task(G, T, values, variance, niterations, ctime, halton_step, halton_offsets, ms))) return rc;
    G.stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return MCK_OK;
}

int slave_work_impl(bool plain, bool sequential, int32_t elem_type, int32_t nvpe, int32_t nelems, int32_t nbasis,
                    int32_t nvertices, const double *vertices, const int32_t *conn, double *values,
                    double *variance, int32_t *niterations, double *ctime, int64_t *halton_step,
                    const int64_t *halton_offsets, int64_t *offsets_out) {
    Task T;
    const int rc = slave_work_body(T, plain, sequential, elem_type, nvpe, nelems, nbasis, nvertices, vertices,
                                   conn, values, variance, niterations, ctime, halton_step, halton_offsets,
                                   offsets_out);
    // an error return can leave copies from the task's host arrays in flight: let them
    // end before the arrays go
    if (rc && g.init) (void)hipStreamSynchronize(g.stream);
    return rc;
}

// point-level evaluation shared by mck_locate_points / mck_load_strain_points /
// mck_base_kernels: npts points, all receivers active
int points_upload(State &G, int npts, const double *xyz) {
    int rc;
    if ((rc = ensure_chunk(G, npts))) return rc;
    HIPCHECK(hipMemcpyAsync(G.xyz.p, xyz, (size_t)npts * 24, hipMemcpyHostToDevice, G.stream));
    HIPCHECK(hipMemsetAsync(G.err.p, 0, 4, G.stream));
    return MCK_OK;
}

int locate_frame(State &G, int frame, int npts) {
    const DbDev &D = frame == 0 ? G.fwd : G.bwd;
    // results for frame f of point p at loc[p] (stride 1)
    hipLaunchKernelGGL(k_locate, dim3((npts + 127) / 128), dim3(128), 0, G.stream, D.d, G.frames.p, frame,
                       1, npts, G.xyz.p, G.S.planet_radius, nullptr, 1, G.R, G.loc.p, G.err.p);
    HIPCHECK(hipGetLastError());
    return MCK_OK;
}

// ---- the differentiable filter-convolve-window-integrate stage ------------------------
size_t ad_smem(int M) { return (2 * (size_t)M + tq_len(M)) * sizeof(float2); }

bool device_ptr(const void *p) {
    hipPointerAttribute_t a;
    if (!p || hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice;
}

int ad_check(const State &G, int npts, std::initializer_list<const void *> ptrs) {
    if (!G.init) return fail(MCK_ENOTINIT, "mck_init has not been called");
    if (npts < 0) return fail(MCK_EINVAL, "negative point count");
    if (npts == 0) return MCK_OK;
    for (const void *p : ptrs)
        if (!device_ptr(p)) return fail(MCK_EINVAL, "arguments must be device memory");
    return MCK_OK;
}

// W_h G_f^2 of every window spectrum and the forward spectra of the traces
hipError_t ad_prologue(State &G, int npts, const float2 *tf, const float *fw, hipStream_t st) {
    const int nw = G.pc_base.nomega, nd = G.pc_base.nd, M = G.pc_base.M;
    hipError_t e = G.ad_Hg.ensure((size_t)G.nH * nw);
    if (e == hipSuccess) e = G.ad_Fs.ensure((size_t)npts * nd * nw);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_ad_hg, dim3((G.nH * nw + 255) / 256), dim3(256), 0, st, G.nH, nw, G.ad_W.p,
                       G.ad_hfilt.p, tf, G.ad_Hg.p);
    const size_t smem = ad_smem(M);
    switch (spt_for(G.pc_base.nt)) {
#define CASE(SP)                                                                                  \
    case SP:                                                                                      \
        if (smem > 65536)                                                                         \
            (void)hipFuncSetAttribute((const void *)&k_ad_fspec<SP>,                              \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);     \
        hipLaunchKernelGGL((k_ad_fspec<SP>), dim3(npts), dim3(kStrainThreads), smem, st, G.pc_base,    \
                           npts, nd, fw, G.twq.p, G.ad_Fs.p);                                    \
        break;
        CASE(1) CASE(2) CASE(4) CASE(8) CASE(16)
#undef CASE
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t ad_launch_eval(State &G, int npts, const float *bw, double *base, hipStream_t st) {
    const size_t smem = ad_smem(G.pc_base.M);
    const int npairs = npts * G.R;
#define EVAL(SP, ND)                                                                              \
    if (spt_for(G.pc_base.nt) == SP && G.pc_base.nd == ND) {                                                \
        if (smem > 65536)                                                                         \
            (void)hipFuncSetAttribute((const void *)&k_ad_eval<SP, ND>,                           \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);     \
        hipLaunchKernelGGL((k_ad_eval<SP, ND>), dim3(npairs), dim3(kStrainThreads), smem, st,     \
                           G.pc_base, npairs, G.R, bw, G.twq.p, G.ad_Fs.p, G.recs_base.p, G.kerns_base.p,       \
                           G.ad_Hg.p, G.nk, base);                                                \
        return hipGetLastError();                                                                 \
    }
    EVAL(1, 1) EVAL(2, 1) EVAL(4, 1) EVAL(8, 1) EVAL(16, 1)
    EVAL(1, 6) EVAL(2, 6) EVAL(4, 6) EVAL(8, 6) EVAL(16, 6)
#undef EVAL
    return hipErrorInvalidValue;
}

// The backward stages (k_ad_bstage, k_ad_cstage, k_ad_bgrad, k_ad_fgrad) over chunks of
// points whose stage buffers (B and dL/dF spectra and U parts of every pair) stay within
// about 8 GB.
hipError_t ad_launch_grad(State &G, int npts, const float *bw, const double *gbase, float *gbw,
                          float *gfw, hipStream_t st) {
    const PathConf &pc = G.pc_base;
    const int R = G.R, nd = pc.nd, nw = pc.nomega, nt = pc.nt, nF = G.S.nfilter, nk = G.nk;
    const size_t smem = ad_smem(pc.M);
    const size_t per_pt = (size_t)R * (2 * (size_t)nd + nF) * nw * sizeof(float2);
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)npts, ((size_t)8 << 30) / per_pt));
    hipError_t e = G.ad_Bs.ensure((size_t)chunk * R * nd * nw);
    if (e == hipSuccess) e = G.ad_dFs.ensure((size_t)chunk * R * nd * nw);
    if (e == hipSuccess) e = G.ad_Ups.ensure((size_t)chunk * R * nF * nw);
    if (e != hipSuccess) return e;
    const int spt = spt_for(nt);
    for (int p0 = 0; p0 < npts && e == hipSuccess; p0 += chunk) {
        const int np = std::min(chunk, npts - p0), nq = np * R;
        const float *bwc = bw + (size_t)p0 * R * nd * nt;
        const double *gbc = gbase + (size_t)p0 * nk * 6;
        const float2 *Fsc = G.ad_Fs.p + (size_t)p0 * nd * nw;
        float2 *Upc = G.ad_U.p + (size_t)p0 * nF * nw;
        float *gbwc = gbw + (size_t)p0 * R * nd * nt, *gfwc = gfw + (size_t)p0 * nd * nt;
        bool done = false;
#define STAGES(SP, ND)                                                                                    \
        if (!done && spt == SP && nd == ND) {                                                             \
            if (smem > 65536) {                                                                           \
                (void)hipFuncSetAttribute((const void *)&k_ad_bstage<SP>,                                 \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);         \
                (void)hipFuncSetAttribute((const void *)&k_ad_bgrad<SP>,                                  \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);         \
                (void)hipFuncSetAttribute((const void *)&k_ad_fgrad<SP>,                                  \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);         \
            }                                                                                             \
            hipLaunchKernelGGL((k_ad_bstage<SP>), dim3(nq), dim3(kStrainThreads), smem, st, pc, nq, nd,   \
                               bwc, G.twq.p, G.ad_Bs.p);                                                  \
            hipLaunchKernelGGL((k_ad_cstage<SP, ND>), dim3(nq), dim3(kStrainThreads), 0, st, pc, nq, R,   \
                               Fsc, G.ad_Bs.p, G.recs_base.p, G.kerns_base.p, G.ad_Hg.p, G.ad_W.p,        \
                               G.ad_hfilt.p, nF, nk, gbc, G.ad_dFs.p, G.ad_Ups.p);                       \
            hipLaunchKernelGGL((k_ad_bgrad<SP>), dim3(nq), dim3(kStrainThreads), smem, st, pc, nq, nd,    \
                               G.ad_Bs.p, G.twq.p, gbwc);                                                 \
            hipLaunchKernelGGL((k_ad_fgrad<SP>), dim3(np), dim3(kStrainThreads), smem, st, pc, np, R, nd, \
                               nF, G.ad_dFs.p, G.ad_Ups.p, G.twq.p, Upc, gfwc);                           \
            e = hipGetLastError();                                                                        \
            done = true;                                                                                  \
        }
        STAGES(1, 1) STAGES(2, 1) STAGES(4, 1) STAGES(8, 1) STAGES(16, 1)
        STAGES(1, 6) STAGES(2, 6) STAGES(4, 6) STAGES(8, 6) STAGES(16, 6)
#undef STAGES
        if (!done) return hipErrorInvalidValue;
    }
    return e;
}

// integrate_1d_model (hetero = false) or integrate_3d_model (hetero = true) for a task:
// one workgroup per element runs that element's whole MC loop
int integrate_task(bool hetero, int32_t elem_type, int32_t nvpe, int32_t nelems, int32_t nbasis,
                   int32_t nvertices, const double *vertices, const int32_t *conn, double *model,
                   int32_t *nsteps, int64_t *halton_step, const int64_t *halton_offsets) {
    State &G = g;
    if (!G.init) return fail(MCK_ENOTINIT, "mck_init has not been called");
    const bool vox = elem_type == 5;
    if (!((elem_type == 4 && nvpe == 4) || (vox && nvpe == 8)))
        return fail(MCK_EUNSUPPORTED, "inversion elements must be tetrahedra (type 4) or voxels (type 5)");
    if ((!vox && nbasis != 1 && nbasis != 4) || (vox && nbasis != 1 && nbasis != 8))
        return fail(MCK_EINVAL, "invalid number of basis functions");
    if (nelems <= 0 || !vertices || !conn || !model || !halton_step) return fail(MCK_EINVAL, "invalid argument");
    const int64_t kStride = hetero ? (int64_t)kHetMaxStep * kHetPts : (int64_t)kModelMaxStep * kModelPts;
    std::vector<double> geom((size_t)nelems * 12, 0.0), abinv;
    std::vector<uint8_t> live(nelems, 1);
    std::vector<int64_t> base(nelems, 0);
    if (nbasis > 1 && !vox) abinv.assign((size_t)nelems * 16, 0.0);
    for (int e = 0; e < nelems; ++e) {
        base[e] = halton_offsets ? halton_offsets[e] : *halton_step + (int64_t)e * kStride;
        if (conn[nvpe * e] <= 0) { live[e] = 0; continue; }
        double ev[24];
        for (int i = 0; i < nvpe; ++i) {
            const int vid = conn[nvpe * e + i] - 1;
            if (vid < 0 || vid >= nvertices) return fail(MCK_EINVAL, "connectivity out of range");
            for (int c = 0; c < 3; ++c) ev[3 * i + c] = vertices[3 * vid + c];
        }
        double *v = &geom[(size_t)e * 12];
        if (vox) {
            vox_range(ev, v);
            if (!(v[0] > 0.0) || v[1] <= v[0] || v[3] <= v[2] || v[5] <= v[4])
                return fail(MCK_EINVAL, "degenerate voxel");
        } else {
            std::copy(ev, ev + 12, v);
            if (nbasis > 1) tet_barycentric_inverse(v, &abinv[(size_t)e * 16]);
        }
    }
    // per-task buffers kept by the library between calls (grown when a task is larger)
    const size_t nout = (size_t)nelems * nbasis * (hetero ? kHetPar : kModelPar);
    std::vector<int> ns(nelems, 0);
    hipStream_t st = G.stream;
    hipError_t e = G.int_geom.ensure(geom.size());
    if (e == hipSuccess && !abinv.empty()) e = G.int_abinv.ensure(abinv.size());
    DBuf<double> &last = hetero ? G.last_hetero : G.last_model;
    (hetero ? G.last_hetero_nelems : G.last_model_nelems) = 0;
    if (e == hipSuccess) e = last.ensure(nout);
    double *d_out = last.p;
    if (e == hipSuccess) e = G.int_base.ensure(nelems);
    if (e == hipSuccess) e = G.int_live.ensure(nelems);
    if (e == hipSuccess) e = G.int_nsteps.ensure(nelems);
    double *d_geom = G.int_geom.p, *d_abinv = abinv.empty() ? nullptr : G.int_abinv.p;
    int64_t *d_base = G.int_base.p;
    uint8_t *d_live = G.int_live.p;
    int *d_nsteps = G.int_nsteps.p;
    if (e == hipSuccess) e = hipMemcpyAsync(d_geom, geom.data(), geom.size() * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && d_abinv)
        e = hipMemcpyAsync(d_abinv, abinv.data(), abinv.size() * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_base, base.data(), (size_t)nelems * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_live, live.data(), (size_t)nelems, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(G.err.p, 0, 4, G.stream);
    if (e == hipSuccess) {
        if (hetero)
            hipLaunchKernelGGL(k_hetero_integral, dim3(nelems), dim3(64), 0, G.stream, nelems, d_geom, vox ? 1 : 0,
                               nbasis, d_abinv, d_base, d_live, G.hgrid, d_out, d_nsteps, G.err.p);
        else
            hipLaunchKernelGGL(k_model_integral, dim3(nelems), dim3(128), 0, G.stream, nelems, d_geom, vox ? 1 : 0,
                               nbasis, d_abinv, d_base, d_live, G.model.p, G.S.model_ndepth, G.S.model_dr, d_out,
                               d_nsteps, G.err.p);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(model, d_out, nout * 8, hipMemcpyDeviceToHost, G.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ns.data(), d_nsteps, (size_t)nelems * 4, hipMemcpyDeviceToHost, G.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(G.stream);
    int err = 0;
    if (e == hipSuccess) e = hipMemcpy(&err, G.err.p, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        // copies from the local arrays may still be in flight
        (void)hipStreamSynchronize(G.stream);
        return fail(MCK_EHIP, std::string(hetero ? "mck_integrate_hetero: " : "mck_integrate_model: ") + hipGetErrorString(e));
    }
    if (err & 4) return fail(MCK_EINVAL, "basis function weight outside [0, 1] (point outside its element)");
    if (err & 8) return fail(MCK_ENAN, "NaN in a model integral (mc_integration.f90:113-130)");
    if (nsteps) std::copy(ns.begin(), ns.end(), nsteps);
    if (!halton_offsets) *halton_step += (int64_t)nelems * kStride;
    (hetero ? G.last_hetero_nelems : G.last_model_nelems) = nelems;
    (hetero ? G.last_hetero_nbasis : G.last_model_nbasis) = nbasis;
    return MCK_OK;
}

}  // namespace

// =====================================================================================
// C-ABI
// =====================================================================================
extern "C" {

int mck_init(const mck_setup_t *setup) {
    if (g.init) mck_finalize();
    int rc = init_impl(setup);
    if (rc) {
        g.init = false;
        return rc;
    }
    const char *t = std::getenv("MCK_TUNE");
    if (t && std::atoi(t) != 0) {
        // 2048 points through the mantle: a Fibonacci sphere at radii between 0.57 and
        // 0.98 planet radii
        const int n = 2048;
        const double R = g.S.planet_radius, gold = kPi * (3.0 - std::sqrt(5.0));
        std::vector<double> xyz(3 * (size_t)n);
        for (int i = 0; i < n; ++i) {
            const double z = 1.0 - 2.0 * (i + 0.5) / n, rho = std::sqrt(1.0 - z * z), ph = gold * i;
            const double r = R * (0.57 + 0.41 * std::fmod(i * 0.6180339887498949, 1.0));
            xyz[3 * i] = r * rho * std::cos(ph);
            xyz[3 * i + 1] = r * rho * std::sin(ph);
            xyz[3 * i + 2] = r * z;
        }
        mck_tuning_t out;
        if ((rc = mck_tune(n, xyz.data(), 2, &out))) return rc;
    }
    return MCK_OK;
}

int mck_set_tuning(const mck_tuning_t *t) {
    State &G = g;
    if (!G.init) return fail(MCK_ENOTINIT, "mck_init has not been called");
    if (!t || t->bwd_path < -1 || t->bwd_path > 1 || (t->mfma_waves != 1 && t->mfma_waves != 2 && t->mfma_waves != 4))
        return fail(MCK_EINVAL, "invalid tuning");
    G.bwd_path_env = t->bwd_path;
    G.mfma_waves = t->mfma_waves;
    G.mfma_xcd = t->mfma_xcd ? 1 : 0;
    if (t->cost_group_us > 0.0 && t->cost_run_us > 0.0 && t->cost_pair_us > 0.0) {
        G.cost_group_us = t->cost_group_us;
        G.cost_run_us = t->cost_run_us;
        G.cost_pair_us = t->cost_pair_us;
    }
    return MCK_OK;
}

int mck_tune(int32_t npts, const double *xyz, int32_t reps, mck_tuning_t *out) {
    State &G = g;
    if (!G.init) return fail(MCK_ENOTINIT, "mck_init has not been called");
    if (npts <= 0 || !xyz || !out) return fail(MCK_EINVAL, "invalid argument");
    reps = std::max(1, reps);
    struct Cand { int path, waves, xcd; };
    const Cand all[5] = {{1, 1, 0}, {1, 1, 1}, {1, 2, 0}, {1, 4, 0}, {0, 1, 0}};
    const bool mfma = G.mfma_ok_base && G.mfma_ok;
    std::vector<int> cands;
    if (mfma) cands = {0, 1, 2, 3};
    cands.push_back(4);
    *out = mck_tuning_t{};
    const int save_path = G.bwd_path_env, save_w = G.mfma_waves, save_x = G.mfma_xcd;
    std::vector<double> base((size_t)npts * G.nk * 6);
    // best of reps timed runs after an untimed one; the statistics of the timed runs
    auto timed = [&](double *best_ms) -> int {
        int rc = mck_base_kernels(npts, xyz, base.data());
        G.stats = mck_stats_t{};
        for (int r = 0; r < reps && rc == MCK_OK; ++r) {
            const auto t0 = std::chrono::steady_clock::now();
            rc = mck_base_kernels(npts, xyz, base.data());
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            *best_ms = r == 0 ? ms : std::min(*best_ms, ms);
        }
        return rc;
    };
    auto restore = [&](int rc) {
        G.bwd_path_env = save_path;
        G.mfma_waves = save_w;
        G.mfma_xcd = save_x;
        return rc;
    };
    double best = 0.0, best_mfma = 0.0;
    int bi = -1, bm = -1;
    for (int ci : cands) {
        const Cand &c = all[ci];
        G.bwd_path_env = c.path;
        G.mfma_waves = c.waves;
        G.mfma_xcd = c.xcd;
        double t = 0.0;
        int rc = timed(&t);
        if (rc) return restore(rc);
        out->candidate_ms[ci] = t;
        // the per-launch costs from the kernels' event times
        if (ci == 0 && G.stats.mfma_groups > 0 && G.stats.mfma_runs > 0) {
            out->cost_group_us = 1e3 * G.stats.mfma_ms / G.stats.mfma_groups;
            out->cost_run_us = 1e3 * G.stats.spectra_ms / (double)G.stats.mfma_runs;
        }
        if (ci == 4 && G.stats.pair_pairs > 0) out->cost_pair_us = 1e3 * G.stats.pair_ms / (double)G.stats.pair_pairs;
        if (bi < 0 || t < best) { best = t; bi = ci; }
        if (ci < 4 && (bm < 0 || t < best_mfma)) { best_mfma = t; bm = ci; }
    }
    out->ncandidates = (int32_t)cands.size();
    out->default_ms = out->candidate_ms[mfma ? 0 : 4];
    Cand chosen = all[bi];
    double chosen_ms = best;
    if (mfma && out->cost_group_us > 0.0 && out->cost_pair_us > 0.0) {
        // the per-launch choice with the measured costs and the fastest matrix-core variant
        G.cost_group_us = out->cost_group_us;
        G.cost_run_us = out->cost_run_us;
        G.cost_pair_us = out->cost_pair_us;
        G.bwd_path_env = -1;
        G.mfma_waves = all[bm].waves;
        G.mfma_xcd = all[bm].xcd;
        double t = 0.0;
        int rc = timed(&t);
        if (rc) return restore(rc);
        out->auto_ms = t;
        // kept unless a fixed candidate is faster by more than the timing noise
        if (t <= 1.02 * best) {
            chosen = Cand{-1, all[bm].waves, all[bm].xcd};
            chosen_ms = t;
        }
    } else {
        out->cost_group_us = G.cost_group_us;
        out->cost_run_us = G.cost_run_us;
        out->cost_pair_us = out->cost_pair_us > 0.0 ? out->cost_pair_us : G.cost_pair_us;
    }
    G.bwd_path_env = chosen.path;
    G.mfma_waves = chosen.waves;
    G.mfma_xcd = chosen.xcd;
    out->bwd_path = chosen.path;
    out->mfma_waves = chosen.waves;
    out->mfma_xcd = chosen.xcd;
    out->best_ms = chosen_ms;
    return MCK_OK;
}

int mck_slave_work(int32_t elem_type, int32_t nvpe, int32_t nelems, int32_t nbasis, int32_t nvertices,
                   const double *vertices, const int32_t *conn, double *kernel_values,
                   double *kernel_variance, int32_t *niterations, double *computation_time,
                   int64_t *halton_step, const int64_t *halton_offsets) {
    return slave_work_impl(false, false, elem_type, nvpe, nelems, nbasis, nvertices, vertices, conn, kernel_values,
                           kernel_variance, niterations, computation_time, halton_step, halton_offsets, nullptr);
}

int mck_slave_work_plain(int32_t elem_type, int32_t nvpe, int32_t nelems, int32_t nbasis,
                         int32_t nvertices, const double *vertices, const int32_t *conn,
                         double *kernel_values, double *kernel_variance, int32_t *niterations,
                         double *computation_time, int64_t *halton_step,
                         const int64_t *halton_offsets) {
    return slave_work_impl(true, false, elem_type, nvpe, nelems, nbasis, nvertices, vertices, conn, kernel_values,
                           kernel_variance, niterations, computation_time, halton_step, halton_offsets, nullptr);
}

int mck_slave_work_sequential(int32_t elem_type, int32_t nvpe, int32_t nelems, int32_t nbasis,
                              int32_t nvertices, const double *vertices, const int32_t *conn,
                              double *kernel_values, double *kernel_variance, int32_t *niterations,
                              double *computation_time, int64_t *halton_step, int64_t *offsets_used) {
    return slave_work_impl(true, true, elem_type, nvpe, nelems, nbasis, nvertices, vertices, conn, kernel_values,
                           kernel_variance, niterations, computation_time, halton_step, nullptr, offsets_used);
}

int mck_locate_points(int32_t frame, int32_t npts, const double *xyz, int32_t *elem, double *xi,
                      double *eta, int32_t *ties) {
    State &G = g;
    if (!G.init) return fail(MCK_ENOTINIT, "mck_init has not been called");
    if (frame < 0 || frame > G.R || npts < 0 || (npts && (!xyz || !elem || !xi || !eta)))
        return fail(MCK_EINVAL, "invalid argument");
    if (npts == 0) return MCK_OK;
    int rc;
    if ((rc = points_upload(G, npts, xyz))) return rc;
    if ((rc = locate_frame(G, frame, npts))) return rc;
    std::vector<Located> L(npts);
    HIPCHECK(hipMemcpyAsync(L.data(), G.loc.p, npts * sizeof(Located), hipMemcpyDeviceToHost, G.stream));
    HIPCHECK(hipStreamSynchronize(G.stream));
    for (int i = 0; i < npts; ++i) {
        elem[i] = L[i].elem + 1;   // 1-based, 0 = not found
        xi[i] = L[i].xi;
        eta[i] = L[i].eta;
        if (ties) ties[i] = L[i].tie;
    }
    return MCK_OK;
}

int mck_load_strain_points(int32_t frame, int32_t npts, const double *xyz, float *traces) {
    State &G = g;
    if (!G.init) return fail(MCK_ENOTINIT, "mck_init has not been called");
    if (frame < 0 || frame > G.R || npts < 0 || (npts && (!xyz || !traces)))
        return fail(MCK_EINVAL, "invalid argument");
    if (npts == 0) return MCK_OK;
    int rc;
    if ((rc = points_upload(G, npts, xyz))) return rc;
    if ((rc = locate_frame(G, frame, npts))) return rc;
    float *d_tr = nullptr;
    const size_t n = (size_t)npts * G.pc.nd * G.pc.nt;
    HIPCHECK(hipMalloc((void **)&d_tr, n * sizeof(float)));
    rc = run_paged(G, frame == 0 ? G.fwd : G.bwd, npts, nullptr, [&](const DevDb &d, int, int) {
        return launch_traces(G, d, frame == 0, frame, npts, G.loc.p, d_tr);
    });
    hipError_t e = hipSuccess;
    if (rc == MCK_OK) e = hipMemcpyAsync(traces, d_tr, n * sizeof(float), hipMemcpyDeviceToHost, G.stream);
    if (rc == MCK_OK && e == hipSuccess) e = hipStreamSynchronize(G.stream);
    (void)hipFree(d_tr);
    if (rc) return rc;
    if (e != hipSuccess) return fail(MCK_EHIP, hipGetErrorString(e));
    int err = 0;
    HIPCHECK(hipMemcpy(&err, G.err.p, 4, hipMemcpyDeviceToHost));
    if (err & 1) return fail(MCK_ENOTFOUND, "element not found");
    return MCK_OK;
}

int mck_base_kernels(int32_t npts, const double *xyz, double *base) {
    State &G = g;
    if (!G.init) return fail(MCK_ENOTINIT, "mck_init has not been called");
    if (npts < 0 || (npts && (!xyz || !base))) return fail(MCK_EINVAL, "invalid argument");
    if (npts == 0) return MCK_OK;
    const int R = G.R;
    int rc;
    if ((rc = points_upload(G, npts, xyz))) return rc;
    // model index of every point (k_points computes it from Halton draws; here from xyz)
    std::vector<int> midx(npts);
    for (int p = 0; p < npts; ++p) {
        double r = norm2_3(xyz[3 * p], xyz[3 * p + 1], xyz[3 * p + 2]);
        int id = (int)(r / (double)G.S.model_dr + 1);
        id = std::max(1, std::min(G.S.model_ndepth, id));
        midx[p] = id - 1;
    }
    HIPCHECK(hipMemcpyAsync(G.model_idx.p, midx.data(), npts * 4, hipMemcpyHostToDevice, G.stream));
    Located *loc_f = G.loc.p, *loc_b = G.loc.p + npts;
    hipLaunchKernelGGL(k_locate, dim3((npts + 127) / 128), dim3(128), 0, G.stream, G.fwd.d, G.frames.p, 0, 1,
                       npts, G.xyz.p, G.S.planet_radius, nullptr, 1, R, loc_f, G.err.p);
    hipLaunchKernelGGL(k_locate, dim3(((long long)npts * R + 127) / 128), dim3(128), 0, G.stream, G.bwd.d,
                       G.frames.p, 1, R, npts, G.xyz.p, G.S.planet_radius, nullptr, 1, R, loc_b, G.err.p);
    HIPCHECK(hipGetLastError());
    double *d_base = nullptr, *d_phys = nullptr;
    HIPCHECK(hipMalloc((void **)&d_base, (size_t)npts * G.nk * 6 * 8));
    if (hipMalloc((void **)&d_phys, (size_t)npts * G.nk * 8) != hipSuccess) {
        (void)hipFree(d_base);
        return fail(MCK_ENOMEM, "device allocation failed");
    }
    hipStream_t st = G.stream;
    auto body = [&]() -> int {
        int rc2;
        HIPCHECK(hipMemsetAsync(d_base, 0, (size_t)npts * G.nk * 6 * 8, st));
        HIPCHECK(launch_setup(G, true, npts, nullptr, nullptr, loc_f));
        if ((rc2 = run_paged(G, G.fwd, npts, nullptr, [&](const DevDb &d, int, int) {
                 return launch_fwd(G, d, npts, nullptr, nullptr, nullptr, loc_f, G.Fspec.p);
             })))
            return rc2;
        if (use_mfma(G, true)) {
            // the located pairs in (element, receiver) order, as slave_work evaluates them
            HIPCHECK(hipMemsetAsync(G.pair_count.p, 0, 4, st));
            hipLaunchKernelGGL(k_keys_pairs, dim3((unsigned)(((long long)npts * R + 255) / 256)), dim3(256), 0, st,
                               npts * R, R, 1, nullptr, loc_b, G.keys_in.p, G.vals_in.p, G.pair_count.p);
            HIPCHECK(hipGetLastError());
            int npairs = 0;
            HIPCHECK(hipMemcpyAsync(&npairs, G.pair_count.p, 4, hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            if ((rc2 = sort_by_element(G, npairs, G.bwd.d.nelem * R))) return rc2;
            HIPCHECK(launch_setup(G, false, npairs, G.vals_out.p, nullptr, loc_b));
            const size_t nwb = (size_t)G.pc.nd_b * G.pc.nvar_bwd * 25;
            return run_paged(G, G.bwd, npairs, G.keys_out.p, [&](const DevDb &d, int i0, int i1) {
                return launch_bwd_mfma(G, d, i1 - i0, G.keys_out.p + i0, G.vals_out.p + i0, loc_b,
                                       G.opWb.p + i0 * nwb, G.opQb.p + (size_t)i0 * 36, d_phys, d_base, nullptr, 1);
            }, R);
        }
        HIPCHECK(launch_setup(G, false, npts * R, nullptr, nullptr, loc_b));
        return run_paged(G, G.bwd, npts * R, nullptr, [&](const DevDb &d, int, int) {
            return launch_bwd(G, d, npts * R, nullptr, loc_b, G.opWb.p, G.opQb.p, d_phys, d_base, nullptr, 1);
        });
    };
    rc = body();
    hipError_t e = hipSuccess;
    if (rc == MCK_OK) e = hipMemcpyAsync(base, d_base, (size_t)npts * G.nk * 6 * 8, hipMemcpyDeviceToHost, st);
    if (rc == MCK_OK && e == hipSuccess) e = hipStreamSynchronize(st);
    else (void)hipStreamSynchronize(st);
    {
        const int r3 = fold_kernel_events(G, rc == MCK_OK && e == hipSuccess);
        if (rc == MCK_OK && e == hipSuccess && r3) rc = r3;
    }
    (void)hipFree(d_base);
    (void)hipFree(d_phys);
    if (rc) return rc;
    if (e != hipSuccess) return fail(MCK_EHIP, hipGetErrorString(e));
    int err = 0;
    HIPCHECK(hipMemcpy(&err, G.err.p, 4, hipMemcpyDeviceToHost));
    if (err & 1) return fail(MCK_ENOTFOUND, "element not found");
    return MCK_OK;
}

int mck_ad_dims(int32_t *nd, int32_t *nt, int32_t *nomega, int32_t *nfilter, int32_t *nrec,
                int32_t *nkernel) {
    const State &G = g;
    if (!G.init) return fail(MCK_ENOTINIT, "mck_init has not been called");
    if (nd) *nd = G.pc.nd;
    if (nt) *nt = G.pc.nt;
    if (nomega) *nomega = G.pc.nomega;
    if (nfilter) *nfilter = G.S.nfilter;
    if (nrec) *nrec = G.R;
    if (nkernel) *nkernel = G.nk;
    return MCK_OK;
}

int mck_ad_transfer(float *tf) {
    const State &G = g;
    if (!G.init) return fail(MCK_ENOTINIT, "mck_init has not been called");
    if (!tf) return fail(MCK_EINVAL, "null output");
    std::memcpy(tf, G.ad_tf_host.data(), G.ad_tf_host.size() * sizeof(float));
    return MCK_OK;
}

int mck_ad_forward(int32_t npts, const float *tf, const float *fw, const float *bw, double *base,
                   void *stream) {
    State &G = g;
    int rc = ad_check(G, npts, {tf, fw, bw, base});
    if (rc || npts == 0) return rc;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = ad_prologue(G, npts, reinterpret_cast<const float2 *>(tf), fw, st);
    if (e == hipSuccess) e = ad_launch_eval(G, npts, bw, base, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(MCK_EHIP, std::string("mck_ad_forward: ") + hipGetErrorString(e));
    return MCK_OK;
}

int mck_ad_backward(int32_t npts, const float *tf, const float *fw, const float *bw,
                    const double *gbase, float *gtf, float *gfw, float *gbw, void *stream) {
    State &G = g;
    int rc = ad_check(G, npts, {tf, fw, bw, gbase, gtf, gfw, gbw});
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const int nw = G.pc.nomega, nF = G.S.nfilter, n = nF * nw;
    if (npts == 0) {
        // no points: the response gradient is zero
        if (!device_ptr(gtf)) return fail(MCK_EINVAL, "arguments must be device memory");
        hipError_t e = hipMemsetAsync(gtf, 0, (size_t)n * sizeof(float2), st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return fail(MCK_EHIP, hipGetErrorString(e));
        return MCK_OK;
    }
    const int chunk = 256, nchunk = (npts + chunk - 1) / chunk;
    hipError_t e = G.ad_U.ensure((size_t)npts * n);
    if (e == hipSuccess) e = G.ad_part.ensure((size_t)nchunk * n);
    if (e == hipSuccess) e = ad_prologue(G, npts, reinterpret_cast<const float2 *>(tf), fw, st);
    if (e == hipSuccess) e = ad_launch_grad(G, npts, bw, gbase, gbw, gfw, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_ad_upart, dim3((n + 255) / 256, nchunk), dim3(256), 0, st, n, npts, chunk,
                           G.ad_U.p, G.ad_part.p);
        hipLaunchKernelGGL(k_ad_tfgrad, dim3((n + 255) / 256), dim3(256), 0, st, n, nchunk, G.ad_part.p,
                           reinterpret_cast<const float2 *>(tf), reinterpret_cast<float2 *>(gtf));
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(MCK_EHIP, std::string("mck_ad_backward: ") + hipGetErrorString(e));
    return MCK_OK;
}

int mck_integrate_model(int32_t elem_type, int32_t nvpe, int32_t nelems, int32_t nbasis, int32_t nvertices,
                        const double *vertices, const int32_t *conn, double *model, int32_t *nsteps,
                        int64_t *halton_step, const int64_t *halton_offsets) {
    return integrate_task(false, elem_type, nvpe, nelems, nbasis, nvertices, vertices, conn, model, nsteps,
                          halton_step, halton_offsets);
}

int mck_set_hetero(int32_t nhet, const double *rtp, const float *dlnhet) {
    State &G = g;
    if (!G.init) return fail(MCK_ENOTINIT, "mck_init has not been called");
    if (nhet <= 0 || !rtp || !dlnhet) return fail(MCK_EINVAL, "invalid heterogeneity model");
    // load_het_rtpv (heterogeneities.f90:63-110): r km, latitude and longitude in degrees
    // to single-precision Cartesian coordinates (spherical_to_cartesian_point in FP64)
    std::vector<float> xyz((size_t)nhet * 3);
    for (int i = 0; i < nhet; ++i) {
        const double r = rtp[3 * i], t = rtp[3 * i + 1], ph = rtp[3 * i + 2];
        if (std::fabs(t) > 90.0) return fail(MCK_EINVAL, "heterogeneity latitude outside [-90, 90]");
        const double rr = r * 1000.0, th = (90.0 - t) * kDeg2Rad, pp = ph * kDeg2Rad;
        xyz[3 * i] = (float)(rr * std::cos(pp) * std::sin(th));
        xyz[3 * i + 1] = (float)(rr * std::sin(pp) * std::sin(th));
        xyz[3 * i + 2] = (float)(rr * std::cos(th));
    }
    // uniform grid of cubic cells, about two model points per cell
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < nhet; ++i)
        for (int c = 0; c < 3; ++c) {
            lo[c] = std::min(lo[c], (double)xyz[3 * i + c]);
            hi[c] = std::max(hi[c], (double)xyz[3 * i + c]);
        }
    double ext[3], vol = 1.0;
    for (int c = 0; c < 3; ++c) { ext[c] = std::max(hi[c] - lo[c], 1.0); vol *= ext[c]; }
    double h = std::cbrt(vol / std::max(1.0, nhet / 2.0));
    int n[3];
    for (int c = 0; c < 3; ++c) n[c] = std::max(1, std::min(512, (int)std::ceil(ext[c] / h)));
    h = std::max({ext[0] / n[0], ext[1] / n[1], ext[2] / n[2]}) * (1.0 + 1e-9);
    const size_t ncell = (size_t)n[0] * n[1] * n[2];
    std::vector<int> cell(nhet), start(ncell + 1, 0);
    for (int i = 0; i < nhet; ++i) {
        int ci[3];
        for (int c = 0; c < 3; ++c) {
            int v = (int)std::floor(((double)xyz[3 * i + c] - lo[c]) / h);
            ci[c] = v < 0 ? 0 : (v >= n[c] ? n[c] - 1 : v);
        }
        cell[i] = (ci[2] * n[1] + ci[1]) * n[0] + ci[0];
        start[cell[i] + 1]++;
    }
    for (size_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
    std::vector<int> fill(start.begin(), start.end() - 1), idx(nhet);
    std::vector<float> sorted((size_t)nhet * 3);
    for (int i = 0; i < nhet; ++i) {
        const int j = fill[cell[i]]++;
        idx[j] = i;
        for (int c = 0; c < 3; ++c) sorted[3 * (size_t)j + c] = xyz[3 * (size_t)i + c];
    }
    HIPCHECK(G.het_xyz.ensure(sorted.size()));
    HIPCHECK(G.het_dln.ensure((size_t)nhet * 3));
    HIPCHECK(G.het_idx.ensure(nhet));
    HIPCHECK(G.het_cells.ensure(start.size()));
    HIPCHECK(hipMemcpy(G.het_xyz.p, sorted.data(), sorted.size() * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(G.het_dln.p, dlnhet, (size_t)nhet * 3 * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(G.het_idx.p, idx.data(), (size_t)nhet * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(G.het_cells.p, start.data(), start.size() * 4, hipMemcpyHostToDevice));
    HetGrid &hg = G.hgrid;
    hg.n = nhet;
    hg.nx = n[0]; hg.ny = n[1]; hg.nz = n[2];
    hg.x0 = lo[0]; hg.y0 = lo[1]; hg.z0 = lo[2]; hg.h = h;
    hg.xyz = G.het_xyz.p; hg.idx = G.het_idx.p; hg.cell_start = G.het_cells.p; hg.dln = G.het_dln.p;
    G.het_ready = true;
    return MCK_OK;
}

int mck_integrate_hetero(int32_t elem_type, int32_t nvpe, int32_t nelems, int32_t nbasis, int32_t nvertices,
                         const double *vertices, const int32_t *conn, double *hetero, int32_t *nsteps,
                         int64_t *halton_step, const int64_t *halton_offsets) {
    if (!g.het_ready) return fail(MCK_ENOTINIT, "mck_set_hetero has not been called");
    return integrate_task(true, elem_type, nvpe, nelems, nbasis, nvertices, vertices, conn, hetero, nsteps,
                          halton_step, halton_offsets);
}

int mck_seismograms(double *seis) {
    State &G = g;
    if (!G.init) return fail(MCK_ENOTINIT, "mck_init has not been called");
    if (!seis) return fail(MCK_EINVAL, "null seis");
    std::memcpy(seis, G.seis_host.data(), G.seis_host.size() * sizeof(double));
    return MCK_OK;
}

int mck_results_init(int32_t int_type, int32_t npoints, int32_t nelements, int32_t nbasis, int32_t nmodel,
                     int32_t nhet, const int32_t *connectivity) {
    State &G = g;
    auto &Rs = G.res;
    if (!G.init) return fail(MCK_ENOTINIT, "mck_init has not been called");
    if (int_type != 0 && int_type != 1) return fail(MCK_EINVAL, "int_type must be 0 (onvertices) or 1 (volumetric)");
    if (npoints <= 0 || nelements <= 0 || nbasis <= 0 || (nmodel != 0 && nmodel != kModelPar) ||
        (nhet != 0 && nhet != kHetPar))
        return fail(MCK_EINVAL, "invalid result dimensions");
    if (int_type == 1 && (npoints != nelements || nbasis != 1))
        return fail(MCK_EINVAL, "volumetric results: one basis function per element and npoints = nelements");
    if (int_type == 0 && !connectivity) return fail(MCK_EINVAL, "onvertices results need the connectivity");
    Rs.release();
    if (int_type == 0) {
        Rs.conn.assign(connectivity, connectivity + (size_t)nelements * nbasis);
        for (int v : Rs.conn)
            if (v < 1 || v > npoints) return fail(MCK_EINVAL, "connectivity out of range");
    }
    Rs.int_type = int_type;
    Rs.npoints = npoints;
    Rs.nelements = nelements;
    Rs.nbasis = nbasis;
    Rs.nmodel = nmodel;
    Rs.nhet = nhet;
    const size_t nk = (size_t)G.nk;
    HIPCHECK(Rs.K.ensure((size_t)npoints * nk));
    HIPCHECK(Rs.V.ensure((size_t)npoints * nk));
    HIPCHECK(Rs.M.ensure((size_t)npoints * std::max(nmodel, 1)));
    HIPCHECK(Rs.H.ensure((size_t)npoints * std::max(nhet, 1)));
    HIPCHECK(Rs.niter.ensure((size_t)nelements * nk));
    HIPCHECK(Rs.ctime.ensure(nelements));
    HIPCHECK(Rs.proc.ensure(nelements));
    HIPCHECK(hipMemsetAsync(Rs.K.p, 0, Rs.K.n * 8, G.stream));
    HIPCHECK(hipMemsetAsync(Rs.V.p, 0, Rs.V.n * 8, G.stream));
    HIPCHECK(hipMemsetAsync(Rs.M.p, 0, Rs.M.n * 8, G.stream));
    HIPCHECK(hipMemsetAsync(Rs.H.p, 0, Rs.H.n * 8, G.stream));
    HIPCHECK(hipMemsetAsync(Rs.niter.p, 0, Rs.niter.n * 4, G.stream));
    HIPCHECK(hipMemsetAsync(Rs.ctime.p, 0, Rs.ctime.n * 8, G.stream));
    HIPCHECK(hipMemsetAsync(Rs.proc.p, 0, Rs.proc.n * 4, G.stream));
    if (int_type == 0) {
        HIPCHECK(Rs.dconn.ensure(Rs.conn.size()));
        HIPCHECK(hipMemcpyAsync(Rs.dconn.p, Rs.conn.data(), Rs.conn.size() * 4, hipMemcpyHostToDevice, G.stream));
    }
    HIPCHECK(hipStreamSynchronize(G.stream));
    Rs.init = true;
    return MCK_OK;
}

}  // extern "C"

// The no-allocation result assembly (mck_results_add_noalloc): its device arrays live in
// `scratch`, a fixed array that every GPU the library runs on holds until the process
// ends.  kNoAllocScratch bytes bound the task: kNoAllocMaxPairs (element, basis function)
// pairs.  Host code takes the array's address per device with hipGetSymbolAddress (the
// name itself is a host-side stand-in).
constexpr size_t kNoAllocScratch = (size_t)16 << 20;
constexpr int kNoAllocMaxPairs = 131072;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

namespace {

// Device arrays of one result-assembly call: the (point, slot) pairs and their sort, the
// runs of equal points, the task's element ids, the cub temporary storage and the staging
// area of host inputs (double) and host iteration counts (int)
struct AddScratch {
    int *pt_in = nullptr, *pt_out = nullptr, *sl_in = nullptr, *sl_out = nullptr;
    int *seg_point = nullptr, *seg_count = nullptr, *seg_start = nullptr, *nseg = nullptr;
    int *elems = nullptr, *elem_idx = nullptr, *stage_i = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    double *stage = nullptr;
    size_t stage_len = 0;
};

std::map<int, unsigned char *> g_scratch_ptr;   // device -> address of its `scratch`

// the address of `scratch` on the current device, taken once per device (mck_init)
int scratch_address(unsigned char **out) {
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
        void *p = nullptr;
        HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        it = g_scratch_ptr.emplace(dev, (unsigned char *)p).first;
    }
    *out = it->second;
    return MCK_OK;
}

// carves n elements of T, 256-byte aligned, from [*cur, end); null when they do not fit
template <typename T>
T *carve(unsigned char **cur, unsigned char *end, size_t n) {
    uintptr_t a = ((uintptr_t)*cur + 255) & ~(uintptr_t)255;
    if (a + n * sizeof(T) > (uintptr_t)end) return nullptr;
    *cur = (unsigned char *)(a + n * sizeof(T));
    return (T *)a;
}

int results_add_impl(int32_t nelems, const int32_t *elems, const double *kernel_values,
                     const double *kernel_variance, const int32_t *niterations, const double *computation_time,
                     const double *model, const double *hetero, int32_t rank, bool noalloc) {
    State &G = g;
    auto &Rs = G.res;
    if (!G.init) return fail(MCK_ENOTINIT, "mck_init has not been called");
    if (!Rs.init) return fail(MCK_ENOTINIT, "mck_results_init has not been called");
    if (nelems < 0 || (nelems > 0 && !elems)) return fail(MCK_EINVAL, "invalid task elements");
    if (nelems == 0) return MCK_OK;
    const int nb = Rs.nbasis, nk = G.nk;
    const bool last = !kernel_values && !kernel_variance && !niterations && !computation_time;
    if (!last && (!kernel_values || !kernel_variance || !niterations))
        return fail(MCK_EINVAL, "kernel_values, kernel_variance and niterations: all given or all NULL");
    if (last && (G.last_nelems != nelems || G.last_nbasis != nb))
        return fail(MCK_EINVAL, "no results of a mck_slave_work call with these nelems and nbasis in device memory");
    const bool with_model = Rs.nmodel > 0;
    if (with_model && !model && (G.last_model_nelems != nelems || G.last_model_nbasis != nb))
        return fail(MCK_EINVAL, "no model integrals of a mck_integrate_model call with these nelems and nbasis");
    const bool with_het = Rs.nhet > 0;
    if (with_het && !hetero && (G.last_hetero_nelems != nelems || G.last_hetero_nbasis != nb))
        return fail(MCK_EINVAL, "no heterogeneity integrals of a mck_integrate_hetero call with these nelems and nbasis");
    const int npair = nelems * nb;
    if (noalloc) {
        if (npair > kNoAllocMaxPairs)
            return fail(MCK_EINVAL, "mck_results_add_noalloc: more than " + std::to_string(kNoAllocMaxPairs)
                                    + " (element, basis function) pairs per task");
        if ((!last && (!device_ptr(kernel_values) || !device_ptr(kernel_variance) || !device_ptr(niterations)))
            || (with_model && model && !device_ptr(model)) || (with_het && hetero && !device_ptr(hetero)))
            return fail(MCK_EINVAL, "mck_results_add_noalloc takes its kernel values, variances, iteration "
                                    "counts and model integrals in device memory (or NULL)");
    }
    // element ids checked on the host; the grouping by point runs on the device
    std::vector<int> eidx(nelems, -1);
    for (int iel = 0; iel < nelems; ++iel) {
        const int e = elems[iel];
        if (e == -1) continue;
        if (e < 1 || e > Rs.nelements) return fail(MCK_EINVAL, "element id out of range");
        eidx[iel] = e - 1;
    }
    hipStream_t st = G.stream;
    int bits = 1;
    while ((1 << bits) <= Rs.npoints) ++bits;
    size_t n1 = 0, n2 = 0, n3 = 0;
    // (size queries: no device work, no allocation)
    HIPCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, n1, (int *)nullptr, (int *)nullptr, (int *)nullptr,
                                                (int *)nullptr, npair, 0, bits, st));
    HIPCHECK(hipcub::DeviceRunLengthEncode::Encode(nullptr, n2, (int *)nullptr, (int *)nullptr, (int *)nullptr,
                                                   (int *)nullptr, npair, st));
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, n3, (int *)nullptr, (int *)nullptr, npair, st));
    const size_t tmp_bytes = std::max({n1, n2, n3, (size_t)1});
    // staging of host inputs
    const size_t nv = (size_t)nelems * nb * nk, nm = with_model ? (size_t)nelems * nb * kModelPar : 0;
    const size_t nh = with_het ? (size_t)nelems * nb * kHetPar : 0;
    const bool sv = !last && !device_ptr(kernel_values), svar = !last && !device_ptr(kernel_variance);
    const bool sm = with_model && model && !device_ptr(model);
    const bool sh = with_het && hetero && !device_ptr(hetero);
    const bool sct = last || (computation_time && !device_ptr(computation_time));
    const bool sni = !last && !device_ptr(niterations);
    const size_t need = (sv ? nv : 0) + (svar ? nv : 0) + (sm ? nm : 0) + (sh ? nh : 0) + (sct ? nelems : 0);
    AddScratch A;
    if (noalloc) {
        unsigned char *base = nullptr;
        int rc;
        if ((rc = scratch_address(&base))) return rc;
        unsigned char *cur = base, *end = base + kNoAllocScratch;
        int **ints[] = {&A.pt_in, &A.pt_out, &A.sl_in, &A.sl_out, &A.seg_point, &A.seg_count, &A.seg_start};
        bool ok = true;
        for (int **q : ints) ok = ok && (*q = carve<int>(&cur, end, npair)) != nullptr;
        ok = ok && (A.nseg = carve<int>(&cur, end, 1)) != nullptr;
        ok = ok && (A.elems = carve<int>(&cur, end, nelems)) != nullptr;
        ok = ok && (A.elem_idx = carve<int>(&cur, end, nelems)) != nullptr;
        ok = ok && (A.stage = carve<double>(&cur, end, std::max<size_t>(need, 1))) != nullptr;
        ok = ok && (A.tmp = carve<unsigned char>(&cur, end, tmp_bytes)) != nullptr;
        if (!ok) return fail(MCK_EINVAL, "mck_results_add_noalloc: the task does not fit the fixed device array");
    } else {
        for (DBuf<int> *b : {&Rs.pt_in, &Rs.pt_out, &Rs.sl_in, &Rs.sl_out, &Rs.seg_point, &Rs.seg_count, &Rs.seg_start})
            HIPCHECK(b->ensure(npair));
        HIPCHECK(Rs.nseg.ensure(1));
        HIPCHECK(Rs.elems.ensure(nelems));
        HIPCHECK(Rs.elem_idx.ensure(nelems));
        HIPCHECK(Rs.tmp.ensure(tmp_bytes));
        HIPCHECK(Rs.stage.ensure(std::max<size_t>(need, 1)));
        if (sni) HIPCHECK(Rs.stage_i.ensure((size_t)nelems * nk));
        A.pt_in = Rs.pt_in.p; A.pt_out = Rs.pt_out.p; A.sl_in = Rs.sl_in.p; A.sl_out = Rs.sl_out.p;
        A.seg_point = Rs.seg_point.p; A.seg_count = Rs.seg_count.p; A.seg_start = Rs.seg_start.p;
        A.nseg = Rs.nseg.p; A.elems = Rs.elems.p; A.elem_idx = Rs.elem_idx.p; A.tmp = Rs.tmp.p;
        A.stage = Rs.stage.p; A.stage_i = Rs.stage_i.p;
    }
    A.tmp_bytes = tmp_bytes;
    HIPCHECK(hipMemcpyAsync(A.elems, elems, (size_t)nelems * 4, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(A.elem_idx, eidx.data(), (size_t)nelems * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_results_pairs, dim3((npair + 255) / 256), dim3(256), 0, st, nelems, nb,
                       Rs.int_type == 0 ? 1 : 0, Rs.npoints, A.elems, Rs.dconn.p, A.pt_in, A.sl_in);
    HIPCHECK(hipGetLastError());
    {
        // stable radix sort by point (task order kept within a point), then the runs;
        // counts past the last run are not read: every segment's start is the sum of the
        // counts before it, whatever follows
        size_t b1 = tmp_bytes, b2 = tmp_bytes, b3 = tmp_bytes;
        HIPCHECK(hipcub::DeviceRadixSort::SortPairs(A.tmp, b1, A.pt_in, A.pt_out, A.sl_in, A.sl_out, npair, 0, bits,
                                                    st));
        HIPCHECK(hipcub::DeviceRunLengthEncode::Encode(A.tmp, b2, A.pt_out, A.seg_point, A.seg_count, A.nseg, npair,
                                                       st));
        HIPCHECK(hipcub::DeviceScan::ExclusiveSum(A.tmp, b3, A.seg_count, A.seg_start, npair, st));
    }
    // inputs: in device memory already, or staged
    const double *dv = G.integral.p, *dvar = G.variance.p, *dm = G.last_model.p, *dh = G.last_hetero.p;
    const double *dct = nullptr;
    const int *dni = G.niter.p;
    const std::vector<double> &ct_last = G.last_ctime;
    {
        double *q = A.stage;
        auto put = [&](const double *src, size_t n) -> const double * {
            double *d = q;
            q += n;
            return hipMemcpyAsync(d, src, n * 8, hipMemcpyHostToDevice, st) == hipSuccess ? d : nullptr;
        };
        if (!last) {
            dv = sv ? put(kernel_values, nv) : kernel_values;
            dvar = svar ? put(kernel_variance, nv) : kernel_variance;
            if (!dv || !dvar) return fail(MCK_EHIP, "mck_results_add: staging copy failed");
            if (!sni) {
                dni = niterations;
            } else {
                HIPCHECK(hipMemcpyAsync(A.stage_i, niterations, (size_t)nelems * nk * 4, hipMemcpyHostToDevice, st));
                dni = A.stage_i;
            }
        }
        if (with_model && model) {
            dm = sm ? put(model, nm) : model;
            if (!dm) return fail(MCK_EHIP, "mck_results_add: staging copy failed");
        }
        if (with_het && hetero) {
            dh = sh ? put(hetero, nh) : hetero;
            if (!dh) return fail(MCK_EHIP, "mck_results_add: staging copy failed");
        }
        if (sct) {
            dct = put(last ? ct_last.data() : computation_time, nelems);
            if (!dct) return fail(MCK_EHIP, "mck_results_add: staging copy failed");
        } else if (computation_time) {
            dct = computation_time;
        }
    }
    {
        // grids for the largest segment count (one per pair); threads past the device
        // count return at once.  K_x and Var in one pass over the segments.
        auto add = [&](int ncol, const double *v0, double *a0, const double *v1, double *a1) {
            const bool two = ncol % 2 == 0 && ((uintptr_t)v0 % 16) == 0 && ((uintptr_t)a0 % 16) == 0 &&
                             (!v1 || (((uintptr_t)v1 % 16) == 0 && ((uintptr_t)a1 % 16) == 0));
            const long long nt = (long long)npair * (two ? ncol / 2 : ncol);
            const dim3 grid((unsigned)((nt + 255) / 256));
            if (two)
                hipLaunchKernelGGL(k_results_add<2>, grid, dim3(256), 0, st, A.nseg, ncol, Rs.npoints,
                                   A.seg_start, A.seg_count, A.seg_point, A.sl_out, v0, a0, v1, a1);
            else
                hipLaunchKernelGGL(k_results_add<1>, grid, dim3(256), 0, st, A.nseg, ncol, Rs.npoints,
                                   A.seg_start, A.seg_count, A.seg_point, A.sl_out, v0, a0, v1, a1);
        };
        add(nk, dv, Rs.K.p, dvar, Rs.V.p);
        if (with_model) add(kModelPar, dm, Rs.M.p, nullptr, nullptr);
        if (with_het) add(kHetPar, dh, Rs.H.p, nullptr, nullptr);
    }
    {
        const bool four = nk % 4 == 0 && ((uintptr_t)dni % 16) == 0;
        const long long nt3 = (long long)nelems * (four ? nk / 4 : nk);
        const dim3 grid((unsigned)((nt3 + 255) / 256));
        if (four)
            hipLaunchKernelGGL(k_results_elem<4>, grid, dim3(256), 0, st, nelems, nk, A.elem_idx, dni, dct, rank,
                               Rs.niter.p, Rs.ctime.p, Rs.proc.p);
        else
            hipLaunchKernelGGL(k_results_elem<1>, grid, dim3(256), 0, st, nelems, nk, A.elem_idx, dni, dct, rank,
                               Rs.niter.p, Rs.ctime.p, Rs.proc.p);
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
    // device-resident results are consumed: a later call must bring its own
    if (last) G.last_nelems = G.last_nbasis = 0;
    if (with_model && !model) G.last_model_nelems = G.last_model_nbasis = 0;
    if (with_het && !hetero) G.last_hetero_nelems = G.last_hetero_nbasis = 0;
    return MCK_OK;
}

}  // namespace

extern "C" {

int mck_results_add(int32_t nelems, const int32_t *elems, const double *kernel_values,
                    const double *kernel_variance, const int32_t *niterations, const double *computation_time,
                    const double *model, const double *hetero, int32_t rank) {
    return results_add_impl(nelems, elems, kernel_values, kernel_variance, niterations, computation_time, model,
                            hetero, rank, false);
}

int mck_results_add_noalloc(int32_t nelems, const int32_t *elems, const double *kernel_values,
                            const double *kernel_variance, const int32_t *niterations,
                            const double *computation_time, const double *model, const double *hetero,
                            int32_t rank) {
    return results_add_impl(nelems, elems, kernel_values, kernel_variance, niterations, computation_time, model,
                            hetero, rank, true);
}

int mck_results_get(double *kernel, double *variance, double *model, double *hetero, int32_t *niterations,
                    double *computation_time, int32_t *element_proc) {
    State &G = g;
    auto &Rs = G.res;
    if (!G.init) return fail(MCK_ENOTINIT, "mck_init has not been called");
    if (!Rs.init) return fail(MCK_ENOTINIT, "mck_results_init has not been called");
    const size_t np = Rs.npoints, ne = Rs.nelements, nk = G.nk;
    if (kernel) HIPCHECK(hipMemcpy(kernel, Rs.K.p, np * nk * 8, hipMemcpyDeviceToHost));
    if (variance) HIPCHECK(hipMemcpy(variance, Rs.V.p, np * nk * 8, hipMemcpyDeviceToHost));
    if (model && Rs.nmodel) HIPCHECK(hipMemcpy(model, Rs.M.p, np * Rs.nmodel * 8, hipMemcpyDeviceToHost));
    if (hetero && Rs.nhet) HIPCHECK(hipMemcpy(hetero, Rs.H.p, np * Rs.nhet * 8, hipMemcpyDeviceToHost));
    if (niterations) HIPCHECK(hipMemcpy(niterations, Rs.niter.p, ne * nk * 4, hipMemcpyDeviceToHost));
    if (computation_time) HIPCHECK(hipMemcpy(computation_time, Rs.ctime.p, ne * 8, hipMemcpyDeviceToHost));
    if (element_proc) HIPCHECK(hipMemcpy(element_proc, Rs.proc.p, ne * 4, hipMemcpyDeviceToHost));
    return MCK_OK;
}

int mck_results_load(const double *kernel, const double *variance, const double *model, const double *hetero,
                     const int32_t *niterations, const double *computation_time, const int32_t *element_proc) {
    State &G = g;
    auto &Rs = G.res;
    if (!G.init) return fail(MCK_ENOTINIT, "mck_init has not been called");
    if (!Rs.init) return fail(MCK_ENOTINIT, "mck_results_init has not been called");
    const size_t np = Rs.npoints, ne = Rs.nelements, nk = G.nk;
    if (kernel) HIPCHECK(hipMemcpy(Rs.K.p, kernel, np * nk * 8, hipMemcpyDefault));
    if (variance) HIPCHECK(hipMemcpy(Rs.V.p, variance, np * nk * 8, hipMemcpyDefault));
    if (model && Rs.nmodel) HIPCHECK(hipMemcpy(Rs.M.p, model, np * Rs.nmodel * 8, hipMemcpyDefault));
    if (hetero && Rs.nhet) HIPCHECK(hipMemcpy(Rs.H.p, hetero, np * Rs.nhet * 8, hipMemcpyDefault));
    if (niterations) HIPCHECK(hipMemcpy(Rs.niter.p, niterations, ne * nk * 4, hipMemcpyDefault));
    if (computation_time) HIPCHECK(hipMemcpy(Rs.ctime.p, computation_time, ne * 8, hipMemcpyDefault));
    if (element_proc) HIPCHECK(hipMemcpy(Rs.proc.p, element_proc, ne * 4, hipMemcpyDefault));
    return MCK_OK;
}

int mck_results_export(int32_t which, float *out) {
    State &G = g;
    auto &Rs = G.res;
    if (!G.init) return fail(MCK_ENOTINIT, "mck_init has not been called");
    if (!Rs.init) return fail(MCK_ENOTINIT, "mck_results_init has not been called");
    if (!out) return fail(MCK_EINVAL, "null output");
    int np = Rs.npoints, ncol = G.nk;
    const void *src = nullptr;
    switch (which) {
    case MCK_RES_KERNEL: src = Rs.K.p; break;
    case MCK_RES_VARIANCE: src = Rs.V.p; break;
    case MCK_RES_ERROR: src = Rs.V.p; break;
    case MCK_RES_MODEL:
        if (!Rs.nmodel) return fail(MCK_EINVAL, "no model integrals accumulated (nmodel = 0)");
        src = Rs.M.p;
        ncol = Rs.nmodel;
        break;
    case MCK_RES_HETERO:
        if (!Rs.nhet) return fail(MCK_EINVAL, "no heterogeneity integrals accumulated (nhet = 0)");
        src = Rs.H.p;
        ncol = Rs.nhet;
        break;
    case MCK_RES_NITER: src = Rs.niter.p; np = Rs.nelements; break;
    default: return fail(MCK_EINVAL, "unknown result variable");
    }
    const size_t n = (size_t)np * ncol;
    const bool dev_out = device_ptr(out);
    float *d = out;
    if (!dev_out) {
        HIPCHECK(Rs.stage.ensure((n * 4 + 7) / 8));
        d = (float *)Rs.stage.p;
    }
    const dim3 grid((np + 31) / 32, (ncol + 31) / 32);
    if (which == MCK_RES_ERROR)
        hipLaunchKernelGGL(k_results_out<1>, grid, dim3(256), 0, G.stream, np, ncol, src, Rs.K.p, d);
    else if (which == MCK_RES_NITER)
        hipLaunchKernelGGL(k_results_out<2>, grid, dim3(256), 0, G.stream, np, ncol, src, nullptr, d);
    else
        hipLaunchKernelGGL(k_results_out<0>, grid, dim3(256), 0, G.stream, np, ncol, src, nullptr, d);
    HIPCHECK(hipGetLastError());
    if (!dev_out) HIPCHECK(hipMemcpyAsync(out, d, n * 4, hipMemcpyDeviceToHost, G.stream));
    HIPCHECK(hipStreamSynchronize(G.stream));
    return MCK_OK;
}

int mck_get_stats(mck_stats_t *stats) {
    if (!stats) return fail(MCK_EINVAL, "null stats");
    *stats = g.stats;
    return MCK_OK;
}

void mck_finalize(void) {
    State &G = g;
    if (G.stream) (void)hipStreamSynchronize(G.stream);
    (void)fold_kernel_events(G, false);
    for (DbDev *D : {&G.fwd, &G.bwd}) {
        if (D->u_owned) (void)hipFree(D->u_owned);
        D->u_owned = nullptr;
        D->release_pages();
        D->mp_s.release(); D->mp_z.release(); D->econ.release(); D->ejac.release(); D->eltype.release();
        D->axis.release(); D->cell_start.release(); D->cell_elems.release();
    }
    G.frames.release(); G.twq.release(); G.twq_d.release(); G.H.release(); G.recs.release();
    G.recs_base.release(); G.kerns_base.release();
    G.worktab.release(); G.Uspec.release(); G.Hp.release();
    G.kerns.release(); G.model.release(); G.d_k_rec.release(); G.d_k_nbk.release();
    G.d_rec_first.release(); G.d_rec_count.release(); G.d_rec_r.release(); G.tet.release(); G.volume.release();
    G.allowed_var.release(); G.hbase.release(); G.masked.release(); G.live.release();
    G.working.release(); G.active.release(); G.work_list.release(); G.nwork.release();
    G.fsum.release(); G.f2sum.release(); G.integral.release(); G.variance.release();
    G.conv.release(); G.done.release(); G.niter.release(); G.draws.release(); G.steps.release();
    G.nmodels.release(); G.xyz.release(); G.model_idx.release(); G.loc.release();
    G.Fspec.release(); G.Ftr.release(); G.phys.release(); G.evals.release(); G.err.release();
    G.opWf.release(); G.opQf.release(); G.opWb.release(); G.opQb.release(); G.page_bounds.release();
    G.nlive.release();
    G.abinv.release(); G.bweights.release();
    G.elem_count.release(); G.keys_in.release(); G.keys_out.release();
    G.vals_in.release(); G.vals_out.release(); G.pair_count.release(); G.sort_tmp.release();
    G.ad_W.release(); G.ad_Hg.release(); G.ad_Fs.release(); G.ad_U.release();
    G.ad_Bs.release(); G.ad_dFs.release(); G.ad_Ups.release();
    G.ad_part.release(); G.ad_hfilt.release();
    G.last_model.release(); G.last_hetero.release();
    G.int_geom.release(); G.int_abinv.release(); G.int_base.release(); G.int_live.release(); G.int_nsteps.release();
    G.het_xyz.release(); G.het_dln.release(); G.het_idx.release(); G.het_cells.release();
    G.het_ready = false;
    G.last_nelems = G.last_nbasis = G.last_model_nelems = G.last_model_nbasis = 0;
    G.last_hetero_nelems = G.last_hetero_nbasis = 0;
    G.res.release();
    G.init = false;
}

const char *mck_last_error(void) { return g_err.c_str(); }

}  // extern "C"
