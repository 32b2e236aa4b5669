/*
 * mck_oracle_path.c - FP64 CPU restatement of load_fw_points / load_bw_points, the
 * kernel set-up (filters, seismogram windows) and slave_work's Monte Carlo loop.
 * TEST INFRASTRUCTURE: the checker for the HIP implementation (see mck_oracle.h).
 */
#include "mck_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define DEG2RAD (ORA_PI / 180.0)

typedef struct { double re, im; } cz;
static cz zmul(cz a, cz b) { cz r; r.re = a.re * b.re - a.im * b.im; r.im = a.re * b.im + a.im * b.re; return r; }
static cz zdiv(cz a, cz b)
{
    cz r;
    if (fabs(b.re) >= fabs(b.im)) {
        double ratio = b.im / b.re, den = b.re + b.im * ratio;
        r.re = (a.re + a.im * ratio) / den; r.im = (a.im - a.re * ratio) / den;
    } else {
        double ratio = b.re / b.im, den = b.im + b.re * ratio;
        r.re = (a.re * ratio + a.im) / den; r.im = (a.im * ratio - a.re) / den;
    }
    return r;
}

typedef struct {
    int nd_needed;
    int needs_bk[6];
    int filter, misfit, param, rec;
    double tw[2];
    double normalization;
    int ifirst, ncut, M;
    double *seis_cut;      /* [ncut] */
    double *seis_cut_fd;   /* [(M/2+1)][2] */
} ora_kernel;

typedef struct {
    double lat, lon, colat, r[3];
    int first, last;
    int needs_bk[6];
    double t_last_p, t_last_s;
} ora_rec;

struct ora_ctx {
    mck_setup_t S;
    int N, nomega;
    double df;
    int full, nd;
    double gll[8], glj[8];
    double *tf_fwd, *tf_bwd;   /* [nfilter][nomega][2] */
    double *tf_g;              /* [nfilter][nomega][2] filter responses alone */
    double *rest_f, *rest_b;   /* [nomega][2] TF_fwd = G rest_f, TF_bwd = G rest_b */
    double *ts_fwd, *ts_bwd;   /* [nomega][2] time shift operators (deconv_stf == 0) */
    ora_kernel *K;
    ora_rec *R;
    double src_lat, src_lon, src_colat, src_r[3];
    double *seis_own;          /* [nrec][nt] seismograms by reciprocity when S.seis is NULL */
    int64_t halton_step;
    ora_fill_fn fill;
    void *fill_user;
    unsigned char *filled[2];
};

void ora_set_lazy_fill(ora_ctx *C, ora_fill_fn fn, void *user)
{
    C->fill = fn;
    C->fill_user = user;
    free(C->filled[0]); free(C->filled[1]);
    C->filled[0] = (unsigned char *)calloc(C->S.fwd.nelem, 1);
    C->filled[1] = (unsigned char *)calloc(C->S.bwd.nelem, 1);
}

static int param_needs(int param, int *nb)
{
    /* tabulate_kernels (kernel.f90:576-628); nb = lambda, mu, rho, a, b, c */
    memset(nb, 0, 6 * sizeof(int));
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

/* Test hook: tabulate_kernels alone */
int ora_param_needs(int param, int32_t *nb)
{
    int b, v[6], rc = param_needs(param, v);
    for (b = 0; b < 6; ++b) nb[b] = v[b];
    return rc;
}

/* apply time shift exp(2 pi i f tau) (filtering.f90:622-652) */
static void make_timeshift(int nomega, double df, double tau, double *ts)
{
    int k;
    for (k = 0; k < nomega; ++k) {
        double th = 2 * ORA_PI * (df * k) * tau;
        ts[2 * k] = cos(th);
        ts[2 * k + 1] = sin(th);
    }
}

/* add_stfs (filtering.f90:241-470) on one transfer function tff[nomega][2] (in: the
 * filter's response; out: the forward one) and tfb (out: the backward one): divide by the
 * spectrum of the database STF (tapered, ntaper 2), scale by the source magnitude, apply
 * the earthquake STF (tapered at its end only; NULL = a unit impulse at t = 0) and its
 * time shift to the forward one, then the squared 12th-order Butterworth low pass at
 * 0.75 x the highest frequency to both; NaNs become 0.  Returns -1 (with the message)
 * when either response at the highest frequency exceeds 1e-2 x its maximum, where the
 * reference stops. */
int ora_add_stfs(int nt, double dt, int N, const double *stf_sem, double amplitude, const double *stf_src,
                 double stf_shift, double *tff, double *tfb, char *err, int errlen)
{
    int nw = N / 2 + 1, k;
    double df = 1.0 / (dt * N);
    double *x = (double *)calloc(N, sizeof(double)), *pad = (double *)calloc(N, sizeof(double));
    double *Fs = (double *)malloc(2 * nw * sizeof(double)), *Fe = (double *)malloc(2 * nw * sizeof(double));
    double *f = (double *)malloc(nw * sizeof(double));
    mck_filter_t lpf;
    double *lp = (double *)malloc(2 * nw * sizeof(double));
    double maxf = 0, maxb = 0, rf, rb;
    for (k = 0; k < nw; ++k) f[k] = df * k;
    ora_taperandzeropad(nt, N, 2, 0, stf_sem, pad);
    ora_rfft(N, dt, pad, Fs);
    for (k = 0; k < nt; ++k) x[k] = stf_src ? stf_src[k] : (k == 0 ? 1.0 : 0.0);
    ora_taperandzeropad(nt, N, 2, 1, x, pad);
    ora_rfft(N, dt, pad, Fe);
    lpf.filterclass = MCK_FILTER_BUTTER_LP;
    lpf.frequencies[0] = 1.0 / (f[nw - 1] * 0.75);
    lpf.frequencies[1] = 0; lpf.frequencies[2] = 12; lpf.frequencies[3] = 0;
    ora_filter_tf(&lpf, nw, df, lp);
    for (k = 0; k < nw; ++k) {
        cz t = {tff[2 * k], tff[2 * k + 1]}, s = {Fs[2 * k], Fs[2 * k + 1]};
        cz e = {Fe[2 * k], Fe[2 * k + 1]}, l = {lp[2 * k], lp[2 * k + 1]}, lc = {lp[2 * k], -lp[2 * k + 1]};
        cz fw, bw, sh;
        double th = 2 * ORA_PI * f[k] * stf_shift;
        sh.re = cos(th); sh.im = sin(th);
        l = zmul(l, lc);
        fw = zdiv(t, s); fw.re *= amplitude; fw.im *= amplitude;
        bw = zdiv(t, s); bw.re *= amplitude; bw.im *= amplitude;
        fw = zmul(fw, e);
        fw = zmul(fw, sh);
        fw = zmul(fw, l);
        bw = zmul(bw, l);
        if (isnan(hypot(fw.re, fw.im))) { fw.re = 0; fw.im = 0; }
        if (isnan(hypot(bw.re, bw.im))) { bw.re = 0; bw.im = 0; }
        tff[2 * k] = fw.re; tff[2 * k + 1] = fw.im;
        tfb[2 * k] = bw.re; tfb[2 * k + 1] = bw.im;
        if (hypot(fw.re, fw.im) > maxf) maxf = hypot(fw.re, fw.im);
        if (hypot(bw.re, bw.im) > maxb) maxb = hypot(bw.re, bw.im);
    }
    rf = hypot(tff[2 * (nw - 1)], tff[2 * (nw - 1) + 1]) / maxf;
    rb = hypot(tfb[2 * (nw - 1)], tfb[2 * (nw - 1) + 1]) / maxb;
    free(x); free(pad); free(Fs); free(Fe); free(f); free(lp);
    if (rf > 1e-2 || rb > 1e-2) {
        if (err) snprintf(err, errlen, "filter not vanishing at the highest frequency (%g, %g)", rf, rb);
        return -1;
    }
    return 0;
}

static int add_stfs(ora_ctx *C, double *tff, double *tfb, char *err, int errlen)
{
    const mck_setup_t *S = &C->S;
    return ora_add_stfs(S->nt, S->dt, C->N, S->stf_fwd, S->fwd.amplitude, S->stf_src, S->stf_src_shift, tff, tfb,
                        err, errlen);
}

/* integrate_parseval (kernel.f90:975-1020): Re[(ab_1 + ab_n + 2 sum ab_mid) df] */
static double parseval(int M, double dt, const double *af, const double *bf)
{
    int nw = M / 2 + 1, k;
    double df = 1. / (dt * M);
    cz s = {0, 0}, a0, an, tot;
    for (k = 1; k < nw - 1; ++k) {
        cz a = {af[2 * k], af[2 * k + 1]}, b = {bf[2 * k], -bf[2 * k + 1]};
        cz p = zmul(a, b);
        s.re += p.re; s.im += p.im;
    }
    { cz a = {af[0], af[1]}, b = {bf[0], -bf[1]}; a0 = zmul(a, b); }
    { cz a = {af[2 * (nw - 1)], af[2 * (nw - 1) + 1]}, b = {bf[2 * (nw - 1)], -bf[2 * (nw - 1) + 1]}; an = zmul(a, b); }
    tot.re = a0.re + an.re + 2 * s.re;
    tot.im = a0.im + an.im + 2 * s.im;
    return tot.re * df;
}

/* Test hook: integrate_parseval of two time series of n samples, zero padded to
 * M = 2 nextpow2(n) as the kernel's own FFT plan is (kernel.f90:975-1020). */
double ora_integrate_parseval(int n, double dt, const double *f, const double *g)
{
    int M = 2 * ora_nextpow2(n), i;
    double *pf = (double *)calloc(2 * (size_t)M, sizeof(double)), *pg = pf + M;
    double *F = (double *)malloc(4 * (size_t)(M / 2 + 1) * sizeof(double)), *G = F + 2 * (M / 2 + 1);
    double v;
    for (i = 0; i < n; ++i) { pf[i] = f[i]; pg[i] = g[i]; }
    ora_rfft(M, dt, pf, F);
    ora_rfft(M, dt, pg, G);
    v = parseval(M, dt, F, G);
    free(pf); free(F);
    return v;
}

/* cut_timewindow (kernel.f90:836-927): samples with w0 <= t_i <= w1 of a monotonic time
 * axis t[n]; first index and count (count 0 = empty window) */
void ora_cut_window(int n, const double *t, double w0, double w1, int *ifirst, int *ncut)
{
    int i;
    *ifirst = -1; *ncut = 0;
    for (i = 0; i < n; ++i) {
        if (t[i] <= w1 && t[i] >= w0) { if (*ifirst < 0) *ifirst = i; (*ncut)++; }
        if (t[i] > w1) break;
    }
}

/* cut_and_add_seismogram (kernel.f90:181-385) */
static int kernel_init(ora_ctx *C, int k, char *err, int errlen)
{
    const mck_setup_t *S = &C->S;
    ora_kernel *K = &C->K[k];
    int N = C->N, nw = C->nomega, i, r = S->kernel_receiver[k];
    const double *seis = S->seis + (size_t)r * S->nt;
    double *disp = (double *)calloc(N, sizeof(double)), *velo = (double *)calloc(N, sizeof(double));
    double *Xd = (double *)malloc(2 * nw * sizeof(double)), *Xv = (double *)malloc(2 * nw * sizeof(double));
    double *d = (double *)malloc(N * sizeof(double)), *v = (double *)malloc(N * sizeof(double));
    const double *tf = C->tf_fwd + (size_t)K->filter * 2 * nw;
    double *pad, *cutfd, term;
    ora_cumsum_trapezoidal(S->nt, S->dt, seis, disp);
    for (i = 0; i < S->nt; ++i) velo[i] = seis[i];
    ora_rfft(N, S->dt, disp, Xd);
    ora_rfft(N, S->dt, velo, Xv);
    for (i = 0; i < nw; ++i) {
        cz a = {Xd[2 * i], Xd[2 * i + 1]}, b = {Xv[2 * i], Xv[2 * i + 1]}, t = {tf[2 * i], tf[2 * i + 1]};
        if (!S->deconv_stf) {
            cz sh = {C->ts_fwd[2 * i], C->ts_fwd[2 * i + 1]};
            a = zmul(a, sh); b = zmul(b, sh);
        }
        a = zmul(a, t); b = zmul(b, t);
        Xd[2 * i] = a.re; Xd[2 * i + 1] = a.im; Xv[2 * i] = b.re; Xv[2 * i + 1] = b.im;
    }
    ora_irfft(N, C->df, Xd, d);
    ora_irfft(N, C->df, Xv, v);
    /* cut_timewindow (kernel.f90:836-927) on t_i = (i-1) dt */
    if (K->tw[0] < 0.0 || K->tw[1] > (N - 1) * S->dt || K->tw[1] - K->tw[0] <= 0) {
        snprintf(err, errlen, "kernel %d: invalid time window", k);
        free(disp); free(velo); free(Xd); free(Xv); free(d); free(v);
        return -1;
    }
    {
        double *tt = (double *)malloc(N * sizeof(double));
        for (i = 0; i < N; ++i) tt[i] = i * S->dt;
        ora_cut_window(N, tt, K->tw[0], K->tw[1], &K->ifirst, &K->ncut);
        free(tt);
    }
    if (K->ncut == 0) {
        snprintf(err, errlen, "kernel %d: time window length was zero", k);
        free(disp); free(velo); free(Xd); free(Xv); free(d); free(v);
        return -1;
    }
    K->seis_cut = (double *)malloc(K->ncut * sizeof(double));
    for (i = 0; i < K->ncut; ++i)
        K->seis_cut[i] = (K->misfit == MCK_MISFIT_CC ? v : d)[K->ifirst + i];
    K->M = 2 * ora_nextpow2(K->ncut);
    pad = (double *)calloc(K->M, sizeof(double));
    for (i = 0; i < K->ncut; ++i) pad[i] = K->seis_cut[i];
    K->seis_cut_fd = (double *)malloc(2 * (K->M / 2 + 1) * sizeof(double));
    ora_rfft(K->M, S->dt, pad, K->seis_cut_fd);
    cutfd = (double *)malloc(2 * (K->M / 2 + 1) * sizeof(double));
    ora_rfft(K->M, S->dt, pad, cutfd);
    term = parseval(K->M, S->dt, cutfd, cutfd);
    K->normalization = (term > 1e-100) ? 1.0 / term : 0.0;
    free(pad); free(cutfd);
    free(disp); free(velo); free(Xd); free(Xv); free(d); free(v);
    return 0;
}

ora_ctx *ora_create(const mck_setup_t *S, char *err, int errlen)
{
    ora_ctx *C = (ora_ctx *)calloc(1, sizeof(ora_ctx));
    int k, r, f;
    C->S = *S;
    C->N = 2 * ora_nextpow2(S->nt);
    C->nomega = C->N / 2 + 1;
    C->df = 1. / (S->dt * C->N);
    ora_gll_points(4, C->gll);
    ora_glj_points(4, C->glj);
    if (S->fwd.npol != 4 || S->bwd.npol != 4) {
        snprintf(err, errlen, "only npol = 4 is supported"); free(C); return NULL;
    }
    /* strain type (type_parameter.f90:477-486) */
    C->full = 0;
    for (k = 0; k < S->nkernel; ++k) {
        int p = S->kernel_param[k];
        if (p == MCK_PARAM_RHO) { snprintf(err, errlen, "rho kernels are not implemented (kernel.f90:554)"); free(C); return NULL; }
        if (!(p == MCK_PARAM_LAM || p == MCK_PARAM_VP || p == MCK_PARAM_VPH || p == MCK_PARAM_VPV)) C->full = 1;
    }
    C->nd = C->full ? 6 : 1;
    C->src_lat = S->src_lat * DEG2RAD;
    C->src_lon = S->src_lon * DEG2RAD;
    C->src_colat = (90 - S->src_lat) * DEG2RAD;
    {
        double rad = S->planet_radius - S->src_depth_km * 1e3;
        C->src_r[0] = cos(C->src_lat) * cos(C->src_lon) * rad;
        C->src_r[1] = cos(C->src_lat) * sin(C->src_lon) * rad;
        C->src_r[2] = sin(C->src_lat) * rad;
    }
    C->R = (ora_rec *)calloc(S->nrec, sizeof(ora_rec));
    for (r = 0; r < S->nrec; ++r) {
        ora_rec *R = &C->R[r];
        R->lat = S->rec_lat[r] * DEG2RAD;
        R->lon = S->rec_lon[r] * DEG2RAD;
        R->colat = (90 - S->rec_lat[r]) * DEG2RAD;
        R->r[0] = cos(R->lat) * cos(R->lon) * S->planet_radius;
        R->r[1] = cos(R->lat) * sin(R->lon) * S->planet_radius;
        R->r[2] = sin(R->lat) * S->planet_radius;
        R->first = -1; R->last = -2;
    }
    /* filters (filtering.f90:78-230) and STF deconvolution (add_stfs) */
    C->tf_fwd = (double *)malloc((size_t)S->nfilter * 2 * C->nomega * sizeof(double));
    C->tf_bwd = (double *)malloc((size_t)S->nfilter * 2 * C->nomega * sizeof(double));
    C->tf_g = (double *)malloc((size_t)S->nfilter * 2 * C->nomega * sizeof(double));
    for (f = 0; f < S->nfilter; ++f) {
        double *a = C->tf_fwd + (size_t)f * 2 * C->nomega, *b = C->tf_bwd + (size_t)f * 2 * C->nomega;
        ora_filter_tf(&S->filters[f], C->nomega, C->df, a);
        memcpy(C->tf_g + (size_t)f * 2 * C->nomega, a, 2 * C->nomega * sizeof(double));
        memcpy(b, a, 2 * C->nomega * sizeof(double));
        if (S->deconv_stf && add_stfs(C, a, b, err, errlen)) { ora_destroy(C); return NULL; }
    }
    /* the deconvolution factors alone: add_stfs of a unit response (its high-frequency
     * check does not apply to them) */
    C->rest_f = (double *)malloc(2 * C->nomega * sizeof(double));
    C->rest_b = (double *)malloc(2 * C->nomega * sizeof(double));
    for (k = 0; k < C->nomega; ++k) {
        C->rest_f[2 * k] = C->rest_b[2 * k] = 1.0;
        C->rest_f[2 * k + 1] = C->rest_b[2 * k + 1] = 0.0;
    }
    if (S->deconv_stf) {
        char e2[256];
        add_stfs(C, C->rest_f, C->rest_b, e2, sizeof e2);
    }
    C->ts_fwd = (double *)malloc(2 * C->nomega * sizeof(double));
    C->ts_bwd = (double *)malloc(2 * C->nomega * sizeof(double));
    make_timeshift(C->nomega, C->df, S->timeshift_fwd, C->ts_fwd);
    make_timeshift(C->nomega, C->df, S->timeshift_bwd, C->ts_bwd);
    /* seismograms: the caller's, or by reciprocity from the bwd database (load_seismogram,
     * slave.f90:94-173 calls it before the kernels are set up) */
    if (!S->seis) {
        C->seis_own = (double *)malloc((size_t)S->nrec * S->nt * sizeof(double));
        if (ora_seismograms(C, C->seis_own)) {
            snprintf(err, errlen, "source point not found in the mesh for the reciprocal seismograms");
            ora_destroy(C); return NULL;
        }
        C->S.seis = C->seis_own;
    }
    /* kernels (type_parameter.f90:521-691) */
    C->K = (ora_kernel *)calloc(S->nkernel, sizeof(ora_kernel));
    for (k = 0; k < S->nkernel; ++k) {
        ora_kernel *K = &C->K[k];
        ora_rec *R;
        int b;
        K->rec = S->kernel_receiver[k];
        K->filter = S->kernel_filter[k];
        K->misfit = S->kernel_misfit[k];
        K->param = S->kernel_param[k];
        K->tw[0] = S->kernel_window[2 * k];
        K->tw[1] = S->kernel_window[2 * k + 1];
        if (param_needs(K->param, K->needs_bk)) { snprintf(err, errlen, "bad model parameter"); ora_destroy(C); return NULL; }
        if (!C->full && (K->needs_bk[1] || K->needs_bk[3] || K->needs_bk[4] || K->needs_bk[5])) {
            snprintf(err, errlen, "kernel %d needs the full strain tensor but only the trace is loaded", k);
            ora_destroy(C); return NULL;
        }
        R = &C->R[K->rec];
        if (R->first < 0) R->first = k;
        R->last = k;
        for (b = 0; b < 6; ++b) R->needs_bk[b] |= K->needs_bk[b];
        if (K->param == MCK_PARAM_VP || K->param == MCK_PARAM_VPH || K->param == MCK_PARAM_VPV || K->param == MCK_PARAM_LAM)
            R->t_last_p = fmax(R->t_last_p, K->tw[1]);
        else
            R->t_last_s = fmax(R->t_last_s, K->tw[1]);
        if (kernel_init(C, k, err, errlen)) { ora_destroy(C); return NULL; }
    }
    C->halton_step = 0;
    return C;
}

void ora_destroy(ora_ctx *C)
{
    int k;
    if (!C) return;
    if (C->K) for (k = 0; k < C->S.nkernel; ++k) { free(C->K[k].seis_cut); free(C->K[k].seis_cut_fd); }
    free(C->seis_own);
    free(C->K); free(C->R); free(C->tf_fwd); free(C->tf_bwd); free(C->ts_fwd); free(C->ts_bwd);
    free(C->tf_g); free(C->rest_f); free(C->rest_b);
    free(C->filled[0]); free(C->filled[1]);
    free(C);
}

int ora_strain_nd(ora_ctx *C) { return C->nd; }

int ora_kernel_info(ora_ctx *C, int k, double *normalization, int *ncut, int *ifirst)
{
    *normalization = C->K[k].normalization;
    *ncut = C->K[k].ncut;
    *ifirst = C->K[k].ifirst;
    return 0;
}

/* ------------------------------------------------------------------------------------
 * find_element (readfields.f90:1052-1134)
 * ---------------------------------------------------------------------------------- */
static int find_element(const mck_db_t *db, double s, double z, double *xi, double *eta,
                        double *nodes, int *id_elem, int *tie)
{
    int32_t idx[12];
    float dist[12];
    int c, i;
    *tie = ora_nearest(db->nelem, db->mp_s, db->mp_z, s, z, 12, idx, dist);
    for (c = 0; c < 12; ++c) {
        int e = idx[c];
        for (i = 0; i < 4; ++i) {
            int p = db->fem_mesh[4 * e + i];
            nodes[i] = (double)db->mesh_s[p];
            nodes[4 + i] = (double)db->mesh_z[p];
        }
        if (ora_inside_element(s, z, nodes, db->eltype[e], 1e-3, xi, eta)) {
            *id_elem = e;
            return 0;
        }
    }
    return -1;
}

static void frame_of(ora_ctx *C, int frame, double *lon, double *colat)
{
    if (frame == 0) { *lon = C->src_lon; *colat = C->src_colat; }
    else { *lon = C->R[frame - 1].lon; *colat = C->R[frame - 1].colat; }
}

int ora_locate(ora_ctx *C, int frame, int npts, const double *xyz, int32_t *elem,
               double *xi, double *eta, int32_t *ties)
{
    const mck_db_t *db = frame == 0 ? &C->S.fwd : &C->S.bwd;
    double lon, colat, R = C->S.planet_radius;
    int p;
    frame_of(C, frame, &lon, &colat);
    for (p = 0; p < npts; ++p) {
        double s, phi, z, d2, nodes[8];
        int id, tie;
        ora_rotate_frame(1, xyz + 3 * p, lon, colat, &s, &phi, &z);
        d2 = s * s + z * z;
        if (d2 > R * R) { s = s / sqrt(d2) * R; z = z / sqrt(d2) * R; }
        if (find_element(db, s, z, &xi[p], &eta[p], nodes, &id, &tie)) return MCK_ENOTFOUND;
        elem[p] = id + 1;
        if (ties) ties[p] = tie;
    }
    return 0;
}

/* azim_factor (rotations.f90:89-131) */
static double azim_factor(double phi, const double *mij, int isim, int ikind)
{
    switch (isim) {
    case 1: return ikind == 1 ? mij[0] : 0.0;
    case 2: return ikind == 1 ? 0.5 * (mij[1] + mij[2]) : 0.0;
    case 3: return ikind == 1 ? mij[3] * cos(phi) + mij[4] * sin(phi)
                              : -mij[3] * sin(phi) + mij[4] * cos(phi);
    default: return ikind == 1 ? 0.5 * (mij[1] - mij[2]) * cos(2.0 * phi) + mij[5] * sin(2.0 * phi)
                               : -0.5 * (mij[1] - mij[2]) * sin(2.0 * phi) + mij[5] * cos(2.0 * phi);
    }
}

/* Test hook: azim_factor alone */
double ora_azim_factor(double phi, const double *mij, int isim, int ikind)
{
    return azim_factor(phi, mij, isim, ikind);
}

/* azim_factor_bw (rotations.f90:197-227) */
static double azim_factor_bw(double phi, const double *fi, int isim, int ikind)
{
    if (isim == 1) return ikind == 1 ? fi[0] : 0.0;
    return ikind == 1 ? fi[1] * cos(phi) + fi[2] * sin(phi) : -fi[1] * sin(phi) + fi[2] * cos(phi);
}

/* rotate_symm_tensor_voigt_src_to_xyz_2d (rotations.f90:382-419) */
static void rot_src_to_xyz(double *t, int nt, double phi)
{
    double sp = sin(phi), cp = cos(phi);
    int i;
    for (i = 0; i < nt; ++i) {
        double a1 = t[0 * nt + i], a2 = t[1 * nt + i], a3 = t[2 * nt + i];
        double a4 = t[3 * nt + i], a5 = t[4 * nt + i], a6 = t[5 * nt + i];
        t[0 * nt + i] = a1 * (cp * cp) + sp * (-2 * a6 * cp + a2 * sp);
        t[1 * nt + i] = a2 * (cp * cp) + sp * (2 * a6 * cp + a1 * sp);
        t[2 * nt + i] = a3;
        t[3 * nt + i] = a4 * cp + a5 * sp;
        t[4 * nt + i] = a5 * cp - a4 * sp;
        t[5 * nt + i] = (a1 - a2) * cp * sp + a6 * (cp * cp - sp * sp);
    }
}

/* rotate_symm_tensor_voigt_xyz_src_to_xyz_earth_2d (rotations.f90:555-612) */
static void rot_xyz_src_to_earth(double *t, int nt, double phi, double theta)
{
    double sp = sin(phi), cp = cos(phi), st = sin(theta), ct = cos(theta);
    int i;
    for (i = 0; i < nt; ++i) {
        double a1 = t[0 * nt + i], a2 = t[1 * nt + i], a3 = t[2 * nt + i];
        double a4 = t[3 * nt + i], a5 = t[4 * nt + i], a6 = t[5 * nt + i];
        t[0 * nt + i] = a1 * (cp * cp) * (ct * ct) - 2 * a6 * cp * ct * sp + a2 * (sp * sp)
                      + 2 * a5 * (cp * cp) * ct * st - 2 * a4 * cp * sp * st
                      + a3 * (cp * cp) * (st * st);
        t[1 * nt + i] = a2 * (cp * cp)
                      + sp * (2 * a6 * cp * ct + a1 * (ct * ct) * sp
                              + st * (2 * a4 * cp + 2 * a5 * ct * sp + a3 * sp * st));
        t[2 * nt + i] = a3 * (ct * ct) + st * (-2 * a5 * ct + a1 * st);
        t[3 * nt + i] = ct * (a4 * cp + a5 * ct * sp + a3 * sp * st)
                      - st * (a6 * cp + a1 * ct * sp + a5 * sp * st);
        t[4 * nt + i] = ct * (a5 * cp * ct - a4 * sp + a3 * cp * st)
                      - st * (a1 * cp * ct - a6 * sp + a5 * cp * st);
        t[5 * nt + i] = cp * st * (a4 * cp + a5 * ct * sp + a3 * sp * st)
                      - sp * (a2 * cp + a6 * ct * sp + a4 * sp * st)
                      + cp * ct * (a6 * cp + a1 * ct * sp + a5 * sp * st);
    }
}

/* lagrange_interpol_2D_td (spectral_basis.f90:181-314) of node values
 * val[c][jpol][ipol][nt] -> out[c][nt] */
/* Test hook: which 0 = src->xyz (azimuth phi), 1 = xyz_src->xyz_earth (phi, theta);
 * t[6][nt] Voigt (xx, yy, zz, yz, xz, xy), rotated in place. */
void ora_rotate_voigt(int which, double *t, int nt, double phi, double theta)
{
    if (which == 0) rot_src_to_xyz(t, nt, phi);
    else rot_xyz_src_to_earth(t, nt, phi, theta);
}

static void interp_nodes(int nc, int nt, const double *val, const double *xp, const double *ep,
                         double x1, double x2, double *out)
{
    double li[5], lj[5];
    int c, i, j, t;
    ora_lagrange_basis(4, xp, x1, li);
    ora_lagrange_basis(4, ep, x2, lj);
    for (c = 0; c < nc; ++c) {
        double *o = out + (size_t)c * nt;
        for (t = 0; t < nt; ++t) o[t] = 0.0;
        for (i = 0; i <= 4; ++i)
            for (j = 0; j <= 4; ++j) {
                const double *v = val + (((size_t)c * 5 + j) * 5 + i) * nt;
                for (t = 0; t < nt; ++t) o[t] = o[t] + v[t] * li[i] * lj[j];
            }
    }
}

/* Strain at (xi, eta) of one element for simulation `sim` (1-based within the DB):
 * load_strain_point_merged_fwd/bwd (readfields.f90:2009-2403); out[nd][nt]. */
static void strain_point(ora_ctx *C, const mck_db_t *db, int e, int pole, int var0,
                         double xi, double eta, const double *nodes, double *out)
{
    int nt = C->S.nt, nd = C->nd, v, k, axial = db->axis[e];
    size_t blk = 25 * (size_t)nt;
    double *u = (double *)malloc(3 * blk * sizeof(double));
    double *st = (double *)malloc(6 * blk * sizeof(double));
    int nv = pole == 0 ? 2 : 3;
    const float *src = db->snapshots + ((size_t)e * db->nvar + var0) * blk;
    if (C->fill) {
        int which = (db == &C->S.fwd) ? 0 : 1;
        if (!C->filled[which][e]) {
            C->fill(which, e, (float *)(db->snapshots + (size_t)e * db->nvar * blk), C->fill_user);
            C->filled[which][e] = 1;
        }
    }
    for (v = 0; v < nv; ++v)
        for (k = 0; k < (int)blk; ++k) u[v * blk + k] = (double)src[v * blk + k];
    if (C->full) ora_strain_pole(pole, u, nt, 4, nodes, db->eltype[e], axial, st);
    else ora_straintrace_pole(pole, u, nt, 4, nodes, db->eltype[e], axial, st);
    interp_nodes(nd, nt, st, axial ? C->glj : C->gll, C->gll, xi, eta, out);
    if (C->full) {
        for (k = 0; k < nt; ++k) { out[3 * nt + k] = -out[3 * nt + k]; out[5 * nt + k] = -out[5 * nt + k]; }
    }
    free(u); free(st);
}

/* load_fw_points (readfields.f90:1138-1308) / load_bw_points (readfields.f90:1312-1465);
 * out[npts][nd][nt] */
static int load_points(ora_ctx *C, int frame, int npts, const double *xyz, double *out)
{
    const mck_setup_t *S = &C->S;
    const mck_db_t *db = frame == 0 ? &S->fwd : &S->bwd;
    int nt = S->nt, nd = C->nd, p, c, k;
    double lon, colat, R = S->planet_radius;
    double *ut = (double *)malloc((size_t)nd * nt * sizeof(double));
    frame_of(C, frame, &lon, &colat);
    for (p = 0; p < npts; ++p) {
        double s, phi, z, d2, xi, eta, nodes[8];
        int e, tie;
        double *o = out + (size_t)p * nd * nt;
        ora_rotate_frame(1, xyz + 3 * p, lon, colat, &s, &phi, &z);
        d2 = s * s + z * z;
        if (d2 > R * R) { s = s / sqrt(d2) * R; z = z / sqrt(d2) * R; }
        if (find_element(db, s, z, &xi, &eta, nodes, &e, &tie)) { free(ut); return MCK_ENOTFOUND; }
        for (k = 0; k < nd * nt; ++k) o[k] = 0.0;
        if (frame == 0) {
            int isim;
            static const int pole_of[4] = {0, 0, 1, 2}, var_of[4] = {0, 2, 4, 7};
            for (isim = 1; isim <= 4; ++isim) {
                double a1 = azim_factor(phi, S->src_mij, isim, 1), a2 = azim_factor(phi, S->src_mij, isim, 2);
                strain_point(C, db, e, pole_of[isim - 1], var_of[isim - 1], xi, eta, nodes, ut);
                for (k = 0; k < nd * nt; ++k) ut[k] = ut[k] / db->amplitude;
                if (!C->full) {
                    for (k = 0; k < nt; ++k) {
                        double v = ut[k];
                        if (v != v) v = 0.0;
                        o[k] = o[k] + v * a1;
                    }
                } else {
                    for (c = 0; c < 6; ++c) {
                        double a = (c == 3 || c == 5) ? a2 : a1;
                        for (k = 0; k < nt; ++k) o[c * nt + k] = o[c * nt + k] + ut[c * nt + k] * a;
                    }
                }
            }
            if (C->full) {
                rot_src_to_xyz(o, nt, phi);
                rot_xyz_src_to_earth(o, nt, C->src_lon, C->src_colat);
            }
        } else {
            double az[6];
            int isim;
            if (S->component == 'Z') {
                isim = 1;
                for (c = 0; c < 6; ++c) az[c] = 1.0;
            } else {
                double fi[3] = {0, 0, 0}, a1, a2;
                fi[S->component == 'R' ? 1 : 2] = 1.0;
                isim = 2;
                a1 = azim_factor_bw(phi, fi, 2, 1);
                a2 = azim_factor_bw(phi, fi, 2, 2);
                az[0] = a1; az[1] = a1; az[2] = a1; az[3] = a2; az[4] = a1 * 2; az[5] = a2 * 2;
            }
            strain_point(C, db, e, isim == 1 ? 0 : 1, isim == 1 ? 0 : 2, xi, eta, nodes, ut);
            for (c = 0; c < nd; ++c)
                for (k = 0; k < nt; ++k) o[c * nt + k] = ut[c * nt + k] * az[c];
            for (k = 0; k < nd * nt; ++k) o[k] = o[k] / db->amplitude;
            if (C->full) {
                rot_src_to_xyz(o, nt, phi);
                rot_xyz_src_to_earth(o, nt, C->R[frame - 1].lon, C->R[frame - 1].colat);
            }
        }
    }
    free(ut);
    return 0;
}

int ora_load_strain(ora_ctx *C, int frame, int npts, const double *xyz, double *traces)
{
    return load_points(C, frame, npts, xyz, traces);
}

/* rotate_symm_tensor_voigt_xyz_earth_to_xyz_src_1d (rotations.f90:615-671) */
static void rot_earth_to_xyz_src(double *t, double phi, double theta)
{
    double sp = sin(phi), cp = cos(phi), st = sin(theta), ct = cos(theta);
    double a1 = t[0], a2 = t[1], a3 = t[2], a4 = t[3], a5 = t[4], a6 = t[5];
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

/* rotate_symm_tensor_voigt_xyz_to_src_1d (rotations.f90:423-455) */
static void rot_xyz_to_src(double *t, double phi)
{
    double sp = sin(phi), cp = cos(phi);
    double a1 = t[0], a2 = t[1], a3 = t[2], a4 = t[3], a5 = t[4], a6 = t[5];
    t[0] = a1 * (cp * cp) + sp * (2 * a6 * cp + a2 * sp);
    t[1] = a2 * (cp * cp) + sp * (-2 * a6 * cp + a1 * sp);
    t[2] = a3;
    t[3] = a4 * cp - a5 * sp;
    t[4] = a5 * cp + a4 * sp;
    t[5] = cp * (a6 * cp + a2 * sp) - sp * (a1 * cp + a6 * sp);
}

/* Test hook: 2 = xyz_earth -> xyz_src (phi, theta), 3 = xyz -> src (phi); t[6] */
void ora_rotate_voigt_inv(int which, double *t, double phi, double theta)
{
    if (which == 2) rot_earth_to_xyz_src(t, phi, theta);
    else rot_xyz_to_src(t, phi);
}

/* The rotated moment tensor of receiver r's reciprocal source, load_fw_points_rdbm
 * (readfields.f90:1514-1530): source Voigt (tt, pp, rr, rp, rt, tp) (source.f90:90-95)
 * src xyz -> earth xyz -> receiver xyz -> receiver (s, phi, z), / bwd amplitude. */
static void seis_moment(ora_ctx *C, int r, double phi, double *mb)
{
    const double *m = C->S.src_mij;
    int i;
    mb[0] = m[1]; mb[1] = m[2]; mb[2] = m[0]; mb[3] = m[4]; mb[4] = m[3]; mb[5] = m[5];
    rot_xyz_src_to_earth(mb, 1, C->src_lon, C->src_colat);
    rot_earth_to_xyz_src(mb, C->R[r].lon, C->R[r].colat);
    rot_xyz_to_src(mb, phi);
    for (i = 0; i < 6; ++i) mb[i] = mb[i] / C->S.bwd.amplitude;
}

/* load_seismogram / load_fw_points_rdbm (readfields.f90:1004-1048, 1469-1604): the
 * seismogram of every receiver by reciprocity, seis[nrec][nt].  The source point is
 * located in the receiver frame in the fwd mesh (without the surface clamp) and the full
 * strain of simulation 1 (Z) or 2 (R, T) of the bwd database is read at that element
 * (load_strain_point_merged_bwd, readfields.f90:2209-2403, comps 4 and 6 negated), then
 * contracted with the rotated moment tensor in the reference's order of operations. */
int ora_seismograms(ora_ctx *C, double *seis)
{
    const mck_setup_t *S = &C->S;
    const mck_db_t *fm = &S->fwd, *bd = &S->bwd;
    const int nt = S->nt, isim = S->component == 'Z' ? 1 : 2;
    const int pole = isim == 1 ? 0 : 1, var0 = isim == 1 ? 0 : 2, nv = isim == 1 ? 2 : 3;
    const size_t blk = 25 * (size_t)nt;
    double *u = (double *)malloc(3 * blk * sizeof(double));
    double *st = (double *)malloc(6 * blk * sizeof(double));
    double *ut = (double *)malloc(6 * (size_t)nt * sizeof(double));
    int r, k, v, rc = 0;
    if (fm->nelem != bd->nelem) { rc = MCK_EINVAL; goto done; }
    for (r = 0; r < S->nrec; ++r) {
        double s, phi, z, xi, eta, nodes[8], mb[6];
        int e, tie, i;
        double *o = seis + (size_t)r * nt;
        const float *src;
        ora_rotate_frame(1, C->src_r, C->R[r].lon, C->R[r].colat, &s, &phi, &z);
        if (find_element(fm, s, z, &xi, &eta, nodes, &e, &tie)) { rc = MCK_ENOTFOUND; goto done; }
        seis_moment(C, r, phi, mb);
        src = bd->snapshots + ((size_t)e * bd->nvar + var0) * blk;
        for (v = 0; v < nv; ++v)
            for (k = 0; k < (int)blk; ++k) u[v * blk + k] = (double)src[v * blk + k];
        ora_strain_pole(pole, u, nt, 4, nodes, fm->eltype[e], fm->axis[e], st);
        interp_nodes(6, nt, st, fm->axis[e] ? C->glj : C->gll, C->gll, xi, eta, ut);
        for (k = 0; k < nt; ++k) { ut[3 * nt + k] = -ut[3 * nt + k]; ut[5 * nt + k] = -ut[5 * nt + k]; }
        for (k = 0; k < nt; ++k) {
            double a = 0.0;
            if (S->component == 'Z') {
                for (i = 0; i < 3; ++i) a = a + mb[i] * ut[i * nt + k];
                a = a + 2 * mb[4] * ut[4 * nt + k];
            } else if (S->component == 'R') {
                a = a + mb[0] * ut[0 * nt + k];
                a = a + mb[1] * ut[1 * nt + k];
                a = a + mb[2] * ut[2 * nt + k];
                a = a + mb[4] * ut[4 * nt + k] * 2;
                a = -a;
            } else {
                a = a + mb[3] * ut[3 * nt + k] * 2;
                a = a + mb[5] * ut[5 * nt + k] * 2;
            }
            o[k] = a;
        }
    }
done:
    free(u); free(st); free(ut);
    return rc;
}

/* rfft(taperandzeropad(field, N, 10)) of every trace (slave.f90:525, 592) */
static void fields_to_fd(ora_ctx *C, int ntr, const double *td, double *fd, const double *ts)
{
    int N = C->N, nw = C->nomega, i, k;
    double *pad = (double *)malloc(N * sizeof(double));
    for (i = 0; i < ntr; ++i) {
        double *X = fd + (size_t)i * 2 * nw;
        ora_taperandzeropad(C->S.nt, N, 10, 0, td + (size_t)i * C->S.nt, pad);
        ora_rfft(N, C->S.dt, pad, X);
        if (ts)
            for (k = 0; k < nw; ++k) {
                cz a = {X[2 * k], X[2 * k + 1]}, b = {ts[2 * k], ts[2 * k + 1]};
                a = zmul(a, b);
                X[2 * k] = a.re; X[2 * k + 1] = a.im;
            }
    }
    free(pad);
}

/* calc_basekernel (kernel.f90:514-570); F, B [nd][nomega][2] -> out [nomega][2];
 * ibk 0..5 = lambda, mu, rho, a, b, c; full: the strain types carry the 6 Voigt
 * components (else the trace alone, nd 1) */
void ora_basekernel(int ibk, int full, int nd, int nw, const double *F, const double *B, double *out)
{
    int k, c;
#define FF(c) (cz){F[((c) * nw + k) * 2], F[((c) * nw + k) * 2 + 1]}
#define BB(c) (cz){B[((c) * nw + k) * 2], B[((c) * nw + k) * 2 + 1]}
    for (k = 0; k < nw; ++k) {
        cz r = {0, 0};
        if (ibk == 0) {
            if (!full) {
                for (c = 0; c < nd; ++c) { cz p = zmul(FF(c), BB(c)); r.re += p.re; r.im += p.im; }
                r.re = -r.re; r.im = -r.im;
            } else {
                cz fs = FF(0), bs = BB(0);
                fs.re = fs.re + FF(1).re; fs.im = fs.im + FF(1).im;
                fs.re = fs.re + FF(2).re; fs.im = fs.im + FF(2).im;
                bs.re = bs.re + BB(1).re; bs.im = bs.im + BB(1).im;
                bs.re = bs.re + BB(2).re; bs.im = bs.im + BB(2).im;
                fs.re = -fs.re; fs.im = -fs.im;
                r = zmul(fs, bs);
            }
        } else if (ibk == 1) {
            for (c = 0; c < nd; ++c) { cz p = zmul(FF(c), BB(c)); r.re += p.re; r.im += p.im; }
            r.re = -r.re * 2.0; r.im = -r.im * 2.0;
        } else if (ibk == 3) {
            cz a = {FF(1).re + FF(0).re, FF(1).im + FF(0).im}, b = {BB(1).re + BB(0).re, BB(1).im + BB(0).im};
            r = zmul(a, b);
        } else if (ibk == 4) {
            cz p = zmul(BB(4), FF(4)), q = zmul(BB(3), FF(3));
            r.re = (p.re + q.re) * 4.0; r.im = (p.im + q.im) * 4.0;
        } else if (ibk == 5) {
            cz a = {FF(0).re + FF(1).re, FF(0).im + FF(1).im};
            cz p = zmul(a, BB(2)), q = zmul(zmul(BB(0), BB(1)), FF(2));
            r.re = p.re + q.re; r.im = p.im + q.im;
        }
        out[2 * k] = r.re; out[2 * k + 1] = r.im;
    }
#undef FF
#undef BB
}

/* apply_filter_2d + irfft + calc_misfit_kernel (kernel.f90:410-510, 946-956) with the
 * transfer functions tf (forward) and tb (backward), [nomega][2] */
static double misfit_value_tf(ora_ctx *C, int k, const double *conv_fd, const double *tf,
                              const double *tb, double *work)
{
    ora_kernel *K = &C->K[k];
    int N = C->N, nw = C->nomega, i;
    double *X = work, *x = work + 2 * nw, *pad, *af, val;
    for (i = 0; i < nw; ++i) {
        cz a = {conv_fd[2 * i], conv_fd[2 * i + 1]}, f = {tf[2 * i], tf[2 * i + 1]}, b = {tb[2 * i], tb[2 * i + 1]};
        a = zmul(a, f);
        a = zmul(a, b);
        X[2 * i] = a.re; X[2 * i + 1] = a.im;
    }
    ora_irfft(N, C->df, X, x);
    if (C->S.int_scheme == MCK_INT_TRAPEZOIDAL) {
        double s = 0.0, f1, fn;
        f1 = x[K->ifirst] * K->seis_cut[0];
        fn = x[K->ifirst + K->ncut - 1] * K->seis_cut[K->ncut - 1];
        for (i = 1; i < K->ncut - 1; ++i) s += x[K->ifirst + i] * K->seis_cut[i];
        val = (f1 + fn + 2 * s) * C->S.dt * 0.5;
    } else {
        pad = (double *)calloc(K->M, sizeof(double));
        af = (double *)malloc(2 * (K->M / 2 + 1) * sizeof(double));
        for (i = 0; i < K->ncut; ++i) pad[i] = x[K->ifirst + i];
        ora_rfft(K->M, C->S.dt, pad, af);
        val = parseval(K->M, C->S.dt, af, K->seis_cut_fd);
        free(pad); free(af);
    }
    return val * K->normalization * (-1.0);
}

static double misfit_value(ora_ctx *C, int k, const double *conv_fd, double *work)
{
    int nw = C->nomega, f = C->K[k].filter;
    return misfit_value_tf(C, k, conv_fd, C->tf_fwd + (size_t)f * 2 * nw, C->tf_bwd + (size_t)f * 2 * nw, work);
}

/* model coefficients at radius r (interpolate_mesh.f90:98-109, background_model.f90:122-198) */
typedef struct { float vp, vs, rho, phi, xi, eta, lam, mu, vsh, vsv, vph, vpv; } ora_model;
static ora_model model_at(const mck_setup_t *S, double r)
{
    ora_model m;
    int idx = (int)(r / (double)S->model_dr + 1);
    if (idx < 1) idx = 1;
    if (idx > S->model_ndepth) idx = S->model_ndepth;
    idx -= 1;
    m.vp = S->model_vp[idx]; m.vs = S->model_vs[idx]; m.rho = S->model_rho[idx];
    m.phi = S->model_phi[idx]; m.xi = S->model_xi[idx]; m.eta = S->model_eta[idx];
    m.lam = m.rho * (m.vp * m.vp - 2 * (m.vs * m.vs));
    m.mu = m.rho * (m.vs * m.vs);
    m.vsh = m.vs; m.vph = m.vp;
    m.vsv = m.vsh / sqrtf(m.xi);
    m.vpv = m.vph * sqrtf(m.phi);
    return m;
}

/* calc_physical_kernels_scalar (kernel.f90:665-832); b = base kernels [6] */
static double physical(int param, const double *b, const ora_model *m, int rel)
{
    double rho = m->rho;
    switch (param) {
    case MCK_PARAM_LAM: return rel ? b[0] * (double)m->lam : b[0];
    case MCK_PARAM_MU:  return rel ? b[1] * (double)m->mu : b[1];
    case MCK_PARAM_VP:  return rel ? 2.0 * rho * (double)(m->vp * m->vp) * b[0] : 2.0 * rho * (double)m->vp * b[0];
    case MCK_PARAM_VS:  return (rel ? 2.0 * rho * (double)(m->vs * m->vs) : 2.0 * rho * (double)m->vs) * (b[1] - 2 * b[0]);
    case MCK_PARAM_VSH: return (rel ? 2.0 * rho * (double)(m->vsh * m->vsh) : 2.0 * rho * (double)m->vsh)
                               * (b[1] - 2 * b[0] - b[4] + 2 * b[5] * (double)(1 - m->eta));
    case MCK_PARAM_VSV: return (rel ? 2.0 * rho * (double)(m->vsv * m->vsv) : 2.0 * rho * (double)m->vsv) * b[4];
    case MCK_PARAM_VPH: return (rel ? 2.0 * rho * (double)(m->vph * m->vph) : 2.0 * rho * (double)m->vph)
                               * (b[3] + b[5] * (double)m->eta);
    case MCK_PARAM_VPV: return (rel ? 2.0 * rho * (double)(m->vpv * m->vpv) : 2.0 * rho * (double)m->vpv)
                               * (b[0] - b[3] - b[5]);
    case MCK_PARAM_ETA: return rel ? 2.0 * rho * (double)(m->vph * m->vph - m->vsh * m->vsh) * b[5] * (double)m->eta
                                   : 2.0 * rho * (double)(m->vph * m->vph - m->vsh * m->vsh) * b[5];
    default: return 0.0;
    }
}

/* Test hook: physical kernels alone; model values as the fp32 table holds them */
double ora_physical(int param, const double *b, float vp, float vs, float rho, float phi, float xi,
                    float eta, int rel)
{
    ora_model m;
    m.vp = vp; m.vs = vs; m.rho = rho; m.phi = phi; m.xi = xi; m.eta = eta;
    m.lam = rho * (vp * vp - 2 * (vs * vs));
    m.mu = rho * (vs * vs);
    m.vsh = vs; m.vph = vp; m.vsv = vs / sqrtf(xi); m.vpv = vp * sqrtf(phi);
    return physical(param, b, &m, rel);
}

/* forward field of a batch of points: load_fw_points, cumsum_trapezoidal, rfft of the
 * tapered field and the optional time shift (slave.f90:515-534); F[npts][nd][nomega][2] */
static int batch_fwd(ora_ctx *C, int npts, const double *pts, double *F)
{
    const mck_setup_t *S = &C->S;
    int nt = S->nt, nd = C->nd, p, rc;
    double *fw = (double *)malloc((size_t)npts * nd * nt * sizeof(double));
    double *tmp = (double *)malloc((size_t)nt * sizeof(double));
    rc = load_points(C, 0, npts, pts, fw);
    if (!rc) {
        for (p = 0; p < npts * nd; ++p) {
            ora_cumsum_trapezoidal(nt, S->dt, fw + (size_t)p * nt, tmp);
            memcpy(fw + (size_t)p * nt, tmp, nt * sizeof(double));
        }
        fields_to_fd(C, npts * nd, fw, F, S->deconv_stf ? NULL : C->ts_fwd);
    }
    free(fw); free(tmp);
    return rc;
}

/* one receiver of slave_work's receiver loop (slave.f90:586-640): load_bw_points, rfft,
 * calc_basekernel per needed base kernel, apply_filter, irfft, calc_misfit_kernel for
 * every kernel of the receiver that is not converged. */
static int receiver_base(ora_ctx *C, int r, int npts, const double *pts, const double *F,
                         const int *conv_flags, double *base, int64_t *evals, int32_t *niter)
{
    const mck_setup_t *S = &C->S;
    ora_rec *R = &C->R[r];
    int nt = S->nt, nd = C->nd, nw = C->nomega, p, b, k, rc;
    double *bw = (double *)malloc((size_t)npts * nd * nt * sizeof(double));
    double *B = (double *)malloc((size_t)npts * nd * nw * 2 * sizeof(double));
    double *conv = (double *)malloc((size_t)nw * 2 * sizeof(double));
    double *work = (double *)malloc(((size_t)2 * nw + C->N) * sizeof(double));
    rc = load_points(C, 1 + r, npts, pts, bw);
    if (rc) goto done;
    fields_to_fd(C, npts * nd, bw, B, S->deconv_stf ? NULL : C->ts_bwd);
    for (b = 0; b < 6; ++b) {
        if (!R->needs_bk[b]) continue;
        for (k = R->first; k <= R->last; ++k) {
            if (!C->K[k].needs_bk[b]) continue;
            if (conv_flags && conv_flags[k]) continue;
            for (p = 0; p < npts; ++p) {
                ora_basekernel(b, C->full, nd, nw, F + (size_t)p * nd * nw * 2, B + (size_t)p * nd * nw * 2, conv);
                base[((size_t)p * S->nkernel + k) * 6 + b] = misfit_value(C, k, conv, work);
            }
            if (niter) niter[k] += 1;
            if (evals) {
                int bb, first = 1;
                for (bb = 0; bb < b; ++bb) if (C->K[k].needs_bk[bb]) first = 0;
                if (first) *evals += npts;
            }
        }
    }
done:
    free(bw); free(B); free(conv); free(work);
    return rc;
}

int ora_base_kernels(ora_ctx *C, int npts, const double *xyz, double *base)
{
    int r, rc;
    double *F = (double *)malloc((size_t)npts * C->nd * C->nomega * 2 * sizeof(double));
    memset(base, 0, (size_t)npts * C->S.nkernel * 6 * sizeof(double));
    rc = batch_fwd(C, npts, xyz, F);
    for (r = 0; r < C->S.nrec && !rc; ++r) rc = receiver_base(C, r, npts, xyz, F, NULL, base, NULL, NULL);
    free(F);
    return rc;
}

/* dampen_field (readfields.f90:3184-3204) on base[npts][nkernel][6] */
void ora_dampen_field(double *base, int npts, int nk, const double *pts, const double *c, double rmax)
{
    int p, i;
    if (!(rmax > 0.0)) return;
    for (p = 0; p < npts; ++p) {
        double d[3] = {pts[3 * p] - c[0], pts[3 * p + 1] - c[1], pts[3 * p + 2] - c[2]};
        double dist = ora_norm2(3, d);
        if (dist < rmax)
            for (i = 0; i < nk * 6; ++i) base[(size_t)p * nk * 6 + i] = base[(size_t)p * nk * 6 + i] * dist / rmax;
    }
}

/* integrated_type (mc_integration.f90:69-187) */
typedef struct {
    int n;
    double volume;
    int64_t nmodels;
    double *fsum, *f2sum, *integral, *variance, allowed_var, allowed_rel;
    int *conv;
} ora_mc;

static void mc_init(ora_mc *m, int n, double volume, double aerr, double arel)
{
    int i;
    m->n = n; m->volume = volume; m->nmodels = 0;
    m->fsum = (double *)calloc(4 * (size_t)n, sizeof(double));
    m->f2sum = m->fsum + n; m->integral = m->f2sum + n; m->variance = m->integral + n;
    m->conv = (int *)calloc(n, sizeof(int));
    m->allowed_var = aerr * aerr;
    m->allowed_rel = arel;
    for (i = 0; i < n; ++i) m->variance[i] = 1e100;
}

static int mc_check(ora_mc *m, int npts, const double *vals /* [npts][n] */)
{
    int i, p;
    m->nmodels += npts;
    for (i = 0; i < m->n; ++i) {
        double s = 0.0, s2 = 0.0;
        if (m->conv[i]) continue;
        for (p = 0; p < npts; ++p) s += vals[(size_t)p * m->n + i];
        for (p = 0; p < npts; ++p) s2 += vals[(size_t)p * m->n + i] * vals[(size_t)p * m->n + i];
        m->fsum[i] = m->fsum[i] + s;
        m->f2sum[i] = m->f2sum[i] + s2;
        m->integral[i] = m->fsum[i] * m->volume / (double)m->nmodels;
        if (m->nmodels > 2) {
            m->variance[i] = m->volume * m->volume / (double)(m->nmodels - 1)
                           * ((m->f2sum[i] / (double)m->nmodels)
                              - (m->fsum[i] / (double)m->nmodels) * (m->fsum[i] / (double)m->nmodels));
            if (m->variance[i] < m->allowed_var) m->conv[i] = 1;
            else if (sqrt(m->variance[i]) < m->allowed_rel * fabs(m->integral[i])) m->conv[i] = 1;
        } else {
            m->variance[i] = m->integral[i] * 1e10;
            m->conv[i] = 0;
        }
        if (m->integral[i] != m->integral[i] || m->variance[i] != m->variance[i]) return MCK_ENAN;
    }
    return 0;
}

static void mc_free(ora_mc *m) { free(m->fsum); free(m->conv); }

/* Test hook: the integrator alone, fed nsteps batches of npts values of n functions
 * (vals[nsteps][npts][n]); returns the final integral/variance/convergence flags. */
int ora_mc_run(int n, double volume, double allowed_error, double allowed_rel, int nsteps,
               int npts, const double *vals, double *integral, double *variance, int32_t *conv)
{
    ora_mc m;
    int s, i, rc = 0;
    mc_init(&m, n, volume, allowed_error, allowed_rel);
    for (s = 0; s < nsteps && !rc; ++s) rc = mc_check(&m, npts, vals + (size_t)s * npts * n);
    for (i = 0; i < n; ++i) { integral[i] = m.integral[i]; variance[i] = m.variance[i]; conv[i] = m.conv[i]; }
    mc_free(&m);
    return rc;
}

/* travel_time_larger_Tmax (slave.f90:928-950) */
static int travel_time_masked(ora_ctx *C, int r, const double *center)
{
    double d1[3], d2[3], dist, tp, ts;
    int i;
    for (i = 0; i < 3; ++i) { d1[i] = C->src_r[i] - center[i]; d2[i] = C->R[r].r[i] - center[i]; }
    dist = ora_norm2(3, d1) + ora_norm2(3, d2);
    tp = dist / (double)C->S.vp_max;
    ts = dist / (double)C->S.vs_max;
    return (tp > C->R[r].t_last_p && ts > C->R[r].t_last_s);
}

/* init_weight_tet_mesh (inversion_mesh.f90:1086-1116): the inverse of
 * ab = [[v_j - v_1], [1]] (columns j = 1..4), by the adjugate; A[4 i + j] = inv(ab)(i, j) */
static double det3(double a, double b, double c, double d, double e, double f, double g, double h, double i)
{
    return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
}

static void tet_abinv(const double *v, double *A)
{
    double ab[4][4], cof[4][4], det = 0.0;
    int r, c, i, j;
    for (c = 0; c < 4; ++c) {
        for (r = 0; r < 3; ++r) ab[r][c] = v[3 * c + r] - v[r];
        ab[3][c] = 1.0;
    }
    for (r = 0; r < 4; ++r)
        for (c = 0; c < 4; ++c) {
            double m[9];
            int n = 0;
            for (i = 0; i < 4; ++i)
                for (j = 0; j < 4; ++j)
                    if (i != r && j != c) m[n++] = ab[i][j];
            cof[r][c] = ((r + c) & 1 ? -1.0 : 1.0) * det3(m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8]);
        }
    for (c = 0; c < 4; ++c) det += ab[0][c] * cof[0][c];
    for (r = 0; r < 4; ++r)
        for (c = 0; c < 4; ++c) A[4 * r + c] = cof[c][r] / det;
}

/* weights (inversion_mesh.f90:530-622): basis function ivertex of a tetrahedron at the
 * points, sum_j abinv(ivertex, j) (dx, dy, dz, 1)_j with d = point - vertex 1 */
static int tet_weights(const double *v, const double *A, int ivertex, int npts, const double *pts, double *w)
{
    int p, bad = 0;
    for (p = 0; p < npts; ++p) {
        double dx = pts[3 * p] - v[0], dy = pts[3 * p + 1] - v[1], dz = pts[3 * p + 2] - v[2];
        w[p] = A[4 * ivertex] * dx + A[4 * ivertex + 1] * dy + A[4 * ivertex + 2] * dz + A[4 * ivertex + 3] * 1;
        if (w[p] < -1e-9 || w[p] > 1.0 + 1e-9) bad = 1;
    }
    return bad;
}

/* Test hook: the four vertex basis functions of a tetrahedron v[4][3] at pts[npts][3]:
 * w[npts][4]; returns 1 if a weight is outside [-1e-9, 1 + 1e-9] */
int ora_tet_weights(const double *v, int npts, const double *pts, double *w)
{
    double A[16], *t = (double *)malloc((size_t)npts * sizeof(double));
    int b, p, bad = 0;
    tet_abinv(v, A);
    for (b = 0; b < 4; ++b) {
        bad |= tet_weights(v, A, b, npts, pts, t);
        for (p = 0; p < npts; ++p) w[4 * p + b] = t[p];
    }
    free(t);
    return bad;
}

/* slave_work (slave.f90:273-810) for tetrahedra (elem_type 4, nvpe 4; volumetric or
 * on-vertex basis functions, nbasis 1 or 4) and voxels (elem_type 5, nvpe 8; nbasis 1 or
 * 8, weight 1 for every basis function as weights() gives non-tetrahedra,
 * inversion_mesh.f90:614-617): one integrator per basis function, values
 * [nelems][nbasis][nkernel].  Tetrahedra take Halton points (quasirandom) or the
 * pseudo-random stream; voxels always the pseudo-random stream (inversion_mesh.f90:662). */
int ora_slave_work(ora_ctx *C, int elem_type, int nvpe, int nelems, int nbasis,
                   const double *vertices, const int32_t *conn,
                   double *values, double *variance, int32_t *niterations,
                   int64_t *halton_step, const int64_t *halton_offsets,
                   int64_t *offsets_used, int64_t *evaluations)
{
    const int vox = elem_type == 5;
    const mck_setup_t *S = &C->S;
    int nk = S->nkernel, P = S->npoints_per_step, e, r, k, i, b, rc = 0;
    double *pts = (double *)malloc((size_t)P * 3 * sizeof(double));
    double *hal = (double *)malloc((size_t)P * 3 * sizeof(double));
    double *base = (double *)calloc((size_t)P * nk * 6, sizeof(double));
    double *wbase = (double *)malloc((size_t)P * nk * 6 * sizeof(double));
    double *phys = (double *)malloc((size_t)P * nk * sizeof(double));
    double *wts = (double *)malloc((size_t)P * sizeof(double));
    double *F = (double *)malloc((size_t)P * C->nd * C->nomega * 2 * sizeof(double));
    int32_t *niter = (int32_t *)malloc(nk * sizeof(int32_t));
    int *conv_all = (int *)malloc(nk * sizeof(int));
    ora_mc mcs[8];                     /* one integrator per basis function (<= 8) */
    static const int64_t seed[3] = {20, 20, 20}, leap[3] = {2, 2, 2};
    static const int32_t bases[3] = {2, 3, 5};
    if (!((elem_type == 4 && nvpe == 4 && (nbasis == 1 || nbasis == 4))
          || (vox && nvpe == 8 && (nbasis == 1 || nbasis == 8)))) { rc = MCK_EINVAL; goto out; }
    if (evaluations) *evaluations = 0;
    C->halton_step = *halton_step;
    for (e = 0; e < nelems; ++e) {
        double v[24], center[3], volume, A[16];
        for (i = 0; i < nvpe; ++i) {
            int vid = conn[e * nvpe + i] - 1;
            v[3 * i] = vertices[3 * vid]; v[3 * i + 1] = vertices[3 * vid + 1]; v[3 * i + 2] = vertices[3 * vid + 2];
        }
        if (vox) {
            ora_vox_center(v, center);
            volume = S->int_over_volume ? ora_vox_volume(v) : 1.0;
        } else {
            for (i = 0; i < 3; ++i) center[i] = (v[i] + v[3 + i] + v[6 + i] + v[9 + i]) / 4.0;
            volume = S->int_over_volume ? ora_tet_volume(v) : 1.0;
            if (nbasis > 1) tet_abinv(v, A);
        }
        if (halton_offsets) C->halton_step = halton_offsets[e];
        if (offsets_used) offsets_used[e] = C->halton_step;
        /* element containing the source: no integration (slave.f90:466-482) */
        if (S->mask_src_rec && (vox ? ora_point_in_vox(v, C->src_r) : ora_point_in_tet(v, C->src_r))) {
            for (k = 0; k < nk * nbasis; ++k) {
                values[(size_t)e * nk * nbasis + k] = 0.0;
                variance[(size_t)e * nk * nbasis + k] = 0.0;
            }
            for (k = 0; k < nk; ++k) niterations[(size_t)e * nk + k] = 0;
            continue;
        }
        for (b = 0; b < nbasis; ++b)
            mc_init(&mcs[b], nk, volume, S->allowed_error * volume, S->allowed_relative_error);
        for (k = 0; k < nk; ++k) niter[k] = 0;
        for (;;) {
            int all = 1, stop = 0;
            /* allallconverged / allisconverged (mc_integration.f90:322-359) */
            for (k = 0; k < nk; ++k) {
                conv_all[k] = 1;
                for (b = 0; b < nbasis; ++b) if (!mcs[b].conv[k]) conv_all[k] = 0;
                if (!conv_all[k]) all = 0;
            }
            if (all) break;
            if (S->quasirandom && !vox) ora_halton(3, seed, bases, leap, C->halton_step, P, hal);
            else ora_prng_points(C->halton_step, P, hal);
            C->halton_step += P;
            if (vox) ora_vox_points(v, P, hal, pts);
            else ora_tet_points(v, P, hal, pts);
            for (k = 0; k < nk; ++k) if (niter[k] == S->max_iter) stop = 1;
            if (stop) break;
            rc = batch_fwd(C, P, pts, F);
            if (rc) { for (b = 0; b < nbasis; ++b) mc_free(&mcs[b]); goto out; }
            for (r = 0; r < S->nrec; ++r) {
                ora_rec *R = &C->R[r];
                int allconv = 1;
                /* receiver inside the element (slave.f90:539-545), then the travel-time
                 * mask (slave.f90:546-557): zero kernels, one iteration each */
                if ((S->mask_src_rec && (vox ? ora_point_in_vox(v, C->R[r].r) : ora_point_in_tet(v, C->R[r].r)))
                    || travel_time_masked(C, r, center)) {
                    for (k = R->first; k <= R->last; ++k) {
                        int b;
                        for (i = 0; i < P; ++i) for (b = 0; b < 6; ++b) base[((size_t)i * nk + k) * 6 + b] = 0.0;
                        niter[k] += 1;
                    }
                    continue;
                }
                for (k = R->first; k <= R->last; ++k) if (!conv_all[k]) allconv = 0;
                if (allconv) continue;
                rc = receiver_base(C, r, P, pts, F, conv_all, base, evaluations, niter);
                if (rc) { for (b = 0; b < nbasis; ++b) mc_free(&mcs[b]); goto out; }
                ora_dampen_field(base, P, nk, pts, C->R[r].r, S->damp_radius);
                ora_dampen_field(base, P, nk, pts, C->src_r, S->damp_radius);
            }
            /* per basis function: weights, mult3d_1d, physical kernels, MC check
             * (slave.f90:723-744) */
            for (b = 0; b < nbasis; ++b) {
                if (nbasis == 1 || vox) {
                    for (i = 0; i < P; ++i) wts[i] = 1.0;
                } else if (tet_weights(v, A, b, P, pts, wts)) {
                    rc = MCK_EINVAL;                  /* weight outside [0, 1]: pabort */
                    for (i = 0; i < nbasis; ++i) mc_free(&mcs[i]);
                    goto out;
                }
                for (i = 0; i < P; ++i)
                    for (k = 0; k < nk * 6; ++k) wbase[(size_t)i * nk * 6 + k] = base[(size_t)i * nk * 6 + k] * wts[i];
                for (i = 0; i < P; ++i) {
                    double rr = ora_norm2(3, pts + 3 * i);
                    ora_model m = model_at(S, rr);
                    for (k = 0; k < nk; ++k)
                        phys[(size_t)i * nk + k] = physical(S->kernel_param[k], wbase + ((size_t)i * nk + k) * 6, &m, S->relative_kernel);
                }
                rc = mc_check(&mcs[b], P, phys);
                if (rc) { for (i = 0; i < nbasis; ++i) mc_free(&mcs[i]); goto out; }
            }
        }
        for (b = 0; b < nbasis; ++b)
            for (k = 0; k < nk; ++k) {
                values[((size_t)e * nbasis + b) * nk + k] = mcs[b].integral[k];
                variance[((size_t)e * nbasis + b) * nk + k] = mcs[b].variance[k];
            }
        for (k = 0; k < nk; ++k) niterations[(size_t)e * nk + k] = niter[k];
        for (b = 0; b < nbasis; ++b) mc_free(&mcs[b]);
    }
out:
    *halton_step = C->halton_step;
    free(pts); free(hal); free(base); free(wbase); free(phys); free(wts); free(F); free(niter); free(conv_all);
    return rc;
}

/* ------------------------------------------------------------------------------------
 * The filter-convolve-window-integrate stage as a function of the filter responses G
 * and the strain traces, and its adjoint.  The stage has no counterpart in the reference;
 * its forward is the receiver-loop body of slave_work (slave.f90:515-534, 586-640:
 * cumsum_trapezoidal and rfft of the tapered fwd traces, rfft of the tapered bwd traces,
 * calc_basekernel kernel.f90:514-570, apply_filter_2d kernel.f90:946-956, irfft,
 * calc_misfit_kernel kernel.f90:410-510) with TF_fwd = G rest_f, TF_bwd = G rest_b, where
 * G is a filter's response (filtering.f90:78-230) and rest_* the deconvolution factors of
 * add_stfs (filtering.f90:241-470; 1 without deconvolution).  The filtered seismogram
 * windows and normalisations are those of the set-up (kernel.f90:181-385), constants of
 * the stage.  The adjoint runs the same steps backwards in FP64.
 * ---------------------------------------------------------------------------------- */

/* integrate_1d_model (slave.f90:815-869): per inversion element and basis function the
 * Monte Carlo integral (volume 1, allowed error 1e-2, relative 1e-2,
 * initialize_montecarlo, mc_integration.f90:147-187) of the 12 background-model
 * parameters in the order of background_model.f90:34-37 (vp vs rho vph vpv vsh vsv eta
 * phi xi lam mu; combine(), background_model.f90:122-198, in fp32), each times the basis
 * function's weight (weight(), background_model.f90:233-261), over 100 points per step:
 * Halton points in tetrahedra whatever quasirandom says (generate_random_points(ielement,
 * nptperstep_model, .true.)), pseudo-random points in voxels; the loop stops when every
 * integrator has converged or after 9999 steps.  Stream positions: halton_offsets[e], or
 * *halton_step + e * 9999 * 100 (then *halton_step advances past the task).
 * model[nelems][nbasis][12]; nsteps[nelems] (nullable).  Elements whose first vertex id is
 * <= 0 (empty slots of a last task) give zeros. */
int ora_integrate_model(ora_ctx *C, int elem_type, int nvpe, int nelems, int nbasis,
                        const double *vertices, const int32_t *conn, double *model,
                        int32_t *nsteps, int64_t *halton_step, const int64_t *halton_offsets)
{
    enum { NP = 100, NF = 12, MAXSTEP = 9999 };
    const int vox = elem_type == 5;
    const mck_setup_t *S = &C->S;
    static const int64_t seed[3] = {20, 20, 20}, leap[3] = {2, 2, 2};
    static const int32_t bases[3] = {2, 3, 5};
    double pts[NP * 3], hal[NP * 3], wts[NP], vals[NP * NF];
    ora_mc mcs[8];
    int e, i, k, b, rc = 0;
    if (!((elem_type == 4 && nvpe == 4 && (nbasis == 1 || nbasis == 4))
          || (vox && nvpe == 8 && (nbasis == 1 || nbasis == 8)))) return MCK_EINVAL;
    for (e = 0; e < nelems; ++e) {
        double v[24], A[16];
        int64_t pos = halton_offsets ? halton_offsets[e] : *halton_step + (int64_t)e * MAXSTEP * NP;
        int istep = 0, all;
        if (conn[e * nvpe] <= 0) {
            for (k = 0; k < nbasis * NF; ++k) model[(size_t)e * nbasis * NF + k] = 0.0;
            if (nsteps) nsteps[e] = 0;
            continue;
        }
        for (i = 0; i < nvpe; ++i) {
            int vid = conn[e * nvpe + i] - 1;
            v[3 * i] = vertices[3 * vid]; v[3 * i + 1] = vertices[3 * vid + 1]; v[3 * i + 2] = vertices[3 * vid + 2];
        }
        if (!vox && nbasis > 1) tet_abinv(v, A);
        for (b = 0; b < nbasis; ++b) mc_init(&mcs[b], NF, 1.0, 1e-2, 1e-2);
        for (;;) {
            /* allallconverged (mc_integration.f90:322-340) */
            all = 1;
            for (b = 0; b < nbasis; ++b) for (k = 0; k < NF; ++k) if (!mcs[b].conv[k]) all = 0;
            if (all) break;
            if (!vox) ora_halton(3, seed, bases, leap, pos, NP, hal);
            else ora_prng_points(pos, NP, hal);
            pos += NP;
            if (vox) ora_vox_points(v, NP, hal, pts);
            else ora_tet_points(v, NP, hal, pts);
            for (b = 0; b < nbasis && !rc; ++b) {
                if (nbasis == 1 || vox) {
                    for (i = 0; i < NP; ++i) wts[i] = 1.0;
                } else if (tet_weights(v, A, b, NP, pts, wts)) {
                    rc = MCK_EINVAL;
                    break;
                }
                for (i = 0; i < NP; ++i) {
                    ora_model m = model_at(S, ora_norm2(3, pts + 3 * i));
                    const float c[NF] = {m.vp, m.vs, m.rho, m.vph, m.vpv, m.vsh, m.vsv, m.eta, m.phi, m.xi,
                                         m.lam, m.mu};
                    for (k = 0; k < NF; ++k) vals[i * NF + k] = (double)c[k] * wts[i];
                }
                rc = mc_check(&mcs[b], NP, vals);
            }
            if (rc) break;
            istep += 1;
            if (istep >= MAXSTEP) break;
        }
        for (b = 0; b < nbasis; ++b) {
            for (k = 0; k < NF; ++k) model[((size_t)e * nbasis + b) * NF + k] = mcs[b].integral[k];
            mc_free(&mcs[b]);
        }
        if (nsteps) nsteps[e] = istep;
        if (rc) return rc;
    }
    if (!halton_offsets) *halton_step += (int64_t)nelems * MAXSTEP * NP;
    return 0;
}

/* load_het_rtpv (heterogeneities.f90:63-110): r km, latitude, longitude in degrees to
 * single-precision Cartesian coordinates (spherical_to_cartesian_point,
 * random_voxel.f90:214-227, in double, then real(sp)); returns nonzero for a latitude
 * outside [-90, 90] */
int ora_het_xyz(int nhet, const double *rtp, float *xyz)
{
    int i;
    for (i = 0; i < nhet; ++i) {
        const double r = rtp[3 * i], t = rtp[3 * i + 1], p = rtp[3 * i + 2];
        double sph[3];
        if (fabs(t) > 90.0) return MCK_EINVAL;
        sph[0] = r * 1000.0;
        sph[1] = (90.0 - t) * DEG2RAD;
        sph[2] = p * DEG2RAD;
        xyz[3 * i] = (float)(sph[0] * cos(sph[2]) * sin(sph[1]));
        xyz[3 * i + 1] = (float)(sph[0] * sin(sph[2]) * sin(sph[1]));
        xyz[3 * i + 2] = (float)(sph[0] * cos(sph[1]));
    }
    return 0;
}

/* the nearest model point of load_model_coeffs (heterogeneities.f90:141-188,
 * kdtree2_n_nearest with nn = 1), by brute force over all points: single-precision
 * squared distance summed as kdtree2.f90:1541-1543 does, the lower index on equal
 * distances; 0-based */
int ora_het_nearest(int nhet, const float *xyz, float qx, float qy, float qz)
{
    float best = 3.0e38f;
    int bi = -1, j;
    for (j = 0; j < nhet; ++j) {
        const float dx = xyz[3 * j] - qx, dy = xyz[3 * j + 1] - qy, dz = xyz[3 * j + 2] - qz;
        const float d = (dx * dx + dy * dy) + dz * dz;
        if (d < best) { best = d; bi = j; }
    }
    return bi;
}

/* integrate_3d_model (slave.f90:874-924) for a task: hetero[nelems][nbasis][3] MC
 * integrals (volume 1, allowed error 1e-2, relative 1e-1) of dlnvp, dlnvs, dlnrho of the
 * nearest model point times the basis function, 10 points per step (Halton in
 * tetrahedra, the pseudo-random stream in voxels), at most 9999 steps */
int ora_integrate_hetero(int nhet, const double *rtp, const float *dlnhet, int elem_type, int nvpe,
                         int nelems, int nbasis, const double *vertices, const int32_t *conn,
                         double *hetero, int32_t *nsteps, int64_t *halton_step,
                         const int64_t *halton_offsets)
{
    enum { NP = 10, NF = 3, MAXSTEP = 9999 };
    const int vox = elem_type == 5;
    static const int64_t seed[3] = {20, 20, 20}, leap[3] = {2, 2, 2};
    static const int32_t bases[3] = {2, 3, 5};
    double pts[NP * 3], hal[NP * 3], wts[NP], vals[NP * NF];
    int near[NP];
    ora_mc mcs[8];
    int e, i, k, b, rc = 0;
    float *xyz;
    if (!((elem_type == 4 && nvpe == 4 && (nbasis == 1 || nbasis == 4))
          || (vox && nvpe == 8 && (nbasis == 1 || nbasis == 8))) || nhet <= 0) return MCK_EINVAL;
    xyz = (float *)malloc((size_t)nhet * 3 * sizeof(float));
    if (ora_het_xyz(nhet, rtp, xyz)) { free(xyz); return MCK_EINVAL; }
    for (e = 0; e < nelems && !rc; ++e) {
        double v[24], A[16];
        int64_t pos = halton_offsets ? halton_offsets[e] : *halton_step + (int64_t)e * MAXSTEP * NP;
        int istep = 0, all;
        if (conn[e * nvpe] <= 0) {
            for (k = 0; k < nbasis * NF; ++k) hetero[(size_t)e * nbasis * NF + k] = 0.0;
            if (nsteps) nsteps[e] = 0;
            continue;
        }
        for (i = 0; i < nvpe; ++i) {
            int vid = conn[e * nvpe + i] - 1;
            v[3 * i] = vertices[3 * vid]; v[3 * i + 1] = vertices[3 * vid + 1]; v[3 * i + 2] = vertices[3 * vid + 2];
        }
        if (!vox && nbasis > 1) tet_abinv(v, A);
        for (b = 0; b < nbasis; ++b) mc_init(&mcs[b], NF, 1.0, 1e-2, 1e-1);
        for (;;) {
            all = 1;
            for (b = 0; b < nbasis; ++b) for (k = 0; k < NF; ++k) if (!mcs[b].conv[k]) all = 0;
            if (all) break;
            if (!vox) ora_halton(3, seed, bases, leap, pos, NP, hal);
            else ora_prng_points(pos, NP, hal);
            pos += NP;
            if (vox) ora_vox_points(v, NP, hal, pts);
            else ora_tet_points(v, NP, hal, pts);
            for (i = 0; i < NP; ++i)
                near[i] = ora_het_nearest(nhet, xyz, (float)pts[3 * i], (float)pts[3 * i + 1], (float)pts[3 * i + 2]);
            for (b = 0; b < nbasis && !rc; ++b) {
                if (nbasis == 1 || vox) {
                    for (i = 0; i < NP; ++i) wts[i] = 1.0;
                } else if (tet_weights(v, A, b, NP, pts, wts)) {
                    rc = MCK_EINVAL;
                    break;
                }
                for (i = 0; i < NP; ++i)
                    for (k = 0; k < NF; ++k) vals[i * NF + k] = (double)dlnhet[3 * (size_t)near[i] + k] * wts[i];
                rc = mc_check(&mcs[b], NP, vals);
            }
            if (rc) break;
            istep += 1;
            if (istep >= MAXSTEP) break;
        }
        for (b = 0; b < nbasis; ++b) {
            for (k = 0; k < NF; ++k) hetero[((size_t)e * nbasis + b) * NF + k] = mcs[b].integral[k];
            mc_free(&mcs[b]);
        }
        if (nsteps) nsteps[e] = istep;
    }
    free(xyz);
    if (rc) return rc;
    if (!halton_offsets) *halton_step += (int64_t)nelems * MAXSTEP * NP;
    return 0;
}

/* G of every filter [nfilter][nomega][2] */
void ora_ad_transfer(ora_ctx *C, double *G)
{
    memcpy(G, C->tf_g, (size_t)C->S.nfilter * 2 * C->nomega * sizeof(double));
}

static void ad_tf(ora_ctx *C, const double *G, int f, double *ta, double *tb)
{
    int k, nw = C->nomega;
    for (k = 0; k < nw; ++k) {
        cz g = {G[((size_t)f * nw + k) * 2], G[((size_t)f * nw + k) * 2 + 1]};
        cz a = {C->rest_f[2 * k], C->rest_f[2 * k + 1]}, b = {C->rest_b[2 * k], C->rest_b[2 * k + 1]};
        a = zmul(g, a); b = zmul(g, b);
        ta[2 * k] = a.re; ta[2 * k + 1] = a.im;
        tb[2 * k] = b.re; tb[2 * k + 1] = b.im;
    }
}

/* F [npts][nd][nomega][2] of fw [npts][nd][nt], B [npts][nrec][nd][nomega][2] of bw */
static void ad_spectra(ora_ctx *C, int npts, const double *fw, const double *bw, double *F, double *B)
{
    const mck_setup_t *S = &C->S;
    int nt = S->nt, nd = C->nd, i;
    double *tmp = (double *)malloc((size_t)npts * nd * nt * sizeof(double));
    for (i = 0; i < npts * nd; ++i) ora_cumsum_trapezoidal(nt, S->dt, fw + (size_t)i * nt, tmp + (size_t)i * nt);
    fields_to_fd(C, npts * nd, tmp, F, S->deconv_stf ? NULL : C->ts_fwd);
    fields_to_fd(C, npts * S->nrec * nd, bw, B, S->deconv_stf ? NULL : C->ts_bwd);
    free(tmp);
}

int ora_ad_forward(ora_ctx *C, int npts, const double *G, const double *fw, const double *bw,
                   double *base)
{
    const mck_setup_t *S = &C->S;
    int nd = C->nd, nw = C->nomega, nk = S->nkernel, R = S->nrec, p, r, b, k;
    size_t fs = (size_t)nd * nw * 2;
    double *F = (double *)malloc((size_t)npts * fs * sizeof(double));
    double *B = (double *)malloc((size_t)npts * R * fs * sizeof(double));
    double *conv = (double *)malloc((size_t)nw * 2 * sizeof(double));
    double *ta = (double *)malloc((size_t)nw * 4 * sizeof(double)), *tb = ta + 2 * nw;
    double *work = (double *)malloc(((size_t)2 * nw + C->N) * sizeof(double));
    memset(base, 0, (size_t)npts * nk * 6 * sizeof(double));
    ad_spectra(C, npts, fw, bw, F, B);
    for (p = 0; p < npts; ++p)
        for (r = 0; r < R; ++r)
            for (b = 0; b < 6; ++b) {
                if (!C->R[r].needs_bk[b]) continue;
                for (k = C->R[r].first; k <= C->R[r].last; ++k) {
                    if (!C->K[k].needs_bk[b]) continue;
                    ora_basekernel(b, C->full, nd, nw, F + p * fs, B + ((size_t)p * R + r) * fs, conv);
                    ad_tf(C, G, C->K[k].filter, ta, tb);
                    base[((size_t)p * nk + k) * 6 + b] = misfit_value_tf(C, k, conv, ta, tb, work);
                }
            }
    free(F); free(B); free(conv); free(ta); free(work);
    return 0;
}

/* gradient of one calc_basekernel product (holomorphic in F and B): gF_c += g conj(dbk/dF_c),
 * gB_c += g conj(dbk/dB_c) at bin q; F, B, gF, gB [nd][nomega][2] */
static void basekernel_adjoint(int ibk, int full, int nd, int nw, int q, const double *F,
                               const double *B, cz g, double *gF, double *gB)
{
    cz dF[6], dB[6];
    int c;
#define FF(c) (cz){F[((c) * nw + q) * 2], F[((c) * nw + q) * 2 + 1]}
#define BB(c) (cz){B[((c) * nw + q) * 2], B[((c) * nw + q) * 2 + 1]}
    for (c = 0; c < 6; ++c) { dF[c].re = dF[c].im = dB[c].re = dB[c].im = 0.0; }
    if ((ibk == 0 && !full) || ibk == 1) {
        double s = ibk == 0 ? -1.0 : -2.0;
        for (c = 0; c < nd; ++c) {
            dF[c].re = s * BB(c).re; dF[c].im = s * BB(c).im;
            dB[c].re = s * FF(c).re; dB[c].im = s * FF(c).im;
        }
    } else if (ibk == 0) {
        cz fs = {-(FF(0).re + FF(1).re + FF(2).re), -(FF(0).im + FF(1).im + FF(2).im)};
        cz bs = {-(BB(0).re + BB(1).re + BB(2).re), -(BB(0).im + BB(1).im + BB(2).im)};
        for (c = 0; c < 3; ++c) { dF[c] = bs; dB[c] = fs; }
    } else if (ibk == 3) {
        cz fs = {FF(0).re + FF(1).re, FF(0).im + FF(1).im}, bs = {BB(0).re + BB(1).re, BB(0).im + BB(1).im};
        dF[0] = dF[1] = bs; dB[0] = dB[1] = fs;
    } else if (ibk == 4) {
        for (c = 3; c < 5; ++c) {
            dF[c].re = 4.0 * BB(c).re; dF[c].im = 4.0 * BB(c).im;
            dB[c].re = 4.0 * FF(c).re; dB[c].im = 4.0 * FF(c).im;
        }
    } else if (ibk == 5) {
        /* (F0 + F1) B2 + (B0 B1) F2 */
        dF[0] = dF[1] = BB(2);
        dF[2] = zmul(BB(0), BB(1));
        dB[2].re = FF(0).re + FF(1).re; dB[2].im = FF(0).im + FF(1).im;
        dB[0] = zmul(BB(1), FF(2));
        dB[1] = zmul(BB(0), FF(2));
    }
    for (c = 0; c < nd; ++c) {
        cz a = {dF[c].re, -dF[c].im}, b = {dB[c].re, -dB[c].im};
        a = zmul(g, a); b = zmul(g, b);
        gF[(c * nw + q) * 2] += a.re; gF[(c * nw + q) * 2 + 1] += a.im;
        gB[(c * nw + q) * 2] += b.re; gB[(c * nw + q) * 2 + 1] += b.im;
    }
#undef FF
#undef BB
}

/* adjoint of x -> rfft(x) (length N, scaled by dt) for real x, restricted to t < nt:
 * xbar[t] = dt sum_q Re(g_q exp(+2 pi i q t / N)), g [nomega][2] gradient w.r.t. X */
static void rfft_adjoint(ora_ctx *C, const double *g, int nt, double *xbar)
{
    int N = C->N, nw = C->nomega, k;
    double *X = (double *)malloc((size_t)nw * 2 * sizeof(double)), *y = (double *)malloc(N * sizeof(double));
    for (k = 0; k < nw; ++k) {
        double h = (k == 0 || k == nw - 1) ? 1.0 : 0.5;
        X[2 * k] = g[2 * k] * h;
        X[2 * k + 1] = g[2 * k + 1] * h;
    }
    ora_irfft(N, 1.0, X, y);
    for (k = 0; k < nt; ++k) xbar[k] = y[k] * C->S.dt;
    free(X); free(y);
}

int ora_ad_backward(ora_ctx *C, int npts, const double *G, const double *fw, const double *bw,
                    const double *gbase, double *gG, double *gfw, double *gbw)
{
    const mck_setup_t *S = &C->S;
    int nd = C->nd, nw = C->nomega, nk = S->nkernel, R = S->nrec, N = C->N, nt = S->nt;
    int p, r, b, k, q, c, n, t;
    size_t fs = (size_t)nd * nw * 2;
    double *F = (double *)malloc((size_t)npts * fs * sizeof(double));
    double *B = (double *)malloc((size_t)npts * R * fs * sizeof(double));
    double *gF = (double *)calloc(fs, sizeof(double)), *gB = (double *)calloc(fs, sizeof(double));
    double *conv = (double *)malloc((size_t)nw * 2 * sizeof(double));
    double *ta = (double *)malloc((size_t)nw * 4 * sizeof(double)), *tb = ta + 2 * nw;
    double *xbar = (double *)malloc(N * sizeof(double)), *Xb = (double *)malloc((size_t)nw * 2 * sizeof(double));
    double *g1 = (double *)malloc((size_t)nw * 2 * sizeof(double));
    double *win = (double *)malloc(N * sizeof(double)), *one = (double *)malloc(nt * sizeof(double));
    double *tr = (double *)malloc(nt * sizeof(double)), *gcs = (double *)malloc(nt * sizeof(double));
    for (t = 0; t < nt; ++t) one[t] = 1.0;
    ora_taperandzeropad(nt, N, 10, 0, one, win);        /* the taper of fields_to_fd */
    memset(gG, 0, (size_t)S->nfilter * nw * 2 * sizeof(double));
    ad_spectra(C, npts, fw, bw, F, B);
    for (p = 0; p < npts; ++p) {
        memset(gF, 0, fs * sizeof(double));
        for (r = 0; r < R; ++r) {
            const double *Fp = F + p * fs, *Bp = B + ((size_t)p * R + r) * fs;
            memset(gB, 0, fs * sizeof(double));
            for (b = 0; b < 6; ++b) {
                if (!C->R[r].needs_bk[b]) continue;
                for (k = C->R[r].first; k <= C->R[r].last; ++k) {
                    ora_kernel *K = &C->K[k];
                    double gv = gbase[((size_t)p * nk + k) * 6 + b];
                    const double *rf = C->rest_f, *rb = C->rest_b;
                    double *gGf = gG + (size_t)K->filter * nw * 2;
                    if (!K->needs_bk[b] || gv == 0.0) continue;
                    ora_basekernel(b, C->full, nd, nw, Fp, Bp, conv);
                    ad_tf(C, G, K->filter, ta, tb);
                    /* value = -norm dt sum_n w_n x(ifirst + n) s(n): the Parseval integral of
                     * the zero-padded window (integrate_parseval_b_complex) and the
                     * trapezoidal rule (kernel.f90:482-491) are both this weighted sum */
                    memset(xbar, 0, N * sizeof(double));
                    for (n = 0; n < K->ncut; ++n) {
                        double w = 1.0;
                        if (S->int_scheme == MCK_INT_TRAPEZOIDAL) {
                            w = (n > 0 && n < K->ncut - 1) ? 1.0 : 0.0;
                            if (n == 0) w += 0.5;
                            if (n == K->ncut - 1) w += 0.5;
                        }
                        xbar[K->ifirst + n] = -gv * K->normalization * S->dt * w * K->seis_cut[n];
                    }
                    /* x = df c2r(X): dL/dX_q = df c_q sum_t xbar_t e^{-2 pi i q t/N}, the
                     * imaginary parts at DC and Nyquist are discarded by the c2r */
                    ora_rfft(N, 1.0, xbar, Xb);
                    for (q = 0; q < nw; ++q) {
                        double cq = (q == 0 || q == nw - 1) ? 1.0 : 2.0;
                        cz gX = {C->df * cq * Xb[2 * q], C->df * cq * Xb[2 * q + 1]};
                        cz cv = {conv[2 * q], conv[2 * q + 1]};
                        cz a = {ta[2 * q], ta[2 * q + 1]}, bb = {tb[2 * q], tb[2 * q + 1]};
                        cz ab, cb, ca, gc, gta, gtb, t1, t2;
                        if (q == 0 || q == nw - 1) gX.im = 0.0;
                        ab = zmul(a, bb); ab.im = -ab.im;
                        cb = zmul(cv, bb); cb.im = -cb.im;
                        ca = zmul(cv, a); ca.im = -ca.im;
                        gc = zmul(gX, ab);
                        gta = zmul(gX, cb);
                        gtb = zmul(gX, ca);
                        t1 = zmul(gta, (cz){rf[2 * q], -rf[2 * q + 1]});
                        t2 = zmul(gtb, (cz){rb[2 * q], -rb[2 * q + 1]});
                        gGf[2 * q] += t1.re + t2.re;
                        gGf[2 * q + 1] += t1.im + t2.im;
                        basekernel_adjoint(b, C->full, nd, nw, q, Fp, Bp, gc, gF, gB);
                    }
                }
            }
            /* B = rfft(taper bw) ts_bwd */
            for (c = 0; c < nd; ++c) {
                for (q = 0; q < nw; ++q) {
                    cz g = {gB[(c * nw + q) * 2], gB[(c * nw + q) * 2 + 1]};
                    if (!S->deconv_stf) g = zmul(g, (cz){C->ts_bwd[2 * q], -C->ts_bwd[2 * q + 1]});
                    g1[2 * q] = g.re; g1[2 * q + 1] = g.im;
                }
                rfft_adjoint(C, g1, nt, tr);
                for (t = 0; t < nt; ++t) gbw[(((size_t)p * R + r) * nd + c) * nt + t] = tr[t] * win[t];
            }
        }
        /* F = rfft(taper cumsum_trapezoidal(fw)) ts_fwd */
        for (c = 0; c < nd; ++c) {
            double *o = gfw + ((size_t)p * nd + c) * nt;
            for (q = 0; q < nw; ++q) {
                cz g = {gF[(c * nw + q) * 2], gF[(c * nw + q) * 2 + 1]};
                if (!S->deconv_stf) g = zmul(g, (cz){C->ts_fwd[2 * q], -C->ts_fwd[2 * q + 1]});
                g1[2 * q] = g.re; g1[2 * q + 1] = g.im;
            }
            rfft_adjoint(C, g1, nt, tr);
            /* cumsum_trapezoidal backwards: out_i = h c_i, c_0 = x_0, c_1 = c_0 + x_1,
             * c_i = c_{i-1} + x_{i-1} + x_i */
            for (t = 0; t < nt; ++t) { gcs[t] = tr[t] * win[t] * S->dt * 0.5; o[t] = 0.0; }
            for (t = nt - 1; t >= 2; --t) { gcs[t - 1] += gcs[t]; o[t - 1] += gcs[t]; o[t] += gcs[t]; }
            if (nt > 1) { gcs[0] += gcs[1]; o[1] += gcs[1]; }
            o[0] += gcs[0];
        }
    }
    free(F); free(B); free(gF); free(gB); free(conv); free(ta); free(xbar); free(Xb); free(g1);
    free(win); free(one); free(tr); free(gcs);
    return 0;
}
