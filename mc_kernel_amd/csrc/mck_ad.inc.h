// mck_ad.inc.h - the filter-convolve-window-integrate stage of the receiver loop
// (slave.f90:586-640) as a differentiable function of the filter responses G and the
// earth-frame strain traces, and its hand-written adjoint.
//
// Forward, per point p, receiver r and slot s = (window spectrum h, base kernel b):
//   F_c = rfft(taper cumsum_trapezoidal(fw_c)), B_c = rfft(taper bw_c)   (slave.f90:515-534, 592)
//   bk_b(q) = calc_basekernel(F, B)                                        (kernel.f90:514-570)
//   base_s = sum_{q=0}^{M} Re(bk_b(q) W_h(q) G_f(q)^2)
// where W_h is the window spectrum of the kernel without its filter (kernel.f90:181-385,
// 410-510, 946-1020): TF_fwd TF_bwd = G^2 x (deconvolution factors, time shifts).  The
// backward pass recomputes the FFTs from the traces instead of storing spectra.
// Gradients of complex inputs are d/dRe + i d/dIm of the real loss.
#pragma once

namespace mck {

// bins of thread tid: k = tid + 256 j for j < SPT (k < M), and k = M for j = SPT on one thread
template <int SPT>
__device__ inline bool ad_bin(int j, int M, int &k)
{
    const int tid = threadIdx.x;
    if (j < SPT) { k = tid + kStrainThreads * j; return k < M; }
    k = M;
    return tid == (M & (kStrainThreads - 1));
}

__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 cscale(float s, float2 a) { return make_float2(s * a.x, s * a.y); }
__device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// one complex value per base kernel lambda, mu, rho, a, b, c; members are addressed with
// compile-time indices only (bk_at<B>), so the values stay in registers
struct Bk6 {
    float2 v0, v1, v2, v3, v4, v5;
};
__device__ inline Bk6 bk_zero()
{
    const float2 z = make_float2(0.f, 0.f);
    return Bk6{z, z, z, z, z, z};
}
template <int B>
__device__ inline float2 &bk_at(Bk6 &a)
{
    static_assert(B >= 0 && B < 6, "base kernel index");
    if constexpr (B == 0) return a.v0;
    else if constexpr (B == 1) return a.v1;
    else if constexpr (B == 2) return a.v2;
    else if constexpr (B == 3) return a.v3;
    else if constexpr (B == 4) return a.v4;
    else return a.v5;
}
template <int B>
__device__ inline float2 bk_at(const Bk6 &a)
{
    return bk_at<B>(const_cast<Bk6 &>(a));
}

// calc_basekernel (kernel.f90:514-570) in the earth frame: trace only (ND 1) or full strain
// (ND 6, Voigt order 11 22 33 23 13 12)
template <int ND>
__device__ inline Bk6 ad_basekernels(const float2 (&f)[ND], const float2 (&b)[ND], int need)
{
    Bk6 bk = bk_zero();
    if constexpr (ND == 1) {
        const float2 pr = cmulf(f[0], b[0]);
        bk.v0 = cscale(-1.0f, pr);
        bk.v1 = cscale(-2.0f, pr);
    } else {
        static_assert(ND == 6, "full strain has six components");
        const float2 fs = cadd(cadd(f[0], f[1]), f[2]), bs = cadd(cadd(b[0], b[1]), b[2]);
        bk.v0 = cscale(-1.0f, cmulf(fs, bs));
        if (need & 2) {
            float2 mu = make_float2(0.f, 0.f);
#pragma unroll
            for (int c = 0; c < 6; ++c) mu = cadd(mu, cmulf(f[c], b[c]));
            bk.v1 = cscale(-2.0f, mu);
        }
        const float2 f01 = cadd(f[0], f[1]), b01 = cadd(b[0], b[1]);
        if (need & 8) bk.v3 = cmulf(f01, b01);
        if (need & 16) bk.v4 = cscale(4.0f, cadd(cmulf(b[4], f[4]), cmulf(b[3], f[3])));
        if (need & 32) bk.v5 = cadd(cmulf(f01, b[2]), cmulf(cmulf(b[0], b[1]), f[2]));   // as written
    }
    return bk;
}

// holomorphic derivatives of sum_b Y_b bk_b: dF_c = sum_b Y_b dbk_b/dF_c, dB_c likewise
template <int ND>
__device__ inline void ad_basekernels_adjoint(const float2 (&f)[ND], const float2 (&b)[ND], const Bk6 &Y,
                                              float2 (&dF)[ND], float2 (&dB)[ND])
{
    if constexpr (ND == 1) {
        const float2 s = cadd(cscale(-1.0f, Y.v0), cscale(-2.0f, Y.v1));
        dF[0] = cmulf(s, b[0]);
        dB[0] = cmulf(s, f[0]);
    } else {
        const float2 fs = cadd(cadd(f[0], f[1]), f[2]), bs = cadd(cadd(b[0], b[1]), b[2]);
        const float2 yl = cscale(-1.0f, Y.v0), ym = cscale(-2.0f, Y.v1);
        // lambda = -(F0 + F1 + F2)(B0 + B1 + B2), mu = -2 sum_c F_c B_c
        const float2 lf = cmulf(yl, bs), lb = cmulf(yl, fs);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            dF[c] = cmulf(ym, b[c]);
            dB[c] = cmulf(ym, f[c]);
            if (c < 3) { dF[c] = cadd(dF[c], lf); dB[c] = cadd(dB[c], lb); }
        }
        const float2 f01 = cadd(f[0], f[1]), b01 = cadd(b[0], b[1]);
        // a = (F0 + F1)(B0 + B1)
        const float2 af = cmulf(Y.v3, b01), ab = cmulf(Y.v3, f01);
        dF[0] = cadd(dF[0], af); dF[1] = cadd(dF[1], af);
        dB[0] = cadd(dB[0], ab); dB[1] = cadd(dB[1], ab);
        // b = 4 (B4 F4 + B3 F3)
        const float2 y4 = cscale(4.0f, Y.v4);
        dF[3] = cadd(dF[3], cmulf(y4, b[3])); dF[4] = cadd(dF[4], cmulf(y4, b[4]));
        dB[3] = cadd(dB[3], cmulf(y4, f[3])); dB[4] = cadd(dB[4], cmulf(y4, f[4]));
        // c = (F0 + F1) B2 + (B0 B1) F2
        const float2 cb2 = cmulf(Y.v5, b[2]);
        dF[0] = cadd(dF[0], cb2); dF[1] = cadd(dF[1], cb2);
        dF[2] = cadd(dF[2], cmulf(Y.v5, cmulf(b[0], b[1])));
        dB[2] = cadd(dB[2], cmulf(Y.v5, f01));
        const float2 yf2 = cmulf(Y.v5, f[2]);
        dB[0] = cadd(dB[0], cmulf(yf2, b[1]));
        dB[1] = cadd(dB[1], cmulf(yf2, b[0]));
    }
}

// Unnormalised length-M complex DFT (sign e^{-2 pi i}) of the data in `in` by Stockham
// passes in LDS: radix 2 first when log2 M is odd, then radix 4.  Returns the buffer
// holding the result; ends with a barrier.
__device__ inline float2 *lds_cfft(float2 *in, float2 *ob, int M, const float2 *Tq)
{
#pragma clang fp contract(fast)
    const int tid = threadIdx.x;
    const int lM = 31 - __clz(M);
    int Ns = 1, lNs = 0;
    if (lM & 1) {
        const int Mh = M >> 1;
        for (int j = tid; j < Mh; j += kStrainThreads) {
            const float2 v0 = in[j], v1 = in[j + Mh];
            ob[2 * j] = make_float2(v0.x + v1.x, v0.y + v1.y);
            ob[2 * j + 1] = make_float2(v0.x - v1.x, v0.y - v1.y);
        }
        __syncthreads();
        float2 *t = in; in = ob; ob = t;
        Ns = 2; lNs = 1;
    }
    const int Mq = M >> 2;
    for (; Ns < M; Ns <<= 2, lNs += 2) {
        const int lstride = lM - lNs - 2;
#pragma unroll 1
        for (int j = tid; j < Mq; j += kStrainThreads) {
            const int kk = j & (Ns - 1);
            const float2 v0 = in[j], v1 = in[j + Mq], v2 = in[j + 2 * Mq], v3 = in[j + 3 * Mq];
            const int ti = kk << (lstride + 1);
            // ti < N/4: w1 lies in the table's quadrant; w2 = w1^2, w3 = w1 w2
            const float2 w1 = Tq[tq_slot(ti)], w2 = cmulf(w1, w1), w3 = cmulf(w1, w2);
            const float2 a1 = cmulf(v1, w1), a2 = cmulf(v2, w2), a3 = cmulf(v3, w3);
            const float2 s02 = make_float2(v0.x + a2.x, v0.y + a2.y), d02 = make_float2(v0.x - a2.x, v0.y - a2.y);
            const float2 s13 = make_float2(a1.x + a3.x, a1.y + a3.y), d13 = make_float2(a1.x - a3.x, a1.y - a3.y);
            const int o = ((j >> lNs) << (lNs + 2)) + kk;
            ob[o] = make_float2(s02.x + s13.x, s02.y + s13.y);
            ob[o + Ns] = make_float2(d02.x + d13.y, d02.y - d13.x);
            ob[o + 2 * Ns] = make_float2(s02.x - s13.x, s02.y - s13.y);
            ob[o + 3 * Ns] = make_float2(d02.x - d13.y, d02.y + d13.x);
        }
        __syncthreads();
        float2 *t = in; in = ob; ob = t;
    }
    return in;
}

// Adjoint of X = dt DFT_N(x) (N = 2M, x real and zero beyond nt) for this thread's samples
// t0..t0+SPT-1: y_t = dt sum_{q=0}^{M} Re(g_q e^{+2 pi i q t/N}), g held by the bin owners
// (ad_bin).  As one real inverse transform: the half-weighted Hermitian spectrum is packed
// into a length-M complex sequence C_k = (Z_k + conj Z_{M-k}) + i (Z_k - conj Z_{M-k}) w_N^{-k}
// whose inverse DFT is y_{2n} + i y_{2n+1}; the inverse runs as conj(DFT(conj C)).
// Uses bufA, bufB and zM; ends with a barrier.
template <int SPT>
__device__ inline void rfft_adjoint(const float2 (&g)[SPT + 1], int t0, int M, float dt, float2 *bufA,
                                    float2 *bufB, float2 *zM, const float2 *Tq, float (&y)[SPT])
{
#pragma clang fp contract(fast)
    const int tid = threadIdx.x, lM = 31 - __clz(M);
#pragma unroll
    for (int j = 0; j <= SPT; ++j) {
        int k;
        if (!ad_bin<SPT>(j, M, k)) continue;
        float2 z = g[j];
        if (k == 0 || k == M) z.y = 0.0f;     // the imaginary parts at DC and Nyquist do not enter
        else z = cscale(0.5f, z);
        if (k < M) bufB[k] = z; else *zM = z;
    }
    __syncthreads();
    for (int k = tid; k < M; k += kStrainThreads) {
        const float2 zk = bufB[k], zc = cconj(k == 0 ? *zM : bufB[M - k]);
        const float2 e = cadd(zk, zc), d = make_float2(zk.x - zc.x, zk.y - zc.y);
        const float2 o = cmulf(d, cconj(twiddle(Tq, k, lM - 1)));
        bufA[k] = make_float2(e.x - o.y, -(e.y + o.x));        // conj(e + i o)
    }
    __syncthreads();
    const float2 *Rz = lds_cfft(bufA, bufB, M, Tq);
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int t = t0 + q, n = t >> 1;
        float v = 0.0f;
        if (n < M) v = (t & 1) ? -Rz[n].y : Rz[n].x;
        y[q] = v * dt;
    }
    __syncthreads();
}

// Adjoint of cumsum_trapezoidal (simple_routines.f90:314-341): out_t = h c_t with
// c_0 = x_0, c_1 = c_0 + x_1, c_t = c_{t-1} + x_{t-1} + x_t.  With S_t = h sum_{u>=t} g_u:
// dx_0 = S_0, dx_t = S_t + S_{t+1} (t >= 1, S_nt = 0).  A block-wide suffix scan.
template <int SPT>
__device__ inline void cumsum_trapezoidal_adjoint(float (&g)[SPT], int t0, int nt, float h, float *edge_x,
                                                  float *wave_tot)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    float gq[SPT], run = 0.0f;
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        gq[q] = (t0 + q < nt) ? g[q] * h : 0.0f;
        run += gq[q];
    }
    float incl = run;                  // sum over lanes >= lane
    for (int o = 1; o < 64; o <<= 1) {
        const float y = __shfl_down(incl, o, 64);
        if (lane + o < 64) incl += y;
    }
    if (lane == 0) wave_tot[wave] = incl;
    __syncthreads();
    float after = incl - run;
    for (int w2 = wave + 1; w2 < kStrainThreads / 64; ++w2) after += wave_tot[w2];
    float S[SPT], acc = after;
#pragma unroll
    for (int q = SPT - 1; q >= 0; --q) { acc += gq[q]; S[q] = acc; }
    edge_x[tid] = S[0];
    if (tid == 0) edge_x[kStrainThreads] = 0.0f;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int t = t0 + q;
        const float nx = q + 1 < SPT ? S[q + 1 < SPT ? q + 1 : 0] : edge_x[tid + 1];
        g[q] = t < nt ? S[q] + (t >= 1 ? nx : 0.0f) : 0.0f;
    }
    __syncthreads();
}

// W_h G_f^2 of every window spectrum: Hg [nh][nw]
__global__ void k_ad_hg(int nh, int nw, const float2 *W, const int *hfilt, const float2 *tf, float2 *Hg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nh * nw) return;
    const int h = i / nw, q = i - (i / nw) * nw;
    const float2 g = tf[(size_t)hfilt[h] * nw + q];
    Hg[i] = cmulf(W[i], cmulf(g, g));
}

// forward spectra of the traces fw [npts][ND][nt]: Fs [npts][ND][nw]
template <int SPT>
__global__ __launch_bounds__(kStrainThreads) void k_ad_fspec(PathConf pc, int npts, int nd, const float *fw,
                                                           const float2 *twq, float2 *Fs)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2 *bufA = reinterpret_cast<float2 *>(smem_raw);
    float2 *bufB = bufA + pc.M;
    float2 *Tq = bufA + 2 * pc.M;
    __shared__ double wave_tot[kStrainThreads / 64];
    __shared__ float edge_x[kStrainThreads + 1];
    const int p = blockIdx.x;
    if (p >= npts) return;
    load_twiddles(twq, pc.M, Tq);
    const int tid = threadIdx.x, nt = pc.nt, M = pc.M, nw = pc.nomega, t0 = tid * SPT;
    const float fdt = (float)pc.dt, h = (float)(pc.dt * 0.5);
    float tap[SPT];
    taper_factors(t0, nt, pc, tap);
    for (int c = 0; c < nd; ++c) {
        const float *src = fw + ((size_t)p * nd + c) * nt;
        float x[SPT];
#pragma unroll
        for (int q = 0; q < SPT; ++q) x[q] = t0 + q < nt ? src[t0 + q] : 0.0f;
        cumsum_trapezoidal(x, t0, nt, h, edge_x, wave_tot);
#pragma unroll
        for (int q = 0; q < SPT; ++q) x[q] *= tap[q];
        const float2 *Z = lds_fft<SPT, false>(x, t0, nt, M, bufA, bufB, Tq);
        float2 *dst = Fs + ((size_t)p * nd + c) * nw;
        for (int k = tid; k <= M; k += kStrainThreads) dst[k] = rfft_bin(Z, k, M, Tq, fdt);
        __syncthreads();
    }
}

// backward spectra of one pair's traces [ND][nt] into registers (bins of ad_bin)
template <int SPT, int ND>
__device__ inline void ad_bspec(const PathConf &pc, const float *src, const float (&tap)[SPT], float2 *bufA,
                                float2 *bufB, const float2 *Tq, float2 (&spec)[ND][SPT + 1])
{
    const int tid = threadIdx.x, nt = pc.nt, M = pc.M, t0 = tid * SPT;
    const float fdt = (float)pc.dt;
#pragma unroll
    for (int c = 0; c < ND; ++c) {
        float x[SPT];
#pragma unroll
        for (int q = 0; q < SPT; ++q) x[q] = t0 + q < nt ? src[(size_t)c * nt + t0 + q] * tap[q] : 0.0f;
        const float2 *Z = lds_fft<SPT, false>(x, t0, nt, M, bufA, bufB, Tq);
#pragma unroll
        for (int j = 0; j < SPT; ++j) {
            const int k = tid + kStrainThreads * j;
            spec[c][j] = k < M ? rfft_bin(Z, k, M, Tq, fdt) : make_float2(0.f, 0.f);
        }
        spec[c][SPT] = rfft_bin(Z, M, M, Tq, fdt);
        __syncthreads();
    }
}

// Forward: one workgroup per (point, receiver) pair q = p R + r with traces
// bw [npts][R][ND][nt]; base [npts][nk][6] (slots a kernel does not use are 0)
template <int SPT, int ND>
__global__ __launch_bounds__(kStrainThreads, SPT <= 8 ? 2 : 1) void k_ad_eval(
    PathConf pc, int npairs, int R, const float *bw, const float2 *twq, const float2 *Fs,
    const RecDesc *recs, const KernDesc *kern, const float2 *Hg, int nk, double *base)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2 *bufA = reinterpret_cast<float2 *>(smem_raw);
    float2 *bufB = bufA + pc.M;
    float2 *Tq = bufA + 2 * pc.M;
    __shared__ double red[kStrainThreads / 64][kMaxSlots];
    __shared__ double slotval[kMaxSlots];
    const int qp = blockIdx.x;
    if (qp >= npairs) return;
    const int p = qp / R, r = qp - (qp / R) * R;
    load_twiddles(twq, pc.M, Tq);
    const int tid = threadIdx.x, nt = pc.nt, M = pc.M, nw = pc.nomega, t0 = tid * SPT;
    const int lane = tid & 63, wave = tid >> 6;
    float tap[SPT];
    taper_factors(t0, nt, pc, tap);
    float2 spec[ND][SPT + 1];
    ad_bspec<SPT, ND>(pc, bw + (size_t)qp * ND * nt, tap, bufA, bufB, Tq, spec);

    // base kernels of every owned bin
    Bk6 bk[SPT + 1];
#pragma unroll
    for (int j = 0; j <= SPT; ++j) {
        int k;
        float2 f[ND], b[ND];
        const bool own = ad_bin<SPT>(j, M, k);
#pragma unroll
        for (int c = 0; c < ND; ++c) {
            f[c] = own ? Fs[((size_t)p * ND + c) * nw + k] : make_float2(0.f, 0.f);
            b[c] = own ? spec[c][j] : make_float2(0.f, 0.f);
        }
        bk[j] = ad_basekernels<ND>(f, b, pc.need_bk);
    }
    // slot sums: for every window spectrum (group) and every base kernel in use, then
    // stored for the group's slot of that base kernel
    const RecDesc &rd = recs[r];
    const int need = pc.need_bk;
    for (int g = 0; g < rd.ngroup; ++g) {
        const int h = rd.grp_h[g], s0 = rd.grp_first[g], ng = rd.grp_n[g];
        float2 hv[SPT + 1];
#pragma unroll
        for (int j = 0; j <= SPT; ++j) {
            int k;
            hv[j] = ad_bin<SPT>(j, M, k) ? Hg[(size_t)h * nw + k] : make_float2(0.f, 0.f);
        }
        auto slot_sum = [&](auto bsel) {
            constexpr int B = decltype(bsel)::value;
            if (!(need & (1 << B))) return;
            double a = 0.0;
#pragma unroll
            for (int j = 0; j <= SPT; ++j) {
                const float2 v = bk_at<B>(bk[j]);
                a += (double)__builtin_fmaf(v.x, hv[j].x, -(v.y * hv[j].y));
            }
            a = wave_sum_lane63(a);
            if (lane == 63)
                for (int t = 0; t < ng; ++t)
                    if (rd.slot_bk[s0 + t] == B) red[wave][s0 + t] = a;
        };
        slot_sum(std::integral_constant<int, 0>{});
        slot_sum(std::integral_constant<int, 1>{});
        slot_sum(std::integral_constant<int, 3>{});
        slot_sum(std::integral_constant<int, 4>{});
        slot_sum(std::integral_constant<int, 5>{});
    }
    __syncthreads();
    if (tid < rd.nslot) {
        double t = 0.0;
        for (int w = 0; w < kStrainThreads / 64; ++w) t += red[w][tid];
        slotval[tid] = t;
    }
    __syncthreads();
    for (int kk = tid; kk < rd.nkernel; kk += kStrainThreads) {
        const int k = rd.first_kernel + kk;
        const KernDesc &kd = kern[k];
        for (int b = 0; b < 6; ++b)
            base[((size_t)p * nk + k) * 6 + b] = kd.slot[b] >= 0 ? slotval[kd.slot[b]] : 0.0;
    }
}

// Backward in four stages, each with few registers, over the pairs q = p R + r of a
// chunk of points (intermediates in global memory):
//   k_ad_bstage  per pair: B = rfft(taper bw) -> Bs [q][ND][nw]
//   k_ad_cstage  per pair: per bin the adjoint of the contraction: slot weights
//                gw = sum of dL/dbase over the receiver's kernels, Y_b = sum gw H, U_f +=
//                W_h sum gw bk_b -> Ups [q][nF][nw]; holomorphic derivatives of bk ->
//                conj(dL/dB) over Bs, conj(dL/dF) -> dFs [q][ND][nw]
//   k_ad_bgrad   per pair: adjoint real FFT of dL/dB and the taper -> dL/dbw
//   k_ad_fgrad   per point: dL/dF and U summed over the receivers in order, adjoint real
//                FFT, taper and cumsum_trapezoidal adjoint -> dL/dfw; U -> Up [p][nF][nw]
template <int SPT>
__global__ __launch_bounds__(kStrainThreads) void k_ad_bstage(PathConf pc, int npairs, int nd, const float *bw,
                                                            const float2 *twq, float2 *Bs)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2 *bufA = reinterpret_cast<float2 *>(smem_raw);
    float2 *bufB = bufA + pc.M;
    float2 *Tq = bufA + 2 * pc.M;
    const int q = blockIdx.x;
    if (q >= npairs) return;
    load_twiddles(twq, pc.M, Tq);
    const int tid = threadIdx.x, nt = pc.nt, M = pc.M, nw = pc.nomega, t0 = tid * SPT;
    const float fdt = (float)pc.dt;
    float tap[SPT];
    taper_factors(t0, nt, pc, tap);
    for (int c = 0; c < nd; ++c) {
        const float *src = bw + ((size_t)q * nd + c) * nt;
        float x[SPT];
#pragma unroll
        for (int i = 0; i < SPT; ++i) x[i] = t0 + i < nt ? src[t0 + i] * tap[i] : 0.0f;
        const float2 *Z = lds_fft<SPT, false>(x, t0, nt, M, bufA, bufB, Tq);
        float2 *dst = Bs + ((size_t)q * nd + c) * nw;
        for (int k = tid; k <= M; k += kStrainThreads) dst[k] = rfft_bin(Z, k, M, Tq, fdt);
        __syncthreads();
    }
}

constexpr int kAdRegFilters = 8;

template <int SPT, int ND>
__global__ __launch_bounds__(kStrainThreads) void k_ad_cstage(
    PathConf pc, int npairs, int R, const float2 *Fs, float2 *Bs, const RecDesc *recs, const KernDesc *kern,
    const float2 *Hg, const float2 *W, const int *hfilt, int nF, int nk, const double *gbase, float2 *dFs,
    float2 *Ups)
{
    __shared__ float gw[kMaxSlots][6];      // [group][base kernel] slot weight of the receiver
    const int q = blockIdx.x;
    if (q >= npairs) return;
    const int p = q / R, r = q - (q / R) * R;
    const int tid = threadIdx.x, M = pc.M, nw = pc.nomega;
    const RecDesc &rd = recs[r];
    for (int i = tid; i < rd.ngroup * 6; i += kStrainThreads) gw[i / 6][i - (i / 6) * 6] = 0.0f;
    __syncthreads();
    if (tid < rd.ngroup) {
        for (int t = 0; t < rd.grp_n[tid]; ++t) {
            const int s = rd.grp_first[tid] + t;
            double gs = 0.0;
            for (int kk = 0; kk < rd.nkernel; ++kk) {
                const int k = rd.first_kernel + kk;
                for (int b = 0; b < 6; ++b)
                    if (kern[k].slot[b] == s) gs += gbase[((size_t)p * nk + k) * 6 + b];
            }
            gw[tid][rd.slot_bk[s]] = (float)gs;
        }
    }
    __syncthreads();
    float2 *U = Ups + (size_t)q * nF * nw;
    float2 *B = Bs + (size_t)q * ND * nw;
    float2 *dF = dFs + (size_t)q * ND * nw;
    const float2 *Fp = Fs + (size_t)p * ND * nw;
    // U parts of up to kAdRegFilters filters accumulate in registers (static indices),
    // more filters in the pair's U buffer
    const bool ureg = nF <= kAdRegFilters;
#pragma unroll 1
    for (int j = 0; j <= SPT; ++j) {
        int k;
        if (!ad_bin<SPT>(j, M, k)) continue;
        float2 ua[kAdRegFilters];
#pragma unroll
        for (int f = 0; f < kAdRegFilters; ++f) ua[f] = make_float2(0.f, 0.f);
        if (!ureg)
            for (int f = 0; f < nF; ++f) U[(size_t)f * nw + k] = make_float2(0.f, 0.f);
        float2 f[ND], b[ND], gF[ND], gB[ND];
#pragma unroll
        for (int c = 0; c < ND; ++c) { f[c] = Fp[(size_t)c * nw + k]; b[c] = B[(size_t)c * nw + k]; }
        const Bk6 bk = ad_basekernels<ND>(f, b, pc.need_bk);
        Bk6 Y = bk_zero();
        for (int g = 0; g < rd.ngroup; ++g) {
            const int h = rd.grp_h[g];
            const float2 hg = Hg[(size_t)h * nw + k], w = W[(size_t)h * nw + k];
            float2 Z = make_float2(0.f, 0.f);
            auto term = [&](auto bsel) {
                constexpr int Bi = decltype(bsel)::value;
                const float gs = gw[g][Bi];
                bk_at<Bi>(Y) = cadd(bk_at<Bi>(Y), cscale(gs, hg));
                Z = cadd(Z, cscale(gs, bk_at<Bi>(bk)));
            };
            term(std::integral_constant<int, 0>{});
            term(std::integral_constant<int, 1>{});
            term(std::integral_constant<int, 3>{});
            term(std::integral_constant<int, 4>{});
            term(std::integral_constant<int, 5>{});
            const float2 wz = cmulf(w, Z);
            const int fh = hfilt[h];
            if (ureg) {
#pragma unroll
                for (int ff = 0; ff < kAdRegFilters; ++ff)
                    if (ff == fh) ua[ff] = cadd(ua[ff], wz);
            } else {
                float2 *u = U + (size_t)fh * nw + k;
                *u = cadd(*u, wz);
            }
        }
        if (ureg)
#pragma unroll
            for (int ff = 0; ff < kAdRegFilters; ++ff)
                if (ff < nF) U[(size_t)ff * nw + k] = ua[ff];
        ad_basekernels_adjoint<ND>(f, b, Y, gF, gB);
#pragma unroll
        for (int c = 0; c < ND; ++c) {
            B[(size_t)c * nw + k] = cconj(gB[c]);
            dF[(size_t)c * nw + k] = cconj(gF[c]);
        }
    }
}

template <int SPT>
__global__ __launch_bounds__(kStrainThreads) void k_ad_bgrad(PathConf pc, int npairs, int nd, const float2 *Bs,
                                                           const float2 *twq, float *gbw)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2 *bufA = reinterpret_cast<float2 *>(smem_raw);
    float2 *bufB = bufA + pc.M;
    float2 *Tq = bufA + 2 * pc.M;
    __shared__ float2 zM;
    const int q = blockIdx.x;
    if (q >= npairs) return;
    load_twiddles(twq, pc.M, Tq);
    const int tid = threadIdx.x, nt = pc.nt, M = pc.M, nw = pc.nomega, t0 = tid * SPT;
    float tap[SPT];
    taper_factors(t0, nt, pc, tap);
    for (int c = 0; c < nd; ++c) {
        const float2 *src = Bs + ((size_t)q * nd + c) * nw;
        float2 g[SPT + 1];
#pragma unroll
        for (int j = 0; j <= SPT; ++j) {
            int k;
            g[j] = ad_bin<SPT>(j, M, k) ? src[k] : make_float2(0.f, 0.f);
        }
        float y[SPT];
        rfft_adjoint<SPT>(g, t0, M, (float)pc.dt, bufA, bufB, &zM, Tq, y);
        float *dst = gbw + ((size_t)q * nd + c) * nt;
#pragma unroll
        for (int i = 0; i < SPT; ++i)
            if (t0 + i < nt) dst[t0 + i] = y[i] * tap[i];
    }
}

template <int SPT>
__global__ __launch_bounds__(kStrainThreads) void k_ad_fgrad(PathConf pc, int npts, int R, int nd, int nF,
                                                           const float2 *dFs, const float2 *Ups,
                                                           const float2 *twq, float2 *Up, float *gfw)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2 *bufA = reinterpret_cast<float2 *>(smem_raw);
    float2 *bufB = bufA + pc.M;
    float2 *Tq = bufA + 2 * pc.M;
    __shared__ float2 zM;
    __shared__ float wave_tot[kStrainThreads / 64];
    __shared__ float edge_x[kStrainThreads + 1];
    const int p = blockIdx.x;
    if (p >= npts) return;
    load_twiddles(twq, pc.M, Tq);
    const int tid = threadIdx.x, nt = pc.nt, M = pc.M, nw = pc.nomega, t0 = tid * SPT;
    // U of the point: the receivers' parts in order
    for (int i = tid; i < nF * nw; i += kStrainThreads) {
        float2 a = make_float2(0.f, 0.f);
        for (int r = 0; r < R; ++r) a = cadd(a, Ups[((size_t)p * R + r) * nF * nw + i]);
        Up[(size_t)p * nF * nw + i] = a;
    }
    float tap[SPT];
    taper_factors(t0, nt, pc, tap);
    const float h = (float)(pc.dt * 0.5);
#pragma unroll 1
    for (int c = 0; c < nd; ++c) {
        float2 g[SPT + 1];
#pragma unroll
        for (int j = 0; j <= SPT; ++j) {
            int k;
            float2 a = make_float2(0.f, 0.f);
            if (ad_bin<SPT>(j, M, k))
                for (int r = 0; r < R; ++r) a = cadd(a, dFs[(((size_t)p * R + r) * nd + c) * nw + k]);
            g[j] = a;
        }
        float y[SPT];
        rfft_adjoint<SPT>(g, t0, M, (float)pc.dt, bufA, bufB, &zM, Tq, y);
#pragma unroll
        for (int i = 0; i < SPT; ++i) y[i] *= tap[i];
        cumsum_trapezoidal_adjoint<SPT>(y, t0, nt, h, edge_x, wave_tot);
        float *dst = gfw + ((size_t)p * nd + c) * nt;
#pragma unroll
        for (int i = 0; i < SPT; ++i)
            if (t0 + i < nt) dst[t0 + i] = y[i];
    }
}

// partial sums over point chunks of Up [npts][n]: part [nchunk][n] (FP64)
__global__ void k_ad_upart(int n, int npts, int chunk, const float2 *Up, double2 *part)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p0 = blockIdx.y * chunk, p1 = min(npts, p0 + chunk);
    double2 a = make_double2(0.0, 0.0);
    for (int p = p0; p < p1; ++p) {
        const float2 v = Up[(size_t)p * n + i];
        a.x += v.x; a.y += v.y;
    }
    part[(size_t)blockIdx.y * n + i] = a;
}

// gradient of the filter responses: d/dG_f(q) of Re(U_f G_f^2) = conj(2 G_f U_f)
__global__ void k_ad_tfgrad(int n, int nchunk, const double2 *part, const float2 *tf, float2 *gtf)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double ux = 0.0, uy = 0.0;
    for (int c = 0; c < nchunk; ++c) { ux += part[(size_t)c * n + i].x; uy += part[(size_t)c * n + i].y; }
    const double gx = tf[i].x, gy = tf[i].y;
    gtf[i] = make_float2((float)(2.0 * (gx * ux - gy * uy)), (float)(-2.0 * (gx * uy + gy * ux)));
}

}  // namespace mck
