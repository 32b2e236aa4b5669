// mck_mfma.inc.h - backward evaluation through element spectra and the matrix cores.
//
// The backward strain of a (point, receiver) pair is a linear map of the nodal traces of
// one backward element: B_c(t) = sum_j W_c,j u_j(t) over the element's traces j (the
// variables the receiver component reads x 25 GLL nodes; the weights hold the strain
// operator, the Lagrange interpolation at (xi, eta) and the azimuthal factors,
// readfields.f90:1359-1465, sem_derivatives.f90, spectral_basis.f90:181-314).  Taper, zero
// padding and the real FFT are linear too (fft.f90:375-413, 546-601), so
//     B_c(k) = sum_j W_c,j U_j(k),    U_j = rfft(taper . u_j)
// The nodal spectra U of every backward element a launch touches are computed once
// (k_elem_spectra, one workgroup per element); the pairs of an element then form their
// spectra as one matrix product [bins x traces] x [traces x pairs] on the matrix cores
// (fp32 operands as split fp16 products, kMfmaSplit), fused with the frequency-domain contraction of
// k_bwd_eval: base kernels per bin (kernel.f90:514-570) against the forward spectra of
// the pair's point, the filtered window integrals sum_k Re(bk_k H_k) accumulated in FP64
// (kernel.f90:410-510, 946-1020), and the physical kernels (kernel.f90:665-832).
//
// Pairs arrive sorted by (element, receiver); a run is the pairs of one element.
#pragma once

namespace mck {


// Work tables of element- and receiver-sorted pair keys (element * R + receiver):
// element runs (the elements whose spectra a launch needs) and groups of at most 16
// consecutive pairs with one key (one wave of k_bwd_mfma each).
// flags: eflag[i] = a new element starts at i, kflag[i] = a new key starts at i
__global__ void k_run_flags(int n, const int *keys, int R, int *eflag, int *kflag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    eflag[i] = (i == 0 || keys[i] / R != keys[i - 1] / R) ? 1 : 0;
    kflag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// from the inclusive scans of the flags (epos, kpos): element runs erun_first[u] (with
// erun_first[nerun] = n), erun_elem[u], *nerun; first item of every key run krun_first[v];
// gflag[i] = a group starts at i (every 16th item of a key run)
__global__ void k_run_table(int n, const int *keys, int R, const int *eflag, const int *epos, const int *kflag,
                            const int *kpos, int *erun_first, int *erun_elem, int *nerun, int *krun_first)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (eflag[i]) {
        const int u = epos[i] - 1;
        erun_first[u] = i;
        erun_elem[u] = keys[i] / R;
    }
    if (kflag[i]) krun_first[kpos[i] - 1] = i;
    if (i == n - 1) {
        *nerun = epos[i];
        erun_first[epos[i]] = n;
    }
}

__global__ void k_group_flags(int n, const int *kpos, const int *krun_first, int *gflag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    gflag[i] = ((i - krun_first[kpos[i] - 1]) & 15) == 0 ? 1 : 0;
}

// groups from the inclusive scan of gflag: group_first[g] (with group_first[ngroups] = n),
// the element run of the group, the first group of every element run (an element run
// starts a group; erun_group_first[nerun] = ngroups), *ngroups
__global__ void k_group_table(int n, const int *gflag, const int *gpos, const int *eflag, const int *epos,
                              int *group_first, int *group_erun, int *erun_group_first, int *ngroups)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (gflag[i]) {
        const int g = gpos[i] - 1;
        group_first[g] = i;
        group_erun[g] = epos[i] - 1;
        if (eflag[i]) erun_group_first[epos[i] - 1] = g;
    }
    if (i == n - 1) {
        *ngroups = gpos[i];
        group_first[gpos[i]] = n;
        erun_group_first[epos[i]] = gpos[i];
    }
}

// ------------------------------------------------------------------------------------
// k_elem_spectra: one workgroup per element run u0 + blockIdx.x; the tapered real
// spectra (x dt) of the element's J = nvar_bwd x 25 nodal traces, U[u - u0][j][k],
// k <= M, rows ustride apart (zero beyond M).  Computed in double precision and rounded
// once: the spectra of an element serve all of its pairs.
// ------------------------------------------------------------------------------------
template <int SPT>
__global__ __launch_bounds__(kStrainThreads) void k_elem_spectra(DevDb db, PathConf pc, int u0, int nrun,
                                                                 const int *run_elem, const double2 *twd,
                                                                 float2 *U)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double2 *bufA = reinterpret_cast<double2 *>(smem_raw);
    double2 *bufB = bufA + pc.M;
    const int u = u0 + (int)blockIdx.x;
    if (u >= nrun) return;
    const int e = run_elem[u];
    const float *ue = elem_u(db, e);
    if (!ue) return;
    const int tid = threadIdx.x, nt = pc.nt, t0 = tid * SPT;
    const int M = fft_size_of_spt<SPT>() ? fft_size_of_spt<SPT>() : pc.M;      // (= pc.M)
    const size_t blk = (size_t)25 * nt;
    float tap[SPT];
    taper_factors(t0, nt, pc, tap);
    const int J = pc.nvar_bwd * 25;
    const bool full_vec = (t0 + SPT <= nt) && ((nt & 3) == 0) && (SPT % 4) == 0;
    float2 *Uu = U + (size_t)(u - u0) * J * pc.ustride;
    for (int j = 0; j < J; ++j) {
        const float *src = ue + (size_t)pc.bwd_vars[j / 25] * blk + (size_t)(j % 25) * nt + t0;
        double x[SPT];
        if (full_vec) {
#pragma unroll
            for (int q = 0; q < SPT; q += 4) {
                const float4 v = *reinterpret_cast<const float4 *>(src + q);
                x[q] = (double)v.x * tap[q]; x[q + 1] = (double)v.y * tap[q + 1];
                x[q + 2] = (double)v.z * tap[q + 2]; x[q + 3] = (double)v.w * tap[q + 3];
            }
        } else {
#pragma unroll
            for (int q = 0; q < SPT; ++q) x[q] = (t0 + q < nt) ? (double)src[q] * tap[q] : 0.0;
        }
        const double2 *Z = lds_fft_d(x, t0, nt, M, bufA, bufB, twd);
        float2 *dst = Uu + (size_t)j * pc.ustride;
        for (int k = tid; k < pc.ustride; k += kStrainThreads)
            dst[k] = k <= M ? rfft_bin_d(Z, k, M, twd, pc.dt) : make_float2(0.0f, 0.0f);
        __syncthreads();
    }
}

constexpr int kMfmaPairs = 16;         // pairs of one wave (the matrix-core columns)

// The contraction U W of the element spectra with the pair weights on the fp16 matrix
// cores (v_mfma_f32_16x16x32_f16, 16 times the fp32 products' rate): every fp32 operand x
// is split into x_hi = fp16(x) and x_lo = fp16(x - x_hi), 22 of its 24 significant bits,
// and U W = U_hi W_hi + U_hi W_lo + U_lo W_hi (fp16 products are exact in fp32; the
// dropped U_lo W_lo is below 2^-22 of a product).  The operands are scaled by powers of
// two into fp16's range first (W per pair, U per tile of 16 bins, to below 2^14) and the
// sums scaled back.  K: the traces in blocks of 32 (8 per lane).  false: the fp32
// products (v_mfma_f32_16x16x4f32, K in steps of 4).
constexpr bool kMfmaSplit = true;
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

// A operands per lane for J traces (trace mfma_trace(s, g) in slot s of lane group g)
__host__ __device__ constexpr int mfma_ks(int J) { return kMfmaSplit ? 8 * ((J + 31) / 32) : (J + 3) / 4; }
__device__ __forceinline__ int mfma_trace(int s, int g)
{
    return kMfmaSplit ? 32 * (s >> 3) + 8 * g + (s & 7) : 4 * s + g;
}

// the largest of a non-negative value over the wave (all 64 lanes active): within each
// row of 16 lanes by DPP (lane swaps 1 and 2, rotations by 4 and 8), then across the 4
// rows on the bit patterns (non-negative floats order as unsigned integers)
__device__ __forceinline__ float mfma_wave_max(float x)
{
    int v = __builtin_bit_cast(int, x);
    v = __builtin_bit_cast(int, fmaxf(__builtin_bit_cast(float, v),
                                      __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false))));
    v = __builtin_bit_cast(int, fmaxf(__builtin_bit_cast(float, v),
                                      __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false))));
    v = __builtin_bit_cast(int, fmaxf(__builtin_bit_cast(float, v),
                                      __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, v, 0x124, 0xF, 0xF, false))));
    v = __builtin_bit_cast(int, fmaxf(__builtin_bit_cast(float, v),
                                      __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false))));
    const unsigned r0 = (unsigned)__builtin_amdgcn_readlane(v, 0), r1 = (unsigned)__builtin_amdgcn_readlane(v, 16);
    const unsigned r2 = (unsigned)__builtin_amdgcn_readlane(v, 32), r3 = (unsigned)__builtin_amdgcn_readlane(v, 48);
    const unsigned m = r0 > r1 ? r0 : r1, n = r2 > r3 ? r2 : r3;
    return __builtin_bit_cast(float, m > n ? m : n);
}

// x (scaled into fp16's range) = hi + lo to 22 bits
__device__ __forceinline__ void mfma_split(float x, _Float16 &hi, _Float16 &lo)
{
    hi = (_Float16)x;
    lo = (_Float16)(x - (float)hi);
}

// LDS bytes of one wave's own arrays: the B operands [KS][NDB][64] of its 16 pairs and the
// FP64 slot sums of its npos class positions (npos / 16 blocks of 16 x 16)
__host__ __device__ constexpr size_t mfma_wave_lds(int ks, int ndb, int npos)
{
    return (size_t)ks * ndb * 64 * sizeof(float) + (size_t)npos * kMfmaPairs * sizeof(double);
}

// the slot sum of class position pos for pair jp, from the blocks of matrix-core results
// (block pos / 16: D[pair][pos % 16] in lane pos % 16 + 16 (pair % 4), element pair / 4)
__device__ __forceinline__ double mfma_slot_sum(const double *sAcc, int pos, int jp)
{
    return sAcc[(size_t)(pos >> 4) * 256 + 4 * ((pos & 15) + 16 * (jp & 3)) + (jp >> 2)];
}

// LDS bytes of k_bwd_mfma_wg with w waves: two 16-bin tiles of the element spectra
// [4 KS][16] and every wave's own arrays
__host__ __device__ constexpr size_t mfma_wg_lds(int ks, int ndb, int npos, int w)
{
    return 2 * (size_t)ks * 4 * 16 * sizeof(float2) + (size_t)w * mfma_wave_lds(ks, ndb, npos);
}

// One wave's group: its pairs (lane l holds pair l % 16), their receiver and the
// per-pair coefficients F.B = f^T Amu b (earth frame through the rotations of the point,
// Qf, and of the pair, Qb) and the traces uf.f, ub.b
template <int NDF, int NDB>
struct MfmaGroup {
    int g0, n, r, p;
    int ew;                             // the pair's weights scaled by 2^ew (split products)
    bool valid;
    float uf[NDF], ub[NDB], Amu[NDF][NDB];
};

// Loads group grp: its B operands into sW (fp32 products: sW[(s NDB + c) 64 + l'] = W of
// pair l' % 16, component c, trace 4 s + l' / 16; split products: the v8h at
// ((kb NDB + c) 2 + part) 64 + l' holds part hi / lo of 2^ew W of pair l' % 16, component c,
// traces 32 kb + 8 (l' / 16) + q, q = 0..7; zero past the group or the traces), its slot
// sums zeroed, its coefficients.  Called by the 64 lanes of the wave; the caller
// synchronises before sW or sAcc is read.
template <int NDF, int NDB, int NVB>
__device__ __forceinline__ void mfma_group_setup(MfmaGroup<NDF, NDB> &G, int grp, const int *group_first,
                                                 const int *order, int R, const float *Wb, const float *Qb,
                                                 const float *Qf, const PathConf &pc, int npos, float *sW,
                                                 double *sAcc, int l)
{
    constexpr int J = NVB * 25, KS = mfma_ks(J);
    const int jp = l & 15;
    G.g0 = group_first[grp];
    G.n = group_first[grp + 1] - G.g0;
    if constexpr (kMfmaSplit) {
        // the pair's scale: its largest weight to below 2^14 (lanes jp, jp + 16, .. share it)
        const float *wp = Wb + (size_t)(G.g0 + jp) * NDB * J;
        float mx = 0.0f;
        if (jp < G.n)
            for (int i = l >> 4; i < NDB * J; i += 4) mx = fmaxf(mx, fabsf(wp[i]));
        mx = fmaxf(mx, __shfl_xor(mx, 16, 64));
        mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
        int e;
        (void)frexpf(mx, &e);
        G.ew = 14 - e;
        v8h *sWh = reinterpret_cast<v8h *>(sW);
        const int g = l >> 4;
        for (int kc = 0; kc < (KS / 8) * NDB; ++kc) {          // kc = kb NDB + c
            const int kb = kc / NDB, c = kc - kb * NDB;
            v8h hi, lo;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int j = 32 * kb + 8 * g + q;
                const float w = (jp < G.n && j < J) ? ldexpf(wp[c * J + j], G.ew) : 0.0f;
                _Float16 h, o;
                mfma_split(w, h, o);
                hi[q] = h;
                lo[q] = o;
            }
            sWh[(2 * kc) * 64 + l] = hi;
            sWh[(2 * kc + 1) * 64 + l] = lo;
        }
    } else {
        G.ew = 0;
        for (int idx = l; idx < KS * NDB * 64; idx += 64) {
            const int s = idx / (NDB * 64), rem = idx - s * (NDB * 64), c = rem >> 6, ll = rem & 63;
            const int pj = ll & 15, j = 4 * s + (ll >> 4);
            float w = 0.0f;
            if (pj < G.n && j < J) w = Wb[(size_t)(G.g0 + pj) * NDB * J + c * J + j];
            sW[idx] = w;
        }
    }
    for (int s = l; s < npos * kMfmaPairs; s += 64) sAcc[s] = 0.0;
    G.valid = jp < G.n;
    // the group's pairs share the receiver
    G.r = order[G.g0] % R;
    G.p = order[G.g0 + (G.valid ? jp : 0)] / R;                 // (no pair: the group's first point)
    if (NDF > 1) {
        float qf[6][NDF], qb[6][NDB];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
#pragma unroll
            for (int m = 0; m < NDF; ++m) qf[i][m] = G.valid ? Qf[(size_t)G.p * 36 + i * 6 + m] : 0.0f;
#pragma unroll
            for (int c = 0; c < NDB; ++c) qb[i][c] = G.valid ? Qb[(size_t)(G.g0 + jp) * 36 + i * 6 + c] : 0.0f;
        }
#pragma unroll
        for (int m = 0; m < NDF; ++m) {
            // the trace is invariant under the rotations: exactly 1 for the diagonal
            // components (ss, pp, zz), 0 for the others
            G.uf[m] = pc.comps_f[m] < 3 ? 1.0f : 0.0f;
#pragma unroll
            for (int c = 0; c < NDB; ++c) {
                float a = 0.0f;
#pragma unroll
                for (int i = 0; i < 6; ++i) a += qf[i][m] * qb[i][c];
                G.Amu[m][c] = a;
            }
        }
#pragma unroll
        for (int c = 0; c < NDB; ++c) G.ub[c] = pc.comps_b[c] < 3 ? 1.0f : 0.0f;
    } else {
        G.uf[0] = 1.0f; G.ub[0] = 1.0f; G.Amu[0][0] = 1.0f;
    }
}

// The slot sums of one tile on the FP64 matrix cores: for every block of 16 class
// positions (one class, columns c = 0..15: window spectrum cls_h[16 b + c]),
// D[pair][c] += sum over the tile's 16 bins of Re(bk H) = bk.x H.x + bk.y (-H.y); the
// fp32 base kernels and window spectra multiply exactly in FP64 and the products are
// summed in FP64.  This lane holds bk of pair l % 16 at bins kb .. kb + 3 (the A
// operands: pair l % 16, k = l / 16) and loads column l % 16's spectra at the same bins
// (the B operands).  D of block b: sAcc[b 256 + 4 l + r] = D[l / 16 + 4 r][l % 16] (the
// FP64 product's result layout).
// blocks of slot sums kept in registers across the tiles (the rest in LDS): 2 where the
// registers still fit two waves per SIMD (4 or 6 forward components, 2 backward variables)
template <int NDF, int NVB>
constexpr int mfma_dreg() { return (NDF >= 4 && NVB == 2) ? 2 : 0; }

// block b of the tile's slot sums: d + the block's products
__device__ __forceinline__ v4d mfma_window_block(const float2 (&bk0)[4], const float2 (&bk1)[4], int b, int l,
                                                 int kb, int fs, const RecDesc &rd, const float2 *Hp, v4d d)
{
    const int p0 = b << 4;
    const int cls = p0 < rd.cls_end[0] ? 0 : (p0 < rd.cls_end[1] ? 1 : 2);
    const float4 *src = reinterpret_cast<const float4 *>(Hp + (size_t)rd.cls_h[p0 + (l & 15)] * fs + kb);
    const float4 q0 = src[0], q1 = src[1];
    const double2 h[4] = {make_double2(q0.x, q0.y), make_double2(q0.z, q0.w), make_double2(q1.x, q1.y),
                          make_double2(q1.z, q1.w)};
    double ax[4], ay[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        if (cls == 0) {
            ax[v] = bk0[v].x; ay[v] = bk0[v].y;
        } else if (cls == 1) {
            ax[v] = bk1[v].x; ay[v] = bk1[v].y;
        } else {                                                // mu - 2 lambda
            ax[v] = (double)bk1[v].x - 2.0 * (double)bk0[v].x;
            ay[v] = (double)bk1[v].y - 2.0 * (double)bk0[v].y;
        }
    }
    v4d dy = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        d = __builtin_amdgcn_mfma_f64_16x16x4f64(ax[v], h[v].x, d, 0, 0, 0);
        dy = __builtin_amdgcn_mfma_f64_16x16x4f64(ay[v], h[v].y, dy, 0, 0, 0);
    }
    return d + dy;
}

// the register blocks into sAcc (before the lanes read each other's sums)
template <int DR>
__device__ __forceinline__ void mfma_store_dreg(const RecDesc &rd, int l, double *sAcc, const v4d (&dreg)[DR ? DR : 1])
{
    const int nblk = rd.cls_end[2] >> 4;
#pragma unroll
    for (int b = 0; b < DR; ++b)
        if (b < nblk) *reinterpret_cast<v4d *>(sAcc + (size_t)b * 256 + 4 * l) = dreg[b];
}

// dreg: blocks 0 .. DR - 1 (registers across the tiles); the others in sAcc
template <int DR>
__device__ __forceinline__ void mfma_window_sums(const float2 (&bk0)[4], const float2 (&bk1)[4], int l, int kb,
                                                 int fs, const RecDesc &rd, const float2 *Hp, double *sAcc,
                                                 v4d (&dreg)[DR ? DR : 1])
{
    const int nblk = rd.cls_end[2] >> 4;
#pragma unroll
    for (int b = 0; b < DR; ++b)
        if (b < nblk) dreg[b] = mfma_window_block(bk0, bk1, b, l, kb, fs, rd, Hp, dreg[b]);
    for (int b = DR; b < nblk; ++b) {
        v4d *acc = reinterpret_cast<v4d *>(sAcc + (size_t)b * 256 + 4 * l);
        *acc = mfma_window_block(bk0, bk1, b, l, kb, fs, rd, Hp, *acc);
    }
}

// One 16-bin tile t of a group, given its A operands a[s] (lane l: trace 4 s + l / 16 at
// bin 16 t + l % 16; bins past nomega and traces past J zero): the matrix cores give every
// pair's NDB carried backward components at the tile's 16 bins (real and imaginary parts
// as two products); lane l then holds pair l % 16 at bins kb = 16 t + 4 (l / 16) .. + 3 and
// forms the base kernels with the forward spectra of the pair's point; then the slot sums
// Re(bk H) with the receiver's window spectra on the FP64 matrix cores.
// this lane's forward spectra of tile t (pair l % 16's point, bins kb .. kb + 3).  A lane
// without a pair loads the group's first point (no branch): its weights are zero, so its
// base kernels and rows of the slot sums are zero, and nothing reads them.
template <int NDF, int NDB>
__device__ __forceinline__ void mfma_load_f(const MfmaGroup<NDF, NDB> &G, const float2 *Fp, int l, int t, int fs,
                                            float4 (&f4)[NDF][2])
{
    const int kb = (t << 4) + 4 * (l >> 4);
#pragma unroll
    for (int m = 0; m < NDF; ++m) {
        const float4 *src = reinterpret_cast<const float4 *>(Fp + (size_t)m * fs + kb);
        f4[m][0] = src[0];
        f4[m][1] = src[1];
    }
}

// this lane's element spectra of tile t (the A operands: trace mfma_trace(s, l / 16), bin
// 16 t + l % 16; traces past J load trace J - 1, whose B operands there are zero)
template <int KS, int J>
__device__ __forceinline__ void mfma_load_u(const float2 *Ue, int us, int l, int t, float2 (&a)[KS])
{
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int j = min(mfma_trace(s, l >> 4), J - 1);
        a[s] = Ue[(size_t)j * us + (t << 4) + (l & 15)];
    }
}

template <int NDF, int NDB, int KS, int DR>
__device__ __forceinline__ void mfma_tile(const MfmaGroup<NDF, NDB> &G, const float2 (&a)[KS],
                                          const float4 (&f4)[NDF][2], const float *sW, double *sAcc, int l, int t,
                                          const PathConf &pc, const RecDesc &rd, const float2 *Hp,
                                          v4d (&dreg)[DR ? DR : 1])
{
    const int bq = l >> 4, kb = (t << 4) + 4 * bq, fs = pc.fstride;
    v4f accR[NDB], accI[NDB];
#pragma unroll
    for (int c = 0; c < NDB; ++c) {
        accR[c] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
        accI[c] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    }
    if constexpr (kMfmaSplit) {
        // the tile's scale: its largest |U| to below 2^14
        float mx = 0.0f;
#pragma unroll
        for (int s = 0; s < KS; ++s) mx = fmaxf(mx, fmaxf(fabsf(a[s].x), fabsf(a[s].y)));
        mx = mfma_wave_max(mx);
        int e;
        (void)frexpf(mx, &e);
        const int eu = __builtin_amdgcn_readfirstlane(14 - e);
        const v8h *sWh = reinterpret_cast<const v8h *>(sW);
#pragma unroll
        for (int kb = 0; kb < KS / 8; ++kb) {
            v8h rh, rl, ih, il;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                _Float16 h, o;
                mfma_split(ldexpf(a[8 * kb + q].x, eu), h, o);
                rh[q] = h;
                rl[q] = o;
                mfma_split(ldexpf(a[8 * kb + q].y, eu), h, o);
                ih[q] = h;
                il[q] = o;
            }
#pragma unroll
            for (int c = 0; c < NDB; ++c) {
                const v8h wh = sWh[(2 * (kb * NDB + c)) * 64 + l], wl = sWh[(2 * (kb * NDB + c) + 1) * 64 + l];
                accR[c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(rh, wh, accR[c], 0, 0, 0);
                accI[c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ih, wh, accI[c], 0, 0, 0);
                accR[c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(rh, wl, accR[c], 0, 0, 0);
                accI[c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ih, wl, accI[c], 0, 0, 0);
                accR[c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(rl, wh, accR[c], 0, 0, 0);
                accI[c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(il, wh, accI[c], 0, 0, 0);
            }
        }
        // back to the spectra's scale (column l % 16: this lane's pair)
        const int es = -(eu + G.ew);
#pragma unroll
        for (int c = 0; c < NDB; ++c)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                accR[c][v] = ldexpf(accR[c][v], es);
                accI[c][v] = ldexpf(accI[c][v], es);
            }
    } else {
#pragma unroll
        for (int s = 0; s < KS; ++s) {
#pragma unroll
            for (int c = 0; c < NDB; ++c) {
                const float w = sW[(s * NDB + c) * 64 + l];
                accR[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s].x, w, accR[c], 0, 0, 0);
                accI[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s].y, w, accI[c], 0, 0, 0);
            }
        }
    }
    // base kernels of this lane's pair at bins kb + v
    float2 bk0[4], bk1[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
#pragma clang fp contract(fast)
        float2 f[NDF], b[NDB];
#pragma unroll
        for (int m = 0; m < NDF; ++m)
            f[m] = (v & 1) ? make_float2(f4[m][v >> 1].z, f4[m][v >> 1].w) : make_float2(f4[m][v >> 1].x, f4[m][v >> 1].y);
#pragma unroll
        for (int c = 0; c < NDB; ++c) b[c] = make_float2(accR[c][v], accI[c][v]);
        if (NDF == 1) {
            const float2 pr = cmulf(f[0], b[0]);
            bk0[v] = make_float2(-pr.x, -pr.y);
            bk1[v] = make_float2(-2.0f * pr.x, -2.0f * pr.y);
        } else {
            float2 trf = make_float2(0.f, 0.f), trb = trf;
#pragma unroll
            for (int m = 0; m < NDF; ++m) { trf.x += G.uf[m] * f[m].x; trf.y += G.uf[m] * f[m].y; }
#pragma unroll
            for (int c = 0; c < NDB; ++c) { trb.x += G.ub[c] * b[c].x; trb.y += G.ub[c] * b[c].y; }
            const float2 lam = cmulf(trf, trb);
            bk0[v] = make_float2(-lam.x, -lam.y);
            bk1[v] = make_float2(0.f, 0.f);
            if (pc.need_bk & 2) {
                float2 mu = make_float2(0.f, 0.f);
#pragma unroll
                for (int m = 0; m < NDF; ++m) {
                    float2 gm = make_float2(0.f, 0.f);
#pragma unroll
                    for (int c = 0; c < NDB; ++c) { gm.x += G.Amu[m][c] * b[c].x; gm.y += G.Amu[m][c] * b[c].y; }
                    const float2 tt = cmulf(f[m], gm);
                    mu.x += tt.x; mu.y += tt.y;
                }
                bk1[v] = make_float2(-2.0f * mu.x, -2.0f * mu.y);
            }
        }
    }
    mfma_window_sums<DR>(bk0, bk1, l, kb, fs, rd, Hp, sAcc, dreg);
}

// Physical kernels (kernel.f90:665-832) of the group's pairs with the model at the point,
// after the damping of every receiver processed in this step (readfields.f90:3184-3204);
// lanes 0..15 (lane l: pair l).  phys: [point][nk]; base_out (optional): [point][nk][6]
template <int NDF, int NDB>
__device__ __forceinline__ void mfma_finish(const MfmaGroup<NDF, NDB> &G, int l, const double *sAcc,
                                            const PathConf &pc, const RecDesc &rd, const KernDesc *kern,
                                            const int *model_idx, const float *model, int ndepth, int relative,
                                            int nk, double *phys, double *base_out, const double *xyz,
                                            const uint8_t *active, int P, int R, const double *rec_xyz)
{
    if (l >= kMfmaPairs || !G.valid) return;
    const int p = G.p, r = G.r, jp = l;
    const int mi = model_idx[p];
    const float vp = model[0 * ndepth + mi], vs = model[1 * ndepth + mi], rho_f = model[2 * ndepth + mi];
    const double rho = rho_f;
    double damp = 1.0;
    if (pc.damp_radius > 0.0 && active) {
        const double x0 = xyz[3 * (size_t)p], x1 = xyz[3 * (size_t)p + 1], x2 = xyz[3 * (size_t)p + 2];
        const double ds = sqrt((x0 - pc.src_r[0]) * (x0 - pc.src_r[0]) + (x1 - pc.src_r[1]) * (x1 - pc.src_r[1])
                               + (x2 - pc.src_r[2]) * (x2 - pc.src_r[2]));
        const double fsrc = ds < pc.damp_radius ? ds / pc.damp_radius : 1.0;
        for (int r2 = r; r2 < R; ++r2) {
            if (!active[(size_t)(p / P) * R + r2]) continue;
            const double *c = rec_xyz + 3 * r2;
            const double dr = sqrt((x0 - c[0]) * (x0 - c[0]) + (x1 - c[1]) * (x1 - c[1]) + (x2 - c[2]) * (x2 - c[2]));
            damp = damp * (dr < pc.damp_radius ? dr / pc.damp_radius : 1.0) * fsrc;
        }
    }
    for (int kk = 0; kk < rd.nkernel; ++kk) {
        const int k = rd.first_kernel + kk;
        const KernDesc &kd = kern[k];
        double b[kNumBk];
#pragma unroll
        for (int qq = 0; qq < kNumBk; ++qq)
            b[qq] = kd.slot[qq] >= 0 ? mfma_slot_sum(sAcc, rd.cls_pos[kd.slot[qq]], jp) * damp : 0.0;
        if (base_out) {
            for (int qq = 0; qq < 6; ++qq) base_out[((size_t)p * nk + k) * 6 + qq] = b[qq];
            continue;
        }
        double v = 0.0;
        switch (kd.param) {
        case 0: v = relative ? b[0] * (double)(rho_f * (vp * vp - 2 * (vs * vs))) : b[0]; break;
        case 3: v = relative ? b[1] * (double)(rho_f * (vs * vs)) : b[1]; break;
        case 1: v = relative ? 2.0 * rho * (double)(vp * vp) * b[0] : 2.0 * rho * (double)vp * b[0]; break;
        case 2: v = (relative ? 2.0 * rho * (double)(vs * vs) : 2.0 * rho * (double)vs) * b[6]; break;
        default: v = 0.0;
        }
        phys[(size_t)p * nk + k] = v;
    }
}

// ------------------------------------------------------------------------------------
// k_bwd_mfma: one wave per group of at most 16 pairs of one (element, receiver), groups of
// element runs u0 .. u0 + ubatch - 1 (the others are left to another launch); the wave
// reads its element's spectra tile by tile from memory.  Only the lambda and mu base
// kernels (and mu - 2 lambda) are formed: the host sends kernel sets with anisotropic
// parameters to k_bwd_eval.  Hp: window spectra, rows fstride apart, zero past nomega;
// F: forward spectra, rows fstride apart, zero past nomega.  xcd: groups in
// XCD-contiguous ranges (the groups of an element on one XCD).
// ------------------------------------------------------------------------------------
template <int NDF, int NDB, int NVB>
__global__ __launch_bounds__(64, NDB <= 4 ? 2 : 1) void k_bwd_mfma(
    PathConf pc, int u0, int ubatch, int ngroups, const int *group_first, const int *group_erun,
    const float2 *U, const int *order, int R, const float *Wb, const float *Qb, const float *Qf,
    const float2 *F, const RecDesc *recs, const KernDesc *kern, const float2 *Hp, int npos,
    const int *model_idx, const float *model, int ndepth, int relative, int nk, double *phys, double *base_out,
    const double *xyz, const uint8_t *active, int P, const double *rec_xyz, int xcd)
{
    constexpr int J = NVB * 25, KS = mfma_ks(J);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int grp = xcd ? xcd_block_n(ngroups) : ((int)blockIdx.x < ngroups ? (int)blockIdx.x : -1);
    if (grp < 0) return;
    const int eu = group_erun[grp];
    if (eu < u0 || eu >= u0 + ubatch) return;
    const int l = threadIdx.x;
    float *sW = reinterpret_cast<float *>(smem_raw);
    double *sAcc = reinterpret_cast<double *>(sW + KS * NDB * 64);
    MfmaGroup<NDF, NDB> G;
    mfma_group_setup<NDF, NDB, NVB>(G, grp, group_first, order, R, Wb, Qb, Qf, pc, npos, sW, sAcc, l);
    __syncthreads();
    const float2 *Ue = U + (size_t)(eu - u0) * J * pc.ustride;
    const RecDesc &rd = recs[G.r];
    const float2 *Fp = F + (size_t)G.p * NDF * pc.fstride;
    const int ntile = (pc.nomega + 15) >> 4;
    constexpr int DR = mfma_dreg<NDF, NVB>();
    v4d dreg[DR ? DR : 1];
#pragma unroll
    for (int b = 0; b < (DR ? DR : 1); ++b) dreg[b] = v4d{0.0, 0.0, 0.0, 0.0};
    if constexpr (NDF <= 4 && NDB <= 4) {
        // the next tile's element and forward spectra are loaded while this one is
        // multiplied (the registers for it fit two waves per SIMD up to 4 components)
        float2 an[KS];
        float4 fn[NDF][2];
        mfma_load_u<KS, J>(Ue, pc.ustride, l, 0, an);
        mfma_load_f<NDF, NDB>(G, Fp, l, 0, pc.fstride, fn);
        for (int t = 0; t < ntile; ++t) {
            float2 a[KS];
            float4 f4[NDF][2];
#pragma unroll
            for (int s = 0; s < KS; ++s) a[s] = an[s];
#pragma unroll
            for (int m = 0; m < NDF; ++m) { f4[m][0] = fn[m][0]; f4[m][1] = fn[m][1]; }
            const int tn = t + 1 < ntile ? t + 1 : t;
            mfma_load_u<KS, J>(Ue, pc.ustride, l, tn, an);
            mfma_load_f<NDF, NDB>(G, Fp, l, tn, pc.fstride, fn);
            mfma_tile<NDF, NDB, KS, DR>(G, a, f4, sW, sAcc, l, t, pc, rd, Hp, dreg);
        }
    } else if constexpr (NVB == 2) {
        // (more components: the element spectra only)
        float2 an[KS];
        mfma_load_u<KS, J>(Ue, pc.ustride, l, 0, an);
        for (int t = 0; t < ntile; ++t) {
            float2 a[KS];
            float4 f4[NDF][2];
#pragma unroll
            for (int s = 0; s < KS; ++s) a[s] = an[s];
            mfma_load_u<KS, J>(Ue, pc.ustride, l, t + 1 < ntile ? t + 1 : t, an);
            mfma_load_f<NDF, NDB>(G, Fp, l, t, pc.fstride, f4);
            mfma_tile<NDF, NDB, KS, DR>(G, a, f4, sW, sAcc, l, t, pc, rd, Hp, dreg);
        }
    } else {
        for (int t = 0; t < ntile; ++t) {
            float2 a[KS];
            float4 f4[NDF][2];
            mfma_load_u<KS, J>(Ue, pc.ustride, l, t, a);
            mfma_load_f<NDF, NDB>(G, Fp, l, t, pc.fstride, f4);
            mfma_tile<NDF, NDB, KS, DR>(G, a, f4, sW, sAcc, l, t, pc, rd, Hp, dreg);
        }
    }
    mfma_store_dreg<DR>(rd, l, sAcc, dreg);
    __syncthreads();                                           // the lanes' blocks of sums
    mfma_finish<NDF, NDB>(G, l, sAcc, pc, rd, kern, model_idx, model, ndepth, relative, nk, phys, base_out, xyz,
                          active, P, R, rec_xyz);
}

// ------------------------------------------------------------------------------------
// k_bwd_mfma_wg: the same evaluation with one workgroup of W waves per element run
// u0 + blockIdx.x: the waves take the run's groups W at a time (wave w: group
// erun_group_first[u] + round + w) and sweep the bins together, each 16-bin tile of the
// element's spectra loaded once into LDS for the W groups, double-buffered (the next
// tile's loads are issued before this tile's products and stored after them).  Waves
// without a group in the last round load and wait with the others.
// ------------------------------------------------------------------------------------
template <int NDF, int NDB, int NVB, int W>
__global__ __launch_bounds__(64 * W) void k_bwd_mfma_wg(
    PathConf pc, int u0, int ubatch, const int *erun_group_first, const int *group_first,
    const float2 *U, const int *order, int R, const float *Wb, const float *Qb, const float *Qf,
    const float2 *F, const RecDesc *recs, const KernDesc *kern, const float2 *Hp, int npos,
    const int *model_idx, const float *model, int ndepth, int relative, int nk, double *phys, double *base_out,
    const double *xyz, const uint8_t *active, int P, const double *rec_xyz)
{
    constexpr int J = NVB * 25, KS = mfma_ks(J), TILE = KS * 4 * 16;   // float2 per tile
    constexpr int PER = (TILE + 64 * W - 1) / (64 * W);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    if ((int)blockIdx.x >= ubatch) return;
    const int u = u0 + (int)blockIdx.x;
    const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, jp = l & 15, bq = l >> 4;
    float2 *sU = reinterpret_cast<float2 *>(smem_raw);                  // [2][4 KS][16]
    float *sW = reinterpret_cast<float *>(sU + 2 * TILE) + (size_t)wv * KS * NDB * 64;
    double *sAcc = reinterpret_cast<double *>(reinterpret_cast<float *>(sU + 2 * TILE) + (size_t)W * KS * NDB * 64)
                 + (size_t)wv * npos * kMfmaPairs;
    const float2 *Ue = U + (size_t)(u - u0) * J * pc.ustride;
    const int gbeg = erun_group_first[u], gend = erun_group_first[u + 1];
    const int ntile = (pc.nomega + 15) >> 4;
    for (int round = gbeg; round < gend; round += W) {
        const int grp = round + wv;
        const bool mine = grp < gend;                                  // uniform in the wave
        MfmaGroup<NDF, NDB> G;
        constexpr int DR = mfma_dreg<NDF, NVB>();
        v4d dreg[DR ? DR : 1];
#pragma unroll
        for (int b = 0; b < (DR ? DR : 1); ++b) dreg[b] = v4d{0.0, 0.0, 0.0, 0.0};
        if (mine)
            mfma_group_setup<NDF, NDB, NVB>(G, grp, group_first, order, R, Wb, Qb, Qf, pc, npos, sW, sAcc, l);
        // tile 0 (the previous round's readers of both buffers are past the last barrier)
        for (int i = 0; i < PER; ++i) {
            const int idx = tid + i * 64 * W;
            if (idx < TILE) {
                const int j = idx >> 4;
                sU[idx] = j < J ? Ue[(size_t)j * pc.ustride + (idx & 15)] : make_float2(0.0f, 0.0f);
            }
        }
        __syncthreads();
        const RecDesc &rd = recs[mine ? G.r : 0];
        const float2 *Fp = F + (size_t)(mine ? G.p : 0) * NDF * pc.fstride;
        for (int t = 0; t < ntile; ++t) {
            float2 nx[PER];
            const bool more = t + 1 < ntile;
            if (more) {
#pragma unroll
                for (int i = 0; i < PER; ++i) {
                    const int idx = tid + i * 64 * W, j = idx >> 4;
                    nx[i] = (idx < TILE && j < J) ? Ue[(size_t)j * pc.ustride + ((t + 1) << 4) + (idx & 15)]
                                                  : make_float2(0.0f, 0.0f);
                }
            }
            if (mine) {
                const float2 *tile = sU + (t & 1) * TILE;
                float2 a[KS];
#pragma unroll
                for (int s = 0; s < KS; ++s) a[s] = tile[mfma_trace(s, bq) * 16 + jp];
                float4 f4[NDF][2];
                mfma_load_f<NDF, NDB>(G, Fp, l, t, pc.fstride, f4);
                mfma_tile<NDF, NDB, KS, DR>(G, a, f4, sW, sAcc, l, t, pc, rd, Hp, dreg);
            }
            if (more) {
                float2 *nt = sU + ((t + 1) & 1) * TILE;
#pragma unroll
                for (int i = 0; i < PER; ++i) {
                    const int idx = tid + i * 64 * W;
                    if (idx < TILE) nt[idx] = nx[i];
                }
            }
            __syncthreads();
        }
        if (mine) mfma_store_dreg<DR>(rd, l, sAcc, dreg);
        __syncthreads();                                               // the lanes' blocks of sums
        if (mine)
            mfma_finish<NDF, NDB>(G, l, sAcc, pc, rd, kern, model_idx, model, ndepth, relative, nk, phys, base_out,
                                  xyz, active, P, R, rec_xyz);
        __syncthreads();
    }
}

}  // namespace mck
