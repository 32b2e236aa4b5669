// mck_device.h - device-side data layout shared by the HIP kernels and the host driver.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mck {

constexpr int kMaxSim = 4;       // fwd simulations MZZ, MXX+MYY, MXZ, MXY
constexpr int kMaxVar = 10;      // merged variables per element (fwd)
constexpr int kMaxKernelsPerRec = 256;
constexpr int kNN = 12;          // nearest midpoints tested (readfields.f90:1067)
constexpr int kStrainThreads = 256;
constexpr int kEvalThreads = 256;
// accumulators of the matrix-core products (fp32 and FP64 16 x 16 blocks: 4 per lane)
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kMaxSlots = 256;   // distinct (window-filter, base kernel) pairs per receiver
// base kernels of the frequency-domain contraction: lambda, mu, rho (unused), a, b, c
// (calc_basekernel, kernel.f90:514-570) and the model-independent combinations that the
// physical kernels use (kernel.f90:665-832): 6 = mu - 2 lambda (vs), 7 = lambda - a - c
// (vpv), 8 = mu - 2 lambda - b (vsh, whose other term is c)
constexpr int kNumBk = 9;
static_assert(kMaxSlots <= kStrainThreads, "one thread per slot in the final slot reduction");

// LDS entries of the quarter twiddle table of an M-point complex FFT: M/2 values, one
// padding slot after every 16 so that strided lookups spread over the LDS banks
__host__ __device__ constexpr int tq_len(int M) { return (M >> 1) + (M >> 5) + 1; }
__host__ __device__ constexpr int tq_slot(int r) { return r + (r >> 4); }

// GLL/GLJ collocation points and Lagrange derivative matrices for npol = 4
// (calc_gradient_terms, readfields.f90:3084-3150); G[i + 5 j] = l_i'(x_j).
struct Basis {
    double gll[5], glj[5];
    double G2[25], G1[25];
};

// One merged database in HBM.
struct DevDb {
    int nelem, nvar, nt;
    const float *u;           // [u_last - u_first][nvar][5][5][nt]: elements u_first ..
    int u_first, u_last;      // u_last - 1 (all elements when resident, one page when paged)
    // paged database whose pages of this launch are all in device buffers: the buffer of
    // each page (null: not held), page_elems elements each; u is then unused
    const float *const *page_u;
    int page_elems;
    const float *mp_s, *mp_z; // element midpoints
    const double *econ;       // [nelem][16] element constants (see host setup)
    const double *ejac;       // [nelem][25][5] inverse Jacobian and 1/s at the nodes
    const int *eltype, *axis;
    // uniform grid over the midpoints for the exact 12-nearest search
    int gnx, gnz;
    double gs0, gz0, gh;      // origin and cell size (m)
    const int *cell_start;    // [gnx*gnz+1]
    const int *cell_elems;    // element ids sorted by cell
    double amplitude;
};

// Observation frame: source (frame 0) or receiver r (frame 1 + r).
struct DevFrame {
    double cos_lon, sin_lon, cos_colat, sin_colat;
};

// Result of locating one point in one frame.
struct Located {
    int elem;                 // 0-based, -1 = not found
    int tie;                  // equal fp32 distances among the 13 nearest
    double xi, eta, phi;
};

// Path-wide configuration passed by value to the kernels.
struct PathConf {
    int nt, M, nomega;        // samples, FFT half length (N/2) and N/2+1
    int ustride;              // row stride of element spectra: nomega rounded up to 16
    int fstride;              // row stride of forward and window spectra (= ustride)
    double dt;
    int nd;                   // 1 (trace) or 6 (full strain)
    int full;
    double planet_radius;
    // fwd frame: active simulations and the variables they read
    int nsim_fwd;
    int fwd_sim[kMaxSim];     // 1..4
    int nvar_fwd;             // variables read from the fwd DB
    int fwd_vars[kMaxVar];
    double mij[6];
    // bwd frame
    int bwd_isim;             // 1 (Z) or 2 (R, T)
    int nvar_bwd;
    int bwd_vars[3];
    double bwd_fi[3];
    float taper[10];          // Gaussian taper samples (fft.f90:586-593), fp32
    int ntaper;
    // strain components carried through the FFT, per frame: the trace (nd 1), or the
    // source-frame Voigt components with a non-zero azimuthal factor (4 for monopole
    // sources ss, pp, zz, sz; 6 otherwise), rotated to the earth frame only in the
    // frequency-domain contraction
    int nd_f, nd_b;
    int comps_f[6], comps_b[6];
    int need_bk;              // bit b set: base kernel b (< 6) must be formed per bin
    int bk_lds[kNumBk];       // LDS array of (combined) base kernel b (-1 = not stored)
    int nbk_lds;
    int max_slots;            // largest slot count of a receiver
    int max_cls;              // largest padded class-position count of a receiver
    int op_bytes_b;           // LDS of a pair's operator set-up in k_bwd_eval
    // dampen_field (readfields.f90:3184-3204): radius (<= 0 disables), source position
    double damp_radius;
    double src_r[3];
};

// LDS layout of the per-point operator set-up
struct OpShared {
    double invj[25][4];     // [ipol + 5 jpol]
    double sinv[25];        // 1 / s at the nodes
    double li[5], lj[5];
    double M[6][kMaxSim][6];// output comp x simulation x strain comp
    float Q[6][6];          // earth Voigt component x carried component (source mode)
    int elem;
    int axial;
    int type;
    float W[6 * kMaxVar * 25];  // last member: a kernel's LDS holds only nd x nvar x 25
};

// LDS bytes of an OpShared whose weight table holds nw entries (16-byte aligned)
__host__ __device__ constexpr size_t op_bytes(int nw)
{
    return ((offsetof(OpShared, W) + 4 * (size_t)nw + 15) / 16) * 16;
}

// Per-pair coefficients of the frequency-domain contraction (backward kernel):
// earth-frame strain = Q x carried components, for the fwd (Qf) and bwd (Qb) fields.
// (The per-slot sums live in dynamic LDS, sized by the largest receiver.)
struct EvalShared {
    float Qf[6][6];
    float Amu[6][6];        // Qf^T Qb: F.B = f^T Amu b
    float uf[6], ub[6];     // trace rows (1,1,1,0,0,0) Q
};

constexpr int kSlotGroup = 6;        // slots sharing one window spectrum, summed together


// receiver: distinct (window spectrum, base kernel) pairs ('slots') and its kernels
struct RecDesc {
    int nslot;
    int first_kernel, nkernel;
    int slot_h[kMaxSlots];
    int slot_bk[kMaxSlots];    // base kernel (or combination) of the slot, < kNumBk
    int slot_lds[kMaxSlots];   // LDS array of the slot's base kernel in the contraction
    // slots ordered by window spectrum, in groups sharing one spectrum (<= kSlotGroup)
    int ngroup;
    int grp_h[kMaxSlots], grp_first[kMaxSlots], grp_n[kMaxSlots];
    // the slots by base-kernel class for the matrix-core kernel: lambda (slot_bk 0), mu
    // (1), combinations; class c holds the positions c ? cls_end[c - 1] : 0 .. cls_end[c] - 1,
    // each class padded to a multiple of 16 positions (blocks of 16 matrix-core columns;
    // padding: another position's window spectrum, a sum nobody reads); cls_h: window
    // spectrum of a position, cls_pos: position of a slot
    int cls_h[kMaxSlots + 48];
    int cls_pos[kMaxSlots];
    int cls_end[3];
};

// kernel: model parameter and the slot of every base kernel it needs (-1 = none)
struct KernDesc {
    int param;                 // MCK_PARAM_*
    int slot[kNumBk];          // slot of each (combined) base kernel it reads (-1 = none)
};

// Monte Carlo state of the voxels of one task: one integrator per basis function
// (1 volumetric, 4 on the vertices of a tetrahedron)
struct VoxState {
    int nbasis;
    double *fsum, *f2sum, *integral, *variance;   // [nvox][nbasis][nk]
    uint8_t *conv;                                // [nvox][nbasis][nk]
    int *niter;                                   // [nvox][nk]
    int64_t *nmodels;                             // [nvox]
    int *draws;                                   // [nvox] Halton draws so far
    uint8_t *done;                                // [nvox]
    int *steps;                                   // [nvox] MC steps evaluated
};

}  // namespace mck
