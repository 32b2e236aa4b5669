// mck_results.inc.h - device kernels of the result assembly: the master's accumulation
// of the slaves' results per inversion-mesh point (extract_receive_buffer,
// src/master_queue.f90:253-323) and the single-precision conversions of finalize
// (src/master_queue.f90:327-581), with the accumulators resident in device memory.
#pragma once

#include <type_traits>

#include "mck_device.h"

namespace mck {

// One (point, slot) pair per basis function of the task: the inversion-mesh point the
// basis function's results go to (its vertex, connectivity(ibasis, ielement), or the
// element itself, master_queue.f90:282-291) and its slot iel * nbasis + ibasis in the
// task's result arrays.  Empty slots (element -1) get the point npoints, which sorts
// after every real point and is skipped.
__global__ __launch_bounds__(256) void k_results_pairs(int nelems, int nbasis, int onvertices, int npoints,
                                                        const int *__restrict__ elems,
                                                        const int *__restrict__ conn, int *__restrict__ point,
                                                        int *__restrict__ slot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nelems * nbasis) return;
    const int iel = i / nbasis, ib = i - iel * nbasis;
    const int e = elems[iel];
    int p = npoints;
    if (e > 0) p = onvertices ? conn[(size_t)(e - 1) * nbasis + ib] - 1 : e - 1;
    point[i] = p;
    slot[i] = i;
}

// One thread per (segment, VEC columns), for up to two arrays at once (K_x and Var share
// the segments).  After a stable sort of the pairs by point, a segment holds the
// contributions of one point in task order (element, then basis function); they are
// added to the point's accumulator one at a time in that order, the order of the
// reference's serial loop (master_queue.f90:275-307, valence 1), so the FP64 sums are the
// same.  The segment of the empty-slot point npoints is skipped.  VEC = 2 needs an even
// column count (16-byte rows of two columns).
template <int VEC>
__global__ __launch_bounds__(256) void k_results_add(const int *__restrict__ nseg_p, int ncol, int npoints,
                                                      const int *__restrict__ seg_start,
                                                      const int *__restrict__ seg_count,
                                                      const int *__restrict__ seg_point,
                                                      const int *__restrict__ slots,
                                                      const double *__restrict__ vals0, double *__restrict__ acc0,
                                                      const double *__restrict__ vals1, double *__restrict__ acc1) {
    using V = typename std::conditional<VEC == 2, double2, double>::type;
    const int nv = ncol / VEC;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nseg = *nseg_p;
    if (i >= (long long)nseg * nv) return;
    const int s = (int)(i / nv), c = (int)(i - (long long)s * nv);
    const int p = seg_point[s];
    if (p >= npoints) return;
    const int j0 = seg_start[s], j1 = j0 + seg_count[s];
    const size_t row = (size_t)p * nv + c;
    V a0 = ((V *)acc0)[row], a1{};
    if (acc1) a1 = ((V *)acc1)[row];
    for (int j = j0; j < j1; ++j) {
        const size_t src = (size_t)slots[j] * nv + c;
        const V v0 = ((const V *)vals0)[src];
        if constexpr (VEC == 2) { a0.x += v0.x; a0.y += v0.y; } else { a0 += v0; }
        if (acc1) {
            const V v1 = ((const V *)vals1)[src];
            if constexpr (VEC == 2) { a1.x += v1.x; a1.y += v1.y; } else { a1 += v1; }
        }
    }
    ((V *)acc0)[row] = a0;
    if (acc1) ((V *)acc1)[row] = a1;
}

// Per element of the task: niterations, computation time and the rank that computed it
// are assigned (master_queue.f90:309-311); elem_idx < 0 marks an empty slot.  One thread
// per (element, VEC columns); VEC = 4 needs nk % 4 == 0 (16-byte rows).
template <int VEC>
__global__ __launch_bounds__(256) void k_results_elem(int nelems, int nk, const int *__restrict__ elem_idx,
                                                       const int *__restrict__ niter_task,
                                                       const double *__restrict__ ctime_task, int rank,
                                                       int *__restrict__ niter, double *__restrict__ ctime,
                                                       int *__restrict__ proc) {
    using V = typename std::conditional<VEC == 4, int4, int>::type;
    const int nv = nk / VEC;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nelems * nv) return;
    const int iel = (int)(i / nv), k = (int)(i - (long long)iel * nv);
    const int e = elem_idx[iel];
    if (e < 0) return;
    ((V *)niter)[(size_t)e * nv + k] = ((const V *)niter_task)[i];
    if (k == 0) {
        ctime[e] = ctime_task ? ctime_task[iel] : 0.0;
        proc[e] = rank;
    }
}

// Accumulator [np][ncol] (FP64, point-major) to the output layout [ncol][np] in single
// precision, through a 32 x 32 LDS tile.  MODE 0: real(a, sp); MODE 1: the relative
// error of finalize, sqrt(V / |K|) where |K| > 1e-30, else 1000 (master_queue.f90:
// 383-387); MODE 2: real(n, sp) of integer counts.
template <int MODE>
__global__ __launch_bounds__(256) void k_results_out(int np, int ncol, const void *__restrict__ a_,
                                                      const double *__restrict__ kx,
                                                      float *__restrict__ out) {
    __shared__ float tile[32][33];
    const int p0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int p = p0 + r, c = c0 + tx;
        float v = 0.f;
        if (p < np && c < ncol) {
            const size_t i = (size_t)p * ncol + c;
            if (MODE == 0) {
                v = (float)((const double *)a_)[i];
            } else if (MODE == 1) {
                const double k = fabs(kx[i]);
                v = k > 1e-30 ? (float)sqrt(((const double *)a_)[i] / k) : 1000.0f;
            } else {
                v = (float)((const int *)a_)[i];
            }
        }
        tile[r][tx] = v;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int c = c0 + r, p = p0 + tx;
        if (c < ncol && p < np) out[(size_t)c * np + p] = tile[tx][r];
    }
}

}  // namespace mck
