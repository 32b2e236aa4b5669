// mck_results.inc.h - device kernels of the result assembly: the master's accumulation
// of the slaves' results per inversion-mesh point (extract_receive_buffer,
// src/master_queue.f90:253-323) and the single-precision conversions of finalize
// (src/master_queue.f90:327-581), with the accumulators resident in device memory.
#pragma once

#include "mck_device.h"

namespace mck {

// One thread per (segment, column).  A segment holds the contributions of one mesh point
// within one task, in task order (element, then basis function); they are added to the
// point's accumulator one at a time in that order, the order of the reference's serial
// loop (master_queue.f90:275-307, valence 1), so the FP64 sums are the same.
__global__ __launch_bounds__(256) void k_results_add(int nseg, int ncol, const int *__restrict__ seg_start,
                                                      const int *__restrict__ seg_point,
                                                      const int *__restrict__ slots,
                                                      const double *__restrict__ vals,
                                                      double *__restrict__ acc) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nseg * ncol) return;
    const int s = (int)(i / ncol), c = (int)(i - (long long)s * ncol);
    double *dst = acc + (size_t)seg_point[s] * ncol + c;
    double a = *dst;
    for (int j = seg_start[s]; j < seg_start[s + 1]; ++j) a += vals[(size_t)slots[j] * ncol + c];
    *dst = a;
}

// Per element of the task: niterations, computation time and the rank that computed it
// are assigned (master_queue.f90:309-311); elem_idx < 0 marks an empty slot.
__global__ __launch_bounds__(256) void k_results_elem(int nelems, int nk, const int *__restrict__ elem_idx,
                                                       const int *__restrict__ niter_task,
                                                       const double *__restrict__ ctime_task, int rank,
                                                       int *__restrict__ niter, double *__restrict__ ctime,
                                                       int *__restrict__ proc) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nelems * nk) return;
    const int iel = (int)(i / nk), k = (int)(i - (long long)iel * nk);
    const int e = elem_idx[iel];
    if (e < 0) return;
    niter[(size_t)e * nk + k] = niter_task[i];
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
