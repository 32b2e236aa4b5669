// synth_gpu.hip - device-side generator of the synthetic merged databases used by the
// benchmark: the wavelet construction of synth.cpp (8 Gaussian-windowed 40 s wavelets
// per GLL point and variable, amplitudes N(0,1), arrival times U(100, 900) s) evaluated
// directly into HBM with the same double-precision arithmetic (synth_math.h), so every
// sample equals the host generator's bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "synth_math.h"

namespace {

// one block per (element, variable); threads stride over 25 nodes x nt samples
__global__ void k_fill(int nelem, const int *sem, int nvar, int nt, double dt, uint64_t seed,
                       float *out) {
    const int ev = blockIdx.x;
    const int e = ev / nvar, v = ev - (ev / nvar) * nvar;
    __shared__ synth_math::Wavelet wl[25][8];
    if (threadIdx.x < 200) {
        const int n = threadIdx.x / 8, w = threadIdx.x % 8;
        wl[n][w] = synth_math::wavelet(seed, (int64_t)sem[25 * e + n], v, w, nt, dt);
    }
    __syncthreads();
    float *dst = out + ((size_t)e * nvar + v) * 25 * nt;
    for (int i = threadIdx.x; i < 25 * nt; i += blockDim.x) {
        const int n = i / nt, t = i - (i / nt) * nt;
        double acc = 0.0;
        for (int w = 0; w < 8; ++w)
            if (t >= wl[n][w].i0 && t <= wl[n][w].i1) acc += synth_math::wavelet_sample(wl[n][w], t, dt);
        dst[i] = (float)acc;
    }
}

// plain streaming copy: the HBM-traffic calibration for hardware counters (a launch
// reads and writes exactly n4 * 16 bytes)
__global__ void k_copy_calibration(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                   size_t n4) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
        dst[i] = src[i];
}

}  // namespace

extern "C" {

// Runs `reps` launches of a plain copy of `bytes` bytes (a multiple of 16) between two
// fresh device buffers; returns the average launch time in ms through *ms.
int synth_gpu_copy_calibration(size_t bytes, int reps, float *ms) {
    float4 *a = nullptr, *b = nullptr;
    size_t n4 = bytes / 16;
    if (bytes % 16 || reps < 1) return 1;
    if (hipMalloc((void **)&a, bytes) != hipSuccess) return 2;
    if (hipMalloc((void **)&b, bytes) != hipSuccess) { (void)hipFree(a); return 2; }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(k_copy_calibration, dim3(4096), dim3(256), 0, 0, a, b, n4);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_copy_calibration, dim3(4096), dim3(256), 0, 0, a, b, n4);
    (void)hipEventRecord(e1, 0);
    hipError_t err = hipEventSynchronize(e1);
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return err == hipSuccess ? 0 : 3;
}

// Allocates [nelem][nvar][5][5][nt] floats in device memory and fills them; sem is the
// host array of GLL point ids [nelem][5][5].  Returns 0 on success.
int synth_gpu_db(int nelem, const int *sem_host, int nvar, int nt, double dt, uint64_t seed,
                 float **d_out) {
    int *d_sem = nullptr;
    float *d = nullptr;
    if (hipMalloc((void **)&d_sem, (size_t)nelem * 25 * sizeof(int)) != hipSuccess) return 1;
    if (hipMemcpy(d_sem, sem_host, (size_t)nelem * 25 * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) return 2;
    if (hipMalloc((void **)&d, (size_t)nelem * nvar * 25 * nt * sizeof(float)) != hipSuccess) {
        (void)hipFree(d_sem);
        return 3;
    }
    hipLaunchKernelGGL(k_fill, dim3(nelem * nvar), dim3(256), 0, 0, nelem, d_sem, nvar, nt, dt, seed, d);
    hipError_t e = hipDeviceSynchronize();
    (void)hipFree(d_sem);
    if (e != hipSuccess) { (void)hipFree(d); return 4; }
    *d_out = d;
    return 0;
}

int synth_gpu_free(float *d) { return hipFree(d) == hipSuccess ? 0 : 1; }

// page-locked host memory for a database that the library pages from the host
int synth_gpu_host_alloc(size_t bytes, float **p) {
    return hipHostMalloc((void **)p, bytes, hipHostMallocDefault) == hipSuccess ? 0 : 1;
}

int synth_gpu_host_free(float *p) { return hipHostFree(p) == hipSuccess ? 0 : 1; }

// copies a device database to host memory (for the CPU baseline)
int synth_gpu_download(const float *d, float *host, size_t nfloat) {
    return hipMemcpy(host, d, nfloat * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess ? 0 : 1;
}

}  // extern "C"
