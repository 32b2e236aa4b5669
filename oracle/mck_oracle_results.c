/*
 * mck_oracle_results.c - FP64 CPU restatement of the master's result assembly.
 *
 * TEST INFRASTRUCTURE ONLY (see mck_oracle.h): the checker of mck_results_*.
 *
 * ora_results_add follows extract_receive_buffer (src/master_queue.f90:253-323): the
 * loop over the task's elements (skipping slots marked -1, :275-277) and their basis
 * functions, the point of a basis function (connectivity(ibasis, ielement) in
 * 'onvertices' mode, the element itself in 'volumetric' mode, :282-291), valence 1, the
 * serial sums K_x, Var and Bg_Model (:293-304), and the assignments of niterations,
 * computation_time and element_proc (:309-311).
 *
 * ora_results_export follows finalize's single-precision conversions
 * (src/master_queue.f90:366-398 for 'onvertices', :409-441 for 'volumetric', and the
 * niterations cell variable :515-516): real(x, kind=sp) of each array, and the error
 * sqrt(Var/|K_x|) where |K_x| > 1e-30, else 1000 (:383-387, :426-430); the result is laid out
 * entry-major, as the output variables (npoints, nentries) of Fortran order.
 *
 * Pins (tests/test_oracle_results.py): hand-summed tasks on a small mesh (shared
 * vertices, empty slots, volumetric mode, reassigned niterations), the error transform at
 * its special cases, and the equality of any task order's sums for exactly representable
 * values.
 */
#include <math.h>
#include <stdint.h>

#include "mck_oracle.h"

int ora_results_add(int int_type, int nbasis, int nk, int nmodel, int npoints, int nelements,
                    const int32_t *conn, int nelems, const int32_t *elems, const double *values,
                    const double *variance, const int32_t *niter, const double *ctime,
                    const double *model, int rank, double *K, double *V, double *M,
                    int32_t *niter_g, double *ctime_g, int32_t *proc) {
    for (int iel = 0; iel < nelems; ++iel) {
        const int ielement = elems[iel];
        if (ielement == -1) continue;
        if (ielement < 1 || ielement > nelements) return 1;
        for (int ib = 0; ib < nbasis; ++ib) {
            int ipoint;
            const double valence = 1.0;
            if (int_type == 0)
                ipoint = conn[(int64_t)(ielement - 1) * nbasis + ib];
            else
                ipoint = ielement;
            if (ipoint < 1 || ipoint > npoints) return 1;
            const int64_t src = ((int64_t)iel * nbasis + ib) * nk;
            const int64_t dst = (int64_t)(ipoint - 1) * nk;
            for (int k = 0; k < nk; ++k) K[dst + k] = K[dst + k] + values[src + k] * valence;
            for (int k = 0; k < nk; ++k) V[dst + k] = V[dst + k] + variance[src + k] * valence;
            if (nmodel > 0) {
                const int64_t ms = ((int64_t)iel * nbasis + ib) * nmodel;
                const int64_t md = (int64_t)(ipoint - 1) * nmodel;
                for (int j = 0; j < nmodel; ++j) M[md + j] = M[md + j] + model[ms + j] * valence;
            }
        }
        for (int k = 0; k < nk; ++k) niter_g[(int64_t)(ielement - 1) * nk + k] = niter[(int64_t)iel * nk + k];
        ctime_g[ielement - 1] = ctime ? ctime[iel] : 0.0;
        proc[ielement - 1] = rank;
    }
    return 0;
}

void ora_results_export(int which, int np, int ncol, const double *a, const double *kx,
                        const int32_t *ia, float *out) {
    /* MCK_RES_KERNEL, _VARIANCE, _MODEL, _HETERO: real(a, sp); _ERROR; _NITER */
    for (int p = 0; p < np; ++p)
        for (int c = 0; c < ncol; ++c) {
            const int64_t i = (int64_t)p * ncol + c;
            float v;
            if (which == MCK_RES_ERROR)
                v = fabs(kx[i]) > 1e-30 ? (float)sqrt(a[i] / fabs(kx[i])) : 1000.0f;
            else if (which == MCK_RES_NITER)
                v = (float)ia[i];
            else
                v = (float)a[i];
            out[(int64_t)c * np + p] = v;
        }
}
