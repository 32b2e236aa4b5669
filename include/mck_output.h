/*
 * mck_output.h - the output files of a kernel run, written from the arrays that
 * mck_results_export (mck.h) produces.  Host code in libmck_dbio.so (HDF5).
 *
 *  - XDMF: <filename>.xdmf and <filename>.nc, as inversion_mesh%dump_data_xdmf writes
 *    them (inversion_mesh.f90:2026-2265): the .nc file (NetCDF-4 layout, i.e. HDF5)
 *    holds points [nvertices][3] f32, grid [nelements][nvpe] i32 (0-based vertex ids),
 *    block [nelements] i32 at its root and the variables in groups node_data and
 *    cell_data, each [nentries][nvertices or nelements] f32 (Fortran (npoints,
 *    nentries), inversion_mesh.f90:1549-1660); the .xdmf file describes one uniform grid
 *    with one hyperslab attribute per entry.
 *  - CSR: <filename>.val, .ind, .pnt, the entries whose |value| > threshold of each
 *    kernel (dump_cell_data_csr / dump_node_data_csr, inversion_mesh.f90:1886-1953),
 *    binary 4-byte records; or the same as text (dump_*_data_ascii, :1957-2022).
 *  - Intermediate results: the checkpoint of the master (create_intermediate /
 *    dump_intermediate / read_intermediate, master_queue.f90:586-705).
 *
 * Every function returns 0 (MCK_OK) or an MCK_* code; mck_dbio_last_error() describes
 * the last failure.
 */
#ifndef MCK_OUTPUT_H
#define MCK_OUTPUT_H

#include <stdint.h>
#include "mck.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One variable of node_data or cell_data (add_node_variable / add_cell_variable and
 * add_*_data, inversion_mesh.f90:1549-1660, 1741-1882). */
typedef struct {
    const char *name;                /* e.g. "kernel", "error", "model", "niterations"   */
    int32_t on_cells;                /* 0: node_data (per vertex), 1: cell_data           */
    int32_t nentries;
    const char *const *entry_names;  /* [nentries] attribute names; NULL = the name      */
    const float *values;             /* [nentries][nvertices or nelements]               */
} mck_outvar_t;

/*
 * dump_data_xdmf: writes <filename>.nc and <filename>.xdmf.
 *  element_type  "tri", "tet", "quad", "hex" or "vox" (inversion_mesh.f90:2058-2073)
 *  vertices      [nvertices][3] metres (written in single precision)
 *  connectivity  [nelements][nvpe] 1-based vertex ids (written 0-based, as the mesh
 *                stores them, inversion_mesh.f90:823)
 *  block_id      [nelements], NULL = every element in block 1
 */
int mck_write_xdmf(const char *filename, const char *element_type, int32_t nvertices,
                   const double *vertices, int32_t nelements, int32_t nvpe,
                   const int32_t *connectivity, const int32_t *block_id, int32_t nvars,
                   const mck_outvar_t *vars);

/*
 * dump_*_data_csr (ascii = 0) or dump_*_data_ascii (ascii = 1) of values
 * [nkernel][npoints]: for each kernel the points with |value| > threshold, their values
 * (.val) and 1-based indices (.ind) in point order, and the running count after each
 * kernel (.pnt, one text line per kernel, a blank and the count in 11 columns as
 * list-directed output writes it).  Binary: 4-byte float and integer records (direct
 * access, recl = 4).  Text: each kernel's values (" ", E15.8) and indices (" ", I15) on
 * one line, the form of dump_node_data_ascii, used for cells too (the reference writes
 * cell values list-directed, one per line).  The node version of the reference writes
 * .pnt unformatted to a formatted unit, which gfortran rejects; both write text here.
 */
int mck_write_csr(const char *filename, int32_t npoints, int32_t nkernel, const float *values,
                  double threshold, int32_t ascii);

/*
 * create_intermediate / dump_intermediate (master_queue.f90:586-705): K_x, Var
 * [nkernel][npoints] f32, Bg_Model [nmodel][npoints] f32, Het_Model [nhet][npoints] f32
 * (nmodel, nhet may be 0 with NULL arrays), niterations [nkernel][nelements] i32 and
 * element_proc [nelements] i32, written to path (replaced if present).
 */
int mck_write_intermediate(const char *path, int32_t npoints, int32_t nelements, int32_t nkernel,
                           int32_t nmodel, int32_t nhet, const float *kernel, const float *variance,
                           const float *bg_model, const float *het_model, const int32_t *niterations,
                           const int32_t *element_proc);

/*
 * read_intermediate (master_queue.f90:615-659): the arrays of mck_write_intermediate
 * read back (each nullable); the dimensions must match the file's.
 */
int mck_read_intermediate(const char *path, int32_t npoints, int32_t nelements, int32_t nkernel,
                          int32_t nmodel, int32_t nhet, float *kernel, float *variance,
                          float *bg_model, float *het_model, int32_t *niterations,
                          int32_t *element_proc);

#ifdef __cplusplus
}
#endif

#endif /* MCK_OUTPUT_H */
