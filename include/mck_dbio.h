/*
 * mck_dbio.h - reading (and writing) merged AxiSEM databases, the files the reference
 * opens as <dir>/merged_instaseis_db.nc4 (readfields.f90:291-467, 499-658, 2925-3078)
 * and UTILS/repack_database.py writes.  A NetCDF-4 file is an HDF5 file: this library
 * reads it through HDF5 (libmck_dbio.so, linked against libhdf5), fills an mck_db_t
 * ready for mck_setup_t, and memory-maps the snapshots when the file stores them
 * contiguously (repack_database.py creates merged_snapshots with contiguous=True), so
 * that opening a database larger than host memory reads nothing and mck_init's paging
 * (mck_setup_t.db_page_elems) streams element pages from the operating system's page
 * cache.
 *
 * Layout read (names as in readfields.f90 and repack_database.py):
 *   root attributes  "file version" (>= 3), "npol", "number of strain dumps",
 *                    "strain dump sampling rate in sec", "planet radius" (km),
 *                    "source shift factor in sec", "scalar source magnitude",
 *                    "source depth in km"
 *   root variables   merged_snapshots [elements][variables][jpol][ipol][snapshots] f32,
 *                    stf_dump, stf_d_dump [snapshots]
 *   group Mesh       mesh_S, mesh_Z [npoints]; mp_mesh_S, mp_mesh_Z, eltype, axis
 *                    [elements]; fem_mesh [elements][4]; sem_mesh [elements][5][5];
 *                    optional mesh_vp, mesh_vs, mesh_rho, mesh_phi, mesh_xi, mesh_eta
 * Numeric values are converted to the types below on reading (as the NetCDF library
 * does for the reference).
 *
 * Every function returns 0 (MCK_OK) or an MCK_* code of mck.h; mck_dbio_last_error()
 * describes the last failure of this library.
 */
#ifndef MCK_DBIO_H
#define MCK_DBIO_H

#include <stdint.h>
#include "mck.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    mck_db_t db;               /* mesh arrays, snapshots, nvar, npol, amplitude          */
    int32_t  nt;               /* "number of strain dumps"                               */
    double   dt;               /* "strain dump sampling rate in sec"                     */
    double   planet_radius;    /* "planet radius" (km in the file), in m                 */
    double   source_shift;     /* "source shift factor in sec"                           */
    double   source_depth_km;  /* "source depth in km"                                   */
    int32_t  file_version;     /* "file version"                                         */
    const double  *stf_dump;   /* [nt] stf_dump (readfields.f90:615-619)                 */
    const double  *stf_d_dump; /* [nt] stf_d_dump                                        */
    const int32_t *sem_mesh;   /* [nelem][5][5] 0-based GLL point ids                    */
    const float   *mesh_vp, *mesh_vs, *mesh_rho,   /* [npoints] background model at the  */
                  *mesh_phi, *mesh_xi, *mesh_eta;  /* GLL points (NULL when absent)      */
    int32_t  mapped;           /* out: 1 = snapshots memory-mapped from the file         */
    void    *priv;             /* library-owned state (NULL for a caller-built struct)   */
} mck_dbfile_t;

/*
 * Opens a merged database; nvar_expected is 10 (forward) or 5 (backward), 0 = any.
 * The arrays of *out stay valid until mck_db_close(out).  map_snapshots = 1 maps the
 * snapshots from the file when they are stored contiguously (else they are read);
 * 0 always reads them into host memory.
 */
int mck_db_open(const char *path, int32_t nvar_expected, int32_t map_snapshots, mck_dbfile_t *out);

/* Releases what mck_db_open allocated or mapped. */
void mck_db_close(mck_dbfile_t *f);

/*
 * Writes db (all arrays of f->db, the scalars and stf arrays of f; sem_mesh and the
 * mesh_* model arrays when not NULL) in the layout above, merged_snapshots contiguous.
 * Used to store generated databases; a file written here opens with mck_db_open.
 */
int mck_db_write(const char *path, const mck_dbfile_t *f);

const char *mck_dbio_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MCK_DBIO_H */
