// mck_dbio.cpp - merged AxiSEM database files through HDF5 (include/mck_dbio.h) and the
// output files of a kernel run (include/mck_output.h).
//
// Readers follow readfields.f90: the root attributes of open_file_read_varids_and_
// attributes (readfields.f90:499-658), the Mesh group of read_meshes (readfields.f90:
// 2925-3078) and the merged_snapshots hyperslabs of read_points_from_disk
// (readfields.f90:2673-2768); the variable layout is the one repack_database.py:151-187
// writes.  Host code only (g++), no device calls.
#include "mck_dbio.h"
#include "mck_output.h"

#include <hdf5.h>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// owner of everything mck_db_open hands out
struct DbFile {
    std::vector<float> mesh_s, mesh_z, mp_s, mp_z, snapshots;
    std::vector<float> vp, vs, rho, phi, xi, eta;
    std::vector<int32_t> fem, sem, eltype, axis;
    std::vector<double> stf, stf_d;
    void *map = nullptr;
    size_t map_len = 0;
};

// silences HDF5's own error printing inside a scope
struct Quiet {
    H5E_auto2_t fn = nullptr;
    void *data = nullptr;
    Quiet() { H5Eget_auto2(H5E_DEFAULT, &fn, &data); H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
    ~Quiet() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

struct Hid {
    hid_t id;
    herr_t (*close)(hid_t);
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Hid() { if (id >= 0) close(id); }
    Hid(const Hid &) = delete;
    Hid &operator=(const Hid &) = delete;
};

bool attr_exists(hid_t loc, const char *name) { return H5Aexists(loc, name) > 0; }

int read_attr_double(hid_t loc, const char *name, double *v) {
    if (!attr_exists(loc, name)) return fail(MCK_EINVAL, std::string("missing attribute '") + name + "'");
    Hid a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
    if (a.id < 0 || H5Aread(a.id, H5T_NATIVE_DOUBLE, v) < 0)
        return fail(MCK_EINVAL, std::string("attribute '") + name + "' is not numeric");
    return MCK_OK;
}

int read_attr_int(hid_t loc, const char *name, int32_t *v) {
    if (!attr_exists(loc, name)) return fail(MCK_EINVAL, std::string("missing attribute '") + name + "'");
    Hid a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
    if (a.id < 0 || H5Aread(a.id, H5T_NATIVE_INT32, v) < 0)
        return fail(MCK_EINVAL, std::string("attribute '") + name + "' is not numeric");
    return MCK_OK;
}

// dimensions of a dataset
int dims_of(hid_t ds, std::vector<hsize_t> &d) {
    Hid sp(H5Dget_space(ds), H5Sclose);
    if (sp.id < 0) return fail(MCK_EINVAL, "dataset without a dataspace");
    int n = H5Sget_simple_extent_ndims(sp.id);
    if (n < 0) return fail(MCK_EINVAL, "dataset without simple extent");
    d.assign(n, 0);
    if (n) H5Sget_simple_extent_dims(sp.id, d.data(), nullptr);
    return MCK_OK;
}

// reads a whole dataset with type conversion; shape must match `want` (product only
// when want is empty)
template <class T>
int read_var(hid_t loc, const char *name, hid_t memtype, const std::vector<hsize_t> &want,
             std::vector<T> &out, bool optional = false, bool *found = nullptr) {
    if (found) *found = false;
    if (H5Lexists(loc, name, H5P_DEFAULT) <= 0) {
        if (optional) return MCK_OK;
        return fail(MCK_EINVAL, std::string("missing variable '") + name + "'");
    }
    Hid ds(H5Dopen2(loc, name, H5P_DEFAULT), H5Dclose);
    if (ds.id < 0) return fail(MCK_EINVAL, std::string("cannot open variable '") + name + "'");
    std::vector<hsize_t> d;
    int rc = dims_of(ds.id, d);
    if (rc) return rc;
    if (d != want) {
        std::string s = std::string("variable '") + name + "' has shape (";
        for (size_t i = 0; i < d.size(); ++i) s += (i ? ", " : "") + std::to_string(d[i]);
        s += "), expected (";
        for (size_t i = 0; i < want.size(); ++i) s += (i ? ", " : "") + std::to_string(want[i]);
        return fail(MCK_EINVAL, s + ")");
    }
    size_t n = 1;
    for (hsize_t x : d) n *= (size_t)x;
    out.resize(n);
    if (n && H5Dread(ds.id, memtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) < 0)
        return fail(MCK_EINVAL, std::string("cannot read variable '") + name + "'");
    if (found) *found = true;
    return MCK_OK;
}

int length_of(hid_t loc, const char *name, hsize_t *n) {
    if (H5Lexists(loc, name, H5P_DEFAULT) <= 0)
        return fail(MCK_EINVAL, std::string("missing variable '") + name + "'");
    Hid ds(H5Dopen2(loc, name, H5P_DEFAULT), H5Dclose);
    std::vector<hsize_t> d;
    int rc = dims_of(ds.id, d);
    if (rc) return rc;
    if (d.size() != 1) return fail(MCK_EINVAL, std::string("variable '") + name + "' is not 1-D");
    *n = d[0];
    return MCK_OK;
}

// ---- writing ---------------------------------------------------------------------
int write_var(hid_t loc, const char *name, hid_t filetype, hid_t memtype,
              const std::vector<hsize_t> &dims, const void *data) {
    Hid sp(H5Screate_simple((int)dims.size(), dims.data(), nullptr), H5Sclose);
    Hid ds(H5Dcreate2(loc, name, filetype, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (sp.id < 0 || ds.id < 0) return fail(MCK_EINVAL, std::string("cannot create variable '") + name + "'");
    if (H5Dwrite(ds.id, memtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
        return fail(MCK_EINVAL, std::string("cannot write variable '") + name + "'");
    return MCK_OK;
}

int write_attr(hid_t loc, const char *name, hid_t filetype, hid_t memtype, const void *v) {
    Hid sp(H5Screate(H5S_SCALAR), H5Sclose);
    Hid a(H5Acreate2(loc, name, filetype, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (a.id < 0 || H5Awrite(a.id, memtype, v) < 0)
        return fail(MCK_EINVAL, std::string("cannot write attribute '") + name + "'");
    return MCK_OK;
}

int write_attr_text(hid_t loc, const char *name, const char *text) {
    // fixed-length text, as NetCDF stores character attributes
    Hid t(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(t.id, std::max<size_t>(1, strlen(text)));
    H5Tset_strpad(t.id, H5T_STR_NULLTERM);
    return write_attr(loc, name, t.id, t.id, text);
}

}  // namespace

extern "C" {

const char *mck_dbio_last_error(void) { return g_err.c_str(); }

void mck_db_close(mck_dbfile_t *f) {
    if (!f || !f->priv) return;
    DbFile *D = static_cast<DbFile *>(f->priv);
    if (D->map) munmap(D->map, D->map_len);
    delete D;
    memset(f, 0, sizeof(*f));
}

int mck_db_open(const char *path, int32_t nvar_expected, int32_t map_snapshots, mck_dbfile_t *out) {
    if (!path || !out) return fail(MCK_EINVAL, "null argument");
    memset(out, 0, sizeof(*out));
    Quiet quiet;
    Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail(MCK_EINVAL, std::string("cannot open '") + path + "' as HDF5 / NetCDF-4");
    int rc;
    int32_t version = 0, npol = 0, nt = 0;
    double dt = 0, radius_km = 0, shift = 0, amp = 0, depth = 0;
    if ((rc = read_attr_int(file.id, "file version", &version))) return rc;
    if (version < 3) return fail(MCK_EINVAL, "AxiSEM file version " + std::to_string(version) + " < 3");
    if ((rc = read_attr_int(file.id, "npol", &npol))) return rc;
    if (npol != 4) return fail(MCK_EINVAL, "only npol = 4 is supported");
    if ((rc = read_attr_int(file.id, "number of strain dumps", &nt))) return rc;
    if ((rc = read_attr_double(file.id, "strain dump sampling rate in sec", &dt))) return rc;
    if ((rc = read_attr_double(file.id, "planet radius", &radius_km))) return rc;
    if ((rc = read_attr_double(file.id, "source shift factor in sec", &shift))) return rc;
    if ((rc = read_attr_double(file.id, "scalar source magnitude", &amp))) return rc;
    if (attr_exists(file.id, "source depth in km") &&
        (rc = read_attr_double(file.id, "source depth in km", &depth))) return rc;

    if (H5Lexists(file.id, "Mesh", H5P_DEFAULT) <= 0) return fail(MCK_EINVAL, "missing group 'Mesh'");
    Hid mesh(H5Gopen2(file.id, "Mesh", H5P_DEFAULT), H5Gclose);
    hsize_t npoints = 0, nelem = 0;
    if ((rc = length_of(mesh.id, "mesh_S", &npoints))) return rc;
    if ((rc = length_of(mesh.id, "mp_mesh_S", &nelem))) return rc;

    DbFile *D = new DbFile;
    auto bail = [&](int code) { delete D; return code; };
    const std::vector<hsize_t> P{npoints}, E{nelem}, E4{nelem, 4}, E55{nelem, 5, 5}, T{(hsize_t)nt};
    if ((rc = read_var(mesh.id, "mesh_S", H5T_NATIVE_FLOAT, P, D->mesh_s))) return bail(rc);
    if ((rc = read_var(mesh.id, "mesh_Z", H5T_NATIVE_FLOAT, P, D->mesh_z))) return bail(rc);
    if ((rc = read_var(mesh.id, "mp_mesh_S", H5T_NATIVE_FLOAT, E, D->mp_s))) return bail(rc);
    if ((rc = read_var(mesh.id, "mp_mesh_Z", H5T_NATIVE_FLOAT, E, D->mp_z))) return bail(rc);
    if ((rc = read_var(mesh.id, "eltype", H5T_NATIVE_INT32, E, D->eltype))) return bail(rc);
    if ((rc = read_var(mesh.id, "axis", H5T_NATIVE_INT32, E, D->axis))) return bail(rc);
    if ((rc = read_var(mesh.id, "fem_mesh", H5T_NATIVE_INT32, E4, D->fem))) return bail(rc);
    if ((rc = read_var(mesh.id, "sem_mesh", H5T_NATIVE_INT32, E55, D->sem))) return bail(rc);
    bool f_vp, f_vs, f_rho, f_phi, f_xi, f_eta;
    if ((rc = read_var(mesh.id, "mesh_vp", H5T_NATIVE_FLOAT, P, D->vp, true, &f_vp))) return bail(rc);
    if ((rc = read_var(mesh.id, "mesh_vs", H5T_NATIVE_FLOAT, P, D->vs, true, &f_vs))) return bail(rc);
    if ((rc = read_var(mesh.id, "mesh_rho", H5T_NATIVE_FLOAT, P, D->rho, true, &f_rho))) return bail(rc);
    if ((rc = read_var(mesh.id, "mesh_phi", H5T_NATIVE_FLOAT, P, D->phi, true, &f_phi))) return bail(rc);
    if ((rc = read_var(mesh.id, "mesh_xi", H5T_NATIVE_FLOAT, P, D->xi, true, &f_xi))) return bail(rc);
    if ((rc = read_var(mesh.id, "mesh_eta", H5T_NATIVE_FLOAT, P, D->eta, true, &f_eta))) return bail(rc);
    if ((rc = read_var(file.id, "stf_dump", H5T_NATIVE_DOUBLE, T, D->stf))) return bail(rc);
    if ((rc = read_var(file.id, "stf_d_dump", H5T_NATIVE_DOUBLE, T, D->stf_d))) return bail(rc);
    for (size_t i = 0; i < D->fem.size(); ++i)
        if (D->fem[i] < 0 || (hsize_t)D->fem[i] >= npoints) return bail(fail(MCK_EINVAL, "fem_mesh id out of range"));

    // snapshots: [elements][variables][jpol][ipol][snapshots] f32
    if (H5Lexists(file.id, "merged_snapshots", H5P_DEFAULT) <= 0)
        return bail(fail(MCK_EINVAL, "missing variable 'merged_snapshots' (not a merged database)"));
    Hid ds(H5Dopen2(file.id, "merged_snapshots", H5P_DEFAULT), H5Dclose);
    std::vector<hsize_t> sd;
    if ((rc = dims_of(ds.id, sd))) return bail(rc);
    if (sd.size() != 5 || sd[0] != nelem || sd[2] != 5 || sd[3] != 5 || sd[4] != (hsize_t)nt)
        return bail(fail(MCK_EINVAL, "merged_snapshots does not have shape (elements, variables, 5, 5, snapshots)"));
    const int nvar = (int)sd[1];
    if (nvar_expected > 0 && nvar != nvar_expected)
        return bail(fail(MCK_EINVAL, "merged_snapshots has " + std::to_string(nvar) + " variables, expected " +
                                         std::to_string(nvar_expected)));
    const size_t nfloat = (size_t)nelem * nvar * 25 * nt;
    const float *snap = nullptr;
    if (map_snapshots) {
        // contiguous little-endian f32 storage: map the file region
        Hid ft(H5Dget_type(ds.id), H5Tclose);
        Hid pl(H5Dget_create_plist(ds.id), H5Pclose);
        const haddr_t off = H5Dget_offset(ds.id);
        if (H5Pget_layout(pl.id) == H5D_CONTIGUOUS && off != HADDR_UNDEF &&
            H5Tequal(ft.id, H5T_IEEE_F32LE) > 0) {
            const int fd = open(path, O_RDONLY);
            if (fd >= 0) {
                const size_t pg = (size_t)sysconf(_SC_PAGESIZE);
                const size_t base = (size_t)off / pg * pg, lead = (size_t)off - base;
                const size_t len = lead + nfloat * sizeof(float);
                void *m = mmap(nullptr, len, PROT_READ, MAP_SHARED, fd, (off_t)base);
                close(fd);
                if (m != MAP_FAILED) {
                    D->map = m;
                    D->map_len = len;
                    snap = reinterpret_cast<const float *>(static_cast<const char *>(m) + lead);
                    out->mapped = 1;
                }
            }
        }
    }
    if (!snap) {
        D->snapshots.resize(nfloat);
        if (nfloat && H5Dread(ds.id, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, D->snapshots.data()) < 0)
            return bail(fail(MCK_EINVAL, "cannot read merged_snapshots"));
        snap = D->snapshots.data();
    }

    mck_db_t &db = out->db;
    db.npol = npol;
    db.nelem = (int32_t)nelem;
    db.npoints = (int32_t)npoints;
    db.mesh_s = D->mesh_s.data();
    db.mesh_z = D->mesh_z.data();
    db.mp_s = D->mp_s.data();
    db.mp_z = D->mp_z.data();
    db.fem_mesh = D->fem.data();
    db.eltype = D->eltype.data();
    db.axis = D->axis.data();
    db.nvar = nvar;
    db.snapshots = snap;
    db.amplitude = amp;
    out->nt = nt;
    out->dt = dt;
    out->planet_radius = radius_km * 1e3;
    out->source_shift = shift;
    out->source_depth_km = depth;
    out->file_version = version;
    out->stf_dump = D->stf.data();
    out->stf_d_dump = D->stf_d.data();
    out->sem_mesh = D->sem.data();
    out->mesh_vp = f_vp ? D->vp.data() : nullptr;
    out->mesh_vs = f_vs ? D->vs.data() : nullptr;
    out->mesh_rho = f_rho ? D->rho.data() : nullptr;
    out->mesh_phi = f_phi ? D->phi.data() : nullptr;
    out->mesh_xi = f_xi ? D->xi.data() : nullptr;
    out->mesh_eta = f_eta ? D->eta.data() : nullptr;
    out->priv = D;
    return MCK_OK;
}

int mck_db_write(const char *path, const mck_dbfile_t *f) {
    if (!path || !f) return fail(MCK_EINVAL, "null argument");
    const mck_db_t &db = f->db;
    if (db.npol != 4 || db.nelem <= 0 || db.npoints <= 0 || f->nt <= 0 || db.nvar <= 0 || !db.snapshots ||
        !db.mesh_s || !db.mesh_z || !db.mp_s || !db.mp_z || !db.fem_mesh || !db.eltype || !db.axis ||
        !f->stf_dump || !f->stf_d_dump)
        return fail(MCK_EINVAL, "incomplete database");
    Quiet quiet;
    Hid file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail(MCK_EINVAL, std::string("cannot create '") + path + "'");
    int rc;
    const int32_t version = f->file_version > 0 ? f->file_version : 3, npol = db.npol, nt = f->nt;
    const int32_t nelem = db.nelem, npoints = db.npoints;
    const double radius_km = f->planet_radius * 1e-3;
    const float shift = (float)f->source_shift, amp = (float)db.amplitude;
    if ((rc = write_attr(file.id, "file version", H5T_STD_I32LE, H5T_NATIVE_INT32, &version))) return rc;
    if ((rc = write_attr(file.id, "npol", H5T_STD_I32LE, H5T_NATIVE_INT32, &npol))) return rc;
    if ((rc = write_attr(file.id, "nelem_kwf_global", H5T_STD_I32LE, H5T_NATIVE_INT32, &nelem))) return rc;
    if ((rc = write_attr(file.id, "npoints", H5T_STD_I32LE, H5T_NATIVE_INT32, &npoints))) return rc;
    if ((rc = write_attr(file.id, "number of strain dumps", H5T_STD_I32LE, H5T_NATIVE_INT32, &nt))) return rc;
    if ((rc = write_attr(file.id, "strain dump sampling rate in sec", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &f->dt)))
        return rc;
    if ((rc = write_attr(file.id, "planet radius", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &radius_km))) return rc;
    if ((rc = write_attr(file.id, "source shift factor in sec", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, &shift)))
        return rc;
    if ((rc = write_attr(file.id, "scalar source magnitude", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, &amp))) return rc;
    if ((rc = write_attr(file.id, "source depth in km", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &f->source_depth_km)))
        return rc;
    if ((rc = write_attr_text(file.id, "source time function", "gauss_0"))) return rc;

    const hsize_t T = (hsize_t)nt, P = (hsize_t)npoints, E = (hsize_t)nelem;
    if ((rc = write_var(file.id, "stf_dump", H5T_IEEE_F32LE, H5T_NATIVE_DOUBLE, {T}, f->stf_dump))) return rc;
    if ((rc = write_var(file.id, "stf_d_dump", H5T_IEEE_F32LE, H5T_NATIVE_DOUBLE, {T}, f->stf_d_dump))) return rc;
    if ((rc = write_var(file.id, "merged_snapshots", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT,
                        {E, (hsize_t)db.nvar, 5, 5, T}, db.snapshots)))
        return rc;
    Hid mesh(H5Gcreate2(file.id, "Mesh", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (mesh.id < 0) return fail(MCK_EINVAL, "cannot create group 'Mesh'");
    struct V { const char *name; hid_t ft, mt; std::vector<hsize_t> d; const void *p; };
    const std::vector<V> vars = {
        {"mesh_S", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, {P}, db.mesh_s},
        {"mesh_Z", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, {P}, db.mesh_z},
        {"mp_mesh_S", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, {E}, db.mp_s},
        {"mp_mesh_Z", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, {E}, db.mp_z},
        {"eltype", H5T_STD_I32LE, H5T_NATIVE_INT32, {E}, db.eltype},
        {"axis", H5T_STD_I32LE, H5T_NATIVE_INT32, {E}, db.axis},
        {"fem_mesh", H5T_STD_I32LE, H5T_NATIVE_INT32, {E, 4}, db.fem_mesh},
        {"sem_mesh", H5T_STD_I32LE, H5T_NATIVE_INT32, {E, 5, 5}, f->sem_mesh},
        {"mesh_vp", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, {P}, f->mesh_vp},
        {"mesh_vs", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, {P}, f->mesh_vs},
        {"mesh_rho", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, {P}, f->mesh_rho},
        {"mesh_phi", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, {P}, f->mesh_phi},
        {"mesh_xi", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, {P}, f->mesh_xi},
        {"mesh_eta", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, {P}, f->mesh_eta},
    };
    for (const V &v : vars) {
        if (!v.p) {
            if (!strcmp(v.name, "sem_mesh")) return fail(MCK_EINVAL, "sem_mesh is required");
            continue;
        }
        if ((rc = write_var(mesh.id, v.name, v.ft, v.mt, v.d, v.p))) return rc;
    }
    if (H5Fflush(file.id, H5F_SCOPE_GLOBAL) < 0) return fail(MCK_EINVAL, "cannot flush the file");
    return MCK_OK;
}

}  // extern "C"

// =====================================================================================
// Output files (include/mck_output.h)
// =====================================================================================
namespace {

const char *xdmf_topology(const char *t) {
    // inversion_mesh.f90:2058-2073
    if (!strcmp(t, "tri")) return "Triangle";
    if (!strcmp(t, "tet")) return "Tetrahedron";
    if (!strcmp(t, "quad")) return "Quadrilateral";
    if (!strcmp(t, "hex") || !strcmp(t, "vox")) return "Hexahedron";
    return nullptr;
}

// Fortran E15.8 of a single-precision value: "0.dddddddd" mantissa and a two-digit
// exponent, right-aligned in 15 characters
std::string fortran_e15_8(float v) {
    char buf[64];
    if (v == 0.0f) {
        snprintf(buf, sizeof buf, "%15s", "0.00000000E+00");
        return buf;
    }
    char d[32];
    snprintf(d, sizeof d, "%.7E", (double)std::fabs(v));   // d.dddddddE+xx
    const int ex = atoi(d + 10) + 1;
    char m[16];
    snprintf(m, sizeof m, "0.%c%.7s", d[0], d + 2);
    char e[16];
    snprintf(e, sizeof e, "E%c%02d", ex < 0 ? '-' : '+', std::abs(ex));
    snprintf(buf, sizeof buf, "%15s", (std::string(v < 0 ? "-" : "") + m + e).c_str());
    return buf;
}

}  // namespace

extern "C" {

int mck_write_xdmf(const char *filename, const char *element_type, int32_t nvertices, const double *vertices,
                   int32_t nelements, int32_t nvpe, const int32_t *connectivity, const int32_t *block_id,
                   int32_t nvars, const mck_outvar_t *vars) {
    if (!filename || !element_type || !vertices || !connectivity || nvertices <= 0 || nelements <= 0 || nvpe <= 0 ||
        nvars < 0 || (nvars && !vars))
        return fail(MCK_EINVAL, "invalid argument");
    const char *topo = xdmf_topology(element_type);
    if (!topo) return fail(MCK_EINVAL, std::string("xdmf output for element type '") + element_type + "' not implemented");
    for (int i = 0; i < nvars; ++i)
        if (!vars[i].name || !vars[i].values || vars[i].nentries <= 0)
            return fail(MCK_EINVAL, "incomplete output variable");
    const std::string base(filename), nc = base + ".nc";
    const size_t slash = base.rfind('/');
    const std::string nc_rel = (slash == std::string::npos ? base : base.substr(slash + 1)) + ".nc";
    {
        Quiet quiet;
        Hid file(H5Fcreate(nc.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
        if (file.id < 0) return fail(MCK_EINVAL, "cannot create '" + nc + "'");
        const hsize_t V = (hsize_t)nvertices, E = (hsize_t)nelements, NV = (hsize_t)nvpe;
        std::vector<float> pts((size_t)nvertices * 3);
        for (size_t i = 0; i < pts.size(); ++i) pts[i] = (float)vertices[i];
        std::vector<int32_t> grid((size_t)nelements * nvpe), block(nelements, 1);
        for (size_t i = 0; i < grid.size(); ++i) {
            if (connectivity[i] < 1 || connectivity[i] > nvertices) return fail(MCK_EINVAL, "connectivity out of range");
            grid[i] = connectivity[i] - 1;
        }
        if (block_id) std::copy(block_id, block_id + nelements, block.begin());
        int rc;
        if ((rc = write_var(file.id, "points", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, {V, 3}, pts.data()))) return rc;
        if ((rc = write_var(file.id, "grid", H5T_STD_I32LE, H5T_NATIVE_INT32, {E, NV}, grid.data()))) return rc;
        if ((rc = write_var(file.id, "block", H5T_STD_I32LE, H5T_NATIVE_INT32, {E}, block.data()))) return rc;
        for (int cells = 0; cells < 2; ++cells) {
            bool any = false;
            for (int i = 0; i < nvars; ++i) any |= (vars[i].on_cells != 0) == (cells != 0);
            if (!any) continue;
            const char *gname = cells ? "cell_data" : "node_data";
            Hid g(H5Gcreate2(file.id, gname, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
            if (g.id < 0) return fail(MCK_EINVAL, std::string("cannot create group '") + gname + "'");
            for (int i = 0; i < nvars; ++i) {
                if ((vars[i].on_cells != 0) != (cells != 0)) continue;
                if ((rc = write_var(g.id, vars[i].name, H5T_IEEE_F32LE, H5T_NATIVE_FLOAT,
                                    {(hsize_t)vars[i].nentries, cells ? E : V}, vars[i].values)))
                    return rc;
            }
        }
        if (H5Fflush(file.id, H5F_SCOPE_GLOBAL) < 0) return fail(MCK_EINVAL, "cannot flush '" + nc + "'");
    }
    // the XDMF description (formats 733, 7341, 73421, 73422, 7343, 736 of
    // inversion_mesh.f90:2163-2232)
    FILE *x = fopen((base + ".xdmf").c_str(), "w");
    if (!x) return fail(MCK_EINVAL, "cannot create '" + base + ".xdmf'");
    fprintf(x,
            "<?xml version=\"1.0\" ?>\n"
            "<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n"
            "<Xdmf xmlns:xi=\"http://www.w3.org/2003/XInclude\" Version=\"2.2\">\n"
            "<Domain>\n\n"
            "<DataItem Name=\"grid\" Dimensions=\"%10d%3d\" NumberType=\"Int\" Format=\"hdf\">\n"
            "  %s:/grid\n</DataItem>\n\n"
            "<DataItem Name=\"points\" Dimensions=\"%10d 3\" NumberType=\"Float\" Format=\"hdf\">\n"
            "  %s:/points\n</DataItem>\n\n"
            "<DataItem Name=\"block\" Dimensions=\"%10d\" NumberType=\"Int\" Format=\"hdf\">\n"
            "  %s:/block\n</DataItem>\n\n"
            "<Grid Name=\"CellsTime\" GridType=\"Collection\" CollectionType=\"Temporal\">\n\n",
            nelements, nvpe, nc_rel.c_str(), nvertices, nc_rel.c_str(), nelements, nc_rel.c_str());
    fprintf(x,
            "    <Grid Name=\"grid\" GridType=\"Uniform\">\n"
            "        <Time Value=\"%8.2f\" />\n"
            "        <Topology TopologyType=\"%s\" NumberOfElements=\"%10d\">\n"
            "            <DataItem Reference=\"XML\">\n"
            "                /Xdmf/Domain/DataItem[@Name=\"grid\"]\n"
            "            </DataItem>\n"
            "        </Topology>\n"
            "        <Geometry GeometryType=\"XYZ\">\n"
            "            <DataItem Reference=\"XML\">\n"
            "                /Xdmf/Domain/DataItem[@Name=\"points\"]\n"
            "            </DataItem>\n"
            "        </Geometry>\n"
            "        <Attribute Name=\"blockid\" AttributeType=\"Scalar\" Center=\"Cell\">\n"
            "            <DataItem Reference=\"XML\">\n"
            "                /Xdmf/Domain/DataItem[@Name=\"block\"]\n"
            "            </DataItem>\n"
            "        </Attribute>\n",
            0.0, topo, nelements);
    // node variables first, then cell variables (inversion_mesh.f90:2092-2118)
    for (int cells = 0; cells < 2; ++cells)
        for (int i = 0; i < nvars; ++i) {
            const mck_outvar_t &v = vars[i];
            if ((v.on_cells != 0) != (cells != 0)) continue;
            const int n = cells ? nelements : nvertices;
            for (int j = 0; j < v.nentries; ++j)
                fprintf(x,
                        "        <Attribute Name=\"%s\" AttributeType=\"Scalar\" Center=\"%s\">\n"
                        "            <DataItem ItemType=\"HyperSlab\" Dimensions=\"%10d\" Type=\"HyperSlab\">\n"
                        "                <DataItem Dimensions=\"3 2\" Format=\"XML\">\n"
                        "                    %10d          0 \n"
                        "                             1          1 \n"
                        "                             1 %10d\n"
                        "                </DataItem>\n"
                        "                <DataItem Dimensions=\"%10d%10d\" NumberType=\"Float\" Format=\"hdf\">\n"
                        "                   %s:/%s/%s\n"
                        "                </DataItem>\n"
                        "            </DataItem>\n"
                        "        </Attribute>\n",
                        v.entry_names ? v.entry_names[j] : v.name, cells ? "Cell" : "Node", n, j, n, v.nentries, n,
                        nc_rel.c_str(), cells ? "cell_data" : "node_data", v.name);
        }
    fprintf(x, "    </Grid>\n\n</Grid>\n</Domain>\n</Xdmf>\n");
    if (fclose(x) != 0) return fail(MCK_EINVAL, "cannot write '" + base + ".xdmf'");
    return MCK_OK;
}

int mck_write_csr(const char *filename, int32_t npoints, int32_t nkernel, const float *values, double threshold,
                  int32_t ascii) {
    if (!filename || npoints <= 0 || nkernel <= 0 || !values) return fail(MCK_EINVAL, "invalid argument");
    const std::string base(filename);
    FILE *fv = fopen((base + ".val").c_str(), ascii ? "w" : "wb");
    FILE *fi = fopen((base + ".ind").c_str(), ascii ? "w" : "wb");
    FILE *fp = fopen((base + ".pnt").c_str(), "w");
    if (!fv || !fi || !fp) {
        if (fv) fclose(fv);
        if (fi) fclose(fi);
        if (fp) fclose(fp);
        return fail(MCK_EINVAL, "cannot create the files '" + base + ".val/.ind/.pnt'");
    }
    // inversion_mesh.f90:1886-2022: kernels outer, points inner, |value| > threshold
    int32_t n = 0;
    bool ok = true;
    for (int k = 0; k < nkernel; ++k) {
        for (int32_t p = 0; p < npoints; ++p) {
            const float v = values[(size_t)k * npoints + p];
            if (!(std::fabs((double)v) > threshold)) continue;
            ++n;
            const int32_t idx = p + 1;
            if (ascii) {
                ok &= fprintf(fv, " %s", fortran_e15_8(v).c_str()) > 0;
                ok &= fprintf(fi, " %15d", idx) > 0;
            } else {
                ok &= fwrite(&v, 4, 1, fv) == 1;
                ok &= fwrite(&idx, 4, 1, fi) == 1;
            }
        }
        if (ascii) {
            ok &= fputc('\n', fv) != EOF;
            ok &= fputc('\n', fi) != EOF;
        }
        // list-directed integer: a blank and the value in 11 columns
        ok &= fprintf(fp, "%12d\n", n) > 0;
    }
    ok &= fclose(fv) == 0;
    ok &= fclose(fi) == 0;
    ok &= fclose(fp) == 0;
    return ok ? MCK_OK : fail(MCK_EINVAL, "cannot write the files '" + base + ".val/.ind/.pnt'");
}

int mck_write_intermediate(const char *path, int32_t npoints, int32_t nelements, int32_t nkernel, int32_t nmodel,
                           int32_t nhet, const float *kernel, const float *variance, const float *bg_model,
                           const float *het_model, const int32_t *niterations, const int32_t *element_proc) {
    if (!path || npoints <= 0 || nelements <= 0 || nkernel <= 0 || nmodel < 0 || nhet < 0 || !kernel ||
        !variance || !niterations || !element_proc || (nmodel && !bg_model) || (nhet && !het_model))
        return fail(MCK_EINVAL, "invalid argument");
    Quiet quiet;
    Hid file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail(MCK_EINVAL, std::string("cannot create '") + path + "'");
    const hsize_t P = (hsize_t)npoints, E = (hsize_t)nelements, K = (hsize_t)nkernel;
    int rc;
    if ((rc = write_var(file.id, "K_x", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, {K, P}, kernel))) return rc;
    if ((rc = write_var(file.id, "Var", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, {K, P}, variance))) return rc;
    if (nmodel && (rc = write_var(file.id, "Bg_Model", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, {(hsize_t)nmodel, P},
                                  bg_model)))
        return rc;
    if (nhet && (rc = write_var(file.id, "Het_Model", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, {(hsize_t)nhet, P},
                                het_model)))
        return rc;
    if ((rc = write_var(file.id, "niterations", H5T_STD_I32LE, H5T_NATIVE_INT32, {K, E}, niterations))) return rc;
    if ((rc = write_var(file.id, "element_proc", H5T_STD_I32LE, H5T_NATIVE_INT32, {E}, element_proc))) return rc;
    if (H5Fflush(file.id, H5F_SCOPE_GLOBAL) < 0) return fail(MCK_EINVAL, "cannot flush the file");
    return MCK_OK;
}

int mck_read_intermediate(const char *path, int32_t npoints, int32_t nelements, int32_t nkernel, int32_t nmodel,
                          int32_t nhet, float *kernel, float *variance, float *bg_model, float *het_model,
                          int32_t *niterations, int32_t *element_proc) {
    if (!path || npoints <= 0 || nelements <= 0 || nkernel <= 0 || nmodel < 0 || nhet < 0)
        return fail(MCK_EINVAL, "invalid argument");
    Quiet quiet;
    Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail(MCK_EINVAL, std::string("cannot open '") + path + "'");
    const hsize_t P = (hsize_t)npoints, E = (hsize_t)nelements, K = (hsize_t)nkernel;
    std::vector<float> f;
    std::vector<int32_t> iv;
    int rc;
    struct FV { const char *name; std::vector<hsize_t> d; float *out; };
    for (const FV &v : {FV{"K_x", {K, P}, kernel}, FV{"Var", {K, P}, variance},
                        FV{"Bg_Model", {(hsize_t)nmodel, P}, nmodel ? bg_model : nullptr},
                        FV{"Het_Model", {(hsize_t)nhet, P}, nhet ? het_model : nullptr}}) {
        if (!v.out) continue;
        if ((rc = read_var(file.id, v.name, H5T_NATIVE_FLOAT, v.d, f))) return rc;
        std::copy(f.begin(), f.end(), v.out);
    }
    if (niterations) {
        if ((rc = read_var(file.id, "niterations", H5T_NATIVE_INT32, {K, E}, iv))) return rc;
        std::copy(iv.begin(), iv.end(), niterations);
    }
    if (element_proc) {
        if ((rc = read_var(file.id, "element_proc", H5T_NATIVE_INT32, {E}, iv))) return rc;
        std::copy(iv.begin(), iv.end(), element_proc);
    }
    return MCK_OK;
}

}  // extern "C"
