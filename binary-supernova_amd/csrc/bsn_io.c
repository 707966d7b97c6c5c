/* bsn_io.c -- snapshot and checkpoint files of Binary-Supernova (HDF5, the
 * reference's layout) streamed between a GPU-resident grid and disk.
 *
 * The grid's state reaches the file slab by slab: libbsn_hip packs each slab
 * of k-planes into the file's array order on the device and hands it over in
 * a pinned buffer (bsn_grid_export_slabs), and this file writes it as a
 * hyperslab of the dataset while the next slab crosses PCIe.  Reading runs the
 * same pipeline backwards.  Plain C on the HDF5 C API.
 */
#include <hdf5.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bsn.h"
#include "bsn_io.h"

#define NG 3
#define CHECKPOINT_FORMAT_VERSION 1   /* src/checkpoint.jl:31 */

static _Thread_local char io_err[512];

const char* bsn_io_last_error(void) { return io_err; }

static int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(io_err, sizeof io_err, fmt, ap);
    va_end(ap);
    return code;
}

static void quiet(void) { H5Eset_auto2(H5E_DEFAULT, NULL, NULL); }

/* ------------------------------------------------------------------------ */
/* attributes                                                                */

static int put_f64(hid_t loc, const char* name, double v) {
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t a = H5Acreate2(loc, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT);
    herr_t e = a < 0 ? -1 : H5Awrite(a, H5T_NATIVE_DOUBLE, &v);
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    return e < 0 ? fail(BSN_ERR_IO, "cannot write attribute %s", name) : BSN_OK;
}

static int put_i64(hid_t loc, const char* name, int64_t v) {
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t a = H5Acreate2(loc, name, H5T_STD_I64LE, sp, H5P_DEFAULT, H5P_DEFAULT);
    herr_t e = a < 0 ? -1 : H5Awrite(a, H5T_NATIVE_INT64, &v);
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    return e < 0 ? fail(BSN_ERR_IO, "cannot write attribute %s", name) : BSN_OK;
}

static int put_str(hid_t loc, const char* name, const char* v) {
    hid_t t = H5Tcopy(H5T_C_S1);
    H5Tset_size(t, strlen(v) + 1);
    H5Tset_cset(t, H5T_CSET_UTF8);
    H5Tset_strpad(t, H5T_STR_NULLTERM);
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t a = H5Acreate2(loc, name, t, sp, H5P_DEFAULT, H5P_DEFAULT);
    herr_t e = a < 0 ? -1 : H5Awrite(a, t, v);
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    H5Tclose(t);
    return e < 0 ? fail(BSN_ERR_IO, "cannot write attribute %s", name) : BSN_OK;
}

static int get_f64(hid_t loc, const char* name, double* v) {
    hid_t a = H5Aopen(loc, name, H5P_DEFAULT);
    if (a < 0) return fail(BSN_ERR_IO, "attribute %s missing", name);
    herr_t e = H5Aread(a, H5T_NATIVE_DOUBLE, v);
    H5Aclose(a);
    return e < 0 ? fail(BSN_ERR_IO, "cannot read attribute %s", name) : BSN_OK;
}

static int get_i64(hid_t loc, const char* name, int64_t* v) {
    hid_t a = H5Aopen(loc, name, H5P_DEFAULT);
    if (a < 0) return fail(BSN_ERR_IO, "attribute %s missing", name);
    herr_t e = H5Aread(a, H5T_NATIVE_INT64, v);
    H5Aclose(a);
    return e < 0 ? fail(BSN_ERR_IO, "cannot read attribute %s", name) : BSN_OK;
}

/* A string attribute, fixed-length or variable-length, into buf. */
static int get_str(hid_t loc, const char* name, char* buf, size_t len) {
    hid_t a = H5Aopen(loc, name, H5P_DEFAULT);
    if (a < 0) return fail(BSN_ERR_IO, "attribute %s missing", name);
    hid_t ft = H5Aget_type(a);
    int rc = BSN_OK;
    if (H5Tget_class(ft) != H5T_STRING) {
        rc = fail(BSN_ERR_IO, "attribute %s is not a string", name);
    } else if (H5Tis_variable_str(ft) > 0) {
        hid_t mt = H5Tcopy(H5T_C_S1);
        H5Tset_size(mt, H5T_VARIABLE);
        H5Tset_cset(mt, H5Tget_cset(ft));     /* no conversion between character sets */
        char* s = NULL;
        if (H5Aread(a, mt, &s) < 0 || !s) rc = fail(BSN_ERR_IO, "cannot read attribute %s", name);
        else { snprintf(buf, len, "%s", s); H5free_memory(s); }
        H5Tclose(mt);
    } else {
        size_t n = H5Tget_size(ft);
        char* tmp = calloc(n + 1, 1);
        hid_t mt = H5Tcopy(H5T_C_S1);
        H5Tset_size(mt, n + 1);
        H5Tset_strpad(mt, H5T_STR_NULLTERM);
        H5Tset_cset(mt, H5Tget_cset(ft));
        if (!tmp || H5Aread(a, mt, tmp) < 0) rc = fail(BSN_ERR_IO, "cannot read attribute %s", name);
        else snprintf(buf, len, "%s", tmp);
        H5Tclose(mt);
        free(tmp);
    }
    H5Tclose(ft);
    H5Aclose(a);
    return rc;
}

#define TRY(x) do { int rc_ = (x); if (rc_ != BSN_OK) { rc = rc_; goto done; } } while (0)

/* ------------------------------------------------------------------------ */
/* datasets                                                                  */

static hid_t make_dataset(hid_t loc, const char* name, int rank, const hsize_t* dims) {
    hid_t sp = H5Screate_simple(rank, dims, NULL);
    hid_t d = H5Dcreate2(loc, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5Sclose(sp);
    return d;
}

static int write_all(hid_t loc, const char* name, int rank, const hsize_t* dims, const double* v) {
    hid_t d = make_dataset(loc, name, rank, dims);
    if (d < 0) return fail(BSN_ERR_IO, "cannot create dataset %s", name);
    herr_t e = H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v);
    H5Dclose(d);
    return e < 0 ? fail(BSN_ERR_IO, "cannot write dataset %s", name) : BSN_OK;
}

/* Opens dataset name and checks it has rank and dims. */
static hid_t open_checked(hid_t loc, const char* name, int rank, const hsize_t* dims) {
    hid_t d = H5Dopen2(loc, name, H5P_DEFAULT);
    if (d < 0) { fail(BSN_ERR_IO, "dataset %s missing", name); return -1; }
    hid_t sp = H5Dget_space(d);
    hsize_t got[8];
    int r = H5Sget_simple_extent_ndims(sp);
    int ok = r == rank && r <= 8;
    if (ok) {
        H5Sget_simple_extent_dims(sp, got, NULL);
        for (int a = 0; a < rank; ++a) ok = ok && got[a] == dims[a];
    }
    H5Sclose(sp);
    if (!ok) {
        H5Dclose(d);
        fail(BSN_ERR_IO, "dataset %s has the wrong shape", name);
        return -1;
    }
    return d;
}

static int read_all(hid_t loc, const char* name, int rank, const hsize_t* dims, double* v) {
    hid_t d = open_checked(loc, name, rank, dims);
    if (d < 0) return BSN_ERR_IO;
    herr_t e = H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v);
    H5Dclose(d);
    return e < 0 ? fail(BSN_ERR_IO, "cannot read dataset %s", name) : BSN_OK;
}

/* Slabs of k-planes (the first, slowest dimension) of datasets[part]. */
struct slab_io {
    hid_t ds[5];
    int rank;
    hsize_t dims[4];
    int writing;
};

static int slab_cb(void* user, int part, int64_t k0, int64_t nk, double* data, int64_t count) {
    struct slab_io* S = (struct slab_io*)user;
    hsize_t start[4] = {0, 0, 0, 0}, cnt[4];
    for (int a = 0; a < S->rank; ++a) cnt[a] = S->dims[a];
    start[0] = (hsize_t)k0;
    cnt[0] = (hsize_t)nk;
    hsize_t n = 1;
    for (int a = 0; a < S->rank; ++a) n *= cnt[a];
    if ((int64_t)n != count) return 1;
    hid_t fs = H5Dget_space(S->ds[part]);
    hid_t ms = H5Screate_simple(1, &n, NULL);
    herr_t e = H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, NULL, cnt, NULL);
    if (e >= 0)
        e = S->writing ? H5Dwrite(S->ds[part], H5T_NATIVE_DOUBLE, ms, fs, H5P_DEFAULT, data)
                       : H5Dread(S->ds[part], H5T_NATIVE_DOUBLE, ms, fs, H5P_DEFAULT, data);
    H5Sclose(ms);
    H5Sclose(fs);
    return e < 0 ? 1 : 0;
}

static int stream_error(int rc, const char* what) {
    return fail(rc, "%s: %s", what, bsn_last_error());
}

/* ------------------------------------------------------------------------ */
/* snapshots (src/io.jl:29-84)                                               */

static const char* GAS[5] = {"rho", "rhovx", "rhovy", "rhovz", "E"};

static int snapshot_attrs(hid_t f, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                          double dz, double t, double gamma) {
    int rc = BSN_OK;
    TRY(put_f64(f, "t", t));
    TRY(put_i64(f, "nx", nx));
    TRY(put_i64(f, "ny", ny));
    TRY(put_i64(f, "nz", nz));
    TRY(put_f64(f, "dx", dx));
    TRY(put_f64(f, "dy", dy));
    TRY(put_f64(f, "dz", dz));
    TRY(put_f64(f, "gamma", gamma));
done:
    return rc;
}

int bsn_write_snapshot(bsn_grid* g, const char* filename, double dx, double dy, double dz,
                       double t, double gamma) {
    if (!g || !filename) return fail(BSN_ERR_ARG, "NULL argument");
    int64_t n[3];
    if (bsn_grid_dims(g, n) != BSN_OK) return stream_error(BSN_ERR_ARG, "grid");
    quiet();
    int rc = BSN_OK;
    struct slab_io S = {{-1, -1, -1, -1, -1}, 3, {(hsize_t)n[2], (hsize_t)n[1], (hsize_t)n[0], 0}, 1};
    hid_t gas = -1;
    hid_t f = H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail(BSN_ERR_IO, "cannot create %s", filename);
    gas = H5Gcreate2(f, "gas", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (gas < 0) { rc = fail(BSN_ERR_IO, "cannot create group gas"); goto done; }
    for (int q = 0; q < 5; ++q) {
        S.ds[q] = make_dataset(gas, GAS[q], 3, S.dims);
        if (S.ds[q] < 0) { rc = fail(BSN_ERR_IO, "cannot create dataset %s", GAS[q]); goto done; }
    }
    int e = bsn_grid_export_slabs(g, BSN_LAYOUT_FIELDS, 0, slab_cb, &S);
    if (e != BSN_OK) { rc = stream_error(e, "streaming the state"); goto done; }
    TRY(snapshot_attrs(f, n[0], n[1], n[2], dx, dy, dz, t, gamma));
done:
    for (int q = 0; q < 5; ++q) if (S.ds[q] >= 0) H5Dclose(S.ds[q]);
    if (gas >= 0) H5Gclose(gas);
    if (H5Fclose(f) < 0 && rc == BSN_OK) rc = fail(BSN_ERR_IO, "cannot close %s", filename);
    return rc;
}

int bsn_write_snapshot_host(const char* filename, const double* U, int64_t nx, int64_t ny,
                            int64_t nz, double dx, double dy, double dz, double t, double gamma) {
    if (!filename || !U || nx < 1 || ny < 1 || nz < 1) return fail(BSN_ERR_ARG, "bad arguments");
    quiet();
    const int64_t Nx = nx + 2 * NG, Ny = ny + 2 * NG;
    double* buf = malloc(sizeof(double) * nx * ny * nz);
    if (!buf) return fail(BSN_ERR_ARG, "out of host memory");
    int rc = BSN_OK;
    hid_t gas = -1;
    hid_t f = H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) { free(buf); return fail(BSN_ERR_IO, "cannot create %s", filename); }
    gas = H5Gcreate2(f, "gas", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (gas < 0) { rc = fail(BSN_ERR_IO, "cannot create group gas"); goto done; }
    const hsize_t dims[3] = {(hsize_t)nz, (hsize_t)ny, (hsize_t)nx};
    for (int q = 0; q < 5; ++q) {
        for (int64_t k = 0; k < nz; ++k)
            for (int64_t j = 0; j < ny; ++j)
                for (int64_t i = 0; i < nx; ++i)
                    buf[(k * ny + j) * nx + i] = U[(((k + NG) * Ny + j + NG) * Nx + i + NG) * 5 + q];
        TRY(write_all(gas, GAS[q], 3, dims, buf));
    }
    TRY(snapshot_attrs(f, nx, ny, nz, dx, dy, dz, t, gamma));
done:
    free(buf);
    if (gas >= 0) H5Gclose(gas);
    if (H5Fclose(f) < 0 && rc == BSN_OK) rc = fail(BSN_ERR_IO, "cannot close %s", filename);
    return rc;
}

int bsn_snapshot_info(const char* filename, int64_t n[3], double meta[5]) {
    if (!filename || !n || !meta) return fail(BSN_ERR_ARG, "NULL argument");
    quiet();
    hid_t f = H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail(BSN_ERR_IO, "cannot open %s", filename);
    int rc = BSN_OK;
    TRY(get_i64(f, "nx", &n[0]));
    TRY(get_i64(f, "ny", &n[1]));
    TRY(get_i64(f, "nz", &n[2]));
    TRY(get_f64(f, "dx", &meta[0]));
    TRY(get_f64(f, "dy", &meta[1]));
    TRY(get_f64(f, "dz", &meta[2]));
    TRY(get_f64(f, "t", &meta[3]));
    TRY(get_f64(f, "gamma", &meta[4]));
done:
    H5Fclose(f);
    return rc;
}

int bsn_read_snapshot(bsn_grid* g, const char* filename) {
    if (!g || !filename) return fail(BSN_ERR_ARG, "NULL argument");
    int64_t n[3], m[3];
    double meta[5];
    int rc = bsn_snapshot_info(filename, m, meta);
    if (rc != BSN_OK) return rc;
    if (bsn_grid_dims(g, n) != BSN_OK) return stream_error(BSN_ERR_ARG, "grid");
    if (n[0] != m[0] || n[1] != m[1] || n[2] != m[2])
        return fail(BSN_ERR_ARG, "snapshot is %lldx%lldx%lld, grid is %lldx%lldx%lld",
                    (long long)m[0], (long long)m[1], (long long)m[2], (long long)n[0],
                    (long long)n[1], (long long)n[2]);
    quiet();
    struct slab_io S = {{-1, -1, -1, -1, -1}, 3, {(hsize_t)n[2], (hsize_t)n[1], (hsize_t)n[0], 0}, 0};
    hid_t f = H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail(BSN_ERR_IO, "cannot open %s", filename);
    hid_t gas = H5Gopen2(f, "gas", H5P_DEFAULT);
    if (gas < 0) { rc = fail(BSN_ERR_IO, "group gas missing"); goto done; }
    for (int q = 0; q < 5; ++q) {
        S.ds[q] = open_checked(gas, GAS[q], 3, S.dims);
        if (S.ds[q] < 0) { rc = BSN_ERR_IO; goto done; }
    }
    int e = bsn_grid_import_slabs(g, BSN_LAYOUT_FIELDS, 0, slab_cb, &S);
    if (e != BSN_OK) rc = stream_error(e, "streaming the state");
done:
    for (int q = 0; q < 5; ++q) if (S.ds[q] >= 0) H5Dclose(S.ds[q]);
    if (gas >= 0) H5Gclose(gas);
    H5Fclose(f);
    return rc;
}

int bsn_read_snapshot_host(const char* filename, double* U_active) {
    if (!filename || !U_active) return fail(BSN_ERR_ARG, "NULL argument");
    int64_t n[3];
    double meta[5];
    int rc = bsn_snapshot_info(filename, n, meta);
    if (rc != BSN_OK) return rc;
    const int64_t nx = n[0], ny = n[1], nz = n[2];
    double* buf = malloc(sizeof(double) * nx * ny * nz);
    if (!buf) return fail(BSN_ERR_ARG, "out of host memory");
    quiet();
    hid_t f = H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) { free(buf); return fail(BSN_ERR_IO, "cannot open %s", filename); }
    hid_t gas = H5Gopen2(f, "gas", H5P_DEFAULT);
    if (gas < 0) { rc = fail(BSN_ERR_IO, "group gas missing"); goto done; }
    const hsize_t dims[3] = {(hsize_t)nz, (hsize_t)ny, (hsize_t)nx};
    for (int q = 0; q < 5; ++q) {
        TRY(read_all(gas, GAS[q], 3, dims, buf));
        for (int64_t c = 0; c < nx * ny * nz; ++c) U_active[c * 5 + q] = buf[c];
    }
done:
    free(buf);
    if (gas >= 0) H5Gclose(gas);
    H5Fclose(f);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* uniform checkpoints (src/checkpoint.jl:202-318)                           */

static const char* BODY_PROPS[4] = {"mass", "eps", "c_code", "r_floor"};

static int write_bodies(hid_t f, int nb, const double* pos, const double* vel, const double* props) {
    int rc = BSN_OK;
    double* col = NULL;
    hid_t b = H5Gcreate2(f, "bhs", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (b < 0) return fail(BSN_ERR_IO, "cannot create group bhs");
    TRY(put_i64(b, "n", nb));
    if (nb > 0) {
        const hsize_t d2[2] = {(hsize_t)nb, 3}, d1[1] = {(hsize_t)nb};
        TRY(write_all(b, "pos", 2, d2, pos));
        TRY(write_all(b, "vel", 2, d2, vel));
        col = malloc(sizeof(double) * nb);
        if (!col) { rc = fail(BSN_ERR_ARG, "out of host memory"); goto done; }
        for (int p = 0; p < 4; ++p) {
            for (int i = 0; i < nb; ++i) col[i] = props[4 * i + p];
            TRY(write_all(b, BODY_PROPS[p], 1, d1, col));
        }
    }
done:
    free(col);
    H5Gclose(b);
    return rc;
}

static int read_bodies(hid_t f, int max_b, int* nb_out, double* pos, double* vel, double* props) {
    int rc = BSN_OK;
    double* col = NULL;
    hid_t b = H5Gopen2(f, "bhs", H5P_DEFAULT);
    if (b < 0) return fail(BSN_ERR_IO, "group bhs missing");
    int64_t n = 0;
    TRY(get_i64(b, "n", &n));
    if (n < 0 || n > max_b) { rc = fail(BSN_ERR_ARG, "%lld bodies stored, room for %d", (long long)n, max_b); goto done; }
    *nb_out = (int)n;
    if (n > 0) {
        const hsize_t d2[2] = {(hsize_t)n, 3}, d1[1] = {(hsize_t)n};
        TRY(read_all(b, "pos", 2, d2, pos));
        TRY(read_all(b, "vel", 2, d2, vel));
        col = malloc(sizeof(double) * n);
        if (!col) { rc = fail(BSN_ERR_ARG, "out of host memory"); goto done; }
        for (int p = 0; p < 4; ++p) {
            TRY(read_all(b, BODY_PROPS[p], 1, d1, col));
            for (int64_t i = 0; i < n; ++i) props[4 * i + p] = col[i];
        }
    }
done:
    free(col);
    H5Gclose(b);
    return rc;
}

static const char* GRID_F64[9] = {"dx", "dy", "dz", "gamma", "x0", "y0", "z0", "rho_floor",
                                  "P_floor"};

int bsn_save_checkpoint_uniform(bsn_grid* g, const double* U_host, int64_t nx, int64_t ny,
                                int64_t nz, const char* path, int64_t step, const double meta[4],
                                const double params[9], int n_bodies, const double* pos,
                                const double* vel, const double* props) {
    if (!path || !meta || !params || n_bodies < 0 || (n_bodies > 0 && (!pos || !vel || !props)))
        return fail(BSN_ERR_ARG, "bad arguments");
    if (g) {
        int64_t n[3];
        if (bsn_grid_dims(g, n) != BSN_OK) return stream_error(BSN_ERR_ARG, "grid");
        nx = n[0]; ny = n[1]; nz = n[2];
    } else if (!U_host || nx < 1 || ny < 1 || nz < 1) {
        return fail(BSN_ERR_ARG, "no state to save");
    }
    quiet();
    int rc = BSN_OK;
    hid_t m = -1, gr = -1;
    struct slab_io S = {{-1, -1, -1, -1, -1}, 4,
                        {(hsize_t)(nz + 2 * NG), (hsize_t)(ny + 2 * NG), (hsize_t)(nx + 2 * NG), 5}, 1};
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail(BSN_ERR_IO, "cannot create %s", path);
    m = H5Gcreate2(f, "manifest", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (m < 0) { rc = fail(BSN_ERR_IO, "cannot create group manifest"); goto done; }
    TRY(put_i64(m, "format_version", CHECKPOINT_FORMAT_VERSION));
    TRY(put_str(m, "kind", "uniform"));
    TRY(put_f64(m, "t", meta[0]));
    TRY(put_i64(m, "step", step));
    TRY(put_f64(m, "dt_last", meta[1]));
    TRY(put_f64(m, "t_snap", meta[2]));
    TRY(put_f64(m, "t_bh2_sink_on", meta[3]));
    gr = H5Gcreate2(f, "grid", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (gr < 0) { rc = fail(BSN_ERR_IO, "cannot create group grid"); goto done; }
    TRY(put_i64(gr, "nx", nx));
    TRY(put_i64(gr, "ny", ny));
    TRY(put_i64(gr, "nz", nz));
    for (int a = 0; a < 9; ++a) TRY(put_f64(gr, GRID_F64[a], params[a]));
    if (g) {
        S.ds[0] = make_dataset(gr, "U", 4, S.dims);
        if (S.ds[0] < 0) { rc = fail(BSN_ERR_IO, "cannot create dataset U"); goto done; }
        int e = bsn_grid_export_slabs(g, BSN_LAYOUT_REF, 0, slab_cb, &S);
        if (e != BSN_OK) { rc = stream_error(e, "streaming the state"); goto done; }
    } else {
        TRY(write_all(gr, "U", 4, S.dims, U_host));
    }
    TRY(write_bodies(f, n_bodies, pos, vel, props));
done:
    if (S.ds[0] >= 0) H5Dclose(S.ds[0]);
    if (gr >= 0) H5Gclose(gr);
    if (m >= 0) H5Gclose(m);
    if (H5Fclose(f) < 0 && rc == BSN_OK) rc = fail(BSN_ERR_IO, "cannot close %s", path);
    return rc;
}

/* Opens a uniform checkpoint after checking its manifest. */
static hid_t open_uniform(const char* path) {
    quiet();
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) { fail(BSN_ERR_IO, "cannot open %s", path); return -1; }
    hid_t m = H5Gopen2(f, "manifest", H5P_DEFAULT);
    int64_t fv = -1;
    char kind[64] = {0};
    int ok = m >= 0 && get_i64(m, "format_version", &fv) == BSN_OK;
    if (ok && fv != CHECKPOINT_FORMAT_VERSION) {
        fail(BSN_ERR_IO, "checkpoint format_version=%lld, expected %d", (long long)fv,
             CHECKPOINT_FORMAT_VERSION);
        ok = 0;
    }
    if (ok && (H5Aexists(m, "kind") <= 0 || get_str(m, "kind", kind, sizeof kind) != BSN_OK ||
               strcmp(kind, "uniform") != 0)) {
        fail(BSN_ERR_IO, "%s is not a uniform-grid checkpoint (kind != \"uniform\")", path);
        ok = 0;
    }
    if (m < 0) fail(BSN_ERR_IO, "group manifest missing");
    if (m >= 0) H5Gclose(m);
    if (!ok) { H5Fclose(f); return -1; }
    return f;
}

int bsn_checkpoint_uniform_info(const char* path, int64_t n[3], int* n_bodies) {
    if (!path || !n || !n_bodies) return fail(BSN_ERR_ARG, "NULL argument");
    hid_t f = open_uniform(path);
    if (f < 0) return BSN_ERR_IO;
    int rc = BSN_OK;
    int64_t nb = 0;
    hid_t gr = H5Gopen2(f, "grid", H5P_DEFAULT), b = H5Gopen2(f, "bhs", H5P_DEFAULT);
    if (gr < 0 || b < 0) { rc = fail(BSN_ERR_IO, "group grid or bhs missing"); goto done; }
    TRY(get_i64(gr, "nx", &n[0]));
    TRY(get_i64(gr, "ny", &n[1]));
    TRY(get_i64(gr, "nz", &n[2]));
    TRY(get_i64(b, "n", &nb));
    *n_bodies = (int)nb;
done:
    if (gr >= 0) H5Gclose(gr);
    if (b >= 0) H5Gclose(b);
    H5Fclose(f);
    return rc;
}

int bsn_load_checkpoint_uniform(bsn_grid* g, double* U_host, const char* path, int64_t* step,
                                double meta[5], double params[9], int max_bodies, int* n_bodies,
                                double* pos, double* vel, double* props) {
    if (!path || !step || !meta || !params || !n_bodies || (!g && !U_host))
        return fail(BSN_ERR_ARG, "bad arguments");
    hid_t f = open_uniform(path);
    if (f < 0) return BSN_ERR_IO;
    int rc = BSN_OK;
    int64_t n[3], fv = 0;
    struct slab_io S = {{-1, -1, -1, -1, -1}, 4, {0, 0, 0, 5}, 0};
    hid_t m = H5Gopen2(f, "manifest", H5P_DEFAULT), gr = H5Gopen2(f, "grid", H5P_DEFAULT);
    if (m < 0 || gr < 0) { rc = fail(BSN_ERR_IO, "group manifest or grid missing"); goto done; }
    TRY(get_f64(m, "t", &meta[0]));
    TRY(get_i64(m, "step", step));
    TRY(get_f64(m, "dt_last", &meta[1]));
    TRY(get_f64(m, "t_snap", &meta[2]));
    TRY(get_f64(m, "t_bh2_sink_on", &meta[3]));
    TRY(get_i64(m, "format_version", &fv));
    meta[4] = (double)fv;
    TRY(get_i64(gr, "nx", &n[0]));
    TRY(get_i64(gr, "ny", &n[1]));
    TRY(get_i64(gr, "nz", &n[2]));
    for (int a = 0; a < 9; ++a) TRY(get_f64(gr, GRID_F64[a], &params[a]));
    S.dims[0] = (hsize_t)(n[2] + 2 * NG);
    S.dims[1] = (hsize_t)(n[1] + 2 * NG);
    S.dims[2] = (hsize_t)(n[0] + 2 * NG);
    if (g) {
        int64_t gn[3];
        if (bsn_grid_dims(g, gn) != BSN_OK) { rc = stream_error(BSN_ERR_ARG, "grid"); goto done; }
        if (gn[0] != n[0] || gn[1] != n[1] || gn[2] != n[2]) {
            rc = fail(BSN_ERR_ARG, "checkpoint is %lldx%lldx%lld, grid is %lldx%lldx%lld",
                      (long long)n[0], (long long)n[1], (long long)n[2], (long long)gn[0],
                      (long long)gn[1], (long long)gn[2]);
            goto done;
        }
        S.ds[0] = open_checked(gr, "U", 4, S.dims);
        if (S.ds[0] < 0) { rc = BSN_ERR_IO; goto done; }
        int e = bsn_grid_import_slabs(g, BSN_LAYOUT_REF, 0, slab_cb, &S);
        if (e != BSN_OK) { rc = stream_error(e, "streaming the state"); goto done; }
    } else {
        TRY(read_all(gr, "U", 4, S.dims, U_host));
    }
    TRY(read_bodies(f, max_bodies, n_bodies, pos, vel, props));
done:
    if (S.ds[0] >= 0) H5Dclose(S.ds[0]);
    if (gr >= 0) H5Gclose(gr);
    if (m >= 0) H5Gclose(m);
    H5Fclose(f);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* two-level FMR checkpoints (src/checkpoint.jl:1-172)                       */

static const char* BOX_ATTR[6] = {"ci_lo", "ci_hi", "cj_lo", "cj_hi", "ck_lo", "ck_hi"};
static const char* LEVEL[2] = {"coarse", "fine"};
static const char* N_ATTR[3] = {"nx", "ny", "nz"};
static const char* D_ATTR[3] = {"dx", "dy", "dz"};

/* One level: /grid/<name> with nx, ny, nz, dx, dy, dz and the dataset U
 * (Julia's (NQ, nx+2NG, ny+2NG, nz+2NG), so (nz+6, ny+6, nx+6, 5) here),
 * streamed from the grid g or written from U_host. */
static int write_level(hid_t gr, int l, bsn_grid* g, const double* U_host, const int64_t n[3],
                       const double d[3]) {
    int rc = BSN_OK;
    struct slab_io S = {{-1, -1, -1, -1, -1}, 4,
                        {(hsize_t)(n[2] + 2 * NG), (hsize_t)(n[1] + 2 * NG), (hsize_t)(n[0] + 2 * NG), 5}, 1};
    hid_t sub = H5Gcreate2(gr, LEVEL[l], H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (sub < 0) return fail(BSN_ERR_IO, "cannot create group grid/%s", LEVEL[l]);
    for (int a = 0; a < 3; ++a) TRY(put_i64(sub, N_ATTR[a], n[a]));
    for (int a = 0; a < 3; ++a) TRY(put_f64(sub, D_ATTR[a], d[a]));
    if (g) {
        S.ds[0] = make_dataset(sub, "U", 4, S.dims);
        if (S.ds[0] < 0) { rc = fail(BSN_ERR_IO, "cannot create dataset grid/%s/U", LEVEL[l]); goto done; }
        int e = bsn_grid_export_slabs(g, BSN_LAYOUT_REF, 0, slab_cb, &S);
        if (e != BSN_OK) { rc = stream_error(e, "streaming a level"); goto done; }
    } else {
        TRY(write_all(sub, "U", 4, S.dims, U_host));
    }
done:
    if (S.ds[0] >= 0) H5Dclose(S.ds[0]);
    H5Gclose(sub);
    return rc;
}

int bsn_save_checkpoint_fmr(bsn_grid* coarse, bsn_grid* fine, const double* Uc_host,
                            const double* Uf_host, const int64_t n[6], const double d[6],
                            const char* path, int64_t step, const double meta[2],
                            const int64_t box[6], int ratio, const double params[3], const char* bc,
                            int n_bodies, const double* pos, const double* vel,
                            const double* props) {
    if (!path || !n || !d || !meta || !box || !params || !bc || ratio < 1 || n_bodies < 0 ||
        (n_bodies > 0 && (!pos || !vel || !props)))
        return fail(BSN_ERR_ARG, "bad arguments");
    if (!((coarse && fine) || (Uc_host && Uf_host)))
        return fail(BSN_ERR_ARG, "no state to save (two grids or two host arrays)");
    if (strcmp(bc, "outflow") != 0 && strcmp(bc, "periodic") != 0)
        return fail(BSN_ERR_ARG, "bc must be \"outflow\" or \"periodic\"");
    for (int a = 0; a < 3; ++a) {
        if (n[a] < 1 || box[2 * a] < 1 || box[2 * a + 1] > n[a] || box[2 * a + 1] < box[2 * a] ||
            n[3 + a] != (box[2 * a + 1] - box[2 * a] + 1) * ratio)
            return fail(BSN_ERR_ARG, "the fine level does not cover the box times the ratio (axis %d)", a);
    }
    if (coarse && fine) {
        int64_t gc[3], gf[3];
        if (bsn_grid_dims(coarse, gc) != BSN_OK || bsn_grid_dims(fine, gf) != BSN_OK)
            return stream_error(BSN_ERR_ARG, "grid");
        for (int a = 0; a < 3; ++a)
            if (gc[a] != n[a] || gf[a] != n[3 + a])
                return fail(BSN_ERR_ARG, "the grids do not have the stated shapes");
    }
    quiet();
    int rc = BSN_OK;
    hid_t m = -1, gr = -1;
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail(BSN_ERR_IO, "cannot create %s", path);
    m = H5Gcreate2(f, "manifest", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (m < 0) { rc = fail(BSN_ERR_IO, "cannot create group manifest"); goto done; }
    TRY(put_i64(m, "format_version", CHECKPOINT_FORMAT_VERSION));
    TRY(put_f64(m, "t", meta[0]));
    TRY(put_i64(m, "step", step));
    TRY(put_f64(m, "dt_last", meta[1]));
    gr = H5Gcreate2(f, "grid", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (gr < 0) { rc = fail(BSN_ERR_IO, "cannot create group grid"); goto done; }
    TRY(put_i64(gr, "ratio", ratio));
    TRY(put_f64(gr, "gamma", params[0]));
    TRY(put_str(gr, "bc", bc));
    TRY(put_f64(gr, "rho_floor", params[1]));
    TRY(put_f64(gr, "P_floor", params[2]));
    for (int a = 0; a < 6; ++a) TRY(put_i64(gr, BOX_ATTR[a], box[a]));
    TRY(write_level(gr, 0, coarse && fine ? coarse : NULL, Uc_host, n, d));
    TRY(write_level(gr, 1, coarse && fine ? fine : NULL, Uf_host, n + 3, d + 3));
    TRY(write_bodies(f, n_bodies, pos, vel, props));
done:
    if (gr >= 0) H5Gclose(gr);
    if (m >= 0) H5Gclose(m);
    if (H5Fclose(f) < 0 && rc == BSN_OK) rc = fail(BSN_ERR_IO, "cannot close %s", path);
    return rc;
}

/* Opens a two-level checkpoint after checking format_version (as
 * load_checkpoint, src/checkpoint.jl:128-130) and that both levels exist. */
static hid_t open_fmr(const char* path) {
    quiet();
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) { fail(BSN_ERR_IO, "cannot open %s", path); return -1; }
    hid_t m = H5Gopen2(f, "manifest", H5P_DEFAULT);
    int64_t fv = -1;
    int ok = m >= 0 && get_i64(m, "format_version", &fv) == BSN_OK;
    if (m >= 0) H5Gclose(m);
    if (ok && fv != CHECKPOINT_FORMAT_VERSION) {
        H5Fclose(f);
        fail(BSN_ERR_IO, "checkpoint format_version=%lld, expected %d", (long long)fv,
             CHECKPOINT_FORMAT_VERSION);
        return -1;
    }
    ok = ok && H5Lexists(f, "grid", H5P_DEFAULT) > 0 && H5Lexists(f, "grid/coarse", H5P_DEFAULT) > 0 &&
         H5Lexists(f, "grid/fine", H5P_DEFAULT) > 0;
    if (!ok) {
        H5Fclose(f);
        fail(BSN_ERR_IO, "%s is not a two-level checkpoint", path);
        return -1;
    }
    return f;
}

int bsn_checkpoint_fmr_info(const char* path, int64_t n[6], double d[6], int64_t box[6],
                            int* ratio, double params[3], char* bc, int bc_len, int* n_bodies) {
    if (!path || !n || !d || !box || !ratio || !params || !bc || bc_len < 1 || !n_bodies)
        return fail(BSN_ERR_ARG, "NULL argument");
    hid_t f = open_fmr(path);
    if (f < 0) return BSN_ERR_IO;
    int rc = BSN_OK;
    int64_t r = 0, nb = 0;
    hid_t gr = H5Gopen2(f, "grid", H5P_DEFAULT), b = H5Gopen2(f, "bhs", H5P_DEFAULT);
    hid_t lv[2] = {-1, -1};
    if (gr < 0 || b < 0) { rc = fail(BSN_ERR_IO, "group grid or bhs missing"); goto done; }
    TRY(get_i64(gr, "ratio", &r));
    *ratio = (int)r;
    TRY(get_f64(gr, "gamma", &params[0]));
    TRY(get_str(gr, "bc", bc, (size_t)bc_len));
    TRY(get_f64(gr, "rho_floor", &params[1]));
    TRY(get_f64(gr, "P_floor", &params[2]));
    for (int a = 0; a < 6; ++a) TRY(get_i64(gr, BOX_ATTR[a], &box[a]));
    for (int l = 0; l < 2; ++l) {
        lv[l] = H5Gopen2(gr, LEVEL[l], H5P_DEFAULT);
        if (lv[l] < 0) { rc = fail(BSN_ERR_IO, "group grid/%s missing", LEVEL[l]); goto done; }
        for (int a = 0; a < 3; ++a) TRY(get_i64(lv[l], N_ATTR[a], &n[3 * l + a]));
        for (int a = 0; a < 3; ++a) TRY(get_f64(lv[l], D_ATTR[a], &d[3 * l + a]));
    }
    TRY(get_i64(b, "n", &nb));
    *n_bodies = (int)nb;
done:
    for (int l = 0; l < 2; ++l) if (lv[l] >= 0) H5Gclose(lv[l]);
    if (gr >= 0) H5Gclose(gr);
    if (b >= 0) H5Gclose(b);
    H5Fclose(f);
    return rc;
}

int bsn_load_checkpoint_fmr(bsn_grid* coarse, bsn_grid* fine, double* Uc_host, double* Uf_host,
                            const char* path, int64_t* step, double meta[3], int max_bodies,
                            int* n_bodies, double* pos, double* vel, double* props) {
    if (!path || !step || !meta || !n_bodies || !((coarse && fine) || (Uc_host && Uf_host)))
        return fail(BSN_ERR_ARG, "bad arguments");
    hid_t f = open_fmr(path);
    if (f < 0) return BSN_ERR_IO;
    int rc = BSN_OK;
    int64_t fv = 0;
    struct slab_io S = {{-1, -1, -1, -1, -1}, 4, {0, 0, 0, 5}, 0};
    hid_t lv = -1;
    hid_t m = H5Gopen2(f, "manifest", H5P_DEFAULT), gr = H5Gopen2(f, "grid", H5P_DEFAULT);
    if (m < 0 || gr < 0) { rc = fail(BSN_ERR_IO, "group manifest or grid missing"); goto done; }
    TRY(get_f64(m, "t", &meta[0]));
    TRY(get_i64(m, "step", step));
    TRY(get_f64(m, "dt_last", &meta[1]));
    TRY(get_i64(m, "format_version", &fv));
    meta[2] = (double)fv;
    for (int l = 0; l < 2; ++l) {
        int64_t n[3];
        lv = H5Gopen2(gr, LEVEL[l], H5P_DEFAULT);
        if (lv < 0) { rc = fail(BSN_ERR_IO, "group grid/%s missing", LEVEL[l]); goto done; }
        for (int a = 0; a < 3; ++a) TRY(get_i64(lv, N_ATTR[a], &n[a]));
        S.dims[0] = (hsize_t)(n[2] + 2 * NG);
        S.dims[1] = (hsize_t)(n[1] + 2 * NG);
        S.dims[2] = (hsize_t)(n[0] + 2 * NG);
        bsn_grid* g = l == 0 ? coarse : fine;
        if (coarse && fine) {
            int64_t gn[3];
            if (bsn_grid_dims(g, gn) != BSN_OK) { rc = stream_error(BSN_ERR_ARG, "grid"); goto done; }
            if (gn[0] != n[0] || gn[1] != n[1] || gn[2] != n[2]) {
                rc = fail(BSN_ERR_ARG, "checkpoint %s level is %lldx%lldx%lld, grid is %lldx%lldx%lld",
                          LEVEL[l], (long long)n[0], (long long)n[1], (long long)n[2],
                          (long long)gn[0], (long long)gn[1], (long long)gn[2]);
                goto done;
            }
            S.ds[0] = open_checked(lv, "U", 4, S.dims);
            if (S.ds[0] < 0) { rc = BSN_ERR_IO; goto done; }
            int e = bsn_grid_import_slabs(g, BSN_LAYOUT_REF, 0, slab_cb, &S);
            H5Dclose(S.ds[0]);
            S.ds[0] = -1;
            if (e != BSN_OK) { rc = stream_error(e, "streaming a level"); goto done; }
        } else {
            TRY(read_all(lv, "U", 4, S.dims, l == 0 ? Uc_host : Uf_host));
        }
        H5Gclose(lv);
        lv = -1;
    }
    TRY(read_bodies(f, max_bodies, n_bodies, pos, vel, props));
done:
    if (S.ds[0] >= 0) H5Dclose(S.ds[0]);
    if (lv >= 0) H5Gclose(lv);
    if (gr >= 0) H5Gclose(gr);
    if (m >= 0) H5Gclose(m);
    H5Fclose(f);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* body trajectories (src/io.jl:86-161)                                      */
/* "time" (N), and per body bh<i>/mass (N), bh<i>/rsink (N), bh<i>/pos and
 * bh<i>/vel (Julia's (N, 3), so (3, N) here), all extendable along N in
 * chunks of 128 records. */

static hid_t make_extendable(hid_t loc, const char* name, int rank) {
    hsize_t dims[2] = {0, 0}, maxd[2], chunk[2];
    if (rank == 1) { maxd[0] = H5S_UNLIMITED; chunk[0] = 128; }
    else { dims[0] = 3; maxd[0] = 3; maxd[1] = H5S_UNLIMITED; chunk[0] = 3; chunk[1] = 128; }
    hid_t sp = H5Screate_simple(rank, dims, maxd);
    hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
    H5Pset_chunk(pl, rank, chunk);
    hid_t d = H5Dcreate2(loc, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, pl, H5P_DEFAULT);
    H5Pclose(pl);
    H5Sclose(sp);
    return d;
}

int bsn_init_trajectory_file(const char* filename, int n_bh) {
    if (!filename || n_bh < 0) return fail(BSN_ERR_ARG, "bad arguments");
    quiet();
    static const char* DS[4] = {"mass", "rsink", "pos", "vel"};
    int rc = BSN_OK;
    hid_t f = H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail(BSN_ERR_IO, "cannot create %s", filename);
    hid_t d = make_extendable(f, "time", 1);
    if (d < 0) { rc = fail(BSN_ERR_IO, "cannot create dataset time"); goto done; }
    H5Dclose(d);
    for (int i = 1; i <= n_bh && rc == BSN_OK; ++i) {
        char name[32];
        snprintf(name, sizeof name, "bh%d", i);
        hid_t g = H5Gcreate2(f, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (g < 0) { rc = fail(BSN_ERR_IO, "cannot create group %s", name); break; }
        for (int k = 0; k < 4 && rc == BSN_OK; ++k) {
            d = make_extendable(g, DS[k], k < 2 ? 1 : 2);
            if (d < 0) rc = fail(BSN_ERR_IO, "cannot create dataset %s/%s", name, DS[k]);
            else H5Dclose(d);
        }
        H5Gclose(g);
    }
done:
    if (H5Fclose(f) < 0 && rc == BSN_OK) rc = fail(BSN_ERR_IO, "cannot close %s", filename);
    return rc;
}

/* Grows dataset d by one record (index n) and writes v (1 or 3 values). */
static int append_record(hid_t d, int rank, hsize_t n, const double* v) {
    hsize_t ext[2], start[2], cnt[2];
    if (rank == 1) { ext[0] = n + 1; start[0] = n; cnt[0] = 1; }
    else { ext[0] = 3; ext[1] = n + 1; start[0] = 0; start[1] = n; cnt[0] = 3; cnt[1] = 1; }
    if (H5Dset_extent(d, ext) < 0) return 1;
    hid_t fs = H5Dget_space(d);
    hsize_t m = rank == 1 ? 1 : 3;
    hid_t ms = H5Screate_simple(1, &m, NULL);
    herr_t e = H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, NULL, cnt, NULL);
    if (e >= 0) e = H5Dwrite(d, H5T_NATIVE_DOUBLE, ms, fs, H5P_DEFAULT, v);
    H5Sclose(ms);
    H5Sclose(fs);
    return e < 0;
}

int bsn_append_trajectory(const char* filename, double t, int n_bh, const double* mass,
                          const double* rsink, const double* pos, const double* vel) {
    if (!filename || n_bh < 0 || (n_bh > 0 && (!mass || !rsink || !pos || !vel)))
        return fail(BSN_ERR_ARG, "bad arguments");
    quiet();
    int rc = BSN_OK;
    hid_t f = H5Fopen(filename, H5F_ACC_RDWR, H5P_DEFAULT);
    if (f < 0) return fail(BSN_ERR_IO, "cannot open %s", filename);
    hid_t d = H5Dopen2(f, "time", H5P_DEFAULT);
    if (d < 0) { rc = fail(BSN_ERR_IO, "dataset time missing"); goto done; }
    hid_t sp = H5Dget_space(d);
    hsize_t n = 0;
    H5Sget_simple_extent_dims(sp, &n, NULL);
    H5Sclose(sp);
    if (append_record(d, 1, n, &t)) rc = fail(BSN_ERR_IO, "cannot extend time");
    H5Dclose(d);
    for (int i = 0; i < n_bh && rc == BSN_OK; ++i) {
        char name[32];
        snprintf(name, sizeof name, "bh%d", i + 1);
        hid_t g = H5Gopen2(f, name, H5P_DEFAULT);
        if (g < 0) { rc = fail(BSN_ERR_IO, "group %s missing", name); break; }
        const double* src[4] = {mass + i, rsink + i, pos + 3 * i, vel + 3 * i};
        static const char* DS[4] = {"mass", "rsink", "pos", "vel"};
        for (int k = 0; k < 4 && rc == BSN_OK; ++k) {
            hid_t dk = H5Dopen2(g, DS[k], H5P_DEFAULT);
            if (dk < 0 || append_record(dk, k < 2 ? 1 : 2, n, src[k]))
                rc = fail(BSN_ERR_IO, "cannot extend %s/%s", name, DS[k]);
            if (dk >= 0) H5Dclose(dk);
        }
        H5Gclose(g);
    }
done:
    if (H5Fclose(f) < 0 && rc == BSN_OK) rc = fail(BSN_ERR_IO, "cannot close %s", filename);
    return rc;
}

int bsn_trajectory_info(const char* filename, int* n_bh, int64_t* n_records) {
    if (!filename || !n_bh || !n_records) return fail(BSN_ERR_ARG, "NULL argument");
    quiet();
    hid_t f = H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail(BSN_ERR_IO, "cannot open %s", filename);
    int rc = BSN_OK;
    hid_t d = H5Dopen2(f, "time", H5P_DEFAULT);
    if (d < 0) {
        rc = fail(BSN_ERR_IO, "dataset time missing");
    } else {
        hid_t sp = H5Dget_space(d);
        hsize_t n = 0;
        H5Sget_simple_extent_dims(sp, &n, NULL);
        H5Sclose(sp);
        H5Dclose(d);
        *n_records = (int64_t)n;
        int i = 1;
        for (;; ++i) {                    /* bh1, bh2, ... as read_trajectory */
            char name[32];
            snprintf(name, sizeof name, "bh%d", i);
            if (H5Lexists(f, name, H5P_DEFAULT) <= 0) break;
        }
        *n_bh = i - 1;
    }
    H5Fclose(f);
    return rc;
}

int bsn_read_trajectory(const char* filename, double* t, double* mass, double* rsink,
                        double* pos, double* vel) {
    int nb = 0;
    int64_t n = 0;
    int rc = bsn_trajectory_info(filename, &nb, &n);
    if (rc != BSN_OK) return rc;
    if (n > 0 && (!t || (nb > 0 && (!mass || !rsink || !pos || !vel))))
        return fail(BSN_ERR_ARG, "NULL output");
    if (n == 0) return BSN_OK;
    quiet();
    double* tmp = malloc(sizeof(double) * 3 * (size_t)n);
    if (!tmp) return fail(BSN_ERR_ARG, "out of host memory");
    hid_t f = H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) { free(tmp); return fail(BSN_ERR_IO, "cannot open %s", filename); }
    const hsize_t d1[1] = {(hsize_t)n}, d2[2] = {3, (hsize_t)n};
    TRY(read_all(f, "time", 1, d1, t));
    for (int i = 0; i < nb; ++i) {
        char name[32];
        snprintf(name, sizeof name, "bh%d", i + 1);
        hid_t g = H5Gopen2(f, name, H5P_DEFAULT);
        if (g < 0) { rc = fail(BSN_ERR_IO, "group %s missing", name); goto done; }
        rc = read_all(g, "mass", 1, d1, mass + (size_t)i * n);
        if (rc == BSN_OK) rc = read_all(g, "rsink", 1, d1, rsink + (size_t)i * n);
        for (int k = 0; k < 2 && rc == BSN_OK; ++k) {
            rc = read_all(g, k ? "vel" : "pos", 2, d2, tmp);
            double* out = (k ? vel : pos) + (size_t)i * n * 3;
            for (int64_t r = 0; r < n && rc == BSN_OK; ++r)
                for (int a = 0; a < 3; ++a) out[3 * r + a] = tmp[(size_t)a * n + r];
        }
        H5Gclose(g);
        if (rc != BSN_OK) goto done;
    }
done:
    free(tmp);
    H5Fclose(f);
    return rc;
}
