// QuickPIC's HDF5 file format (hdf5io_class.f03), written with the serial
// HDF5 C library.  One process writes a whole file: the pipelined writers
// of the reference (each stage opening the file in turn) reduce to this at
// one stage.
//
// The reference writes through HDF5's Fortran API, so its layouts are
// Fortran's: a dataset of Fortran dimensions (d1, d2, d3) appears with C
// dimensions (d3, d2, d1) and the same bytes, and its strings are the
// Fortran character type (space padded).  Attributes given as the
// character(len=100) members of hdf5file (NAME, TYPE, TIME UNITS, UNITS,
// LONG_NAME) keep their padding to 100 characters, as len() of those
// members is 100; attributes given as trimmed literals do not.
#include <hdf5.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/qpio.h"

namespace {

thread_local char g_err[512] = "";

int fail(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return -1;
}

constexpr int PADDED = 100;   // character(len=100) members of hdf5file

std::string str_or(const char *s, const char *def) { return s ? std::string(s) : std::string(def); }

// add_h5_atribute_str (hdf5io_class.f03:147-172): a one-element dataspace
// of a fixed-length Fortran string; pad > 0 pads the value to that length
bool attr_str(hid_t obj, const char *name, std::string v, int pad)
{
    if (pad > 0 && (int)v.size() < pad) v.append(pad - v.size(), ' ');
    hsize_t dims[1] = {1};
    hid_t sp = H5Screate_simple(1, dims, nullptr);
    hid_t ty = H5Tcopy(H5T_FORTRAN_S1);
    H5Tset_size(ty, v.empty() ? 1 : v.size());
    hid_t a = H5Acreate2(obj, name, ty, sp, H5P_DEFAULT, H5P_DEFAULT);
    bool ok = a >= 0 && H5Awrite(a, ty, v.empty() ? " " : v.data()) >= 0;
    if (a >= 0) H5Aclose(a);
    H5Tclose(ty);
    H5Sclose(sp);
    return ok;
}

// add_h5_atribute_single / _int (:197-284): a one-element dataspace
bool attr_real(hid_t obj, const char *name, double v)
{
    hsize_t dims[1] = {1};
    hid_t sp = H5Screate_simple(1, dims, nullptr);
    hid_t a = H5Acreate2(obj, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT);
    bool ok = a >= 0 && H5Awrite(a, H5T_NATIVE_DOUBLE, &v) >= 0;
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    return ok;
}

bool attr_int(hid_t obj, const char *name, int v)
{
    hsize_t dims[1] = {1};
    hid_t sp = H5Screate_simple(1, dims, nullptr);
    hid_t a = H5Acreate2(obj, name, H5T_NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT);
    bool ok = a >= 0 && H5Awrite(a, H5T_NATIVE_INT, &v) >= 0;
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    return ok;
}

std::string ty_of(const qpio_file *f) { return str_or(f->ty, "grid"); }

// wrattr_file (:299-347)
bool write_file_attrs(const qpio_file *f, hid_t file)
{
    hid_t root = H5Gopen2(file, "/", H5P_DEFAULT);
    if (root < 0) return false;
    bool ok = attr_str(root, "NAME", str_or(f->dataname, "Data"), PADDED) &&
              attr_str(root, "TYPE", ty_of(f), PADDED) && attr_real(root, "TIME", f->t) &&
              attr_int(root, "ITER", f->n) && attr_real(root, "DT", f->dt) &&
              attr_str(root, "TIME UNITS", str_or(f->timeunits, "a.u."), PADDED);
    if (ok && ty_of(f) == "grid") {
        static const char *defname[3] = {"x1", "x2", "x3"};
        hid_t ax = H5Gcreate2(root, "AXIS", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        hsize_t dims[1] = {2};
        hid_t sp = H5Screate_simple(1, dims, nullptr);
        for (int i = 0; ok && i < f->rank; ++i) {
            char nm[8];
            snprintf(nm, sizeof nm, "AXIS%d", i + 1);
            hid_t d = H5Dcreate2(ax, nm, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT,
                                 H5P_DEFAULT);
            ok = d >= 0;
            if (ok) {
                ok = attr_str(d, "TYPE", "linear", 0) &&
                     attr_str(d, "UNITS", str_or(f->axisunits[i], "a.u."), 0) &&
                     attr_str(d, "NAME", str_or(f->axisname[i], defname[i]), 0) &&
                     attr_str(d, "LONG_NAME", str_or(f->axislabel[i], defname[i]), 0);
                const double mm[2] = {f->axismin[i], f->axismax[i]};
                ok = ok && H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, mm) >= 0;
                H5Dclose(d);
            }
        }
        H5Sclose(sp);
        H5Gclose(ax);
    }
    H5Gclose(root);
    return ok;
}

// wrattr_dataset (:349-369)
bool write_dset_attrs(const qpio_file *f, hid_t d, const char *unit = nullptr,
                      const char *name = nullptr)
{
    return attr_str(d, "UNITS", unit ? std::string(unit) : str_or(f->units, "a.u."),
                    unit ? 0 : PADDED) &&
           attr_str(d, "LONG_NAME", name ? std::string(name) : str_or(f->label, "Data"),
                    name ? 0 : PADDED);
}

struct File {
    hid_t id = -1;
    ~File() { if (id >= 0) H5Fclose(id); }
};

bool check(const qpio_file *f)
{
    if (!f) return false;
    if (f->rank < 0 || f->rank > 3) return false;
    return true;
}

}  // namespace

extern "C" {

const char *qpio_last_error(void) { return g_err; }

int qpio_path(const qpio_file *f, char *out, size_t cap)
{
    if (!f || !out) return fail("qpio_path: null argument");
    std::string s = str_or(f->filename, "file.h5");
    char st[32];
    snprintf(st, sizeof st, "%08d", f->n);
    s += str_or(f->dataname, "Data") + "_" + st + ".h5";
    if (s.size() + 1 > cap) return fail("qpio_path: %zu characters do not fit", s.size());
    std::memcpy(out, s.c_str(), s.size() + 1);
    return 0;
}

int qpio_write_grid(const qpio_file *f, int rank, const int64_t *gsize, const double *data)
{
    if (!check(f) || !gsize || !data || rank < 2 || rank > 3)
        return fail("qpio_write_grid: bad arguments (rank %d)", rank);
    char path[4096];
    if (qpio_path(f, path, sizeof path)) return -1;
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    File fl;
    fl.id = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (fl.id < 0) return fail("qpio_write_grid: cannot create %s", path);
    if (!write_file_attrs(f, fl.id)) return fail("qpio_write_grid: attributes of %s", path);
    hsize_t dims[3];
    for (int i = 0; i < rank; ++i) {
        if (gsize[i] <= 0) return fail("qpio_write_grid: size %lld", (long long)gsize[i]);
        dims[rank - 1 - i] = (hsize_t)gsize[i];     // Fortran order reversed
    }
    hid_t sp = H5Screate_simple(rank, dims, nullptr);
    hid_t d = H5Dcreate2(fl.id, str_or(f->dataname, "Data").c_str(), H5T_NATIVE_DOUBLE, sp,
                         H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    bool ok = d >= 0 && write_dset_attrs(f, d) &&
              H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
    if (d >= 0) H5Dclose(d);
    H5Sclose(sp);
    return ok ? 0 : fail("qpio_write_grid: writing %s", path);
}

int qpio_write_raw(const qpio_file *f, const double *part, int64_t np, int dspl,
                   const double delta[3], const double origin[3])
{
    if (!check(f) || (np > 0 && !part) || dspl < 1 || !delta || !origin)
        return fail("qpio_write_raw: bad arguments");
    char path[4096];
    if (qpio_path(f, path, sizeof path)) return -1;
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    File fl;
    fl.id = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (fl.id < 0) return fail("qpio_write_raw: cannot create %s", path);
    const int64_t tp = np / dspl;      // tnpp = int(npp/dspl)
    if (!write_file_attrs(f, fl.id)) return fail("qpio_write_raw: attributes of %s", path);
    hid_t root = H5Gopen2(fl.id, "/", H5P_DEFAULT);
    bool ok = attr_int(root, "tp", (int)tp);
    std::vector<double> buf((size_t)tp);
    static const char *names[7] = {"x1", "x2", "x3", "p1", "p2", "p3", "q"};
    static const char *longs[7] = {"x_1", "x_2", "x_3", "p_1", "p_2", "p_3", "q"};
    for (int c = 0; ok && tp > 0 && c < 7; ++c) {
        for (int64_t i = 0; i < tp; ++i) {
            const double v = part[(size_t)(i * dspl) * 7 + c];
            buf[(size_t)i] = c < 3 ? v * delta[c] + origin[c] : v;
        }
        // an extendable chunked dataset created at 1 and extended to tp
        hsize_t one[1] = {1}, maxd[1] = {H5S_UNLIMITED}, n[1] = {(hsize_t)tp};
        hid_t sp = H5Screate_simple(1, one, maxd);
        hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
        H5Pset_chunk(pl, 1, n);
        hid_t d = H5Dcreate2(root, names[c], H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, pl, H5P_DEFAULT);
        ok = d >= 0 && H5Dset_extent(d, n) >= 0 &&
             H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) >= 0;
        // units: c/\omega_p for x and p (:1177-1178, :1232-1233), a.u. for q (:1269)
        ok = ok && write_dset_attrs(f, d, c < 6 ? "c/\\omega_p" : "a.u.", longs[c]);
        if (d >= 0) H5Dclose(d);
        H5Pclose(pl);
        H5Sclose(sp);
    }
    H5Gclose(root);
    return ok ? 0 : fail("qpio_write_raw: writing %s", path);
}

int qpio_write_rst(const qpio_file *f, const double *part, int64_t np)
{
    if (!check(f) || np < 0 || (np > 0 && !part)) return fail("qpio_write_rst: bad arguments");
    char path[4096];
    if (qpio_path(f, path, sizeof path)) return -1;
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    File fl;
    fl.id = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (fl.id < 0) return fail("qpio_write_rst: cannot create %s", path);
    if (!write_file_attrs(f, fl.id)) return fail("qpio_write_rst: attributes of %s", path);
    hid_t root = H5Gopen2(fl.id, "/", H5P_DEFAULT);
    bool ok = attr_int(root, "tp", (int)np);
    if (ok && np > 0) {
        hsize_t dims[2] = {(hsize_t)np, 7};       // Fortran (7, tp)
        hid_t sp = H5Screate_simple(2, dims, nullptr);
        hid_t d = H5Dcreate2(root, str_or(f->dataname, "Data").c_str(), H5T_NATIVE_DOUBLE, sp,
                             H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        ok = d >= 0 && write_dset_attrs(f, d) &&
             H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, part) >= 0;
        if (d >= 0) H5Dclose(d);
        H5Sclose(sp);
    }
    H5Gclose(root);
    return ok ? 0 : fail("qpio_write_rst: writing %s", path);
}

int qpio_read_rst(const qpio_file *f, double *part, int64_t max, int64_t *np)
{
    if (!check(f) || !np || (max > 0 && !part)) return fail("qpio_read_rst: bad arguments");
    char path[4096];
    if (qpio_path(f, path, sizeof path)) return -1;
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    File fl;
    fl.id = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (fl.id < 0) return fail("qpio_read_rst: cannot open %s", path);
    hid_t root = H5Gopen2(fl.id, "/", H5P_DEFAULT);
    hid_t a = H5Aopen(root, "tp", H5P_DEFAULT);
    int tp = -1;
    bool ok = a >= 0 && H5Aread(a, H5T_NATIVE_INT, &tp) >= 0;
    if (a >= 0) H5Aclose(a);
    if (!ok) { H5Gclose(root); return fail("qpio_read_rst: no tp attribute in %s", path); }
    *np = tp;
    if (tp > max) {
        H5Gclose(root);
        return fail("qpio_read_rst: %d records in %s, room for %lld", tp, path, (long long)max);
    }
    if (tp > 0) {
        hid_t d = H5Dopen2(root, str_or(f->dataname, "Data").c_str(), H5P_DEFAULT);
        ok = d >= 0;
        if (ok) {
            hid_t sp = H5Dget_space(d);
            hsize_t dims[2] = {0, 0};
            ok = H5Sget_simple_extent_ndims(sp) == 2 && H5Sget_simple_extent_dims(sp, dims, nullptr) == 2 &&
                 dims[0] == (hsize_t)tp && dims[1] == 7;
            ok = ok && H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, part) >= 0;
            H5Sclose(sp);
            H5Dclose(d);
        }
    }
    H5Gclose(root);
    return ok ? 0 : fail("qpio_read_rst: reading %s", path);
}

}  // extern "C"
