// mdx_io.cpp -- checkpoint files of the flow state (libmobydiff_io.so, host code over the HDF5 C
// library).  The file format is the reference's (field_hdf5.c:149-204): four ghosted FP64 datasets
// stored [i][j][k] with k fastest, twelve scalar attributes on the root group.  The fields arrive
// in the Fortran layout (i fastest), so each one is transposed through a staging buffer; the
// transpose walks 32x32 (i, k) tiles of one j row, so both the strided reads and the strided writes
// stay within a few hundred cache lines.
#include "../../include/mobydiff_io.h"

#include <hdf5.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const std::string &msg) {
    g_err = msg;
    return 1;
}

// HDF5 identifiers closed on scope exit
struct Id {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Id() {
        if (id >= 0 && close) close(id);
    }
    Id(const Id &) = delete;
    Id &operator=(const Id &) = delete;
    bool ok() const { return id >= 0; }
    // hand the identifier to an explicit close whose status matters (the file's)
    hid_t release() {
        hid_t i = id;
        id = -1;
        return i;
    }
};

// The library's automatic error stack printing is switched off while a call runs: failures are
// reported through the return code and mdx_io_last_error instead.
struct QuietErrors {
    H5E_auto2_t fn = nullptr;
    void *data = nullptr;
    QuietErrors() {
        H5Eget_auto2(H5E_DEFAULT, &fn, &data);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~QuietErrors() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

constexpr size_t TILE = 32;

// Fortran q[i + ni*(j + nj*k)]  ->  file order b[(i*nj + j)*nk + k]
void to_file_order(const double *q, double *b, size_t ni, size_t nj, size_t nk) {
    for (size_t j = 0; j < nj; ++j)
        for (size_t i0 = 0; i0 < ni; i0 += TILE)
            for (size_t k0 = 0; k0 < nk; k0 += TILE) {
                const size_t i1 = i0 + TILE < ni ? i0 + TILE : ni, k1 = k0 + TILE < nk ? k0 + TILE : nk;
                for (size_t i = i0; i < i1; ++i)
                    for (size_t k = k0; k < k1; ++k) b[(i * nj + j) * nk + k] = q[i + ni * (j + nj * k)];
            }
}

void from_file_order(const double *b, double *q, size_t ni, size_t nj, size_t nk) {
    for (size_t j = 0; j < nj; ++j)
        for (size_t k0 = 0; k0 < nk; k0 += TILE)
            for (size_t i0 = 0; i0 < ni; i0 += TILE) {
                const size_t i1 = i0 + TILE < ni ? i0 + TILE : ni, k1 = k0 + TILE < nk ? k0 + TILE : nk;
                for (size_t k = k0; k < k1; ++k)
                    for (size_t i = i0; i < i1; ++i) q[i + ni * (j + nj * k)] = b[(i * nj + j) * nk + k];
            }
}

int put_attr(hid_t file, const char *name, hid_t type, const void *value) {
    Id space(H5Screate(H5S_SCALAR), H5Sclose);
    if (!space.ok()) return fail(std::string("attribute ") + name + ": dataspace");
    Id attr(H5Acreate2(file, name, type, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!attr.ok()) return fail(std::string("attribute ") + name + ": create");
    if (H5Awrite(attr.id, type, value) < 0) return fail(std::string("attribute ") + name + ": write");
    return 0;
}

int get_attr(hid_t file, const char *name, hid_t type, void *value) {
    Id attr(H5Aopen(file, name, H5P_DEFAULT), H5Aclose);
    if (!attr.ok()) return fail(std::string("attribute ") + name + ": missing");
    if (H5Aread(attr.id, type, value) < 0) return fail(std::string("attribute ") + name + ": read");
    return 0;
}

int put_field(hid_t file, const char *name, const double *q, size_t ni, size_t nj, size_t nk,
              std::vector<double> &buf) {
    if (!q) return fail(std::string("dataset ") + name + ": null array");
    to_file_order(q, buf.data(), ni, nj, nk);
    const hsize_t dims[3] = {ni, nj, nk};
    Id space(H5Screate_simple(3, dims, nullptr), H5Sclose);
    if (!space.ok()) return fail(std::string("dataset ") + name + ": dataspace");
    Id dset(H5Dcreate2(file, name, H5T_NATIVE_DOUBLE, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!dset.ok()) return fail(std::string("dataset ") + name + ": create");
    if (H5Dwrite(dset.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0)
        return fail(std::string("dataset ") + name + ": write");
    return 0;
}

int get_field(hid_t file, const char *name, double *q, size_t ni, size_t nj, size_t nk, std::vector<double> &buf) {
    if (!q) return fail(std::string("dataset ") + name + ": null array");
    Id dset(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose);
    if (!dset.ok()) return fail(std::string("dataset ") + name + ": missing");
    Id space(H5Dget_space(dset.id), H5Sclose);
    if (!space.ok() || H5Sget_simple_extent_ndims(space.id) != 3)
        return fail(std::string("dataset ") + name + ": not three-dimensional");
    hsize_t dims[3] = {0, 0, 0};
    H5Sget_simple_extent_dims(space.id, dims, nullptr);
    if (dims[0] != ni || dims[1] != nj || dims[2] != nk) {
        char msg[160];
        std::snprintf(msg, sizeof msg, "dataset %s: shape [%llu][%llu][%llu], expected [%zu][%zu][%zu]", name,
                      (unsigned long long)dims[0], (unsigned long long)dims[1], (unsigned long long)dims[2], ni,
                      nj, nk);
        return fail(msg);
    }
    if (H5Dread(dset.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0)
        return fail(std::string("dataset ") + name + ": read");
    from_file_order(buf.data(), q, ni, nj, nk);
    return 0;
}

}  // namespace

extern "C" const char *mdx_io_last_error(void) { return g_err.c_str(); }

extern "C" int mdx_h5_write_field(const char *path, int nx, int ny, int nz, double lx, double ly, double lz,
                                  double dx, double dy, double dz, double re, double dt, double t_current,
                                  const double *un, const double *vn, const double *wn, const double *pn) {
    if (!path) return fail("mdx_h5_write_field: null path");
    if (nx < 1 || ny < 1 || nz < 1) return fail("mdx_h5_write_field: grid sizes must be positive");
    QuietErrors quiet;
    const size_t ni = (size_t)nx + 2, nj = (size_t)ny + 2, nk = (size_t)nz + 2;
    Id file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!file.ok()) return fail(std::string("cannot create ") + path);
    const struct {
        const char *name;
        double v;
    } scal[] = {{"lx", lx}, {"ly", ly}, {"lz", lz}, {"dx", dx}, {"dy", dy},
                {"dz", dz}, {"re", re}, {"dt", dt}, {"t_current", t_current}};
    int err = put_attr(file.id, "nx", H5T_NATIVE_INT, &nx) || put_attr(file.id, "ny", H5T_NATIVE_INT, &ny) ||
              put_attr(file.id, "nz", H5T_NATIVE_INT, &nz);
    for (const auto &s : scal)
        if (!err) err = put_attr(file.id, s.name, H5T_NATIVE_DOUBLE, &s.v);
    if (err) return 1;
    std::vector<double> buf(ni * nj * nk);
    if (put_field(file.id, "un", un, ni, nj, nk, buf) || put_field(file.id, "vn", vn, ni, nj, nk, buf) ||
        put_field(file.id, "wn", wn, ni, nj, nk, buf) || put_field(file.id, "pn", pn, ni, nj, nk, buf))
        return 1;
    if (H5Fclose(file.release()) < 0) return fail(std::string("cannot close ") + path);
    return 0;
}

extern "C" int mdx_h5_read_field(const char *path, int nx, int ny, int nz, double *un, double *vn, double *wn,
                                 double *pn) {
    if (!path) return fail("mdx_h5_read_field: null path");
    if (nx < 1 || ny < 1 || nz < 1) return fail("mdx_h5_read_field: grid sizes must be positive");
    QuietErrors quiet;
    const size_t ni = (size_t)nx + 2, nj = (size_t)ny + 2, nk = (size_t)nz + 2;
    Id file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!file.ok()) return fail(std::string("cannot open ") + path);
    std::vector<double> buf(ni * nj * nk);
    if (get_field(file.id, "un", un, ni, nj, nk, buf) || get_field(file.id, "vn", vn, ni, nj, nk, buf) ||
        get_field(file.id, "wn", wn, ni, nj, nk, buf) || get_field(file.id, "pn", pn, ni, nj, nk, buf))
        return 1;
    if (H5Fclose(file.release()) < 0) return fail(std::string("cannot close ") + path);
    return 0;
}

extern "C" int mdx_h5_read_field_info(const char *path, int32_t *nx, int32_t *ny, int32_t *nz, double *dt,
                                      double *t_current) {
    if (!path) return fail("mdx_h5_read_field_info: null path");
    QuietErrors quiet;
    Id file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!file.ok()) return fail(std::string("cannot open ") + path);
    int n[3] = {0, 0, 0};
    double t[2] = {0.0, 0.0};
    if (get_attr(file.id, "nx", H5T_NATIVE_INT, &n[0]) || get_attr(file.id, "ny", H5T_NATIVE_INT, &n[1]) ||
        get_attr(file.id, "nz", H5T_NATIVE_INT, &n[2]) || get_attr(file.id, "dt", H5T_NATIVE_DOUBLE, &t[0]) ||
        get_attr(file.id, "t_current", H5T_NATIVE_DOUBLE, &t[1]))
        return 1;
    if (nx) *nx = n[0];
    if (ny) *ny = n[1];
    if (nz) *nz = n[2];
    if (dt) *dt = t[0];
    if (t_current) *t_current = t[1];
    return 0;
}
