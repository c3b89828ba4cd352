// qpc_io.cpp -- 1-D HDF5 datasets with the driver's element types
// (include/qpc_io.h).  Written against the HDF5 C API directly.
#include "../../include/qpc_io.h"

#include <hdf5.h>

#include <algorithm>
#include <string>

struct qpc_h5 {
    hid_t file = -1;
};

namespace {
hid_t native(int type) {
    switch (type) {
        case QPC_H5_F64: return H5T_NATIVE_DOUBLE;
        case QPC_H5_U64: return H5T_NATIVE_ULONG;
        case QPC_H5_I64: return H5T_NATIVE_LONG;
        case QPC_H5_I32: return H5T_NATIVE_INT;
        default: return -1;
    }
}
// classify a stored type as one of the driver's
int classify(hid_t t) {
    const H5T_class_t cls = H5Tget_class(t);
    const size_t sz = H5Tget_size(t);
    if (cls == H5T_FLOAT && sz == 8) return QPC_H5_F64;
    if (cls == H5T_INTEGER) {
        const H5T_sign_t sg = H5Tget_sign(t);
        if (sz == 8) return sg == H5T_SGN_NONE ? QPC_H5_U64 : QPC_H5_I64;
        if (sz == 4 && sg != H5T_SGN_NONE) return QPC_H5_I32;
    }
    return -1;
}
struct Silence {  // no HDF5 error stack on stderr for the expected misses
    H5E_auto2_t fn;
    void* data;
    Silence() {
        H5Eget_auto2(H5E_DEFAULT, &fn, &data);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~Silence() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};
herr_t list_one(hid_t, const char* name, const H5L_info_t*, void* op) {
    std::string* s = static_cast<std::string*>(op);
    *s += name;
    *s += '\n';
    return 0;
}
}  // namespace

extern "C" {

int qpc_h5_open(const char* path, char mode, qpc_h5** out) {
    if (!path || !out) return -1;
    Silence quiet;
    hid_t f = -1;
    switch (mode) {
        case 'w': f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT); break;
        case 'r': f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT); break;
        case 'u': f = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT); break;
        default: return -1;
    }
    if (f < 0) return -2;
    *out = new qpc_h5{f};
    return 0;
}

int qpc_h5_close(qpc_h5* f) {
    if (!f) return -1;
    const herr_t st = H5Fclose(f->file);
    delete f;
    return st < 0 ? -2 : 0;
}

int qpc_h5_save(qpc_h5* f, const char* name, int type, const void* data, int64_t n) {
    const hid_t t = native(type);
    if (!f || !name || t < 0 || n < 0 || (n > 0 && !data)) return -1;
    Silence quiet;
    hsize_t dims[1] = {(hsize_t)n};
    const hid_t sp = H5Screate_simple(1, dims, nullptr);
    if (sp < 0) return -2;
    const hid_t ds = H5Dcreate2(f->file, name, t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    herr_t st = ds < 0 ? -1 : H5Dwrite(ds, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(sp);
    return st < 0 ? -2 : 0;
}

int qpc_h5_info(qpc_h5* f, const char* name, int64_t* n, int* type) {
    if (!f || !name) return -1;
    Silence quiet;
    const hid_t ds = H5Dopen2(f->file, name, H5P_DEFAULT);
    if (ds < 0) return -2;
    const hid_t sp = H5Dget_space(ds);
    hsize_t dims[1] = {0};
    const int nd = H5Sget_simple_extent_ndims(sp);
    if (nd == 1) H5Sget_simple_extent_dims(sp, dims, nullptr);
    const hid_t t = H5Dget_type(ds);
    if (n) *n = nd == 1 ? (int64_t)dims[0] : -1;
    if (type) *type = classify(t);
    H5Tclose(t);
    H5Sclose(sp);
    H5Dclose(ds);
    return 0;
}

int qpc_h5_load(qpc_h5* f, const char* name, int type, void* data, int64_t n) {
    const hid_t t = native(type);
    if (!f || !name || t < 0 || !data) return -1;
    int64_t have = 0;
    int rc = qpc_h5_info(f, name, &have, nullptr);
    if (rc) return rc;
    if (have != n) return -3;
    Silence quiet;
    const hid_t ds = H5Dopen2(f->file, name, H5P_DEFAULT);
    if (ds < 0) return -2;
    const herr_t st = H5Dread(ds, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    H5Dclose(ds);
    return st < 0 ? -2 : 0;
}

int qpc_h5_list(qpc_h5* f, char* buf, int64_t cap) {
    if (!f || (!buf && cap > 0)) return -1;
    std::string s;
    hsize_t idx = 0;
    if (H5Literate(f->file, H5_INDEX_NAME, H5_ITER_INC, &idx, list_one, &s) < 0) return -2;
    int count = 0;
    for (char ch : s) count += ch == '\n';
    if (cap > 0) {
        const size_t k = std::min((size_t)cap - 1, s.size());
        s.copy(buf, k);
        buf[k] = '\0';
    }
    return count;
}

}  // extern "C"
