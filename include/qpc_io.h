/* qpc_io.h -- HDF5 datasets in the layout the QPC-TDSE driver writes and
 * reads (support_hdf.h hdf5_file::save/load: every dataset is 1-D, one
 * native type).  Built as its own library, lib/libqpc_io.so, against the
 * HDF5 C library, so that the propagator library does not depend on it. */
#ifndef QPC_IO_H
#define QPC_IO_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qpc_h5 qpc_h5;

/* dataset element types (support_hdf.h:76-90) */
#define QPC_H5_F64  0   /* H5T_NATIVE_DOUBLE (double)      */
#define QPC_H5_U64  1   /* H5T_NATIVE_ULONG  (size_t)      */
#define QPC_H5_I64  2   /* H5T_NATIVE_LONG   (long)        */
#define QPC_H5_I32  3   /* H5T_NATIVE_INT    (int, gauge)  */

/* hdf5_file::open (support_hdf.h:34-61): mode 'w' truncates or creates,
 * 'r' read-only, 'u' read-write.  0 on success, -1 bad mode, -2 HDF5 error. */
int qpc_h5_open(const char* path, char mode, qpc_h5** out);
int qpc_h5_close(qpc_h5* f);
/* hdf5_file::save (support_hdf.h:76-110): a 1-D dataset of n elements of
 * `type` named `name` (the driver's names, e.g. "/coef1", and "gauge"
 * without a slash); 0 or -2. */
int qpc_h5_save(qpc_h5* f, const char* name, int type, const void* data, int64_t n);
/* hdf5_file::load (support_hdf.h:111-140): reads the whole dataset
 * converted to `type` into data (n elements); -3 when the dataset holds a
 * different number of elements, -2 on an HDF5 error (missing dataset). */
int qpc_h5_load(qpc_h5* f, const char* name, int type, void* data, int64_t n);
/* elements of a dataset and its stored type (QPC_H5_* or -1 for another
 * type); -2 when it does not exist */
int qpc_h5_info(qpc_h5* f, const char* name, int64_t* n, int* type);
/* names of the root group's datasets, '\n'-separated, into buf (cap bytes,
 * truncated); returns the count */
int qpc_h5_list(qpc_h5* f, char* buf, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
