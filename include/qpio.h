/* libqpio.so: QuickPIC's HDF5 output and beam restart files, written and
 * read in the reference's on-disk format (hdf5io_class.f03) with the serial
 * HDF5 C library.  Host-side only: the caller downloads device data first
 * (qp_field3d_download, qp_beam_download).
 *
 * Every file is named <filename><dataname>_<n as %08d>.h5
 * (hdf5io_class.f03:393-395) and carries the root attributes of wrattr_file
 * (:299-347): NAME, TYPE, TIME, ITER, DT, TIME UNITS, and for TYPE "grid" an
 * AXIS group with one 2-element dataset [min, max] per axis (attributes TYPE
 * "linear", UNITS, NAME, LONG_NAME).  Reals are written as FP64 (the
 * reference's double-precision build, detect_precision :1423-1436).
 * Returns 0, or a negative code with a message from qpio_last_error(). */
#ifndef QPIO_H
#define QPIO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* type hdf5file (hdf5io_class.f03:18-35); null strings take the reference's
 * defaults ('file.h5', 'a.u.', 'Data', 'grid', axis names x1..x3) */
typedef struct qpio_file {
    const char *filename;    /* directory prefix, e.g. "./Fields/Ez/" */
    const char *dataname;    /* dataset name, e.g. "ez" */
    const char *timeunits;
    const char *ty;          /* "grid", "particles" or "rst" */
    const char *units;       /* dataset UNITS */
    const char *label;       /* dataset LONG_NAME */
    int n;                   /* ITER, and the file number */
    int rank;                /* axes written to the AXIS group */
    double t, dt;
    const char *axisname[3], *axislabel[3], *axisunits[3];
    double axismin[3], axismax[3];
} qpio_file;

const char *qpio_last_error(void);

/* the file name <filename><dataname>_<n>.h5 into out (size cap) */
int qpio_path(const qpio_file *f, char *out, size_t cap);

/* pwfield_3d / pwfield_2d and their pipelined forms (hdf5io_class.f03:
 * 371-516, 517-760): a rank-2 or rank-3 grid dataset named dataname with
 * Fortran dimensions gsize[0..rank-1] (gsize[0] fastest), data contiguous
 * in that order */
int qpio_write_grid(const qpio_file *f, int rank, const int64_t *gsize, const double *data);

/* pwpart_3d_pipe (:1018-1299) for one rank: every dspl-th of np records
 * (7 values each: x, y, z in grid units, px, py, pz, q) as the extendable
 * datasets x1..x3 (position * delta + origin), p1..p3 and q, with the root
 * attribute tp (records written) */
int qpio_write_raw(const qpio_file *f, const double *part, int64_t np, int dspl,
                   const double delta[3], const double origin[3]);

/* wpart (:1301-1360): the first np records, all 7 values, as the dataset
 * dataname of Fortran dimensions (7, np), with the root attribute tp */
int qpio_write_rst(const qpio_file *f, const double *part, int64_t np);

/* rpart (:1362-1421): reads tp, then up to max records into part (7 per
 * record); *np gets tp.  QPIO error if tp exceeds max. */
int qpio_read_rst(const qpio_file *f, double *part, int64_t max, int64_t *np);

#ifdef __cplusplus
}
#endif
#endif /* QPIO_H */
