/*
 * weldform_io.h -- C-ABI of the output writers (libweldform_io.so, host
 * only).  They take the tables that wf_pack_xdmf / wf_pack_csv
 * (weldform_hip.h) copy from the device, or the same tables packed on the
 * host, and write the reference's files:
 *
 *   wfio_write_xdmf <- Domain::WriteXDMF (Source/Output.cpp:75-518):
 *       FileKey.hdf5 with the datasets /NP, Position, ..., Contact Force
 *       (1-D, n x width values each, float32 or int32, created in the
 *       reference's order) and FileKey.xmf, the XDMF 2.0 description of a
 *       Polyvertex grid whose DataItems point at FileKey.hdf5:/<dataset>.
 *       Needs the HDF5 library the build found; without it the call fails
 *       with WFIO_ENOHDF5 and writes nothing.
 *   wfio_write_csv  <- Domain::WriteCSV (Source/Output.cpp:521-554):
 *       FileKey.csv, the header line and one line per particle, values
 *       printed as std::ostream prints a double by default (%g, 6 digits),
 *       ID as an integer.  `threads` (>= 1) format rows in parallel; the
 *       file is the same for every thread count.
 *
 * f / iv / rows are laid out as wf_pack_xdmf / wf_pack_csv fill them
 * (WF_XDMF_* offsets, WF_CSV_NCOL columns).  prop_names: the names of the
 * Property1..3 datasets (Domain::OutputName[3]); NULL for the defaults.
 * Returns 0 or a negative code; wfio_last_error() gives the message.
 */
#ifndef WELDFORM_IO_H
#define WELDFORM_IO_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WFIO_OK 0
#define WFIO_EINVAL (-1)
#define WFIO_EIO (-2)       /* file could not be created or written          */
#define WFIO_ENOHDF5 (-3)   /* library built without HDF5                    */

int wfio_has_hdf5(void);
const char *wfio_last_error(void);
int wfio_write_xdmf(const char *file_key, int64_t n, const float *f, const int32_t *iv,
                    const char *const prop_names[3]);
int wfio_write_csv(const char *file_key, int64_t n, const double *rows, int32_t threads);

#ifdef __cplusplus
}
#endif
#endif
