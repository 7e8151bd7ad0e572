/* libnpde_io.so: HDF5 output of the dataset pipeline, in the reference's file layouts.
 *
 * Replaces the h5py writer of the reference's data/pack.py:55-80 (one file per split and resolution)
 * for callers that have no h5py: the packer in npde/dataset.py reaches it through ctypes.
 */
#ifndef NPDE_IO_H
#define NPDE_IO_H

#ifdef __cplusplus
extern "C" {
#endif

/* 1 when the library was built with HDF5. */
int npde_io_available(void);
/* Message of the last failed call on this thread ("" after a success). */
const char* npde_io_last_error(void);
/* Write <path>: U (N, F, 4, n, n) float32 (chunks of one frame, shuffle + deflate 1), t (N, F) binary64,
 * family (N) strings, index (N) int32, attributes variables, split, n, frames, dt.  dt is the frame
 * spacing every row shares (the reference writes its constant 0.005, data/pack.py:80); pass NaN when
 * the rows have windows of their own, and the attribute is left out (the t dataset gives every row's
 * frame times).  Returns 0 on success, 1 for a bad argument, 2 for an HDF5 error. */
int npde_io_write_packed(const char* path, int N, int F, int n, const float* U, const double* t,
                         const char* const* family, const int* index, const char* split, double dt);

#ifdef __cplusplus
}
#endif
#endif /* NPDE_IO_H */
