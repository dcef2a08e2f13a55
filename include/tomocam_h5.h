/* Real-data front end: HDF5 projection I/O (host side, libtomocam_h5.so).
 *
 * The device side of the front end, the (proj, slice, col) -> (slice, proj,
 * col) transpose with the even-width column drop, is tc_projections_to_
 * sinogram in tomocam_hip.h.
 *
 * Conventions: every function returns 0 on success or a non-zero status
 * (1: bad argument / shape / type, 2: HDF5 error); tc_h5_last_error() gives
 * the message of the calling thread's last failure.  Datasets hold float32
 * (the reference reads its float instantiation and rejects other types).
 */
#ifndef TOMOCAM_H5_H
#define TOMOCAM_H5_H

#ifdef __cplusplus
extern "C" {
#endif

/* Rank and extents of a dataset (tomocam::h5::Reader::dims,
 * src/hdf5/reader.h:43-56, which requires rank 3; rank 1 and 2 are accepted
 * here for the angle arrays).  dims receives up to 3 extents. */
int tc_h5_dims(const char* file, const char* dataset, int* ndim, long* dims);

/* Projection block of slices [begin, end) of a [nproj][nslices][ncols]
 * dataset, as stored: h_out [nproj][end - begin][ncols].  The reference's
 * read_sinogram (src/hdf5/reader.h:66-122) reads the same hyperslab and
 * transposes it on the host; here the transpose runs on the GPU
 * (tc_projections_to_sinogram).  end < 0 means all slices. */
int tc_h5_read_projections(const char* file, const char* dataset, long begin,
                           long end, float* h_out);

/* Slices [begin, end) along axis 0 of a rank-3 dataset already in slice
 * order (Reader::read2, src/hdf5/reader.h:124-167): h_out [end - begin]
 * [d1][d2].  end < 0 means all. */
int tc_h5_read_slices(const char* file, const char* dataset, long begin,
                      long end, float* h_out);

/* A whole rank-1 dataset of n floats (Reader::read, src/hdf5/reader.h:
 * 169-190), e.g. the angles. */
int tc_h5_read_1d(const char* file, const char* dataset, float* h_out,
                  long n);

/* Writes a float32 dataset of rank ndim (1-3) and extents dims
 * (tomocam::h5::Writer::write, src/hdf5/writer.h:46-70, writes rank 3; rank
 * 1 also serves the angle arrays); truncate != 0 creates or truncates the
 * file, else the dataset is added to an existing file. */
int tc_h5_write(const char* file, const char* dataset, const float* h_data,
                int ndim, const long* dims, int truncate);

const char* tc_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TOMOCAM_H5_H */
