/* On-disk formats of the reference, read and written with the HDF5 C library (host memory, no GPU).
 *
 * The reference (Julia, HDF5.jl) writes column-major arrays; a Julia array of size (n1, ..., nk) is an
 * HDF5 dataset of dims (nk, ..., n1) with the same bytes, i.e. exactly this library's C-contiguous
 * arrays of the reversed shape (Q: (6, Nz_tot, Ny_tot, Nx_tot), a metric: (Nz_tot, Ny_tot, Nx_tot)).
 *
 *   mesh.h5     (parse_mesh.jl:262-273; read at run.jl:79-82, solver.jl:253-273)
 *               "NG", "Nx", "Ny", "Nz": Int64 scalars; "coords": Float32 Julia (3, Nx, Ny, Nz), the
 *               interior mesh points (no ghosts).
 *   metrics.h5  (parse_mesh.jl:250-261; read at solver.jl:238-251)
 *               "dξdx", "dξdy", "dξdz", "dηdx", ..., "dζdz", "J" (UTF-8 names): Float32 Julia
 *               (Nx_tot, Ny_tot, Nz_tot), ghosts included.
 *   chk-*.h5    (IO.jl:152-179; restart read at solver.jl:211-222)
 *               "Q_h": Float32 Julia (Nx_tot, Ny_tot, Nz_tot, Nprim = 6, nranks); one rank here.
 *
 * Every entry returns 0 on success, a negative F3D_E_* code on a bad argument or a file whose content
 * does not match the params (message in f3d_h5_last_error()), F3D_E_IO on an HDF5 failure.  Arrays are
 * host pointers.  One process, no MPI-IO.  The library is separate from libf3d_hip.so (it links the
 * HDF5 library; the GPU path does not). */
#ifndef F3D_H5_H
#define F3D_H5_H
#include <stdint.h>

#include "f3d.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failing call on this thread. */
const char* f3d_h5_last_error(void);

/* NG, Nx, Ny, Nz of a mesh.h5 (run.jl:79-82) into dims[4]. */
int f3d_h5_mesh_dims(const char* path, int64_t* dims);

/* Mesh points with ghosts: the interior from "coords" (solver.jl:253-271), the ghost layers by the
 * reference's linear extrapolation through the boundary point (`extrapolation`, solver.jl:682-829:
 * x and y faces, the xy corners, then z over the whole xy extent; in double where the reference's
 * arrays are Float32).  p: Nxp = Nx, Nyp = Ny, Nzp = Nz, NG as in the file, each at least NG+1 (the
 * outermost ghost reflects the (NG+1)-th point).  x, y, z: double, (Nz_tot, Ny_tot, Nx_tot). */
int f3d_h5_read_mesh(const char* path, const f3d_params* p, double* x, double* y, double* z);

/* mesh.h5 from mesh points with ghosts (their interior is written), as parse_mesh.jl:262-273.
 * file_precision: F3D_PREC_F32B writes Float32 coordinates as the reference does, F3D_PREC_F64
 * Float64.  compress: deflate level 0-9 (0: contiguous, uncompressed). */
int f3d_h5_write_mesh(const char* path, const f3d_params* p, const double* x, const double* y,
                      const double* z, int32_t file_precision, int32_t compress);

/* The 9 metrics (reference order dξdx, dξdy, dξdz, dηdx, ..., dζdz) and J with ghosts from metrics.h5
 * (solver.jl:238-251), converted to double. */
int f3d_h5_read_metrics(const char* path, const f3d_params* p, double* const* metrics, double* J);

/* metrics.h5 as parse_mesh.jl:250-261 (file_precision, compress as f3d_h5_write_mesh). */
int f3d_h5_write_metrics(const char* path, const f3d_params* p, const double* const* metrics, const double* J,
                         int32_t file_precision, int32_t compress);

/* Primitive state Q (6 variables, ghosts included) of rank 1 from a checkpoint's "Q_h"
 * (solver.jl:211-222); precision: the element type of Q (F3D_PREC_F64 double, F3D_PREC_F32B float). */
int f3d_h5_read_checkpoint(const char* path, const f3d_params* p, void* Q, int32_t precision);

/* A one-rank checkpoint as IO.jl:152-179: "Q_h" Float32 (Nx_tot, Ny_tot, Nz_tot, 6, 1), shuffle +
 * deflate at level `compress` (0: contiguous, uncompressed); Q of element type `precision`. */
int f3d_h5_write_checkpoint(const char* path, const f3d_params* p, const void* Q, int32_t precision,
                            int32_t compress);

#ifdef __cplusplus
}
#endif
#endif /* F3D_H5_H */
