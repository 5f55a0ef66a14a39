// On-disk formats of the reference through the HDF5 C library (host code; built apart from the HIP
// library because it links libhdf5).  Layout conventions in include/f3d_h5.h.
#include <hdf5.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/f3d_h5.h"

#define F3D_H5_API extern "C" __attribute__((visibility("default")))

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

// Owned HDF5 identifier, closed with the matching H5?close.
struct Id {
  hid_t id = -1;
  herr_t (*closer)(hid_t) = nullptr;
  Id() = default;
  Id(hid_t i, herr_t (*c)(hid_t)) : id(i), closer(c) {}
  Id(const Id&) = delete;
  Id& operator=(const Id&) = delete;
  Id(Id&& o) noexcept : id(o.id), closer(o.closer) { o.id = -1; }
  ~Id() {
    if (id >= 0 && closer) closer(id);
  }
  bool ok() const { return id >= 0; }
};

// Silences HDF5's own error stack printing for the duration of a call (failures become return codes).
struct Quiet {
  H5E_auto2_t fn = nullptr;
  void* data = nullptr;
  Quiet() {
    H5Eget_auto2(H5E_DEFAULT, &fn, &data);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  }
  ~Quiet() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

struct Grid {
  long Nx, Ny, Nz, NG, Xt, Yt, Zt;
  explicit Grid(const f3d_params* p)
      : Nx(p->Nxp), Ny(p->Nyp), Nz(p->Nzp), NG(p->NG), Xt(p->Nxp + 2 * p->NG), Yt(p->Nyp + 2 * p->NG),
        Zt(p->Nzp + 2 * p->NG) {}
  size_t cells() const { return (size_t)Xt * Yt * Zt; }
  // 1-based (i, j, k) with ghosts, the reference's indexing
  size_t at(long i, long j, long k) const { return (size_t)(i - 1) + (size_t)Xt * ((size_t)(j - 1) + (size_t)Yt * (k - 1)); }
};

int check_params(const f3d_params* p, const char* who) {
  if (!p) return fail(F3D_E_NULL, "%s: params is NULL", who);
  if (p->NG != 4) return fail(F3D_E_NG, "%s: NG = %d (the stencils need 4)", who, p->NG);
  if (p->Nxp < 1 || p->Nyp < 1 || p->Nzp < 1) return fail(F3D_E_DIMS, "%s: bad grid dims", who);
  return F3D_OK;
}

hid_t file_type(int32_t precision) { return precision == F3D_PREC_F64 ? H5T_IEEE_F64LE : H5T_IEEE_F32LE; }

// Link creation with UTF-8 names (the reference's metric names are Greek; HDF5.jl marks them UTF-8).
Id utf8_lcpl() {
  Id l(H5Pcreate(H5P_LINK_CREATE), H5Pclose);
  if (l.ok()) H5Pset_char_encoding(l.id, H5T_CSET_UTF8);
  return l;
}

// Dataset creation: chunked + shuffle + deflate at `compress` (> 0), else contiguous.
Id dcpl_for(int rank, const hsize_t* dims, int compress, bool shuffle) {
  Id d(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
  if (!d.ok() || compress <= 0) return d;
  // chunk: the whole extent of the fastest dims, the leading ones cut so a chunk stays below 64 MiB
  std::vector<hsize_t> ch(dims, dims + rank);
  hsize_t bytes = 4;
  for (int r = 0; r < rank; ++r) bytes *= ch[r];
  for (int r = 0; r < rank && bytes > (64ull << 20); ++r) {
    bytes /= ch[r];
    ch[r] = 1;
  }
  H5Pset_chunk(d.id, rank, ch.data());
  if (shuffle) H5Pset_shuffle(d.id);
  H5Pset_deflate(d.id, (unsigned)compress);
  return d;
}

// Write a whole dataset of `rank` dims from a double buffer, stored as `ftype`.
int write_array(hid_t file, const char* name, int rank, const hsize_t* dims, const double* data, hid_t ftype,
                int compress, bool shuffle, const char* who) {
  Id space(H5Screate_simple(rank, dims, nullptr), H5Sclose);
  Id lcpl = utf8_lcpl();
  Id dcpl = dcpl_for(rank, dims, compress, shuffle);
  if (!space.ok() || !lcpl.ok() || !dcpl.ok()) return fail(F3D_E_IO, "%s: HDF5 property setup failed", who);
  Id ds(H5Dcreate2(file, name, ftype, space.id, lcpl.id, dcpl.id, H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return fail(F3D_E_IO, "%s: cannot create dataset %s", who, name);
  if (H5Dwrite(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    return fail(F3D_E_IO, "%s: writing %s failed", who, name);
  return F3D_OK;
}

int write_i64_scalar(hid_t file, const char* name, int64_t v, const char* who) {
  Id space(H5Screate(H5S_SCALAR), H5Sclose);
  Id ds(H5Dcreate2(file, name, H5T_STD_I64LE, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return fail(F3D_E_IO, "%s: cannot create %s", who, name);
  if (H5Dwrite(ds.id, H5T_NATIVE_INT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v) < 0)
    return fail(F3D_E_IO, "%s: writing %s failed", who, name);
  return F3D_OK;
}

int read_i64_scalar(hid_t file, const char* name, int64_t* v, const char* who) {
  Id ds(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return fail(F3D_E_IO, "%s: no dataset %s", who, name);
  Id sp(H5Dget_space(ds.id), H5Sclose);
  if (H5Sget_simple_extent_npoints(sp.id) != 1) return fail(F3D_E_IO, "%s: %s is not a scalar", who, name);
  if (H5Dread(ds.id, H5T_NATIVE_INT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0)
    return fail(F3D_E_IO, "%s: reading %s failed", who, name);
  return F3D_OK;
}

// Open a dataset and check its dims against `want` (rank dims, C order).
int open_checked(hid_t file, const char* name, int rank, const hsize_t* want, Id& ds, const char* who) {
  ds.id = H5Dopen2(file, name, H5P_DEFAULT);
  ds.closer = H5Dclose;
  if (!ds.ok()) return fail(F3D_E_IO, "%s: no dataset %s", who, name);
  Id sp(H5Dget_space(ds.id), H5Sclose);
  const int r = H5Sget_simple_extent_ndims(sp.id);
  if (r != rank) return fail(F3D_E_DIMS, "%s: %s has rank %d, expected %d", who, name, r, rank);
  hsize_t d[8];
  H5Sget_simple_extent_dims(sp.id, d, nullptr);
  for (int q = 0; q < rank; ++q)
    if (d[q] != want[q])
      return fail(F3D_E_DIMS, "%s: %s dim %d is %llu, the params need %llu", who, name, q, (unsigned long long)d[q],
                  (unsigned long long)want[q]);
  return F3D_OK;
}

Id open_file(const char* path, bool write) {
  if (write) return Id(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  return Id(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
}

const char* const kMetricNames[9] = {"dξdx", "dξdy", "dξdz", "dηdx", "dηdy", "dηdz", "dζdx", "dζdy", "dζdz"};

// Ghost mesh points by linear extrapolation through the boundary point, in the reference's order
// (solver.jl:682-829, one rank: every face is physical): x faces, y faces, the four xy corners over
// the interior k, then the z faces over the whole xy extent.
void extrapolate(const Grid& g, double* X[3]) {
  const long NG = g.NG, Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
  auto set = [&](long i, long j, long k, auto&& f) {
    for (int c = 0; c < 3; ++c) X[c][g.at(i, j, k)] = f(X[c]);
  };
  for (long k = NG + 1; k <= Nz + NG; ++k)
    for (long j = NG + 1; j <= Ny + NG; ++j) {
      for (long i = 1; i <= NG; ++i)
        set(i, j, k, [&](double* a) { return 2 * a[g.at(NG + 1, j, k)] - a[g.at(2 * NG + 2 - i, j, k)]; });
      for (long i = Nx + NG + 1; i <= Nx + 2 * NG; ++i)
        set(i, j, k, [&](double* a) { return 2 * a[g.at(Nx + NG, j, k)] - a[g.at(2 * NG + 2 * Nx - i, j, k)]; });
    }
  for (long k = NG + 1; k <= Nz + NG; ++k)
    for (long i = NG + 1; i <= Nx + NG; ++i) {
      for (long j = 1; j <= NG; ++j)
        set(i, j, k, [&](double* a) { return 2 * a[g.at(i, NG + 1, k)] - a[g.at(i, 2 * NG + 2 - j, k)]; });
      for (long j = Ny + NG + 1; j <= Ny + 2 * NG; ++j)
        set(i, j, k, [&](double* a) { return 2 * a[g.at(i, Ny + NG, k)] - a[g.at(i, 2 * NG + 2 * Ny - j, k)]; });
    }
  for (long k = NG + 1; k <= Nz + NG; ++k) {
    for (int cx = 0; cx < 2; ++cx)
      for (int cy = 0; cy < 2; ++cy) {
        const long ib = cx ? Nx + NG : NG + 1, jb = cy ? Ny + NG : NG + 1;
        const long i0 = cx ? Nx + NG + 1 : 1, j0 = cy ? Ny + NG + 1 : 1;
        for (long j = j0; j < j0 + NG; ++j)
          for (long i = i0; i < i0 + NG; ++i)
            set(i, j, k, [&](double* a) { return a[g.at(i, jb, k)] + a[g.at(ib, j, k)] - a[g.at(ib, jb, k)]; });
      }
  }
  for (long j = 1; j <= g.Yt; ++j)
    for (long i = 1; i <= g.Xt; ++i) {
      for (long k = 1; k <= NG; ++k)
        set(i, j, k, [&](double* a) { return 2 * a[g.at(i, j, NG + 1)] - a[g.at(i, j, 2 * NG + 2 - k)]; });
      for (long k = Nz + NG + 1; k <= Nz + 2 * NG; ++k)
        set(i, j, k, [&](double* a) { return 2 * a[g.at(i, j, Nz + NG)] - a[g.at(i, j, 2 * NG + 2 * Nz - k)]; });
    }
}

}  // namespace

F3D_H5_API const char* f3d_h5_last_error(void) { return g_err; }

F3D_H5_API int f3d_h5_mesh_dims(const char* path, int64_t* dims) {
  if (!path || !dims) return fail(F3D_E_NULL, "h5_mesh_dims: NULL argument");
  Quiet q;
  Id f = open_file(path, false);
  if (!f.ok()) return fail(F3D_E_IO, "h5_mesh_dims: cannot open %s", path);
  const char* names[4] = {"NG", "Nx", "Ny", "Nz"};
  for (int n = 0; n < 4; ++n) {
    const int rc = read_i64_scalar(f.id, names[n], &dims[n], "h5_mesh_dims");
    if (rc) return rc;
  }
  return F3D_OK;
}

F3D_H5_API int f3d_h5_read_mesh(const char* path, const f3d_params* p, double* x, double* y, double* z) {
  int rc = check_params(p, "h5_read_mesh");
  if (rc) return rc;
  if (!path || !x || !y || !z) return fail(F3D_E_NULL, "h5_read_mesh: NULL argument");
  int64_t d[4];
  rc = f3d_h5_mesh_dims(path, d);
  if (rc) return rc;
  if (d[0] != p->NG || d[1] != p->Nxp || d[2] != p->Nyp || d[3] != p->Nzp)
    return fail(F3D_E_DIMS, "h5_read_mesh: file NG=%lld Nx=%lld Ny=%lld Nz=%lld, params %d %d %d %d", (long long)d[0],
                (long long)d[1], (long long)d[2], (long long)d[3], p->NG, p->Nxp, p->Nyp, p->Nzp);
  // the outermost ghost reflects the (NG+1)-th point from the boundary (solver.jl:706-708)
  if (p->Nxp < p->NG + 1 || p->Nyp < p->NG + 1 || p->Nzp < p->NG + 1)
    return fail(F3D_E_DIMS, "h5_read_mesh: the ghost extrapolation needs at least NG+1 = %d points per direction",
                p->NG + 1);
  Quiet q;
  Id f = open_file(path, false);
  if (!f.ok()) return fail(F3D_E_IO, "h5_read_mesh: cannot open %s", path);
  const Grid g(p);
  const hsize_t want[4] = {(hsize_t)g.Nz, (hsize_t)g.Ny, (hsize_t)g.Nx, 3};
  Id ds;
  rc = open_checked(f.id, "coords", 4, want, ds, "h5_read_mesh");
  if (rc) return rc;
  std::vector<double> buf((size_t)g.Nx * g.Ny * g.Nz * 3);
  if (H5Dread(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0)
    return fail(F3D_E_IO, "h5_read_mesh: reading coords failed");
  double* X[3] = {x, y, z};
  for (int c = 0; c < 3; ++c) std::memset(X[c], 0, g.cells() * sizeof(double));
  for (long k = 0; k < g.Nz; ++k)
    for (long j = 0; j < g.Ny; ++j)
      for (long i = 0; i < g.Nx; ++i)
        for (int c = 0; c < 3; ++c)
          X[c][g.at(i + g.NG + 1, j + g.NG + 1, k + g.NG + 1)] = buf[(((size_t)k * g.Ny + j) * g.Nx + i) * 3 + c];
  extrapolate(g, X);
  return F3D_OK;
}

F3D_H5_API int f3d_h5_write_mesh(const char* path, const f3d_params* p, const double* x, const double* y,
                                 const double* z, int32_t file_precision, int32_t compress) {
  int rc = check_params(p, "h5_write_mesh");
  if (rc) return rc;
  if (!path || !x || !y || !z) return fail(F3D_E_NULL, "h5_write_mesh: NULL argument");
  if (compress < 0 || compress > 9) return fail(F3D_E_ARG, "h5_write_mesh: compress %d not in 0..9", compress);
  Quiet q;
  Id f = open_file(path, true);
  if (!f.ok()) return fail(F3D_E_IO, "h5_write_mesh: cannot create %s", path);
  const Grid g(p);
  const int64_t sc[4] = {g.NG, g.Nx, g.Ny, g.Nz};
  const char* names[4] = {"NG", "Nx", "Ny", "Nz"};
  for (int n = 0; n < 4; ++n)
    if ((rc = write_i64_scalar(f.id, names[n], sc[n], "h5_write_mesh"))) return rc;
  std::vector<double> buf((size_t)g.Nx * g.Ny * g.Nz * 3);
  const double* X[3] = {x, y, z};
  for (long k = 0; k < g.Nz; ++k)
    for (long j = 0; j < g.Ny; ++j)
      for (long i = 0; i < g.Nx; ++i)
        for (int c = 0; c < 3; ++c)
          buf[(((size_t)k * g.Ny + j) * g.Nx + i) * 3 + c] = X[c][g.at(i + g.NG + 1, j + g.NG + 1, k + g.NG + 1)];
  const hsize_t dims[4] = {(hsize_t)g.Nz, (hsize_t)g.Ny, (hsize_t)g.Nx, 3};
  return write_array(f.id, "coords", 4, dims, buf.data(), file_type(file_precision), compress, false,
                     "h5_write_mesh");
}

F3D_H5_API int f3d_h5_read_metrics(const char* path, const f3d_params* p, double* const* metrics, double* J) {
  int rc = check_params(p, "h5_read_metrics");
  if (rc) return rc;
  if (!path || !metrics || !J) return fail(F3D_E_NULL, "h5_read_metrics: NULL argument");
  for (int n = 0; n < 9; ++n)
    if (!metrics[n]) return fail(F3D_E_NULL, "h5_read_metrics: metric %d is NULL", n);
  Quiet q;
  Id f = open_file(path, false);
  if (!f.ok()) return fail(F3D_E_IO, "h5_read_metrics: cannot open %s", path);
  const Grid g(p);
  const hsize_t want[3] = {(hsize_t)g.Zt, (hsize_t)g.Yt, (hsize_t)g.Xt};
  for (int n = 0; n < 10; ++n) {
    const char* name = n < 9 ? kMetricNames[n] : "J";
    double* dst = n < 9 ? metrics[n] : J;
    Id ds;
    if ((rc = open_checked(f.id, name, 3, want, ds, "h5_read_metrics"))) return rc;
    if (H5Dread(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, dst) < 0)
      return fail(F3D_E_IO, "h5_read_metrics: reading %s failed", name);
  }
  return F3D_OK;
}

F3D_H5_API int f3d_h5_write_metrics(const char* path, const f3d_params* p, const double* const* metrics,
                                    const double* J, int32_t file_precision, int32_t compress) {
  int rc = check_params(p, "h5_write_metrics");
  if (rc) return rc;
  if (!path || !metrics || !J) return fail(F3D_E_NULL, "h5_write_metrics: NULL argument");
  for (int n = 0; n < 9; ++n)
    if (!metrics[n]) return fail(F3D_E_NULL, "h5_write_metrics: metric %d is NULL", n);
  if (compress < 0 || compress > 9) return fail(F3D_E_ARG, "h5_write_metrics: compress %d not in 0..9", compress);
  Quiet q;
  Id f = open_file(path, true);
  if (!f.ok()) return fail(F3D_E_IO, "h5_write_metrics: cannot create %s", path);
  const Grid g(p);
  const hsize_t dims[3] = {(hsize_t)g.Zt, (hsize_t)g.Yt, (hsize_t)g.Xt};
  for (int n = 0; n < 10; ++n) {
    rc = write_array(f.id, n < 9 ? kMetricNames[n] : "J", 3, dims, n < 9 ? metrics[n] : J, file_type(file_precision),
                     compress, false, "h5_write_metrics");
    if (rc) return rc;
  }
  return F3D_OK;
}

F3D_H5_API int f3d_h5_read_checkpoint(const char* path, const f3d_params* p, void* Q, int32_t precision) {
  int rc = check_params(p, "h5_read_checkpoint");
  if (rc) return rc;
  if (!path || !Q) return fail(F3D_E_NULL, "h5_read_checkpoint: NULL argument");
  if (precision != F3D_PREC_F64 && precision != F3D_PREC_F32B)
    return fail(F3D_E_ARG, "h5_read_checkpoint: unknown precision %d", precision);
  Quiet q;
  Id f = open_file(path, false);
  if (!f.ok()) return fail(F3D_E_IO, "h5_read_checkpoint: cannot open %s", path);
  const Grid g(p);
  Id ds(H5Dopen2(f.id, "Q_h", H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return fail(F3D_E_IO, "h5_read_checkpoint: no dataset Q_h");
  Id sp(H5Dget_space(ds.id), H5Sclose);
  hsize_t d[5];
  if (H5Sget_simple_extent_ndims(sp.id) != 5) return fail(F3D_E_DIMS, "h5_read_checkpoint: Q_h is not 5-D");
  H5Sget_simple_extent_dims(sp.id, d, nullptr);
  if (d[1] != 6 || d[2] != (hsize_t)g.Zt || d[3] != (hsize_t)g.Yt || d[4] != (hsize_t)g.Xt || d[0] < 1)
    return fail(F3D_E_DIMS, "h5_read_checkpoint: Q_h dims (%llu, %llu, %llu, %llu, %llu) do not fit the params",
                (unsigned long long)d[0], (unsigned long long)d[1], (unsigned long long)d[2],
                (unsigned long long)d[3], (unsigned long long)d[4]);
  // rank 1's slab: Q_h[:, :, :, :, 1] (solver.jl:215)
  const hsize_t start[5] = {0, 0, 0, 0, 0}, count[5] = {1, 6, d[2], d[3], d[4]};
  H5Sselect_hyperslab(sp.id, H5S_SELECT_SET, start, nullptr, count, nullptr);
  Id mem(H5Screate_simple(5, count, nullptr), H5Sclose);
  const hid_t mt = precision == F3D_PREC_F64 ? H5T_NATIVE_DOUBLE : H5T_NATIVE_FLOAT;
  if (H5Dread(ds.id, mt, mem.id, sp.id, H5P_DEFAULT, Q) < 0) return fail(F3D_E_IO, "h5_read_checkpoint: read failed");
  return F3D_OK;
}

F3D_H5_API int f3d_h5_write_checkpoint(const char* path, const f3d_params* p, const void* Q, int32_t precision,
                                       int32_t compress) {
  int rc = check_params(p, "h5_write_checkpoint");
  if (rc) return rc;
  if (!path || !Q) return fail(F3D_E_NULL, "h5_write_checkpoint: NULL argument");
  if (precision != F3D_PREC_F64 && precision != F3D_PREC_F32B)
    return fail(F3D_E_ARG, "h5_write_checkpoint: unknown precision %d", precision);
  if (compress < 0 || compress > 9) return fail(F3D_E_ARG, "h5_write_checkpoint: compress %d not in 0..9", compress);
  Quiet q;
  Id f = open_file(path, true);
  if (!f.ok()) return fail(F3D_E_IO, "h5_write_checkpoint: cannot create %s", path);
  const Grid g(p);
  const hsize_t dims[5] = {1, 6, (hsize_t)g.Zt, (hsize_t)g.Yt, (hsize_t)g.Xt};
  Id space(H5Screate_simple(5, dims, nullptr), H5Sclose);
  Id dcpl = dcpl_for(5, dims, compress, true);
  Id ds(H5Dcreate2(f.id, "Q_h", H5T_IEEE_F32LE, space.id, H5P_DEFAULT, dcpl.id, H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return fail(F3D_E_IO, "h5_write_checkpoint: cannot create Q_h");
  const hid_t mt = precision == F3D_PREC_F64 ? H5T_NATIVE_DOUBLE : H5T_NATIVE_FLOAT;
  if (H5Dwrite(ds.id, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, Q) < 0)
    return fail(F3D_E_IO, "h5_write_checkpoint: write failed");
  return F3D_OK;
}
