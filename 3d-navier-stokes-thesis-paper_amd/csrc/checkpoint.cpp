// Velocity checkpoints in the reference's format: savevel.out (read) and
// newvel.out (written) by read_velocity / SAVE_VELOCITY
// (PFC/read_write.f90:21-176, 180-292), Fortran sequential unformatted:
//   record 1        : NX, NY, NZ (int32), LX, LZ, RE, DT, T_TOTAL (float64)
//   records 2..nx/2+2: for x = 1..nx/2+1, slice = data1(1:nz, x, 1:ny+1, 1:3)
//                     complex128, nz*(ny+1)*3 values
// each record framed by its byte length as a 4-byte little-endian integer
// before and after (the convention of gfortran, ifort and flang for records
// below 2 GiB).  Passive-scalar and subgrid records (nscalar = 0,
// subgrid = 0 here) are neither written nor read.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "chan_internal.h"

namespace {

struct File {
  FILE* f = nullptr;
  ~File() { if (f) std::fclose(f); }
};

void put_record(FILE* f, const void* p, size_t n, const std::string& path) {
  if (n > (size_t)std::numeric_limits<int32_t>::max())
    throw chan_error("checkpoint record above 2 GiB: " + path);
  const int32_t m = (int32_t)n;
  if (std::fwrite(&m, 4, 1, f) != 1 || (n && std::fwrite(p, 1, n, f) != n) || std::fwrite(&m, 4, 1, f) != 1)
    throw chan_error("write failed: " + path);
}

void get_record(FILE* f, void* p, size_t n, const std::string& path, const char* what) {
  int32_t a = 0, b = 0;
  if (std::fread(&a, 4, 1, f) != 1) throw chan_error(std::string("unexpected end of file reading ") + what + ": " + path);
  if (a < 0 || (size_t)a != n)
    throw chan_error(std::string("record length ") + std::to_string(a) + " where " + std::to_string(n) +
                     " bytes were expected (" + what + "): " + path);
  if (n && std::fread(p, 1, n, f) != n) throw chan_error(std::string("short record (") + what + "): " + path);
  if (std::fread(&b, 4, 1, f) != 1 || b != a) throw chan_error(std::string("bad record trailer (") + what + "): " + path);
}

struct Header {
  int32_t nx, ny, nz;
  double lx, lz, re, dt, t_total;
};

void write_file(const char* path, const Header& h, const double* data1) {
  if (!path || !data1) throw chan_error("null argument");
  if (h.nx < 2 || h.ny < 1 || h.nz < 1) throw chan_error("invalid grid size");
  File fh;
  fh.f = std::fopen(path, "wb");
  if (!fh.f) throw chan_error(std::string("cannot open ") + path);
  unsigned char rec[52];
  std::memcpy(rec, &h.nx, 4); std::memcpy(rec + 4, &h.ny, 4); std::memcpy(rec + 8, &h.nz, 4);
  std::memcpy(rec + 12, &h.lx, 8); std::memcpy(rec + 20, &h.lz, 8); std::memcpy(rec + 28, &h.re, 8);
  std::memcpy(rec + 36, &h.dt, 8); std::memcpy(rec + 44, &h.t_total, 8);
  put_record(fh.f, rec, sizeof(rec), path);
  const long nxh = h.nx / 2 + 1, nyp = h.ny + 1, nz = h.nz;
  std::vector<double> slice((size_t)2 * nz * nyp * 3);
  for (long x = 0; x < nxh; ++x) {
    // slice(1:nz, 1:ny+1, 1:3) = data1(1:nz, x, 1:ny+1, 1:3)
    for (long c = 0; c < 3; ++c)
      for (long j = 0; j < nyp; ++j)
        std::memcpy(&slice[(size_t)2 * ((c * nyp + j) * nz)], data1 + 2 * (((c * nyp + j) * nxh + x) * nz),
                    sizeof(double) * 2 * nz);
    put_record(fh.f, slice.data(), sizeof(double) * slice.size(), path);
  }
  if (std::fflush(fh.f) != 0) throw chan_error(std::string("write failed: ") + path);
}

Header read_header(FILE* f, const char* path) {
  unsigned char rec[52];
  get_record(f, rec, sizeof(rec), path, "header");
  Header h;
  std::memcpy(&h.nx, rec, 4); std::memcpy(&h.ny, rec + 4, 4); std::memcpy(&h.nz, rec + 8, 4);
  std::memcpy(&h.lx, rec + 12, 8); std::memcpy(&h.lz, rec + 20, 8); std::memcpy(&h.re, rec + 28, 8);
  std::memcpy(&h.dt, rec + 36, 8); std::memcpy(&h.t_total, rec + 44, 8);
  return h;
}

void read_slices(FILE* f, const char* path, const Header& h, double* data1) {
  const long nxh = h.nx / 2 + 1, nyp = h.ny + 1, nz = h.nz;
  std::vector<double> slice((size_t)2 * nz * nyp * 3);
  for (long x = 0; x < nxh; ++x) {
    get_record(f, slice.data(), sizeof(double) * slice.size(), path, "velocity slice");
    for (long c = 0; c < 3; ++c)
      for (long j = 0; j < nyp; ++j)
        std::memcpy(data1 + 2 * (((c * nyp + j) * nxh + x) * nz), &slice[(size_t)2 * ((c * nyp + j) * nz)],
                    sizeof(double) * 2 * nz);
  }
}

// The reference's consistency test (read_velocity :36-47): grid sizes
// exactly, box lengths to one unit of round-off.
void check_consistent(const Header& h, const chan_params& p, const char* path) {
  std::string bad;
  if (h.nx != p.nx) bad += " NX " + std::to_string(h.nx) + " (run " + std::to_string(p.nx) + ")";
  if (h.ny != p.ny) bad += " NY " + std::to_string(h.ny) + " (run " + std::to_string(p.ny) + ")";
  if (h.nz != p.nz) bad += " NZ " + std::to_string(h.nz) + " (run " + std::to_string(p.nz) + ")";
  const double eps = std::numeric_limits<double>::epsilon();
  if (std::fabs(p.lx - h.lx) > p.lx * eps) bad += " LX";
  if (std::fabs(p.lz - h.lz) > p.lz * eps) bad += " LZ";
  if (!bad.empty()) throw chan_error(std::string("INCONSISTENT PARAMETERS in ") + path + ":" + bad);
}

thread_local std::string g_io_error;

template <class F>
int io_guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const chan_error& e) {
    g_io_error = e.msg;
  } catch (const std::exception& e) {
    g_io_error = e.what();
  }
  return 1;
}

}  // namespace

extern "C" {

const char* chan_io_last_error(void) { return g_io_error.c_str(); }

int chan_savevel_write(const char* path, int nx, int ny, int nz, double lx, double lz, double re, double dt,
                       double t_total, const double* data1) {
  return io_guarded([&] { write_file(path, Header{nx, ny, nz, lx, lz, re, dt, t_total}, data1); });
}

int chan_savevel_read_header(const char* path, int* nx, int* ny, int* nz, double* lx, double* lz, double* re,
                             double* dt, double* t_total) {
  return io_guarded([&] {
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) throw chan_error(std::string("cannot open ") + (path ? path : "(null)"));
    const Header h = read_header(fh.f, path);
    if (nx) *nx = h.nx;
    if (ny) *ny = h.ny;
    if (nz) *nz = h.nz;
    if (lx) *lx = h.lx;
    if (lz) *lz = h.lz;
    if (re) *re = h.re;
    if (dt) *dt = h.dt;
    if (t_total) *t_total = h.t_total;
  });
}

int chan_savevel_read(const char* path, int nx, int ny, int nz, double* data1) {
  return io_guarded([&] {
    if (!data1) throw chan_error("null argument");
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) throw chan_error(std::string("cannot open ") + (path ? path : "(null)"));
    const Header h = read_header(fh.f, path);
    if (h.nx != nx || h.ny != ny || h.nz != nz)
      throw chan_error(std::string("INCONSISTENT PARAMETERS in ") + path + ": grid " + std::to_string(h.nx) + "x" +
                       std::to_string(h.ny) + "x" + std::to_string(h.nz));
    read_slices(fh.f, path, h, data1);
  });
}

int chan_save_velocity(chan_ctx* c, const char* path) {
  if (!c) return 1;
  const chan_params p = c->prm;
  std::vector<double> d1((size_t)2 * 3 * (p.ny + 1) * (p.nx / 2 + 1) * p.nz);
  double dt = 0, t = 0;
  if (int rc = chan_get_state(c, d1.data(), nullptr, nullptr)) return rc;
  if (int rc = chan_get_scalars(c, &dt, &t, nullptr, nullptr)) return rc;
  return io_guarded([&] { write_file(path, Header{p.nx, p.ny, p.nz, p.lx, p.lz, p.re, dt, t}, d1.data()); })
             ? (c->last_error = g_io_error, 1)
             : 0;
}

int chan_read_velocity(chan_ctx* c, const char* path) {
  if (!c) return 1;
  const chan_params p = c->prm;
  std::vector<double> d1((size_t)2 * 3 * (p.ny + 1) * (p.nx / 2 + 1) * p.nz);
  Header h{};
  if (io_guarded([&] {
        File fh;
        fh.f = std::fopen(path, "rb");
        if (!fh.f) throw chan_error(std::string("cannot open ") + (path ? path : "(null)"));
        h = read_header(fh.f, path);
        check_consistent(h, p, path);
        read_slices(fh.f, path, h, d1.data());
      })) {
    c->last_error = g_io_error;
    return 1;
  }
  if (int rc = chan_set_state(c, d1.data(), nullptr, nullptr)) return rc;
  // dt is taken from the file only for CFL-adaptive runs (:60-63); t_total always
  return chan_set_scalars(c, p.dt_flag == 1 ? &h.dt : nullptr, &h.t_total, nullptr, nullptr);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Legacy VTK output as output_hdf_dfsd writes it with xskip = yskip = zskip = 1
// (PFC/read_write.f90:296-500): a big-endian stream file with a
// RECTILINEAR_GRID of nx x (ny+1) x nz points (x = lx (i-1)/nx, y = yh,
// z = lz (k-1)/nz) and one float32 vector field "flowvec" ordered
// ((((u_i, i), x), plane), z).  The 30-column header lines are Fortran
// CHARACTER(30) variables with an I5/I3/I9 field written into them.
// ---------------------------------------------------------------------------
namespace {

void put_be32(std::vector<unsigned char>& out, float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  out.push_back((unsigned char)(u >> 24));
  out.push_back((unsigned char)(u >> 16));
  out.push_back((unsigned char)(u >> 8));
  out.push_back((unsigned char)u);
}

// CHARACTER(LEN=30) initialised to `text` with an integer written into
// columns a..b (1-based) right-justified, then a line feed.
std::string line30(const std::string& text, int a, int b, const std::vector<long>& vals, int width) {
  std::string s(30, ' ');
  for (size_t i = 0; i < text.size() && i < 30; ++i) s[i] = text[i];
  std::string f;
  for (long v : vals) {
    std::string t = std::to_string(v);
    if ((int)t.size() > width) t = std::string(width, '*');
    f += std::string(width - t.size(), ' ') + t;
  }
  for (int i = a - 1, k = 0; i < b && k < (int)f.size(); ++i, ++k) s[i] = f[k];
  return s + "\n";
}

void write_vtk(const char* path, int nx, int ny, int nz, double lx, double lz, const double* yh, const double* phys) {
  if (!path || !yh || !phys) throw chan_error("null argument");
  std::vector<unsigned char> b;
  auto put = [&](const std::string& s) { b.insert(b.end(), s.begin(), s.end()); };
  put("# vtk DataFile Version 2.0\n");
  put("(This is header line) vtk rectilinear grid with binary data\n");
  put("BINARY\n");
  put("DATASET RECTILINEAR_GRID\n");
  put(line30("DIMENSIONS", 12, 26, {nx, ny + 1, nz}, 5));
  put(line30("X_COORDINATES     float", 15, 17, {nx}, 3));
  for (int x = 0; x < nx; ++x) put_be32(b, (float)(lx * (double)x / (double)nx));
  b.push_back('\n');
  put(line30("Y_COORDINATES     float", 15, 17, {ny + 1}, 3));
  for (int j = 0; j <= ny; ++j) put_be32(b, (float)yh[j]);
  b.push_back('\n');
  put(line30("Z_COORDINATES     float", 15, 17, {nz}, 3));
  for (int z = 0; z < nz; ++z) put_be32(b, (float)(lz * (double)z / (double)nz));
  b.push_back('\n');
  put(line30("POINT_DATA ", 12, 20, {(long)nx * (ny + 1) * nz}, 9));
  put("VECTORS flowvec float\n");
  const long ps = (long)nx * nz, cs = (long)(ny + 1) * ps;
  for (int z = 0; z < nz; ++z)
    for (int j = 0; j <= ny; ++j)
      for (int x = 0; x < nx; ++x)
        for (int c = 0; c < 3; ++c) put_be32(b, (float)phys[c * cs + j * ps + (long)z * nx + x]);
  b.push_back('\n');
  File fh;
  fh.f = std::fopen(path, "wb");
  if (!fh.f) throw chan_error(std::string("cannot open ") + path);
  if (std::fwrite(b.data(), 1, b.size(), fh.f) != b.size()) throw chan_error(std::string("write failed: ") + path);
}

}  // namespace

extern "C" {

int chan_vtk_write(const char* path, int nx, int ny, int nz, double lx, double lz, const double* yh,
                   const double* phys) {
  return io_guarded([&] { write_vtk(path, nx, ny, nz, lx, lz, yh, phys); });
}

}  // extern "C"
