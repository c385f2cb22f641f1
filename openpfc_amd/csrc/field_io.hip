// field_io.hip -- raw binary field files, read and written from device memory.
//
// File format (the reference's BinaryWriter / BinaryReader,
// include/openpfc/frontend/io/binary_writer.hpp:76-205 and
// include/openpfc/kernel/simulation/binary_reader.hpp:79-200): the global
// array of gx*gy*gz values, x fastest (MPI_ORDER_FORTRAN subarray of
// dims {gx, gy, gz}), each value a native little-endian f64 (complex: f64 re,
// f64 im), no header.  A caller's buffer holds one brick local[3] at
// offset[3] of that array, x fastest.  The whole file is one contiguous span
// when the brick is the global array; otherwise every x run of the brick is
// its own span (adjacent spans are merged).
//
// Device buffers are staged through two pinned chunks: the copy of one chunk
// overlaps the file IO of the other.  FP32 buffers are widened to f64 on
// write and rounded on read, so one file serves both precisions.
#include "../../include/pfc_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cstring>
#include <fcntl.h>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

namespace pfc_amd {
extern thread_local std::string g_last_error; // pfc_hip_last_error()
}

namespace {

int io_fail(int code, const std::string &msg) {
  pfc_amd::g_last_error = msg;
  return code;
}

constexpr size_t kChunkBytes = size_t(32) << 20; // pinned staging chunk

struct Brick {
  long gx, gy, gz, lx, ly, lz, ox, oy, oz;
  int cplx, prec;
  size_t fesz() const { return cplx ? 16 : 8; }                   // file bytes/value
  size_t besz() const { return (cplx ? 2 : 1) * (prec == PFC_HIP_F64 ? 8 : 4); }
  size_t rows() const { return (size_t)ly * lz; }
  size_t count() const { return (size_t)lx * ly * lz; }
  // file byte offset of brick row r (j = r % ly, k = r / ly)
  off_t row_offset(size_t r) const {
    const long j = (long)(r % ly), k = (long)(r / ly);
    return (off_t)((((oz + k) * gy + (oy + j)) * gx + ox) * (long)fesz());
  }
  off_t file_end() const { return (off_t)(gx * gy * gz * (long)fesz()); }
};

// include/openpfc/kernel/mpi/domain_geometry.hpp:36-60
int make_brick(const int *g, const int *l, const int *o, int prec, int kind,
               const char *ctx, Brick &b) {
  if (!g || !l || !o)
    return io_fail(PFC_HIP_EINVAL, std::string(ctx) + ": null domain array");
  if (prec != PFC_HIP_F64 && prec != PFC_HIP_F32)
    return io_fail(PFC_HIP_EINVAL, std::string(ctx) + ": unknown precision");
  if (kind != PFC_HIP_IO_REAL && kind != PFC_HIP_IO_COMPLEX)
    return io_fail(PFC_HIP_EINVAL, std::string(ctx) + ": unknown field kind");
  for (int i = 0; i < 3; ++i) {
    if (g[i] <= 0 || l[i] <= 0)
      return io_fail(PFC_HIP_EINVAL,
                     std::string(ctx) + ": global/local dimensions must be positive");
    if (o[i] < 0)
      return io_fail(PFC_HIP_EINVAL, std::string(ctx) + ": offsets must be non-negative");
    if ((long long)o[i] + l[i] > g[i])
      return io_fail(PFC_HIP_EINVAL,
                     std::string(ctx) +
                         ": piece does not lie inside global domain (check offset + "
                         "local_size vs global)");
  }
  b = Brick{g[0], g[1], g[2], l[0], l[1], l[2], o[0], o[1], o[2],
            kind == PFC_HIP_IO_COMPLEX, prec};
  return 0;
}

int full_pwrite(int fd, const char *p, size_t n, off_t off) {
  while (n) {
    ssize_t w = pwrite(fd, p, n, off);
    if (w < 0) {
      if (errno == EINTR) continue;
      return -1;
    }
    p += w;
    n -= (size_t)w;
    off += w;
  }
  return 0;
}

int full_pread(int fd, char *p, size_t n, off_t off) {
  while (n) {
    ssize_t r = pread(fd, p, n, off);
    if (r < 0) {
      if (errno == EINTR) continue;
      return -1;
    }
    if (r == 0) return -2; // end of file
    p += r;
    n -= (size_t)r;
    off += r;
  }
  return 0;
}

// Rows [r0, r0 + nr) of the brick as f64 values in `f64` (file order within
// each row) <-> the file, merging rows that are adjacent in the file.
int rows_io(int fd, const Brick &b, size_t r0, size_t nr, char *f64, bool write,
            const std::string &path) {
  const size_t row_bytes = (size_t)b.lx * b.fesz();
  size_t r = 0;
  while (r < nr) {
    const off_t off = b.row_offset(r0 + r);
    size_t run = 1;
    while (r + run < nr &&
           b.row_offset(r0 + r + run) == off + (off_t)(run * row_bytes))
      ++run;
    char *p = f64 + r * row_bytes;
    const int rc = write ? full_pwrite(fd, p, run * row_bytes, off)
                         : full_pread(fd, p, run * row_bytes, off);
    if (rc == -2)
      return io_fail(PFC_HIP_EIO, "Short read from \"" + path + "\": the file is shorter "
                                  "than the global array of the domain");
    if (rc)
      return io_fail(PFC_HIP_EIO, (write ? "write to \"" : "read from \"") + path +
                                      "\": " + std::strerror(errno));
    r += run;
  }
  return 0;
}

// f32 <-> f64 conversion of n scalars
void widen(const float *s, double *d, size_t n) {
  for (size_t i = 0; i < n; ++i) d[i] = (double)s[i];
}
void narrow(const double *s, float *d, size_t n) {
  for (size_t i = 0; i < n; ++i) d[i] = (float)s[i];
}

struct Fd {
  int fd = -1;
  ~Fd() {
    if (fd >= 0) close(fd);
  }
};

int open_file(const std::string &path, bool write, bool truncate, const Brick &b, Fd &f) {
  f.fd = write ? open(path.c_str(), O_WRONLY | O_CREAT | (truncate ? O_TRUNC : 0), 0644)
               : open(path.c_str(), O_RDONLY);
  if (f.fd < 0)
    return io_fail(PFC_HIP_EIO, "cannot open \"" + path + "\" for " +
                                    (write ? "writing" : "reading") + ": " +
                                    std::strerror(errno));
  if (!write) {
    struct stat st {};
    if (fstat(f.fd, &st) == 0 && st.st_size < b.row_offset(b.rows() - 1) +
                                                  (off_t)((size_t)b.lx * b.fesz())) {
      const long got = (long)(st.st_size / (off_t)b.fesz());
      return io_fail(PFC_HIP_EIO, "Short read from \"" + path + "\": got " +
                                      std::to_string(got) + " values, the domain needs " +
                                      std::to_string((long)(b.file_end() / (off_t)b.fesz())));
    }
  }
  return 0;
}

// Host buffers: rows in chunks through one f64 staging buffer when narrowing
// or widening is needed, directly otherwise.
int host_io(const std::string &path, char *buf, const Brick &b, bool write, bool truncate) {
  Fd f;
  if (int rc = open_file(path, write, truncate, b, f)) return rc;
  const size_t row_scalars = (size_t)b.lx * (b.cplx ? 2 : 1);
  if (b.prec == PFC_HIP_F64) return rows_io(f.fd, b, 0, b.rows(), buf, write, path);
  const size_t crow = std::max<size_t>(1, kChunkBytes / (row_scalars * 8));
  std::vector<double> tmp(crow * row_scalars);
  for (size_t r0 = 0; r0 < b.rows(); r0 += crow) {
    const size_t nr = std::min(crow, b.rows() - r0);
    float *s = reinterpret_cast<float *>(buf) + r0 * row_scalars;
    if (write) widen(s, tmp.data(), nr * row_scalars);
    if (int rc = rows_io(f.fd, b, r0, nr, reinterpret_cast<char *>(tmp.data()), write, path))
      return rc;
    if (!write) narrow(tmp.data(), s, nr * row_scalars);
  }
  return 0;
}

struct Pinned {
  void *p[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  ~Pinned() {
    for (int i = 0; i < 2; ++i) {
      if (ev[i]) (void)hipEventDestroy(ev[i]);
      if (p[i]) (void)hipHostFree(p[i]);
    }
  }
};

#define IO_HIP(expr, what)                                                     \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return io_fail(PFC_HIP_EDEVICE, std::string(what) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Device buffers: chunk c is copied into pinned slot c % 2 while the other
// slot's chunk goes to or comes from the file.
int device_io(const std::string &path, char *dev, const Brick &b, bool write, bool truncate,
              hipStream_t s) {
  Fd f;
  if (int rc = open_file(path, write, truncate, b, f)) return rc;
  const size_t row_scalars = (size_t)b.lx * (b.cplx ? 2 : 1);
  const size_t row_buf = (size_t)b.lx * b.besz();
  const size_t crow = std::max<size_t>(1, kChunkBytes / (row_scalars * 8));
  const size_t nchunks = (b.rows() + crow - 1) / crow;
  Pinned pin;
  for (int i = 0; i < 2; ++i) {
    IO_HIP(hipHostMalloc(&pin.p[i], crow * row_buf), "hipHostMalloc (IO staging)");
    IO_HIP(hipEventCreateWithFlags(&pin.ev[i], hipEventDisableTiming), "hipEventCreate");
  }
  std::vector<double> wide(b.prec == PFC_HIP_F32 ? crow * row_scalars : 0);
  auto chunk_rows = [&](size_t c) { return std::min(crow, b.rows() - c * crow); };
  auto as_f64 = [&](int slot) {
    return b.prec == PFC_HIP_F64 ? static_cast<char *>(pin.p[slot])
                                 : reinterpret_cast<char *>(wide.data());
  };
  if (write) {
    auto d2h = [&](size_t c) -> int {
      const int slot = (int)(c % 2);
      IO_HIP(hipMemcpyAsync(pin.p[slot], dev + c * crow * row_buf, chunk_rows(c) * row_buf,
                            hipMemcpyDeviceToHost, s),
             "device to host copy (IO)");
      IO_HIP(hipEventRecord(pin.ev[slot], s), "hipEventRecord");
      return 0;
    };
    if (nchunks)
      if (int rc = d2h(0)) return rc;
    for (size_t c = 0; c < nchunks; ++c) {
      const int slot = (int)(c % 2);
      if (c + 1 < nchunks)
        if (int rc = d2h(c + 1)) return rc;
      IO_HIP(hipEventSynchronize(pin.ev[slot]), "hipEventSynchronize");
      if (b.prec == PFC_HIP_F32)
        widen(static_cast<float *>(pin.p[slot]), wide.data(), chunk_rows(c) * row_scalars);
      if (int rc = rows_io(f.fd, b, c * crow, chunk_rows(c), as_f64(slot), true, path))
        return rc;
    }
  } else {
    for (size_t c = 0; c < nchunks; ++c) {
      const int slot = (int)(c % 2);
      // the copy that used this slot two chunks ago must be done
      if (c >= 2) IO_HIP(hipEventSynchronize(pin.ev[slot]), "hipEventSynchronize");
      if (int rc = rows_io(f.fd, b, c * crow, chunk_rows(c), as_f64(slot), false, path))
        return rc;
      if (b.prec == PFC_HIP_F32)
        narrow(wide.data(), static_cast<float *>(pin.p[slot]), chunk_rows(c) * row_scalars);
      IO_HIP(hipMemcpyAsync(dev + c * crow * row_buf, pin.p[slot], chunk_rows(c) * row_buf,
                            hipMemcpyHostToDevice, s),
             "host to device copy (IO)");
      IO_HIP(hipEventRecord(pin.ev[slot], s), "hipEventRecord");
    }
    IO_HIP(hipStreamSynchronize(s), "hipStreamSynchronize (IO)");
  }
  return 0;
}

int io_entry(const char *path, const void *field, int prec, int kind, const int *g,
             const int *l, const int *o, int truncate, void *stream, bool write,
             bool device, const char *ctx) {
  Brick b{};
  if (int rc = make_brick(g, l, o, prec, kind, ctx, b)) return rc;
  if (!path || !field) return io_fail(PFC_HIP_EINVAL, std::string(ctx) + ": null argument");
  char *buf = const_cast<char *>(static_cast<const char *>(field));
  if (device) return device_io(path, buf, b, write, truncate != 0, (hipStream_t)stream);
  return host_io(path, buf, b, write, truncate != 0);
}

} // namespace

extern "C" {

int pfc_hip_field_write(const char *path, const void *field, int precision, int kind,
                        const int *global3, const int *local3, const int *offset3,
                        int truncate, void *stream) {
  return io_entry(path, field, precision, kind, global3, local3, offset3, truncate, stream,
                  true, true, "pfc_hip_field_write");
}

int pfc_hip_field_read(const char *path, void *field, int precision, int kind,
                       const int *global3, const int *local3, const int *offset3,
                       void *stream) {
  return io_entry(path, field, precision, kind, global3, local3, offset3, 0, stream, false,
                  true, "pfc_hip_field_read");
}

int pfc_hip_field_write_host(const char *path, const void *field, int precision, int kind,
                             const int *global3, const int *local3, const int *offset3,
                             int truncate) {
  return io_entry(path, field, precision, kind, global3, local3, offset3, truncate, nullptr,
                  true, false, "pfc_hip_field_write_host");
}

int pfc_hip_field_read_host(const char *path, void *field, int precision, int kind,
                            const int *global3, const int *local3, const int *offset3) {
  return io_entry(path, field, precision, kind, global3, local3, offset3, 0, nullptr, false,
                  false, "pfc_hip_field_read_host");
}

} // extern "C"
