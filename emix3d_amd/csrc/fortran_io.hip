// The reference's on-disk formats: Fortran unformatted sequential files
// (OPEN(form='unformatted')), as GRAVA writes them (solver/gravar_incomp.f90:
// 2-53) and LEER_MALLA / LEER_CAMPOS / LEER_TIME / LEER_MASCARA read them
// (solver/leer_incomp.f90:1-116,215-262).  Each record is a 4-byte length,
// the data and the same length again (little-endian); integers are 4 bytes
// and reals 8 (the reference is built with -r8).  Host code only: the files
// feed the device operator through the caller's copies.
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "emx_internal.hpp"

namespace emx {
namespace {

struct FileCloser {
  FILE* f;
  ~FileCloser() {
    if (f) fclose(f);
  }
};

int open_file(const char* path, const char* mode, const char* who, FILE** f) {
  if (!path) {
    set_error("%s: null path", who);
    return EMX_ERR_ARG;
  }
  *f = fopen(path, mode);
  if (!*f) {
    set_error("%s: cannot open %s", who, path);
    return EMX_ERR_IO;
  }
  return EMX_OK;
}

// One record of bytes; records longer than a 4-byte marker holds (the
// compilers' subrecords) are refused.
int write_record(FILE* f, const void* data, long long bytes, const char* who) {
  if (bytes < 0 || bytes > INT32_MAX) {
    set_error("%s: record of %lld bytes: records above 2^31-1 bytes are not supported", who,
              bytes);
    return EMX_ERR_SIZE;
  }
  const int32_t m = (int32_t)bytes;
  if (fwrite(&m, 4, 1, f) != 1 || (bytes && fwrite(data, 1, (size_t)bytes, f) != (size_t)bytes) ||
      fwrite(&m, 4, 1, f) != 1) {
    set_error("%s: write failed", who);
    return EMX_ERR_IO;
  }
  return EMX_OK;
}

// One record into buf, which must hold exactly want bytes.
int read_record(FILE* f, void* buf, long long want, const char* who) {
  int32_t m = 0, t = 0;
  if (fread(&m, 4, 1, f) != 1) {
    set_error("%s: file ends before a record of %lld bytes", who, want);
    return EMX_ERR_IO;
  }
  if (m < 0) {
    set_error("%s: continued record (marker %d): records above 2^31-1 bytes are not supported",
              who, (int)m);
    return EMX_ERR_IO;
  }
  if ((long long)m != want) {
    set_error("%s: record of %d bytes where %lld were expected", who, (int)m, want);
    return EMX_ERR_SIZE;
  }
  if ((want && fread(buf, 1, (size_t)want, f) != (size_t)want) || fread(&t, 4, 1, f) != 1) {
    set_error("%s: file ends inside a record", who);
    return EMX_ERR_IO;
  }
  if (t != m) {
    set_error("%s: trailing record marker %d does not match %d", who, (int)t, (int)m);
    return EMX_ERR_IO;
  }
  return EMX_OK;
}

int check_dims(int nx, int ny, int nz, const char* who) {
  if (nx < 1 || ny < 1 || nz < 1) {
    set_error("%s: bad grid %dx%dx%d", who, nx, ny, nz);
    return EMX_ERR_ARG;
  }
  return EMX_OK;
}

// The (nvar) and (nx, ny, nz) header records of field and grid files.
int read_header(FILE* f, int* nvar, int dims[3], const char* who) {
  EMX_TRY(read_record(f, nvar, 4, who));
  return read_record(f, dims, 12, who);
}

int expect_dims(const int dims[3], int nx, int ny, int nz, const char* who) {
  if (dims[0] != nx || dims[1] != ny || dims[2] != nz) {
    set_error("%s: file holds %dx%dx%d, the caller expects %dx%dx%d", who, dims[0], dims[1],
              dims[2], nx, ny, nz);
    return EMX_ERR_SIZE;
  }
  return EMX_OK;
}

}  // namespace
}  // namespace emx

using namespace emx;

extern "C" {

int emx_io_write_field(const char* path, int nd, int nx, int ny, int nz, const double* u,
                       const double* temp, const double* pres) {
  const char* who = "emx_io_write_field";
  EMX_TRY(check_dims(nx, ny, nz, who));
  if (!u || !temp || !pres) {
    set_error("%s: null array", who);
    return EMX_ERR_ARG;
  }
  FILE* f = nullptr;
  EMX_TRY(open_file(path, "wb", who, &f));
  FileCloser c{f};
  const long long N = (long long)nx * ny * nz, B = 8 * N;
  const int dims[3] = {nx, ny, nz};
  EMX_TRY(write_record(f, &nd, 4, who));
  EMX_TRY(write_record(f, dims, 12, who));
  for (int k = 0; k < 3; k++) EMX_TRY(write_record(f, u + k * N, B, who));
  EMX_TRY(write_record(f, temp, B, who));
  return write_record(f, pres, B, who);
}

int emx_io_field_dims(const char* path, int* nd, int* nx, int* ny, int* nz) {
  const char* who = "emx_io_field_dims";
  if (!nd || !nx || !ny || !nz) {
    set_error("%s: null output", who);
    return EMX_ERR_ARG;
  }
  FILE* f = nullptr;
  EMX_TRY(open_file(path, "rb", who, &f));
  FileCloser c{f};
  int dims[3];
  EMX_TRY(read_header(f, nd, dims, who));
  *nx = dims[0];
  *ny = dims[1];
  *nz = dims[2];
  return EMX_OK;
}

int emx_io_read_field(const char* path, int nx, int ny, int nz, double* u, double* temp,
                      double* pres) {
  const char* who = "emx_io_read_field";
  EMX_TRY(check_dims(nx, ny, nz, who));
  if (!u || !temp || !pres) {
    set_error("%s: null array", who);
    return EMX_ERR_ARG;
  }
  FILE* f = nullptr;
  EMX_TRY(open_file(path, "rb", who, &f));
  FileCloser c{f};
  int nd = 0, dims[3];
  EMX_TRY(read_header(f, &nd, dims, who));
  if (nd != 5) {  // LEER_CAMPOS checks Nd (5 equations)
    set_error("%s: %s holds %d variables, 5 expected", who, path, nd);
    return EMX_ERR_SIZE;
  }
  EMX_TRY(expect_dims(dims, nx, ny, nz, who));
  const long long N = (long long)nx * ny * nz, B = 8 * N;
  for (int k = 0; k < 3; k++) EMX_TRY(read_record(f, u + k * N, B, who));
  EMX_TRY(read_record(f, temp, B, who));
  return read_record(f, pres, B, who);
}

int emx_io_write_grid(const char* path, int nvar, int nx, int ny, int nz, const double* x) {
  const char* who = "emx_io_write_grid";
  EMX_TRY(check_dims(nx, ny, nz, who));
  if (!x) {
    set_error("%s: null array", who);
    return EMX_ERR_ARG;
  }
  FILE* f = nullptr;
  EMX_TRY(open_file(path, "wb", who, &f));
  FileCloser c{f};
  const long long N = (long long)nx * ny * nz;
  const int dims[3] = {nx, ny, nz};
  EMX_TRY(write_record(f, &nvar, 4, who));
  EMX_TRY(write_record(f, dims, 12, who));
  for (int k = 0; k < 3; k++) EMX_TRY(write_record(f, x + k * N, 8 * N, who));
  return EMX_OK;
}

int emx_io_read_grid(const char* path, int nx, int ny, int nz, double* x) {
  const char* who = "emx_io_read_grid";
  EMX_TRY(check_dims(nx, ny, nz, who));
  if (!x) {
    set_error("%s: null array", who);
    return EMX_ERR_ARG;
  }
  FILE* f = nullptr;
  EMX_TRY(open_file(path, "rb", who, &f));
  FileCloser c{f};
  int nvar = 0, dims[3];  // LEER_MALLA does not check the variable count
  EMX_TRY(read_header(f, &nvar, dims, who));
  EMX_TRY(expect_dims(dims, nx, ny, nz, who));
  const long long N = (long long)nx * ny * nz;
  for (int k = 0; k < 3; k++) EMX_TRY(read_record(f, x + k * N, 8 * N, who));
  return EMX_OK;
}

int emx_io_write_time(const char* path, double dto, int it) {
  const char* who = "emx_io_write_time";
  FILE* f = nullptr;
  EMX_TRY(open_file(path, "wb", who, &f));
  FileCloser c{f};
  unsigned char rec[12];
  memcpy(rec, &dto, 8);
  memcpy(rec + 8, &it, 4);
  return write_record(f, rec, 12, who);
}

int emx_io_read_time(const char* path, double* dto, int* it) {
  const char* who = "emx_io_read_time";
  if (!dto || !it) {
    set_error("%s: null output", who);
    return EMX_ERR_ARG;
  }
  FILE* f = nullptr;
  EMX_TRY(open_file(path, "rb", who, &f));
  FileCloser c{f};
  unsigned char rec[12];
  EMX_TRY(read_record(f, rec, 12, who));
  memcpy(dto, rec, 8);
  memcpy(it, rec + 8, 4);
  return EMX_OK;
}

int emx_io_write_scalar(const char* path, int header, int nx, int ny, int nz, const double* a) {
  const char* who = "emx_io_write_scalar";
  EMX_TRY(check_dims(nx, ny, nz, who));
  if (!a || header < 0 || header > 2) {
    set_error("%s: null array or header kind %d outside 0..2", who, header);
    return EMX_ERR_ARG;
  }
  FILE* f = nullptr;
  EMX_TRY(open_file(path, "wb", who, &f));
  FileCloser c{f};
  const int dims[3] = {nx, ny, nz};
  if (header == 2) {
    const int one = 1;
    EMX_TRY(write_record(f, &one, 4, who));
  }
  if (header >= 1) EMX_TRY(write_record(f, dims, 12, who));
  return write_record(f, a, 8LL * nx * ny * nz, who);
}

int emx_io_read_scalar(const char* path, int header, int nx, int ny, int nz, double* a) {
  const char* who = "emx_io_read_scalar";
  EMX_TRY(check_dims(nx, ny, nz, who));
  if (!a || header < 0 || header > 2) {
    set_error("%s: null array or header kind %d outside 0..2", who, header);
    return EMX_ERR_ARG;
  }
  FILE* f = nullptr;
  EMX_TRY(open_file(path, "rb", who, &f));
  FileCloser c{f};
  if (header == 2) {
    int count = 0;
    EMX_TRY(read_record(f, &count, 4, who));
  }
  if (header >= 1) {
    int dims[3];
    EMX_TRY(read_record(f, dims, 12, who));
    EMX_TRY(expect_dims(dims, nx, ny, nz, who));
  }
  return read_record(f, a, 8LL * nx * ny * nz, who);
}

}  // extern "C"
