// Restart files of the boundary-layer DNS in the reference's HDF5 layout
// (host code; see opentbl_io.h): the serial writer escribezy + writeheader
// (escru.F90:159-251, 950-997) and the reader getstartzy + readheader
// (getstartzy-bg.F90:180-295, 490-581).
#include "opentbl_io.h"

#include <hdf5.h>
#include <hdf5_hl.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct H5Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void need(bool ok, const std::string &msg) {
  if (!ok) throw std::invalid_argument(msg);
}

void h5ok(herr_t e, const std::string &what) {
  if (e < 0) throw H5Error("HDF5: " + what);
}

// an open file, closed on scope exit
struct File {
  hid_t id = -1;
  File(const char *path, bool create) {
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); // errors come back as statuses
    id = create ? H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT) : H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (id < 0) throw H5Error(std::string("HDF5: cannot ") + (create ? "create " : "open ") + path);
  }
  ~File() {
    if (id >= 0) H5Fclose(id);
  }
};

// one-element datasets of the header (h5ltmake_dataset_double_f /
// _int_f with hdims = 1, escru.F90:975-988)
void put1(hid_t f, const char *name, double v) {
  const hsize_t d = 1;
  h5ok(H5LTmake_dataset_double(f, name, 1, &d, &v), std::string("write ") + name);
}
void put1(hid_t f, const char *name, int v) {
  const hsize_t d = 1;
  h5ok(H5LTmake_dataset_int(f, name, 1, &d, &v), std::string("write ") + name);
}

// the Fortran wrapper stores a string of exactly len(field) characters
void put_string(hid_t f, const char *name, const char *s) {
  const size_t n = std::max<size_t>(1, std::strlen(s));
  hid_t t = H5Tcopy(H5T_C_S1);
  hid_t sp = H5Screate(H5S_SCALAR);
  herr_t e = H5Tset_size(t, n);
  hid_t ds = e < 0 ? -1 : H5Dcreate2(f, name, t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  e = ds < 0 ? -1 : H5Dwrite(ds, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, s);
  if (ds >= 0) H5Dclose(ds);
  H5Sclose(sp);
  H5Tclose(t);
  h5ok(e, std::string("write ") + name);
}

double get1d(hid_t f, const char *name) {
  double v = 0.0;
  h5ok(H5LTread_dataset_double(f, name, &v), std::string("read ") + name);
  return v;
}
int get1i(hid_t f, const char *name) {
  int v = 0;
  h5ok(H5LTread_dataset_int(f, name, &v), std::string("read ") + name);
  return v;
}

// shape of a dataset (rank and dims, C order)
std::vector<hsize_t> shape(hid_t f, const char *name) {
  int rank = 0;
  h5ok(H5LTget_dataset_ndims(f, name, &rank), std::string("shape of ") + name);
  std::vector<hsize_t> d(rank);
  H5T_class_t cls;
  size_t sz;
  h5ok(H5LTget_dataset_info(f, name, d.data(), &cls, &sz), std::string("shape of ") + name);
  return d;
}

// writeheader (escru.F90:950-997)
void write_header(hid_t f, const char *variable, const otbl_io_header *h, const double *y, const double *um) {
  put_string(f, "Variable", variable);
  put1(f, "tiempo", h->tiempo);
  put1(f, "cfl", h->cfl);
  put1(f, "Re", h->re);
  put1(f, "lx", h->lx);
  put1(f, "ly", h->ly);
  put1(f, "lz", h->lz);
  put1(f, "nx", h->nx);
  put1(f, "ny", h->ny);
  put1(f, "nz2", h->nz2);
  put1(f, "xout", h->xout);
  put1(f, "procs", h->procs);
  put1(f, "timeinit", h->timeinit);
  put1(f, "dt", h->dt);
  const hsize_t dy = (hsize_t)h->ny + 2, du = (hsize_t)h->ny + 1;
  h5ok(H5LTmake_dataset_double(f, "y", 1, &dy, y), "write y");
  h5ok(H5LTmake_dataset_double(f, "um", 1, &du, um), "write um");
}

// Fortran sequential unformatted records, big-endian (open(...,
// form='unformatted', convert='Big_endian'), escru.F90:300, 466): each record
// is its byte count (4-byte integer), the data, the byte count again.
struct Records {
  FILE *f = nullptr;
  std::vector<unsigned char> buf;
  explicit Records(const char *path) {
    f = std::fopen(path, "wb");
    if (!f) throw std::runtime_error(std::string("cannot create ") + path);
  }
  ~Records() {
    if (f) std::fclose(f);
  }
  void put_be(const void *p, size_t n) { // n-byte value, host (little-endian) order in
    const unsigned char *c = (const unsigned char *)p;
    for (size_t q = 0; q < n; ++q) buf.push_back(c[n - 1 - q]);
  }
  void d(double v) { put_be(&v, 8); }
  void i(int v) { put_be(&v, 4); }
  void end() {
    const int n = (int)buf.size();
    unsigned char m[4];
    for (int q = 0; q < 4; ++q) m[q] = (unsigned char)((unsigned)n >> (8 * (3 - q)));
    if (std::fwrite(m, 1, 4, f) != 4 || std::fwrite(buf.data(), 1, buf.size(), f) != buf.size() ||
        std::fwrite(m, 1, 4, f) != 4)
      throw std::runtime_error("short write");
    buf.clear();
  }
};

int fail(const std::exception &e, int code) {
  g_err = e.what();
  return code;
}

} // namespace

extern "C" {

const char *otbl_io_last_error(void) { return g_err.c_str(); }

int otbl_io_write_field(const char *path, const char *variable, const otbl_io_header *h, const double *y,
                        const double *um, const double *field, int R) {
  try {
    need(path && variable && h && y && um && field, "otbl_io_write_field: null argument");
    need(h->nx >= 1 && h->ny >= 1 && h->nz2 >= 0, "otbl_io_write_field: bad sizes in the header");
    need(R == h->ny || R == h->ny + 1, "otbl_io_write_field: R must be ny or ny+1");
    const size_t nz1 = 2 * (size_t)(h->nz2 + 1);
    const size_t n = (size_t)h->nx * R * nz1;
    // the R4 buffer of escru.F90:180-181
    std::vector<float> r4(n);
    for (size_t q = 0; q < n; ++q) r4[q] = (float)field[q];
    File f(path, true);
    const hsize_t dims[3] = {(hsize_t)h->nx, (hsize_t)R, (hsize_t)nz1};
    h5ok(H5LTmake_dataset_float(f.id, "value", 3, dims, r4.data()), "write value");
    write_header(f.id, variable, h, y, um);
    return 0;
  } catch (const std::invalid_argument &e) {
    return fail(e, 1);
  } catch (const std::exception &e) {
    return fail(e, 2);
  }
}

int otbl_io_read_header(const char *path, otbl_io_header *h, char *variable, size_t varcap, double *y,
                        double *um) {
  try {
    need(path && h, "otbl_io_read_header: null argument");
    File f(path, false);
    otbl_io_header o{};
    o.tiempo = get1d(f.id, "tiempo");
    o.cfl = get1d(f.id, "cfl");
    o.re = get1d(f.id, "Re");
    o.lx = get1d(f.id, "lx");
    o.ly = get1d(f.id, "ly");
    o.lz = get1d(f.id, "lz");
    o.nx = get1i(f.id, "nx");
    o.ny = get1i(f.id, "ny");
    o.nz2 = get1i(f.id, "nz2");
    o.xout = get1i(f.id, "xout");
    o.procs = get1i(f.id, "procs");
    o.timeinit = get1d(f.id, "timeinit");
    o.dt = get1d(f.id, "dt");
    need(o.ny >= 1, "otbl_io_read_header: bad ny in the file");
    if (variable && varcap > 0) {
      hid_t ds = H5Dopen2(f.id, "Variable", H5P_DEFAULT);
      if (ds < 0) throw H5Error("HDF5: read Variable");
      hid_t t = H5Dget_type(ds);
      const size_t n = H5Tget_size(t);
      std::vector<char> buf(n + 1, '\0');
      hid_t mt = H5Tcopy(H5T_C_S1);
      H5Tset_size(mt, n);
      const herr_t e = H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data());
      H5Tclose(mt);
      H5Tclose(t);
      H5Dclose(ds);
      h5ok(e, "read Variable");
      std::string s(buf.data(), strnlen(buf.data(), n));
      while (!s.empty() && s.back() == ' ') s.pop_back(); // Fortran blank padding
      std::snprintf(variable, varcap, "%s", s.c_str());
    }
    if (y) {
      need(shape(f.id, "y") == std::vector<hsize_t>{(hsize_t)o.ny + 2}, "otbl_io_read_header: y is not ny+2 long");
      h5ok(H5LTread_dataset_double(f.id, "y", y), "read y");
    }
    if (um) {
      need(shape(f.id, "um") == std::vector<hsize_t>{(hsize_t)o.ny + 1}, "otbl_io_read_header: um is not ny+1 long");
      h5ok(H5LTread_dataset_double(f.id, "um", um), "read um");
    }
    *h = o;
    return 0;
  } catch (const std::invalid_argument &e) {
    return fail(e, 1);
  } catch (const std::exception &e) {
    return fail(e, 2);
  }
}

int otbl_io_read_field(const char *path, double *field, int nx, int R, int nz1) {
  try {
    need(path && field && nx >= 1 && R >= 1 && nz1 >= 2, "otbl_io_read_field: bad argument");
    File f(path, false);
    const std::vector<hsize_t> d = shape(f.id, "value");
    need(d.size() == 3, "otbl_io_read_field: value is not a 3-d array");
    need(d[1] == (hsize_t)R, "otbl_io_read_field: the file has " + std::to_string(d[1]) + " rows per plane, not " +
                                 std::to_string(R) + " (a different ny is converted separately)");
    need(d[0] == (hsize_t)nx, "otbl_io_read_field: the file has " + std::to_string(d[0]) + " planes, not " +
                                  std::to_string(nx));
    const size_t fz = d[2], rows = (size_t)nx * R, nzz = std::min<size_t>(fz, (size_t)nz1);
    std::vector<float> r4(rows * fz);
    h5ok(H5LTread_dataset_float(f.id, "value", r4.data()), "read value");
    for (size_t r = 0; r < rows; ++r) {
      double *o = field + r * nz1;
      const float *s = r4.data() + r * fz;
      for (size_t k = 0; k < nzz; ++k) o[k] = (double)s[k];
      for (size_t k = nzz; k < (size_t)nz1; ++k) o[k] = 0.0;
    }
    return 0;
  } catch (const std::invalid_argument &e) {
    return fail(e, 1);
  } catch (const std::exception &e) {
    return fail(e, 2);
  }
}

int otbl_io_write_ens(const char *path, const otbl_io_header *h, const double *y, const double *um, const float *ens,
                      int nz) {
  try {
    need(path && h && y && um && ens && nz >= 1, "otbl_io_write_ens: null argument or nz < 1");
    need(h->nx >= 1 && h->ny >= 1, "otbl_io_write_ens: bad sizes in the header");
    File f(path, true);
    const hsize_t dims[3] = {(hsize_t)h->nx, (hsize_t)h->ny, (hsize_t)nz};
    h5ok(H5LTmake_dataset_float(f.id, "value", 3, dims, ens), "write value");
    write_header(f.id, "enstrophy", h, y, um);
    return 0;
  } catch (const std::invalid_argument &e) {
    return fail(e, 1);
  } catch (const std::exception &e) {
    return fail(e, 2);
  }
}

int otbl_io_write_stats(const char *st_path, const char *budget_path, const otbl_io_stats_header *h,
                        const double *y, const double *stats) {
  try {
    need(h && y && stats, "otbl_io_write_stats: null argument");
    need(h->nx >= 1 && h->ny >= 4, "otbl_io_write_stats: bad sizes in the header");
    const int nx = h->nx, ny = h->ny;
    const size_t R = (size_t)ny + 1, S = (size_t)nx * R;
    // accumulator id, plane i (1-based), row j (1-based)
    auto at = [&](int id, int i, int j) { return stats[(size_t)id * S + (size_t)(i - 1) * R + (size_t)(j - 1)]; };
    auto head = [&](Records &r) {
      for (double v : {h->tiempo, h->cfl, h->re, h->ax, h->ay, h->az}) r.d(v);
      for (int v : {nx, ny, h->nz2, h->ical}) r.i(v);
      r.end();
      for (int j = 0; j <= ny + 1; ++j) r.d(y[j]);
      r.end();
    };
    auto rows = [&](Records &r, int id, int i, int n) {
      for (int j = 1; j <= n; ++j) r.d(at(id, i, j));
    };
    if (st_path) {
      // escru.F90:300-341: header, y, a record holding 0d0, then one record
      // per plane: us(ny+1) vs uv... as listed (ua, ws, us, wa: ny+1 rows)
      Records r(st_path);
      head(r);
      r.d(0.0);
      r.end();
      // (id, rows) in record order; ids as in OTBL stats order
      static const int kSt[][2] = {{0, 1}, {1, 0}, {2, 1}, {3, 0}, {4, 1}, {5, 0}, {6, 1}, {7, 0}, {8, 0},
                                   {9, 0}, {10, 0}, {11, 0}, {12, 0}, {13, 0}, {14, 0}, {15, 0}, {16, 0}};
      for (int i = 1; i <= nx; ++i) {
        for (const auto &e : kSt) rows(r, e[0], i, ny + e[1]);
        r.end();
      }
    }
    if (budget_path) {
      // escru.F90:466-516: header, y, then per plane buf_bud(ny, 21) and
      // buf_bud2(ny+1, 2)
      Records r(budget_path);
      head(r);
      static const int kBud[21] = {17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37};
      for (int i = 1; i <= nx; ++i) {
        for (int id : kBud) rows(r, id, i, ny);
        rows(r, 38, i, ny + 1);
        rows(r, 39, i, ny + 1);
        r.end();
      }
    }
    return 0;
  } catch (const std::invalid_argument &e) {
    return fail(e, 1);
  } catch (const std::exception &e) {
    return fail(e, 2);
  }
}

} // extern "C"
