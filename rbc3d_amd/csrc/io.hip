// File input and output of a simulation's state (common/ModIO.F90), host code:
//   ReadWallMesh (:547-605)   an Exodus II Tri3 wall mesh.  Exodus files are netCDF; the
//                             classic (CDF-1), 64-bit-offset (CDF-2) and 64-bit-data (CDF-5)
//                             encodings are parsed here directly (big-endian header of
//                             dimensions, attributes and variables with their file offsets).
//   WriteRestart (:742-786) / ReadRestart (:792-921)   the restart file: Fortran sequential
//                             unformatted records, each framed by its byte length as a 4-byte
//                             integer before and after, 4-byte integers and 8-byte reals (the
//                             reference's build promotes REAL to 8 bytes).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"

namespace {

int io_error(const std::string &msg) {
  rbc_set_error(msg);
  return RBC3D_EIO;
}

bool read_file(const char *fn, std::vector<unsigned char> &buf) {
  FILE *f = fopen(fn, "rb");
  if (!f) return false;
  if (fseek(f, 0, SEEK_END) != 0) {
    fclose(f);
    return false;
  }
  const long n = ftell(f);
  if (n < 0 || fseek(f, 0, SEEK_SET) != 0) {
    fclose(f);
    return false;
  }
  buf.resize((size_t)n);
  const bool ok = n == 0 || fread(buf.data(), 1, (size_t)n, f) == (size_t)n;
  fclose(f);
  return ok;
}

// --- netCDF ---------------------------------------------------------------------------
enum { NC_DIMENSION = 10, NC_VARIABLE = 11, NC_ATTRIBUTE = 12 };

struct NcVar {
  std::string name;
  std::vector<long long> dims;  // dimension ids
  int type = 0;
  unsigned long long begin = 0;
};

struct NcFile {
  int version = 0;
  std::vector<std::string> dim_name;
  std::vector<long long> dim_len;  // 0: the record dimension
  std::vector<NcVar> vars;
  const unsigned char *p = nullptr;
  size_t n = 0, pos = 0;
  bool bad = false;

  unsigned long long be(int bytes) {
    if (pos + bytes > n) {
      bad = true;
      return 0;
    }
    unsigned long long v = 0;
    for (int k = 0; k < bytes; k++) v = (v << 8) | p[pos + k];
    pos += bytes;
    return v;
  }
  long long i32() { return (long long)(int32_t)(uint32_t)be(4); }
  long long count() { return version == 5 ? (long long)be(8) : i32(); }  // nelems, lengths
  std::string name() {
    const long long len = count();
    if (len < 0 || pos + (size_t)len > n) {
      bad = true;
      return "";
    }
    std::string s((const char *)p + pos, (size_t)len);
    pos += ((size_t)len + 3) & ~(size_t)3;
    return s;
  }
  static int type_size(int t) {
    switch (t) {
      case 1: case 2: case 7: return 1;   // byte, char, ubyte
      case 3: case 8: return 2;           // short, ushort
      case 4: case 5: case 9: return 4;   // int, float, uint
      case 6: case 10: case 11: return 8; // double, int64, uint64
      default: return 0;
    }
  }
  void skip_atts() {
    const long long tag = i32();
    const long long ne = count();
    if (tag == 0 && ne == 0) return;
    if (tag != NC_ATTRIBUTE || ne < 0) {
      bad = true;
      return;
    }
    for (long long k = 0; k < ne && !bad; k++) {
      name();
      const int t = (int)i32();
      const long long m = count();
      const int sz = type_size(t);
      if (!sz || m < 0 || (size_t)m > n) {
        bad = true;
        return;
      }
      pos += ((size_t)m * sz + 3) & ~(size_t)3;
      if (pos > n) bad = true;
    }
  }
  bool parse(const std::vector<unsigned char> &buf) {
    p = buf.data();
    n = buf.size();
    if (n < 4 || p[0] != 'C' || p[1] != 'D' || p[2] != 'F') return false;
    version = p[3];
    if (version != 1 && version != 2 && version != 5) return false;
    pos = 4;
    if (version == 5) be(8); else be(4);  // numrecs
    long long tag = i32(), ne = count();
    if (!(tag == 0 && ne == 0)) {
      if (tag != NC_DIMENSION || ne < 0) return false;
      for (long long k = 0; k < ne && !bad; k++) {
        dim_name.push_back(name());
        dim_len.push_back(count());
      }
    }
    skip_atts();  // global attributes
    tag = i32();
    ne = count();
    if (!(tag == 0 && ne == 0)) {
      if (tag != NC_VARIABLE || ne < 0) return false;
      for (long long k = 0; k < ne && !bad; k++) {
        NcVar v;
        v.name = name();
        const long long nd = count();
        if (nd < 0) return false;
        for (long long d = 0; d < nd && !bad; d++) v.dims.push_back(count());
        skip_atts();
        v.type = (int)i32();
        if (version == 5) be(8); else be(4);  // vsize
        v.begin = version == 1 ? (unsigned long long)(uint32_t)be(4) : be(8);
        vars.push_back(v);
      }
    }
    return !bad;
  }
  long long dim(const char *nm) const {
    for (size_t k = 0; k < dim_name.size(); k++)
      if (dim_name[k] == nm) return dim_len[k];
    return -1;
  }
  const NcVar *var(const char *nm) const {
    for (const NcVar &v : vars)
      if (v.name == nm) return &v;
    return nullptr;
  }
  // a non-record variable of `want` values, converted to double
  bool get(const NcVar &v, long long want, double *out) {
    long long cnt = 1;
    for (long long d : v.dims) {
      if (d < 0 || d >= (long long)dim_len.size() || dim_len[d] <= 0) return false;
      cnt *= dim_len[d];
      if (cnt > (long long)n) return false;
    }
    const int sz = type_size(v.type);
    if (cnt != want || !sz || v.begin + (unsigned long long)cnt * sz > n) return false;
    pos = v.begin;
    for (long long k = 0; k < cnt; k++) {
      const unsigned long long b = be(sz);
      double x;
      switch (v.type) {
        case 1: x = (double)(int8_t)b; break;
        case 3: x = (double)(int16_t)b; break;
        case 4: x = (double)(int32_t)(uint32_t)b; break;
        case 5: { uint32_t u = (uint32_t)b; float f; memcpy(&f, &u, 4); x = f; break; }
        case 6: { double d; memcpy(&d, &b, 8); x = d; break; }
        case 7: x = (double)(uint8_t)b; break;
        case 8: x = (double)(uint16_t)b; break;
        case 9: x = (double)(uint32_t)b; break;
        case 10: x = (double)(int64_t)b; break;
        case 11: x = (double)b; break;
        default: return false;
      }
      out[k] = x;
    }
    return true;
  }
};

int open_mesh(const char *fn, std::vector<unsigned char> &buf, NcFile &nc, long long &nvert,
              long long &nele) {
  if (!fn) return RBC3D_EINVAL;
  if (!read_file(fn, buf)) return io_error(std::string("ReadWallMesh: file ") + fn + " does not exist");
  if (!nc.parse(buf)) return io_error(std::string("ReadWallMesh: ") + fn + " is not a netCDF file");
  if (nc.dim("num_nod_per_el1") != 3)
    return io_error(std::string("ReadWallMesh: the mesh in ") + fn + " is not of Tri3 type");
  nvert = nc.dim("num_nodes");
  nele = nc.dim("num_elem");
  if (nvert <= 0 || nele <= 0 || nvert > INT32_MAX / 3 || nele > INT32_MAX / 3)
    return io_error(std::string("ReadWallMesh: no or too many nodes or elements in ") + fn);
  if (nc.dim("num_el_in_blk1") != nele)
    return io_error(std::string("ReadWallMesh: ") + fn + " has more than one element block");
  return RBC3D_OK;
}

// --- Fortran sequential unformatted records --------------------------------------------
struct RecWriter {
  FILE *f;
  bool ok = true;
  void rec(const void *data, size_t bytes) {
    const int32_t m = (int32_t)bytes;
    ok = ok && fwrite(&m, 4, 1, f) == 1 && (bytes == 0 || fwrite(data, 1, bytes, f) == bytes) &&
         fwrite(&m, 4, 1, f) == 1;
  }
};

struct RecReader {
  const std::vector<unsigned char> &b;
  size_t pos = 0;
  bool ok = true;
  explicit RecReader(const std::vector<unsigned char> &buf) : b(buf) {}
  // the next record, which must hold exactly `bytes` bytes
  const unsigned char *rec(size_t bytes) {
    if (!ok || pos + 4 > b.size()) return fail();
    int32_t m;
    memcpy(&m, &b[pos], 4);
    if (m < 0 || (size_t)m != bytes || pos + 8 + bytes > b.size()) return fail();
    int32_t m2;
    memcpy(&m2, &b[pos + 4 + bytes], 4);
    if (m2 != m) return fail();
    const unsigned char *p = &b[pos + 4];
    pos += 8 + bytes;
    return p;
  }
  const unsigned char *fail() {
    ok = false;
    return nullptr;
  }
};

bool rec_size_ok(long long bytes) { return bytes >= 0 && bytes <= INT32_MAX; }

// The grid ReadRestart re-creates a cell on: it infers the dealiasing factor from the
// file's nlat0 and nlat (common/ModIO.F90:858-862: 100 when nlat/nlat0 < 1.5, else the
// integer quotient) and RBC_Create sets nlat from it (common/ModRbc.F90:55-78: nlat0 + 2
// for 100, factor x nlat0 above 1, nlat0 for 1), nlon = 2 nlat.
int restart_created_nlat(int nlat0, int nlat) {
  if ((double)nlat / (double)nlat0 < 1.5) return nlat0 + 2;
  const int k = nlat / nlat0;
  return k > 1 ? k * nlat0 : nlat0;
}

}  // namespace

extern "C" int rbc3d_wall_mesh_info(const char *fn, int *nvert, int *nele) {
  if (!nvert || !nele) return RBC3D_EINVAL;
  std::vector<unsigned char> buf;
  NcFile nc;
  long long nv = 0, ne = 0;
  RBC_TRY(open_mesh(fn, buf, nc, nv, ne));
  *nvert = (int)nv;
  *nele = (int)ne;
  return RBC3D_OK;
}

extern "C" int rbc3d_read_wall_mesh(const char *fn, double *x, int *e2v) {
  if (!x || !e2v) return RBC3D_EINVAL;
  std::vector<unsigned char> buf;
  NcFile nc;
  long long nv = 0, ne = 0;
  RBC_TRY(open_mesh(fn, buf, nc, nv, ne));
  const char *cn[3] = {"coordx", "coordy", "coordz"};
  for (int d = 0; d < 3; d++) {
    const NcVar *v = nc.var(cn[d]);
    if (!v || !nc.get(*v, nv, x + (size_t)nv * d))
      return io_error(std::string("ReadWallMesh: no readable ") + cn[d] + " in " + fn);
  }
  // connect1(num_el_in_blk1, num_nod_per_el1) in netCDF's row-major order is the
  // reference's connect(3, nele); e2v = transpose(connect) is (nele, 3), column-major
  const NcVar *v = nc.var("connect1");
  std::vector<double> c((size_t)3 * ne);
  if (!v || !nc.get(*v, 3 * ne, c.data()))
    return io_error(std::string("ReadWallMesh: no readable connect1 in ") + fn);
  for (long long e = 0; e < ne; e++)
    for (int l = 0; l < 3; l++) {
      const double t = c[3 * (size_t)e + l];
      if (!(t >= 1 && t <= (double)nv))
        return io_error(std::string("ReadWallMesh: connect1 names a node outside 1..num_nodes in ") + fn);
      e2v[e + (size_t)ne * l] = (int)t;
    }
  return RBC3D_OK;
}

extern "C" int rbc3d_restart_write(const char *fn, const rbc3d_restart_info *info, const int *nlat0,
                                   const int *nlon0, const int *nlat, const int *nlon,
                                   const int *celltype, const double *starting_area,
                                   const double *x, const int *nvert, const int *nele,
                                   const double *wx, const double *wf, const int *e2v) {
  if (!fn || !info || info->nrbc < 0 || info->nwall < 0) return RBC3D_EINVAL;
  if (info->nrbc > 0 && (!nlat0 || !nlon0 || !nlat || !nlon || !celltype || !starting_area || !x))
    return RBC3D_EINVAL;
  if (info->nwall > 0 && (!nvert || !nele || !wx || !wf || !e2v)) return RBC3D_EINVAL;
  for (int i = 0; i < info->nrbc; i++)
    if (nlat[i] < 1 || nlon[i] < 1 || !rec_size_ok(24LL * nlat[i] * nlon[i])) return RBC3D_EINVAL;
  for (int k = 0; k < info->nwall; k++)
    if (nvert[k] < 1 || nele[k] < 1 || !rec_size_ok(24LL * nvert[k]) || !rec_size_ok(12LL * nele[k]))
      return RBC3D_EINVAL;
  FILE *f = fopen(fn, "wb");
  if (!f) return io_error(std::string("WriteRestart: cannot open ") + fn);
  RecWriter w{f};
  const int32_t lt = info->lt, nrbc = info->nrbc, nwall = info->nwall;
  w.rec(info->Lb, 24);
  w.rec(&lt, 4);
  w.rec(&info->time, 8);
  w.rec(info->vBkg, 24);
  w.rec(&nrbc, 4);
  size_t off = 0;
  for (int i = 0; i < nrbc; i++) {
    const int32_t a[2] = {nlat0[i], nlon0[i]}, b[2] = {nlat[i], nlon[i]}, t = celltype[i];
    w.rec(a, 8);
    w.rec(b, 8);
    w.rec(&t, 4);
    w.rec(&starting_area[i], 8);
    const size_t n = (size_t)3 * nlat[i] * nlon[i];
    w.rec(x + off, 8 * n);
    off += n;
  }
  w.rec(&nwall, 4);
  size_t vo = 0, eo = 0;
  for (int k = 0; k < nwall; k++) {
    const int32_t a[2] = {nvert[k], nele[k]};
    w.rec(a, 8);
    w.rec(wx + 3 * vo, 24 * (size_t)nvert[k]);
    w.rec(wf + 3 * vo, 24 * (size_t)nvert[k]);
    w.rec(e2v + 3 * eo, 12 * (size_t)nele[k]);
    vo += nvert[k];
    eo += nele[k];
  }
  const bool ok = w.ok && fclose(f) == 0;
  return ok ? RBC3D_OK : io_error(std::string("WriteRestart: writing ") + fn + " failed");
}

// Walks the records of a restart file held in buf; with fill, also copies the arrays out.
static int restart_parse(const char *fn, const std::vector<unsigned char> &buf, rbc3d_restart_info *info,
                         bool fill, int *nlat0, int *nlon0, int *nlat, int *nlon, int *celltype,
                         double *starting_area, double *x, int *nvert, int *nele, double *wx,
                         double *wf, int *e2v) {
  RecReader r(buf);
  const std::string bad = std::string("ReadRestart: ") + fn + " is not a restart file or is truncated";
  const unsigned char *p;
  if (!(p = r.rec(24))) return io_error(bad);
  memcpy(info->Lb, p, 24);
  int32_t i4;
  if (!(p = r.rec(4))) return io_error(bad);
  memcpy(&i4, p, 4);
  info->lt = i4;
  if (!(p = r.rec(8))) return io_error(bad);
  memcpy(&info->time, p, 8);
  if (!(p = r.rec(24))) return io_error(bad);
  memcpy(info->vBkg, p, 24);
  if (!(p = r.rec(4))) return io_error(bad);
  memcpy(&i4, p, 4);
  if (i4 < 0) return io_error(bad);
  info->nrbc = i4;
  info->npoint = 0;
  for (int i = 0; i < info->nrbc; i++) {
    int32_t a[2], b[2], t;
    double s;
    if (!(p = r.rec(8))) return io_error(bad);
    memcpy(a, p, 8);
    if (!(p = r.rec(8))) return io_error(bad);
    memcpy(b, p, 8);
    if (!(p = r.rec(4))) return io_error(bad);
    memcpy(&t, p, 4);
    if (!(p = r.rec(8))) return io_error(bad);
    memcpy(&s, p, 8);
    if (b[0] < 1 || b[1] < 1 || !rec_size_ok(24LL * b[0] * b[1])) return io_error(bad);
    const size_t n = (size_t)3 * b[0] * b[1];
    if (!(p = r.rec(8 * n))) return io_error(bad);
    // the reference stops here when the re-created grid differs (common/ModIO.F90:866-875)
    const int made = a[0] >= 1 ? restart_created_nlat(a[0], b[0]) : -1;
    if (made != b[0] || 2 * made != b[1])
      return io_error(std::string("ReadRestart: invalid array dimension in ") + fn + ": cell " +
                      std::to_string(i + 1) + " has nlat0 " + std::to_string(a[0]) + ", nlat " +
                      std::to_string(b[0]) + ", nlon " + std::to_string(b[1]) +
                      ", which RBC_Create does not reproduce");
    if (fill) {
      nlat0[i] = a[0];
      nlon0[i] = a[1];
      nlat[i] = b[0];
      nlon[i] = b[1];
      celltype[i] = t;
      starting_area[i] = s;
      memcpy(x + 3 * info->npoint, p, 8 * n);
    }
    info->npoint += (long long)b[0] * b[1];
  }
  if (!(p = r.rec(4))) return io_error(bad);
  memcpy(&i4, p, 4);
  if (i4 < 0) return io_error(bad);
  info->nwall = i4;
  info->nvert = info->nele = 0;
  for (int k = 0; k < info->nwall; k++) {
    int32_t a[2];
    if (!(p = r.rec(8))) return io_error(bad);
    memcpy(a, p, 8);
    if (a[0] < 1 || a[1] < 1 || !rec_size_ok(24LL * a[0]) || !rec_size_ok(12LL * a[1]))
      return io_error(bad);
    const unsigned char *px, *pf, *pe;
    if (!(px = r.rec(24 * (size_t)a[0])) || !(pf = r.rec(24 * (size_t)a[0])) ||
        !(pe = r.rec(12 * (size_t)a[1])))
      return io_error(bad);
    if (fill) {
      nvert[k] = a[0];
      nele[k] = a[1];
      memcpy(wx + 3 * info->nvert, px, 24 * (size_t)a[0]);
      memcpy(wf + 3 * info->nvert, pf, 24 * (size_t)a[0]);
      memcpy(e2v + 3 * info->nele, pe, 12 * (size_t)a[1]);
    }
    info->nvert += a[0];
    info->nele += a[1];
  }
  return RBC3D_OK;
}

extern "C" int rbc3d_restart_info_read(const char *fn, rbc3d_restart_info *info) {
  if (!fn || !info) return RBC3D_EINVAL;
  std::vector<unsigned char> buf;
  if (!read_file(fn, buf)) return io_error(std::string("ReadRestart: cannot read ") + fn);
  return restart_parse(fn, buf, info, false, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
}

extern "C" int rbc3d_restart_read(const char *fn, rbc3d_restart_info *info, int *nlat0, int *nlon0,
                                  int *nlat, int *nlon, int *celltype, double *starting_area,
                                  double *x, int *nvert, int *nele, double *wx, double *wf,
                                  int *e2v) {
  if (!fn || !info) return RBC3D_EINVAL;
  std::vector<unsigned char> buf;
  if (!read_file(fn, buf)) return io_error(std::string("ReadRestart: cannot read ") + fn);
  rbc3d_restart_info hdr;
  RBC_TRY(restart_parse(fn, buf, &hdr, false, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                        nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
  // the arrays were sized by the counts in *info (from rbc3d_restart_info_read)
  if (hdr.nrbc != info->nrbc || hdr.npoint != info->npoint || hdr.nwall != info->nwall ||
      hdr.nvert != info->nvert || hdr.nele != info->nele) {
    rbc_set_error(std::string("ReadRestart: the counts of ") + fn +
                  " differ from those the arrays were sized for");
    return RBC3D_EINVAL;
  }
  if ((hdr.nrbc > 0 && (!nlat0 || !nlon0 || !nlat || !nlon || !celltype || !starting_area || !x)) ||
      (hdr.nwall > 0 && (!nvert || !nele || !wx || !wf || !e2v)))
    return RBC3D_EINVAL;
  return restart_parse(fn, buf, info, true, nlat0, nlon0, nlat, nlon, celltype, starting_area, x,
                       nvert, nele, wx, wf, e2v);
}
