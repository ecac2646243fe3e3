// rrtmgsw_nc.cpp -- netCDF classic-format reader (CDF-1, CDF-2, CDF-5), written from the published
// format specification: a big-endian header (magic "CDF" + version byte, record count, dimension,
// global attribute and variable lists) followed by the fixed-size variables and then the records,
// each record holding one slab of every record variable in header order.
#include "rrtmgsw_nc.h"

#include <cstdio>
#include <cstring>

namespace rrtmgsw {

namespace {

const uint32_t TAG_DIM = 0x0A, TAG_VAR = 0x0B, TAG_ATT = 0x0C;

int type_size(int t) {
  switch (t) {
    case 1: case 2: case 7: return 1;  // byte, char, ubyte
    case 3: case 8: return 2;          // short, ushort
    case 4: case 5: case 9: return 4;  // int, float, uint
    case 6: case 10: case 11: return 8;  // double, int64, uint64
    default: return 0;
  }
}

// Bounds-checked big-endian reader over the header bytes.
struct Cursor {
  const uint8_t *p;
  size_t n, i = 0;
  bool bad = false;
  uint64_t be(int bytes) {
    if (bad || n - i < (size_t)bytes) return bad = true, 0;
    uint64_t v = 0;
    for (int k = 0; k < bytes; k++) v = v << 8 | p[i + k];
    i += bytes;
    return v;
  }
  void skip(uint64_t bytes) {
    if (bad || n - i < bytes) bad = true;
    else i += bytes;
  }
  std::string name(int nn) {
    uint64_t len = be(nn);
    if (bad || len > 1u << 16 || n - i < len) return bad = true, std::string();
    std::string s(reinterpret_cast<const char *>(p + i), len);
    skip((len + 3) & ~uint64_t(3));
    return s;
  }
};

double decode(const uint8_t *q, int type) {
  uint64_t v = 0;
  int sz = type_size(type);
  for (int k = 0; k < sz; k++) v = v << 8 | q[k];
  switch (type) {
    case 1: return (double)(int8_t)v;
    case 3: return (double)(int16_t)v;
    case 4: return (double)(int32_t)v;
    case 5: { uint32_t u = (uint32_t)v; float f; memcpy(&f, &u, 4); return (double)f; }
    case 6: { double d; memcpy(&d, &v, 8); return d; }
    case 7: return (double)(uint8_t)v;
    case 8: return (double)(uint16_t)v;
    case 9: return (double)(uint32_t)v;
    case 10: return (double)(int64_t)v;
    default: return (double)v;  // 11: uint64
  }
}

}  // namespace

bool NcFile::open(const std::string &path, std::string &err) {
  path_ = path;
  vars_.clear();
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return err = "cannot open the k-table file " + path, false;
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  buf_.resize(sz > 0 ? (size_t)sz : 0);
  size_t got = sz > 0 ? fread(buf_.data(), 1, buf_.size(), f) : 0;
  fclose(f);
  if (got != buf_.size()) return err = "cannot read the k-table file " + path, false;
  static const uint8_t hdf5[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
  if (buf_.size() >= 8 && !memcmp(buf_.data(), hdf5, 8))
    return err = path + " is a netCDF-4 (HDF5) file; convert it to the classic format "
                        "(nccopy -k classic) to read it",
           false;
  if (buf_.size() < 4 || memcmp(buf_.data(), "CDF", 3))
    return err = path + " is not a netCDF file", false;
  int version = buf_[3];
  if (version != 1 && version != 2 && version != 5)
    return err = path + ": unknown netCDF format version " + std::to_string(version), false;
  const int nn = version == 5 ? 8 : 4;  // NON_NEG width
  const int off = version == 1 ? 4 : 8;  // OFFSET width
  Cursor c{buf_.data(), buf_.size()};
  c.i = 4;
  uint64_t numrecs = c.be(nn);
  const bool streaming = numrecs == (nn == 8 ? ~uint64_t(0) : 0xFFFFFFFFu);

  // dimensions
  std::vector<uint64_t> dimlen;
  int recdim = -1;
  uint32_t tag = (uint32_t)c.be(4);
  uint64_t ndims = c.be(nn);
  if (tag != TAG_DIM && !(tag == 0 && ndims == 0)) c.bad = true;
  for (uint64_t d = 0; d < ndims && !c.bad && d < 1024; d++) {
    c.name(nn);
    uint64_t len = c.be(nn);
    if (len == 0) {
      if (recdim >= 0) c.bad = true;
      recdim = (int)d;
    }
    dimlen.push_back(len);
  }
  if (ndims >= 1024) c.bad = true;

  // attributes (global and per variable) are skipped: the reader needs none of them
  auto skip_atts = [&]() {
    uint32_t t = (uint32_t)c.be(4);
    uint64_t na = c.be(nn);
    if (t != TAG_ATT && !(t == 0 && na == 0)) c.bad = true;
    for (uint64_t a = 0; a < na && !c.bad; a++) {
      c.name(nn);
      int type = (int)c.be(4);
      uint64_t ne = c.be(nn);
      int es = type_size(type);
      if (!es || ne > buf_.size()) c.bad = true;
      else c.skip((ne * es + 3) & ~uint64_t(3));
    }
  };
  skip_atts();

  // variables
  tag = (uint32_t)c.be(4);
  uint64_t nvars = c.be(nn);
  if (tag != TAG_VAR && !(tag == 0 && nvars == 0)) c.bad = true;
  int nrecvars = 0;
  for (uint64_t v = 0; v < nvars && !c.bad; v++) {
    Var var;
    var.name = c.name(nn);
    uint64_t rank = c.be(nn);
    if (rank > 64) c.bad = true;
    for (uint64_t r = 0; r < rank && !c.bad; r++) {
      uint64_t id = c.be(nn);
      if (id >= dimlen.size()) c.bad = true;
      else {
        if ((int)id == recdim) {
          if (r != 0) c.bad = true;  // only the first dimension may be the unlimited one
          var.record = true;
        }
        var.shape.push_back(dimlen[id]);
      }
    }
    skip_atts();
    var.type = (int)c.be(4);
    c.be(nn);  // vsize: recomputed below (it saturates for variables over 4 GiB)
    var.begin = c.be(off);
    if (!type_size(var.type)) c.bad = true;
    nrecvars += var.record;
    vars_.push_back(var);
  }
  if (c.bad) return err = path + ": malformed netCDF header", false;

  // record size: every record variable's slab, padded to 4 bytes unless it is the only one
  recsize_ = 0;
  for (const Var &v : vars_) {
    if (!v.record) continue;
    uint64_t s = type_size(v.type);
    for (size_t d = 1; d < v.shape.size(); d++) s *= v.shape[d];
    recsize_ += nrecvars == 1 ? s : (s + 3) & ~uint64_t(3);
  }
  if (streaming) {
    uint64_t first = buf_.size();
    for (const Var &v : vars_)
      if (v.record && v.begin < first) first = v.begin;
    numrecs = recsize_ ? (buf_.size() - first) / recsize_ : 0;
  }
  numrecs_ = numrecs;
  for (Var &v : vars_)
    if (v.record) v.shape[0] = numrecs_;

  // every variable's data lies inside the file
  for (const Var &v : vars_) {
    uint64_t elems = 1;
    for (size_t d = v.record ? 1 : 0; d < v.shape.size(); d++) {
      if (v.shape[d] && elems > ~uint64_t(0) / v.shape[d]) return err = path + ": variable too large", false;
      elems *= v.shape[d];
    }
    if (elems > buf_.size() || (v.record && recsize_ && numrecs_ > buf_.size() / recsize_ + 1))
      return err = path + ": data of variable " + v.name + " runs past the end of the file", false;
    uint64_t bytes = elems * type_size(v.type);
    uint64_t end = v.record && numrecs_ ? v.begin + (numrecs_ - 1) * recsize_ + bytes
                                        : v.begin + (v.record ? 0 : bytes);
    if (end > buf_.size() || end < v.begin)
      return err = path + ": data of variable " + v.name + " runs past the end of the file", false;
  }
  return true;
}

const NcFile::Var *NcFile::var(const std::string &name) const {
  for (const Var &v : vars_)
    if (v.name == name) return &v;
  return nullptr;
}

bool NcFile::get(const Var &v, const uint64_t *start, const uint64_t *count, double *out,
                 std::string &err) const {
  if (v.type == 2) return err = path_ + ": variable " + v.name + " holds characters", false;
  const int R = (int)v.shape.size();
  for (int d = 0; d < R; d++)
    if (start[d] > v.shape[d] || count[d] > v.shape[d] - start[d])
      return err = path_ + ": start/count out of range for variable " + v.name + " (dimension " +
                   std::to_string(d) + ": start " + std::to_string(start[d]) + ", count " +
                   std::to_string(count[d]) + ", length " + std::to_string(v.shape[d]) + ")",
             false;
  uint64_t total = 1;
  for (int d = 0; d < R; d++) total *= count[d];
  if (!total) return true;
  const int es = type_size(v.type);
  // element strides in bytes (a record variable's first dimension steps by the record size)
  std::vector<uint64_t> stride(R ? R : 1, es);
  for (int d = R - 2; d >= 0; d--) stride[d] = stride[d + 1] * v.shape[d + 1];
  if (v.record) stride[0] = recsize_;
  std::vector<uint64_t> idx(R ? R : 1, 0);
  // the innermost dimension is read as one contiguous run, unless it is the record dimension
  const bool contig = R > 0 && !(v.record && R == 1);
  const uint64_t run = contig ? count[R - 1] : 1;
  const int outer = contig ? R - 1 : R;   // dimensions the odometer steps
  for (uint64_t done = 0; done < total; done += run) {
    uint64_t o = v.begin;
    for (int d = 0; d < R; d++) o += (start[d] + idx[d]) * stride[d];
    for (uint64_t k = 0; k < run; k++) out[done + k] = decode(buf_.data() + o + k * es, v.type);
    for (int d = outer - 1; d >= 0; d--) {
      if (++idx[d] < count[d]) break;
      idx[d] = 0;
    }
  }
  return true;
}

}  // namespace rrtmgsw
