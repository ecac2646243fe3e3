// rrtmgsw_nc.h -- reader of netCDF classic-format files, for the k-table file rrtmg_sw.nc.
//
// Covers what src/rrtmg_sw_read_nc.f90 asks of netcdf-fortran (nf90_open, nf90_inq_varid,
// nf90_get_var with start/count into real(kind=rb) arrays) for the three classic formats: CDF-1
// (classic), CDF-2 (64-bit offsets) and CDF-5 (64-bit data). A netCDF-4 (HDF5) file is recognised
// and refused with a message naming the conversion (nccopy -k classic).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace rrtmgsw {

class NcFile {
 public:
  struct Var {
    std::string name;
    std::vector<uint64_t> shape;  // dimension lengths, slowest first; a record variable's first is numrecs
    int type = 0;                 // nc_type: 1 byte, 2 char, 3 short, 4 int, 5 float, 6 double, 7-11 CDF-5
    uint64_t begin = 0;           // file offset of the data (of record 0 for a record variable)
    bool record = false;          // first dimension is the unlimited one
  };

  // Reads and parses the whole file; false with a message on any malformed or unsupported input.
  bool open(const std::string &path, std::string &err);
  const Var *var(const std::string &name) const;
  const std::vector<Var> &vars() const { return vars_; }
  // Values start[i] .. start[i] + count[i] - 1 along each dimension (slowest first, 0-based), as
  // doubles in row-major order over `count` -- nc_get_vara_double. Integer and float types are
  // converted exactly; char variables are refused.
  bool get(const Var &v, const uint64_t *start, const uint64_t *count, double *out,
           std::string &err) const;

 private:
  std::string path_;
  std::vector<uint8_t> buf_;
  std::vector<Var> vars_;
  uint64_t numrecs_ = 0, recsize_ = 0;
};

}  // namespace rrtmgsw
