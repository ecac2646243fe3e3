// rrtmgsw_tables.h -- host copy of the reduced tables built by rrtmgsw_init.
#pragma once
#include <string>
#include <vector>

#include "rrtmgsw_internal.h"

namespace rrtmgsw {

struct HostTables {
  std::vector<double> tab;       // k-tables (g fastest), solar tables, cloud / aerosol optics
  std::vector<double> exp_tbl;   // NTBL + 1
  std::vector<GPoint> gp;        // NGPT
  std::vector<double> rwgt;      // 224 reduction weights
  std::vector<double> pthr;      // NPTHR pressure thresholds of setcoef_sw's index decisions
  int32_t extliq1, ssaliq1, asyliq1, extice2, ssaice2, asyice2, extice3, ssaice3, asyice3, fdlice3;
  int32_t abari, bbari, cbari, dbari, ebari, fbari, rsrtaua, rsrpiza, rsrasya;
};

// Builds the tables from the data compiled into the library or, when `ktable_file` is given,
// with the k-distribution and solar-source data read from that netCDF file as
// src/rrtmg_sw_read_nc.f90 reads rrtmg_sw.nc (cloud and aerosol optics stay compiled in, as the
// reference keeps them in module data under either build).
bool build_tables(HostTables &h, std::string &err, const char *ktable_file = nullptr);
// The pressures (hPa) at which setcoef_sw's jp, troposphere and lowest-layer decisions switch,
// on the host libm's log (layout PTHR_*).
bool pressure_thresholds(std::vector<double> &pthr, std::string &err);

}  // namespace rrtmgsw
