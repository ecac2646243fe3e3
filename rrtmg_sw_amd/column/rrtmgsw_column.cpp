// rrtmgsw_column.cpp -- the reference's column-mode program (src/rrtmg_sw.1col.f90) on the MI355X
// library: reads INPUT_RRTM, IN_CLD_RRTM (icld >= 1) and IN_AER_RRTM (iaer = 10) from a
// directory, runs rrtmgsw_column_sw, and writes OUTPUT_RRTM there in the reference's layout.
//
//   rrtmgsw_column [--ktable rrtmg_sw.nc] [directory]   (default: the current directory)
//
// --ktable takes the k-distribution and solar-source data from that netCDF file, as the
// reference's build with src/rrtmg_sw_read_nc.f90 does.
//
// Covered: layer input given directly (IATM = 0; src/rrtmg_sw.1col.f90:1693-1766), clear or
// overcast clouds without McICA (imca = 0; readcld :1884-1958), McICA sampling (imca = 1: 200
// samples with the Mersenne Twister, irng = 1 as the program sets it, averaged; :566-600,
// :874-946), the overlap parameters of icld 4/5 (get_alpha, idcor 0/1), the aerosol file for
// iaer = 10 (readaer :1961-2127, nstr = 2), output blocks for iout = 0, 16..29 and 98
// (:824-997). Not covered, with an error message: IATM = 1 (the atmosphere generator
// src/rrtatm.f), iaer = 6, cross-section molecules.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "rrtmgsw.h"

namespace {

constexpr int NB = 14;
constexpr double WAVENUM1[NB] = {2600., 3250., 4000., 4650., 5150., 6150., 7700.,
                                 8050., 12850., 16000., 22650., 29000., 38000., 820.};
constexpr double WAVENUM2[NB] = {3250., 4000., 4650., 5150., 6150., 7700., 8050.,
                                 12850., 16000., 22650., 29000., 38000., 50000., 2600.};
constexpr double CPDAIR = 1.004e3;   // the program's cpdair (src/rrtmg_sw.1col.f90:366)

[[noreturn]] void die(const std::string &msg) {
  std::fprintf(stderr, "rrtmgsw_column: %s\n", msg.c_str());
  std::exit(1);
}

// Fixed-column formatted input: `width` characters from `pos` (blank beyond the line's end).
std::string field(const std::string &line, size_t &pos, size_t width) {
  std::string f = pos < line.size() ? line.substr(pos, width) : std::string();
  pos += width;
  return f;
}
bool blank(const std::string &f) { return f.find_first_not_of(" \t\r") == std::string::npos; }
int ifield(const std::string &line, size_t &pos, size_t width) {
  std::string f = field(line, pos, width);
  return blank(f) ? 0 : std::atoi(f.c_str());
}
// F/E/G input editing with d digits after an implied point: a field without a decimal point
// holds an integer mantissa scaled by 10^-d.
double ffield(const std::string &line, size_t &pos, size_t width, int d) {
  std::string f = field(line, pos, width);
  if (blank(f)) return 0.0;
  std::string t;
  for (char ch : f)
    if (ch != ' ' && ch != '\r') t += (ch == 'D' || ch == 'd') ? 'E' : ch;
  double v = std::strtod(t.c_str(), nullptr);
  if (t.find('.') == std::string::npos) {
    size_t e = t.find_first_of("eE");
    double mant = std::strtod(t.substr(0, e).c_str(), nullptr) * std::pow(10.0, -d);
    v = e == std::string::npos ? mant : mant * std::pow(10.0, std::atoi(t.c_str() + e + 1));
  }
  return v;
}
void skip(size_t &pos, size_t n) { pos += n; }

std::vector<std::string> read_lines(const std::string &path) {
  std::ifstream in(path);
  if (!in) die("cannot open " + path);
  std::vector<std::string> lines;
  for (std::string l; std::getline(in, l);) lines.push_back(l);
  return lines;
}

struct Profile {
  int iaer = 0, iatm = 0, iscat = 0, istrm = 0, iout = 0, imca = 0, icld = 0, idelm = 0, icos = 0;
  int juldat = 0, isolvar = 0, nlay = 0, idcor = 0, juldat_alpha = 0;
  double sza = 0, scon = 0, solcycfrac = 0, zenith = 0, decorr_con = 0, lat = 0;
  double solvar[NB] = {}, semiss[NB] = {};
  double bndsolvar[NB], indsolvar[2];
  std::vector<double> pavel, tavel, pz, coldry, wkl, altz;   // wkl (7, nlay), altz (nlay+1) km
};

// readprof for IATM = 0 (src/rrtmg_sw.1col.f90:1364-1766, formats :1331-1336, :1870-1877)
Profile read_input(const std::string &path) {
  auto lines = read_lines(path);
  Profile p;
  size_t k = 0;
  while (k < lines.size() && (lines[k].empty() || lines[k][0] != '$')) k++;
  if (k + 1 >= lines.size()) die("INVALID INPUT_RRTM: no record 1.1 ('$')");
  size_t pos = 0;
  const std::string &r12 = lines.at(++k);
  skip(pos, 18), p.iaer = ifield(r12, pos, 2), skip(pos, 29), p.iatm = ifield(r12, pos, 1);
  skip(pos, 32), p.iscat = ifield(r12, pos, 1), skip(pos, 1), p.istrm = ifield(r12, pos, 1);
  skip(pos, 2), p.iout = ifield(r12, pos, 3), skip(pos, 3), p.imca = ifield(r12, pos, 1);
  p.icld = ifield(r12, pos, 1), skip(pos, 3), p.idelm = ifield(r12, pos, 1), p.icos = ifield(r12, pos, 1);
  if (p.idelm > 1 || p.idelm < 0 || p.icos != 0) die("INVALID MEASUREMENT COMPARISON FLAG");
  if (p.iscat != 1) die("INVALID SCATTERING OPTION CHOSEN");
  if (p.istrm != 0) die("INVALID VALUE FOR ISTRM");
  pos = 0;
  const std::string &r121 = lines.at(++k);
  skip(pos, 12), p.juldat = ifield(r121, pos, 3), skip(pos, 3), p.sza = ffield(r121, pos, 7, 4);
  skip(pos, 3), p.isolvar = ifield(r121, pos, 2), p.scon = ffield(r121, pos, 10, 4);
  p.solcycfrac = ffield(r121, pos, 10, 5);
  for (int b = 0; b < NB; b++) p.solvar[b] = ffield(r121, pos, 10, 5);
  for (int b = 0; b < NB; b++) p.bndsolvar[b] = 1.0;
  p.indsolvar[0] = p.indsolvar[1] = 1.0;
  if (p.isolvar == 1 || p.isolvar == 2) p.indsolvar[0] = p.solvar[0], p.indsolvar[1] = p.solvar[1];
  if (p.isolvar == -1 || p.isolvar == 3)
    for (int b = 0; b < NB; b++) p.bndsolvar[b] = p.solvar[b];
  p.zenith = std::cos(p.sza * (2.0 * std::asin(1.0)) / 180.0);
  pos = 0;
  const std::string &r14 = lines.at(++k);
  skip(pos, 11);
  int iemis = ifield(r14, pos, 1);
  skip(pos, 2), ifield(r14, pos, 1);
  for (int b = 0; b < NB; b++) p.semiss[b] = ffield(r14, pos, 5, 3);
  if (iemis == 0)
    for (double &e : p.semiss) e = 1.0;
  else if (iemis == 1)
    for (double &e : p.semiss) e = p.semiss[0];
  else if (iemis != 2)
    die("IEMIS = " + std::to_string(iemis) + " NOT A VALID INPUT VALUE");
  if (p.icld == 4 || p.icld == 5) {   // :1664-1690
    pos = 0;
    const std::string &rc = lines.at(++k);
    skip(pos, 8), p.idcor = ifield(rc, pos, 2);
    pos = 0;
    if (p.idcor == 0) p.decorr_con = ffield(lines.at(++k), pos, 10, 3);
    if (p.idcor == 1) p.lat = ffield(lines.at(++k), pos, 10, 3);
    // the decorrelation length takes day 94 for juldat = 0 (:1692); the Earth-Sun factor was
    // already taken from juldat itself (:1431-1435)
    p.juldat_alpha = (p.juldat == 0 && p.idcor == 1) ? 94 : p.juldat;
  }
  if (p.iatm != 0) die("IATM = 1 (the rrtatm atmosphere generator) is not available; give the layers (IATM = 0)");
  if (p.iaer == 6) die("iaer = 6 is not available in this program");
  pos = 0;
  const std::string &r21 = lines.at(++k);
  skip(pos, 1);
  int iform = ifield(r21, pos, 1);
  p.nlay = ifield(r21, pos, 3);
  int nmol = ifield(r21, pos, 5);
  if (nmol == 0) nmol = 7;
  if (nmol != 7) die("only the seven standard molecules are read (nmol = 7)");
  if (iform != 0 && iform != 1) die("IFORM must be 0 or 1");
  const int L = p.nlay;
  if (L < 4 || L > 256) die("the library takes 4..256 layers");
  p.pavel.assign(L, 0.), p.tavel.assign(L, 0.), p.pz.assign(L + 1, 0.), p.coldry.assign(L, 0.);
  p.wkl.assign(7 * L, 0.);
  p.altz.assign(L + 1, 0.);
  std::vector<double> wbrodl(L);
  // iform 0: (3f10.4,a3,i2,1x|23x,...(f7.2,f8.3,f7.2)), (8e10.3); iform 1: g15.7 / g10.4 fields
  const int w1 = iform ? 15 : 10, d1 = iform ? 7 : 4, w2 = 10, d2 = iform ? 4 : 4;
  const int wm = iform ? 15 : 10, dm = iform ? 7 : 3;
  for (int l = 0; l < L; l++) {
    pos = 0;
    const std::string &ra = lines.at(++k);
    p.pavel[l] = ffield(ra, pos, w1, d1);
    p.tavel[l] = ffield(ra, pos, w2, d2);
    ffield(ra, pos, w2, d2);          // secntk
    field(ra, pos, 3), ifield(ra, pos, 2);   // cinp, ipthak
    if (l == 0) {
      skip(pos, 1);
      p.altz[0] = ffield(ra, pos, 7, 2);
      p.pz[0] = ffield(ra, pos, 8, 3);
      ffield(ra, pos, 7, 2);
    } else {
      skip(pos, 23);
    }
    p.altz[l + 1] = ffield(ra, pos, 7, 2);
    p.pz[l + 1] = ffield(ra, pos, 8, 3);
    ffield(ra, pos, 7, 2);
    pos = 0;
    const std::string &rb = lines.at(++k);
    for (int m = 0; m < 7; m++) p.wkl[m + 7 * l] = ffield(rb, pos, wm, dm);
    wbrodl[l] = ffield(rb, pos, wm, dm);
  }
  // mixing ratios or column amounts (:1733-1752)
  bool imix = true;
  for (int m = 0; m < nmol; m++)
    if (p.wkl[m] > 1.0) imix = false;
  for (int l = 0; l < L; l++) {
    double summol = 0.0;
    for (int m = 1; m < nmol; m++) summol = summol + p.wkl[m + 7 * l];
    if (imix) {
      p.coldry[l] = wbrodl[l] / (1.0 - summol);
      for (int m = 0; m < nmol; m++) p.wkl[m + 7 * l] = p.coldry[l] * p.wkl[m + 7 * l];
    } else {
      p.coldry[l] = wbrodl[l] + summol;
    }
  }
  return p;
}

struct Clouds {
  int inflag = 0, iceflag = 0, liqflag = 0;
  std::vector<double> cldfrac, ciwp, clwp, rei, rel, tauc, ssac, asmc, fsfc;   // tauc.. (14, nlay)
};

// readcld (src/rrtmg_sw.1col.f90:1884-1958) and the conversion of readprof :1807-1834 (nstr = 2)
Clouds read_cld(const std::string &path, int L) {
  auto lines = read_lines(path);
  Clouds c;
  size_t pos = 0;
  skip(pos, 3), c.inflag = ifield(lines.at(0), pos, 2), skip(pos, 4);
  c.iceflag = ifield(lines.at(0), pos, 1), skip(pos, 4), c.liqflag = ifield(lines.at(0), pos, 1);
  c.cldfrac.assign(L, 0.), c.ciwp.assign(L, 0.), c.clwp.assign(L, 0.), c.rei.assign(L, 0.);
  c.rel.assign(L, 0.), c.tauc.assign(NB * L, 0.), c.ssac.assign(NB * L, 1.), c.asmc.assign(NB * L, 0.);
  c.fsfc.assign(NB * L, 0.);
  std::vector<double> d1(L, 0.), d2(L, 0.), d3(L, 0.), d4(L, 0.), mom1(L, 0.);
  for (size_t k = 1; k < lines.size(); k++) {
    const std::string &ln = lines[k];
    if (!ln.empty() && ln[0] == '%') break;
    if (blank(ln)) break;
    pos = 0;
    field(ln, pos, 1), skip(pos, 1);
    int lay = ifield(ln, pos, 3);
    if (lay < 1 || lay > L) die("IN_CLD_RRTM: layer " + std::to_string(lay) + " out of range");
    c.cldfrac[lay - 1] = ffield(ln, pos, 10, 5);
    d1[lay - 1] = ffield(ln, pos, 10, 5);
    d2[lay - 1] = ffield(ln, pos, 10, 5);
    if (c.inflag == 0) {
      ffield(ln, pos, 10, 5);             // moment 0
      mom1[lay - 1] = ffield(ln, pos, 10, 5);
    } else {
      d3[lay - 1] = ffield(ln, pos, 10, 5);
      d4[lay - 1] = ffield(ln, pos, 10, 5);
    }
  }
  for (int l = 0; l < L; l++) {
    if (c.inflag == 0) {
      for (int b = 0; b < NB; b++) {
        c.tauc[b + NB * l] = d1[l], c.ssac[b + NB * l] = d2[l], c.asmc[b + NB * l] = mom1[l];
        c.fsfc[b + NB * l] = mom1[l] * mom1[l];
      }
    } else {
      const double cwp = d1[l], fice = d2[l];
      c.ciwp[l] = cwp * fice;
      c.clwp[l] = cwp * (1.0 - fice);
      c.rei[l] = d3[l], c.rel[l] = d4[l];
    }
  }
  return c;
}

// readaer for iaer = 10 and nstr = 2 (src/rrtmg_sw.1col.f90:1961-2127): per-layer aerosol optical
// depth (spectral shape from aerpar, or per band), single-scattering albedo and asymmetry.
void read_aer(const std::string &path, int L, double zenith, std::vector<double> &tau,
              std::vector<double> &ssa, std::vector<double> &asm_) {
  const double rnu0[NB] = {2903., 3601., 4310., 4892., 5623., 6872., 7872.,
                           10590., 14420., 18970., 25015., 30390., 43507., 1412.};
  const double rnu1[4] = {10530.7, 14293.3, 18678.0, 24475.1};   // bands 23..26
  const double f_od0[4] = {0.1084, 0.167, 0.245, 0.3611};
  const double f_od1[4] = {0.3144, 0.4822, 0.7013, 1.0239};
  const double eps = 1.e-10;
  auto lines = read_lines(path);
  tau.assign((size_t)L * NB, 0.), ssa.assign((size_t)L * NB, 0.), asm_.assign((size_t)L * NB, 0.);
  std::vector<double> aod(L + 1, 0.);
  size_t pos = 0, k = 0;
  skip(pos, 3);
  int naer = ifield(lines.at(k++), pos, 2);
  for (int ia = 0; ia < naer; ia++) {
    pos = 0;
    const std::string &h = lines.at(k++);
    skip(pos, 2);
    int nl = ifield(h, pos, 3);
    skip(pos, 4);
    int iaod = ifield(h, pos, 1);
    skip(pos, 4);
    int issa = ifield(h, pos, 1);
    skip(pos, 4);
    int iasym = ifield(h, pos, 1);
    double aerpar[3];
    for (double &x : aerpar) x = ffield(h, pos, 8, 3);
    double specfac[NB] = {};
    if (iaod == 0) {
      if (aerpar[1] < eps) aerpar[1] = 1.0;
      for (int b = 0; b < NB; b++) {
        double rl = 10000.0 / rnu0[b];
        specfac[b] = (aerpar[1] + aerpar[2] * rl) / ((aerpar[1] + aerpar[2] - 1.0) + std::pow(rl, aerpar[0]));
      }
    }
    std::vector<int> lays;
    for (int i = 0; i < nl; i++) {
      pos = 0;
      const std::string &r = lines.at(k++);
      skip(pos, 2);
      int lay = ifield(r, pos, 3);
      double aod1[NB];
      for (double &x : aod1) x = ffield(r, pos, 7, 4);
      if (lay < 1 || lay > L) die("IN_AER_RRTM: layer out of range");
      lays.push_back(lay);
      if (iaod == 0) {
        aod[lay] = aod1[0];
        for (int b = 0; b < NB; b++) tau[(lay - 1) + (size_t)L * b] = aod[lay] * specfac[b];
      } else {
        for (int b = 0; b < NB; b++) {
          aod[lay] = std::fmax(aod[lay], aod1[b]);
          tau[(lay - 1) + (size_t)L * b] = aod1[b];
        }
      }
    }
    if (iaod == 0)   // bands 23-26: the spectral shape at the optical-depth-weighted wavenumber
      for (int bb = 0; bb < 4; bb++) {
        const int b = 7 + bb;
        double od = 0.0;
        for (int lay : lays) od += tau[(lay - 1) + (size_t)L * b];
        od = od / zenith;
        double rnu = rnu0[b] + (rnu1[bb] - rnu0[b]) * (od - f_od0[bb]) / (f_od1[bb] - f_od0[bb]);
        double rl = 10000.0 / rnu;
        double sf = (aerpar[1] + aerpar[2] * rl) / ((aerpar[1] + aerpar[2] - 1.) + std::pow(rl, aerpar[0]));
        for (int lay : lays) tau[(lay - 1) + (size_t)L * b] = tau[(lay - 1) + (size_t)L * b] * sf / specfac[b];
      }
    double s[NB], g[NB];
    pos = 0;
    for (double &x : s) x = ffield(lines.at(k), pos, 5, 3);
    k++;
    pos = 0;
    for (double &x : g) x = ffield(lines.at(k), pos, 5, 3);
    k++;
    if (iasym >= 2) die("IN_AER_RRTM: phase-function moments (iasym = 2) are not read");
    for (int lay : lays)
      for (int b = 0; b < NB; b++) {
        ssa[(lay - 1) + (size_t)L * b] = issa == 0 ? s[0] : s[b];
        asm_[(lay - 1) + (size_t)L * b] = iasym == 0 ? g[0] : g[b];
      }
  }
}

// Fortran F editing of a value in `w` characters with `d` decimals: the leading zero of |v| < 1
// is dropped when the field is too narrow for it.
std::string fmt_f(double v, int w, int d) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.*f", d, v);
  std::string s(buf);
  if (d == 0) s += '.';   // F editing always writes the decimal point
  if ((int)s.size() > w) {
    if (s.rfind("0.", 0) == 0)
      s.erase(0, 1);
    else if (s.rfind("-0.", 0) == 0)
      s.erase(1, 1);
  }
  if ((int)s.size() > w) return std::string(w, '*');
  return std::string(w - s.size(), ' ') + s;
}

// One output block (src/rrtmg_sw.1col.f90:835-873): levels from the top, pressure in the
// format its magnitude selects (outform, :498-505).
void write_rows(FILE *f, const Profile &p, double w1, double w2, const double *up, const double *dif,
                const double *dir, const double *dn, const double *net, const double *htr) {
  const int L = p.nlay;
  std::fprintf(f, "\n");
  if (p.idelm == 0)
    std::fprintf(f, "\n");
  else
    std::fprintf(f, " The downwelling direct and diffuse fluxes have been computed using the delta-M scaling approximation.\n");
  std::fprintf(f, " Wavenumbers: %s - %s cm-1, ATM %6d\n", fmt_f(w1, 6, 0).c_str(), fmt_f(w2, 6, 0).c_str(), 1);
  std::fprintf(f, " LEVEL PRESSURE   UPWARD FLUX   DIFDOWN FLUX  DIRDOWN FLUX  DOWNWARD FLUX   NET FLUX    HEATING RATE\n");
  std::fprintf(f, "          mb          W/m2          W/m2          W/m2        W/m2          W/m2       degree/day\n");
  for (int i = L; i >= 0; i--) {
    const double pzi = p.pz[i];
    std::string ps;
    if (pzi < 1.e-2)
      ps = "   " + fmt_f(pzi, 7, 6);
    else
      ps = "    " + fmt_f(pzi, 6, pzi < 1.e-1 ? 5 : pzi < 1. ? 4 : pzi < 10. ? 3 : pzi < 100. ? 2 : 1);
    std::fprintf(f, " %3d%s    %s    %s    %s    %s    %s    %s\n", i, ps.c_str(), fmt_f(up[i], 10, 4).c_str(),
                 fmt_f(dif[i], 10, 4).c_str(), fmt_f(dir[i], 10, 4).c_str(), fmt_f(dn[i], 10, 4).c_str(),
                 fmt_f(net[i], 11, 6).c_str(), fmt_f(htr[i], 10, 5).c_str());
  }
  std::fprintf(f, "\f\n");
}

// One sample's block: difdown = down - direct, net = down - up, heating rate from the net flux
// (:784-800), as the program forms them.
void write_block(FILE *f, const Profile &p, double w1, double w2, const double *up, const double *dn,
                 const double *dir, double heatfac) {
  const int L = p.nlay;
  std::vector<double> dif(L + 1), net(L + 1), htr(L + 1, 0.0);
  for (int i = 0; i <= L; i++) dif[i] = dn[i] - dir[i], net[i] = dn[i] - up[i];
  for (int i = 0; i < L; i++) htr[i] = (net[i + 1] - net[i]) * (heatfac / (p.pz[i] - p.pz[i + 1]));
  write_rows(f, p, w1, w2, up, dif.data(), dir, dn, net.data(), htr.data());
}

}  // namespace

void dump(const char *name, const double *v, size_t n) {
  std::printf("%s", name);
  for (size_t i = 0; i < n; i++) std::printf(" %.17g", v[i]);
  std::printf("\n");
}

int main(int argc, char **argv) {
  // --dump: print what was read (no GPU work), for checking the readers;
  // --ktable FILE: the k-tables from a netCDF file, as the reference's rrtmg_sw_read_nc.f90
  // build reads rrtmg_sw.nc (else the compiled-in data); then the input directory, if any
  bool dump_only = false;
  const char *ktable = nullptr;
  int a0 = 1;
  for (; a0 < argc; a0++) {
    if (std::strcmp(argv[a0], "--dump") == 0) dump_only = true;
    else if (std::strcmp(argv[a0], "--ktable") == 0 && a0 + 1 < argc) ktable = argv[++a0];
    else break;
  }
  const std::string dir = argc > a0 ? std::string(argv[a0]) + "/" : std::string();
  Profile p = read_input(dir + "INPUT_RRTM");
  const int L = p.nlay;
  Clouds cl;
  if (p.icld >= 1) cl = read_cld(dir + "IN_CLD_RRTM", L);
  std::vector<double> tauaer, ssaaer, asmaer;
  if (p.iaer == 10) read_aer(dir + "IN_AER_RRTM", L, p.zenith, tauaer, ssaaer, asmaer);
  if (dump_only) {
    const double sc[] = {(double)p.iaer, (double)p.iout, (double)p.icld, (double)p.idelm,
                         (double)p.juldat, (double)p.isolvar, p.scon, p.solcycfrac, p.zenith};
    dump("scalars", sc, 9);
    dump("pavel", p.pavel.data(), L), dump("tavel", p.tavel.data(), L);
    dump("pz", p.pz.data(), L + 1), dump("coldry", p.coldry.data(), L);
    dump("wkl", p.wkl.data(), 7 * (size_t)L), dump("semiss", p.semiss, NB);
    dump("bndsolvar", p.bndsolvar, NB), dump("indsolvar", p.indsolvar, 2);
    const double ov[] = {(double)p.imca, (double)p.idcor, p.decorr_con, p.lat, (double)p.juldat_alpha};
    dump("overlap", ov, 5), dump("altz", p.altz.data(), L + 1);
    if (p.icld >= 1) {
      const double fl[] = {(double)cl.inflag, (double)cl.iceflag, (double)cl.liqflag};
      dump("cldflags", fl, 3);
      dump("cldfrac", cl.cldfrac.data(), L), dump("ciwp", cl.ciwp.data(), L);
      dump("clwp", cl.clwp.data(), L), dump("rei", cl.rei.data(), L), dump("rel", cl.rel.data(), L);
    }
    if (p.iaer == 10) {
      dump("tauaer", tauaer.data(), tauaer.size()), dump("ssaaer", ssaaer.data(), ssaaer.size());
      dump("asmaer", asmaer.data(), asmaer.size());
    }
    return 0;
  }
  std::vector<double> alb(NB);
  for (int b = 0; b < NB; b++) alb[b] = 1.0 - p.semiss[b];
  const double cosz = p.zenith;
  if (ktable ? rrtmgsw_init_nc(CPDAIR, ktable) : rrtmgsw_init(CPDAIR)) die(rrtmgsw_last_error());
  const size_t nlev = (size_t)L + 1;
  std::vector<double> fup(nlev * 15), fdn(nlev * 15), fdir(nlev * 15);
  const bool c = p.icld >= 1;
  const double heatfac = 9.8066 * 8.6400e4 / (CPDAIR * 1.e2);   // src/rrtmg_sw_init.f90:258
  if (p.imca == 1 && c) {
    // McICA: 200 samples (nmca, :571) averaged level by level in sample order (:874-894)
    const int nmca = 200, irng = 1;
    std::vector<double> alpha(L, 0.0);
    if (p.icld == 4 || p.icld == 5) {
      std::vector<double> dz(L);
      for (int l = 0; l < L; l++) dz[l] = (p.altz[l + 1] - p.altz[l]) * 1000.0;   // km -> m (:1796)
      if (rrtmgsw_get_alpha(1, 1, L, p.icld, p.idcor, p.decorr_con, dz.data(), &p.lat, p.juldat_alpha,
                            cl.cldfrac.data(), alpha.data()))
        die(rrtmgsw_last_error());
    }
    std::vector<double> su(nmca * nlev * 15), sd(nmca * nlev * 15), sdir(nmca * nlev * 15);
    int rc = rrtmgsw_column_mcica_sw(L, nmca, p.icld, irng, p.idelm, p.pavel.data(), p.tavel.data(),
                                     p.pz.data(), p.coldry.data(), p.wkl.data(), alb.data(), cosz, 1.0,
                                     p.juldat, p.scon, p.isolvar, p.bndsolvar, p.indsolvar,
                                     &p.solcycfrac, cl.inflag, cl.iceflag, cl.liqflag, cl.cldfrac.data(),
                                     cl.ciwp.data(), cl.clwp.data(), cl.rei.data(), cl.rel.data(),
                                     cl.tauc.data(), cl.ssac.data(), cl.asmc.data(), cl.fsfc.data(),
                                     alpha.data(), p.iaer == 10 ? tauaer.data() : nullptr,
                                     p.iaer == 10 ? ssaaer.data() : nullptr,
                                     p.iaer == 10 ? asmaer.data() : nullptr, su.data(), sd.data(), sdir.data());
    if (rc) die(rrtmgsw_last_error());
    FILE *f = std::fopen((dir + "OUTPUT_RRTM").c_str(), "w");
    if (!f) die("cannot write OUTPUT_RRTM");
    std::vector<int> blocks;
    if (p.iout == 98) {
      blocks.push_back(0);
      for (int b = 1; b <= NB; b++) blocks.push_back(b);
    } else if (p.iout >= 16 && p.iout <= 29) {
      blocks.push_back(p.iout - 15);
    } else if (p.iout == 0) {
      blocks.push_back(0);
    }
    for (int b : blocks) {
      std::vector<double> up(nlev, 0.), dn(nlev, 0.), di(nlev, 0.), df(nlev, 0.), nt(nlev, 0.), hr(nlev, 0.);
      for (int s = 0; s < nmca; s++) {
        // sample s, block b: (sample, level) fastest in the library's (nsamples, nlay+1, 15)
        auto at = [&](const std::vector<double> &x, int i) { return x[s + (size_t)nmca * (i + nlev * b)]; };
        for (int i = 0; i <= L; i++) {
          double u = at(su, i), d = at(sd, i), r = at(sdir, i);
          up[i] += u, dn[i] += d, di[i] += r, df[i] += d - r, nt[i] += d - u;
        }
        for (int i = 0; i < L; i++) {
          double n1 = at(sd, i + 1) - at(su, i + 1), n0 = at(sd, i) - at(su, i);
          hr[i] += (n1 - n0) * (heatfac / (p.pz[i] - p.pz[i + 1]));
        }
      }
      for (size_t i = 0; i < nlev; i++) {
        up[i] /= nmca, dn[i] /= nmca, di[i] /= nmca, df[i] /= nmca, nt[i] /= nmca, hr[i] /= nmca;
      }
      const double w1 = b == 0 ? WAVENUM1[NB - 1] : WAVENUM1[b - 1];
      const double w2 = b == 0 ? WAVENUM2[NB - 2] : WAVENUM2[b - 1];
      write_rows(f, p, w1, w2, up.data(), df.data(), di.data(), dn.data(), nt.data(), hr.data());
    }
    std::fprintf(f, "  Computed by rrtmgsw_column (librrtmgsw_hip.so, MI355X): column program, McICA, %d samples.\n", nmca);
    std::fclose(f);
    return 0;
  }
  int rc = rrtmgsw_column_sw(1, L, p.icld, p.idelm, p.pavel.data(), p.tavel.data(), p.pz.data(),
                             p.coldry.data(), p.wkl.data(), alb.data(), &cosz, 1.0, p.juldat, p.scon,
                             p.isolvar, p.bndsolvar, p.indsolvar, &p.solcycfrac, cl.inflag,
                             cl.iceflag, cl.liqflag, c ? cl.cldfrac.data() : nullptr,
                             c ? cl.ciwp.data() : nullptr, c ? cl.clwp.data() : nullptr,
                             c ? cl.rei.data() : nullptr, c ? cl.rel.data() : nullptr,
                             c ? cl.tauc.data() : nullptr, c ? cl.ssac.data() : nullptr,
                             c ? cl.asmc.data() : nullptr, c ? cl.fsfc.data() : nullptr,
                             p.iaer == 10 ? tauaer.data() : nullptr, p.iaer == 10 ? ssaaer.data() : nullptr,
                             p.iaer == 10 ? asmaer.data() : nullptr, fup.data(), fdn.data(), fdir.data());
  if (rc) die(rrtmgsw_last_error());
  FILE *f = std::fopen((dir + "OUTPUT_RRTM").c_str(), "w");
  if (!f) die("cannot write OUTPUT_RRTM");
  // blocks: iout = 0 broadband; 16..29 that band; 98 broadband then every band (:560-575, :950-960)
  std::vector<int> blocks;
  if (p.iout == 98) {
    blocks.push_back(0);
    for (int b = 1; b <= NB; b++) blocks.push_back(b);
  } else if (p.iout >= 16 && p.iout <= 29) {
    blocks.push_back(p.iout - 15);
  } else if (p.iout == 0) {
    blocks.push_back(0);
  }
  for (int b : blocks) {
    const double w1 = b == 0 ? WAVENUM1[NB - 1] : WAVENUM1[b - 1];
    const double w2 = b == 0 ? WAVENUM2[NB - 2] : WAVENUM2[b - 1];
    write_block(f, p, w1, w2, fup.data() + nlev * b, fdn.data() + nlev * b, fdir.data() + nlev * b, heatfac);
  }
  std::fprintf(f, "  Computed by rrtmgsw_column (librrtmgsw_hip.so, MI355X): column program without McICA.\n");
  std::fclose(f);
  return 0;
}
