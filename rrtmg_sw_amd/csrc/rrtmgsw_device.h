// rrtmgsw_device.h -- device building blocks shared by the forward and adjoint kernels of the
// GCM-mode McICA shortwave path (gfx950): per-column LDS staging, inatm_sw + setcoef_sw, the
// gas-optics template over per-g-point records (taumol_sw), cldprmc_sw, exponential-table
// transmittances, reftra_sw, the adding-method elimination step (vrtqdr_sw), wavefront
// reductions, the FP32 pieces of the mixed form, and the McICA sub-column mask generators
// (KISS, Mersenne Twister).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/rrtmgsw.h"
#include "rrtmgsw_internal.h"

namespace rrtmgsw {

constexpr double BPADE = 1.0 / PADE;
constexpr double TBLINT = 10000.0;
constexpr double ONEMINUS = 1.0 - 1.e-06;   // src/rrtmg_sw_rad.f90:563-565
constexpr double REPCLC = 1.e-12;           // src/rrtmg_sw_spcvmc.f90:295
constexpr double CLDMIN = 1.e-20;           // src/rrtmg_sw_cldprmc.f90:118

// ------------------------------------------------------------------------------------------
// Per-column LDS staging. Offsets are computed for the call's nlay (dynamic shared memory).
// ------------------------------------------------------------------------------------------
struct ColLayout {
  int L;
  int o_ix, o_misc;   // packed setcoef indices (L words), misc (32 ints)
  int o_fac, o_col, o_red, o_reda, o_cf, o_mask, total;   // double blocks / mask words
  // fluxes = false leaves out the per-level flux sums (red, reda) of the forward kernels; mixed
  // holds them in FP32, with the layers' absorbed fluxes (red: 6 a level, reda: 4 a layer)
  // stage = false leaves out the staged cloud fractions (cf, al) of the mask drawing
  __host__ __device__ ColLayout(int nlay, int nmaskw, bool fluxes = true, bool mixed = false,
                                bool stage = true) {
    L = nlay;
    int ib = 0;
    o_ix = ib, ib += L;
    o_misc = ib, ib += 32;
    int bytes = ((ib * 4 + 15) / 16) * 16;
    o_fac = bytes, bytes += 8 * L * 8;     // fac00 fac01 fac10 fac11 selffac selffrac forfac forfrac
    o_col = bytes, bytes += 6 * L * 8;     // h2o co2 o3 ch4 o2 colmol
    o_red = bytes, bytes += !fluxes ? 0 : mixed ? 2 * (L + 1) * 6 * 4 : 2 * (L + 1) * 4 * 8;
    o_reda = bytes, bytes += !fluxes ? 0 : mixed ? 2 * L * 4 * 4 : 2 * L * 2 * 8;
    bytes = ((bytes + 15) / 16) * 16;
    o_cf = bytes, bytes += nmaskw > 0 && stage ? 2 * L * 8 : 0;   // masks drawn here: cldfrac, alpha
    o_mask = bytes, bytes += nmaskw * NGPT * 4;
    total = bytes;
  }
};

struct ColView {
  int L;
  uint32_t *ix;   // a layer's jp | jt << 8 | jt1 << 12 | indself << 16 | indfor << 24
  int *misc;
  double *fac, *col, *red, *reda;
  double *cf, *al;       // the column's cloud fractions and overlap parameters (mask drawing)
  float *redf, *redaf;   // the mixed form's FP32 sums, in place of red / reda
  uint32_t *mask;
  __device__ ColView(char *smem, const ColLayout &lo) {
    L = lo.L;
    int *ip = reinterpret_cast<int *>(smem);
    ix = reinterpret_cast<uint32_t *>(ip + lo.o_ix), misc = ip + lo.o_misc;
    fac = reinterpret_cast<double *>(smem + lo.o_fac);
    col = reinterpret_cast<double *>(smem + lo.o_col);
    red = reinterpret_cast<double *>(smem + lo.o_red);
    reda = reinterpret_cast<double *>(smem + lo.o_reda);
    redf = reinterpret_cast<float *>(smem + lo.o_red);
    redaf = reinterpret_cast<float *>(smem + lo.o_reda);
    cf = reinterpret_cast<double *>(smem + lo.o_cf), al = cf + L;
    mask = reinterpret_cast<uint32_t *>(smem + lo.o_mask);
  }
  // layer-major: a layer's eight fractions (its six columns) are adjacent, so the gas optics
  // reads them from one address with immediate offsets
  __device__ double &f(int k, int lay) { return fac[lay * 8 + k]; }   // 0..7 as listed above
  __device__ double &c(int s, int lay) { return col[lay * 6 + s]; }   // 0..4 species, 5 colmol
  // the packed indices: 0 jp, 1 jt, 2 jt1, 3 indself, 4 indfor (all small and non-negative)
  __device__ static int idx(uint32_t w, int k) {
    return k == 0 ? w & 0xff : k == 1 ? (w >> 8) & 0xf : k == 2 ? (w >> 12) & 0xf : k == 3 ? (w >> 16) & 0xff : w >> 24;
  }
  __device__ int jp(int lay) const { return ix[lay] & 0xff; }
};
// misc slots
constexpr int M_LAYTROP = 0, M_LAYSOL = 1;   // laysolfr of band b at misc[M_LAYSOL + b]
constexpr int M_PARTIAL = 15;                 // non-McICA: a layer has a partial cloud fraction
constexpr int M_LAYLOW = 16;                  // setcoef_sw's laylow (layers with log p >= 6.62)
constexpr int M_CLDTOP = 17;                  // highest layer with cloud fraction >= cldmin, or -1

// ------------------------------------------------------------------------------------------
// inatm_sw + setcoef_sw for one layer (src/rrtmg_sw_rad.f90:1531-1586,
// src/rrtmg_sw_setcoef.f90:163-303). Contraction is off so that the interpolation indices
// round exactly as the reference's separate multiply and add do.
// ------------------------------------------------------------------------------------------
__device__ static const double PREFLOG[59] = {
    6.9600e+00, 6.7600e+00, 6.5600e+00, 6.3600e+00, 6.1600e+00, 5.9600e+00, 5.7600e+00,
    5.5600e+00, 5.3600e+00, 5.1600e+00, 4.9600e+00, 4.7600e+00, 4.5600e+00, 4.3600e+00,
    4.1600e+00, 3.9600e+00, 3.7600e+00, 3.5600e+00, 3.3600e+00, 3.1600e+00, 2.9600e+00,
    2.7600e+00, 2.5600e+00, 2.3600e+00, 2.1600e+00, 1.9600e+00, 1.7600e+00, 1.5600e+00,
    1.3600e+00, 1.1600e+00, 9.6000e-01, 7.6000e-01, 5.6000e-01, 3.6000e-01, 1.6000e-01,
    -4.0000e-02, -2.4000e-01, -4.4000e-01, -6.4000e-01, -8.4000e-01, -1.0400e+00, -1.2400e+00,
    -1.4400e+00, -1.6400e+00, -1.8400e+00, -2.0400e+00, -2.2400e+00, -2.4400e+00, -2.6400e+00,
    -2.8400e+00, -3.0400e+00, -3.2400e+00, -3.4400e+00, -3.6400e+00, -3.8400e+00, -4.0400e+00,
    -4.2400e+00, -4.4400e+00, -4.6400e+00};
__device__ static const double TREF[59] = {
    2.9420e+02, 2.8799e+02, 2.7894e+02, 2.6925e+02, 2.5983e+02, 2.5017e+02, 2.4077e+02,
    2.3179e+02, 2.2306e+02, 2.1578e+02, 2.1570e+02, 2.1570e+02, 2.1570e+02, 2.1706e+02,
    2.1858e+02, 2.2018e+02, 2.2174e+02, 2.2328e+02, 2.2479e+02, 2.2655e+02, 2.2834e+02,
    2.3113e+02, 2.3401e+02, 2.3703e+02, 2.4022e+02, 2.4371e+02, 2.4726e+02, 2.5085e+02,
    2.5457e+02, 2.5832e+02, 2.6216e+02, 2.6606e+02, 2.6999e+02, 2.7340e+02, 2.7536e+02,
    2.7568e+02, 2.7372e+02, 2.7163e+02, 2.6955e+02, 2.6593e+02, 2.6211e+02, 2.5828e+02,
    2.5360e+02, 2.4854e+02, 2.4348e+02, 2.3809e+02, 2.3206e+02, 2.2603e+02, 2.2000e+02,
    2.1435e+02, 2.0887e+02, 2.0340e+02, 1.9792e+02, 1.9290e+02, 1.8809e+02, 1.8329e+02,
    1.7849e+02, 1.7394e+02, 1.7212e+02};

// jp and the troposphere / lowest-layer tests are taken from the pressure thresholds of the host
// libm's log (DeviceTables::pthr): the device log estimates jp, the thresholds correct it, so the
// indices are the reference's bit for bit even where p lies within an ulp of a switch. The
// fraction fp uses the device log (within an ulp or two of the host's).
__device__ inline int jp_index(double p, double plog, const double *__restrict__ pthr) {
#pragma clang fp contract(off)
  int jp = (int)(36. - 5 * (plog + 0.04));
  jp = jp < 1 ? 1 : (jp > 58 ? 58 : jp);
  // the reference's jp is 1 + #{k = 2..58 : p < pthr[k - 2]} (the thresholds decrease with k)
  while (jp < 58 && p < pthr[PTHR_JP + jp - 1]) jp++;
  while (jp > 1 && !(p < pthr[PTHR_JP + jp - 2])) jp--;
  return jp;
}
__device__ static void setcoef_layer(const RtArgs &a, const DeviceTables &T, int c, int lay, ColView &v,
                                     int *ntrop) {
#pragma clang fp contract(off)
  const int C = a.ncol;
  const double amd = 28.9660, amw = 18.0160, avogad = 6.02214199e+23, grav = 9.8066;
  double p = a.play[c + (size_t)C * lay], t = a.tlay[c + (size_t)C * lay];
  double pz0 = a.plev[c + (size_t)C * lay], pz1 = a.plev[c + (size_t)C * (lay + 1)];
  double coldry, w_h2o, w_co2, w_o3, w_ch4, w_o2;
  if (a.coldry_in) {
    // column program: dry-air and molecular amounts (molecules/cm2) given per layer
    // (src/rrtmg_sw.1col.f90:1733-1752); wkl rows H2O, CO2, O3, N2O, CO, CH4, O2
    const size_t q7 = (size_t)7 * (c + (size_t)C * lay);
    coldry = a.coldry_in[c + (size_t)C * lay];
    w_h2o = a.wkl_in[q7 + 0], w_co2 = a.wkl_in[q7 + 1], w_o3 = a.wkl_in[q7 + 2];
    w_ch4 = a.wkl_in[q7 + 5], w_o2 = a.wkl_in[q7 + 6];
  } else {
    double q = a.h2ovmr[c + (size_t)C * lay];
    double amm = (1. - q) * amd + q * amw;
    coldry = (pz0 - pz1) * 1.e3 * avogad / (1.e2 * grav * amm * (1. + q));
    w_h2o = coldry * q, w_co2 = coldry * a.co2vmr[c + (size_t)C * lay];
    w_o3 = coldry * a.o3vmr[c + (size_t)C * lay], w_ch4 = coldry * a.ch4vmr[c + (size_t)C * lay];
    w_o2 = coldry * a.o2vmr[c + (size_t)C * lay];
  }

  double plog = log(p);
  int jp = jp_index(p, plog, T.pthr);
  int jp1 = jp + 1;
  double fp = 5. * (PREFLOG[jp - 1] - plog);
  int jt = (int)(3. + (t - TREF[jp - 1]) / 15.);
  jt = jt < 1 ? 1 : (jt > 4 ? 4 : jt);
  double ft = ((t - TREF[jp - 1]) / 15.) - (double)(jt - 3);
  int jt1 = (int)(3. + (t - TREF[jp1 - 1]) / 15.);
  jt1 = jt1 < 1 ? 1 : (jt1 > 4 ? 4 : jt1);
  double ft1 = ((t - TREF[jp1 - 1]) / 15.) - (double)(jt1 - 3);
  double water = w_h2o / coldry;
  double scalefac = p * (296. / 1013.) / t;
  double forfac = scalefac / (1. + water), forfrac, selffac, selffrac;
  int indfor, indself;
  if (p >= T.pthr[PTHR_TROP]) {   // log p > 4.56
    atomicAdd(ntrop, 1);
    if (p >= T.pthr[PTHR_LOW]) atomicAdd(&v.misc[M_LAYLOW], 1);   // log p >= 6.62
    double factor = (332.0 - t) / 36.0;
    indfor = (int)factor;
    indfor = indfor < 1 ? 1 : (indfor > 2 ? 2 : indfor);
    forfrac = factor - (double)indfor;
    selffac = water * forfac;
    factor = (t - 188.0) / 7.2;
    indself = (int)factor - 7;
    indself = indself < 1 ? 1 : (indself > 9 ? 9 : indself);
    selffrac = factor - (double)(indself + 7);
  } else {
    double factor = (t - 188.0) / 36.0;
    indfor = 3;
    forfrac = factor - 1.0;
    selffac = 0.;
    selffrac = 0.;
    indself = 0;
  }
  double colh2o = 1.e-20 * w_h2o, colco2 = 1.e-20 * w_co2, colo3 = 1.e-20 * w_o3;
  double colch4 = 1.e-20 * w_ch4, colo2 = 1.e-20 * w_o2;
  double colmol = 1.e-20 * coldry + colh2o;
  if (colco2 == 0.) colco2 = 1.e-32 * coldry;
  if (colch4 == 0.) colch4 = 1.e-32 * coldry;
  if (colo2 == 0.) colo2 = 1.e-32 * coldry;
  double compfp = 1. - fp;
  v.ix[lay] = (uint32_t)jp | (uint32_t)jt << 8 | (uint32_t)jt1 << 12 | (uint32_t)indself << 16 |
              (uint32_t)indfor << 24;
  v.f(0, lay) = compfp * (1. - ft);   // fac00
  v.f(1, lay) = fp * (1. - ft1);      // fac01
  v.f(2, lay) = compfp * ft;          // fac10
  v.f(3, lay) = fp * ft1;             // fac11
  v.f(4, lay) = selffac, v.f(5, lay) = selffrac, v.f(6, lay) = forfac, v.f(7, lay) = forfrac;
  v.c(SP_H2O, lay) = colh2o, v.c(SP_CO2, lay) = colco2, v.c(SP_O3, lay) = colo3;
  v.c(SP_CH4, lay) = colch4, v.c(SP_O2, lay) = colo2, v.c(5, lay) = colmol;
}

// laysolfr of one band: the layer whose species ratio sets the band's solar source
// (e.g. src/rrtmg_sw_taumol.f90:351-356 upper loop, :559-564 lower loop). 1-based.
__device__ static int find_laysolfr(const GPoint &g, const ColView &v, int laytrop, int L) {
  int ref = g.layreffr;
  if (!g.sol_upper) {
    int ls = laytrop;
    for (int lay = 1; lay <= laytrop; lay++)
      if (lay + 1 <= L && v.jp(lay - 1) < ref && v.jp(lay) >= ref) ls = min(lay + 1, laytrop);
    return ls < 1 ? 1 : ls;
  }
  int ls = L;
  for (int lay = laytrop + 1; lay <= L; lay++)
    if (lay >= 2 && v.jp(lay - 2) < ref && v.jp(lay - 1) >= ref) ls = lay;
  return ls;
}

// ------------------------------------------------------------------------------------------
// Gas optics template: taug and taur of one g-point at one layer.
// ------------------------------------------------------------------------------------------
template <typename R>
struct SpecT {
  R speccomb, fs;
  int js;
};
using Spec = SpecT<double>;
// Binary species parameter (e.g. src/rrtmg_sw_taumol.f90:287-292): the species ratio, clipped
// at oneminus, its bin js and the fraction fs (mod(specmult, 1) = specmult - trunc(specmult)
// exactly for the non-negative values that occur).
template <typename R>
__device__ inline SpecT<R> key_spec_t(R c1, R c2, R strrat, R mult, bool bin) {
#pragma clang fp contract(off)
  SpecT<R> s;
  R speccomb = c1 + strrat * c2;
  R specparm = c1 / speccomb;
  if (specparm >= R(ONEMINUS)) specparm = R(ONEMINUS);
  R specmult = mult * specparm;
  int jsb = 1 + (int)specmult;
  R fsb = specmult - trunc(specmult);
  s.speccomb = speccomb;
  s.js = bin ? jsb : 1;
  s.fs = bin ? fsb : R(0);
  return s;
}
__device__ inline Spec key_spec(double c1, double c2, double strrat, double mult, bool bin) {
  return key_spec_t<double>(c1, c2, strrat, mult, bin);
}

// Solar source of the lane's g-point at its band's laysolfr (src/rrtmg_sw_taumol.f90, e.g.
// :366-375 plain, :495-520 interpolated in the binary species parameter), times
// adjflux * cos(sza): zincflx of src/rrtmg_sw_spcvmc.f90:349-356.
__device__ inline double incident_flux(const GPoint &g, const double *__restrict__ tab, ColView &v,
                                       const SwScalars &sc, double mu0) {
  int ls = v.misc[M_LAYSOL + g.ibnd] - 1;  // 0-based layer
  int js = 1;
  double fs = 0.0;
  if (g.sol_interp) {
    Spec s = g.sol_upper ? key_spec(v.c(g.hi_s1, ls), v.c(g.hi_s2, ls), g.hi_strrat, 4.0, true)
                         : key_spec(v.c(g.lo_s1, ls), v.c(g.lo_s2, ls), g.lo_strrat, 8.0, true);
    js = s.js, fs = s.fs;
  }
  auto interp = [&](int off) {
    double t0 = tab[off + (js - 1) * g.ssf], t1 = tab[off + js * g.ssf];
    return t0 + fs * (t1 - t0);
  };
  double src;
  if (sc.isolvar < 0) {
    src = g.sol_scale * interp(g.sflux);
  } else {
    double fb = interp(g.facb), sn = interp(g.snsp), ir = interp(g.irrad);
    if (sc.isolvar == 3)
      src = sc.svar_bnd[0][g.ibnd] * fb + sc.svar_bnd[1][g.ibnd] * sn + sc.svar_bnd[2][g.ibnd] * ir;
    else
      src = sc.svar_f * fb + sc.svar_s * sn + sc.svar_i * ir;
  }
  return sc.adjflux[g.ibnd] * src * mu0;
}

// ------------------------------------------------------------------------------------------
// cldprmc_sw for one (g, layer): src/rrtmg_sw_cldprmc.f90:141-347. Returns the delta-scaled
// cloud optical depth, single-scattering albedo and asymmetry; raises *err on a STOP condition.
// ------------------------------------------------------------------------------------------
struct CloudOpt {
  double tau, ssa, asm_;
  double orig;   // optical depth before delta scaling (taormc / taucldorig)
};
__device__ inline double tab2(const double *tab, int off, int n, int i, int band) {
  return tab[off + (i - 1) + n * (band - 16)];
}
// Ice and liquid single-particle optics of one band at the layer's particle sizes
// (src/rrtmg_sw_cldprmc.f90:176-295): extinction per unit water path, single-scattering albedo,
// asymmetry and forward-scattering fraction; sets err on the reference's STOP conditions, the
// first one in the reference's order of tests (radius, then the derived properties; ice before
// liquid) unless err is already set.
struct CloudProps {
  double extcoice, ssacoice, gice, forwice, extcoliq, ssacoliq, gliq, forwliq;
};
__device__ static CloudProps cloud_props(const DeviceTables &T, const SwScalars &sc, int band,
                                         double ciwp, double clwp, double rei, double rel, int &err) {
  CloudProps p = {0, 0, 0, 0, 0, 0, 0, 0};
  const double *tab = T.tab;
  auto stop = [&](int code, int which) {
    if (!err) err = cloud_stop(code, which);
  };
  if (ciwp != 0.0) {
    double radice = rei;
    if (sc.iceflag == 1) {  // :182-208
      if (radice < 13.0 || radice > 130.) stop(RRTMGSW_ERR_RADIUS, CS_ICE_RADIUS);
      const double w2[14] = {3250., 4000., 4650., 5150., 6150., 7700., 8050.,
                             12850., 16000., 22650., 29000., 38000., 50000., 2600.};
      double wn = w2[band - 16];
      int icx = wn > 1.43e04 ? 1 : wn > 7.7e03 ? 2 : wn > 5.3e03 ? 3 : wn > 4.0e03 ? 4 : 5;
      p.extcoice = tab[T.abari + icx - 1] + tab[T.bbari + icx - 1] / radice;
      p.ssacoice = 1. - tab[T.cbari + icx - 1] - tab[T.dbari + icx - 1] * radice;
      p.gice = tab[T.ebari + icx - 1] + tab[T.fbari + icx - 1] * radice;
      if (p.gice >= 1.) p.gice = 1. - 1.e-06;
      p.forwice = p.gice * p.gice;
    } else if (sc.iceflag == 2) {  // :212-231
      if (radice < 5.0 || radice > 131.0) stop(RRTMGSW_ERR_RADIUS, CS_ICE_RADIUS);
      double factor = (radice - 2.) / 3.;
      int index = (int)factor;
      if (index == 43) index = 42;
      index = index < 1 ? 1 : (index > 42 ? 42 : index);
      double fint = factor - (double)index;
      double e0 = tab2(tab, T.extice2, 43, index, band), e1 = tab2(tab, T.extice2, 43, index + 1, band);
      double s0 = tab2(tab, T.ssaice2, 43, index, band), s1 = tab2(tab, T.ssaice2, 43, index + 1, band);
      double g0 = tab2(tab, T.asyice2, 43, index, band), g1 = tab2(tab, T.asyice2, 43, index + 1, band);
      p.extcoice = e0 + fint * (e1 - e0);
      p.ssacoice = s0 + fint * (s1 - s0);
      p.gice = g0 + fint * (g1 - g0);
      p.forwice = p.gice * p.gice;
    } else if (sc.iceflag == 3) {  // :235-260
      if (radice < 5.0 || radice > 140.0) stop(RRTMGSW_ERR_RADIUS, CS_ICE_GEN_SIZE);
      double factor = (radice - 2.) / 3.;
      int index = (int)factor;
      if (index == 46) index = 45;
      index = index < 1 ? 1 : (index > 45 ? 45 : index);
      double fint = factor - (double)index;
      double e0 = tab2(tab, T.extice3, 46, index, band), e1 = tab2(tab, T.extice3, 46, index + 1, band);
      double s0 = tab2(tab, T.ssaice3, 46, index, band), s1 = tab2(tab, T.ssaice3, 46, index + 1, band);
      double g0 = tab2(tab, T.asyice3, 46, index, band), g1 = tab2(tab, T.asyice3, 46, index + 1, band);
      double d0 = tab2(tab, T.fdlice3, 46, index, band), d1 = tab2(tab, T.fdlice3, 46, index + 1, band);
      p.extcoice = e0 + fint * (e1 - e0);
      p.ssacoice = s0 + fint * (s1 - s0);
      p.gice = g0 + fint * (g1 - g0);
      double fdelta = d0 + fint * (d1 - d0);
      if (fdelta < 0.0) stop(RRTMGSW_ERR_OPTICS, CS_FDELTA_LO);
      if (fdelta > 1.0) stop(RRTMGSW_ERR_OPTICS, CS_FDELTA_HI);
      p.forwice = fdelta + 0.5 / p.ssacoice;
      if (p.forwice > p.gice) p.forwice = p.gice;
    }
    if (p.extcoice < 0.0) stop(RRTMGSW_ERR_OPTICS, CS_ICE_EXT_LO);
    if (p.ssacoice > 1.0) stop(RRTMGSW_ERR_OPTICS, CS_ICE_SSA_HI);
    if (p.ssacoice < 0.0) stop(RRTMGSW_ERR_OPTICS, CS_ICE_SSA_LO);
    if (p.gice > 1.0) stop(RRTMGSW_ERR_OPTICS, CS_ICE_ASY_HI);
    if (p.gice < 0.0) stop(RRTMGSW_ERR_OPTICS, CS_ICE_ASY_LO);
  }
  if (clwp != 0.0 && sc.liqflag == 1) {  // :265-295
    double radliq = rel;
    if (radliq < 2.5 || radliq > 60.) stop(RRTMGSW_ERR_RADIUS, CS_LIQ_RADIUS);
    int index = (int)(radliq - 1.5);
    if (index == 0) index = 1;
    if (index == 58) index = 57;
    index = index < 1 ? 1 : (index > 57 ? 57 : index);
    double fint = radliq - 1.5 - (double)index;
    double e0 = tab2(tab, T.extliq1, 58, index, band), e1 = tab2(tab, T.extliq1, 58, index + 1, band);
    double s0 = tab2(tab, T.ssaliq1, 58, index, band), s1 = tab2(tab, T.ssaliq1, 58, index + 1, band);
    double g0 = tab2(tab, T.asyliq1, 58, index, band), g1 = tab2(tab, T.asyliq1, 58, index + 1, band);
    p.extcoliq = e0 + fint * (e1 - e0);
    p.ssacoliq = s0 + fint * (s1 - s0);
    if (fint < 0. && p.ssacoliq > 1.) p.ssacoliq = s0;
    p.gliq = g0 + fint * (g1 - g0);
    p.forwliq = p.gliq * p.gliq;
    if (p.extcoliq < 0.0) stop(RRTMGSW_ERR_OPTICS, CS_LIQ_EXT_LO);
    if (p.ssacoliq > 1.0) stop(RRTMGSW_ERR_OPTICS, CS_LIQ_SSA_HI);
    if (p.ssacoliq < 0.0) stop(RRTMGSW_ERR_OPTICS, CS_LIQ_SSA_LO);
    if (p.gliq > 1.0) stop(RRTMGSW_ERR_OPTICS, CS_LIQ_ASY_HI);
    if (p.gliq < 0.0) stop(RRTMGSW_ERR_OPTICS, CS_LIQ_ASY_LO);
  }
  return p;
}

// The optics of a cloudy cell once the cloud test has passed, inflag 0 (optical properties
// given, delta-M scaled here) or 2 (from water paths and particle sizes). cldprmc_sw and
// cldprop_sw evaluate the same expressions here (src/rrtmg_sw_cldprmc.f90:150-163, :297-337;
// src/rrtmg_sw_cldprop.f90:161-174, :316-357).
__device__ static CloudOpt cloud_optics_in(const DeviceTables &T, const SwScalars &sc, int band,
                                           double ciwp, double clwp, double rei, double rel,
                                           double tauc, double ssac, double asmc, double fsfc,
                                           int &err) {
  CloudOpt o;
  if (sc.inflag == 0) {
    double ffp = fsfc, ffp1 = 1.0 - ffp, ffpssa = 1.0 - ffp * ssac;
    o.orig = tauc;
    o.ssa = ffp1 * ssac / ffpssa;
    o.tau = ffpssa * tauc;
    o.asm_ = (asmc - ffp) / ffp1;
    return o;
  }
  CloudProps p = cloud_props(T, sc, band, ciwp, clwp, rei, rel, err);
  const double extcoice = p.extcoice, ssacoice = p.ssacoice, gice = p.gice, forwice = p.forwice;
  const double extcoliq = p.extcoliq, ssacoliq = p.ssacoliq, gliq = p.gliq, forwliq = p.forwliq;
  // :297-337
  double tauliqorig = clwp * extcoliq, tauiceorig = ciwp * extcoice;
  o.orig = tauliqorig + tauiceorig;
  double ssaliq = ssacoliq * (1. - forwliq) / (1. - forwliq * ssacoliq);
  double tauliq = (1. - forwliq * ssacoliq) * tauliqorig;
  double ssaice = ssacoice * (1. - forwice) / (1. - forwice * ssacoice);
  double tauice = (1. - forwice * ssacoice) * tauiceorig;
  double scatliq = ssaliq * tauliq, scatice = ssaice * tauice;
  double tau = tauliq + tauice;
  if (tau == 0.) tau = CLDMIN;
  if (scatice == 0.) scatice = CLDMIN;
  o.tau = tau;
  o.ssa = (scatliq + scatice) / tau;
  if (sc.iceflag == 3)
    o.asm_ = (1.0 / (scatliq + scatice)) *
             (scatliq * (gliq - forwliq) / (1.0 - forwliq) + scatice * ((gice - forwice) / (1.0 - forwice)));
  else
    o.asm_ = (scatliq * (gliq - forwliq) / (1.0 - forwliq) + scatice * (gice - forwice) / (1.0 - forwice)) /
             (scatliq + scatice);
  return o;
}

// cldprmc_sw for one (g, layer) (src/rrtmg_sw_cldprmc.f90:141-347): a cell failing the cloud test
// keeps its McICA inputs; a cell passing it with inflag = 1 is the reference's STOP (a call whose
// cells all fail the test runs); another inflag than 0, 1 and 2 keeps the inputs.
__device__ static CloudOpt cloud_optics(const DeviceTables &T, const SwScalars &sc, int band,
                                        double cldf, double ciwp, double clwp, double rei,
                                        double rel, double tauc, double ssac, double asmc,
                                        double fsfc, int &err) {
  const CloudOpt keep = {tauc, ssac, asmc, tauc};
  double cwp = ciwp + clwp;
  if (!(cldf >= CLDMIN && (cwp >= CLDMIN || tauc >= CLDMIN))) return keep;
  if (sc.inflag == 1) err = RRTMGSW_ERR_INFLAG;
  if (sc.inflag != 0 && sc.inflag != 2) return keep;
  return cloud_optics_in(T, sc, band, ciwp, clwp, rei, rel, tauc, ssac, asmc, fsfc, err);
}

// cldprop_sw for one (layer, band) of the non-McICA path (src/rrtmg_sw_cldprop.f90:141-362): the
// cloud test uses the layer's band-summed tauc (tauctot); a layer failing it, or an inflag other
// than 0 and 2, gives the clear defaults tau = 0, ssa = 1, g = 0.
__device__ static CloudOpt cloud_optics_layer(const DeviceTables &T, const SwScalars &sc, int band,
                                              double cldf, double ciwp, double clwp, double rei,
                                              double rel, double tauc, double ssac, double asmc,
                                              double fsfc, double tauctot, int &err) {
  const CloudOpt none = {0.0, 1.0, 0.0, tauc};
  double cwp = ciwp + clwp;
  if (!(cldf >= CLDMIN && (cwp >= CLDMIN || tauctot >= CLDMIN))) return none;
  if (sc.inflag != 0 && sc.inflag != 2) return none;
  return cloud_optics_in(T, sc, band, ciwp, clwp, rei, rel, tauc, ssac, asmc, fsfc, err);
}

// Adjoint of cloud_optics (inflag = 2) with respect to the in-cloud water paths, given the
// adjoints of the returned (tau, ssa, asm). tau, scatliq and scatice are linear in the water
// paths; ssa = (scatliq + scatice) / tau and asm is the scattering-weighted mean of the two
// delta-scaled asymmetries.
__device__ static void cloud_optics_adj(const DeviceTables &T, const SwScalars &sc, int band,
                                        double cldf, double ciwp, double clwp, double rei,
                                        double rel, double tau_b, double ssa_b, double asm_b,
                                        double &clwp_b, double &ciwp_b) {
  clwp_b = 0.0, ciwp_b = 0.0;
  double cwp = ciwp + clwp;
  if (!(cldf >= CLDMIN && cwp >= CLDMIN) || sc.inflag != 2) return;
  int err = 0;
  CloudProps p = cloud_props(T, sc, band, ciwp, clwp, rei, rel, err);
  double al = (1. - p.forwliq * p.ssacoliq) * p.extcoliq;             // d tauliq / d clwp
  double ai = (1. - p.forwice * p.ssacoice) * p.extcoice;             // d tauice / d ciwp
  double bl = p.ssacoliq * (1. - p.forwliq) / (1. - p.forwliq * p.ssacoliq) * al;   // d scatliq / d clwp
  double bi = p.ssacoice * (1. - p.forwice) / (1. - p.forwice * p.ssacoice) * ai;   // d scatice / d ciwp
  double scatliq = bl * clwp, scatice = bi * ciwp;
  double tau = al * clwp + ai * ciwp;
  if (tau == 0.) return;
  double gl = p.gliq != 0. || p.forwliq != 0. ? (p.gliq - p.forwliq) / (1.0 - p.forwliq) : 0.;
  double gi = p.gice != 0. || p.forwice != 0. ? (p.gice - p.forwice) / (1.0 - p.forwice) : 0.;
  double S = scatliq + (scatice == 0. ? CLDMIN : scatice);
  double asmv = (scatliq * gl + scatice * gi) / S;
  double S_b = ssa_b / tau;
  double t_b = tau_b - ssa_b * S / (tau * tau);
  double sl_b = S_b + asm_b * (gl - asmv) / S;
  double si_b = S_b + asm_b * (gi - asmv) / S;
  clwp_b = t_b * al + sl_b * bl;
  ciwp_b = t_b * ai + si_b * bi;
}

// FP64 reciprocal from the hardware estimate refined by two Newton-Raphson steps (an IEEE
// division adds operand rescaling for denormal and overflowing operands, which the
// denominators it replaces never are, and a final correction: 11 instructions instead of 5).
__device__ inline double frcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
// One Newton-Raphson step (relative error the square of the estimate's, below 1e-13): for the
// forward forms' optics and adding-method denominators, whose reciprocals only scale fluxes;
// the reciprocals that pick a table entry or a branch keep frcp.
__device__ inline double frcp1(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// ------------------------------------------------------------------------------------------
// Layer reflectance, diffuse reflectance, transmittance and diffuse transmittance.
struct RT4 {
  double r, rd, t, td;
};

// ------------------------------------------------------------------------------------------
// Exact-exponential reftra_sw: the smooth function of (tau, w, g) that the adjoint
// differentiates and that finite differences probe.
// ------------------------------------------------------------------------------------------
// exp(y) for |y| <= 700 (the exact-exponential forms clamp their arguments at 500): Cody-Waite
// reduction y = k ln2 + r with ln2 split in a high part of 32 significant bits and a low part
// (the fdlibm constants), |r| <= ln2/2, exp(r) by its Taylor series to degree 12 (truncation
// below 2e-16 relative), scaled by 2^k. No special cases: arguments stay in range.
__device__ inline double exp_cw(double y) {
  const double LOG2E = 1.4426950408889634074;
  const double LN2_HI = 6.93147180369123816490e-01, LN2_LO = 1.90821492927058770002e-10;
  double k = rint(y * LOG2E);
  double r = fma(-k, LN2_HI, y);
  r = fma(-k, LN2_LO, r);
  double p = 1.0 / 479001600.0;          // 1/12!
  p = fma(p, r, 1.0 / 39916800.0);       // 1/11!
  p = fma(p, r, 1.0 / 3628800.0);        // 1/10!
  p = fma(p, r, 1.0 / 362880.0);         // 1/9!
  p = fma(p, r, 1.0 / 40320.0);          // 1/8!
  p = fma(p, r, 1.0 / 5040.0);           // 1/7!
  p = fma(p, r, 1.0 / 720.0);            // 1/6!
  p = fma(p, r, 1.0 / 120.0);            // 1/5!
  p = fma(p, r, 1.0 / 24.0);             // 1/4!
  p = fma(p, r, 1.0 / 6.0);              // 1/3!
  p = fma(p, r, 0.5);                    // 1/2!
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)k);
}
// reftra_sw with exp() in place of the lookup table and its expansion: the reference's branches
// (src/rrtmg_sw_reftra.f90:164-322), the general solution multiplied through by e1 as in
// reftra_e. e2 = exp(-min(tau / mu0, 500)) is also the layer's direct-beam transmittance (the
// reference evaluates the same exponential separately, src/rrtmg_sw_spcvmc.f90:558-585).
struct RTx {
  double r, rd, t, td, e2;
};
// phi(x) = (exp(x) - 1) / x and its derivative for |x| <= 0.51, by their Taylor series to degree
// 15 (truncation below 1e-18): sum x^n / (n + 1)! and sum n x^(n - 1) / (n + 1)!.
__device__ inline double phi_em1(double x) {
  double p = 1.0 / 20922789888000.0;             // 1/16!
  const double c[15] = {1.0 / 1307674368000.0, 1.0 / 87178291200.0, 1.0 / 6227020800.0, 1.0 / 479001600.0,
                        1.0 / 39916800.0, 1.0 / 3628800.0, 1.0 / 362880.0, 1.0 / 40320.0, 1.0 / 5040.0,
                        1.0 / 720.0, 1.0 / 120.0, 1.0 / 24.0, 1.0 / 6.0, 0.5, 1.0};   // 1/15! .. 1/1!
#pragma unroll
  for (int i = 0; i < 15; i++) p = fma(p, x, c[i]);
  return p;
}
__device__ inline double dphi_em1(double x) {
  // n / (n + 1)! for n = 15 .. 1
  double p = 15.0 / 20922789888000.0;
  const double c[14] = {14.0 / 1307674368000.0, 13.0 / 87178291200.0, 12.0 / 6227020800.0, 11.0 / 479001600.0,
                        10.0 / 39916800.0, 9.0 / 3628800.0, 8.0 / 362880.0, 7.0 / 40320.0, 6.0 / 5040.0,
                        5.0 / 720.0, 4.0 / 120.0, 3.0 / 24.0, 2.0 / 6.0, 1.0 / 2.0};
#pragma unroll
  for (int i = 0; i < 14; i++) p = fma(p, x, c[i]);
  return p;
}
// The exact-exponential forms evaluate the general solution in a cancellation-free form where
// 1 - k mu0 lies within this window of 0 (the removable singularity of the direct-beam particular
// solution): with eps = 1 - k mu0, 1 - (k mu0)^2 = eps (2 - eps), and the numerators of R and T
// (multiplied through by e1) are eps times smooth terms plus 2B e1 (e2 - e1) and 2C (e2 - e1),
// where e2 - e1 = -e1 z phi(-z), z = k tau eps / (1 - eps) (= tau / mu0 - k tau):
//   R = w [A + B e1^2 - 2 a2 e1 e2 - 2 B e1^2 q] / ((2 - eps) D),
//   T = e2 - w [2 a1 e1 - C e2 - E e1^2 e2 - 2 C e1 q] / ((2 - eps) D),
// A = a2 + k g3, B = a2 - k g3, C = a1 + k g4, E = a1 - k g4, q = k tau phi(-z) / (1 - eps).
// The closed form loses about 1e-16 / |1 - (k mu0)^2| of relative precision there (and the
// reference's 1e-8 guard sets R = 1e-8 where its denominator vanishes); this form keeps the
// exact forward smooth, so that finite differences and the adjoint see the same function.
constexpr double KMU_WIN = 1.e-3;

// Intermediates of one reftra_xt evaluation, kept for its reverse sweep (reftra_xt_adj). The
// conservative branch keeps za1 in a1, zgt in a2 and 1 / (1 + zgt) in rD; the near-resonance form
// (reg) keeps 1 / ((2 - eps) D) in rD, 1 / D in P, eps in eps and its numerators in NR, NT.
struct RtTape {
  double g1, g2, g3, a1, a2, k, e1, e2, D, P, rD, NR, NT, eps;
  bool cons, tiny, reg;
};
__device__ inline RTx reftra_xt(double zto1, double zw, double zg, double prmuz, double rmu, RtTape &tp) {
#pragma clang fp contract(off)
  const double zwcrit = 0.9999995, eps = 1.e-08;
  RTx o;
  double zg3 = 3. * zg;
  double g1 = (8. - zw * (5. + zg3)) * 0.25;
  double g2 = 3. * (zw * (1. - zg)) * 0.25;
  double g3 = (2. - zg3 * prmuz) * 0.25;
  double g4 = 1. - g3;
  double omg = 1. - zg;
  double den2 = omg * omg - (1. - zw) * zg * zg;
  bool cons = zg != 1. ? (zw > 0. && den2 > 0. && zw * (omg * omg) >= zwcrit * den2) : (zw >= zwcrit);
  const double e2 = exp_cw(-fmin(zto1 * rmu, 500.));
  o.e2 = e2;
  tp.g1 = g1, tp.g2 = g2, tp.g3 = g3, tp.e2 = e2, tp.cons = cons, tp.tiny = false, tp.reg = false;
  tp.k = 0., tp.e1 = 1., tp.D = 1., tp.P = 1., tp.NR = 0., tp.NT = 0., tp.eps = 0.;
  if (cons) {
    double za1 = g1 * prmuz - g3;
    double zgt = g1 * zto1;
    double inv = frcp1(1. + zgt);
    o.r = (zgt - za1 * (1. - e2)) * inv;
    o.t = 1. - o.r;
    o.rd = zgt * inv;
    o.td = 1. - o.rd;
    if (e2 == 1.0) o.r = 0., o.t = 1., o.rd = 0., o.td = 1.;
    tp.a1 = za1, tp.a2 = zgt, tp.rD = inv;
  } else {
    double a1 = g1 * g4 + g2 * g3;
    double a2 = g1 * g3 + g2 * g4;
    double k = sqrt(g1 * g1 - g2 * g2);
    double kmu = k * prmuz;
    const double x1 = k * zto1;
    double e1 = exp_cw(-fmin(x1, 500.));
    double ee = e1 * e1;
    double D = (k + g1) - (g1 - k) * ee;
    const double ke = 1. - kmu;
    tp.a1 = a1, tp.a2 = a2, tp.k = k, tp.e1 = e1, tp.D = D, tp.eps = ke;
    if (fabs(ke) < KMU_WIN && x1 < 500. && zto1 * rmu < 500.) {
      // the near-resonance form (KMU_WIN)
      const double s1 = frcp1(1. - ke);
      const double q = x1 * phi_em1(-(x1 * ke * s1)) * s1;
      const double B = a2 - k * g3, C = a1 + k * g4;
      const double NR = (a2 + k * g3) + B * ee - 2. * a2 * e1 * e2 - 2. * B * ee * q;
      const double NT = 2. * a1 * e1 - C * e2 - (a1 - k * g4) * ee * e2 - 2. * C * e1 * q;
      const double rDd = frcp1(D);
      const double rW = rDd * frcp1(2. - ke);
      o.r = zw * NR * rW;
      o.t = e2 - zw * NT * rW;
      o.rd = g2 * (1. - ee) * rDd;
      o.td = 2. * k * e1 * rDd;
      tp.P = rDd, tp.rD = rW, tp.NR = NR, tp.NT = NT, tp.reg = true;
    } else {
      double P = 1. - kmu * kmu;
      bool tiny = fabs(P * D) <= eps * e1;
      P = tiny ? 1. : P;
      double rD = frcp1(P * D);
      double NR = (1. - kmu) * (a2 + k * g3) - (1. + kmu) * (a2 - k * g3) * ee - 2. * k * (g3 - a2 * prmuz) * (e1 * e2);
      double NT = e2 * ((1. + kmu) * (a1 + k * g4) - (1. - kmu) * (a1 - k * g4) * ee) - 2. * k * (g4 + a1 * prmuz) * e1;
      o.r = tiny ? eps : zw * NR * rD;
      o.t = tiny ? e2 : e2 - zw * NT * rD;
      double PrD = P * rD;
      o.rd = g2 * (1. - ee) * PrD;
      o.td = 2. * k * e1 * PrD;
      tp.P = P, tp.rD = rD, tp.NR = NR, tp.NT = NT, tp.tiny = tiny;
    }
  }
  return o;
}
__device__ inline RTx reftra_x(double zto1, double zw, double zg, double prmuz, double rmu) {
  RtTape tp;
  return reftra_xt(zto1, zw, zg, prmuz, rmu, tp);
}

// Reverse-mode derivative of reftra_xt from its tape: given the adjoints of (r, rd, t, td, e2),
// returns the adjoints of (tau, w, g). General branch, with X = (1 + kmu)(a1 + k g4) -
// (1 - kmu)(a1 - k g4) ee, Y = g4 + a1 mu0, Z = g3 - a2 mu0 (so NT = e2 X - 2k Y e1 and
// NR = (1 - kmu)(a2 + k g3) - (1 + kmu)(a2 - k g3) ee - 2k Z e1 e2): the adjoints flow from
// (r, t) = (w NR rD, e2 - w NT rD) and (rd, td) = (g2 (1 - ee), 2k e1) P rD back through
// rD = 1 / (P D), P = 1 - kmu^2 (a constant 1 where the denominator is tiny), D, NR, NT,
// ee = e1^2, e1 = exp(-min(k tau, 500)), k = sqrt(g1^2 - g2^2), a1, a2 and the gammas; within
// KMU_WIN of the resonance, through the near-resonance form's expressions instead.
__device__ inline void reftra_xt_adj(const RtTape &tp, double zto1, double zw, double zg, double prmuz,
                                     double rmu, double rb, double rdb, double tb, double tdb, double e2b,
                                     double &taub, double &wb, double &gb) {
  const double g1 = tp.g1, g2 = tp.g2, g3 = tp.g3, g4 = 1. - g3;
  double g1b = 0., g2b = 0., g3b = 0., g4b = 0.;
  taub = 0., wb = 0., gb = 0.;
  if (tp.cons) {
    const double za1 = tp.a1, zgt = tp.a2, inv = tp.rD, ze2 = tp.e2;
    if (ze2 != 1.0) {
      // r = (zgt - za1 (1 - e2)) inv, t = 1 - r, rd = zgt inv, td = 1 - rd
      const double Reff = rb - tb, Rdeff = rdb - tdb;
      const double Rn = zgt - za1 * (1. - ze2);
      const double Rnb = Reff * inv;
      const double invb = Reff * Rn + Rdeff * zgt;
      const double zgtb = Rdeff * inv + Rnb - invb * inv * inv;
      const double za1b = -Rnb * (1. - ze2);
      e2b += Rnb * za1;
      g1b += zgtb * zto1 + za1b * prmuz;
      taub += zgtb * g1;
      g3b -= za1b;
    }
  } else {
    const double k = tp.k, e1 = tp.e1, e2 = tp.e2, a1 = tp.a1, a2 = tp.a2;
    const double ee = e1 * e1, D = tp.D;
    double kb = 0., e1b = 0., eeb = 0., a1b = 0., a2b = 0., Db = 0.;
    if (tp.reg) {
      // the near-resonance form: R = w NR rW, T = e2 - w NT rW, rW = 1 / ((2 - eps) D),
      // (rd, td) = (g2 (1 - ee), 2k e1) / D, NR = A + B ee - 2 a2 e1 e2 - 2 B ee q,
      // NT = 2 a1 e1 - C e2 - E ee e2 - 2 C e1 q, q = x phi(-z) s, z = x eps s, s = 1 / (1 - eps),
      // x = k tau, eps = 1 - k mu0
      const double ke = tp.eps, x1 = k * zto1, s1 = frcp1(1. - ke), z = x1 * ke * s1;
      const double ph = phi_em1(-z), dph = dphi_em1(-z), q = x1 * ph * s1;
      const double B = a2 - k * g3, C = a1 + k * g4, E = a1 - k * g4;
      const double rDd = tp.P, rW = tp.rD, r2e = frcp1(2. - ke);
      wb += (rb * tp.NR - tb * tp.NT) * rW;
      const double NRb = rb * zw * rW, NTb = -tb * zw * rW, rWb = zw * (rb * tp.NR - tb * tp.NT);
      e2b += tb;
      double rDdb = rdb * g2 * (1. - ee) + tdb * 2. * k * e1 + rWb * r2e;
      double epsb = rWb * rW * r2e;
      g2b += rdb * (1. - ee) * rDd;
      eeb -= rdb * g2 * rDd;
      kb += tdb * 2. * e1 * rDd;
      e1b += tdb * 2. * k * rDd;
      Db = -rDdb * rDd * rDd;
      // NR
      const double Ab = NRb, Bb0 = NRb * ee * (1. - 2. * q);
      eeb += NRb * B * (1. - 2. * q);
      a2b -= NRb * 2. * e1 * e2;
      e1b -= NRb * 2. * a2 * e2;
      e2b -= NRb * 2. * a2 * e1;
      double qb = -NRb * 2. * B * ee;
      // NT
      a1b += NTb * 2. * e1;
      e1b += NTb * 2. * (a1 - C * q);
      const double Cb = -NTb * (e2 + 2. * e1 * q), Eb = -NTb * ee * e2;
      e2b -= NTb * (C + E * ee);
      eeb -= NTb * E * e2;
      qb -= NTb * 2. * C * e1;
      // q
      const double x1b = qb * s1 * (ph - z * dph);
      epsb += qb * x1 * s1 * s1 * (ph - x1 * s1 * dph);
      kb += x1b * zto1, taub += x1b * k;
      // A, B, C, E, eps
      a2b += Ab + Bb0;
      kb += (Ab - Bb0) * g3;
      g3b += (Ab - Bb0) * k;
      a1b += Cb + Eb;
      kb += (Cb - Eb) * g4;
      g4b += (Cb - Eb) * k;
      kb -= epsb * prmuz;
    } else {
      const double kmu = k * prmuz, P = tp.P, rD = tp.rD, PrD = P * rD;
      double kmub = 0., rDb = 0., NRb = 0., NTb = 0.;
      // rd = g2 (1 - ee) P rD, td = 2k e1 P rD
      const double PrDb = rdb * g2 * (1. - ee) + tdb * 2. * k * e1;
      g2b += rdb * (1. - ee) * PrD;
      eeb -= rdb * g2 * PrD;
      kb += tdb * 2. * e1 * PrD;
      e1b += tdb * 2. * k * PrD;
      // r = w NR rD, t = e2 - w NT rD (r = eps, t = e2 where the denominator is tiny)
      if (!tp.tiny) {
        wb += (rb * tp.NR - tb * tp.NT) * rD;
        NRb = rb * zw * rD;
        NTb = -tb * zw * rD;
        rDb = zw * (rb * tp.NR - tb * tp.NT);
      }
      e2b += tb;
      rDb += PrDb * P;
      const double PDb = -rDb * rD * rD;
      Db = PDb * P;
      if (!tp.tiny) kmub -= (PrDb * rD + PDb * D) * 2. * kmu;
      // NT = e2 X - 2k Y e1
      const double Xp = a1 + k * g4, Xm = a1 - k * g4;
      const double Y = g4 + a1 * prmuz;
      e2b += NTb * ((1. + kmu) * Xp - (1. - kmu) * Xm * ee);
      const double Xb = NTb * e2, Yb = -NTb * 2. * k * e1;
      kb -= NTb * 2. * Y * e1;
      e1b -= NTb * 2. * k * Y;
      const double xs = (1. + kmu) + (1. - kmu) * ee;
      kmub += Xb * (Xp + Xm * ee);
      a1b += Xb * ((1. + kmu) - (1. - kmu) * ee) + Yb * prmuz;
      kb += Xb * g4 * xs;
      g4b += Xb * k * xs + Yb;
      eeb -= Xb * (1. - kmu) * Xm;
      // NR = (1 - kmu)(a2 + k g3) - (1 + kmu)(a2 - k g3) ee - 2k Z e1 e2
      const double Rp = a2 + k * g3, Rm = a2 - k * g3, Z = g3 - a2 * prmuz;
      const double rs = (1. - kmu) + (1. + kmu) * ee;
      const double Zb = -NRb * 2. * k * e1 * e2;
      kmub -= NRb * (Rp + Rm * ee);
      a2b += NRb * ((1. - kmu) - (1. + kmu) * ee) - Zb * prmuz;
      kb += NRb * (g3 * rs - 2. * Z * e1 * e2);
      g3b += NRb * k * rs + Zb;
      eeb -= NRb * (1. + kmu) * Rm;
      e1b -= NRb * 2. * k * Z * e2;
      e2b -= NRb * 2. * k * Z * e1;
      kb += kmub * prmuz;
    }
    // D = (k + g1) - (g1 - k) ee, ee = e1^2, e1 = exp(-min(k tau, 500))
    kb += Db * (1. + ee);
    g1b += Db * (1. - ee);
    eeb -= Db * (g1 - k);
    e1b += eeb * 2. * e1;
    const double x1b = -e1b * e1;
    if (k * zto1 < 500.) kb += x1b * zto1, taub += x1b * k;
    // k = sqrt(g1^2 - g2^2); a1 = g1 g4 + g2 g3, a2 = g1 g3 + g2 g4
    if (k > 0.) {
      const double rk = frcp1(k);
      g1b += kb * g1 * rk, g2b -= kb * g2 * rk;
    }
    g1b += a1b * g4 + a2b * g3, g3b += a1b * g2 + a2b * g1;
    g2b += a1b * g3 + a2b * g4, g4b += a1b * g1 + a2b * g2;
  }
  // e2 = exp(-min(tau / mu0, 500))
  const double x2b = -e2b * tp.e2;
  if (zto1 * rmu < 500.) taub += x2b * rmu;
  // the gammas: g4 = 1 - g3, g3 = (2 - 3 g mu0) / 4, g2 = 3 w (1 - g) / 4, g1 = (8 - w (5 + 3g)) / 4
  g3b -= g4b;
  double g3p_b = -g3b * prmuz * 0.25;      // d/d(3g) through gamma3
  wb += g2b * 3. * (1. - zg) * 0.25;
  gb -= g2b * 3. * zw * 0.25;
  wb -= g1b * (5. + 3. * zg) * 0.25;
  g3p_b -= g1b * zw * 0.25;
  gb += 3. * g3p_b;
}

// Downward half of the adding method for one sky (vrtqdr_sw :144-159), written as forward
// elimination: with D_dif(j) = (ztdn_j - tdbt_j) + prdnd_j U_j from the stack above and
// U_j = R_j tdbt_j + Rd_j D_dif(j) + Td_j U_(j+1) for layer j,
//   U_j = C_j + E_j U_(j+1),  C_j = (Rd_j a_j + R_j tdbt_j) / (1 - Rd_j prdnd_j),
//   E_j = Td_j / (1 - Rd_j prdnd_j),  a_j = ztdn_j - tdbt_j,  D_j = ztdn_j + prdnd_j U_j.
struct Sweep {
  double tdbt, ztdn, prdnd;
};
__device__ inline void sweep_layer(Sweep &s, const RT4 &l, double dbt, double &A, double &B,
                                   double &Cc, double &E) {
#pragma clang fp contract(off)
  A = s.ztdn;
  B = s.prdnd;
  double a = s.ztdn - s.tdbt;
  double zreflect = frcp1(1. - l.rd * s.prdnd);
  Cc = (l.rd * a + l.r * s.tdbt) * zreflect;
  E = l.td * zreflect;
  double ztdn = s.tdbt * l.t + (l.td * (a + s.tdbt * l.r * s.prdnd)) * zreflect;
  double prdnd = l.rd + l.td * l.td * s.prdnd * zreflect;
  s.tdbt = dbt * s.tdbt;
  s.ztdn = ztdn;
  s.prdnd = prdnd;
}

// ------------------------------------------------------------------------------------------
// Pieces of the FP64 forward kernel shaped for gfx950 issue:
//  * reciprocals from frcp;
//  * table reads at 32-bit offsets from a uniform base (one full-rate add per corner instead of
//    a quarter-rate 32-bit multiply and 64-bit address arithmetic);
//  * wavefront sums by DPP row operations, with one LDS-pipe and one VALU exchange across rows.
// ------------------------------------------------------------------------------------------

template <typename R>
__device__ inline R ldt(const R *__restrict__ base, uint32_t i) {
  return *reinterpret_cast<const R *>(reinterpret_cast<const char *>(base) +
                                      (size_t)(i * (uint32_t)sizeof(R)));
}

// Exponential-table transmittance (e.g. src/rrtmg_sw_spcvmc.f90:558-566, reftra :203-209): the
// table entry at int(1e4 ze / (bpade + ze) + 0.5) (the quotient from frcp), or 1 - ze + ze^2/2
// below od_lo.
__device__ inline double tbl_exp1(double ze, const double *__restrict__ exp_tbl) {
  double tblind = ze * frcp(BPADE + ze);
  int itind = (int)(TBLINT * tblind + 0.5);
  itind = itind < 0 ? 0 : (itind > NTBL ? NTBL : itind);
  double te = ldt(exp_tbl, (uint32_t)itind);
  return ze <= OD_LO ? 1. - ze + 0.5 * ze * ze : te;
}

// reftra_sw (PIFM, table exponentials) from e1 = exp(-k tau) and e2 = exp(-tau / mu0) alone.
// The reference's general solution (src/rrtmg_sw_reftra.f90:233-322) divides numerators holding
// 1/e1 and 1/e2 by zdenr = (1 - k^2 mu0^2)((k + g1)/e1 + (k - g1) e1); multiplied through by
// e1 (and, in T's last term, e2/e2) every term is a product of e1, e2 and the coefficients:
//   D  = (k + g1) - (g1 - k) e1^2,  rD = 1 / ((1 - k^2 mu0^2) D)
//   R  = w [(1 - k mu0)(a2 + k g3) - (1 + k mu0)(a2 - k g3) e1^2 - 2k (g3 - a2 mu0) e1 e2] rD
//   T  = e2 - w [e2 ((1 + k mu0)(a1 + k g4) - (1 - k mu0)(a1 - k g4) e1^2) - 2k (g4 + a1 mu0) e1] rD
//   Rd = g2 (1 - e1^2) (1 - k^2 mu0^2) rD,  Td = 2k e1 (1 - k^2 mu0^2) rD,
// with the reference's |zdenr| <= 1e-8 test as |(1 - k^2 mu0^2) D| <= 1e-8 e1 (R = 1e-8, T = e2,
// and then rD = 1 / D). e2 = table(min(tau / mu0, 500)) is also the layer's direct-beam
// transmittance: every table entry past index 9275 is the floor 1e-20, reached at tau/mu0 > 46.
// The FP64 part of reftra_e that decides its table lookups: the PIFM coefficients, the
// conservative-scattering decision and, for the general branch, k and e1 = exp(-k tau). The
// mixed form runs this same code, so its lookups pick the FP64 form's table entries.
struct RtFront {
  double g1, g2, g3, g4, k, kmu, e1;
  bool cons;
};
__device__ inline RtFront reftra_front(double zto1, double zw, double zg, double prmuz,
                                       const double *__restrict__ exp_tbl) {
  const double zwcrit = 0.9999995;
  RtFront f;
  double zg3 = 3. * zg;
  f.g1 = (8. - zw * (5. + zg3)) * 0.25;
  f.g2 = 3. * (zw * (1. - zg)) * 0.25;
  f.g3 = (2. - zg3 * prmuz) * 0.25;
  f.g4 = 1. - f.g3;
  double omg = 1. - zg;
  double den2 = omg * omg - (1. - zw) * zg * zg;
  f.cons = zg != 1. ? (zw > 0. && den2 > 0. && zw * (omg * omg) >= zwcrit * den2) : (zw >= zwcrit);
  f.k = 0., f.kmu = 0., f.e1 = 1.;
  if (!f.cons) {
    f.k = sqrt(f.g1 * f.g1 - f.g2 * f.g2);
    f.kmu = f.k * prmuz;
    f.e1 = tbl_exp1(fmin(f.k * zto1, 500.), exp_tbl);
  }
  return f;
}
// The rest of reftra_e in FP64, from the front.
__device__ inline RT4 reftra_back(const RtFront &f, double zto1, double zw, double prmuz, double e2) {
  const double eps = 1.e-08;
  RT4 o;
  const double g1 = f.g1, g2 = f.g2, g3 = f.g3, g4 = f.g4;
  if (f.cons) {  // conservative scattering, :188-228
    double za1 = g1 * prmuz - g3;
    double zgt = g1 * zto1;
    double inv = frcp1(1. + zgt);
    o.r = (zgt - za1 * (1. - e2)) * inv;
    o.t = 1. - o.r;
    o.rd = zgt * inv;
    o.td = 1. - o.rd;
    if (e2 == 1.0) o.r = 0., o.t = 1., o.rd = 0., o.td = 1.;
  } else {
    const double k = f.k, kmu = f.kmu, e1 = f.e1;
    double a1 = g1 * g4 + g2 * g3;
    double a2 = g1 * g3 + g2 * g4;
    double ee = e1 * e1;
    double D = (k + g1) - (g1 - k) * ee;
    double P = 1. - kmu * kmu;
    bool tiny = fabs(P * D) <= eps * e1;
    P = tiny ? 1. : P;
    double rD = frcp1(P * D);
    double NR = (1. - kmu) * (a2 + k * g3) - (1. + kmu) * (a2 - k * g3) * ee - 2. * k * (g3 - a2 * prmuz) * (e1 * e2);
    double NT = e2 * ((1. + kmu) * (a1 + k * g4) - (1. - kmu) * (a1 - k * g4) * ee) - 2. * k * (g4 + a1 * prmuz) * e1;
    o.r = tiny ? eps : zw * NR * rD;
    o.t = tiny ? e2 : e2 - zw * NT * rD;
    double PrD = P * rD;
    o.rd = g2 * (1. - ee) * PrD;
    o.td = 2. * k * e1 * PrD;
  }
  return o;
}
__device__ inline RT4 reftra_e(double zto1, double zw, double zg, double prmuz, double e2,
                               const double *__restrict__ exp_tbl) {
  return reftra_back(reftra_front(zto1, zw, zg, prmuz, exp_tbl), zto1, zw, prmuz, e2);
}

// sweep_layer with a reciprocal from frcp; A (the level's ztdn) and B (prdnd) are the state
// before the layer.
__device__ inline void sweep_r(Sweep &s, const RT4 &l, double dbt, double &Cc, double &E) {
  double a = s.ztdn - s.tdbt;
  double zreflect = frcp1(1. - l.rd * s.prdnd);
  Cc = (l.rd * a + l.r * s.tdbt) * zreflect;
  E = l.td * zreflect;
  double ztdn = s.tdbt * l.t + (l.td * (a + s.tdbt * l.r * s.prdnd)) * zreflect;
  double prdnd = l.rd + l.td * l.td * s.prdnd * zreflect;
  s.tdbt = dbt * s.tdbt;
  s.ztdn = ztdn;
  s.prdnd = prdnd;
}

// Gas optics of one regime (lower: layers below laytrop, upper: the rest) with the lane's
// constants resolved once per kernel from its g-point record: element offsets of the k-table
// corners and of the continuum, Rayleigh and species rows (the same taumol template as
// gas_optics, e.g. src/rrtmg_sw_taumol.f90:287-375 lower and :357-383 upper).
struct GasLane {
  uint32_t kb, kjs, kt, kp, ks;   // key table: base, species step, temperature, pressure, row
  int s1, s2, sm;                 // species of the key pair and the minor absorber (0..4)
  double strrat, kmult, mcoef, rayl;
  bool bin, rayl_js, cont;        // binary key; Rayleigh interpolated in js; continuum present
  uint32_t selfb, selfs, forb, fors, rab, ras;
};
// key_spec_t with the species ratio from a reciprocal (frcp or the FP32 estimate): js and fs
// can differ from the divided ratio's only where specmult lies within an ulp of an integer,
// where the interpolation is continuous across the bin edge.
template <typename R>
__device__ inline SpecT<R> key_spec_r(R c1, R c2, R strrat, R mult, bool bin) {
#pragma clang fp contract(off)
  SpecT<R> s;
  R speccomb = c1 + strrat * c2;
  R specparm = c1 * (sizeof(R) == 8 ? R(frcp(double(speccomb))) : R(__builtin_amdgcn_rcpf(float(speccomb))));
  specparm = specparm >= R(ONEMINUS) ? R(ONEMINUS) : specparm;
  R specmult = mult * specparm;
  s.speccomb = speccomb;
  s.js = bin ? 1 + (int)specmult : 1;
  s.fs = bin ? specmult - trunc(specmult) : R(0);
  return s;
}
template <bool LOWER>
__device__ inline GasLane gas_lane(const GPoint &g, int L) {
  GasLane q;
  const int st = LOWER ? g.sa : g.sb, nsp = LOWER ? g.lo_nsp : g.hi_nsp;
  q.bin = (LOWER ? g.lo_mode : g.hi_mode) == KEY_BIN;
  q.kb = LOWER ? g.absa : g.absb;
  q.ks = st;
  q.kjs = q.bin ? st : 0;
  q.kt = nsp * st;
  q.kp = 5 * nsp * st;
  q.s1 = LOWER ? g.lo_s1 : g.hi_s1;
  q.s2 = LOWER ? g.lo_s2 : g.hi_s2;
  q.sm = LOWER ? g.lo_minor : g.hi_minor;
  q.strrat = LOWER ? g.lo_strrat : g.hi_strrat;
  q.kmult = LOWER ? g.lo_mult : g.hi_mult;
  q.mcoef = LOWER ? g.lo_minor_coef : g.hi_minor_coef;
  q.rayl = LOWER ? g.rayl_lo : g.rayl_hi;
  q.rayl_js = LOWER && g.rayl_js;
  q.cont = LOWER || g.hi_for;
  q.selfb = g.selfref, q.selfs = g.sself, q.forb = g.forref, q.fors = g.sfor;
  q.rab = g.rayla, q.ras = g.sra;
  return q;
}
// A value that is the same in every lane (read from LDS at a uniform address), moved to scalar
// registers so that it occupies no vector registers while the table reads are in flight.
__device__ inline double uni(double x) {
  int lo = __builtin_amdgcn_readfirstlane(__double2loint(x));
  int hi = __builtin_amdgcn_readfirstlane(__double2hiint(x));
  return __hiloint2double(hi, lo);
}
__device__ inline float uni(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ inline int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

// The gas optics in two halves: gas_fetch computes the key-species parameter and issues every
// table read, gas_finish combines them (callers place their pending stores in between).
template <typename R>
struct GasFetch {
  R a0, a1, a2, a3, b0, b1, b2, b3, fo0, fo1, s0, s1, r0, r1;
  R fs, kmul;
};
template <bool LOWER, typename R>
__device__ inline GasFetch<R> gas_fetch(const GasLane &q, const R *__restrict__ tab, ColView &v, int lay) {
#pragma clang fp contract(off)
  GasFetch<R> f;
  const uint32_t w = uni((int)v.ix[lay]);
  const int jp = LOWER ? min(ColView::idx(w, 0), 12) : max(ColView::idx(w, 0), 13);
  const int jt = ColView::idx(w, 1), jt1 = ColView::idx(w, 2);
  const R c1 = R(v.c(q.s1, lay)), c2 = R(v.c(q.s2, lay));
  SpecT<R> s = key_spec_r<R>(c1, c2, R(q.strrat), R(LOWER ? 8 : 4), q.bin);
  f.kmul = q.bin ? s.speccomb : c1 * R(q.kmult);
  f.fs = s.fs;
  const uint32_t jpr = LOWER ? jp - 1 : jp - 13;
  const uint32_t o0 = q.kb + __umul24(jpr, q.kp) + __umul24(jt - 1, q.kt) + __umul24(s.js - 1, q.ks);
  const uint32_t o1 = q.kb + __umul24(jpr + 1, q.kp) + __umul24(jt1 - 1, q.kt) + __umul24(s.js - 1, q.ks);
  f.a0 = ldt(tab, o0), f.a1 = ldt(tab, o0 + q.kjs), f.a2 = ldt(tab, o0 + q.kt), f.a3 = ldt(tab, o0 + q.kt + q.kjs);
  f.b0 = ldt(tab, o1), f.b1 = ldt(tab, o1 + q.kjs), f.b2 = ldt(tab, o1 + q.kt), f.b3 = ldt(tab, o1 + q.kt + q.kjs);
  const int jf = LOWER ? min(ColView::idx(w, 4), 2) : min(max(ColView::idx(w, 4), 1), 3);
  f.fo0 = ldt(tab, q.forb + __umul24(jf - 1, q.fors)), f.fo1 = ldt(tab, q.forb + __umul24(jf, q.fors));
  f.s0 = R(0), f.s1 = R(0), f.r0 = R(0), f.r1 = R(0);
  if (LOWER) {
    const int is = max(ColView::idx(w, 3), 1);
    f.s0 = ldt(tab, q.selfb + __umul24(is - 1, q.selfs)), f.s1 = ldt(tab, q.selfb + __umul24(is, q.selfs));
    f.r0 = ldt(tab, q.rab + __umul24(s.js - 1, q.ras)), f.r1 = ldt(tab, q.rab + __umul24(s.js, q.ras));
  }
  return f;
}
template <bool LOWER, typename R>
__device__ inline void gas_finish(const GasLane &q, const GasFetch<R> &f, ColView &v, int lay, R &taug,
                                  R &taur) {
  // no implicit contraction: every caller (forward sweep, checkpoint recomputation) rounds alike
#pragma clang fp contract(off)
  const R fac00 = uni(R(v.f(0, lay))), fac01 = uni(R(v.f(1, lay))), fac10 = uni(R(v.f(2, lay))), fac11 = uni(R(v.f(3, lay)));
  const R colh2o = uni(R(v.c(SP_H2O, lay))), colmol = uni(R(v.c(5, lay)));
  const R fs = f.fs, f0 = R(1) - fs;
  R k = fs * fac11 * f.b3;
  k = fma(f0 * fac11, f.b2, k);
  k = fma(fs * fac01, f.b1, k);
  k = fma(f0 * fac01, f.b0, k);
  k = fma(fs * fac10, f.a3, k);
  k = fma(f0 * fac10, f.a2, k);
  k = fma(fs * fac00, f.a1, k);
  k = fma(f0 * fac00, f.a0, k);
  R cont = uni(R(v.f(6, lay))) * (f.fo0 + uni(R(v.f(7, lay))) * (f.fo1 - f.fo0));
  if (LOWER) cont = uni(R(v.f(4, lay))) * (f.s0 + uni(R(v.f(5, lay))) * (f.s1 - f.s0)) + cont;
  cont = q.cont ? cont : R(0);
  taug = f.kmul * k + colh2o * cont + R(v.c(q.sm, lay)) * R(q.mcoef);
  R ray = R(q.rayl);
  if (LOWER) ray = q.rayl_js ? f.r0 + fs * (f.r1 - f.r0) : ray;
  taur = colmol * ray;
}

// Wavefront sums: exchanges across the wave halves and row pairs (v_permlane32_swap,
// v_permlane16_swap), then DPP exchanges inside each row of 16 lanes.
template <int CTRL>
__device__ inline double dpp_d(double x) {
  int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xf, 0xf, false);
  int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
constexpr int DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_MIRROR = 0x140, DPP_HALF_MIRROR = 0x141;
// Every lane of a 16-lane row ends with the row's total of x (row mirror, half-row mirror, then
// the two quad exchanges: each step pairs every lane with another).
__device__ inline double row_sum_d(double x) {
  x += dpp_d<DPP_MIRROR>(x);
  x += dpp_d<DPP_HALF_MIRROR>(x);
  x += dpp_d<DPP_XOR1>(x);
  x += dpp_d<DPP_XOR2>(x);
  return x;
}
// a + b after exchanging the upper wave half of a with the lower half of b (v_permlane32_swap:
// lanes 0-31 end with a's sum over l and l + 32, lanes 32-63 with b's), or the odd rows of a with
// the even rows of b (v_permlane16_swap: even rows end with a's sum over rows 2r and 2r + 1, odd
// rows with b's).
__device__ inline void swap_add(double &a, double b, bool rows16) {
  const unsigned alo = __double2loint(a), ahi = __double2hiint(a);
  const unsigned blo = __double2loint(b), bhi = __double2hiint(b);
  auto lo = rows16 ? __builtin_amdgcn_permlane16_swap(alo, blo, false, false)
                   : __builtin_amdgcn_permlane32_swap(alo, blo, false, false);
  auto hi = rows16 ? __builtin_amdgcn_permlane16_swap(ahi, bhi, false, false)
                   : __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false);
  a = __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
}
// Wavefront sums by recursive halving. Four values: the total of v_k ends in every lane of row k
// (lanes 16k .. 16k + 15).
__device__ inline double wave_sum4d(double v0, double v1, double v2, double v3) {
  swap_add(v0, v2, false);   // upper half: v2
  swap_add(v1, v3, false);   // upper half: v3
  swap_add(v0, v1, true);    // rows 0, 1, 2, 3: v0, v1, v2, v3
  return row_sum_d(v0);
}
// Two values: v0's total in lanes 0-31, v1's in lanes 32-63.
__device__ inline double wave_sum2d(double v0, double v1) {
  swap_add(v0, v1, false);
  swap_add(v0, v0, true);
  return row_sum_d(v0);
}
// Sixteen values: swaps of half the values across the wave halves and across row pairs leave
// four per lane, each summed over its four rows; a row mirror and a half-row mirror, each
// exchanging one of two selected values, and the quad exchanges finish within the row. Lane l
// ends with the total of x[l >> 2] (lanes 0, 4, ..., 60 hold x[0] ... x[15]).
__device__ inline double wave_sum16d(double (&x)[16], int lane) {
#pragma unroll
  for (int p = 0; p < 8; p++) swap_add(x[p], x[p + 8], false);   // lanes >= 32: x[p + 8]
#pragma unroll
  for (int p = 0; p < 4; p++) swap_add(x[p], x[p + 4], true);    // odd rows: x[p + 4]
  const bool b3 = lane & 8, b2 = lane & 4;
  const double w0 = (b3 ? x[2] : x[0]) + dpp_d<DPP_MIRROR>(b3 ? x[0] : x[2]);
  const double w1 = (b3 ? x[3] : x[1]) + dpp_d<DPP_MIRROR>(b3 ? x[1] : x[3]);
  double s = (b2 ? w1 : w0) + dpp_d<DPP_HALF_MIRROR>(b2 ? w0 : w1);
  s += dpp_d<DPP_XOR1>(s);
  s += dpp_d<DPP_XOR2>(s);
  return s;
}

// ------------------------------------------------------------------------------------------
// The mixed form's layer. Its FP64 front -- the delta-scaled optics, the two-stream
// coefficients and k, the conservative-scattering decision and every exponential-table lookup --
// is the FP64 form's own arithmetic (reftra_front, tbl_exp1), so its lookups pick the FP64
// form's table entries. Its FP32 back evaluates the PIFM solution (src/rrtmg_sw_reftra.f90:233-322)
// in the deficits d = 1 - e of the two transmittances. With e1 = 1 - d1 (e1 = exp(-k tau)),
// e2 = 1 - d2 (e2 = exp(-tau / mu0)), s1 = 1 - e1^2 = d1 (2 - d1), the reference's numerators,
// multiplied through by e1, lose the terms that cancel at tau = 0:
//   R    =  w [B2 s1 + C2 (d1 + d2 - d1 d2)] rD,
//   Tdif = -w [C1 (d1 - d2) + (1 - d2) B1 s1] rD,       T = e2 + Tdif,
//   B2 = (1 + k mu0)(a2 - k g3), C2 = 2k (g3 - a2 mu0), B1 = (1 - k mu0)(a1 - k g4),
//   C1 = 2k (g4 + a1 mu0), rD = 1 / ((1 - k^2 mu0^2) D'),  D' = k (1 + e1^2) + g1 s1,
//   Rd = g2 s1 / D',  Td = 2k e1 / D'
// (every term O(tau) with no cancellation), and the absorptances of the direct and the diffuse
// beam, with c = 1 - w carried explicitly (gamma1 - gamma2 = 2c),
//   a  = 1 - R - T   = d2 - (R + Tdif),
//   ad = 1 - Rd - Td = d1 (k d1 + 2c (2 - d1)) / D',
// keep their accuracy relative to tau / mu0 however thin the layer. FP64 remains in the
// conservative branch and on lanes with |1 - k mu0| < 2e-3, where 1 - (k mu0)^2 vanishes with
// the numerators and the table's steps make the quotient large (they run reftra_back).
// ------------------------------------------------------------------------------------------
__device__ inline float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
struct RTm {
  float r, rd, t, td, a, ad;
};
constexpr double SING_BAND = 2.e-3;   // |1 - k mu0| below which a lane runs the FP64 back
// 1 - e of a table transmittance e = tbl_exp1(ze): the expansion's ze - ze^2 / 2 below od_lo
__device__ inline float tbl_deficit(double ze, double e) {
  return float(ze <= OD_LO ? ze - 0.5 * ze * ze : 1. - e);
}
__device__ inline RTm reftra_mixed(const RtFront &f, double zto1, double zw, float zc, double prmuz,
                                   double e2, float d2) {
  RTm o;
  if (f.cons || fabs(1. - f.kmu) < SING_BAND) {
    const RT4 q = reftra_back(f, zto1, zw, prmuz, e2);
    o.r = float(q.r), o.rd = float(q.rd), o.t = float(q.t), o.td = float(q.td);
    o.a = float(1. - q.r - q.t), o.ad = float(1. - q.rd - q.td);
    return o;
  }
  const float w = float(zw), mu = float(prmuz), k = float(f.k), kmu = float(f.kmu);
  const float g1 = float(f.g1), g2 = float(f.g2), g3 = float(f.g3), g4 = float(f.g4);
  const float a1 = g1 * g4 + g2 * g3, a2 = g1 * g3 + g2 * g4;
  const float d1 = tbl_deficit(fmin(f.k * zto1, 500.), f.e1), e1 = float(f.e1), e2f = float(e2);
  const float s1 = d1 * (2.f - d1);
  const float Dp = k * (1.f + e1 * e1) + g1 * s1;
  const float rDp = rcpf(Dp);
  const float rD = rcpf(1.f - kmu * kmu) * rDp;
  const float B2 = (1.f + kmu) * (a2 - k * g3), C2 = 2.f * k * (g3 - a2 * mu);
  const float B1 = (1.f - kmu) * (a1 - k * g4), C1 = 2.f * k * (g4 + a1 * mu);
  const float NR = B2 * s1 + C2 * (d1 + d2 - d1 * d2);
  const float NT = C1 * (d1 - d2) + (1.f - d2) * B1 * s1;
  o.r = w * NR * rD;
  const float tdif = -w * NT * rD;
  o.t = e2f + tdif;
  o.rd = g2 * s1 * rDp;
  o.td = 2.f * k * e1 * rDp;
  o.a = d2 - (o.r + tdif);
  o.ad = d1 * (k * d1 + 2.f * zc * (2.f - d1)) * rDp;
  return o;
}

// The elimination step of sweep_r in FP32 (state and coefficients), B = prdnd before the layer.
struct SweepF {
  float tdbt, ztdn, prdnd;
};
__device__ inline void sweep_f(SweepF &s, const RTm &l, float dbt, float &Cc, float &E) {
  float a = s.ztdn - s.tdbt;
  float zreflect = rcpf(1.f - l.rd * s.prdnd);
  Cc = (l.rd * a + l.r * s.tdbt) * zreflect;
  E = l.td * zreflect;
  float ztdn = s.tdbt * l.t + (l.td * (a + s.tdbt * l.r * s.prdnd)) * zreflect;
  float prdnd = l.rd + l.td * l.td * s.prdnd * zreflect;
  s.tdbt = dbt * s.tdbt;
  s.ztdn = ztdn;
  s.prdnd = prdnd;
}

// Wavefront sums in FP32, as wave_sum4d / wave_sum2d (one 32-bit exchange a value).
template <int CTRL>
__device__ inline float dpp_f(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xf, 0xf, false));
}
__device__ inline float row_sum_f(float x) {
  x += dpp_f<DPP_MIRROR>(x);
  x += dpp_f<DPP_HALF_MIRROR>(x);
  x += dpp_f<DPP_XOR1>(x);
  x += dpp_f<DPP_XOR2>(x);
  return x;
}
__device__ inline void swap_add_f(float &a, float b, bool rows16) {
  const unsigned ua = __float_as_uint(a), ub = __float_as_uint(b);
  auto p = rows16 ? __builtin_amdgcn_permlane16_swap(ua, ub, false, false)
                  : __builtin_amdgcn_permlane32_swap(ua, ub, false, false);
  a = __uint_as_float(p[0]) + __uint_as_float(p[1]);
}
__device__ inline float wave_sum4f(float v0, float v1, float v2, float v3) {
  swap_add_f(v0, v2, false);
  swap_add_f(v1, v3, false);
  swap_add_f(v0, v1, true);
  return row_sum_f(v0);
}
__device__ inline float wave_sum2f(float v0, float v1) {
  swap_add_f(v0, v1, false);
  swap_add_f(v0, v0, true);
  return row_sum_f(v0);
}

// ------------------------------------------------------------------------------------------
// KISS and the sub-column masks (src/mcica_subcol_gen_sw.f90:468-659, :721-753).
// ------------------------------------------------------------------------------------------
struct Kiss {
  uint32_t s1, s2, s3, s4;
};
__device__ inline double kiss_draw(Kiss &k) {
#pragma clang fp contract(off)
  k.s1 = 69069u * k.s1 + 1327217885u;
  uint32_t x = k.s2;
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  k.s2 = x;
  k.s3 = 18000u * (k.s3 & 65535u) + (k.s3 >> 16);
  k.s4 = 30903u * (k.s4 & 65535u) + (k.s4 >> 16);
  int32_t kiss = (int32_t)(k.s1 + k.s2 + (k.s3 << 16) + k.s4);
  return (double)kiss * 2.328306e-10 + 0.5;
}
// Seeds of a column: fractional Pa of the bottom four layers times 1e9 (:471-479).
__device__ inline uint32_t kiss_seed(double play_hpa) {
#pragma clang fp contract(off)
  double p = play_hpa * 1.e2;
  return (uint32_t)(int32_t)((p - (double)(int32_t)p) * 1000000000.0);
}
__device__ inline Kiss kiss_jump(Kiss k, const KissJump &j) {
  Kiss o;
  o.s1 = j.a1 * k.s1 + j.c1;
  uint32_t r = 0;
  for (int i = 0; i < 32; i++) r ^= ((k.s2 >> i) & 1u) ? j.m2[i] : 0u;
  o.s2 = r;
  o.s3 = (uint32_t)(((uint64_t)k.s3 * j.p3) % 1179647999ull);   // 18000*2^16 - 1
  o.s4 = (uint32_t)(((uint64_t)k.s4 * j.p4) % 2025259007ull);   // 30903*2^16 - 1
  return o;
}
__device__ inline double mt_real(uint32_t u) { return (double)u / (4294967296.0 - 1.0); }

// Sub-column cloudiness of g-point g in every layer of column c (bottom to top), written as
// bits into mask[w * NGPT + g] (bit l%32 of word l/32). Returns an error code.
// Sample mode (the column program's McICA loop, src/rrtmg_sw.1col.f90:577-600, with every sample
// a workgroup): col_jump[c] advances column c's KISS streams by its own permuteseed, and each
// column reads an MT19937 stream of its own, mt + c * mt_stride, indexed as a single column.
// cf[l * cs] and al[l * cs] are the column's cloud fraction and overlap parameter of layer l
// (the caller's arrays at column c with stride ncol, or the copies staged in LDS with stride 1).
// ltop is the highest layer whose fraction is at least cldmin (-1 if none, L - 1 if not known).
// KISS draws lie strictly inside (0, 1), so a layer below cldmin (fraction taken as 0) is never
// cloudy and the overlap recurrence above ltop reaches no cloudy layer: the KISS path draws up
// to ltop and leaves the layers above clear. A Mersenne Twister draw can be exactly 1, which
// makes a clear layer cloudy (cdf >= 1 - 0), so that path draws every layer.
__device__ static int column_masks(int icld, int irng, int permuteseed, int C, int L, int c, int g,
                                   const double *play, const double *cf, const double *al, size_t cs,
                                   const uint32_t *mt, const KissJump *jumps, uint32_t *mask, int ltop,
                                   const KissJump *col_jump = nullptr, size_t mt_stride = 0) {
#pragma clang fp contract(off)
  Kiss k = {0, 0, 0, 0};
  if (irng == 0) {
    if (play[c] * 1.e2 < play[c + C] * 1.e2) return RRTMGSW_ERR_KISS_SEED;
    k.s1 = kiss_seed(play[c]);
    k.s2 = kiss_seed(play[c + (size_t)C]);
    k.s3 = kiss_seed(play[c + 2 * (size_t)C]);
    k.s4 = kiss_seed(play[c + 3 * (size_t)C]);
    if (col_jump) k = kiss_jump(k, col_jump[c]);
    k = kiss_jump(k, jumps[g]);
  }
  // MT draw index: ((g C + c) L + l) for one stream per call, (g L + l) in column c's own stream
  const int mC = mt_stride ? 1 : C, mc = mt_stride ? 0 : c;
  if (mt_stride) mt += (size_t)c * mt_stride;
  double cdf_prev = 0.0, r3 = 0.0;
  if (icld == 3) r3 = irng == 0 ? kiss_draw(k) : mt_real(mt[(size_t)g * mC + mc]);
  uint32_t word = 0;
  double cldf_prev = 0.0;
  const int lend = irng == 0 ? min(ltop, L - 1) : L - 1;   // the last layer drawn
  // one layer: its draws (cdf, cdf2), the overlap recurrence, the mask bit
  auto layer = [&](int l, double cdf, double cdf2) {
    double cldf = cf[l * cs];
    if (cldf < CLDMIN) cldf = 0.0;
    if (l > 0) {
      if (icld == 2) {  // :541-551
        const double cprev = cldf_prev;   // layer l - 1's fraction, clamped as above
        if (cdf_prev > 1. - cprev) cdf = cdf_prev;
        else cdf = cdf * (1. - cprev);
      } else if (icld >= 4) {  // :607-611, :647-651
        if (cdf2 < al[l * cs]) cdf = cdf_prev;
      }
    }
    cdf_prev = cdf;
    cldf_prev = cldf;
    if (cdf >= 1. - cldf) word |= 1u << (l & 31);
    if ((l & 31) == 31 || l == lend) {
      mask[(l >> 5) * NGPT + g] = word;
      word = 0;
    }
  };
  for (int w = (lend + 32) >> 5; w < (L + 31) >> 5; w++) mask[w * NGPT + g] = 0u;   // above lend
  if (irng == 0 || icld == 3) {
    for (int l = 0; l <= lend; l++) {
      double cdf = r3, cdf2 = 0.0;
      if (icld != 3) {
        cdf = kiss_draw(k);
        if (icld >= 4) cdf2 = kiss_draw(k);
      }
      layer(l, cdf, cdf2);
    }
  } else {
    // the Mersenne Twister stream's words of four layers requested together (they are
    // consecutive: draw index ((g C + c) L + l), pairs for icld 4/5)
    for (int l0 = 0; l0 < L; l0 += 4) {
      const int nl = min(4, L - l0);
      const size_t base = ((size_t)g * mC + mc) * L + l0;
      uint32_t w[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
      if (icld >= 4) {
#pragma unroll
        for (int u = 0; u < 8; u++)
          if (u < 2 * nl) w[u] = mt[2 * base + u];
      } else {
#pragma unroll
        for (int u = 0; u < 4; u++)
          if (u < nl) w[u] = mt[base + u];
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (u < nl) {
          if (icld >= 4) layer(l0 + u, mt_real(w[2 * u]), mt_real(w[2 * u + 1]));
          else layer(l0 + u, mt_real(w[u]), 0.0);
        }
    }
  }
  return 0;
}

}  // namespace rrtmgsw
