// rrtmgsw_internal.h -- shared host/device definitions of the MI355X shortwave library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace rrtmgsw {

constexpr int NGPT = 112;      // reduced g-points (modules/parrrsw.f90:36)
constexpr int NBND = 14;       // shortwave bands 16..29
constexpr int NAER = 6;        // ECMWF aerosol types (modules/parrrsw.f90:31)
constexpr int WAVE_G = 56;     // g-points per wavefront: wave 0 = bands 16-22, wave 1 = bands 23-29
constexpr int MAX_LAY = 256;   // layer limit of the per-column LDS staging
constexpr int NTBL = 10000;    // exp lookup table, modules/rrsw_tbl.f90:28
constexpr int SCR_COEFS = 6;   // elimination coefficients kept per (layer, g-point): B C E for two skies
// Per-column scratch of the forward kernels, in doubles: the elimination coefficients
// (nlay, SCR_COEFS, NGPT), then the cloud optics (tau, ssa, asymmetry, unscaled tau) of each
// (layer, band) (nlay, NBND, 4). After the launch's columns come each column's per-level flux
// sums over g (scr_red_doubles: 2 waves x (nlay + 1) x 4, then 2 x nlay x 2; the mixed form keeps
// FP32 sums in the same space), used instead of LDS where a column's LDS would otherwise limit
// the workgroups per CU below what the registers allow (137 layers, for instance).
constexpr size_t scr_col_doubles(int nlay) { return (size_t)nlay * (SCR_COEFS * NGPT + 4 * NBND); }
constexpr size_t scr_red_doubles(int nlay) { return 8 * ((size_t)nlay + 1) + 4 * (size_t)nlay; }
// what one column of a launch takes of the scratch buffer
constexpr size_t scr_alloc_doubles(int nlay) { return scr_col_doubles(nlay) + scr_red_doubles(nlay); }
// Pressure thresholds of setcoef_sw's decisions (src/rrtmg_sw_setcoef.f90:169-176, :208-210):
// pthr[PTHR_JP + k - 2] (k = 2..58) is the smallest pressure at which jp < k, pthr[PTHR_TROP] the
// smallest at which log p > 4.56, pthr[PTHR_LOW] the smallest at which log p >= 6.62.
constexpr int PTHR_JP = 0, PTHR_TROP = 57, PTHR_LOW = 58, NPTHR = 64;
constexpr double PADE = 0.278;
constexpr double OD_LO = 0.06;

// Which of the reference's cloud-optics STOP messages an error is (src/rrtmg_sw_cldprmc.f90:
// 183-294, src/rrtmg_sw_cldprop.f90:194-311): kernels report RRTMGSW_ERR_RADIUS or
// RRTMGSW_ERR_OPTICS in the low byte of their status word and one of these above it.
enum CloudStop : int {
  CS_ICE_RADIUS = 1, CS_ICE_GEN_SIZE, CS_LIQ_RADIUS, CS_FDELTA_LO, CS_FDELTA_HI,
  CS_ICE_EXT_LO, CS_ICE_SSA_HI, CS_ICE_SSA_LO, CS_ICE_ASY_HI, CS_ICE_ASY_LO,
  CS_LIQ_EXT_LO, CS_LIQ_SSA_HI, CS_LIQ_SSA_LO, CS_LIQ_ASY_HI, CS_LIQ_ASY_LO, CS_COUNT
};
constexpr int cloud_stop(int code, int which) { return code | (which << 8); }

// Key-species interpolation modes of the unified gas-optics template (one record per g-point
// replaces the fourteen taumolNN routines of src/rrtmg_sw_taumol.f90).
enum KeyMode : int8_t { KEY_NONE = 0, KEY_BIN = 1, KEY_MONO = 2 };
// Species columns staged per layer: index into ColumnLds::col.
enum Species : int8_t { SP_H2O = 0, SP_CO2 = 1, SP_O3 = 2, SP_CH4 = 3, SP_O2 = 4, SP_NONE = -1 };

// Per-g-point parameters. Offsets are element offsets into the device table (doubles).
// Table layouts are transposed so that lanes of one band read consecutive addresses:
// absa (ind, ig) at absa + (ind-1)*ng + ig, and likewise for absb, selfref, forref, rayla,
// and the solar tables (js, ig).
struct GPoint {
  int8_t band, ibnd, ig, ng;
  int8_t lo_mode, lo_s1, lo_s2, lo_cont;   // lower atmosphere key species and continuum
  int8_t hi_mode, hi_s1, hi_s2, hi_for;    // upper atmosphere
  int8_t lo_minor, hi_minor;               // minor absorber species (SP_NONE if none)
  int8_t rayl_js;                          // 1: lower Rayleigh interpolated in js (band 24)
  int8_t sol_interp;                       // 1: solar source interpolated in js
  int8_t sol_upper;                        // 1: laysolfr found in the upper loop
  int8_t layreffr;                         // reference-layer index for laysolfr (0: none)
  int8_t lo_nsp, hi_nsp;                   // binary-species steps (9 or 1 / 5 or 1)
  int8_t nfor;
  int8_t sa, sb, sself, sfor, ssf, sra;    // row strides of the tables (0: term unused -> zero pad)
  int8_t pad[3];
  double lo_strrat, lo_mult, hi_strrat, hi_mult;
  double lo_minor_coef, hi_minor_coef;
  double rayl_lo, rayl_hi, sol_scale;
  int32_t absa, absb, selfref, forref, rayla;
  int32_t sflux, irrad, facb, snsp;        // solar source tables (ng, nsf) transposed
  int32_t nsf;
  int32_t zero;                            // offset of a zero pad (targets of unused terms)
};

// Device copies of all read-only tables (built once by rrtmgsw_init).
struct DeviceTables {
  const double *tab;        // k-tables, solar tables, cloud optics
  const double *exp_tbl;    // NTBL+1 entries
  const GPoint *gp;         // NGPT records
  const double *pthr;       // NPTHR pressure thresholds of setcoef_sw's index decisions
  // cloud optics offsets (tables (n, 16:29) column-major: value(i, band) at off + (i-1) + n*(band-16))
  int32_t extliq1, ssaliq1, asyliq1, extice2, ssaice2, asyice2, extice3, ssaice3, asyice3, fdlice3;
  int32_t abari, bbari, cbari, dbari, ebari, fbari;
  int32_t rsrtaua, rsrpiza, rsrasya;   // (14, 6)
  const uint32_t *kiss_jump;           // per-call KISS jump-ahead records (see KissJump)
};

// Jump-ahead of one KISS stream to the first draw of sub-column g:
// s1 -> a1*s1 + c1, s2 -> M*s2 over GF(2) (columns m2[0..31]), s3 -> s3*p3 mod m3,
// s4 -> s4*p4 mod m4 (the multiply-with-carry components are multiplications by 2^-16 modulo
// 18000*2^16-1 and 30903*2^16-1; seeds below those moduli stay below them).
struct KissJump {
  uint32_t a1, c1, p3, p4;
  uint32_t m2[32];
};

// Scalar inputs of one rrtmg_sw call, prepared on the host (inatm_sw's solar variability and
// flux adjustment, src/rrtmg_sw_rad.f90:1226-1512, and the driver's flag handling).
struct SwScalars {
  double adjflux[NBND];
  double svar_f, svar_s, svar_i;
  double svar_bnd[3][NBND];  // f, s, i by band
  double heatfac;
  int isolvar;
  int icld;                  // after the driver's clamp (0..3)
  int iaer;                  // 0, 6 or 10
  int inflag, iceflag, liqflag;
};

// Arguments of the radiative-transfer kernel.
struct RtArgs {
  int ncol, nlay, col0, ncol_chunk;
  // profiles (ncol, nlay) / (ncol, nlay+1) / (ncol)
  const double *play, *plev, *tlay, *h2ovmr, *o3vmr, *co2vmr, *ch4vmr, *o2vmr;
  const double *asdir, *asdif, *aldir, *aldif, *coszen;
  // materialized McICA inputs (112, ncol, nlay) -- ABI path
  const double *cldfmcl, *taucmcl, *ssacmcl, *asmcmcl, *fsfcmcl, *ciwpmcl, *clwpmcl;
  const double *reicmcl, *relqmcl;
  // fused path: physical cloud inputs of mcica_subcol_sw
  int fused, mc_icld, mc_irng, permuteseed;
  const double *cldfrac, *ciwp, *clwp, *rei, *rel, *tauc, *ssac, *asmc, *fsfc, *alpha;
  // non-McICA path: the physical inputs above (no alpha) are the layer clouds of cldprop_sw
  int layer_clouds;
  // column program (non-McICA kernel with spectral output): layer amounts given directly,
  // coldry (ncol, nlay) and wkl (7, ncol, nlay); per-band surface albedo albbnd (14, ncol);
  // per-g level terms gspec (per chunk column (nlay+1) x 4 x 112) reduced into the per-band
  // total-sky fluxes bandu, bandd, banddir (ncol, nlay+1, 14); idelm 0/1 selects the direct
  // beam without / with delta scaling (src/rrtmg_sw_spcvmc.f90:449-470, :623-628)
  const double *coldry_in, *wkl_in, *albbnd;
  double *gspec, *bandu, *bandd, *banddir;
  int idelm;
  // column program with McICA: one sample per column, each with its own KISS jump (col_jump,
  // KissJump per column) or MT19937 stream (mt_stream + c * mt_stride)
  const uint32_t *col_jump;
  size_t mt_stride;
  const uint32_t *mt_stream;     // Mersenne-Twister draws for irng = 1 (fused path)
  // aerosol
  const double *tauaer, *ssaaer, *asmaer, *ecaer;
  // outputs (ncol, nlay+1), (ncol, nlay)
  double *swuflx, *swdflx, *swhr, *swuflxc, *swdflxc, *swhrc;
  double *scratch;               // per (column, layer, coefficient, g) elimination coefficients
  int *status;                   // first error code raised inside a kernel
  SwScalars sc;
};

struct McicaArgs {
  int ncol, nlay, icld, irng, permuteseed;
  const double *play, *cldfrac, *ciwp, *clwp, *rei, *rel, *tauc, *ssac, *asmc, *fsfc, *alpha;
  const uint32_t *mt_stream;
  double *cldfmcl, *ciwpmcl, *clwpmcl, *reicmcl, *relqmcl, *taucmcl, *ssacmcl, *asmcmcl, *fsfcmcl;
  int *status;
};

// Outputs of rrtmgsw_setcoef_dev (NULL: not written): laytrop, laylow (ncol); laysolfr (ncol, 14);
// the rest (ncol, nlay).
struct SetcoefOut {
  int *laytrop, *laylow, *laysolfr;
  int *jp, *jt, *jt1, *indself, *indfor;
  double *fac00, *fac01, *fac10, *fac11, *selffac, *selffrac, *forfac, *forfrac;
  double *colh2o, *colco2, *colo3, *colch4, *colo2, *colmol;
};

// Adjoint of the fused path (rrtmgsw_adjoint.hip): seeds of the total-sky TOA upward flux and
// heating rates in, gradients with respect to T, H2O, O3, in-cloud water paths and albedos out.
struct AdjArgs {
  RtArgs f;                       // forward inputs (fused form), scalars, status word
  const double *seed_toa;         // (ncol)        cotangent of swuflx at the top level
  const double *seed_hr;          // (ncol, nlay)  cotangent of swhr
  double *d_tlay, *d_h2ovmr, *d_o3vmr, *d_clwp, *d_ciwp;   // (ncol, nlay)
  double *d_alb;                  // (ncol, 4): asdir, asdif, aldir, aldif
  double *work;                   // per column adj_work_bytes(nlay, ckpt)
  int ckpt;                       // 0: store every sweep state; K > 0: checkpoint every K layers
};
size_t adj_work_bytes(int nlay, int ckpt);   // work area per column
size_t adj_lds_bytes(int nlay);               // dynamic LDS of one adjoint workgroup

// Kernel launchers (rrtmgsw_kernels.hip)
hipError_t launch_rt_fp64(const RtArgs &a, const DeviceTables &t, hipStream_t s);
hipError_t launch_rt_mixed(const RtArgs &a, const DeviceTables &t, hipStream_t s);
hipError_t launch_rt_exact(const RtArgs &a, const DeviceTables &t, hipStream_t s);
hipError_t launch_rt_spectral(const RtArgs &a, const DeviceTables &t, hipStream_t s);
hipError_t launch_adjoint(const AdjArgs &a, const DeviceTables &t, hipStream_t s);
hipError_t launch_mcica(const McicaArgs &a, const DeviceTables &t, hipStream_t s);
hipError_t launch_setcoef(const RtArgs &a, const SetcoefOut &o, const DeviceTables &t, hipStream_t s);
// The set bits of x^(k J) mod P (J = 2^log2_seg, P the MT19937 characteristic polynomial) for
// k = 1 .. count, on the device: bits[offs[k - 1] .. offs[k]).
struct MtJumpTable {
  int log2_seg, count;
  const uint32_t *bits, *offs;
};
// MT19937 stream of ndraw tempered words to out, in segments of 2^mt_log2_seg(ndraw) words when
// jt (for that segment length, count >= segments - 1) is given and the stream holds more than two;
// then out must hold ndraw + mt_extra_words(ndraw) words: segment start states go after the stream.
int mt_log2_seg(long long ndraw);
size_t mt_extra_words(long long ndraw);
hipError_t launch_mt_stream(uint32_t seed, long long ndraw, uint32_t *out, const MtJumpTable *jt, hipStream_t s);
std::vector<uint32_t> mt_jump_coefficients(int log2_jump);          // host, rrtmgsw_mtjump.cpp
std::vector<uint32_t> mt_jump_multiples(int log2_jump, int count);  // host, rrtmgsw_mtjump.cpp

}  // namespace rrtmgsw
