// rrtmgsw_adjoint.hip -- hand-written adjoint of the fused McICA shortwave path (gfx950).
//
// Differentiates J = sum_c [ seed_toa(c) * swuflx(c, top) + sum_k seed_hr(c, k) * swhr(c, k) ]
// (total sky) with respect to layer temperature, H2O and O3 volume mixing ratios, the
// in-cloud liquid and ice water paths and the four surface albedos, for the exact-exponential
// FP64 forward (the reference's algorithm with exp() in place of the 10001-entry lookup table,
// whose piecewise-constant values have no useful derivative). McICA sub-column masks are held
// fixed (they are integer streams of the inputs' pressures). The forward is the one of
// rt_kernel in its exact-exponential mode: setcoef_sw, taumol_sw, cldprmc_sw, delta scaling,
// reftra_sw, vrtqdr_sw, the incident solar flux and the heating-rate differences.
//
// Mapping: as the forward, one workgroup per column, one lane per g-point. Per lane:
//   F1  top -> bottom: layer optics and the total-sky downward elimination (vrtqdr_sw
//       :144-159 as U_j = C_j + E_j U_(j+1), D_j = ztdn_j + prdnd_j U_j); since J reaches U_j
//       through U_j itself, D_j and U_(j-1), ubar_j = dJ/dU_j = ubar_(j-1) E_(j-1) +
//       zinc (wU_j + wD_j prdnd_j) is known as soon as layer j is: stored per layer, with the
//       gas optics (taug, taur) of every layer, with the sweep state s_j = (tdbt, ztdn, prdnd)
//       of every layer (store mode) or of every K-th layer (checkpoint mode); then the surface
//       adjoint (albedos, s_L);
//   R1  bottom -> top: s_j (loaded, or recomputed per K-layer segment from its checkpoint and
//       the kept gas optics), the layer optics and reftra recomputed (cheaper than storing and
//       reloading them), U_j by the back substitution and the objective per unit incident flux
//       Phi on the way up, the adjoint of the elimination step, reftra_sw's reverse-mode
//       derivative, the reverse of the layer-optics combination, and the adjoint of cldprmc_sw
//       (per-(layer, band) water-path Jacobians from the prologue); the adjoints of taug and
//       taur are left in the layer's slot;
//   R2  every layer: the hand-written adjoint of the gas optics (taumol template) from them;
//       per-layer sums over g go to LDS accumulators;
//   end the incident-flux adjoint (solar source interpolated at laysolfr), and per layer the
//       adjoint of inatm_sw + setcoef_sw onto T, H2O, O3.
// Memory per (column, layer, g-point): ubar_j, s_j and the gas and Rayleigh optical depths, 6
// doubles (store mode), or ubar_j, the optical depths and 3/K for the checkpoints plus a 3K
// segment buffer per column (checkpoint mode, every layer's sweep recomputed); plus 9 doubles
// per (layer, band) for the cloud table.
#include <type_traits>

#include "rrtmgsw_device.h"

namespace rrtmgsw {

// per-layer accumulator slots (sums over g of the adjoints of setcoef_sw's outputs)
enum : int {
  AC_COL = 0,        // 0..4 species columns (SP_H2O .. SP_O2), 5 colmol
  AC_FAC = 6,        // fac00 fac01 fac10 fac11
  AC_SELFFAC = 10,
  AC_SELFFRAC = 11,
  AC_FORFAC = 12,
  AC_FORFRAC = 13,
  AC_CLWP = 14,
  AC_CIWP = 15,
  AC_N = 16
};

// Adjoint of the key-species parameter (specparm = c1 / (c1 + strrat c2), fs = frac(mult *
// specparm), clipped at oneminus) given d fs and d speccomb: adds to the adjoints of c1 and c2.
__device__ inline void key_spec_adj(double c1, double c2, double strrat, double mult, bool bin,
                                    double fs_b, double speccomb_b, double &c1_b, double &c2_b) {
  double speccomb = c1 + strrat * c2;
  if (bin) {
    // only binned bands use the ratio; the others can have speccomb = 0 (no water vapour in a
    // band keyed on it alone), where its reciprocal is infinite
    const double rsc = frcp1(speccomb);
    const double specparm = c1 * rsc;
    const double specparm_b = specparm < ONEMINUS ? fs_b * mult : 0.0;
    speccomb_b -= specparm_b * c1 * rsc * rsc;
    c1_b += specparm_b * rsc;
  }
  c1_b += speccomb_b;
  c2_b += speccomb_b * strrat;
}

// Adjoint of gas_fetch + gas_finish for one regime (the lane's constants in q): re-reads the
// table entries and accumulates d/d(setcoef outputs) given the adjoints of taug and taur. The
// column adjoints are gathered per role (key species 1 and 2, minor absorber, H2O, colmol) and
// added to the species slots once.
template <bool LOWER>
__device__ inline void gas_adj_r(const GasLane &q, const double *__restrict__ tab, ColView &v, int lay,
                                 double tg_b, double tr_b, double *acc) {
  const GasFetch<double> f = gas_fetch<LOWER, double>(q, tab, v, lay);
  const double fac00 = uni(v.f(0, lay)), fac01 = uni(v.f(1, lay)), fac10 = uni(v.f(2, lay)), fac11 = uni(v.f(3, lay));
  const double colh2o = uni(v.c(SP_H2O, lay)), colmol = uni(v.c(5, lay));
  const double forfac = uni(v.f(6, lay)), forfrac = uni(v.f(7, lay));
  const double fs = f.fs, f0 = 1.0 - fs;
  const double k = f0 * fac00 * f.a0 + fs * fac00 * f.a1 + f0 * fac10 * f.a2 + fs * fac10 * f.a3 +
                   f0 * fac01 * f.b0 + fs * fac01 * f.b1 + f0 * fac11 * f.b2 + fs * fac11 * f.b3;
  // taug = kmul k + colh2o cont + c_minor mcoef; taur = colmol ray
  const double kmul_b = tg_b * k, k_b = tg_b * f.kmul;
  double fs_b = k_b * (fac00 * (f.a1 - f.a0) + fac10 * (f.a3 - f.a2) + fac01 * (f.b1 - f.b0) + fac11 * (f.b3 - f.b2));
  acc[AC_FAC + 0] += k_b * (f0 * f.a0 + fs * f.a1);
  acc[AC_FAC + 2] += k_b * (f0 * f.a2 + fs * f.a3);
  acc[AC_FAC + 1] += k_b * (f0 * f.b0 + fs * f.b1);
  acc[AC_FAC + 3] += k_b * (f0 * f.b2 + fs * f.b3);
  const double fori = f.fo0 + forfrac * (f.fo1 - f.fo0);
  const double tgc_b = q.cont ? tg_b : 0.0;
  double h2o_b, mol_b;
  if (LOWER) {
    const double selffac = uni(v.f(4, lay)), selffrac = uni(v.f(5, lay));
    const double selfi = f.s0 + selffrac * (f.s1 - f.s0);
    h2o_b = tgc_b * (selffac * selfi + forfac * fori);
    const double cont_b = tgc_b * colh2o;
    acc[AC_SELFFAC] += cont_b * selfi;
    acc[AC_SELFFRAC] += cont_b * selffac * (f.s1 - f.s0);
    acc[AC_FORFAC] += cont_b * fori;
    acc[AC_FORFRAC] += cont_b * forfac * (f.fo1 - f.fo0);
    const double ray = q.rayl_js ? f.r0 + fs * (f.r1 - f.r0) : q.rayl;
    if (q.rayl_js) fs_b += tr_b * colmol * (f.r1 - f.r0);
    mol_b = tr_b * ray;
  } else {
    h2o_b = tgc_b * forfac * fori;
    acc[AC_FORFAC] += tgc_b * colh2o * fori;
    acc[AC_FORFRAC] += tgc_b * colh2o * forfac * (f.fo1 - f.fo0);
    mol_b = tr_b * q.rayl;
  }
  const double cm_b = tg_b * q.mcoef;
  const double c1 = v.c(q.s1, lay), c2 = v.c(q.s2, lay);
  double c1_b = q.bin ? 0.0 : kmul_b * q.kmult, c2_b = 0.0;
  key_spec_adj(c1, c2, q.strrat, LOWER ? 8.0 : 4.0, q.bin, fs_b, q.bin ? kmul_b : 0.0, c1_b, c2_b);
#pragma unroll
  for (int i = 0; i < 5; i++)
    acc[AC_COL + i] += (q.s1 == i ? c1_b : 0.0) + (q.s2 == i ? c2_b : 0.0) + (q.sm == i ? cm_b : 0.0);
  acc[AC_COL + SP_H2O] += h2o_b;
  acc[AC_COL + 5] += mol_b;
}

// Everything a layer of one sub-column needs besides the gas optics (aerosol and cloud optics),
// as the forward computes it. Cloud optics come from the per-(layer, band) table CJ built in the
// prologue: every cloudy sub-column cell of a band has the same cloud inputs.
struct LayerIn {
  double ta, wa, ga, cf, ct, cw, cg;
};
// Layer optics of one sub-column (src/rrtmg_sw_spcvmc.f90:442-505): clear sky with delta
// scaling (g0 = ga wa ta / s and w0 = s / tau0 share r1 = 1 / (s tau0); (w0 - w0 f)/(1 - w0 f) and
// (g0 - f)/(1 - f) = g0 / (1 + g0) share r2 = 1 / ((1 - w0 f)(1 + g0)), f = g0^2), and the total
// sky (clear + delta-scaled cloud), as the forward kernel evaluates them.
struct SkyOpt {
  double tau, w, g;
};
__device__ inline SkyOpt clear_optics(double taug, double taur, double ta, double wa, double ga, double &r1,
                                      double &r2) {
#pragma clang fp contract(off)
  const double tau0 = taur + taug + ta;
  const double sct = taur * 1.0 + ta * wa;
  r1 = frcp1(sct * tau0);
  const double g0 = ga * wa * ta * tau0 * r1;
  const double w0 = sct * sct * r1;
  const double zwf = w0 * (g0 * g0);
  r2 = frcp1((1.0 - zwf) * (1.0 + g0));
  return {(1.0 - zwf) * tau0, (w0 - zwf) * (1.0 + g0) * r2, g0 * (1.0 - zwf) * r2};
}

// The layer of one sub-column cell (r, rd, t, td, direct-beam transmittance) in the
// exact-exponential forward. McICA cells are clear (cf = 0) or overcast (cf = 1), so the
// total-sky layer is the clear solution or the cloudy one and each lane evaluates one reftra;
// the forward kernel's blend (1 - cf) rc + cf ro (:539-549) gives the same bits (1 * rc, or ro
// alone). The tape keeps what the reverse sweep needs.
struct LayerV {
  double r, rd, t, td, dbt;
};
struct CellTape {
  SkyOpt c;            // clear optics
  SkyOpt s;            // reftra's inputs: the clear optics, or the total-sky optics of a cloudy cell
  double r1, r2, r3;   // the delta scaling's reciprocals, and 1 / (so tauo) of a cloudy cell
  RtTape rt;
};
__device__ inline LayerV cell_layer(double taug, double taur, const LayerIn &in, double mu0, double rmu,
                                    CellTape &tp) {
#pragma clang fp contract(off)
  tp.c = clear_optics(taug, taur, in.ta, in.wa, in.ga, tp.r1, tp.r2);
  tp.s = tp.c;
  tp.r3 = 0.;
  if (in.cf > 0.0) {
    // total-sky optics (:500-505): go = (ct cw cg + tauc omc gcc) / so, omo = so / tauo
    const double tauo = tp.c.tau + in.ct;
    const double so = tp.c.tau * tp.c.w + in.ct * in.cw;
    const double r = frcp1(so * tauo);
    tp.s = {tauo, so * so * r, (in.ct * in.cw * in.cg + tp.c.tau * tp.c.w * tp.c.g) * tauo * r};
    tp.r3 = r;
  }
  const RTx o = reftra_xt(tp.s.tau, tp.s.w, tp.s.g, mu0, rmu, tp.rt);
  return {o.r, o.rd, o.t, o.td, o.e2};
}

// Adjoint of cell_layer from its tape: given the adjoints lb of (r, rd, t, td, dbt), returns in
// xb the adjoints of (taug, taur, ct, cw, cg): reftra in reverse mode, then the reverse of the
// total-sky combination (cloudy cells) and of the clear-sky delta scaling, written out:
//   clear: tauc0 = taur + taug + ta, s = taur + ta wa, g0 = ga wa ta / s, w0 = s / tauc0,
//          f = g0^2, tauc = (1 - w0 f) tauc0, wc = w0 (1 - f) / (1 - w0 f), gc = g0 / (1 + g0);
//   total: tauo = tauc + ct, so = tauc wc + ct cw, go = (ct cw cg + tauc wc gc) / so,
//          wo = so / tauo.
__device__ inline void cell_layer_adj(const CellTape &tp, double taug, double taur, const LayerIn &in,
                                      double mu0, double rmu, const double *lb, double *xb) {
  double sbt, sbw, sbg;
  reftra_xt_adj(tp.rt, tp.s.tau, tp.s.w, tp.s.g, mu0, rmu, lb[0], lb[1], lb[2], lb[3], lb[4], sbt, sbw, sbg);
  double tcb = sbt, wcb = sbw, gcb = sbg;     // adjoints of the clear optics (tauc, wc, gc)
  double ctb = 0., cwb = 0., cgb = 0.;
  if (in.cf > 0.0) {
    const SkyOpt &c = tp.c, &t = tp.s;
    const double ct = in.ct, cw = in.cw, cg = in.cg;
    const double so = c.tau * c.w + ct * cw;
    const double rtt = so * tp.r3, rso = t.tau * tp.r3;   // 1 / tauo, 1 / so
    const double taob = sbt - sbw * t.w * rtt;
    const double Nb = sbg * rso;
    const double sob = sbw * rtt - sbg * t.g * rso;
    ctb = taob + sob * cw + Nb * cw * cg;
    cwb = sob * ct + Nb * ct * cg;
    cgb = Nb * ct * cw;
    tcb = taob + sob * c.w + Nb * c.w * c.g;
    wcb = sob * c.tau + Nb * c.tau * c.g;
    gcb = Nb * c.tau * c.w;
  }
  // clear-sky delta scaling: 1 / s = tauc0 r1, 1 / tauc0 = s r1, 1 / (1 + g0) = r2 D, 1 / D = r2 (1 + g0)
  const double tauc0 = taur + taug + in.ta;
  const double sc = taur * 1.0 + in.ta * in.wa;
  const double rsc = tauc0 * tp.r1, rtau0 = sc * tp.r1;
  const double g0 = in.ga * in.wa * in.ta * rsc;
  const double w0 = sc * rtau0;
  const double f = g0 * g0;
  const double D = 1.0 - w0 * f;
  const double rg = tp.r2 * D, rD = tp.r2 * (1.0 + g0);
  double g0b = gcb * rg * rg;
  double w0b = wcb * (1.0 - f) * rD * rD - tcb * f * tauc0;
  double fb = -wcb * w0 * (1.0 - w0) * rD * rD - tcb * w0 * tauc0;
  double tauc0b = tcb * D;
  g0b += fb * 2.0 * g0;
  double scb = w0b * rtau0 - g0b * g0 * rsc;
  tauc0b -= w0b * w0 * rtau0;
  xb[0] = tauc0b;            // taug
  xb[1] = tauc0b + scb;      // taur
  xb[2] = ctb, xb[3] = cwb, xb[4] = cgb;
}

// per (layer, band): cloud tau, ssa, asymmetry, then their derivatives with respect to the
// in-cloud liquid water path and to the ice water path
constexpr int CJ_N = 9;
__device__ static LayerIn layer_inputs(const RtArgs &a, const DeviceTables &T, int ibnd, ColView &v,
                                       const double *CJ, int c, int g, int lay, bool active) {
#pragma clang fp contract(off)
  const int C = a.ncol, L = a.nlay;
  LayerIn in = {0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0};
  if (a.sc.iaer == 10) {
    size_t q = c + (size_t)C * (lay + (size_t)L * ibnd);
    in.ta = a.tauaer[q], in.wa = a.ssaaer[q], in.ga = a.asmaer[q];
  } else if (a.sc.iaer == 6) {
    double t6 = 0.0, w6 = 0.0, g6 = 0.0;
    for (int ia = 0; ia < NAER; ia++) {
      double ec = a.ecaer[c + (size_t)C * (lay + (size_t)L * ia)];
      double rt = T.tab[T.rsrtaua + ibnd + NBND * ia], rp = T.tab[T.rsrpiza + ibnd + NBND * ia];
      double rg = T.tab[T.rsrasya + ibnd + NBND * ia];
      t6 = t6 + rt * ec;
      w6 = w6 + rt * ec * rp;
      g6 = g6 + rt * ec * rp * rg;
    }
    if (t6 == 0.) {
      in.ta = 0., in.ga = 0., in.wa = 1.;
    } else {
      in.ga = w6 != 0. ? g6 / w6 : g6;
      in.wa = w6 / t6;
      in.ta = t6;
    }
  }
  if (a.sc.icld >= 1 && active && ((v.mask[(lay >> 5) * NGPT + g] >> (lay & 31)) & 1u)) {
    const double *o = CJ + ((size_t)lay * NBND + ibnd) * CJ_N;
    in.cf = 1.0, in.ct = o[0], in.cw = o[1], in.cg = o[2];
  }
  return in;
}

// One step of the back substitution and its adjoint. With the sweep state s at the top of layer
// j and its solution l: C_j = P rho, E_j = td rho (rho = 1 / (1 - rd prdnd), P = rd a + r tdbt,
// a = ztdn - tdbt) give U_j = C_j + E_j U_(j+1) (u: U_(j+1) in, U_j out). J reaches the step
// through U_j (adjoint ub, the part carried from above), through U_j's weight in D_j = ztdn +
// prdnd U_j (zwD = zinc wD_j, adjoints A = zwD of ztdn and B = zwD U_j of prdnd) and through the
// new state (adjoint sb, replaced by the adjoint of s). Returns the adjoints of the layer's
// r, rd, t, td, dbt in lb[0..4].
__device__ inline void sweep_step_adj(const Sweep &s, const RT4 &l, double dbt, double ub, double zwD,
                                      double &u, Sweep &sb, double *lb) {
  const double t = s.tdbt, z = s.ztdn, p = s.prdnd;
  const double a = z - t;
  const double den = 1. - l.rd * p, rho = frcp1(den);
  const double P = l.rd * a + l.r * t;
  const double Q = a + t * l.r * p;
  const double Uj1 = u;
  u = P * rho + l.td * rho * Uj1;
  const double C_b = ub, E_b = ub * Uj1, A_b = zwD, B_b = zwD * u;
  double t_b = 0., z_b = A_b, p_b = B_b, a_b = 0., rho_b = 0.;
  double r_b = 0., rd_b = 0., tr_b = 0., td_b = 0., dbt_b = 0.;
  // tdbt' = dbt t
  t_b += sb.tdbt * dbt;
  dbt_b += sb.tdbt * t;
  // prdnd' = rd + td^2 p rho
  rd_b += sb.prdnd;
  td_b += sb.prdnd * 2. * l.td * p * rho;
  p_b += sb.prdnd * l.td * l.td * rho;
  rho_b += sb.prdnd * l.td * l.td * p;
  // ztdn' = t tr + td Q rho
  t_b += sb.ztdn * l.t;
  tr_b += sb.ztdn * t;
  td_b += sb.ztdn * Q * rho;
  double Q_b = sb.ztdn * l.td * rho;
  rho_b += sb.ztdn * l.td * Q;
  a_b += Q_b;
  t_b += Q_b * l.r * p;
  r_b += Q_b * t * p;
  p_b += Q_b * t * l.r;
  // E = td rho
  td_b += E_b * rho;
  rho_b += E_b * l.td;
  // C = P rho
  double P_b = C_b * rho;
  rho_b += C_b * P;
  rd_b += P_b * a;
  a_b += P_b * l.rd;
  r_b += P_b * t;
  t_b += P_b * l.r;
  // rho = 1 / (1 - rd p)
  double den_b = -rho_b * rho * rho;
  rd_b -= den_b * p;
  p_b -= den_b * l.rd;
  // a = z - t
  z_b += a_b;
  t_b -= a_b;
  sb.tdbt = t_b, sb.ztdn = z_b, sb.prdnd = p_b;
  lb[0] = r_b, lb[1] = rd_b, lb[2] = tr_b, lb[3] = td_b, lb[4] = dbt_b;
}

// Adjoint of inatm_sw's layer amounts and setcoef_sw for one layer (src/rrtmg_sw_rad.f90:1531-1586,
// src/rrtmg_sw_setcoef.f90:163-303) onto T, H2O and O3 given the summed adjoints of its outputs.
__device__ static void setcoef_adj(const RtArgs &a, int c, int lay, const double *acc, double &t_b,
                                   double &q_b, double &o3_b) {
  const int C = a.ncol;
  const double amd = 28.9660, amw = 18.0160, avogad = 6.02214199e+23, grav = 9.8066;
  double p = a.play[c + (size_t)C * lay], t = a.tlay[c + (size_t)C * lay];
  double pz0 = a.plev[c + (size_t)C * lay], pz1 = a.plev[c + (size_t)C * (lay + 1)];
  double q = a.h2ovmr[c + (size_t)C * lay];
  double co2 = a.co2vmr[c + (size_t)C * lay], o3 = a.o3vmr[c + (size_t)C * lay];
  double ch4 = a.ch4vmr[c + (size_t)C * lay], o2 = a.o2vmr[c + (size_t)C * lay];
  double amm = (1. - q) * amd + q * amw;
  double coldry = (pz0 - pz1) * 1.e3 * avogad / (1.e2 * grav * amm * (1. + q));
  double plog = log(p);
  int jp = (int)(36. - 5 * (plog + 0.04));
  jp = jp < 1 ? 1 : (jp > 58 ? 58 : jp);
  double fp = 5. * (PREFLOG[jp - 1] - plog);
  double water = q;   // w_h2o / coldry
  double scalefac = p * (296. / 1013.) / t;
  double forfac = scalefac / (1. + water);
  t_b = 0.0, q_b = 0.0, o3_b = 0.0;
  // fac00..fac11 through ft, ft1 (d/dT = 1/15)
  t_b += (-(1. - fp) * acc[AC_FAC + 0] + (1. - fp) * acc[AC_FAC + 2] - fp * acc[AC_FAC + 1] +
          fp * acc[AC_FAC + 3]) / 15.;
  double forfac_b = acc[AC_FORFAC], water_b = 0.0;
  if (plog > 4.56) {
    t_b += acc[AC_SELFFRAC] / 7.2 - acc[AC_FORFRAC] / 36.;
    // selffac = water forfac
    water_b += acc[AC_SELFFAC] * forfac;
    forfac_b += acc[AC_SELFFAC] * water;
  } else {
    t_b += acc[AC_FORFRAC] / 36.;
  }
  // forfac = scalefac / (1 + water)
  double sf_b = forfac_b / (1. + water);
  water_b -= forfac_b * forfac / (1. + water);
  t_b -= sf_b * scalefac / t;
  q_b += water_b;
  // column amounts: colX = 1e-20 coldry vmrX (zero amounts replaced by 1e-32 coldry)
  double coldry_b = 1.e-20 * acc[AC_COL + 5];                    // colmol = 1e-20 coldry + colh2o
  double h2o_b = acc[AC_COL + SP_H2O] + acc[AC_COL + 5];
  coldry_b += 1.e-20 * h2o_b * q;
  q_b += 1.e-20 * h2o_b * coldry;
  coldry_b += 1.e-20 * acc[AC_COL + SP_O3] * o3;
  o3_b += 1.e-20 * acc[AC_COL + SP_O3] * coldry;
  coldry_b += (co2 * coldry != 0. ? 1.e-20 * co2 : 1.e-32) * acc[AC_COL + SP_CO2];
  coldry_b += (ch4 * coldry != 0. ? 1.e-20 * ch4 : 1.e-32) * acc[AC_COL + SP_CH4];
  coldry_b += (o2 * coldry != 0. ? 1.e-20 * o2 : 1.e-32) * acc[AC_COL + SP_O2];
  // coldry = K dp / (amm (1 + q))
  q_b -= coldry_b * coldry * ((amw - amd) / amm + 1. / (1. + q));
}

// A layer's record for the reverse sweep, per g-point: the sweep state at its top (tdbt, ztdn,
// prdnd) and its gas optics (taug, taur). The reverse sweep forms the layer's two-stream
// solution again from them, keeping the intermediates its own reverse needs, and leaves the
// adjoints of taug and taur in place of the gas optics for the gas-optics pass that follows.
// Doubles of work area per column, times NGPT: UB (L), then in store mode every layer's record
// (5 L); in checkpoint mode every layer's gas optics (2 L), the sweep state at the top of every
// K-th layer (3 ceil(L/K)) and the sweep states of one K-layer segment (3 K). Then the cloud
// table CJ (L, NBND, CJ_N).
constexpr int REC_S = 5;
__host__ __device__ inline size_t adj_work_doubles(int L, int K) {
  size_t states = K == 0 ? (size_t)L * REC_S : (size_t)L * 2 + (size_t)((L + K - 1) / K) * 3 + (size_t)K * 3;
  return ((size_t)L + states) * NGPT + (size_t)L * NBND * CJ_N;
}

// CKPT: checkpoint mode (every K-th sweep state kept, K = A.ckpt >= 1), else store mode.
template <bool CKPT>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2))) adj_kernel(AdjArgs A, DeviceTables T) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const RtArgs &a = A.f;
  const int L = a.nlay, C = a.ncol;
  const int nmw = (L + 31) / 32;
  ColLayout lo(L, nmw, false, false, false);
  ColView v(smem, lo);
  double *acc = reinterpret_cast<double *>(smem + ((lo.total + 15) & ~15));   // (L, AC_N)
  double *wU = acc + (size_t)L * AC_N, *wD = wU + (L + 1);     // level weights, top-based
  double *albacc = wD + (L + 1);                                // asdir asdif aldir aldif
  const int per = (a.ncol_chunk + 7) / 8;
  const int cl = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (cl >= a.ncol_chunk) return;
  const int c = a.col0 + cl;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int g = wave * WAVE_G + lane;
  const bool active = lane < WAVE_G;

  // ---- prologue (as the forward) + adjoint accumulators and level weights ----
  if (tid == 0) v.misc[M_LAYTROP] = 0, v.misc[M_LAYLOW] = 0, v.misc[M_CLDTOP] = -1;
  for (int i = tid; i < L * AC_N; i += blockDim.x) acc[i] = 0.0;
  if (tid < 4) albacc[tid] = 0.0;
  __syncthreads();
  for (int lay = tid; lay < L; lay += blockDim.x) {
    setcoef_layer(a, T, c, lay, v, &v.misc[M_LAYTROP]);
    if (a.sc.icld >= 1 && a.cldfrac[c + (size_t)C * lay] >= CLDMIN) atomicMax(&v.misc[M_CLDTOP], lay);
  }
  for (int j = tid; j <= L; j += blockDim.x) {
    // top-based level j is bottom-based level i = L - j; swhr(k) = (net(k+1) - net(k)) hf / pdp(k)
    // for k < L - 1 (the top layer's heating rate is zero)
    const int i = L - j;
    double wnet = 0.0;
    if (i >= 1 && i - 1 < L - 1) {
      double pdp = a.plev[c + (size_t)C * (i - 1)] - a.plev[c + (size_t)C * i];
      wnet += A.seed_hr[c + (size_t)C * (i - 1)] * a.sc.heatfac / pdp;
    }
    if (i < L - 1) {
      double pdp = a.plev[c + (size_t)C * i] - a.plev[c + (size_t)C * (i + 1)];
      wnet -= A.seed_hr[c + (size_t)C * i] * a.sc.heatfac / pdp;
    }
    wD[j] = wnet;
    wU[j] = -wnet + (j == 0 ? A.seed_toa[c] : 0.0);
  }
  __syncthreads();
  const int laytrop = v.misc[M_LAYTROP];
  const int first[NBND + 1] = {0, 6, 18, 26, 34, 44, 54, 56, 66, 74, 80, 86, 94, 100, NGPT};   // g of each band
  if (tid < NBND) {
    GPoint gb = T.gp[first[tid]];
    v.misc[M_LAYSOL + tid] = find_laysolfr(gb, v, laytrop, L);
  }
  int err = 0;
  if (a.sc.icld >= 1 && active)
    // the masks read the cloud fractions from memory: staged in LDS they would take the
    // workgroup past the 40 KB at which four fit a CU at 137 layers
    err = column_masks(a.mc_icld, a.mc_irng, a.permuteseed, C, L, c, g, a.play, a.cldfrac + c,
                       a.alpha ? a.alpha + c : nullptr, (size_t)C,
                       a.mt_stream, reinterpret_cast<const KissJump *>(T.kiss_jump), v.mask, v.misc[M_CLDTOP]);
  __syncthreads();
  // cloud optics and their water-path derivatives for each (layer, band) with a cloudy sub-column
  const int K = CKPT ? A.ckpt : 0;
  double *UB = A.work + (size_t)cl * adj_work_doubles(L, K);
  double *CJ = UB + adj_work_doubles(L, K) - (size_t)L * NBND * CJ_N;
  if (a.sc.icld >= 1) {
    const int nlc = a.mc_irng == 0 ? v.misc[M_CLDTOP] + 1 : L;   // KISS masks: clear above the cloud top
    for (int idx = tid; idx < nlc * NBND; idx += blockDim.x) {
      const int lay = idx / NBND, b = idx - lay * NBND;
      uint32_t any = 0;
      for (int gg = first[b]; gg < first[b + 1]; gg++) any |= v.mask[(lay >> 5) * NGPT + gg];
      if ((any >> (lay & 31)) & 1u) {
        const size_t q1 = c + (size_t)C * lay, qb = b + NBND * q1;
        const double ciwp = a.ciwp[q1], clwp = a.clwp[q1], rei = a.rei[q1], rel = a.rel[q1];
        CloudOpt co = cloud_optics(T, a.sc, b + 16, 1.0, ciwp, clwp, rei, rel, a.tauc[qb], a.ssac[qb],
                                   a.asmc[qb], a.fsfc[qb], err);
        double *o = CJ + (size_t)idx * CJ_N;
        o[0] = co.tau, o[1] = co.ssa, o[2] = co.asm_;
        for (int k = 0; k < 3; k++) {
          double lw_b, iw_b;
          cloud_optics_adj(T, a.sc, b + 16, 1.0, ciwp, clwp, rei, rel, k == 0 ? 1. : 0., k == 1 ? 1. : 0.,
                           k == 2 ? 1. : 0., lw_b, iw_b);
          o[3 + k] = lw_b, o[6 + k] = iw_b;
        }
      }
    }
    __syncthreads();
  }

  // the lane's g-point record is read field by field where used (a register copy held across
  // the layer loops would take ~40 VGPRs per lane)
  const GPoint &gp = T.gp[active ? g : 0];
  const double mu0 = fmax(a.coszen[c], 1.e-10);
  const double rmu = 1.0 / mu0;
  const bool vis = gp.ibnd >= 9 && gp.ibnd <= 12;
  const double albdir = vis ? a.asdir[c] : a.aldir[c];
  const double albdif = vis ? a.asdif[c] : a.aldif[c];
  const double zinc = active ? incident_flux(gp, T.tab, v, a.sc, mu0) : 0.0;

  // Per-column work area: UB (L, NGPT) = dJ/dU_j, then either every layer's record
  // S (L, REC_S, NGPT) (store mode) or every layer's gas optics GT (L, 2, NGPT), one sweep state
  // per K layers CK (ceil(L/K), 3, NGPT) and the sweep states of one K-layer segment
  // SG (K, 3, NGPT) (checkpoint mode, K = A.ckpt).
  double *S = UB + (size_t)L * NGPT;
  double *GT = S;
  double *CK = GT + (size_t)L * 2 * NGPT;
  double *SG = CK + (size_t)((L + K - 1) / max(K, 1)) * 3 * NGPT;
  auto ck = [&](int i, int k) -> double & { return CK[((size_t)i * 3 + k) * NGPT + g]; };
  // the lane's sweep state for layer j (fields NGPT apart): layer j of S in store mode, slot
  // j - jseg0 of the current segment in SG in checkpoint mode (a uniform base plus the lane)
  auto statep = [&](int i) -> double * { return (CKPT ? SG + (size_t)i * 3 * NGPT : S + (size_t)i * REC_S * NGPT) + g; };
  // the lane's gas optics of layer j: in its record (store mode) or in GT (checkpoint mode)
  auto gasp = [&](int j) -> double * { return (CKPT ? GT + (size_t)j * 2 * NGPT : S + ((size_t)j * REC_S + 3) * NGPT) + g; };
  auto put_state = [&](int i, const Sweep &s) {
    double *r = statep(i);
    r[0] = s.tdbt, r[NGPT] = s.ztdn, r[2 * NGPT] = s.prdnd;
  };

  const int jsplit = L - laytrop;   // layers j >= jsplit lie below laytrop

  // ---- F1: top -> bottom: sweep states and ubar_j = dJ/dU_j. The back substitution
  // U_j = C_j + E_j U_(j+1) feeds J through U_j and D_j = ztdn_j + prdnd_j U_j, so
  // ubar_j = ubar_(j-1) E_(j-1) + zinc (wU_j + wD_j prdnd_j), known as soon as layer j is. ----
  Sweep st = {1.0, 1.0, 0.0};
  double carry = 0.0;
  // A layer's record and ubar are stored once the next layer's table reads have been issued (the
  // vector memory counter retires in issue order, so earlier stores would lengthen those waits).
  double pr[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int pj = -1;
  auto flush_rec = [&]() {
    if (active && pj >= 0) {
      if (!CKPT) {
        double *r = statep(pj);
        r[0] = pr[0], r[NGPT] = pr[1], r[2 * NGPT] = pr[2];
      }
      double *t = gasp(pj);
      t[0] = pr[3], t[NGPT] = pr[4];
      UB[(size_t)pj * NGPT + g] = pr[5];
    }
  };
  int jk = 0, ci = 0;   // checkpoint mode: layer within the segment, checkpoint index
  auto f1_layer = [&](auto lower_tag, const GasLane &q, int j) {
    constexpr bool LOWER = decltype(lower_tag)::value;
    if (CKPT) {
      if (jk == 0 && active) ck(ci, 0) = st.tdbt, ck(ci, 1) = st.ztdn, ck(ci, 2) = st.prdnd;
      ci += jk == 0;
      jk = jk + 1 == K ? 0 : jk + 1;
    }
    const int lay = L - 1 - j;
    const GasFetch<double> gf = gas_fetch<LOWER, double>(q, T.tab, v, lay);
    flush_rec();
    double taug, taur;
    gas_finish<LOWER, double>(q, gf, v, lay, taug, taur);
    const LayerIn in = layer_inputs(a, T, gp.ibnd, v, CJ, c, g, lay, active);
    CellTape tp;
    const LayerV ly = cell_layer(taug, taur, in, mu0, rmu, tp);
    pr[0] = st.tdbt, pr[1] = st.ztdn, pr[2] = st.prdnd, pr[3] = taug, pr[4] = taur;
    RT4 l = {ly.r, ly.rd, ly.t, ly.td};
    double Aa, Bb, Cc, Ee;
    sweep_layer(st, l, ly.dbt, Aa, Bb, Cc, Ee);
    double ub = carry + zinc * (wU[j] + wD[j] * Bb);
    pr[5] = ub, pj = j;
    carry = ub * Ee;
  };
  {
    const GasLane q = gas_lane<false>(gp, L);
    for (int j = 0; j < jsplit; j++) f1_layer(std::integral_constant<bool, false>(), q, j);
  }
  {
    const GasLane q = gas_lane<true>(gp, L);
    for (int j = jsplit; j < L; j++) f1_layer(std::integral_constant<bool, true>(), q, j);
  }
  flush_rec();
  if (err) atomicCAS(a.status, 0, err);

  // ---- surface: U_L = (adir tdbt + adif (ztdn - tdbt)) / (1 - adif prdnd), D_L = ztdn + prdnd U_L ----
  const double sig = 1. / (1. - albdif * st.prdnd);
  const double Nsf = albdir * st.tdbt + albdif * (st.ztdn - st.tdbt);
  double u = Nsf * sig;                                   // U_L, then U_j going up
  double phi = wU[L] * u + wD[L] * (st.ztdn + st.prdnd * u);
  Sweep sb = {0.0, 0.0, 0.0};
  {
    double ubL = carry + zinc * (wU[L] + wD[L] * st.prdnd);
    sb.ztdn += zinc * wD[L];
    sb.prdnd += zinc * wD[L] * u;
    double N_b = ubL * sig, sig_b = ubL * Nsf;
    double adir_b = N_b * st.tdbt, adif_b = N_b * (st.ztdn - st.tdbt);
    sb.tdbt += N_b * (albdir - albdif);
    sb.ztdn += N_b * albdif;
    double den_b = -sig_b * sig * sig;
    adif_b -= den_b * st.prdnd;
    sb.prdnd -= den_b * albdif;
    // reduce over g: vis bands -> asdir / asdif, the others -> aldir / aldif
    double s = wave_sum4d(vis ? adir_b : 0.0, vis ? adif_b : 0.0, vis ? 0.0 : adir_b, vis ? 0.0 : adif_b);
    if ((lane & 15) == 0) atomicAdd(&albacc[lane >> 4], s);
  }

  // ---- R1: bottom -> top: the reverse of the two-stream layers and of the elimination; the
  // gas optics' adjoints wait in the layers' slots for R2. ----
  // Checkpoint mode: the sweep states of the segment jseg0..j, re-swept from its checkpoint with
  // the forward sweep's own functions on the gas optics F1 kept (the same values as F1's, so the
  // same as store mode's); a state is stored once the next layer's gas optics have been read.
  ci = CKPT ? (L - 1) / max(K, 1) : 0;
  int jseg0 = ci * K;
  auto resweep = [&](int j) {
    Sweep ss = {active ? ck(ci, 0) : 1.0, active ? ck(ci, 1) : 1.0, active ? ck(ci, 2) : 0.0};
    // two layers a step: their two-stream solutions are independent, only the sweep is serial
    // (measured 1.5 % faster at 137 layers than one a step; three or four are no faster)
#pragma unroll 2
    for (int jj = jseg0; jj < j; jj++) {
      const double *t = gasp(jj);
      const double taug = active ? t[0] : 1., taur = active ? t[NGPT] : 0.1;
      if (active) put_state(jj - jseg0, ss);
      const LayerIn in = layer_inputs(a, T, gp.ibnd, v, CJ, c, g, L - 1 - jj, active);
      CellTape tp;
      const LayerV ly = cell_layer(taug, taur, in, mu0, rmu, tp);
      RT4 l = {ly.r, ly.rd, ly.t, ly.td};
      double Aa, Bb, Cc, Ee;
      sweep_layer(ss, l, ly.dbt, Aa, Bb, Cc, Ee);
    }
    if (active) put_state(j - jseg0, ss);
  };
  for (int j = L - 1; j >= 0; j--) {
    if (CKPT && (j == L - 1 || j < jseg0)) {
      if (j < jseg0) ci -= 1, jseg0 -= K;
      resweep(j);
    }
    // the layer's record (sweep state at its top, gas optics); its two-stream solution formed
    // again with the intermediates the reverse needs (taped)
    const double *r = statep(CKPT ? j - jseg0 : j), *t = gasp(j);
    const Sweep sj = {active ? r[0] : 1.0, active ? r[NGPT] : 1.0, active ? r[2 * NGPT] : 0.0};
    const double taug = active ? t[0] : 1., taur = active ? t[NGPT] : 0.1;
    const double ub = active ? UB[(size_t)j * NGPT + g] : 0.0;
    const int lay = L - 1 - j;
    const LayerIn in = layer_inputs(a, T, gp.ibnd, v, CJ, c, g, lay, active);
    CellTape tp;
    const LayerV ly = cell_layer(taug, taur, in, mu0, rmu, tp);
    const RT4 l = {ly.r, ly.rd, ly.t, ly.td};
    // U_j by the back substitution, the objective per unit incident flux, and the step's adjoint
    const double zwD = zinc * wD[j];
    double lb[5], xb[5];
    sweep_step_adj(sj, l, ly.dbt, ub, zwD, u, sb, lb);
    phi += wU[j] * u + wD[j] * (sj.ztdn + sj.prdnd * u);
    cell_layer_adj(tp, taug, taur, in, mu0, rmu, lb, xb);
    // the adjoints of taug and taur replace the layer's gas optics in its slot, for R2
    if (active) {
      double *t = gasp(j);
      t[0] = xb[0], t[NGPT] = xb[1];
    }
    double lw = 0.0, li = 0.0;
    if (active && in.cf > 0.0) {
      const double *o = CJ + ((size_t)lay * NBND + gp.ibnd) * CJ_N;
      lw = xb[2] * o[3] + xb[3] * o[4] + xb[4] * o[5];
      li = xb[2] * o[6] + xb[3] * o[7] + xb[4] * o[8];
    }
    static_assert(AC_CIWP == AC_CLWP + 1, "water paths adjacent");
    const double s = wave_sum2d(lw, li);
    // one term per wave and entry, onto zero: a + b rounds as b + a, whichever wave is first
    if ((lane & 31) == 0) atomicAdd(&acc[(size_t)lay * AC_N + AC_CLWP + (lane >> 5)], s);
  }

  // ---- R2: the gas-optics adjoint of every layer (onto the setcoef outputs), from the adjoints
  // of taug and taur R1 left in the layers' slots. Its layers are independent of one another, so
  // their table reads overlap; one regime at a time with the lane's constants resolved once. ----
  auto r2 = [&](auto lower_tag, int j0, int j1) {
    constexpr bool LOWER = decltype(lower_tag)::value;
    if (j0 >= j1) return;
    const GasLane q = gas_lane<LOWER>(gp, L);
    for (int j = j0; j < j1; j++) {
      const int lay = L - 1 - j;
      double la[AC_N];
#pragma unroll
      for (int k = 0; k < AC_N; k++) la[k] = 0.0;
      if (active) {
        const double *t = gasp(j);
        gas_adj_r<LOWER>(q, T.tab, v, lay, t[0], t[NGPT], la);
      }
      static_assert(AC_N == 16 && AC_CLWP == 14, "one wave_sum16d per layer, gas terms in 0..13");
      const double s = wave_sum16d(la, lane);
      if ((lane & 3) == 0 && (lane >> 2) < AC_CLWP) atomicAdd(&acc[(size_t)lay * AC_N + (lane >> 2)], s);
    }
  };
  r2(std::integral_constant<bool, false>(), 0, jsplit);
  r2(std::integral_constant<bool, true>(), jsplit, L);

  // ---- incident flux: zinc = adjflux src mu0, src interpolated in the key-species parameter
  // at laysolfr for some bands (src/rrtmg_sw_taumol.f90, e.g. :495-520). The g-points of a band
  // add to the same two accumulators (laysolfr and the key species are band constants), and
  // bands can share one. Each wave sums its bands over lanes in lane order (shuffles), and once
  // both waves' layer terms are in, wave 0 then wave 1 adds its band sums in band order: the
  // rounding does not depend on wave timing, and no LDS beyond the accumulators is needed. ----
  double c1_b = 0.0, c2_b = 0.0;
  int t1 = -1, t2 = -1;
  if (active && gp.sol_interp) {
    const int ls = v.misc[M_LAYSOL + gp.ibnd] - 1;
    const int s1 = gp.sol_upper ? gp.hi_s1 : gp.lo_s1, s2 = gp.sol_upper ? gp.hi_s2 : gp.lo_s2;
    const double strrat = gp.sol_upper ? gp.hi_strrat : gp.lo_strrat, mult = gp.sol_upper ? 4.0 : 8.0;
    Spec s = key_spec(v.c(s1, ls), v.c(s2, ls), strrat, mult, true);
    auto dif = [&](int off) { return T.tab[off + s.js * gp.ssf] - T.tab[off + (s.js - 1) * gp.ssf]; };
    double dsrc;
    if (a.sc.isolvar < 0) {
      dsrc = gp.sol_scale * dif(gp.sflux);
    } else if (a.sc.isolvar == 3) {
      dsrc = a.sc.svar_bnd[0][gp.ibnd] * dif(gp.facb) + a.sc.svar_bnd[1][gp.ibnd] * dif(gp.snsp) +
             a.sc.svar_bnd[2][gp.ibnd] * dif(gp.irrad);
    } else {
      dsrc = a.sc.svar_f * dif(gp.facb) + a.sc.svar_s * dif(gp.snsp) + a.sc.svar_i * dif(gp.irrad);
    }
    double fs_b = phi * a.sc.adjflux[gp.ibnd] * mu0 * dsrc;
    key_spec_adj(v.c(s1, ls), v.c(s2, ls), strrat, mult, true, fs_b, 0.0, c1_b, c2_b);
    t1 = ls * AC_N + AC_COL + s1, t2 = ls * AC_N + AC_COL + s2;
  }
  __syncthreads();
  for (int w = 0; w < 2; w++) {
    if (wave == w) {
      for (int b = w * (NBND / 2); b < (w + 1) * (NBND / 2); b++) {   // bands of this wave
        const int l0 = first[b] - w * WAVE_G, l1 = first[b + 1] - w * WAVE_G;
        double s1 = 0.0, s2 = 0.0;
        for (int l = l0; l < l1; l++) s1 += __shfl(c1_b, l), s2 += __shfl(c2_b, l);
        const int to1 = __shfl(t1, l0), to2 = __shfl(t2, l0);
        if (lane == 0 && s1 != 0.0) acc[to1] += s1;
        if (lane == 0 && s2 != 0.0) acc[to2] += s2;
      }
    }
    __syncthreads();
  }

  // ---- per layer: adjoint of inatm_sw + setcoef_sw; outputs ----
  for (int lay = tid; lay < L; lay += blockDim.x) {
    double tb, qb, ob;
    setcoef_adj(a, c, lay, acc + (size_t)lay * AC_N, tb, qb, ob);
    size_t q = c + (size_t)C * lay;
    A.d_tlay[q] = tb;
    A.d_h2ovmr[q] = qb;
    A.d_o3vmr[q] = ob;
    A.d_clwp[q] = acc[(size_t)lay * AC_N + AC_CLWP];
    A.d_ciwp[q] = acc[(size_t)lay * AC_N + AC_CIWP];
  }
  if (tid < 4) A.d_alb[c + (size_t)C * tid] = albacc[tid];
}

size_t adj_lds_bytes(int nlay) {
  ColLayout lo(nlay, (nlay + 31) / 32, false, false, false);
  return ((lo.total + 15) & ~15) + sizeof(double) * ((size_t)nlay * AC_N + 2 * (nlay + 1) + 4);
}

size_t adj_work_bytes(int nlay, int ckpt) { return adj_work_doubles(nlay, ckpt) * sizeof(double); }

hipError_t launch_adjoint(const AdjArgs &a, const DeviceTables &t, hipStream_t s) {
  int per = (a.f.ncol_chunk + 7) / 8;
  if (a.ckpt > 0)
    hipLaunchKernelGGL(adj_kernel<true>, dim3(per * 8), dim3(128), adj_lds_bytes(a.f.nlay), s, a, t);
  else
    hipLaunchKernelGGL(adj_kernel<false>, dim3(per * 8), dim3(128), adj_lds_bytes(a.f.nlay), s, a, t);
  return hipGetLastError();
}

}  // namespace rrtmgsw
