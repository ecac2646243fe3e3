// rrtmgsw_kernels.hip -- CDNA4 (gfx950) kernels of the GCM-mode McICA shortwave path.
//
// Mapping: one workgroup per column, 128 threads = 2 wavefronts, one lane per g-point.
// Wave 0 holds g-points 1-56 (bands 16-22), wave 1 holds 57-112 (bands 23-29); lanes 56-63
// of each wave idle but take part in the cross-lane reductions. Column-level data (inatm_sw /
// setcoef_sw results, sub-column masks, flux sums) live in LDS; per-g work is a register-resident
// loop over layers:
//   prologue  inatm_sw + setcoef_sw for every layer of the column (lanes = layers), laytrop,
//             laysolfr of each band (src/rrtmg_sw_rad.f90:1531-1586, src/rrtmg_sw_setcoef.f90)
//   masks     (fused entry) McICA sub-column cloud masks for the lane's g-point, bottom to top
//             (src/mcica_subcol_gen_sw.f90:468-659); the non-McICA entry instead takes each
//             layer's cloud fraction for every g-point (src/rrtmg_sw_spcvrt.f90:429, :453-454)
//   pass 1    top to bottom: gas optics (one data-driven template for the fourteen taumolNN
//             routines), cloud optics (cldprmc_sw), delta scaling, two-stream layer
//             reflectance / transmittance (reftra_sw), direct beam, and the downward half of
//             the adding method (vrtqdr_sw :144-159), storing four elimination coefficients
//             per layer and sky
//   pass 2    bottom to top: upward fluxes by back substitution, level fluxes for both skies,
//             weighted by the g-point's incident flux and summed over g with wavefront
//             reductions (spcvmc_sw :607-622)
//   epilogue  broadband fluxes and heating rates (src/rrtmg_sw_rad.f90:805-836)
#include <type_traits>

#include "rrtmgsw_device.h"

namespace rrtmgsw {

// Broadband fluxes at the nlay+1 levels and heating rates (src/rrtmg_sw_rad.f90:805-836,
// top layer 0) from the per-level sums over g of both waves (red, reda: in LDS or the column's
// scratch).
__device__ inline void rt_epilogue(const RtArgs &a, const double *red, const double *reda, int c) {
  const int L = a.nlay, C = a.ncol, tid = threadIdx.x;
  // level j (top-based): up = red[.][0], down = red[.][1] (+ reda total), clear in 2, 3 (+ reda clear)
  auto level = [&](int j, double &fu, double &fd, double &cu, double &cd) {
    const double *r0 = red + (0 * (L + 1) + j) * 4, *r1 = red + (1 * (L + 1) + j) * 4;
    fu = r0[0] + r1[0], fd = r0[1] + r1[1], cu = r0[2] + r1[2], cd = r0[3] + r1[3];
    if (j < L) {
      fd += reda[(0 * L + j) * 2 + 0] + reda[(1 * L + j) * 2 + 0];
      cd += reda[(0 * L + j) * 2 + 1] + reda[(1 * L + j) * 2 + 1];
    }
  };
  for (int j = tid; j <= L; j += blockDim.x) {
    double fu, fd, cu, cd;
    level(j, fu, fd, cu, cd);
    size_t q = c + (size_t)C * (L - j);
    a.swuflx[q] = fu, a.swdflx[q] = fd, a.swuflxc[q] = cu, a.swdflxc[q] = cd;
  }
  for (int i = tid; i < L; i += blockDim.x) {
    // layer i (bottom-based) lies between output levels i and i+1 = top-based j and j-1
    int j = L - i;
    double fu0, fd0, cu0, cd0, fu1, fd1, cu1, cd1;
    level(j, fu0, fd0, cu0, cd0);
    level(j - 1, fu1, fd1, cu1, cd1);
    double pdp = a.plev[c + (size_t)C * i] - a.plev[c + (size_t)C * (i + 1)];
    double zdpgcp = a.sc.heatfac / pdp;
    size_t q = c + (size_t)C * i;
    a.swhr[q] = i == L - 1 ? 0.0 : ((fd1 - fu1) - (fd0 - fu0)) * zdpgcp;
    a.swhrc[q] = i == L - 1 ? 0.0 : ((cd1 - cu1) - (cd0 - cu0)) * zdpgcp;
  }
}

// ------------------------------------------------------------------------------------------
// Pass 2 and epilogue of the mixed form (FP32 state and coefficients, FP32 sums over g).
// Level j (top-based) gets the sums of zinc U, zinc B U (clear and total) and layer j the rest
// of its absorbed flux, zinc ad (B_j U_j + U_(j+1)); the heating rate of a layer is heatfac / pdp
// times its absorbed flux (pass 1's part + this one) summed over g: the layer's net-flux
// divergence formed without the difference of two nearly equal fluxes.
// ------------------------------------------------------------------------------------------
__device__ inline void rt_pass2_mixed(const RtArgs &a, float *redf, const float *redaf, int c,
                                      const double *scr, int gi, int lane, int wave, int ibnd,
                                      float zinc, int jfirst, const SweepF &scf, const SweepF &stf) {
  const int L = a.nlay, C = a.ncol;
  const bool vis = ibnd >= 9 && ibnd <= 12;
  const float albdir = float(vis ? a.asdir[c] : a.aldir[c]);
  const float albdif = float(vis ? a.asdif[c] : a.aldif[c]);
  float uc = (albdir * scf.tdbt + albdif * (scf.ztdn - scf.tdbt)) * rcpf(1.f - albdif * scf.prdnd);
  float ut = (albdir * stf.tdbt + albdif * (stf.ztdn - stf.tdbt)) * rcpf(1.f - albdif * stf.prdnd);
  {
    const float dc = scf.ztdn + scf.prdnd * uc, dt = stf.ztdn + stf.prdnd * ut;
    const float s = wave_sum4f(zinc * ut, zinc * dt, zinc * uc, zinc * dc);
    float *r = redf + (wave * (L + 1) + L) * 6;
    if ((lane & 15) == 0) r[lane >> 4] = s;
    if (lane == 0) r[4] = 0.f, r[5] = 0.f;
  }
  struct Co {
    float cB, cC, cE, cA, tB, tC, tE, tA;
  };
  auto fetch = [&](int j) {
    Co o;
    const int jj = j < 0 ? 0 : j;
    const float *p = reinterpret_cast<const float *>(scr - gi + (size_t)jj * SCR_COEFS * NGPT) + gi;
    o.cB = p[0], o.cC = p[NGPT], o.cE = p[2 * NGPT], o.cA = p[3 * NGPT];
    o.tB = 0.f, o.tC = 0.f, o.tE = 0.f, o.tA = 0.f;
    if (jj >= jfirst) o.tB = p[4 * NGPT], o.tC = p[5 * NGPT], o.tE = p[6 * NGPT], o.tA = p[7 * NGPT];
    return o;
  };
  auto step = [&](int j, const Co &o) {
    const bool tot = j >= jfirst;
    const float ub_t = ut, ub_c = uc;   // upward flux at the layer's bottom
    ut = tot ? o.tC + o.tE * ut : o.cC + o.cE * ut;
    const float bu_t = (tot ? o.tB : o.cB) * ut, ad_t = tot ? o.tA : o.cA;
    uc = o.cC + o.cE * uc;
    const float bu_c = o.cB * uc;
    const float s4 = wave_sum4f(zinc * ut, zinc * bu_t, zinc * uc, zinc * bu_c);
    const float s2 = wave_sum2f(zinc * (ad_t * (bu_t + ub_t)), zinc * (o.cA * (bu_c + ub_c)));
    float *r = redf + (wave * (L + 1) + j) * 6;
    if ((lane & 15) == 0) r[lane >> 4] = s4;
    if ((lane & 31) == 0) r[4 + (lane >> 5)] = s2;
  };
  Co r0 = fetch(L - 1), r1 = fetch(L - 2), r2 = fetch(L - 3), r3 = fetch(L - 4);
  int j = L - 1;
  for (; j >= 3; j -= 4) {
    step(j, r0);
    r0 = fetch(j - 4);
    step(j - 1, r1);
    r1 = fetch(j - 5);
    step(j - 2, r2);
    r2 = fetch(j - 6);
    step(j - 3, r3);
    r3 = fetch(j - 7);
  }
  if (j >= 0) step(j, r0);
  if (j >= 1) step(j - 1, r1);
  if (j >= 2) step(j - 2, r2);
  __syncthreads();
  // epilogue: fluxes at the nlay+1 levels; heating rates from the layers' absorbed fluxes
  const int tid = threadIdx.x;
  for (int jl = tid; jl <= L; jl += blockDim.x) {
    const float *q0 = redf + (0 * (L + 1) + jl) * 6, *q1 = redf + (1 * (L + 1) + jl) * 6;
    double fu = double(q0[0]) + q1[0], fd = double(q0[1]) + q1[1];
    double cu = double(q0[2]) + q1[2], cd = double(q0[3]) + q1[3];
    if (jl < L) {
      const float *p0 = redaf + (0 * L + jl) * 4, *p1 = redaf + (1 * L + jl) * 4;
      fd += double(p0[0]) + p1[0];
      cd += double(p0[1]) + p1[1];
    }
    const size_t q = c + (size_t)C * (L - jl);
    a.swuflx[q] = fu, a.swdflx[q] = fd, a.swuflxc[q] = cu, a.swdflxc[q] = cd;
  }
  for (int i = tid; i < L; i += blockDim.x) {
    const int jl = L - 1 - i;   // layer i (bottom-based) is layer jl from the top
    const float *q0 = redf + (0 * (L + 1) + jl) * 6, *q1 = redf + (1 * (L + 1) + jl) * 6;
    const float *p0 = redaf + (0 * L + jl) * 4, *p1 = redaf + (1 * L + jl) * 4;
    const double abs_t = (double(p0[2]) + p1[2]) + (double(q0[4]) + q1[4]);
    const double abs_c = (double(p0[3]) + p1[3]) + (double(q0[5]) + q1[5]);
    const double pdp = a.plev[c + (size_t)C * i] - a.plev[c + (size_t)C * (i + 1)];
    const double zdpgcp = a.sc.heatfac / pdp;
    const size_t q = c + (size_t)C * i;
    a.swhr[q] = i == L - 1 ? 0.0 : abs_t * zdpgcp;
    a.swhrc[q] = i == L - 1 ? 0.0 : abs_c * zdpgcp;
  }
}

// ------------------------------------------------------------------------------------------
// FP64 forms: MODE 0 = the reference's exponential table, MODE 2 = exact exponentials (the
// smooth forward the adjoint differentiates). CLD, where the clouds come from:
//   CLD_SUBCOL  the materialised (112, ncol, nlay) McICA sub-column arrays (rrtmg_sw), cloud
//               optics per cell (cldprmc_sw);
//   CLD_FUSED   McICA masks generated here from the physical cloud inputs (mcica_subcol_sw +
//               rrtmg_sw), cloud optics per (layer, band);
//   CLD_LAYER   the non-McICA driver (src/rrtmg_sw_rad.nomcica.f90): every g-point sees the
//               layer's cloud fraction and its band's cldprop_sw optics (spcvrt_sw).
//
// Pass 1 runs the layers above laytrop and those below it as two loops, each with only its
// regime's gas-optics constants (GasLane) live. Per layer and g: gas optics, aerosol, cloud
// optics, delta-scaled clear optics (two reciprocals instead of four divisions), reftra from the
// two transmittances alone (the direct-beam one shared), the elimination step, and the flux sum
// over g by DPP.
// Pass 2 requests the coefficients of layer j-1 while layer j is processed.
// ------------------------------------------------------------------------------------------
// Occupancy target: four waves per SIMD for the table form with per-(layer, band) cloud optics
// (measured fastest, a few VGPRs spilled outside the layer loop), three for the others.
// SPEC (rt_column_kernel: the column program, with CLD_LAYER): per-g level terms written to
// a.gspec for per-band fluxes, the direct beam with or without delta scaling, per-band surface
// albedo, layer amounts given directly.
enum : int { CLD_SUBCOL = 0, CLD_FUSED = 1, CLD_LAYER = 2 };
// GSUM: the waves' per-level flux sums go to the column's scratch instead of LDS (a compile-time
// choice, so the layer loops carry no branch or extra registers for it).
template <int MODE, int CLD, bool SPEC, bool GSUM = false>
__device__ __attribute__((always_inline)) inline void rt_body(const RtArgs &a, const DeviceTables &T) {
  static_assert(!SPEC || (MODE == 0 && CLD != CLD_SUBCOL), "spectral output: FP64 layer-cloud or fused form");
  constexpr bool EXACT = MODE == 2;
  constexpr bool MIXED = MODE == 1;
  constexpr bool FUSED = CLD == CLD_FUSED;
  constexpr bool LAYERCLD = CLD == CLD_LAYER;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int L = a.nlay, C = a.ncol;
  static_assert(!MIXED || CLD == CLD_FUSED, "the mixed form is the fused entry's");
  constexpr bool gsum = GSUM;
  ColLayout lo(L, FUSED ? (L + 31) / 32 : 0, !gsum, MIXED);
  ColView v(smem, lo);
  // XCD-aware column order: consecutive columns run on the same XCD (its own L2), so the
  // strided (ncol, nlay) reads and writes of neighbouring columns share cache lines.
  const int per = (a.ncol_chunk + 7) / 8;
  const int cl = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (cl >= a.ncol_chunk) return;
  const int c = a.col0 + cl;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int g = wave * WAVE_G + lane;
  const bool active = lane < WAVE_G;
  // the waves' per-level flux sums: in LDS (v.red ...), or in the column's scratch (gsum)
  double *const gred = a.scratch + (size_t)a.ncol_chunk * scr_col_doubles(L) + (size_t)cl * scr_red_doubles(L);
  double *const greda = gred + 8 * (L + 1);
  // Idle lanes (56-63) mirror their own wave's first g-point: every scratch location a lane
  // reads in pass 2 is written by its own wave in pass 1 (there is no barrier between the passes).
  const int gi = active ? g : wave * WAVE_G;

  // ---- prologue: inatm_sw + setcoef_sw per layer, laytrop, laysolfr per band, masks ----
  if (tid == 0) v.misc[M_LAYTROP] = 0, v.misc[M_PARTIAL] = 0, v.misc[M_LAYLOW] = 0, v.misc[M_CLDTOP] = -1;
  __syncthreads();
  for (int lay = tid; lay < L; lay += blockDim.x) {
    setcoef_layer(a, T, c, lay, v, &v.misc[M_LAYTROP]);
    // the column's cloud fractions and overlap parameters, for the mask drawing below (its
    // layer-by-layer reads then wait on LDS, not on a load from memory), and its cloud top
    if (FUSED && a.sc.icld >= 1) {
      v.cf[lay] = a.cldfrac[c + (size_t)C * lay];
      v.al[lay] = a.mc_icld >= 4 ? a.alpha[c + (size_t)C * lay] : 0.0;
      if (v.cf[lay] >= CLDMIN) atomicMax(&v.misc[M_CLDTOP], lay);
    }
  }
  __syncthreads();
  const int laytrop = v.misc[M_LAYTROP];
  if (tid < NBND) {
    const int first[NBND] = {0, 6, 18, 26, 34, 44, 54, 56, 66, 74, 80, 86, 94, 100};
    v.misc[M_LAYSOL + tid] = find_laysolfr(T.gp[first[tid]], v, laytrop, L);
  }
  int err = 0;
  if (FUSED && a.sc.icld >= 1 && active) {
    if (SPEC)   // the column program's McICA samples
      err = column_masks(a.mc_icld, a.mc_irng, a.permuteseed, C, L, c, g, a.play, v.cf, v.al, 1,
                         a.mt_stream, reinterpret_cast<const KissJump *>(T.kiss_jump), v.mask,
                         v.misc[M_CLDTOP], reinterpret_cast<const KissJump *>(a.col_jump), a.mt_stride);
    else
      err = column_masks(a.mc_icld, a.mc_irng, a.permuteseed, C, L, c, g, a.play, v.cf, v.al, 1,
                         a.mt_stream, reinterpret_cast<const KissJump *>(T.kiss_jump), v.mask,
                         v.misc[M_CLDTOP]);
  }
  if (LAYERCLD && a.sc.icld >= 1)   // src/rrtmg_sw_rad.nomcica.f90:617-622
    for (int lay = tid; lay < L; lay += blockDim.x) {
      const double f = a.cldfrac[c + (size_t)C * lay];
      if (f > 1.e-06 && f < 1.0 - 1.e-06) v.misc[M_PARTIAL] = 1;
    }
  __syncthreads();
  // Fused entry: every cloudy sub-column cell of a band sees the same cloud inputs (the layer's
  // water paths and radii, the band's tauc/ssac/asmc/fsfc), so cldprmc_sw's result depends on
  // (layer, band) only. It is evaluated here for each (layer, band) with at least one cloudy
  // sub-column (the cells cldprmc_sw visits, so its STOP conditions fire as in the reference).
  double *csc = a.scratch + (size_t)cl * scr_col_doubles(L) + (size_t)L * SCR_COEFS * NGPT;
  if (LAYERCLD && a.sc.icld >= 1) {
    // cldprop_sw (src/rrtmg_sw_cldprop.f90:141-362) for every (layer, band): its cloud test
    // reads the layer's tauc summed over the bands in band order (tauctot). The reference stops
    // on a partial cloud before calling it, so that error takes precedence in the column.
    const bool partial = v.misc[M_PARTIAL] != 0;
    if (partial && tid == 0) err = RRTMGSW_ERR_PARTIAL_CLOUD;
    for (int idx = partial ? L * NBND : tid; idx < L * NBND; idx += blockDim.x) {
      const int lay = idx / NBND, b = idx - lay * NBND;
      const size_t q1 = c + (size_t)C * lay;
      double tauctot = 0.0;
      for (int bb = 0; bb < NBND; bb++) tauctot = tauctot + a.tauc[bb + NBND * q1];
      const size_t qb = b + NBND * q1;
      CloudOpt co = cloud_optics_layer(T, a.sc, b + 16, a.cldfrac[q1], a.ciwp[q1], a.clwp[q1], a.rei[q1],
                                       a.rel[q1], a.tauc[qb], a.ssac[qb], a.asmc[qb], a.fsfc[qb], tauctot, err);
      double *o = csc + (size_t)idx * 3;
      o[0] = co.tau, o[1] = co.ssa, o[2] = co.asm_;
      if (SPEC) csc[(size_t)L * NBND * 3 + idx] = co.orig;   // taucldorig for the idelm = 0 beam
    }
    __syncthreads();
  }
  if (FUSED && a.sc.icld >= 1) {
    const int first[NBND + 1] = {0, 6, 18, 26, 34, 44, 54, 56, 66, 74, 80, 86, 94, 100, NGPT};
    const int nlc = a.mc_irng == 0 ? v.misc[M_CLDTOP] + 1 : L;   // KISS masks: clear above the cloud top
    for (int idx = tid; idx < nlc * NBND; idx += blockDim.x) {
      const int lay = idx / NBND, b = idx - lay * NBND;
      uint32_t any = 0;
      for (int gg = first[b]; gg < first[b + 1]; gg++) any |= v.mask[(lay >> 5) * NGPT + gg];
      if ((any >> (lay & 31)) & 1u) {
        const size_t q1 = c + (size_t)C * lay, qb = b + NBND * q1;
        CloudOpt co = cloud_optics(T, a.sc, b + 16, 1.0, a.ciwp[q1], a.clwp[q1], a.rei[q1], a.rel[q1],
                                   a.tauc[qb], a.ssac[qb], a.asmc[qb], a.fsfc[qb], err);
        double *o = csc + (size_t)idx * 3;
        o[0] = co.tau, o[1] = co.ssa, o[2] = co.asm_;
        if (SPEC) csc[(size_t)L * NBND * 3 + idx] = co.orig;   // taormc for the idelm = 0 beam
      }
    }
    __syncthreads();
  }

  const GPoint &gp = T.gp[gi];
  const int band = gp.band, ibnd = gp.ibnd;
  const double mu0 = fmax(a.coszen[c], 1.e-10);
  const double zinc = active ? incident_flux(gp, T.tab, v, a.sc, mu0) : 0.0;
  const double rmu = 1.0 / mu0;

  // ---- pass 1: top to bottom ----
  // Per layer and sky the lane stores B = prdnd, C and E of the elimination (slots 0-2 clear,
  // 3-5 total sky); the A = ztdn term of the downward flux is summed over g here. A sub-column
  // stays identical to the clear sky until its first cloudy layer (jfirst): above it the
  // total-sky sweep is not computed and nothing is stored for it.
  Sweep sc_ = {1.0, 1.0, 0.0}, st_ = {1.0, 1.0, 0.0};
  // the mixed form's FP32 sweep state (clear, total)
  SweepF scf = {1.f, 1.f, 0.f}, stf = {1.f, 1.f, 0.f};
  int jfirst = L;
  double *scr = a.scratch + (size_t)cl * scr_col_doubles(L) + gi;
  // A layer's coefficients are stored once the next layer's table reads have been issued, so that
  // the waits for those reads do not include the stores (the vector memory counter retires in
  // issue order). The mixed form stores FP32 B, C, E and the diffuse absorptance ad per sky in
  // the same slots (8 floats in the 6 doubles of a layer).
  double pend[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  float pendf[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  int pend_j = -1;
  bool pend_t = false;
  auto flush = [&]() {
    if (active && pend_j >= 0) {
      if constexpr (MIXED) {
        float *p = reinterpret_cast<float *>(scr - gi + (size_t)pend_j * SCR_COEFS * NGPT) + gi;
        p[0] = pendf[0], p[NGPT] = pendf[1], p[2 * NGPT] = pendf[2], p[3 * NGPT] = pendf[3];
        if (pend_t)
          p[4 * NGPT] = pendf[4], p[5 * NGPT] = pendf[5], p[6 * NGPT] = pendf[6], p[7 * NGPT] = pendf[7];
      } else {
        double *p = scr + (size_t)pend_j * SCR_COEFS * NGPT;
        p[0] = pend[0], p[NGPT] = pend[1], p[2 * NGPT] = pend[2];
        if (pend_t) p[3 * NGPT] = pend[3], p[4 * NGPT] = pend[4], p[5 * NGPT] = pend[5];
      }
    }
  };
  // aerosol (src/rrtmg_sw_rad.f90:709-765) and this sub-column's cloud optics (cldprmc_sw) and
  // total-sky inputs (src/rrtmg_sw_spcvmc.f90:500-505)
  struct LayIn {
    double ta, wa, ga, cf, ct, cw, cg, co;
  };
  auto inputs = [&](int lay) {
    LayIn in = {0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0};
    if (a.sc.iaer == 10) {
      size_t qa = c + (size_t)C * (lay + (size_t)L * ibnd);
      in.ta = a.tauaer[qa], in.wa = a.ssaaer[qa], in.ga = a.asmaer[qa];
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
    if (FUSED && a.sc.icld >= 1 && active) {
      if ((v.mask[(lay >> 5) * NGPT + g] >> (lay & 31)) & 1u) {
        const double *o = csc + ((size_t)lay * NBND + ibnd) * 3;
        in.cf = 1.0, in.ct = o[0], in.cw = o[1], in.cg = o[2];
        if (SPEC) in.co = csc[(size_t)L * NBND * 3 + (size_t)lay * NBND + ibnd];
      }
    } else if (LAYERCLD && a.sc.icld >= 1 && active) {
      const double f = a.cldfrac[c + (size_t)C * lay];
      if (f != 0.0) {
        const double *o = csc + ((size_t)lay * NBND + ibnd) * 3;
        in.cf = f, in.ct = o[0], in.cw = o[1], in.cg = o[2];
        if (SPEC) in.co = csc[(size_t)L * NBND * 3 + (size_t)lay * NBND + ibnd];
      }
    } else if (CLD == CLD_SUBCOL && a.sc.icld >= 1 && active) {
      const size_t q1 = c + (size_t)C * lay, qg = (size_t)g + NGPT * q1;
      in.cf = a.cldfmcl[qg];
      const double ciwp = a.ciwpmcl[qg], clwp = a.clwpmcl[qg];
      const double tin = a.taucmcl[qg], win = a.ssacmcl[qg], gin = a.asmcmcl[qg], fin = a.fsfcmcl[qg];
      if (in.cf >= CLDMIN) {
        CloudOpt co = cloud_optics(T, a.sc, band, in.cf, ciwp, clwp, a.reicmcl[q1], a.relqmcl[q1], tin,
                                   win, gin, fin, err);
        in.ct = co.tau, in.cw = co.ssa, in.cg = co.asm_;
      } else {
        in.ct = tin, in.cw = win, in.cg = gin;
      }
    }
    return in;
  };
  // The FP64 layer: clear-sky optics and delta scaling (src/rrtmg_sw_spcvmc.f90:442-496):
  // g0 = ga wa ta / s and w0 = s / tau0 share 1 / (s tau0); omc = (w0 - w0 f)/(1 - w0 f) and
  // gcc = (g0 - f)/(1 - f) = g0 / (1 + g0) share 1 / ((1 - w0 f)(1 + g0)), f = g0^2; reftra;
  // total sky; the elimination steps. Leaves the coefficients in pend and returns the ztdn
  // terms (total, clear) of the level's downward flux.
  auto body64 = [&](int j, double taug, double taur, const LayIn &in, double &At, double &Ac) {
    const double ta = in.ta, wa = in.wa, ga = in.ga, cf = in.cf, ct = in.ct, cw = in.cw, cg = in.cg;
    const double tau0 = taur + taug + ta;
    const double sct = taur * 1.0 + ta * wa;
    const double r1 = frcp1(sct * tau0);
    const double g0 = ga * wa * ta * tau0 * r1;
    const double w0 = sct * sct * r1;
    const double zwf = w0 * (g0 * g0);
    const double r2 = frcp1((1.0 - zwf) * (1.0 + g0));
    const double tauc = (1.0 - zwf) * tau0;
    const double omc = (w0 - zwf) * (1.0 + g0) * r2;
    const double gcc = g0 * (1.0 - zwf) * r2;
    RT4 rc;
    double dbc;
    if (EXACT) {
      RTx o = reftra_x(tauc, omc, gcc, mu0, rmu);
      rc = {o.r, o.rd, o.t, o.td};
      dbc = o.e2;
    } else {
      const double e2 = tbl_exp1(fmin(tauc * rmu, 500.), T.exp_tbl);
      rc = reftra_e(tauc, omc, gcc, mu0, e2, T.exp_tbl);
      dbc = e2;
    }
    Ac = sc_.ztdn;
    const double Bc = sc_.prdnd;
    // a cloud fraction other than zero (negative ones included) enters the clear / cloudy blend
    if (cf != 0.0 && jfirst == L) {
      jfirst = j;
      st_ = sc_;
    }
    double Cc, Ec;
    sweep_r(sc_, rc, dbc, Cc, Ec);
    At = Ac;
    double Bt = Bc, Ct = Cc, Et = Ec;
    if (jfirst <= j) {
      At = st_.ztdn, Bt = st_.prdnd;
      RT4 rt = rc;
      double dbt = dbc;
      if (cf != 0.0) {
        // total-sky optics (:500-505): go = (ct cw cg + tauc omc gcc) / so, omo = so / tauo
        const double tauo = tauc + ct;
        const double so = tauc * omc + ct * cw;
        const double r3 = frcp1(so * tauo);
        const double go = (ct * cw * cg + tauc * omc * gcc) * tauo * r3;
        const double omo = so * so * r3;
        RT4 ro;
        double dbo;
        if (EXACT) {
          RTx o = reftra_x(tauo, omo, go, mu0, rmu);
          ro = {o.r, o.rd, o.t, o.td};
          dbo = o.e2;
        } else {
          const double e2 = tbl_exp1(fmin(tauo * rmu, 500.), T.exp_tbl);
          ro = reftra_e(tauo, omo, go, mu0, e2, T.exp_tbl);
          dbo = e2;
        }
        if (!(cf > REPCLC)) ro = {0.0, 0.0, 1.0, 1.0};
        const double zcl = 1.0 - cf;
        rt.r = zcl * rc.r + cf * ro.r;
        rt.rd = zcl * rc.rd + cf * ro.rd;
        rt.t = zcl * rc.t + cf * ro.t;
        rt.td = zcl * rc.td + cf * ro.td;
        dbt = zcl * dbc + cf * dbo;
      }
      sweep_r(st_, rt, dbt, Ct, Et);
    }
    pend[0] = Bc, pend[1] = Cc, pend[2] = Ec, pend[3] = Bt, pend[4] = Ct, pend[5] = Et;
  };
  // The mixed form's layer: the FP64 front of body64 (delta-scaled optics, reftra_front, every
  // table lookup), the FP32 back (reftra_mixed) and the FP32 elimination step. Besides the
  // coefficients it leaves in pendf the layer's absorbed flux known in pass 1,
  // P = a S + ad (ztdn - S) (S = tdbt, the direct beam at the layer top), per sky; pass 2 adds
  // ad (B U_top + U_bottom). Returns ztdn (total, clear) at the layer top.
  auto bodymx = [&](int j, double taug, double taur, const LayIn &in, float &At, float &Ac, float &Pt,
                    float &Pc) {
    const double ta = in.ta, wa = in.wa, ga = in.ga, cf = in.cf, ct = in.ct, cw = in.cw, cg = in.cg;
    const double tau0 = taur + taug + ta;
    const double sct = taur * 1.0 + ta * wa;
    const double r1 = frcp1(sct * tau0);
    const double g0 = ga * wa * ta * tau0 * r1;
    const double w0 = sct * sct * r1;
    const double zwf = w0 * (g0 * g0);
    const double r2 = frcp1((1.0 - zwf) * (1.0 + g0));
    const double tauc = (1.0 - zwf) * tau0;
    const double omc = (w0 - zwf) * (1.0 + g0) * r2;
    const double gcc = g0 * (1.0 - zwf) * r2;
    // co-albedo 1 - omc = (1 - w0)(1 + g0) r2 with 1 - w0 = (taug + ta (1 - wa)) / tau0
    const double ccl = (taug + ta * (1.0 - wa)) * sct * r1 * (1.0 + g0) * r2;
    const double ze2 = fmin(tauc * rmu, 500.);
    const double e2 = tbl_exp1(ze2, T.exp_tbl);
    const RTm rc = reftra_mixed(reftra_front(tauc, omc, gcc, mu0, T.exp_tbl), tauc, omc, float(ccl), mu0,
                                e2, tbl_deficit(ze2, e2));
    const float dbc = float(e2);
    Ac = scf.ztdn;
    Pc = rc.a * scf.tdbt + rc.ad * (scf.ztdn - scf.tdbt);
    const float Bc = scf.prdnd;
    if (cf != 0.0 && jfirst == L) {
      jfirst = j;
      stf = scf;
    }
    float Cc, Ec;
    sweep_f(scf, rc, dbc, Cc, Ec);
    At = Ac, Pt = Pc;
    float Bt = Bc, Ct = Cc, Et = Ec, adt = rc.ad;
    if (jfirst <= j) {
      At = stf.ztdn, Bt = stf.prdnd;
      RTm rt = rc;
      float dbt = dbc;
      if (cf != 0.0) {   // a cloudy sub-column cell (cf = 1): the total sky is the cloudy one
        const double tauo = tauc + ct;
        const double so = tauc * omc + ct * cw;
        const double r3 = frcp1(so * tauo);
        const double go = (ct * cw * cg + tauc * omc * gcc) * tauo * r3;
        const double omo = so * so * r3;
        const double cco = (tauc * ccl + ct * (1.0 - cw)) * so * r3;   // 1 - omo
        const double zo = fmin(tauo * rmu, 500.);
        const double e2o = tbl_exp1(zo, T.exp_tbl);
        rt = reftra_mixed(reftra_front(tauo, omo, go, mu0, T.exp_tbl), tauo, omo, float(cco), mu0, e2o,
                          tbl_deficit(zo, e2o));
        dbt = float(e2o);
      }
      Pt = rt.a * stf.tdbt + rt.ad * (stf.ztdn - stf.tdbt);
      sweep_f(stf, rt, dbt, Ct, Et);
      adt = rt.ad;
    }
    pendf[0] = Bc, pendf[1] = Cc, pendf[2] = Ec, pendf[3] = rc.ad;
    pendf[4] = Bt, pendf[5] = Ct, pendf[6] = Et, pendf[7] = adt;
  };
  // column program: the per-g level terms of the per-band fluxes, and the direct beam without
  // delta scaling (src/rrtmg_sw_spcvmc.f90:449-470: the table transmittance of the unscaled
  // clear and cloudy optical depths, blended by the cloud fraction)
  double *gs = SPEC ? a.gspec + (size_t)cl * (L + 1) * 4 * NGPT + g : nullptr;
  double tnd = 1.0;
  auto spec_level = [&](int j, double A, double Tdir) {
    if (SPEC && active) {
      gs[(size_t)(j * 4 + 0) * NGPT] = zinc * A;
      gs[(size_t)(j * 4 + 1) * NGPT] = zinc * (a.idelm == 0 ? tnd : Tdir);
    }
  };
  auto layer = [&](auto lower_tag, const GasLane &q, int j) {
    constexpr bool LOWER = decltype(lower_tag)::value;
    const int lay = L - 1 - j;
    if constexpr (MIXED) {
      const GasFetch<double> gf = gas_fetch<LOWER, double>(q, T.tab, v, lay);
      flush();
      double taug, taur;
      gas_finish<LOWER, double>(q, gf, v, lay, taug, taur);
      float At, Ac, Pt, Pc;
      bodymx(j, taug, taur, inputs(lay), At, Ac, Pt, Pc);
      pend_j = j, pend_t = jfirst <= j;
      // sums over g of the incident-flux-weighted ztdn and pass-1 absorbed flux, both skies
      const float zf = float(zinc);
      const float sa = wave_sum4f(zf * At, zf * Ac, zf * Pt, zf * Pc);
      if ((lane & 15) == 0) {
        const int q = (wave * L + j) * 4 + (lane >> 4);
        if (gsum) reinterpret_cast<float *>(greda)[q] = sa;
        else v.redaf[q] = sa;
      }
      return;
    }
    double At, Ac;
    if (SPEC) {
      const double Tdir = jfirst <= j ? st_.tdbt : sc_.tdbt;   // total-sky beam at level j
      const GasFetch<double> gf = gas_fetch<LOWER, double>(q, T.tab, v, lay);
      flush();
      double taug, taur;
      gas_finish<LOWER, double>(q, gf, v, lay, taug, taur);
      const LayIn in = inputs(lay);
      body64(j, taug, taur, in, At, Ac);
      spec_level(j, At, Tdir);
      const double tau0 = taur + taug + in.ta;
      const double dmc = tbl_exp1(tau0 / mu0, T.exp_tbl);
      const double dmo = tbl_exp1((tau0 + in.co) / mu0, T.exp_tbl);
      tnd = (in.cf != 0.0 ? (1.0 - in.cf) * dmc + in.cf * dmo : dmc) * tnd;
    } else {
      const GasFetch<double> gf = gas_fetch<LOWER, double>(q, T.tab, v, lay);
      flush();
      double taug, taur;
      gas_finish<LOWER, double>(q, gf, v, lay, taug, taur);
      body64(j, taug, taur, inputs(lay), At, Ac);
    }
    pend_j = j, pend_t = jfirst <= j;
    // sum over g of the incident-flux-weighted ztdn of level j, both skies
    const double sa = wave_sum2d(zinc * At, zinc * Ac);
    if ((lane & 31) == 0) {
      const int q = (wave * L + j) * 2 + (lane >> 5);
      if (gsum) greda[q] = sa;
      else v.reda[q] = sa;
    }
  };
  // Each layer issues its gas-optics table reads, then stores the coefficients of the layer
  // above.
  const int jsplit = L - laytrop;   // layers j >= jsplit lie below laytrop
  auto regime = [&](auto lower_tag, int j0, int j1) {
    constexpr bool LOWER = decltype(lower_tag)::value;
    const GasLane q = gas_lane<LOWER>(gp, L);
    for (int j = j0; j < j1; j++) layer(lower_tag, q, j);
  };
  regime(std::integral_constant<bool, false>(), 0, jsplit);
  regime(std::integral_constant<bool, true>(), jsplit, L);
  flush();
  if (jfirst == L) st_ = sc_, stf = scf;
  spec_level(L, st_.ztdn, st_.tdbt);
  if (err) atomicCAS(a.status, 0, err);
  if constexpr (MIXED) {
    if (gsum)
      rt_pass2_mixed(a, reinterpret_cast<float *>(gred), reinterpret_cast<const float *>(greda), c, scr,
                     gi, lane, wave, ibnd, float(zinc), jfirst, scf, stf);
    else
      rt_pass2_mixed(a, v.redf, v.redaf, c, scr, gi, lane, wave, ibnd, float(zinc), jfirst, scf, stf);
    return;
  }

  // ---- pass 2: bottom to top ----
  // surface (src/rrtmg_sw_vrtqdr.f90:163-169 with prup = albdir, prupd = albdif)
  const bool vis = ibnd >= 9 && ibnd <= 12;
  double albdir, albdif;
  if (SPEC) {
    albdir = albdif = a.albbnd[ibnd + NBND * (size_t)c];   // column program: 1 - semiss per band
  } else {
    albdir = vis ? a.asdir[c] : a.aldir[c];
    albdif = vis ? a.asdif[c] : a.aldif[c];
  }
  double uc = (albdir * sc_.tdbt + albdif * (sc_.ztdn - sc_.tdbt)) / (1. - albdif * sc_.prdnd);
  double dc = sc_.ztdn + sc_.prdnd * uc;
  double ut = (albdir * st_.tdbt + albdif * (st_.ztdn - st_.tdbt)) / (1. - albdif * st_.prdnd);
  double dt = st_.ztdn + st_.prdnd * ut;
  {
    const double s = wave_sum4d(zinc * ut, zinc * dt, zinc * uc, zinc * dc);
    if ((lane & 15) == 0) {
      const int q = (wave * (L + 1) + L) * 4 + (lane >> 4);
      if (gsum) gred[q] = s;
      else v.red[q] = s;
    }
  }
  if (SPEC && active) {
    gs[(size_t)(L * 4 + 2) * NGPT] = zinc * ut;
    gs[(size_t)(L * 4 + 3) * NGPT] = zinc * (st_.prdnd * ut);
  }
  // Coefficients are requested four layers ahead (a ring of four register sets); the total-sky
  // slots only by lanes whose sub-column is cloudy at or above that layer (the others were
  // never written).
  struct Co {
    double cB, cC, cE, tB, tC, tE;
  };
  auto fetch = [&](int j) {
    Co o;
    const int jj = j < 0 ? 0 : j;
    o.tB = 0.0, o.tC = 0.0, o.tE = 0.0;
    const double *p = scr + (size_t)jj * SCR_COEFS * NGPT;
    o.cB = p[0], o.cC = p[NGPT], o.cE = p[2 * NGPT];
    if (jj >= jfirst) o.tB = p[3 * NGPT], o.tC = p[4 * NGPT], o.tE = p[5 * NGPT];
    return o;
  };
  auto step = [&](int j, const Co &o) {
    const bool tot = j >= jfirst;
    const double ut_t = o.tC + o.tE * ut, ut_c = o.cC + o.cE * ut;
    ut = tot ? ut_t : ut_c;
    const double bu_t = (tot ? o.tB : o.cB) * ut;
    uc = o.cC + o.cE * uc;
    const double bu_c = o.cB * uc;
    // downward flux = sum(zinc ztdn) (pass 1) + sum(zinc prdnd U)
    if (SPEC && active) {
      gs[(size_t)(j * 4 + 2) * NGPT] = zinc * ut;
      gs[(size_t)(j * 4 + 3) * NGPT] = zinc * bu_t;
    }
    const double s = wave_sum4d(zinc * ut, zinc * bu_t, zinc * uc, zinc * bu_c);
    if ((lane & 15) == 0) {
      const int q = (wave * (L + 1) + j) * 4 + (lane >> 4);
      if (gsum) gred[q] = s;
      else v.red[q] = s;
    }
  };
  Co r0 = fetch(L - 1), r1 = fetch(L - 2), r2 = fetch(L - 3), r3 = fetch(L - 4);
  int j = L - 1;
  for (; j >= 3; j -= 4) {
    step(j, r0);
    r0 = fetch(j - 4);
    step(j - 1, r1);
    r1 = fetch(j - 5);
    step(j - 2, r2);
    r2 = fetch(j - 6);
    step(j - 3, r3);
    r3 = fetch(j - 7);
  }
  // at most three layers remain: j, j - 1, j - 2 in r0, r1, r2
  if (j >= 0) step(j, r0);
  if (j >= 1) step(j - 1, r1);
  if (j >= 2) step(j - 2, r2);
  __syncthreads();
  if (gsum)
    rt_epilogue(a, gred, greda, c);
  else
    rt_epilogue(a, v.red, v.reda, c);
}

template <int MODE, int CLD, bool GSUM>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu((MODE == 0 && CLD != CLD_SUBCOL) || MODE == 1 ? 4 : 3)))
rt_kernel(RtArgs a, DeviceTables T) {
  rt_body<MODE, CLD, false, GSUM>(a, T);
}

// The column program's forms: per-band and direct-beam outputs, with the non-McICA clouds
// (imca = 0) or McICA samples (imca = 1, one sample per workgroup).
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4)))
rt_column_kernel(RtArgs a, DeviceTables T) {
  rt_body<0, CLD_LAYER, true>(a, T);
}
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4)))
rt_column_mcica_kernel(RtArgs a, DeviceTables T) {
  rt_body<0, CLD_FUSED, true>(a, T);
}

// The fused forms' flux sums go to the column's scratch when keeping them in LDS would leave room
// for fewer than the eight workgroups per CU that their registers allow (4 waves per SIMD); the
// other forms keep them in LDS.
static bool sums_global(int nlay, int nmw, bool mixed) {
  const int per_cu = 160 * 1024 / 8;
  return ColLayout(nlay, nmw, true, mixed).total > per_cu;
}

template <int MODE, int CLD>
static hipError_t launch_rt(const RtArgs &a, const DeviceTables &t, hipStream_t s) {
  const int nmw = CLD == CLD_FUSED ? (a.nlay + 31) / 32 : 0;
  const bool gsum = CLD == CLD_FUSED && sums_global(a.nlay, nmw, MODE == 1);
  ColLayout lo(a.nlay, nmw, !gsum, MODE == 1);
  int per = (a.ncol_chunk + 7) / 8;
  if (gsum)
    hipLaunchKernelGGL((rt_kernel<MODE, CLD, CLD == CLD_FUSED>), dim3(per * 8), dim3(128), lo.total, s, a, t);
  else
    hipLaunchKernelGGL((rt_kernel<MODE, CLD, false>), dim3(per * 8), dim3(128), lo.total, s, a, t);
  return hipGetLastError();
}

// Per-band total-sky fluxes of one chunk from the per-g level terms (column program output):
// one thread per (column, level, band), g summed in order within the band.
__global__ void __launch_bounds__(256) spec_reduce_kernel(RtArgs a, DeviceTables T) {
  const int L = a.nlay, C = a.ncol;
  const long long n = (long long)a.ncol_chunk * (L + 1) * NBND;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int b = (int)(t % NBND);
  const long long r = t / NBND;
  const int j = (int)(r % (L + 1)), cl = (int)(r / (L + 1));
  const int first[NBND + 1] = {0, 6, 18, 26, 34, 44, 54, 56, 66, 74, 80, 86, 94, 100, NGPT};
  const double *gs = a.gspec + ((size_t)cl * (L + 1) + j) * 4 * NGPT;
  double up = 0.0, dn = 0.0, dir = 0.0;
  for (int g = first[b]; g < first[b + 1]; g++) {
    up += gs[2 * NGPT + g];
    dn += gs[0 * NGPT + g] + gs[3 * NGPT + g];
    dir += gs[1 * NGPT + g];
  }
  (void)T;
  const size_t q = (size_t)(a.col0 + cl) + (size_t)C * ((L - j) + (size_t)(L + 1) * b);
  a.bandu[q] = up, a.bandd[q] = dn, a.banddir[q] = dir;
}

hipError_t launch_rt_spectral(const RtArgs &a, const DeviceTables &t, hipStream_t s) {
  const bool mcica = a.fused != 0;
  ColLayout lo(a.nlay, mcica ? (a.nlay + 31) / 32 : 0);
  int per = (a.ncol_chunk + 7) / 8;
  if (mcica)
    hipLaunchKernelGGL(rt_column_mcica_kernel, dim3(per * 8), dim3(128), lo.total, s, a, t);
  else
    hipLaunchKernelGGL(rt_column_kernel, dim3(per * 8), dim3(128), lo.total, s, a, t);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const long long n = (long long)a.ncol_chunk * (a.nlay + 1) * NBND;
  hipLaunchKernelGGL(spec_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, t);
  return hipGetLastError();
}

hipError_t launch_rt_fp64(const RtArgs &a, const DeviceTables &t, hipStream_t s) {
  if (a.layer_clouds) return launch_rt<0, CLD_LAYER>(a, t, s);
  return a.fused ? launch_rt<0, CLD_FUSED>(a, t, s) : launch_rt<0, CLD_SUBCOL>(a, t, s);
}
hipError_t launch_rt_mixed(const RtArgs &a, const DeviceTables &t, hipStream_t s) {
  // the mixed form is offered by the fused entry point only
  return a.fused && !a.layer_clouds ? launch_rt<1, CLD_FUSED>(a, t, s) : hipErrorInvalidValue;
}
hipError_t launch_rt_exact(const RtArgs &a, const DeviceTables &t, hipStream_t s) {
  // the exact-exponential form is offered by the fused entry point only
  return a.fused && !a.layer_clouds ? launch_rt<2, CLD_FUSED>(a, t, s) : hipErrorInvalidValue;
}

// ------------------------------------------------------------------------------------------
// setcoef_sw (src/rrtmg_sw_setcoef.f90:49-305) for every column, fed from inatm_sw's layer
// amounts: the forward kernels' own prologue (setcoef_layer for every layer, laytrop,
// find_laysolfr of each band), one workgroup per column, its LDS results written out.
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(128) setcoef_kernel(RtArgs a, DeviceTables T, SetcoefOut o) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int L = a.nlay, C = a.ncol;
  ColLayout lo(L, 0, false);
  ColView v(smem, lo);
  const int c = a.col0 + blockIdx.x, tid = threadIdx.x;
  if (tid == 0) v.misc[M_LAYTROP] = 0, v.misc[M_LAYLOW] = 0;
  __syncthreads();
  for (int lay = tid; lay < L; lay += blockDim.x) setcoef_layer(a, T, c, lay, v, &v.misc[M_LAYTROP]);
  __syncthreads();
  const int laytrop = v.misc[M_LAYTROP];
  if (tid < NBND) {
    const int first[NBND] = {0, 6, 18, 26, 34, 44, 54, 56, 66, 74, 80, 86, 94, 100};
    v.misc[M_LAYSOL + tid] = find_laysolfr(T.gp[first[tid]], v, laytrop, L);
  }
  __syncthreads();
  if (tid == 0) {
    if (o.laytrop) o.laytrop[c] = laytrop;
    if (o.laylow) o.laylow[c] = v.misc[M_LAYLOW];
  }
  if (tid < NBND && o.laysolfr) o.laysolfr[c + (size_t)C * tid] = v.misc[M_LAYSOL + tid];
  int *const iv[5] = {o.jp, o.jt, o.jt1, o.indself, o.indfor};
  // fac00 fac01 fac10 fac11 selffac selffrac forfac forfrac, then h2o co2 o3 ch4 o2 colmol
  double *const dv[14] = {o.fac00,   o.fac01,  o.fac10,  o.fac11,  o.selffac, o.selffrac, o.forfac,
                          o.forfrac, o.colh2o, o.colco2, o.colo3,  o.colch4,  o.colo2,    o.colmol};
  for (int lay = tid; lay < L; lay += blockDim.x) {
    const size_t q = c + (size_t)C * lay;
    for (int k = 0; k < 5; k++)
      if (iv[k]) iv[k][q] = ColView::idx(v.ix[lay], k);
    for (int k = 0; k < 8; k++)
      if (dv[k]) dv[k][q] = v.f(k, lay);
    for (int k = 0; k < 6; k++)
      if (dv[8 + k]) dv[8 + k][q] = v.c(k, lay);
  }
}

hipError_t launch_setcoef(const RtArgs &a, const SetcoefOut &o, const DeviceTables &t, hipStream_t s) {
  ColLayout lo(a.nlay, 0, false);
  hipLaunchKernelGGL(setcoef_kernel, dim3(a.ncol_chunk), dim3(128), lo.total, s, a, t, o);
  return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// mcica_subcol_sw with materialized outputs: one thread per (column, g-point).
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(128) mcica_kernel(McicaArgs a, DeviceTables T) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  uint32_t *mask = reinterpret_cast<uint32_t *>(smem);
  const int c = blockIdx.x, g = threadIdx.x;
  const int C = a.ncol, L = a.nlay;
  const bool active = g < NGPT;
  if (active) {
    int e = column_masks(a.icld, a.irng, a.permuteseed, C, L, c, g, a.play, a.cldfrac + c,
                         a.alpha ? a.alpha + c : nullptr, (size_t)C, a.mt_stream,
                         reinterpret_cast<const KissJump *>(T.kiss_jump), mask, L - 1);
    if (e) atomicCAS(a.status, 0, e);
    // band of g (ngb, src/rrtmg_sw_init.f90:340-353)
    int ibnd = T.gp[g].ibnd;
    for (int l = 0; l < L; l++) {
      bool cloudy = (mask[(l >> 5) * NGPT + g] >> (l & 31)) & 1u;
      size_t q = c + (size_t)C * l, qg = (size_t)g + NGPT * q, qb = ibnd + NBND * q;
      a.cldfmcl[qg] = cloudy ? 1.0 : 0.0;
      a.clwpmcl[qg] = cloudy ? a.clwp[q] : 0.0;
      a.ciwpmcl[qg] = cloudy ? a.ciwp[q] : 0.0;
      a.taucmcl[qg] = cloudy ? a.tauc[qb] : 0.0;
      a.ssacmcl[qg] = cloudy ? a.ssac[qb] : 1.0;
      a.asmcmcl[qg] = cloudy ? a.asmc[qb] : 0.0;
      a.fsfcmcl[qg] = cloudy ? a.fsfc[qb] : 0.0;
    }
  }
  if (g < L) {
    for (int l = g; l < L; l += blockDim.x) {
      size_t q = c + (size_t)C * l;
      a.reicmcl[q] = a.rei[q];
      a.relqmcl[q] = a.rel[q];
    }
  }
}

hipError_t launch_mcica(const McicaArgs &a, const DeviceTables &t, hipStream_t s) {
  int nmw = (a.nlay + 31) / 32;
  hipLaunchKernelGGL(mcica_kernel, dim3(a.ncol), dim3(128), nmw * NGPT * 4, s, a, t);
  return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Mersenne Twister MT19937 stream (src/mcica_random_numbers.f90:77-306): the single stream of a
// call, generated in order by one workgroup, or for long streams in segments of 2^19 words or
// more, one workgroup each, from start states found by jump-ahead.
// ------------------------------------------------------------------------------------------
// Barrier that orders this workgroup's LDS accesses only: the wave's LDS operations complete,
// then all waves meet. Unlike __syncthreads it does not wait for the wave's global stores, which
// no other wave reads. The memory clobber keeps the compiler from moving accesses across it.
__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// The stream continued from the 624 raw words in ring[0 .. 624) (ring of 2048 words in LDS):
// `count` tempered outputs to out. MT19937 is run as the recurrence over raw words
// x[r] = x[r - 227] ^ twist(x[r - 624], x[r - 623]), output r - 624 = temper(x[r]). A word depends
// on words at least 227 back, so thread t computes words r = 624 + base + t and r + 227 (the
// second takes the first as its x[r - 227]) from words written before the batch: 454 words per
// barrier, the batch's written and read windows [624 + base, 1078 + base) and [base, 624 + base)
// never meeting in the ring.
__device__ inline uint32_t mt_twist(uint32_t a, uint32_t b) {
  const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
  return (y >> 1) ^ ((b & 1u) ? 0x9908b0dfu : 0u);
}
__device__ inline uint32_t mt_temper(uint32_t y) {
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  return y ^ (y >> 18);
}
__device__ inline void mt_ring_run(uint32_t *ring, long long count, uint32_t *out) {
  constexpr uint32_t MSK = 2047;
  const int t = threadIdx.x;
  for (long long base = 0; base < count; base += 454) {
    if (t < 227) {
      const uint32_t r = (uint32_t)(624 + base + t);   // ring positions are r mod 2048
      const uint32_t a1 = ring[(r - 624) & MSK], b1 = ring[(r - 623) & MSK], c1 = ring[(r - 227) & MSK];
      const uint32_t a2 = ring[(r - 397) & MSK], b2 = ring[(r - 396) & MSK];
      const uint32_t x1 = c1 ^ mt_twist(a1, b1);
      const uint32_t x2 = x1 ^ mt_twist(a2, b2);
      ring[r & MSK] = x1;
      ring[(r + 227) & MSK] = x2;
      const long long o1 = base + t, o2 = o1 + 227;
      if (o1 < count) out[o1] = mt_temper(x1);
      if (o2 < count) out[o2] = mt_temper(x2);
    }
    lds_barrier();
  }
}
__device__ inline void mt_init(uint32_t seed, uint32_t *w) {   // init_genrand
  uint32_t x = seed;
  w[0] = x;
  for (int i = 1; i < 624; i++) {
    x = 1812433253u * (x ^ (x >> 30)) + (uint32_t)i;
    w[i] = x;
  }
}

// The whole stream in one workgroup.
__global__ void __launch_bounds__(256) mt_kernel(uint32_t seed, long long ndraw, uint32_t *out) {
  __shared__ uint32_t ring[2048];
  if (threadIdx.x == 0) mt_init(seed, ring);
  __syncthreads();
  mt_ring_run(ring, ndraw, out);
}

// Jump-ahead (Haramoto et al. 2008): with x^(k J) = sum_i c_i x^i mod P (the set bits i listed by
// the host), the 624-word window k J raw words ahead is the XOR of the windows i ahead. Every
// segment start is formed at once from the stream's own start: workgroup k - 1 extends the raw
// words from the seed to 624 + 19936, then forms start k (thread w owns word w).
constexpr int MT_EXT = 624 + 19936;   // raw words a jump reads
// bits[offs[k - 1] .. offs[k]) are the set bits of x^(k J) mod P, each below 19937.
// Segment k's start window: the XOR of the windows i words into the stream over the set bits i
// of x^(k J) mod P. MT_PARTS workgroups share a segment, each taking a slice of its bits and
// XOR-ing its partial window into the start (zeroed before the launch; XOR makes the order of
// the parts immaterial). Each workgroup generates the stream's first MT_EXT words in LDS and
// stages its bit list there in chunks (the threads of a wave read the same bit).
constexpr int MT_PARTS = 4, MT_BCHUNK = 4096;
__global__ void __launch_bounds__(640) mt_starts_kernel(uint32_t seed, const uint32_t *__restrict__ bits,
                                                        const uint32_t *__restrict__ offs, uint32_t *starts) {
  extern __shared__ __attribute__((aligned(16))) uint32_t msm[];
  uint32_t *buf = msm, *bl = msm + MT_EXT;
  const int t = threadIdx.x, k = blockIdx.x / MT_PARTS + 1, part = blockIdx.x % MT_PARTS;
  if (t == 0) mt_init(seed, buf);
  __syncthreads();
  if (k == 1 && part == 0)
    for (int i = t; i < 624; i += blockDim.x) starts[i] = buf[i];
  for (int r0 = 624; r0 < MT_EXT; r0 += 227) {
    const int r = r0 + t;
    if (t < 227 && r < MT_EXT) buf[r] = buf[r - 227] ^ mt_twist(buf[r - 624], buf[r - 623]);
    __syncthreads();
  }
  const uint32_t b0 = offs[k - 1], n = offs[k] - b0;
  const uint32_t p0 = b0 + (uint32_t)((uint64_t)n * part / MT_PARTS);
  const uint32_t p1 = b0 + (uint32_t)((uint64_t)n * (part + 1) / MT_PARTS);
  uint32_t acc = 0;
  for (uint32_t c0 = p0; c0 < p1; c0 += MT_BCHUNK) {
    const uint32_t cn = min((uint32_t)MT_BCHUNK, p1 - c0);
    __syncthreads();   // the previous chunk is consumed
    for (uint32_t i = t; i < cn; i += blockDim.x) bl[i] = bits[c0 + i];
    __syncthreads();
    if (t < 624) {
      // eight independent LDS reads in flight per step; a bit is at most 19936, so bit + t stays
      // below MT_EXT
      uint32_t j = 0;
      for (; j + 8 <= cn; j += 8) {
        uint32_t v = 0;
#pragma unroll
        for (int q = 0; q < 8; q++) v ^= buf[bl[j + q] + t];
        acc ^= v;
      }
      for (; j < cn; j++) acc ^= buf[bl[j] + t];
    }
  }
  if (t < 624) atomicXor(&starts[(size_t)k * 624 + t], acc);
}

__global__ void __launch_bounds__(256) zero_words(uint32_t *p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// Segment k of a jumped stream: outputs [k J, min((k + 1) J, ndraw)) from start window k.
__global__ void __launch_bounds__(256) mt_segment_kernel(const uint32_t *starts, long long ndraw, int log2_seg,
                                                         uint32_t *out) {
  __shared__ uint32_t ring[2048];
  const long long k = blockIdx.x, o0 = k << log2_seg, J = 1LL << log2_seg;
  const long long cnt = ndraw - o0 < J ? ndraw - o0 : J;
  for (int i = threadIdx.x; i < 624; i += blockDim.x) ring[i] = starts[k * 624 + i];
  __syncthreads();
  mt_ring_run(ring, cnt, out + o0);
}

// Segments of 2^19 words or more: at most 256 of them.
int mt_log2_seg(long long ndraw) {
  int l = 19;
  while ((ndraw >> l) > 256) l++;
  return l;
}

size_t mt_extra_words(long long ndraw) {
  const long long J = 1LL << mt_log2_seg(ndraw);
  return ndraw > 2 * J ? (size_t)((ndraw + J - 1) / J) * 624 : 0;
}

hipError_t launch_mt_stream(uint32_t seed, long long ndraw, uint32_t *out, const MtJumpTable *jt,
                            hipStream_t s) {
  const int lj = mt_log2_seg(ndraw);
  const long long J = 1LL << lj;
  if (jt && ndraw > 2 * J) {
    const int nseg = (int)((ndraw + J - 1) / J);
    if (jt->log2_seg != lj || jt->count < nseg - 1) return hipErrorInvalidValue;
    uint32_t *starts = out + ndraw;   // the caller sized out for ndraw + mt_extra_words(ndraw)
    const size_t nz = (size_t)(nseg - 1) * 624;
    hipLaunchKernelGGL(zero_words, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, s, starts + 624, nz);
    hipLaunchKernelGGL(mt_starts_kernel, dim3((nseg - 1) * MT_PARTS), dim3(640),
                       (MT_EXT + MT_BCHUNK) * sizeof(uint32_t), s, seed, jt->bits, jt->offs, starts);
    hipLaunchKernelGGL(mt_segment_kernel, dim3(nseg), dim3(256), 0, s, starts, ndraw, lj, out);
  } else {
    hipLaunchKernelGGL(mt_kernel, dim3(1), dim3(256), 0, s, seed, ndraw, out);
  }
  return hipGetLastError();
}

}  // namespace rrtmgsw
