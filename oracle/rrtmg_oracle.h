/*
 * rrtmg_oracle.h -- CPU restatement of RRTMG_SW's GCM-mode McICA shortwave path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product path (rrtmg_sw_amd, librrtmgsw_hip.so) never links or calls it.
 *
 * Every routine restates one reference routine in plain C, FP64, in the reference's
 * evaluation order, and cites the file:line it follows. Arrays use the reference's
 * Fortran column-major layouts; 2-D (ncol, nlay) arrays are addressed a[i + ncol*k].
 */
#ifndef RRTMG_ORACLE_H
#define RRTMG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* rrtmg_sw_ini(cpdair) -- src/rrtmg_sw_init.f90:47-173. blob = raw data tables. */
int orc_init(const char *blob_path, double cpdair);

/* Last error message (set when a routine returns non-zero, as the reference STOPs). */
const char *orc_last_error(void);

/* rrtmg_sw(...) -- src/rrtmg_sw_rad.f90:99-841; argument order of the Fortran dummies.
 * bndsolvar (14), indsolvar (2), solcycfrac (1) may be NULL (absent optionals). */
int orc_rrtmg_sw(int ncol, int nlay, int *icld, int *iaer,
                 const double *play, const double *plev, const double *tlay, const double *tlev,
                 const double *tsfc, const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                 const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                 const double *asdir, const double *asdif, const double *aldir, const double *aldif,
                 const double *coszen, double adjes, int dyofyr, double scon, int isolvar,
                 int inflgsw, int iceflgsw, int liqflgsw, const double *cldfmcl,
                 const double *taucmcl, const double *ssacmcl, const double *asmcmcl,
                 const double *fsfcmcl, const double *ciwpmcl, const double *clwpmcl,
                 const double *reicmcl, const double *relqmcl,
                 const double *tauaer, const double *ssaaer, const double *asmaer,
                 const double *ecaer,
                 double *swuflx, double *swdflx, double *swhr,
                 double *swuflxc, double *swdflxc, double *swhrc,
                 const double *bndsolvar, const double *indsolvar, const double *solcycfrac);

/* Same as orc_rrtmg_sw, and also returns per-band flux blocks (ncol, nlay+1, 14) for
 * total-sky up/down when bandu/bandd are non-NULL (the column program's IOUT=98 output,
 * src/rrtmg_sw.1col.f90:824-997), and the direct-beam downward flux dirdflx (ncol, nlay+1). */
int orc_rrtmg_sw_ext(int ncol, int nlay, int *icld, int *iaer,
                     const double *play, const double *plev, const double *tlay, const double *tlev,
                     const double *tsfc, const double *h2ovmr, const double *o3vmr,
                     const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                     const double *o2vmr, const double *asdir, const double *asdif,
                     const double *aldir, const double *aldif, const double *coszen, double adjes,
                     int dyofyr, double scon, int isolvar, int inflgsw, int iceflgsw, int liqflgsw,
                     const double *cldfmcl, const double *taucmcl, const double *ssacmcl,
                     const double *asmcmcl, const double *fsfcmcl, const double *ciwpmcl,
                     const double *clwpmcl, const double *reicmcl, const double *relqmcl,
                     const double *tauaer, const double *ssaaer, const double *asmaer,
                     const double *ecaer, double *swuflx, double *swdflx, double *swhr,
                     double *swuflxc, double *swdflxc, double *swhrc, const double *bndsolvar,
                     const double *indsolvar, const double *solcycfrac, double *bandu,
                     double *bandd, double *dirdflx, int idelm);

/* Column-mode driver (src/rrtmg_sw.1col.f90:634-821) for clear skies on layer amounts:
 * pavel/tavel/coldry (nlay), pz (nlay+1, surface first), wkl7 (7, nlay) molecules/cm2,
 * semiss (14). out: (6, nlay+1, 15) = up, difdown, dirdown, down, net, heating rate for the
 * broadband (block 0) and bands 16..29 (blocks 1..14), levels from the surface. */
int orc_column_sw(int nlay, const double *pavel, const double *tavel, const double *pz,
                  const double *coldry, const double *wkl7, const double *semiss, double zenith,
                  int juldat, double scon, int isolvar, const double *bndsolvar,
                  const double *indsolvar, const double *solcycfrac, int iaer,
                  const double *tauaer, const double *ssaaer, const double *asmaer, int idelm,
                  double cpdair, double *out);

/* orc_column_sw with the column program's imca = 0 clouds (cldprop_sw + spcvrt_sw,
 * src/rrtmg_sw.1col.f90:612-621, :669-680): cldfrac, ciwp, clwp, rei, rel (nlay, from the
 * surface); tauc, ssac, asmc, fsfc (14, nlay). Returns 6 for a fraction other than 0 or 1. */
int orc_column_sw_cld(int nlay, const double *pavel, const double *tavel, const double *pz,
                      const double *coldry, const double *wkl7, const double *semiss, double zenith,
                      int juldat, double scon, int isolvar, const double *bndsolvar,
                      const double *indsolvar, const double *solcycfrac, int iaer,
                      const double *tauaer, const double *ssaaer, const double *asmaer, int idelm,
                      double cpdair, int inflag, int iceflag, int liqflag, const double *cldfrac,
                      const double *ciwp, const double *clwp, const double *rei, const double *rel,
                      const double *tauc, const double *ssac, const double *asmc,
                      const double *fsfc, double *out);

/* orc_column_sw with the column program's McICA loop (imca = 1, src/rrtmg_sw.1col.f90:577-610,
 * :874-894): nmca samples, sample ims drawn with permuteseed (ims + 1) * 112 and the Mersenne
 * Twister (irng = 1, as the program sets it), overlap icld with alpha (nlay) for icld 4/5;
 * out holds the outputs averaged over the samples as the program prints them. */
int orc_column_sw_mcica(int nlay, const double *pavel, const double *tavel, const double *pz,
                        const double *coldry, const double *wkl7, const double *semiss, double zenith,
                        int juldat, double scon, int isolvar, const double *bndsolvar,
                        const double *indsolvar, const double *solcycfrac, int iaer,
                        const double *tauaer, const double *ssaaer, const double *asmaer, int idelm,
                        double cpdair, int inflag, int iceflag, int liqflag, const double *cldfrac,
                        const double *ciwp, const double *clwp, const double *rei, const double *rel,
                        const double *tauc, const double *ssac, const double *asmc,
                        const double *fsfc, int icld, int nmca, const double *alpha, double *out);

/* mcica_subcol_sw(...) -- src/mcica_subcol_gen_sw.f90:181-296 (+ :300-753). */
int orc_mcica_subcol_sw(int iplon, int ncol, int nlay, int icld, int permuteseed, int *irng,
                        const double *play, const double *cldfrac, const double *ciwp,
                        const double *clwp, const double *rei, const double *rel,
                        const double *tauc, const double *ssac, const double *asmc,
                        const double *fsfc, const double *alpha,
                        double *cldfmcl, double *ciwpmcl, double *clwpmcl,
                        double *reicmcl, double *relqmcl, double *taucmcl, double *ssacmcl,
                        double *asmcmcl, double *fsfcmcl);

/* get_alpha(...) -- src/mcica_subcol_gen_sw.f90:66-178. */
int orc_get_alpha(int iplon, int ncol, int nlayers, int icld, int idcor, double decorr_con,
                  const double *dz, const double *lat, int juldat, const double *cldfrac,
                  double *alpha);

/* Exact-exponential mode on (1) or off (0, the reference): every exponential-table transmittance
 * of reftra_sw (src/rrtmg_sw_reftra.f90:203-209, :277-295) and of spcvmc_sw's direct beam
 * (src/rrtmg_sw_spcvmc.f90:558-582) becomes exp(-x), its Taylor branch included. Not part of
 * the reference: the smooth forward that the adjoint differentiates. */
void orc_set_exact_exp(int on);

/* ---- per-routine entry points (one column), used by the unit-level tests ---- */

/* setcoef_sw -- src/rrtmg_sw_setcoef.f90:49-305. wkl is (39, nlay).
 * iout: laytrop, layswtch, laylow, then jp, jt, jt1, indself, indfor (nlay each).
 * dout: 17 arrays of nlay: colh2o colco2 colo3 coln2o colch4 colo2 colmol co2mult
 *       fac00 fac01 fac10 fac11 selffac selffrac forfac forfrac (16) + spare. */
void orc_setcoef(int nlay, const double *pavel, const double *tavel, const double *coldry,
                 const double *wkl, int *iout, double *dout);

/* setcoef_sw for ncol columns fed as rrtmg_sw feeds it (inatm_sw's coldry / wkl from the profiles,
 * src/rrtmg_sw_rad.f90:1531-1586), with the laysolfr of each band that taumol_sw finds.
 * iout: laytrop (ncol), laylow (ncol), laysolfr (ncol, 14), jp, jt, jt1, indself, indfor
 * (ncol, nlay each). dout: colh2o colco2 colo3 colch4 colo2 colmol fac00 fac01 fac10 fac11 selffac
 * selffrac forfac forfrac (ncol, nlay each). Column index fastest. */
void orc_setcoef_gcm(int ncol, int nlay, const double *play, const double *plev, const double *tlay,
                     const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                     const double *ch4vmr, const double *o2vmr, int *iout, double *dout);

/* taumol_sw -- src/rrtmg_sw_taumol.f90:50-1787 on the outputs of orc_setcoef.
 * taug, taur: (nlay, 112); ssi, sfluxzen: (112). svar_bnd: 3 x 29 (f, s, i). */
void orc_taumol(int nlay, const int *iout, const double *dout, int isolvar, double svar_f,
                double svar_s, double svar_i, const double *svar_bnd, double *ssi,
                double *sfluxzen, double *taug, double *taur);

/* reftra_sw -- src/rrtmg_sw_reftra.f90:48-329 for n layers. */
void orc_reftra(int n, const int *lrtchk, const double *pgg, double prmuz, const double *ptau,
                const double *pw, double *pref, double *prefd, double *ptra, double *ptrad);

/* vrtqdr_sw -- src/rrtmg_sw_vrtqdr.f90:47-171 for klev layers (arrays klev+1 long). */
void orc_vrtqdr(int klev, const double *pref, const double *prefd, const double *ptra,
                const double *ptrad, const double *pdbt, double *prdnd, double *prup,
                double *prupd, const double *ptdbt, double *pfd, double *pfu);

/* cldprmc_sw -- src/rrtmg_sw_cldprmc.f90:53-349 for one column; (112, nlay) in/out. */
int orc_cldprmc(int nlay, int inflag, int iceflag, int liqflag, const double *cldfmc,
                const double *ciwpmc, const double *clwpmc, const double *reicmc,
                const double *relqmc, double *taormc, double *taucmc, double *ssacmc,
                double *asmcmc, const double *fsfcmc);

/* cldprop_sw -- src/rrtmg_sw_cldprop.f90:53-365 for one column. cldfrac, ciwp, clwp, rei, rel
 * (nlay); tauc, ssac, asmc, fsfc (14, nlay); outputs (nlay, 14), column b = band 16 + b. */
int orc_cldprop(int nlay, int inflag, int iceflag, int liqflag, const double *cldfrac,
                const double *tauc, const double *ssac, const double *asmc, const double *fsfc,
                const double *ciwp, const double *clwp, const double *rei, const double *rel,
                double *taucldorig, double *taucloud, double *ssacloud, double *asmcloud);

/* rrtmg_sw of the non-McICA driver, src/rrtmg_sw_rad.nomcica.f90:97-816 (spcvrt_sw,
 * cldprop_sw). Cloud inputs: cldfr, cicewp, cliqwp, reice, reliq (ncol, nlay); taucld, ssacld,
 * asmcld, fsfcld (14, ncol, nlay). Returns 6 ("PARTIAL CLOUD NOT ALLOWED") for a layer
 * cloud fraction strictly between 1e-6 and 1 - 1e-6. */
int orc_rrtmg_sw_nomcica(int ncol, int nlay, int *icld, int *iaer, const double *play,
                         const double *plev, const double *tlay, const double *tlev, const double *tsfc,
                         const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                         const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                         const double *asdir, const double *asdif, const double *aldir,
                         const double *aldif, const double *coszen, double adjes, int dyofyr,
                         double scon, int isolvar, int inflgsw, int iceflgsw, int liqflgsw,
                         const double *cldfr, const double *taucld, const double *ssacld,
                         const double *asmcld, const double *fsfcld, const double *cicewp,
                         const double *cliqwp, const double *reice, const double *reliq,
                         const double *tauaer, const double *ssaaer, const double *asmaer,
                         const double *ecaer, double *swuflx, double *swdflx, double *swhr,
                         double *swuflxc, double *swdflxc, double *swhrc, const double *bndsolvar,
                         const double *indsolvar, const double *solcycfrac);

/* exp lookup table value and the reduced-table accessor (for table-level tests). */
double orc_exp_tbl(int i);
/* Copy a reduced table (after the 224->112 reduction) by name, e.g. "ka16", "kb17",
 * "selfref20", "forref29", "sfluxref17", "irradnce17", "facbrght17", "snsptdrk17",
 * "rayl23", "rayla24", "raylb24", "abso3a24", "abso3b25", "absch420", "absh2o29",
 * "absco229", "rwgt". Returns element count; copies when out != NULL. */
int orc_table(const char *name, double *out);

/* KISS generator step (src/mcica_subcol_gen_sw.f90:721-753) on one state; returns the
 * real draw. MT19937 (src/mcica_random_numbers.f90:77-306): seed and draw n reals. */
double orc_kiss_step(unsigned int *s1, unsigned int *s2, unsigned int *s3, unsigned int *s4);
void orc_mt_reals(int seed, long n, double *out);
unsigned int orc_mt_first_int(int seed, long n);

#ifdef __cplusplus
}
#endif
#endif
