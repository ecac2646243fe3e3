/*
 * rrtmgsw.h -- C-ABI of the MI355X shortwave library (librrtmgsw_hip.so).
 *
 * Drop-in boundary for RRTMG_SW's GCM-mode McICA path. Every array is a plain pointer to
 * column-major FP64 data in the reference's Fortran layout (the column index fastest), exactly
 * the dummies of the reference routines each entry point replaces, in the same order; integer
 * flags are int32 (kind_im). A Fortran host binds these through iso_c_binding
 * (rrtmg_sw_amd/fortran/rrtmg_sw_hip.f90 keeps the reference's module and routine names).
 *
 * Return value: 0 on success, otherwise an error code; rrtmgsw_last_error() returns the
 * reference's STOP message for the condition (e.g. "ICE RADIUS OUT OF BOUNDS"). Flags the
 * reference clamps silently are clamped the same way and written back (icld, iaer, irng).
 *
 * Host-array entry points copy inputs to the GPU and outputs back (synchronous). The *_dev
 * entry points take device pointers and a hipStream_t (passed as void*, NULL = default stream)
 * and are asynchronous except for the final status check, which synchronizes the stream.
 * Threads: after rrtmgsw_init every entry point may be called from several host threads at
 * once. Calls on different streams overlap; calls on one stream (the host-array and column
 * entry points use the default stream) run one at a time, since a call's scratch, status and
 * generator buffers belong to its stream until it returns. rrtmgsw_init and rrtmgsw_finalize
 * must not overlap other calls.
 */
#ifndef RRTMGSW_H
#define RRTMGSW_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes */
#define RRTMGSW_OK 0
#define RRTMGSW_ERR_NOT_INITIALIZED 1
#define RRTMGSW_ERR_ARGUMENT 2
#define RRTMGSW_ERR_INFLAG 3       /* "INFLAG = 1 OPTION NOT AVAILABLE WITH MCICA" */
#define RRTMGSW_ERR_RADIUS 4       /* ice / liquid particle size out of the table bounds */
#define RRTMGSW_ERR_OPTICS 5       /* cloud optical property outside its physical limits */
#define RRTMGSW_ERR_PARTIAL_CLOUD 6 /* "PARTIAL CLOUD NOT ALLOWED" (non-McICA driver) */
#define RRTMGSW_ERR_ICLD 10        /* "MCICA_SUBCOL: INVALID ICLD" */
#define RRTMGSW_ERR_KISS_SEED 11   /* "... REQUIRES PMID FROM BOTTOM FOUR LAYERS." */
#define RRTMGSW_ERR_HIP 20         /* a HIP runtime call failed */

/* Precision forms of the device-resident entry points. */
#define RRTMGSW_FP64 0
#define RRTMGSW_MIXED 1
/* FP64 with exact exponentials in place of the reference's lookup table: the smooth forward that
 * rrtmgsw_mcica_sw_adjoint_dev differentiates (for finite-difference checks). */
#define RRTMGSW_FP64_EXP 2

/* Replaces rrtmg_sw_ini(cpdair), src/rrtmg_sw_init.f90:47-173: builds the 112-point k-tables
 * (224 -> 112 reduction), the exponential lookup table and the cloud optics tables, and
 * uploads them to the current GPU. Must precede every other call. */
int rrtmgsw_init(double cpdair);

/* Replaces rrtmg_sw_ini(cpdair) of a build with src/rrtmg_sw_read_nc.f90 in place of
 * src/rrtmg_sw_k_g.f90: the k-distribution and solar-source data of the fourteen bands are read
 * from the netCDF file `path` (NULL or "": "rrtmg_sw.nc" in the working directory, the name
 * read_nc opens) with the start/count of each nf90_get_var there (sw_kgb16 .. sw_kgb29,
 * :46-1047); the rest as rrtmgsw_init. Classic, 64-bit-offset and 64-bit-data files are read; a
 * netCDF-4 (HDF5) file is refused with a message. A missing variable or a dimension too short
 * for a read fails with "Error reading band NN variables from file", read_nc's STOP text. The
 * tables of one process come from one source: rrtmgsw_init and rrtmgsw_init_nc with another
 * source fail until rrtmgsw_finalize. */
int rrtmgsw_init_nc(double cpdair, const char *path);

/* Host-only (no GPU call): builds the reduced tables from `ktable_file` (NULL: the compiled-in
 * data) and writes up to `cap` doubles of them to `out` -- the k, solar, cloud and aerosol table
 * image, then per g-point (112) the Rayleigh coefficients of the lower and upper atmosphere and
 * the minor-absorber coefficients of both -- and their total count to *n. No counterpart in the
 * reference, whose tables are module variables a Fortran program reads directly; it lets a
 * k-table file be checked without a GPU. */
int rrtmgsw_host_tables(const char *ktable_file, double *out, size_t cap, size_t *n);

/* Host-only: reads the hyperslab start[d] .. start[d] + count[d] - 1 (dimensions slowest first,
 * 0-based) of variable `name` of the netCDF file `path` as doubles, row-major over count -- the
 * nc_get_vara_double under the nf90_get_var calls of src/rrtmg_sw_read_nc.f90, for the classic
 * formats that rrtmgsw_init_nc reads. `ndims` must equal the variable's rank. */
int rrtmgsw_nc_get_vara_double(const char *path, const char *name, int ndims, const size_t *start,
                               const size_t *count, double *out);

/* Frees the device tables (no counterpart in the reference, whose tables are module state). */
int rrtmgsw_finalize(void);

/* Message for the last non-zero return on this thread. */
const char *rrtmgsw_last_error(void);

/* Replaces rrtmg_sw(...), src/rrtmg_sw_rad.f90:99-111 (dummies :220-388). bndsolvar (14),
 * indsolvar (2) and solcycfrac (1) are the optional arguments: NULL when absent. */
int rrtmgsw_sw(int ncol, int nlay, int *icld, int *iaer,
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
               const double *indsolvar, const double *solcycfrac);

/* Same as rrtmgsw_sw with device pointers (outputs and inputs resident in HBM). The scalar
 * optional inputs bndsolvar / indsolvar / solcycfrac stay host pointers (NULL = absent). */
int rrtmgsw_sw_dev(int ncol, int nlay, int *icld, int *iaer,
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
                   const double *indsolvar, const double *solcycfrac, void *stream);

/* Replaces rrtmg_sw(...) of the non-McICA driver, src/rrtmg_sw_rad.nomcica.f90:97-111
 * (dummies :209-376; cldprop_sw, spcvrt_sw). Layer clouds: cldfr, cicewp, cliqwp, reice,
 * reliq (ncol, nlay); taucld, ssacld, asmcld, fsfcld (nbndsw = 14, ncol, nlay). A layer cloud
 * fraction strictly between 1e-6 and 1 - 1e-6 returns RRTMGSW_ERR_PARTIAL_CLOUD (the
 * reference's STOP 'PARTIAL CLOUD NOT ALLOWED', :617-622). icld outside 0..3 becomes 2 and iaer
 * outside {0, 6, 10} becomes 0, written back (:560, :568). Host arrays, as rrtmgsw_sw. */
int rrtmgsw_sw_nomcica(int ncol, int nlay, int *icld, int *iaer,
                       const double *play, const double *plev, const double *tlay,
                       const double *tlev, const double *tsfc, const double *h2ovmr,
                       const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                       const double *n2ovmr, const double *o2vmr, const double *asdir,
                       const double *asdif, const double *aldir, const double *aldif,
                       const double *coszen, double adjes, int dyofyr, double scon, int isolvar,
                       int inflgsw, int iceflgsw, int liqflgsw, const double *cldfr,
                       const double *taucld, const double *ssacld, const double *asmcld,
                       const double *fsfcld, const double *cicewp, const double *cliqwp,
                       const double *reice, const double *reliq, const double *tauaer,
                       const double *ssaaer, const double *asmaer, const double *ecaer,
                       double *swuflx, double *swdflx, double *swhr, double *swuflxc,
                       double *swdflxc, double *swhrc, const double *bndsolvar,
                       const double *indsolvar, const double *solcycfrac);

/* Same as rrtmgsw_sw_nomcica with device pointers (inputs and outputs resident in HBM); the
 * optional solar-variability inputs stay host pointers (NULL = absent). */
int rrtmgsw_sw_nomcica_dev(int ncol, int nlay, int *icld, int *iaer,
                           const double *play, const double *plev, const double *tlay,
                           const double *tlev, const double *tsfc, const double *h2ovmr,
                           const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                           const double *n2ovmr, const double *o2vmr, const double *asdir,
                           const double *asdif, const double *aldir, const double *aldif,
                           const double *coszen, double adjes, int dyofyr, double scon,
                           int isolvar, int inflgsw, int iceflgsw, int liqflgsw,
                           const double *cldfr, const double *taucld, const double *ssacld,
                           const double *asmcld, const double *fsfcld, const double *cicewp,
                           const double *cliqwp, const double *reice, const double *reliq,
                           const double *tauaer, const double *ssaaer, const double *asmaer,
                           const double *ecaer, double *swuflx, double *swdflx, double *swhr,
                           double *swuflxc, double *swdflxc, double *swhrc,
                           const double *bndsolvar, const double *indsolvar,
                           const double *solcycfrac, void *stream);

/* The column program's radiative transfer without McICA (src/rrtmg_sw.1col.f90:540-1000 with
 * imca = 0: cldprop_sw, setcoef_sw, spcvrt_sw), for the column-mode I/O driver
 * (rrtmg_sw_amd/column/rrtmgsw_column.cpp). Inputs are the layer quantities that program hands to
 * setcoef_sw / spcvrt_sw: pavel, tavel, coldry (ncol, nlay) and pz (ncol, nlay+1), surface first;
 * wkl (7, ncol, nlay) molecules/cm2 (H2O, CO2, O3, N2O, CO, CH4, O2); albedo (14, ncol) =
 * 1 - semiss; coszen (ncol) = cos(SZA) (0 becomes 1e-10, :644-645); the solar inputs as for
 * rrtmgsw_sw (juldat as dyofyr, adjes = 1); clouds as cldprop_sw takes them, (ncol, nlay) and
 * (14, ncol, nlay), read when icld >= 1, fractions 0 or 1 only (:612-616); aerosol optics
 * (ncol, nlay, 14) or NULL. idelm 0/1: direct beam without / with delta scaling. Outputs fup,
 * fdn, fdir (ncol, nlay+1, 15): total-sky upward, downward and direct downward flux, block 0
 * the broadband and blocks 1..14 bands 16..29 (the program's iout = 98 blocks, :824-997). Host
 * arrays. */
int rrtmgsw_column_sw(int ncol, int nlay, int icld, int idelm, const double *pavel,
                      const double *tavel, const double *pz, const double *coldry,
                      const double *wkl, const double *albedo, const double *coszen, double adjes,
                      int dyofyr, double scon, int isolvar, const double *bndsolvar,
                      const double *indsolvar, const double *solcycfrac, int inflag, int iceflag,
                      int liqflag, const double *cldfrac, const double *ciwp, const double *clwp,
                      const double *rei, const double *rel, const double *tauc, const double *ssac,
                      const double *asmc, const double *fsfc, const double *tauaer,
                      const double *ssaaer, const double *asmaer, double *fup, double *fdn,
                      double *fdir);

/* The column program's McICA loop (src/rrtmg_sw.1col.f90:566-610 with imca = 1): nsamples draws
 * of one column, sample s (ims = s + 1) with permuteseed (ims + 1) * 112 as
 * src/mcica_subcol_gen_sw.1col.f90 sets it, irng 0 (KISS) or 1 (Mersenne Twister, what the
 * program uses), overlap icld 1-5 with alpha (nlay) for icld 4/5, then cldprmc_sw and spcvmc_sw;
 * all samples run in one launch. Column inputs as rrtmgsw_column_sw for one column (coszen a
 * scalar); fup, fdn, fdir (nsamples, nlay+1, 15) per sample, which the caller averages as the
 * program does (:874-894). Host arrays. */
int rrtmgsw_column_mcica_sw(int nlay, int nsamples, int icld, int irng, int idelm,
                            const double *pavel, const double *tavel, const double *pz,
                            const double *coldry, const double *wkl, const double *albedo,
                            double coszen, double adjes, int dyofyr, double scon, int isolvar,
                            const double *bndsolvar, const double *indsolvar,
                            const double *solcycfrac, int inflag, int iceflag, int liqflag,
                            const double *cldfrac, const double *ciwp, const double *clwp,
                            const double *rei, const double *rel, const double *tauc,
                            const double *ssac, const double *asmc, const double *fsfc,
                            const double *alpha, const double *tauaer, const double *ssaaer,
                            const double *asmaer, double *fup, double *fdn, double *fdir);

/* Replaces mcica_subcol_sw(...), src/mcica_subcol_gen_sw.f90:181-185 (dummies :200-246).
 * Writes the seven (112, ncol, nlay) sub-column arrays and reicmcl / relqmcl (ncol, nlay).
 * irng is clamped (non-zero -> 1) and written back, as src/mcica_subcol_gen_sw.f90:450 does.
 * The overlap order of icld 4/5 with irng = 0 follows the compiling single-column generator
 * (src/mcica_subcol_gen_sw.1col.f90:593-600): per column, CDF then CDF2 for each layer. */
int rrtmgsw_mcica_subcol(int iplon, int ncol, int nlay, int icld, int permuteseed, int *irng,
                         const double *play, const double *cldfrac, const double *ciwp,
                         const double *clwp, const double *rei, const double *rel,
                         const double *tauc, const double *ssac, const double *asmc,
                         const double *fsfc, const double *alpha, double *cldfmcl,
                         double *ciwpmcl, double *clwpmcl, double *reicmcl, double *relqmcl,
                         double *taucmcl, double *ssacmcl, double *asmcmcl, double *fsfcmcl);

/* Same as rrtmgsw_mcica_subcol with device pointers. */
int rrtmgsw_mcica_subcol_dev(int iplon, int ncol, int nlay, int icld, int permuteseed, int *irng,
                             const double *play, const double *cldfrac, const double *ciwp,
                             const double *clwp, const double *rei, const double *rel,
                             const double *tauc, const double *ssac, const double *asmc,
                             const double *fsfc, const double *alpha, double *cldfmcl,
                             double *ciwpmcl, double *clwpmcl, double *reicmcl,
                             double *relqmcl, double *taucmcl, double *ssacmcl,
                             double *asmcmcl, double *fsfcmcl, void *stream);

/* Replaces get_alpha(...), src/mcica_subcol_gen_sw.f90:66-67 (host arrays; host-side work in
 * the reference's API as well). The reference's missing THEN at :147 is read as intended. */
int rrtmgsw_get_alpha(int iplon, int ncol, int nlayers, int icld, int idcor, double decorr_con,
                      const double *dz, const double *lat, int juldat, const double *cldfrac,
                      double *alpha);

/* Replaces setcoef_sw(...), src/rrtmg_sw_setcoef.f90:49-305, for ncol columns at once on device
 * pointers, fed as rrtmg_sw feeds it: inatm_sw's dry-air and molecular amounts
 * (src/rrtmg_sw_rad.f90:1531-1586) are formed from play, plev (ncol, nlay+1), tlay and the volume
 * mixing ratios (ncol, nlay), as the forward kernels form them. It runs the forward kernels' own
 * setcoef code and writes its results, in the reference's dummy order: laytrop and laylow (ncol),
 * then jp, jt, jt1 (1-based), the column amounts (molecules / cm2 times 1e-20), fac00..fac11, the
 * self- and foreign-continuum factors, fractions and indices (ncol, nlay) each; and last laysolfr
 * (ncol, 14, 1-based), the layer at which taumol_sw takes each band's solar source
 * (src/rrtmg_sw_taumol.f90, e.g. :351-356, :559-564). Any output may be NULL (not written).
 * Not produced: layswtch (always 0 in the reference), coln2o and co2mult, which no later step of
 * the path reads. Synchronizes the stream before returning. */
int rrtmgsw_setcoef_dev(int ncol, int nlay, const double *play, const double *plev, const double *tlay,
                        const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                        const double *ch4vmr, const double *o2vmr, int *laytrop, int *laylow,
                        int *jp, int *jt, int *jt1, double *colch4, double *colco2, double *colh2o,
                        double *colmol, double *colo2, double *colo3, double *fac00, double *fac01,
                        double *fac10, double *fac11, double *selffac, double *selffrac,
                        int *indself, double *forfac, double *forfrac, int *indfor, int *laysolfr,
                        void *stream);

/* mcica_subcol_sw followed by rrtmg_sw on device pointers, without materializing the seven
 * (112, ncol, nlay) sub-column arrays: the sub-column masks are generated inside the radiative
 * transfer kernel from the same KISS / Mersenne-Twister streams (no counterpart in the
 * reference, which always materializes them; the results are those of calling
 * rrtmgsw_mcica_subcol_dev then rrtmgsw_sw_dev). icld is the mcica_subcol_sw overlap flag
 * (0..5); precision is RRTMGSW_FP64, RRTMGSW_MIXED or RRTMGSW_FP64_EXP. */
int rrtmgsw_mcica_sw_dev(int ncol, int nlay, int icld, int iaer, int permuteseed, int *irng,
                         const double *play, const double *plev, const double *tlay,
                         const double *tlev, const double *tsfc, const double *h2ovmr,
                         const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                         const double *n2ovmr, const double *o2vmr, const double *asdir,
                         const double *asdif, const double *aldir, const double *aldif,
                         const double *coszen, double adjes, int dyofyr, double scon,
                         int isolvar, int inflgsw, int iceflgsw, int liqflgsw,
                         const double *cldfrac, const double *ciwp, const double *clwp,
                         const double *rei, const double *rel, const double *tauc,
                         const double *ssac, const double *asmc, const double *fsfc,
                         const double *alpha, const double *tauaer, const double *ssaaer,
                         const double *asmaer, const double *ecaer, double *swuflx,
                         double *swdflx, double *swhr, double *swuflxc, double *swdflxc,
                         double *swhrc, const double *bndsolvar, const double *indsolvar,
                         const double *solcycfrac, int precision, void *stream);

/* Bytes of the device array rrtmgsw_mcica_sw_dev_noalloc works in (one per GPU, 256 MiB). */
#define RRTMGSW_NOALLOC_SCRATCH_BYTES ((size_t)256 * 1024 * 1024)

/* rrtmgsw_mcica_sw_dev that allocates nothing: no counterpart in the reference. It is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments, results and errors as rrtmgsw_mcica_sw_dev;
 * the caller still passes every input and output. Its working memory -- the status word, the
 * KISS jump records, the Mersenne Twister stream of irng = 1 and the per-column elimination
 * coefficients -- lives in one file-scope device array of RRTMGSW_NOALLOC_SCRATCH_BYTES per GPU,
 * set up by rrtmgsw_init on the device current then (the entry point fails on any other); the
 * columns run in launches of as many as the array holds (640 at 72 layers). A Mersenne Twister
 * stream (112 x ncol x nlay words, twice that for icld 4/5) that leaves no room for one column
 * is refused with RRTMGSW_ERR_ARGUMENT. The array keeps the last call's contents; calls share
 * it, so they run one at a time (a lock), and the call synchronizes its stream before returning. */
int rrtmgsw_mcica_sw_dev_noalloc(int ncol, int nlay, int icld, int iaer, int permuteseed, int *irng,
                                 const double *play, const double *plev, const double *tlay,
                                 const double *tlev, const double *tsfc, const double *h2ovmr,
                                 const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                                 const double *n2ovmr, const double *o2vmr, const double *asdir,
                                 const double *asdif, const double *aldir, const double *aldif,
                                 const double *coszen, double adjes, int dyofyr, double scon,
                                 int isolvar, int inflgsw, int iceflgsw, int liqflgsw,
                                 const double *cldfrac, const double *ciwp, const double *clwp,
                                 const double *rei, const double *rel, const double *tauc,
                                 const double *ssac, const double *asmc, const double *fsfc,
                                 const double *alpha, const double *tauaer, const double *ssaaer,
                                 const double *asmaer, const double *ecaer, double *swuflx,
                                 double *swdflx, double *swhr, double *swuflxc, double *swdflxc,
                                 double *swhrc, const double *bndsolvar, const double *indsolvar,
                                 const double *solcycfrac, int precision, void *stream);

/* Adjoint of rrtmgsw_mcica_sw_dev (precision RRTMGSW_FP64_EXP) for the objective
 *   J = sum_c seed_toa(c) * swuflx(c, nlay+1) + sum_(c,k) seed_hr(c, k) * swhr(c, k)
 * (total sky; seeds are device arrays (ncol) and (ncol, nlay)). Writes dJ/d tlay, h2ovmr,
 * o3vmr, clwp and ciwp (the in-cloud water paths of mcica_subcol_sw), each (ncol, nlay), and
 * dJ/d (asdir, asdif, aldir, aldif) as d_alb (ncol, 4). McICA masks are held fixed. ckpt = 0
 * stores every layer's gas and Rayleigh optical depths, adding-method state and state adjoint
 * (6 doubles per layer and g-point); ckpt = K > 0 stores the optical depths and the state
 * adjoint, keeps the adding-method state of one layer in K and re-sweeps each K-layer segment
 * from it in the reverse sweep (3 + 3/K doubles per layer and g-point plus 3 K per g-point, one
 * more two-stream evaluation per layer). Both add 9 doubles per layer and band. No counterpart
 * in the reference; inputs as rrtmgsw_mcica_sw_dev (same argument order). */
int rrtmgsw_mcica_sw_adjoint_dev(int ncol, int nlay, int icld, int iaer, int permuteseed, int *irng,
                                 const double *play, const double *plev, const double *tlay,
                                 const double *tlev, const double *tsfc, const double *h2ovmr,
                                 const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                                 const double *n2ovmr, const double *o2vmr, const double *asdir,
                                 const double *asdif, const double *aldir, const double *aldif,
                                 const double *coszen, double adjes, int dyofyr, double scon,
                                 int isolvar, int inflgsw, int iceflgsw, int liqflgsw,
                                 const double *cldfrac, const double *ciwp, const double *clwp,
                                 const double *rei, const double *rel, const double *tauc,
                                 const double *ssac, const double *asmc, const double *fsfc,
                                 const double *alpha, const double *tauaer, const double *ssaaer,
                                 const double *asmaer, const double *ecaer,
                                 const double *bndsolvar, const double *indsolvar,
                                 const double *solcycfrac, const double *seed_toa,
                                 const double *seed_hr, double *d_tlay, double *d_h2ovmr,
                                 double *d_o3vmr, double *d_clwp, double *d_ciwp, double *d_alb,
                                 int ckpt, void *stream);

/* Bytes of device work area per column that rrtmgsw_mcica_sw_adjoint_dev uses for nlay layers
 * and checkpoint interval ckpt (it processes columns in chunks of at most 4 GiB of work area).
 * No counterpart in the reference. */
size_t rrtmgsw_adjoint_work_bytes(int nlay, int ckpt);

#ifdef __cplusplus
}
#endif
#endif
