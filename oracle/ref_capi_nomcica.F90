! ref_capi_nomcica.F90 -- C-callable entry points into the compiled reference's non-McICA driver.
!
! TEST INFRASTRUCTURE ONLY: compiled by oracle/build_ref.sh together with the reference's own
! sources (read in place under the reference tree) into oracle/_ref/librrtmg_ref_nomcica.so, a
! library of its own because src/rrtmg_sw_rad.nomcica.f90 defines the same module name
! (rrtmg_sw_rad) as the McICA driver. The tests use it to pin the C oracle's non-McICA path.
! Nothing here is part of the product.
module ref_capi_nomcica
  use iso_c_binding
  use parkind, only : im => kind_im, rb => kind_rb
  use parrrsw, only : nbndsw, naerec
  use rrtmg_sw_init, only : rrtmg_sw_ini
  use rrtmg_sw_rad, only : rrtmg_sw
  use rrtmg_sw_cldprop, only : cldprop_sw
  implicit none
contains

  ! rrtmg_sw_ini(cpdair), src/rrtmg_sw_init.f90:47
  subroutine ref_nm_init(cpdair) bind(C, name="ref_nm_init")
    real(c_double), value :: cpdair
    call rrtmg_sw_ini(cpdair)
  end subroutine ref_nm_init

  ! rrtmg_sw(...), src/rrtmg_sw_rad.nomcica.f90:97; the optional solar-variability arguments
  ! are passed when has_solvar /= 0.
  subroutine ref_rrtmg_sw_nomcica(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, h2ovmr, &
       o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, &
       scon, isolvar, inflgsw, iceflgsw, liqflgsw, cldfr, taucld, ssacld, asmcld, fsfcld, &
       cicewp, cliqwp, reice, reliq, tauaer, ssaaer, asmaer, ecaer, swuflx, swdflx, swhr, &
       swuflxc, swdflxc, swhrc, has_solvar, bndsolvar, indsolvar, solcycfrac) &
       bind(C, name="ref_rrtmg_sw_nomcica")
    integer(c_int), value :: ncol, nlay, dyofyr, isolvar, inflgsw, iceflgsw, liqflgsw, has_solvar
    integer(c_int) :: icld, iaer
    real(c_double), value :: adjes, scon
    real(c_double) :: play(ncol,nlay), plev(ncol,nlay+1), tlay(ncol,nlay), tlev(ncol,nlay+1)
    real(c_double) :: tsfc(ncol), h2ovmr(ncol,nlay), o3vmr(ncol,nlay), co2vmr(ncol,nlay)
    real(c_double) :: ch4vmr(ncol,nlay), n2ovmr(ncol,nlay), o2vmr(ncol,nlay)
    real(c_double) :: asdir(ncol), asdif(ncol), aldir(ncol), aldif(ncol), coszen(ncol)
    real(c_double) :: cldfr(ncol,nlay), taucld(nbndsw,ncol,nlay), ssacld(nbndsw,ncol,nlay)
    real(c_double) :: asmcld(nbndsw,ncol,nlay), fsfcld(nbndsw,ncol,nlay)
    real(c_double) :: cicewp(ncol,nlay), cliqwp(ncol,nlay), reice(ncol,nlay), reliq(ncol,nlay)
    real(c_double) :: tauaer(ncol,nlay,nbndsw), ssaaer(ncol,nlay,nbndsw)
    real(c_double) :: asmaer(ncol,nlay,nbndsw), ecaer(ncol,nlay,naerec)
    real(c_double) :: swuflx(ncol,nlay+1), swdflx(ncol,nlay+1), swhr(ncol,nlay)
    real(c_double) :: swuflxc(ncol,nlay+1), swdflxc(ncol,nlay+1), swhrc(ncol,nlay)
    real(c_double) :: bndsolvar(nbndsw), indsolvar(2), solcycfrac
    integer(im) :: ic, ia
    ic = icld
    ia = iaer
#define ARGS ncol, nlay, ic, ia, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, \
  n2ovmr, o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, isolvar, inflgsw, \
  iceflgsw, liqflgsw, cldfr, taucld, ssacld, asmcld, fsfcld, cicewp, cliqwp, reice, reliq, \
  tauaer, ssaaer, asmaer, ecaer, swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc
    if (has_solvar /= 0) then
      call rrtmg_sw(ARGS, bndsolvar=bndsolvar, indsolvar=indsolvar, solcycfrac=solcycfrac)
    else
      call rrtmg_sw(ARGS)
    end if
#undef ARGS
    icld = ic
    iaer = ia
  end subroutine ref_rrtmg_sw_nomcica

  ! cldprop_sw(...), src/rrtmg_sw_cldprop.f90:53. Its assumed-shape outputs are indexed by band
  ! number (16..29), so they are passed with 29 columns as the driver declares them
  ! (src/rrtmg_sw_rad.nomcica.f90:467-470) and columns 16..29 are returned as (nlay, 14).
  subroutine ref_cldprop(nlay, inflag, iceflag, liqflag, cldfrac, tauc, ssac, asmc, fsfc, ciwp, &
       clwp, rei, rel, taucldorig, taucloud, ssacloud, asmcloud) bind(C, name="ref_cldprop")
    integer(c_int), value :: nlay, inflag, iceflag, liqflag
    real(c_double) :: cldfrac(nlay), ciwp(nlay), clwp(nlay), rei(nlay), rel(nlay)
    real(c_double) :: tauc(nbndsw,nlay), ssac(nbndsw,nlay), asmc(nbndsw,nlay), fsfc(nbndsw,nlay)
    real(c_double) :: taucldorig(nlay,nbndsw), taucloud(nlay,nbndsw)
    real(c_double) :: ssacloud(nlay,nbndsw), asmcloud(nlay,nbndsw)
    real(rb) :: tco(nlay,29), tcl(nlay,29), scl(nlay,29), acl(nlay,29)
    call cldprop_sw(nlay, inflag, iceflag, liqflag, cldfrac, tauc, ssac, asmc, fsfc, ciwp, clwp, &
                    rei, rel, tco, tcl, scl, acl)
    taucldorig = tco(:,16:29)
    taucloud = tcl(:,16:29)
    ssacloud = scl(:,16:29)
    asmcloud = acl(:,16:29)
  end subroutine ref_cldprop
end module ref_capi_nomcica
