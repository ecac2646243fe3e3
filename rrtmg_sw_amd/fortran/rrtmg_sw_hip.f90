! rrtmg_sw_hip.f90 -- Fortran drop-in for RRTMG_SW's GCM-mode McICA interface on the MI355X library.
!
! Keeps the reference's module names, routine names and dummy arguments (order, kinds, intents,
! optional arguments) so that a host GCM recompiles against these modules unchanged:
!   module rrtmg_sw_init       rrtmg_sw_ini(cpdair)            src/rrtmg_sw_init.f90:47
!   module mcica_subcol_gen_sw get_alpha(...), mcica_subcol_sw(...)  src/mcica_subcol_gen_sw.f90:66,181
!   module rrtmg_sw_rad        rrtmg_sw(...)                   src/rrtmg_sw_rad.f90:99
! Each routine forwards to the C-ABI of librrtmgsw_hip.so (include/rrtmgsw.h) through
! iso_c_binding. Dummies are declared contiguous, so non-contiguous actual arguments are copied
! in and out by the compiler. A non-zero return stops with the reference's STOP message.

module rrtmgsw_capi
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function c_rrtmgsw_init(cpdair) bind(C, name="rrtmgsw_init")
      import :: c_int, c_double
      real(c_double), value :: cpdair
    end function
    type(c_ptr) function c_rrtmgsw_last_error() bind(C, name="rrtmgsw_last_error")
      import :: c_ptr
    end function
    integer(c_int) function c_rrtmgsw_get_alpha(iplon, ncol, nlayers, icld, idcor, decorr_con, &
         dz, lat, juldat, cldfrac, alpha) bind(C, name="rrtmgsw_get_alpha")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: iplon, ncol, nlayers, icld, idcor, juldat
      real(c_double), value :: decorr_con
      type(c_ptr), value :: dz, lat, cldfrac, alpha
    end function
    integer(c_int) function c_rrtmgsw_mcica_subcol(iplon, ncol, nlay, icld, permuteseed, irng, &
         play, cldfrac, ciwp, clwp, rei, rel, tauc, ssac, asmc, fsfc, alpha, cldfmcl, ciwpmcl, &
         clwpmcl, reicmcl, relqmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl) &
         bind(C, name="rrtmgsw_mcica_subcol")
      import :: c_int, c_ptr
      integer(c_int), value :: iplon, ncol, nlay, icld, permuteseed
      integer(c_int) :: irng
      type(c_ptr), value :: play, cldfrac, ciwp, clwp, rei, rel, tauc, ssac, asmc, fsfc, alpha
      type(c_ptr), value :: cldfmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl, taucmcl, ssacmcl
      type(c_ptr), value :: asmcmcl, fsfcmcl
    end function
    integer(c_int) function c_rrtmgsw_sw(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, &
         h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, &
         dyofyr, scon, isolvar, inflgsw, iceflgsw, liqflgsw, cldfmcl, taucmcl, ssacmcl, asmcmcl, &
         fsfcmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl, tauaer, ssaaer, asmaer, ecaer, swuflx, &
         swdflx, swhr, swuflxc, swdflxc, swhrc, bndsolvar, indsolvar, solcycfrac) &
         bind(C, name="rrtmgsw_sw")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: ncol, nlay, dyofyr, isolvar, inflgsw, iceflgsw, liqflgsw
      integer(c_int) :: icld, iaer
      real(c_double), value :: adjes, scon
      type(c_ptr), value :: play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr
      type(c_ptr), value :: o2vmr, asdir, asdif, aldir, aldif, coszen, cldfmcl, taucmcl, ssacmcl
      type(c_ptr), value :: asmcmcl, fsfcmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl, tauaer, ssaaer
      type(c_ptr), value :: asmaer, ecaer, swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc
      type(c_ptr), value :: bndsolvar, indsolvar, solcycfrac
    end function
  end interface
contains
  ! Stop with the library's message, as the reference's STOP statements do.
  subroutine rrtmgsw_check(rc)
    integer(c_int), intent(in) :: rc
    character(kind=c_char), pointer :: msg(:)
    character(len=256) :: text
    integer :: i
    if (rc == 0) return
    call c_f_pointer(c_rrtmgsw_last_error(), msg, [256])
    text = ' '
    do i = 1, 256
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    error stop trim(text)
  end subroutine rrtmgsw_check
end module rrtmgsw_capi

module rrtmg_sw_init
  use iso_c_binding
  use rrtmgsw_capi
  implicit none
  integer, parameter, private :: rb = selected_real_kind(12)
contains
  subroutine rrtmg_sw_ini(cpdair)
    real(kind=rb), intent(in) :: cpdair
    call rrtmgsw_check(c_rrtmgsw_init(real(cpdair, c_double)))
  end subroutine rrtmg_sw_ini
end module rrtmg_sw_init

module mcica_subcol_gen_sw
  use iso_c_binding
  use rrtmgsw_capi
  implicit none
  integer, parameter, private :: rb = selected_real_kind(12), im = selected_int_kind(8)
contains
  subroutine get_alpha(iplon, ncol, nlayers, icld, idcor, decorr_con, dz, lat, juldat, cldfrac, alpha)
    integer(kind=im), intent(in) :: iplon, ncol, nlayers, icld, idcor, juldat
    real(kind=rb), intent(in) :: decorr_con
    real(kind=rb), intent(in), target, contiguous :: dz(:,:), lat(:), cldfrac(:,:)
    real(kind=rb), intent(out), target, contiguous :: alpha(:,:)
    call rrtmgsw_check(c_rrtmgsw_get_alpha(iplon, ncol, nlayers, icld, idcor, decorr_con, &
         c_loc(dz), c_loc(lat), juldat, c_loc(cldfrac), c_loc(alpha)))
  end subroutine get_alpha

  subroutine mcica_subcol_sw(iplon, ncol, nlay, icld, permuteseed, irng, play, &
                             cldfrac, ciwp, clwp, rei, rel, tauc, ssac, asmc, fsfc, &
                             alpha, &
                             cldfmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl, &
                             taucmcl, ssacmcl, asmcmcl, fsfcmcl)
    integer(kind=im), intent(in) :: iplon, ncol, nlay, icld, permuteseed
    integer(kind=im), intent(inout) :: irng
    real(kind=rb), intent(in), target, contiguous :: play(:,:), cldfrac(:,:), tauc(:,:,:)
    real(kind=rb), intent(in), target, contiguous :: ssac(:,:,:), asmc(:,:,:), fsfc(:,:,:)
    real(kind=rb), intent(in), target, contiguous :: ciwp(:,:), clwp(:,:), rei(:,:), rel(:,:)
    real(kind=rb), intent(in), target, contiguous :: alpha(:,:)
    real(kind=rb), intent(out), target, contiguous :: cldfmcl(:,:,:), ciwpmcl(:,:,:), clwpmcl(:,:,:)
    real(kind=rb), intent(out), target, contiguous :: relqmcl(:,:), reicmcl(:,:)
    real(kind=rb), intent(out), target, contiguous :: taucmcl(:,:,:), ssacmcl(:,:,:)
    real(kind=rb), intent(out), target, contiguous :: asmcmcl(:,:,:), fsfcmcl(:,:,:)
    integer(c_int) :: ir
    ir = irng
    call rrtmgsw_check(c_rrtmgsw_mcica_subcol(iplon, ncol, nlay, icld, permuteseed, ir, &
         c_loc(play), c_loc(cldfrac), c_loc(ciwp), c_loc(clwp), c_loc(rei), c_loc(rel), &
         c_loc(tauc), c_loc(ssac), c_loc(asmc), c_loc(fsfc), c_loc(alpha), c_loc(cldfmcl), &
         c_loc(ciwpmcl), c_loc(clwpmcl), c_loc(reicmcl), c_loc(relqmcl), c_loc(taucmcl), &
         c_loc(ssacmcl), c_loc(asmcmcl), c_loc(fsfcmcl)))
    irng = ir
  end subroutine mcica_subcol_sw
end module mcica_subcol_gen_sw

module rrtmg_sw_rad
  use iso_c_binding
  use rrtmgsw_capi
  implicit none
  integer, parameter, private :: rb = selected_real_kind(12), im = selected_int_kind(8)
  public :: rrtmg_sw
contains
  subroutine rrtmg_sw &
            (ncol    ,nlay    ,icld    ,iaer    , &
             play    ,plev    ,tlay    ,tlev    ,tsfc   , &
             h2ovmr , o3vmr   ,co2vmr  ,ch4vmr  ,n2ovmr ,o2vmr , &
             asdir   ,asdif   ,aldir   ,aldif   , &
             coszen  ,adjes   ,dyofyr  ,scon    ,isolvar, &
             inflgsw ,iceflgsw,liqflgsw,cldfmcl , &
             taucmcl ,ssacmcl ,asmcmcl ,fsfcmcl , &
             ciwpmcl ,clwpmcl ,reicmcl ,relqmcl , &
             tauaer  ,ssaaer  ,asmaer  ,ecaer   , &
             swuflx  ,swdflx  ,swhr    ,swuflxc ,swdflxc ,swhrc, &
             bndsolvar,indsolvar,solcycfrac)
    integer(kind=im), intent(in) :: ncol, nlay
    integer(kind=im), intent(inout) :: icld, iaer
    real(kind=rb), intent(in), target, contiguous :: play(:,:), plev(:,:), tlay(:,:), tlev(:,:)
    real(kind=rb), intent(in), target, contiguous :: tsfc(:), h2ovmr(:,:), o3vmr(:,:)
    real(kind=rb), intent(in), target, contiguous :: co2vmr(:,:), ch4vmr(:,:), n2ovmr(:,:)
    real(kind=rb), intent(in), target, contiguous :: o2vmr(:,:), asdir(:), aldir(:), asdif(:)
    real(kind=rb), intent(in), target, contiguous :: aldif(:), coszen(:)
    integer(kind=im), intent(in) :: dyofyr, isolvar, inflgsw, iceflgsw, liqflgsw
    real(kind=rb), intent(in) :: adjes, scon
    real(kind=rb), intent(in), optional, target :: indsolvar(:), bndsolvar(:), solcycfrac
    real(kind=rb), intent(in), target, contiguous :: cldfmcl(:,:,:), taucmcl(:,:,:)
    real(kind=rb), intent(in), target, contiguous :: ssacmcl(:,:,:), asmcmcl(:,:,:)
    real(kind=rb), intent(in), target, contiguous :: fsfcmcl(:,:,:), ciwpmcl(:,:,:), clwpmcl(:,:,:)
    real(kind=rb), intent(in), target, contiguous :: reicmcl(:,:), relqmcl(:,:)
    real(kind=rb), intent(in), target, contiguous :: tauaer(:,:,:), ssaaer(:,:,:)
    real(kind=rb), intent(in), target, contiguous :: asmaer(:,:,:), ecaer(:,:,:)
    real(kind=rb), intent(out), target, contiguous :: swuflx(:,:), swdflx(:,:), swhr(:,:)
    real(kind=rb), intent(out), target, contiguous :: swuflxc(:,:), swdflxc(:,:), swhrc(:,:)
    real(kind=rb), target :: bsv(14), isv(2), scf
    type(c_ptr) :: pb, pi, pf
    integer(c_int) :: ic, ia
    pb = c_null_ptr
    pi = c_null_ptr
    pf = c_null_ptr
    if (present(bndsolvar)) then
      bsv = bndsolvar(1:14)
      pb = c_loc(bsv)
    end if
    if (present(indsolvar)) then
      isv = indsolvar(1:2)
      pi = c_loc(isv)
    end if
    if (present(solcycfrac)) then
      scf = solcycfrac
      pf = c_loc(scf)
    end if
    ic = icld
    ia = iaer
    call rrtmgsw_check(c_rrtmgsw_sw(ncol, nlay, ic, ia, c_loc(play), c_loc(plev), c_loc(tlay), &
         c_loc(tlev), c_loc(tsfc), c_loc(h2ovmr), c_loc(o3vmr), c_loc(co2vmr), c_loc(ch4vmr), &
         c_loc(n2ovmr), c_loc(o2vmr), c_loc(asdir), c_loc(asdif), c_loc(aldir), c_loc(aldif), &
         c_loc(coszen), adjes, dyofyr, scon, isolvar, inflgsw, iceflgsw, liqflgsw, &
         c_loc(cldfmcl), c_loc(taucmcl), c_loc(ssacmcl), c_loc(asmcmcl), c_loc(fsfcmcl), &
         c_loc(ciwpmcl), c_loc(clwpmcl), c_loc(reicmcl), c_loc(relqmcl), c_loc(tauaer), &
         c_loc(ssaaer), c_loc(asmaer), c_loc(ecaer), c_loc(swuflx), c_loc(swdflx), c_loc(swhr), &
         c_loc(swuflxc), c_loc(swdflxc), c_loc(swhrc), pb, pi, pf))
    icld = ic
    iaer = ia
  end subroutine rrtmg_sw
end module rrtmg_sw_rad
