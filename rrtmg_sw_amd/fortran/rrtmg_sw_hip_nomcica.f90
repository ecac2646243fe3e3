! rrtmg_sw_hip_nomcica.f90 -- Fortran drop-in for RRTMG_SW's non-McICA GCM interface on the MI355X
! library. The reference ships this driver as its own source (src/rrtmg_sw_rad.nomcica.f90) that
! declares the same module name as the McICA driver, so a GCM builds with one or the other; this
! file likewise builds into its own library (librrtmgsw_fortran_nomcica.so, module files in
! mod_nomcica/) with the same module and routine names and dummies:
!   module rrtmg_sw_init   rrtmg_sw_ini(cpdair)   src/rrtmg_sw_init.f90:47
!   module rrtmg_sw_rad    rrtmg_sw(...)          src/rrtmg_sw_rad.nomcica.f90:97-111
! Each routine forwards to the C-ABI of librrtmgsw_hip.so (include/rrtmgsw.h) through
! iso_c_binding; a non-zero return stops with the reference's STOP message.

module rrtmgsw_capi_nomcica
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
    integer(c_int) function c_rrtmgsw_sw_nomcica(ncol, nlay, icld, iaer, play, plev, tlay, tlev, &
         tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, asdir, asdif, aldir, aldif, coszen, &
         adjes, dyofyr, scon, isolvar, inflgsw, iceflgsw, liqflgsw, cldfr, taucld, ssacld, asmcld, &
         fsfcld, cicewp, cliqwp, reice, reliq, tauaer, ssaaer, asmaer, ecaer, swuflx, swdflx, &
         swhr, swuflxc, swdflxc, swhrc, bndsolvar, indsolvar, solcycfrac) &
         bind(C, name="rrtmgsw_sw_nomcica")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: ncol, nlay, dyofyr, isolvar, inflgsw, iceflgsw, liqflgsw
      integer(c_int) :: icld, iaer
      real(c_double), value :: adjes, scon
      type(c_ptr), value :: play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr
      type(c_ptr), value :: o2vmr, asdir, asdif, aldir, aldif, coszen, cldfr, taucld, ssacld
      type(c_ptr), value :: asmcld, fsfcld, cicewp, cliqwp, reice, reliq, tauaer, ssaaer
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
end module rrtmgsw_capi_nomcica

module rrtmg_sw_init
  use iso_c_binding
  use rrtmgsw_capi_nomcica
  implicit none
  integer, parameter, private :: rb = selected_real_kind(12)
contains
  subroutine rrtmg_sw_ini(cpdair)
    real(kind=rb), intent(in) :: cpdair
    call rrtmgsw_check(c_rrtmgsw_init(real(cpdair, c_double)))
  end subroutine rrtmg_sw_ini
end module rrtmg_sw_init

module rrtmg_sw_rad
  use iso_c_binding
  use rrtmgsw_capi_nomcica
  implicit none
  integer, parameter, private :: rb = selected_real_kind(12), im = selected_int_kind(8)
  public :: rrtmg_sw
contains
  subroutine rrtmg_sw &
            (ncol    ,nlay    ,icld    ,iaer    , &
             play    ,plev    ,tlay    ,tlev    ,tsfc    , &
             h2ovmr  ,o3vmr   ,co2vmr  ,ch4vmr  ,n2ovmr  ,o2vmr, &
             asdir   ,asdif   ,aldir   ,aldif   , &
             coszen  ,adjes   ,dyofyr  ,scon    ,isolvar, &
             inflgsw ,iceflgsw,liqflgsw,cldfr   , &
             taucld  ,ssacld  ,asmcld  ,fsfcld  , &
             cicewp  ,cliqwp  ,reice   ,reliq   , &
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
    real(kind=rb), intent(in), target, contiguous :: cldfr(:,:), taucld(:,:,:), ssacld(:,:,:)
    real(kind=rb), intent(in), target, contiguous :: asmcld(:,:,:), fsfcld(:,:,:)
    real(kind=rb), intent(in), target, contiguous :: cicewp(:,:), cliqwp(:,:), reice(:,:), reliq(:,:)
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
    call rrtmgsw_check(c_rrtmgsw_sw_nomcica(ncol, nlay, ic, ia, c_loc(play), c_loc(plev), &
         c_loc(tlay), c_loc(tlev), c_loc(tsfc), c_loc(h2ovmr), c_loc(o3vmr), c_loc(co2vmr), &
         c_loc(ch4vmr), c_loc(n2ovmr), c_loc(o2vmr), c_loc(asdir), c_loc(asdif), c_loc(aldir), &
         c_loc(aldif), c_loc(coszen), adjes, dyofyr, scon, isolvar, inflgsw, iceflgsw, liqflgsw, &
         c_loc(cldfr), c_loc(taucld), c_loc(ssacld), c_loc(asmcld), c_loc(fsfcld), c_loc(cicewp), &
         c_loc(cliqwp), c_loc(reice), c_loc(reliq), c_loc(tauaer), c_loc(ssaaer), c_loc(asmaer), &
         c_loc(ecaer), c_loc(swuflx), c_loc(swdflx), c_loc(swhr), c_loc(swuflxc), c_loc(swdflxc), &
         c_loc(swhrc), pb, pi, pf))
    icld = ic
    iaer = ia
  end subroutine rrtmg_sw
end module rrtmg_sw_rad
