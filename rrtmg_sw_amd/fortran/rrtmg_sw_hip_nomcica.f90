! rrtmg_sw_hip_nomcica.f90 -- Fortran drop-in for RRTMG_SW's non-McICA GCM interface on the MI355X
! library. The reference ships this driver as its own source (src/rrtmg_sw_rad.nomcica.f90) that
! declares the same module name as the McICA driver, so a GCM builds with one or the other; this
! file likewise builds into its own library (librrtmgsw_fortran_nomcica.so, module files in
! mod_nomcica/) with the same module and routine names and dummies:
!   module rrtmg_sw_init   rrtmg_sw_ini(cpdair)   src/rrtmg_sw_init.f90:47
!   module rrtmg_sw_rad    rrtmg_sw(...)          src/rrtmg_sw_rad.nomcica.f90:97-111
! Each routine forwards to the C-ABI of librrtmgsw_hip.so (include/rrtmgsw.h) through
! iso_c_binding; arrays larger than the counts (pcols-sized arrays with ncol columns in use) pass
! their leading blocks, which the reference reads (rrtmgsw_lead). A non-zero return stops with
! the reference's STOP message.

module rrtmgsw_capi_nomcica
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function c_rrtmgsw_init(cpdair) bind(C, name="rrtmgsw_init")
      import :: c_int, c_double
      real(c_double), value :: cpdair
    end function
    integer(c_int) function c_rrtmgsw_init_nc(cpdair, path) bind(C, name="rrtmgsw_init_nc")
      import :: c_int, c_double, c_char
      real(c_double), value :: cpdair
      character(kind=c_char) :: path(*)
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
  ! ktable_file (an addition, optional): read the k-distribution and solar-source data from that
  ! netCDF file as the reference's src/rrtmg_sw_read_nc.f90 build reads 'rrtmg_sw.nc'.
  subroutine rrtmg_sw_ini(cpdair, ktable_file)
    real(kind=rb), intent(in) :: cpdair
    character(len=*), intent(in), optional :: ktable_file
    if (present(ktable_file)) then
      call rrtmgsw_check(c_rrtmgsw_init_nc(real(cpdair, c_double), trim(ktable_file) // c_null_char))
    else
      call rrtmgsw_check(c_rrtmgsw_init(real(cpdair, c_double)))
    end if
  end subroutine rrtmg_sw_ini
end module rrtmg_sw_init

module rrtmg_sw_rad
  use iso_c_binding
  use rrtmgsw_capi_nomcica
  use rrtmgsw_lead
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
    integer, parameter :: nbndsw = 14, naerec = 6
    type(buf1), target :: i1(6)
    type(buf2), target :: i2(15), o2(6)
    type(buf3), target :: i3(8)
    type(c_ptr) :: p(35), q(6)
    logical :: uc, ua, ue, u
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
    ! which arrays the call reads: icld outside 0..3 becomes 2, iaer outside 0, 6, 10 becomes 0
    ! (src/rrtmg_sw_rad.nomcica.f90:560, :568); every cloud array when icld >= 1 (:1579)
    u = ncol > 0
    uc = icld /= 0 .and. u
    ua = iaer == 10 .and. u
    ue = iaer == 6 .and. u
    p(1) = lead2(play, ncol, nlay, i2(1), u)
    p(2) = lead2(plev, ncol, nlay + 1, i2(2), u)
    p(3) = lead2(tlay, ncol, nlay, i2(3), u)
    p(4) = lead2(tlev, ncol, nlay + 1, i2(4), .false.)
    p(5) = lead1(tsfc, ncol, i1(1), .false.)
    p(6) = lead2(h2ovmr, ncol, nlay, i2(5), u)
    p(7) = lead2(o3vmr, ncol, nlay, i2(6), u)
    p(8) = lead2(co2vmr, ncol, nlay, i2(7), u)
    p(9) = lead2(ch4vmr, ncol, nlay, i2(8), u)
    p(10) = lead2(n2ovmr, ncol, nlay, i2(9), .false.)
    p(11) = lead2(o2vmr, ncol, nlay, i2(10), u)
    p(12) = lead1(asdir, ncol, i1(2), u)
    p(13) = lead1(asdif, ncol, i1(3), u)
    p(14) = lead1(aldir, ncol, i1(4), u)
    p(15) = lead1(aldif, ncol, i1(5), u)
    p(16) = lead1(coszen, ncol, i1(6), u)
    p(17) = lead2(cldfr, ncol, nlay, i2(11), uc)
    p(18) = lead3(taucld, nbndsw, ncol, nlay, i3(1), uc)
    p(19) = lead3(ssacld, nbndsw, ncol, nlay, i3(2), uc)
    p(20) = lead3(asmcld, nbndsw, ncol, nlay, i3(3), uc)
    p(21) = lead3(fsfcld, nbndsw, ncol, nlay, i3(4), uc)
    p(22) = lead2(cicewp, ncol, nlay, i2(12), uc)
    p(23) = lead2(cliqwp, ncol, nlay, i2(13), uc)
    p(24) = lead2(reice, ncol, nlay, i2(14), uc)
    p(25) = lead2(reliq, ncol, nlay, i2(15), uc)
    p(26) = lead3(tauaer, ncol, nlay, nbndsw, i3(5), ua)
    p(27) = lead3(ssaaer, ncol, nlay, nbndsw, i3(6), ua)
    p(28) = lead3(asmaer, ncol, nlay, nbndsw, i3(7), ua)
    p(29) = lead3(ecaer, ncol, nlay, naerec, i3(8), ue)
    q(1) = outb2(swuflx, ncol, nlay + 1, o2(1))
    q(2) = outb2(swdflx, ncol, nlay + 1, o2(2))
    q(3) = outb2(swhr, ncol, nlay, o2(3))
    q(4) = outb2(swuflxc, ncol, nlay + 1, o2(4))
    q(5) = outb2(swdflxc, ncol, nlay + 1, o2(5))
    q(6) = outb2(swhrc, ncol, nlay, o2(6))
    call rrtmgsw_check(c_rrtmgsw_sw_nomcica(ncol, nlay, ic, ia, p(1), p(2), p(3), p(4), p(5), &
         p(6), p(7), p(8), p(9), p(10), p(11), p(12), p(13), p(14), p(15), p(16), adjes, dyofyr, &
         scon, isolvar, inflgsw, iceflgsw, liqflgsw, p(17), p(18), p(19), p(20), p(21), p(22), &
         p(23), p(24), p(25), p(26), p(27), p(28), p(29), q(1), q(2), q(3), q(4), q(5), q(6), &
         pb, pi, pf))
    icld = ic
    iaer = ia
    call back2(swuflx, o2(1))
    call back2(swdflx, o2(2))
    call back2(swhr, o2(3))
    call back2(swuflxc, o2(4))
    call back2(swdflxc, o2(5))
    call back2(swhrc, o2(6))
  end subroutine rrtmg_sw
end module rrtmg_sw_rad
