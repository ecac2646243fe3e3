! rrtmg_sw_hip.f90 -- Fortran drop-in for RRTMG_SW's GCM-mode McICA interface on the MI355X library.
!
! Keeps the reference's module names, routine names and dummy arguments (order, kinds, intents,
! optional arguments) so that a host GCM recompiles against these modules unchanged:
!   module rrtmg_sw_init       rrtmg_sw_ini(cpdair)            src/rrtmg_sw_init.f90:47
!   module mcica_subcol_gen_sw get_alpha(...), mcica_subcol_sw(...)  src/mcica_subcol_gen_sw.f90:66,181
!   module rrtmg_sw_rad        rrtmg_sw(...)                   src/rrtmg_sw_rad.f90:99
! Each routine forwards to the C-ABI of librrtmgsw_hip.so (include/rrtmgsw.h) through
! iso_c_binding. Dummies are declared contiguous, so non-contiguous actual arguments are copied
! in and out by the compiler; arrays larger than the counts (a GCM's pcols-sized arrays with ncol
! columns in use) pass their leading blocks, which the reference reads (rrtmgsw_lead). A non-zero
! return stops with the reference's STOP message.

module rrtmgsw_capi
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

module mcica_subcol_gen_sw
  use iso_c_binding
  use rrtmgsw_capi
  use rrtmgsw_lead
  implicit none
  integer, parameter, private :: rb = selected_real_kind(12), im = selected_int_kind(8)
  integer, parameter, private :: ngptsw = 112, nbndsw = 14
contains
  subroutine get_alpha(iplon, ncol, nlayers, icld, idcor, decorr_con, dz, lat, juldat, cldfrac, alpha)
    integer(kind=im), intent(in) :: iplon, ncol, nlayers, icld, idcor, juldat
    real(kind=rb), intent(in) :: decorr_con
    real(kind=rb), intent(in), target, contiguous :: dz(:,:), lat(:), cldfrac(:,:)
    real(kind=rb), intent(out), target, contiguous :: alpha(:,:)
    type(buf1), target :: tl
    type(buf2), target :: td, tc, ta
    type(c_ptr) :: pd, pl, pc, pa
    logical :: used
    used = icld == 4 .or. icld == 5
    if (.not. used) return
    pd = lead2(dz, ncol, nlayers, td, used)
    pl = lead1(lat, ncol, tl, used .and. idcor == 1)
    pc = lead2(cldfrac, ncol, nlayers, tc, used)
    pa = outb2(alpha, ncol, nlayers, ta)
    call rrtmgsw_check(c_rrtmgsw_get_alpha(iplon, ncol, nlayers, icld, idcor, decorr_con, &
         pd, pl, juldat, pc, pa))
    call back2(alpha, ta)
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
    type(buf2), target :: i2(7), o2(2)
    type(buf3), target :: i3(4), o3(7)
    type(c_ptr) :: p(11), q(9)
    logical :: u
    ir = irng
    ! the reference returns at once for icld = 0 and stops for icld outside 0..5 (the C-ABI's
    ! checks): only a call that draws reads the arrays
    u = icld >= 1 .and. icld <= 5 .and. ncol > 0
    p(1) = lead2(play, ncol, nlay, i2(1), u)
    p(2) = lead2(cldfrac, ncol, nlay, i2(2), u)
    p(3) = lead2(ciwp, ncol, nlay, i2(3), u)
    p(4) = lead2(clwp, ncol, nlay, i2(4), u)
    p(5) = lead2(rei, ncol, nlay, i2(5), u)
    p(6) = lead2(rel, ncol, nlay, i2(6), u)
    p(7) = lead3(tauc, nbndsw, ncol, nlay, i3(1), u)
    p(8) = lead3(ssac, nbndsw, ncol, nlay, i3(2), u)
    p(9) = lead3(asmc, nbndsw, ncol, nlay, i3(3), u)
    p(10) = lead3(fsfc, nbndsw, ncol, nlay, i3(4), u)
    p(11) = lead2(alpha, ncol, nlay, i2(7), u .and. icld >= 4)
    if (u) then
      q(1) = outb3(cldfmcl, ngptsw, ncol, nlay, o3(1))
      q(2) = outb3(ciwpmcl, ngptsw, ncol, nlay, o3(2))
      q(3) = outb3(clwpmcl, ngptsw, ncol, nlay, o3(3))
      q(4) = outb2(reicmcl, ncol, nlay, o2(1))
      q(5) = outb2(relqmcl, ncol, nlay, o2(2))
      q(6) = outb3(taucmcl, ngptsw, ncol, nlay, o3(4))
      q(7) = outb3(ssacmcl, ngptsw, ncol, nlay, o3(5))
      q(8) = outb3(asmcmcl, ngptsw, ncol, nlay, o3(6))
      q(9) = outb3(fsfcmcl, ngptsw, ncol, nlay, o3(7))
    else
      q = [c_loc(cldfmcl), c_loc(ciwpmcl), c_loc(clwpmcl), c_loc(reicmcl), c_loc(relqmcl), &
           c_loc(taucmcl), c_loc(ssacmcl), c_loc(asmcmcl), c_loc(fsfcmcl)]
    end if
    call rrtmgsw_check(c_rrtmgsw_mcica_subcol(iplon, ncol, nlay, icld, permuteseed, ir, &
         p(1), p(2), p(3), p(4), p(5), p(6), p(7), p(8), p(9), p(10), p(11), &
         q(1), q(2), q(3), q(4), q(5), q(6), q(7), q(8), q(9)))
    irng = ir
    if (u) then
      call back3(cldfmcl, o3(1))
      call back3(ciwpmcl, o3(2))
      call back3(clwpmcl, o3(3))
      call back2(reicmcl, o2(1))
      call back2(relqmcl, o2(2))
      call back3(taucmcl, o3(4))
      call back3(ssacmcl, o3(5))
      call back3(asmcmcl, o3(6))
      call back3(fsfcmcl, o3(7))
    end if
  end subroutine mcica_subcol_sw
end module mcica_subcol_gen_sw

module rrtmg_sw_rad
  use iso_c_binding
  use rrtmgsw_capi
  use rrtmgsw_lead
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
    integer, parameter :: ngptsw = 112, nbndsw = 14, naerec = 6
    type(buf1), target :: i1(6)
    type(buf2), target :: i2(12), o2(6)
    type(buf3), target :: i3(11)
    type(c_ptr) :: p(30), q(6)
    logical :: uc, ua, ue
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
    ! (src/rrtmg_sw_rad.f90:589, :597)
    uc = icld /= 0 .and. ncol > 0
    ua = iaer == 10 .and. ncol > 0
    ue = iaer == 6 .and. ncol > 0
    p(1) = lead2(play, ncol, nlay, i2(1), ncol > 0)
    p(2) = lead2(plev, ncol, nlay + 1, i2(2), ncol > 0)
    p(3) = lead2(tlay, ncol, nlay, i2(3), ncol > 0)
    p(4) = lead2(tlev, ncol, nlay + 1, i2(4), .false.)
    p(5) = lead1(tsfc, ncol, i1(1), .false.)
    p(6) = lead2(h2ovmr, ncol, nlay, i2(5), ncol > 0)
    p(7) = lead2(o3vmr, ncol, nlay, i2(6), ncol > 0)
    p(8) = lead2(co2vmr, ncol, nlay, i2(7), ncol > 0)
    p(9) = lead2(ch4vmr, ncol, nlay, i2(8), ncol > 0)
    p(10) = lead2(n2ovmr, ncol, nlay, i2(9), .false.)
    p(11) = lead2(o2vmr, ncol, nlay, i2(10), ncol > 0)
    p(12) = lead1(asdir, ncol, i1(2), ncol > 0)
    p(13) = lead1(asdif, ncol, i1(3), ncol > 0)
    p(14) = lead1(aldir, ncol, i1(4), ncol > 0)
    p(15) = lead1(aldif, ncol, i1(5), ncol > 0)
    p(16) = lead1(coszen, ncol, i1(6), ncol > 0)
    p(17) = lead3(cldfmcl, ngptsw, ncol, nlay, i3(1), uc)
    p(18) = lead3(taucmcl, ngptsw, ncol, nlay, i3(2), uc)
    p(19) = lead3(ssacmcl, ngptsw, ncol, nlay, i3(3), uc)
    p(20) = lead3(asmcmcl, ngptsw, ncol, nlay, i3(4), uc)
    p(21) = lead3(fsfcmcl, ngptsw, ncol, nlay, i3(5), uc)
    p(22) = lead3(ciwpmcl, ngptsw, ncol, nlay, i3(6), uc)
    p(23) = lead3(clwpmcl, ngptsw, ncol, nlay, i3(7), uc)
    p(24) = lead2(reicmcl, ncol, nlay, i2(11), uc)
    p(25) = lead2(relqmcl, ncol, nlay, i2(12), uc)
    p(26) = lead3(tauaer, ncol, nlay, nbndsw, i3(8), ua)
    p(27) = lead3(ssaaer, ncol, nlay, nbndsw, i3(9), ua)
    p(28) = lead3(asmaer, ncol, nlay, nbndsw, i3(10), ua)
    p(29) = lead3(ecaer, ncol, nlay, naerec, i3(11), ue)
    q(1) = outb2(swuflx, ncol, nlay + 1, o2(1))
    q(2) = outb2(swdflx, ncol, nlay + 1, o2(2))
    q(3) = outb2(swhr, ncol, nlay, o2(3))
    q(4) = outb2(swuflxc, ncol, nlay + 1, o2(4))
    q(5) = outb2(swdflxc, ncol, nlay + 1, o2(5))
    q(6) = outb2(swhrc, ncol, nlay, o2(6))
    call rrtmgsw_check(c_rrtmgsw_sw(ncol, nlay, ic, ia, p(1), p(2), p(3), p(4), p(5), p(6), p(7), &
         p(8), p(9), p(10), p(11), p(12), p(13), p(14), p(15), p(16), adjes, dyofyr, scon, isolvar, &
         inflgsw, iceflgsw, liqflgsw, p(17), p(18), p(19), p(20), p(21), p(22), p(23), p(24), p(25), &
         p(26), p(27), p(28), p(29), q(1), q(2), q(3), q(4), q(5), q(6), pb, pi, pf))
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
