! example_gcm.f90 -- a host program that calls the drop-in modules exactly as a GCM calls the
! reference: rrtmg_sw_ini once, then mcica_subcol_sw and rrtmg_sw per chunk of columns
! (README.md of the reference, "GCM" usage). Inputs are read from a raw stream file and the six
! outputs written to another, so that tests can drive it:
!   example_gcm <input.bin> <output.bin>
! input: 11 int32 (ncol nlay icld iaer permuteseed irng dyofyr isolvar inflgsw iceflgsw liqflgsw),
!        2 real64 (adjes scon), then play plev tlay tlev tsfc h2o o3 co2 ch4 n2o o2 asdir asdif
!        aldir aldif coszen cldfrac ciwp clwp rei rel tauc ssac asmc fsfc alpha tauaer ssaaer
!        asmaer ecaer in Fortran order.
program example_gcm
  use rrtmg_sw_init, only : rrtmg_sw_ini
  use mcica_subcol_gen_sw, only : mcica_subcol_sw
  use rrtmg_sw_rad, only : rrtmg_sw
  implicit none
  integer, parameter :: rb = selected_real_kind(12), ngptsw = 112, nbndsw = 14, naerec = 6
  integer :: hdr(11), ncol, nlay, icld, iaer, permuteseed, irng, dyofyr, isolvar
  integer :: inflgsw, iceflgsw, liqflgsw, u
  real(rb) :: sc(2)
  character(len=512) :: fin, fout
  real(rb), allocatable :: play(:,:), plev(:,:), tlay(:,:), tlev(:,:), tsfc(:), h2o(:,:), o3(:,:)
  real(rb), allocatable :: co2(:,:), ch4(:,:), n2o(:,:), o2(:,:), asdir(:), asdif(:), aldir(:)
  real(rb), allocatable :: aldif(:), coszen(:), cldfrac(:,:), ciwp(:,:), clwp(:,:), rei(:,:)
  real(rb), allocatable :: rel(:,:), tauc(:,:,:), ssac(:,:,:), asmc(:,:,:), fsfc(:,:,:), alpha(:,:)
  real(rb), allocatable :: tauaer(:,:,:), ssaaer(:,:,:), asmaer(:,:,:), ecaer(:,:,:)
  real(rb), allocatable :: cldfmcl(:,:,:), ciwpmcl(:,:,:), clwpmcl(:,:,:), taucmcl(:,:,:)
  real(rb), allocatable :: ssacmcl(:,:,:), asmcmcl(:,:,:), fsfcmcl(:,:,:), reicmcl(:,:), relqmcl(:,:)
  real(rb), allocatable :: swuflx(:,:), swdflx(:,:), swhr(:,:), swuflxc(:,:), swdflxc(:,:), swhrc(:,:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr, sc
  ncol = hdr(1); nlay = hdr(2); icld = hdr(3); iaer = hdr(4); permuteseed = hdr(5)
  irng = hdr(6); dyofyr = hdr(7); isolvar = hdr(8); inflgsw = hdr(9); iceflgsw = hdr(10)
  liqflgsw = hdr(11)
  allocate(play(ncol,nlay), plev(ncol,nlay+1), tlay(ncol,nlay), tlev(ncol,nlay+1), tsfc(ncol))
  allocate(h2o(ncol,nlay), o3(ncol,nlay), co2(ncol,nlay), ch4(ncol,nlay), n2o(ncol,nlay))
  allocate(o2(ncol,nlay), asdir(ncol), asdif(ncol), aldir(ncol), aldif(ncol), coszen(ncol))
  allocate(cldfrac(ncol,nlay), ciwp(ncol,nlay), clwp(ncol,nlay), rei(ncol,nlay), rel(ncol,nlay))
  allocate(tauc(nbndsw,ncol,nlay), ssac(nbndsw,ncol,nlay), asmc(nbndsw,ncol,nlay))
  allocate(fsfc(nbndsw,ncol,nlay), alpha(ncol,nlay))
  allocate(tauaer(ncol,nlay,nbndsw), ssaaer(ncol,nlay,nbndsw), asmaer(ncol,nlay,nbndsw))
  allocate(ecaer(ncol,nlay,naerec))
  read(u) play, plev, tlay, tlev, tsfc, h2o, o3, co2, ch4, n2o, o2, asdir, asdif, aldir, aldif, &
          coszen, cldfrac, ciwp, clwp, rei, rel, tauc, ssac, asmc, fsfc, alpha, tauaer, ssaaer, &
          asmaer, ecaer
  close(u)
  allocate(cldfmcl(ngptsw,ncol,nlay), ciwpmcl(ngptsw,ncol,nlay), clwpmcl(ngptsw,ncol,nlay))
  allocate(taucmcl(ngptsw,ncol,nlay), ssacmcl(ngptsw,ncol,nlay), asmcmcl(ngptsw,ncol,nlay))
  allocate(fsfcmcl(ngptsw,ncol,nlay), reicmcl(ncol,nlay), relqmcl(ncol,nlay))
  allocate(swuflx(ncol,nlay+1), swdflx(ncol,nlay+1), swhr(ncol,nlay))
  allocate(swuflxc(ncol,nlay+1), swdflxc(ncol,nlay+1), swhrc(ncol,nlay))
  cldfmcl = 0; ciwpmcl = 0; clwpmcl = 0; taucmcl = 0; ssacmcl = 1; asmcmcl = 0; fsfcmcl = 0
  reicmcl = 0; relqmcl = 0

  call rrtmg_sw_ini(1004.64_rb)
  if (icld > 0) then
    call mcica_subcol_sw(1, ncol, nlay, icld, permuteseed, irng, play, cldfrac, ciwp, clwp, &
                         rei, rel, tauc, ssac, asmc, fsfc, alpha, cldfmcl, ciwpmcl, clwpmcl, &
                         reicmcl, relqmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl)
  end if
  call rrtmg_sw(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, h2o, o3, co2, ch4, n2o, &
                o2, asdir, asdif, aldir, aldif, coszen, sc(1), dyofyr, sc(2), isolvar, &
                inflgsw, iceflgsw, liqflgsw, cldfmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl, &
                ciwpmcl, clwpmcl, reicmcl, relqmcl, tauaer, ssaaer, asmaer, ecaer, &
                swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc)
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) icld, iaer, swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc
  close(u)
end program example_gcm
