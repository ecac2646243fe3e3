! example_gcm.f90 -- a host program that calls the drop-in modules exactly as a GCM calls the
! reference: rrtmg_sw_ini once, then mcica_subcol_sw and rrtmg_sw per chunk of columns
! (README.md of the reference, "GCM" usage). Inputs are read from a raw stream file and the six
! outputs written to another, so that tests can drive it:
!   example_gcm <input.bin> <output.bin> [pad [ktable.nc]]
! With pad > 0 every array is re-homed in one with pcols = ncol + pad columns (the extra columns
! hold junk), as a GCM's pcols-sized arrays are, and the outputs' first ncol columns are written.
! With ktable.nc the k-tables come from that netCDF file (rrtmg_sw_ini's ktable_file).
! input: 11 int32 (ncol nlay icld iaer permuteseed irng dyofyr isolvar inflgsw iceflgsw liqflgsw),
!        2 real64 (adjes scon), then play plev tlay tlev tsfc h2o o3 co2 ch4 n2o o2 asdir asdif
!        aldir aldif coszen cldfrac ciwp clwp rei rel tauc ssac asmc fsfc alpha tauaer ssaaer
!        asmaer ecaer in Fortran order.
program example_gcm
  use rrtmg_sw_init, only : rrtmg_sw_ini
  use mcica_subcol_gen_sw, only : mcica_subcol_sw
  use rrtmg_sw_rad, only : rrtmg_sw
  use example_pad
  implicit none
  integer, parameter :: rb = selected_real_kind(12), ngptsw = 112, nbndsw = 14, naerec = 6
  integer :: hdr(11), ncol, nlay, icld, iaer, permuteseed, irng, dyofyr, isolvar
  integer :: inflgsw, iceflgsw, liqflgsw, u, pad, pcols
  real(rb) :: sc(2)
  character(len=512) :: fin, fout, arg
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
  pad = 0
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg)
    read(arg, *) pad
  end if
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
  pcols = ncol + pad
  if (pad > 0) then
    call pad2(play, pcols); call pad2(plev, pcols); call pad2(tlay, pcols)
    call pad2(tlev, pcols); call pad1(tsfc, pcols); call pad2(h2o, pcols); call pad2(o3, pcols)
    call pad2(co2, pcols); call pad2(ch4, pcols); call pad2(n2o, pcols); call pad2(o2, pcols)
    call pad1(asdir, pcols); call pad1(asdif, pcols); call pad1(aldir, pcols)
    call pad1(aldif, pcols); call pad1(coszen, pcols); call pad2(cldfrac, pcols)
    call pad2(ciwp, pcols); call pad2(clwp, pcols); call pad2(rei, pcols); call pad2(rel, pcols)
    call pad3b(tauc, pcols); call pad3b(ssac, pcols); call pad3b(asmc, pcols)
    call pad3b(fsfc, pcols); call pad2(alpha, pcols); call pad3a(tauaer, pcols)
    call pad3a(ssaaer, pcols); call pad3a(asmaer, pcols); call pad3a(ecaer, pcols)
  end if
  allocate(cldfmcl(ngptsw,pcols,nlay), ciwpmcl(ngptsw,pcols,nlay), clwpmcl(ngptsw,pcols,nlay))
  allocate(taucmcl(ngptsw,pcols,nlay), ssacmcl(ngptsw,pcols,nlay), asmcmcl(ngptsw,pcols,nlay))
  allocate(fsfcmcl(ngptsw,pcols,nlay), reicmcl(pcols,nlay), relqmcl(pcols,nlay))
  allocate(swuflx(pcols,nlay+1), swdflx(pcols,nlay+1), swhr(pcols,nlay))
  allocate(swuflxc(pcols,nlay+1), swdflxc(pcols,nlay+1), swhrc(pcols,nlay))
  cldfmcl = 0; ciwpmcl = 0; clwpmcl = 0; taucmcl = 0; ssacmcl = 1; asmcmcl = 0; fsfcmcl = 0
  reicmcl = 0; relqmcl = 0

  if (command_argument_count() >= 4) then   ! a k-table file, read as read_nc reads rrtmg_sw.nc
    call get_command_argument(4, arg)
    call rrtmg_sw_ini(1004.64_rb, ktable_file=trim(arg))
  else
    call rrtmg_sw_ini(1004.64_rb)
  end if
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
  write(u) icld, iaer, swuflx(1:ncol,:), swdflx(1:ncol,:), swhr(1:ncol,:), swuflxc(1:ncol,:), &
           swdflxc(1:ncol,:), swhrc(1:ncol,:)
  close(u)
end program example_gcm
