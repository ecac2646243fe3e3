! ref_capi.F90 -- C-callable entry points into the compiled reference Fortran (oracle/_ref).
!
! TEST INFRASTRUCTURE ONLY: this file is compiled together with the reference's own sources
! (read in place under the reference tree by oracle/build_ref.sh) into oracle/_ref/librrtmg_ref.so,
! which the tests use to pin the C oracle (oracle/rrtmg_oracle.c) and to produce golden fixtures.
! Nothing here is part of the product. Each wrapper forwards to the reference's public module
! procedure named in its comment; arrays are passed with explicit shapes.
module ref_capi
  use iso_c_binding
  use parkind, only : im => kind_im, rb => kind_rb
  use parrrsw, only : ngptsw, nbndsw, naerec, mxmol
  use rrsw_con, only : oneminus
  use rrsw_tbl, only : exp_tbl
  use rrsw_wvn, only : rwgt
  use rrtmg_sw_init, only : rrtmg_sw_ini
  use rrtmg_sw_rad, only : rrtmg_sw
  use mcica_subcol_gen_sw, only : mcica_subcol_sw, get_alpha
  use rrtmg_sw_setcoef, only : setcoef_sw
  use rrtmg_sw_taumol, only : taumol_sw
  use rrtmg_sw_reftra, only : reftra_sw
  use rrtmg_sw_vrtqdr, only : vrtqdr_sw
  use rrtmg_sw_cldprmc, only : cldprmc_sw
  implicit none
contains

  ! rrtmg_sw_ini(cpdair), src/rrtmg_sw_init.f90:47
  subroutine ref_init(cpdair) bind(C, name="ref_init")
    real(c_double), value :: cpdair
    call rrtmg_sw_ini(cpdair)
  end subroutine ref_init

  ! rrtmg_sw(...), src/rrtmg_sw_rad.f90:99; has_* select the optional arguments.
  subroutine ref_rrtmg_sw(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, &
       co2vmr, ch4vmr, n2ovmr, o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, &
       isolvar, inflgsw, iceflgsw, liqflgsw, cldfmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl, &
       ciwpmcl, clwpmcl, reicmcl, relqmcl, tauaer, ssaaer, asmaer, ecaer, swuflx, swdflx, swhr, &
       swuflxc, swdflxc, swhrc, has_bnd, bndsolvar, has_ind, indsolvar, has_frac, solcycfrac) &
       bind(C, name="ref_rrtmg_sw")
    integer(c_int), value :: ncol, nlay, dyofyr, isolvar, inflgsw, iceflgsw, liqflgsw
    integer(c_int), value :: has_bnd, has_ind, has_frac
    integer(c_int) :: icld, iaer
    real(c_double), value :: adjes, scon
    real(c_double) :: play(ncol,nlay), plev(ncol,nlay+1), tlay(ncol,nlay), tlev(ncol,nlay+1)
    real(c_double) :: tsfc(ncol), h2ovmr(ncol,nlay), o3vmr(ncol,nlay), co2vmr(ncol,nlay)
    real(c_double) :: ch4vmr(ncol,nlay), n2ovmr(ncol,nlay), o2vmr(ncol,nlay)
    real(c_double) :: asdir(ncol), asdif(ncol), aldir(ncol), aldif(ncol), coszen(ncol)
    real(c_double) :: cldfmcl(ngptsw,ncol,nlay), taucmcl(ngptsw,ncol,nlay)
    real(c_double) :: ssacmcl(ngptsw,ncol,nlay), asmcmcl(ngptsw,ncol,nlay)
    real(c_double) :: fsfcmcl(ngptsw,ncol,nlay), ciwpmcl(ngptsw,ncol,nlay)
    real(c_double) :: clwpmcl(ngptsw,ncol,nlay), reicmcl(ncol,nlay), relqmcl(ncol,nlay)
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
  iceflgsw, liqflgsw, cldfmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl, ciwpmcl, clwpmcl, reicmcl, \
  relqmcl, tauaer, ssaaer, asmaer, ecaer, swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc
    if (has_bnd /= 0 .and. has_ind /= 0 .and. has_frac /= 0) then
      call rrtmg_sw(ARGS, bndsolvar=bndsolvar, indsolvar=indsolvar, solcycfrac=solcycfrac)
    else if (has_bnd /= 0 .and. has_ind /= 0) then
      call rrtmg_sw(ARGS, bndsolvar=bndsolvar, indsolvar=indsolvar)
    else if (has_bnd /= 0 .and. has_frac /= 0) then
      call rrtmg_sw(ARGS, bndsolvar=bndsolvar, solcycfrac=solcycfrac)
    else if (has_ind /= 0 .and. has_frac /= 0) then
      call rrtmg_sw(ARGS, indsolvar=indsolvar, solcycfrac=solcycfrac)
    else if (has_bnd /= 0) then
      call rrtmg_sw(ARGS, bndsolvar=bndsolvar)
    else if (has_ind /= 0) then
      call rrtmg_sw(ARGS, indsolvar=indsolvar)
    else if (has_frac /= 0) then
      call rrtmg_sw(ARGS, solcycfrac=solcycfrac)
    else
      call rrtmg_sw(ARGS)
    end if
#undef ARGS
    icld = ic
    iaer = ia
  end subroutine ref_rrtmg_sw

  ! mcica_subcol_sw(...), src/mcica_subcol_gen_sw.f90:181
  subroutine ref_mcica_subcol_sw(iplon, ncol, nlay, icld, permuteseed, irng, play, cldfrac, ciwp, &
       clwp, rei, rel, tauc, ssac, asmc, fsfc, alpha, cldfmcl, ciwpmcl, clwpmcl, reicmcl, &
       relqmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl) bind(C, name="ref_mcica_subcol_sw")
    integer(c_int), value :: iplon, ncol, nlay, icld, permuteseed
    integer(c_int) :: irng
    real(c_double) :: play(ncol,nlay), cldfrac(ncol,nlay), ciwp(ncol,nlay), clwp(ncol,nlay)
    real(c_double) :: rei(ncol,nlay), rel(ncol,nlay), alpha(ncol,nlay)
    real(c_double) :: tauc(nbndsw,ncol,nlay), ssac(nbndsw,ncol,nlay)
    real(c_double) :: asmc(nbndsw,ncol,nlay), fsfc(nbndsw,ncol,nlay)
    real(c_double) :: cldfmcl(ngptsw,ncol,nlay), ciwpmcl(ngptsw,ncol,nlay)
    real(c_double) :: clwpmcl(ngptsw,ncol,nlay), taucmcl(ngptsw,ncol,nlay)
    real(c_double) :: ssacmcl(ngptsw,ncol,nlay), asmcmcl(ngptsw,ncol,nlay)
    real(c_double) :: fsfcmcl(ngptsw,ncol,nlay), reicmcl(ncol,nlay), relqmcl(ncol,nlay)
    integer(im) :: ir
    ir = irng
    call mcica_subcol_sw(iplon, ncol, nlay, icld, permuteseed, ir, play, cldfrac, ciwp, clwp, &
         rei, rel, tauc, ssac, asmc, fsfc, alpha, cldfmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl, &
         taucmcl, ssacmcl, asmcmcl, fsfcmcl)
    irng = ir
  end subroutine ref_mcica_subcol_sw

  ! get_alpha(...), src/mcica_subcol_gen_sw.f90:66
  subroutine ref_get_alpha(iplon, ncol, nlayers, icld, idcor, decorr_con, dz, lat, juldat, &
       cldfrac, alpha) bind(C, name="ref_get_alpha")
    integer(c_int), value :: iplon, ncol, nlayers, icld, idcor, juldat
    real(c_double), value :: decorr_con
    real(c_double) :: dz(ncol,nlayers), lat(ncol), cldfrac(ncol,nlayers), alpha(ncol,nlayers)
    call get_alpha(iplon, ncol, nlayers, icld, idcor, decorr_con, dz, lat, juldat, cldfrac, alpha)
  end subroutine ref_get_alpha

  ! setcoef_sw(...), src/rrtmg_sw_setcoef.f90:49. iout/dout use the C oracle's packing.
  subroutine ref_setcoef(nlay, pavel, tavel, coldry, wkl, iout, dout) bind(C, name="ref_setcoef")
    integer(c_int), value :: nlay
    real(c_double) :: pavel(nlay), tavel(nlay), coldry(nlay), wkl(mxmol,nlay)
    integer(c_int) :: iout(3 + 5*nlay)
    real(c_double) :: dout(nlay, 16)
    real(rb) :: pz(0:nlay), tz(0:nlay), tbound
    integer(im) :: laytrop, layswtch, laylow
    integer(im) :: jp(nlay), jt(nlay), jt1(nlay), indself(nlay), indfor(nlay)
    pz = 0._rb
    tz = 250._rb
    tbound = 250._rb
    call setcoef_sw(nlay, pavel, tavel, pz, tz, tbound, coldry, wkl, laytrop, layswtch, laylow, &
         jp, jt, jt1, dout(:,8), dout(:,5), dout(:,2), dout(:,1), dout(:,7), dout(:,4), &
         dout(:,6), dout(:,3), dout(:,9), dout(:,10), dout(:,11), dout(:,12), dout(:,13), &
         dout(:,14), indself, dout(:,15), dout(:,16), indfor)
    iout(1) = laytrop
    iout(2) = layswtch
    iout(3) = laylow
    iout(4:3+nlay) = jp
    iout(4+nlay:3+2*nlay) = jt
    iout(4+2*nlay:3+3*nlay) = jt1
    iout(4+3*nlay:3+4*nlay) = indself
    iout(4+4*nlay:3+5*nlay) = indfor
  end subroutine ref_setcoef

  ! taumol_sw(...), src/rrtmg_sw_taumol.f90:50, on the packing of ref_setcoef.
  subroutine ref_taumol(nlay, iout, dout, isolvar, svar_f, svar_s, svar_i, svar_bnd, ssi, &
       sfluxzen, taug, taur) bind(C, name="ref_taumol")
    integer(c_int), value :: nlay, isolvar
    integer(c_int) :: iout(3 + 5*nlay)
    real(c_double) :: dout(nlay, 16)
    real(c_double), value :: svar_f, svar_s, svar_i
    real(c_double) :: svar_bnd(29, 3), ssi(ngptsw), sfluxzen(ngptsw)
    real(c_double) :: taug(nlay, ngptsw), taur(nlay, ngptsw)
    oneminus = 1.0_rb - 1.e-06_rb
    call taumol_sw(nlay, dout(:,1), dout(:,2), dout(:,5), dout(:,6), dout(:,3), dout(:,7), &
         iout(1), iout(4:3+nlay), iout(4+nlay:3+2*nlay), iout(4+2*nlay:3+3*nlay), &
         dout(:,9), dout(:,10), dout(:,11), dout(:,12), dout(:,13), dout(:,14), &
         iout(4+3*nlay:3+4*nlay), dout(:,15), dout(:,16), iout(4+4*nlay:3+5*nlay), &
         isolvar, svar_f, svar_s, svar_i, svar_bnd(:,1), svar_bnd(:,2), svar_bnd(:,3), &
         ssi, sfluxzen, taug, taur)
  end subroutine ref_taumol

  ! reftra_sw(...), src/rrtmg_sw_reftra.f90:48
  subroutine ref_reftra(n, lrtchk, pgg, prmuz, ptau, pw, pref, prefd, ptra, ptrad) &
       bind(C, name="ref_reftra")
    integer(c_int), value :: n
    integer(c_int) :: lrtchk(n)
    real(c_double), value :: prmuz
    real(c_double) :: pgg(n), ptau(n), pw(n), pref(n+1), prefd(n+1), ptra(n+1), ptrad(n+1)
    logical :: l(n)
    l = lrtchk /= 0
    call reftra_sw(n, l, pgg, prmuz, ptau, pw, pref, prefd, ptra, ptrad)
  end subroutine ref_reftra

  ! vrtqdr_sw(...), src/rrtmg_sw_vrtqdr.f90:47
  subroutine ref_vrtqdr(klev, pref, prefd, ptra, ptrad, pdbt, prdnd, prup, prupd, ptdbt, pfd, pfu) &
       bind(C, name="ref_vrtqdr")
    integer(c_int), value :: klev
    real(c_double) :: pref(klev+1), prefd(klev+1), ptra(klev+1), ptrad(klev+1), pdbt(klev+1)
    real(c_double) :: prdnd(klev+1), prup(klev+1), prupd(klev+1), ptdbt(klev+1)
    real(c_double) :: pfd(klev+1, 1), pfu(klev+1, 1)
    call vrtqdr_sw(klev, 1, pref, prefd, ptra, ptrad, pdbt, prdnd, prup, prupd, ptdbt, pfd, pfu)
  end subroutine ref_vrtqdr

  ! cldprmc_sw(...), src/rrtmg_sw_cldprmc.f90:53
  subroutine ref_cldprmc(nlay, inflag, iceflag, liqflag, cldfmc, ciwpmc, clwpmc, reicmc, relqmc, &
       taormc, taucmc, ssacmc, asmcmc, fsfcmc) bind(C, name="ref_cldprmc")
    integer(c_int), value :: nlay, inflag, iceflag, liqflag
    real(c_double) :: cldfmc(ngptsw,nlay), ciwpmc(ngptsw,nlay), clwpmc(ngptsw,nlay)
    real(c_double) :: reicmc(nlay), relqmc(nlay), taormc(ngptsw,nlay), taucmc(ngptsw,nlay)
    real(c_double) :: ssacmc(ngptsw,nlay), asmcmc(ngptsw,nlay), fsfcmc(ngptsw,nlay)
    call cldprmc_sw(nlay, inflag, iceflag, liqflag, cldfmc, ciwpmc, clwpmc, reicmc, relqmc, &
         taormc, taucmc, ssacmc, asmcmc, fsfcmc)
  end subroutine ref_cldprmc

  ! exp_tbl (modules/rrsw_tbl.f90) and rwgt (modules/rrsw_wvn.f90) after rrtmg_sw_ini.
  subroutine ref_exp_tbl(out) bind(C, name="ref_exp_tbl")
    real(c_double) :: out(0:10000)
    out = exp_tbl
  end subroutine ref_exp_tbl
  subroutine ref_rwgt(out) bind(C, name="ref_rwgt")
    real(c_double) :: out(224)
    out = rwgt
  end subroutine ref_rwgt

  ! Reduced 112-point tables of the rrsw_kgNN modules after rrtmg_sw_ini. kind:
  ! 1 absa, 2 absb, 3 selfref, 4 forref, 5 sfluxref, 6 irradnce, 7 facbrght, 8 snsptdrk,
  ! 9 rayl, 10 rayla, 11 raylb, 12 abso3a, 13 abso3b, 14 absch4, 15 absh2o, 16 absco2.
  ! Returns the element count (0 when the band has no such array).
  function ref_table(band, kind, out) result(n) bind(C, name="ref_table")
    integer(c_int), value :: band, kind
    real(c_double) :: out(*)
    integer(c_int) :: n
    n = 0
    select case (band)
#define PUT(arr) n = size(arr); out(1:n) = reshape(arr, (/n/))
#define PUT0(sc) n = 1; out(1) = sc
    case (16)
      block
        use rrsw_kg16
        select case (kind)
        case (1); PUT(absa)
        case (2); PUT(absb)
        case (3); PUT(selfref)
        case (4); PUT(forref)
        case (5); PUT(sfluxref)
        case (6); PUT(irradnce)
        case (7); PUT(facbrght)
        case (8); PUT(snsptdrk)
        case (9); PUT0(rayl)
        end select
      end block
    case (17)
      block
        use rrsw_kg17
        select case (kind)
        case (1); PUT(absa)
        case (2); PUT(absb)
        case (3); PUT(selfref)
        case (4); PUT(forref)
        case (5); PUT(sfluxref)
        case (6); PUT(irradnce)
        case (7); PUT(facbrght)
        case (8); PUT(snsptdrk)
        case (9); PUT0(rayl)
        end select
      end block
    case (18)
      block
        use rrsw_kg18
        select case (kind)
        case (1); PUT(absa)
        case (2); PUT(absb)
        case (3); PUT(selfref)
        case (4); PUT(forref)
        case (5); PUT(sfluxref)
        case (6); PUT(irradnce)
        case (7); PUT(facbrght)
        case (8); PUT(snsptdrk)
        case (9); PUT0(rayl)
        end select
      end block
    case (19)
      block
        use rrsw_kg19
        select case (kind)
        case (1); PUT(absa)
        case (2); PUT(absb)
        case (3); PUT(selfref)
        case (4); PUT(forref)
        case (5); PUT(sfluxref)
        case (6); PUT(irradnce)
        case (7); PUT(facbrght)
        case (8); PUT(snsptdrk)
        case (9); PUT0(rayl)
        end select
      end block
    case (20)
      block
        use rrsw_kg20
        select case (kind)
        case (1); PUT(absa)
        case (2); PUT(absb)
        case (3); PUT(selfref)
        case (4); PUT(forref)
        case (5); PUT(sfluxref)
        case (6); PUT(irradnce)
        case (7); PUT(facbrght)
        case (8); PUT(snsptdrk)
        case (9); PUT0(rayl)
        case (14); PUT(absch4)
        end select
      end block
    case (21)
      block
        use rrsw_kg21
        select case (kind)
        case (1); PUT(absa)
        case (2); PUT(absb)
        case (3); PUT(selfref)
        case (4); PUT(forref)
        case (5); PUT(sfluxref)
        case (6); PUT(irradnce)
        case (7); PUT(facbrght)
        case (8); PUT(snsptdrk)
        case (9); PUT0(rayl)
        end select
      end block
    case (22)
      block
        use rrsw_kg22
        select case (kind)
        case (1); PUT(absa)
        case (2); PUT(absb)
        case (3); PUT(selfref)
        case (4); PUT(forref)
        case (5); PUT(sfluxref)
        case (6); PUT(irradnce)
        case (7); PUT(facbrght)
        case (8); PUT(snsptdrk)
        case (9); PUT0(rayl)
        end select
      end block
    case (23)
      block
        use rrsw_kg23
        select case (kind)
        case (1); PUT(absa)
        case (3); PUT(selfref)
        case (4); PUT(forref)
        case (5); PUT(sfluxref)
        case (6); PUT(irradnce)
        case (7); PUT(facbrght)
        case (8); PUT(snsptdrk)
        case (9); PUT(rayl)
        end select
      end block
    case (24)
      block
        use rrsw_kg24
        select case (kind)
        case (1); PUT(absa)
        case (2); PUT(absb)
        case (3); PUT(selfref)
        case (4); PUT(forref)
        case (5); PUT(sfluxref)
        case (6); PUT(irradnce)
        case (7); PUT(facbrght)
        case (8); PUT(snsptdrk)
        case (10); PUT(rayla)
        case (11); PUT(raylb)
        case (12); PUT(abso3a)
        case (13); PUT(abso3b)
        end select
      end block
    case (25)
      block
        use rrsw_kg25
        select case (kind)
        case (1); PUT(absa)
        case (5); PUT(sfluxref)
        case (6); PUT(irradnce)
        case (7); PUT(facbrght)
        case (8); PUT(snsptdrk)
        case (9); PUT(rayl)
        case (12); PUT(abso3a)
        case (13); PUT(abso3b)
        end select
      end block
    case (26)
      block
        use rrsw_kg26
        select case (kind)
        case (5); PUT(sfluxref)
        case (6); PUT(irradnce)
        case (7); PUT(facbrght)
        case (8); PUT(snsptdrk)
        case (9); PUT(rayl)
        end select
      end block
    case (27)
      block
        use rrsw_kg27
        select case (kind)
        case (1); PUT(absa)
        case (2); PUT(absb)
        case (5); PUT(sfluxref)
        case (6); PUT(irradnce)
        case (7); PUT(facbrght)
        case (8); PUT(snsptdrk)
        case (9); PUT(rayl)
        end select
      end block
    case (28)
      block
        use rrsw_kg28
        select case (kind)
        case (1); PUT(absa)
        case (2); PUT(absb)
        case (5); PUT(sfluxref)
        case (6); PUT(irradnce)
        case (7); PUT(facbrght)
        case (8); PUT(snsptdrk)
        case (9); PUT0(rayl)
        end select
      end block
    case (29)
      block
        use rrsw_kg29
        select case (kind)
        case (1); PUT(absa)
        case (2); PUT(absb)
        case (3); PUT(selfref)
        case (4); PUT(forref)
        case (5); PUT(sfluxref)
        case (6); PUT(irradnce)
        case (7); PUT(facbrght)
        case (8); PUT(snsptdrk)
        case (9); PUT0(rayl)
        case (15); PUT(absh2o)
        case (16); PUT(absco2)
        end select
      end block
    end select
#undef PUT
#undef PUT0
  end function ref_table
end module ref_capi
