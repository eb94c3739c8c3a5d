! C-callable driver around the reference RRTMG_LW column code, compiled from
! the sources under the MiMA reference tree by oracle/build_ref.sh into
! oracle/_ref/librrtmg_lw_ref.so.  TEST INFRASTRUCTURE ONLY: it produces the
! golden vectors the LW oracle is pinned against; nothing in the product loads it.
!
! What is compiled from the reference: the LW modules (LW/modules/*.f90) and
! rrtmg_lw_setcoef.f90, rrtmg_lw_taumol.f90, rrtmg_lw_rtrnmr.f90,
! rrtmg_lw_rtrn.f90, rrtmg_lw_cldprop.f90, rrtmg_lw_rad.nomcica.f90.
! What is not: rrtmg_lw_init.f90 (its rrtmg_lw_ini calls lw_kgb01..16 of the
! absent AER data file rrtmg_lw_k_g.f90).  Instead of running the init, this
! driver fills the module state the column code reads:
!   * ref_lw_set: the reduced k-tables of rrlw_kg01..16 (ka/absa, kb/absb,
!     selfref, forref, fracrefa/b, minor-gas tables), given by the caller;
!   * ref_lw_set_bands: ngc/ngs/ngb, nspa/nspb, delwave, as
!     lwdatinit/lwcmbdat define them (rrtmg_lw_init.f90:178-363);
!   * ref_lw_init: grav, avogad, heatfac (rrtmg_lw_init.f90:228-281), the
!     transmittance tables (rrtmg_lw_init.f90:106-123), and the reference's own
!     lwatmref / lwavplank / lwavplankderiv (rrtmg_lw_setcoef.f90:418-1990).
! The reference column code then runs unchanged:
!   ref_lw_run    -> rrtmg_lw (rrtmg_lw_rad.nomcica.f90:80), clear sky, with
!                    the cloud arguments MiMA passes (RR/rrtm_radiation.f90:730-749)
!   ref_lw_run_cloud -> rrtmg_lw with icld = 1..3 and inflglw = 0 (cldprop,
!                    then rtrn or rtrnmr with the caller's cloud profiles)
!   ref_lw_set_cld + ref_lw_run_cloud_wp -> the same with inflglw = 1, 2
!                    (cldprop's water-path optics on lwcldpr's coefficients,
!                    which the caller passes as extracted from rrtmg_lw_init.f90)
!   ref_lw_column -> inatm + setcoef + taumol for one column
module ref_lw_driver
  use iso_c_binding
  use parkind, only : im => kind_im, rb => kind_rb
  use parrrtm, only : nbndlw, ngptlw, mxmol, maxxsec
  use rrtmg_lw_rad, only : rrtmg_lw, inatm
  use rrtmg_lw_setcoef, only : setcoef, lwatmref, lwavplank, lwavplankderiv
  use rrtmg_lw_taumol, only : taumol
  use rrlw_cld, only : abscld1, absice0, absice1, absice2, absice3, absliq0, absliq1
  include 'ref_lw_setters.inc'
  implicit none
contains

  include 'ref_lw_common.inc'

  subroutine ref_lw_run(ncol, nlay, play, plev, tlay, tlev, tsfc, h2o, o3, co2, ch4, n2o, o2, &
                        cfc11, cfc12, cfc22, ccl4, emis, tauaer, &
                        uflx, dflx, hr, uflxc, dflxc, hrc) bind(C, name="ref_lw_run")
    integer(c_int), value :: ncol, nlay
    real(c_double), intent(in) :: play(ncol, nlay), plev(ncol, nlay + 1), tlay(ncol, nlay)
    real(c_double), intent(in) :: tlev(ncol, nlay + 1), tsfc(ncol)
    real(c_double), intent(in) :: h2o(ncol, nlay), o3(ncol, nlay), co2(ncol, nlay)
    real(c_double), intent(in) :: ch4(ncol, nlay), n2o(ncol, nlay), o2(ncol, nlay)
    real(c_double), intent(in) :: cfc11(ncol, nlay), cfc12(ncol, nlay), cfc22(ncol, nlay)
    real(c_double), intent(in) :: ccl4(ncol, nlay), emis(ncol, nbndlw), tauaer(ncol, nlay, nbndlw)
    real(c_double), intent(out) :: uflx(ncol, nlay + 1), dflx(ncol, nlay + 1), hr(ncol, nlay)
    real(c_double), intent(out) :: uflxc(ncol, nlay + 1), dflxc(ncol, nlay + 1), hrc(ncol, nlay)
    integer(kind=im) :: icld, idrv
    real(kind=rb), allocatable :: zeros(:, :), tencl(:, :), taucld(:, :, :)
    allocate(zeros(ncol, nlay), tencl(ncol, nlay), taucld(nbndlw, ncol, nlay))
    zeros = 0.0_rb
    tencl = 10.0_rb
    taucld = 0.0_rb
    icld = 0
    idrv = 0
    call rrtmg_lw(ncol, nlay, icld, idrv, play, plev, tlay, tlev, tsfc, &
                  h2o, o3, co2, ch4, n2o, o2, cfc11, cfc12, cfc22, ccl4, emis, &
                  0, 0, 0, zeros, taucld, zeros, zeros, tencl, tencl, tauaer, &
                  uflx, dflx, hr, uflxc, dflxc, hrc)
  end subroutine ref_lw_run

  ! rrtmg_lw with clouds: icld (1 random overlap -> rtrn, 2/3 maximum-random
  ! -> rtrnmr), inflglw = 0 (cldprop takes the in-cloud optical depths
  ! taucld(nbndlw, ncol, nlay) as given), the caller's cloud fraction and
  ! water paths; effective radii 10 um as MiMA passes them
  subroutine ref_lw_run_cloud(ncol, nlay, icld_in, play, plev, tlay, tlev, tsfc, h2o, o3, co2, ch4, n2o, &
                              o2, cfc11, cfc12, cfc22, ccl4, emis, cldfr, taucld, cicewp, cliqwp, tauaer, &
                              uflx, dflx, hr, uflxc, dflxc, hrc) bind(C, name="ref_lw_run_cloud")
    integer(c_int), value :: ncol, nlay, icld_in
    real(c_double), intent(in) :: play(ncol, nlay), plev(ncol, nlay + 1), tlay(ncol, nlay)
    real(c_double), intent(in) :: tlev(ncol, nlay + 1), tsfc(ncol)
    real(c_double), intent(in) :: h2o(ncol, nlay), o3(ncol, nlay), co2(ncol, nlay)
    real(c_double), intent(in) :: ch4(ncol, nlay), n2o(ncol, nlay), o2(ncol, nlay)
    real(c_double), intent(in) :: cfc11(ncol, nlay), cfc12(ncol, nlay), cfc22(ncol, nlay)
    real(c_double), intent(in) :: ccl4(ncol, nlay), emis(ncol, nbndlw), tauaer(ncol, nlay, nbndlw)
    real(c_double), intent(in) :: cldfr(ncol, nlay), taucld(nbndlw, ncol, nlay)
    real(c_double), intent(in) :: cicewp(ncol, nlay), cliqwp(ncol, nlay)
    real(c_double), intent(out) :: uflx(ncol, nlay + 1), dflx(ncol, nlay + 1), hr(ncol, nlay)
    real(c_double), intent(out) :: uflxc(ncol, nlay + 1), dflxc(ncol, nlay + 1), hrc(ncol, nlay)
    integer(kind=im) :: icld, idrv
    real(kind=rb), allocatable :: tencl(:, :)
    allocate(tencl(ncol, nlay))
    tencl = 10.0_rb
    icld = icld_in
    idrv = 0
    call rrtmg_lw(ncol, nlay, icld, idrv, play, plev, tlay, tlev, tsfc, &
                  h2o, o3, co2, ch4, n2o, o2, cfc11, cfc12, cfc22, ccl4, emis, &
                  0, 0, 0, cldfr, taucld, cicewp, cliqwp, tencl, tencl, tauaer, &
                  uflx, dflx, hr, uflxc, dflxc, hrc)
  end subroutine ref_lw_run_cloud

  ! lwcldpr's cloud absorption coefficients (rrtmg_lw_init.f90, module
  ! rrlw_cld), which this build does not run: the caller passes the values
  ! extracted from that file (rrtmg_lw_cld.tbl)
  subroutine ref_lw_set_cld(c1, i0, i1, i2, i3, l0, l1) bind(C, name="ref_lw_set_cld")
    real(c_double), value :: c1, l0
    real(c_double), intent(in) :: i0(2), i1(2, 5), i2(43, 16), i3(46, 16), l1(58, 16)
    abscld1 = c1
    absice0 = i0
    absice1 = i1
    absice2 = i2
    absice3 = i3
    absliq0 = l0
    absliq1 = l1
  end subroutine ref_lw_set_cld

  ! rrtmg_lw with clouds given by water paths: inflglw, iceflglw, liqflglw
  ! and the effective radii / sizes reice, reliq (ncol, nlay)
  subroutine ref_lw_run_cloud_wp(ncol, nlay, icld_in, inflag_in, iceflag_in, liqflag_in, play, plev, tlay, &
                                 tlev, tsfc, h2o, o3, co2, ch4, n2o, o2, cfc11, cfc12, cfc22, ccl4, emis, &
                                 cldfr, taucld, cicewp, cliqwp, reice, reliq, tauaer, &
                                 uflx, dflx, hr, uflxc, dflxc, hrc) bind(C, name="ref_lw_run_cloud_wp")
    integer(c_int), value :: ncol, nlay, icld_in, inflag_in, iceflag_in, liqflag_in
    real(c_double), intent(in) :: play(ncol, nlay), plev(ncol, nlay + 1), tlay(ncol, nlay)
    real(c_double), intent(in) :: tlev(ncol, nlay + 1), tsfc(ncol)
    real(c_double), intent(in) :: h2o(ncol, nlay), o3(ncol, nlay), co2(ncol, nlay)
    real(c_double), intent(in) :: ch4(ncol, nlay), n2o(ncol, nlay), o2(ncol, nlay)
    real(c_double), intent(in) :: cfc11(ncol, nlay), cfc12(ncol, nlay), cfc22(ncol, nlay)
    real(c_double), intent(in) :: ccl4(ncol, nlay), emis(ncol, nbndlw), tauaer(ncol, nlay, nbndlw)
    real(c_double), intent(in) :: cldfr(ncol, nlay), taucld(nbndlw, ncol, nlay)
    real(c_double), intent(in) :: cicewp(ncol, nlay), cliqwp(ncol, nlay), reice(ncol, nlay), reliq(ncol, nlay)
    real(c_double), intent(out) :: uflx(ncol, nlay + 1), dflx(ncol, nlay + 1), hr(ncol, nlay)
    real(c_double), intent(out) :: uflxc(ncol, nlay + 1), dflxc(ncol, nlay + 1), hrc(ncol, nlay)
    integer(kind=im) :: icld, idrv, inflag, iceflag, liqflag
    icld = icld_in
    idrv = 0
    inflag = inflag_in
    iceflag = iceflag_in
    liqflag = liqflag_in
    call rrtmg_lw(ncol, nlay, icld, idrv, play, plev, tlay, tlev, tsfc, &
                  h2o, o3, co2, ch4, n2o, o2, cfc11, cfc12, cfc22, ccl4, emis, &
                  inflag, iceflag, liqflag, cldfr, taucld, cicewp, cliqwp, reice, reliq, tauaer, &
                  uflx, dflx, hr, uflxc, dflxc, hrc)
  end subroutine ref_lw_run_cloud_wp

  subroutine ref_lw_column(nlay, play, plev, tlay, tlev, tsfc, h2o, o3, co2, ch4, n2o, o2, &
                           cfc11, cfc12, cfc22, ccl4, &
                           laytrop, jp, jt, jt1, fac00, fac01, fac10, fac11, pwvcm, &
                           planklay, planklev, plankbnd, taug, fracs) bind(C, name="ref_lw_column")
    integer(c_int), value :: nlay
    real(c_double), intent(in) :: play(1, nlay), plev(1, nlay + 1), tlay(1, nlay), tlev(1, nlay + 1)
    real(c_double), intent(in) :: tsfc(1), h2o(1, nlay), o3(1, nlay), co2(1, nlay), ch4(1, nlay)
    real(c_double), intent(in) :: n2o(1, nlay), o2(1, nlay), cfc11(1, nlay), cfc12(1, nlay)
    real(c_double), intent(in) :: cfc22(1, nlay), ccl4(1, nlay)
    integer(c_int), intent(out) :: laytrop, jp(nlay), jt(nlay), jt1(nlay)
    real(c_double), intent(out) :: fac00(nlay), fac01(nlay), fac10(nlay), fac11(nlay), pwvcm
    real(c_double), intent(out) :: planklay(nlay, nbndlw), planklev(0:nlay, nbndlw), plankbnd(nbndlw)
    real(c_double), intent(out) :: taug(nlay, ngptlw), fracs(nlay, ngptlw)
    integer(kind=im) :: nlayers, inflag, iceflag, liqflag, lt
    integer(kind=im) :: ijp(nlay + 1), ijt(nlay + 1), ijt1(nlay + 1)
    integer(kind=im) :: indself(nlay + 1), indfor(nlay + 1), indminor(nlay + 1)
    real(kind=rb) :: pavel(nlay + 1), tavel(nlay + 1), pz(0:nlay + 1), tz(0:nlay + 1), tbound
    real(kind=rb) :: coldry(nlay + 1), wbrodl(nlay + 1), wkl(mxmol, nlay + 1), wx(maxxsec, nlay + 1)
    real(kind=rb) :: semiss(nbndlw), pw, emis(1, nbndlw)
    real(kind=rb) :: cldfrac(nlay + 1), tauc(nbndlw, nlay + 1), ciwp(nlay + 1), clwp(nlay + 1)
    real(kind=rb) :: rei(nlay + 1), rel(nlay + 1), taua(nlay + 1, nbndlw)
    real(kind=rb) :: plkl(nlay + 1, nbndlw), plkv(0:nlay + 1, nbndlw), plkb(nbndlw), dplk(nbndlw)
    real(kind=rb), dimension(nlay + 1) :: colh2o, colco2, colo3, coln2o, colco, colch4, colo2, colbrd
    real(kind=rb), dimension(nlay + 1) :: f00, f01, f10, f11
    real(kind=rb), dimension(nlay + 1) :: rh2oco2, rh2oco2_1, rh2oo3, rh2oo3_1, rh2on2o, rh2on2o_1
    real(kind=rb), dimension(nlay + 1) :: rh2och4, rh2och4_1, rn2oco2, rn2oco2_1, ro3co2, ro3co2_1
    real(kind=rb), dimension(nlay + 1) :: selffac, selffrac, forfac, forfrac
    real(kind=rb), dimension(nlay + 1) :: minorfrac, scaleminor, scaleminorn2
    real(kind=rb) :: tg(nlay + 1, ngptlw), fr(nlay + 1, ngptlw)
    real(kind=rb) :: zeros(1, nlay), tencl(1, nlay), taucld(nbndlw, 1, nlay), tauaer(1, nlay, nbndlw)
    zeros = 0.0_rb
    tencl = 10.0_rb
    taucld = 0.0_rb
    tauaer = 0.0_rb
    emis = 1.0_rb
    call inatm(1, nlay, 0, 10, play, plev, tlay, tlev, tsfc, h2o, o3, co2, ch4, n2o, o2, &
               cfc11, cfc12, cfc22, ccl4, emis, 0, 0, 0, &
               zeros, taucld, zeros, zeros, tencl, tencl, tauaer, &
               nlayers, pavel, pz, tavel, tz, tbound, semiss, coldry, &
               wkl, wbrodl, wx, pw, inflag, iceflag, liqflag, &
               cldfrac, tauc, ciwp, clwp, rei, rel, taua)
    call setcoef(nlayers, 1, pavel, tavel, tz, tbound, semiss, coldry, wkl, wbrodl, &
                 lt, ijp, ijt, ijt1, plkl, plkv, plkb, 0, dplk, &
                 colh2o, colco2, colo3, coln2o, colco, colch4, colo2, colbrd, f00, f01, f10, f11, &
                 rh2oco2, rh2oco2_1, rh2oo3, rh2oo3_1, rh2on2o, rh2on2o_1, &
                 rh2och4, rh2och4_1, rn2oco2, rn2oco2_1, ro3co2, ro3co2_1, &
                 selffac, selffrac, indself, forfac, forfrac, indfor, &
                 minorfrac, scaleminor, scaleminorn2, indminor)
    call taumol(nlayers, pavel, wx, coldry, lt, ijp, ijt, ijt1, plkl, plkv, plkb, &
                colh2o, colco2, colo3, coln2o, colco, colch4, colo2, colbrd, f00, f01, f10, f11, &
                rh2oco2, rh2oco2_1, rh2oo3, rh2oo3_1, rh2on2o, rh2on2o_1, &
                rh2och4, rh2och4_1, rn2oco2, rn2oco2_1, ro3co2, ro3co2_1, &
                selffac, selffrac, indself, forfac, forfrac, indfor, &
                minorfrac, scaleminor, scaleminorn2, indminor, fr, tg)
    laytrop = lt
    jp = ijp(1:nlay)
    jt = ijt(1:nlay)
    jt1 = ijt1(1:nlay)
    fac00 = f00(1:nlay)
    fac01 = f01(1:nlay)
    fac10 = f10(1:nlay)
    fac11 = f11(1:nlay)
    pwvcm = pw
    planklay = plkl(1:nlay, :)
    planklev = plkv(0:nlay, :)
    plankbnd = plkb
    taug = tg(1:nlay, :)
    fracs = fr(1:nlay, :)
  end subroutine ref_lw_column

end module ref_lw_driver
