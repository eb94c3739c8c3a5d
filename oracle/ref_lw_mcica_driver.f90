! C-callable driver around the reference RRTMG_LW McICA code, compiled from
! the sources under the MiMA reference tree by oracle/build_ref.sh into
! oracle/_ref/librrtmg_lw_mcica_ref.so.  TEST INFRASTRUCTURE ONLY: it
! produces the golden vectors the oracle's McICA path is pinned against.
!
! Compiled from the reference: the LW modules, rrtmg_lw_setcoef.f90,
! rrtmg_lw_taumol.f90, rrtmg_lw_cldprmc.f90, rrtmg_lw_rtrnmc.f90,
! mcica_random_numbers.f90, mcica_subcol_gen_lw.f90 and the McICA driver
! rrtmg_lw_rad.f90 (whose module shares its name with the nomcica driver's,
! hence a library of its own).  The module state of the absent init comes
! from the same setters as in ref_lw_driver.f90 (ref_lw_common.inc).
!   ref_lw_mcica_subcol -> mcica_subcol_lw (mcica_subcol_gen_lw.f90:48)
!   ref_lw_mcica_run    -> rrtmg_lw of rrtmg_lw_rad.f90 (:80) on subcolumns
module ref_lw_mcica_driver
  use iso_c_binding
  use parkind, only : im => kind_im, rb => kind_rb
  use parrrtm, only : nbndlw, ngptlw, mxmol, maxxsec
  use rrtmg_lw_rad, only : rrtmg_lw
  use mcica_subcol_gen_lw, only : mcica_subcol_lw
  use rrtmg_lw_setcoef, only : lwatmref, lwavplank, lwavplankderiv
  use rrlw_cld, only : abscld1, absice0, absice1, absice2, absice3, absliq0, absliq1
  include 'ref_lw_setters.inc'
  implicit none
contains

  include 'ref_lw_common.inc'

  ! the stochastic subcolumns of one call: icld 1..3, irng (0 KISS, 1
  ! Mersenne twister), permuteseed; play (ncol, nlay) in hPa
  subroutine ref_lw_mcica_subcol(ncol, nlay, icld_in, permuteseed, irng_in, play, cldfrac, ciwp, clwp, rei, rel, &
                                 tauc, cldfmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl, taucmcl) &
                                 bind(C, name="ref_lw_mcica_subcol")
    integer(c_int), value :: ncol, nlay, icld_in, permuteseed, irng_in
    real(c_double), intent(in) :: play(ncol, nlay), cldfrac(ncol, nlay), ciwp(ncol, nlay), clwp(ncol, nlay)
    real(c_double), intent(in) :: rei(ncol, nlay), rel(ncol, nlay), tauc(nbndlw, ncol, nlay)
    real(c_double), intent(out) :: cldfmcl(ngptlw, ncol, nlay), ciwpmcl(ngptlw, ncol, nlay)
    real(c_double), intent(out) :: clwpmcl(ngptlw, ncol, nlay), taucmcl(ngptlw, ncol, nlay)
    real(c_double), intent(out) :: reicmcl(ncol, nlay), relqmcl(ncol, nlay)
    integer(kind=im) :: irng
    irng = irng_in
    call mcica_subcol_lw(1, ncol, nlay, icld_in, permuteseed, irng, play, cldfrac, ciwp, clwp, rei, rel, tauc, &
                         cldfmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl, taucmcl)
  end subroutine ref_lw_mcica_subcol

  ! rrtmg_lw (McICA) on given subcolumns, inflglw = 0, idrv = 0
  subroutine ref_lw_mcica_run(ncol, nlay, icld_in, play, plev, tlay, tlev, tsfc, h2o, o3, co2, ch4, n2o, o2, &
                              cfc11, cfc12, cfc22, ccl4, emis, cldfmcl, taucmcl, ciwpmcl, clwpmcl, reicmcl, &
                              relqmcl, tauaer, uflx, dflx, hr, uflxc, dflxc, hrc) bind(C, name="ref_lw_mcica_run")
    integer(c_int), value :: ncol, nlay, icld_in
    real(c_double), intent(in) :: play(ncol, nlay), plev(ncol, nlay + 1), tlay(ncol, nlay)
    real(c_double), intent(in) :: tlev(ncol, nlay + 1), tsfc(ncol)
    real(c_double), intent(in) :: h2o(ncol, nlay), o3(ncol, nlay), co2(ncol, nlay)
    real(c_double), intent(in) :: ch4(ncol, nlay), n2o(ncol, nlay), o2(ncol, nlay)
    real(c_double), intent(in) :: cfc11(ncol, nlay), cfc12(ncol, nlay), cfc22(ncol, nlay)
    real(c_double), intent(in) :: ccl4(ncol, nlay), emis(ncol, nbndlw), tauaer(ncol, nlay, nbndlw)
    real(c_double), intent(in) :: cldfmcl(ngptlw, ncol, nlay), taucmcl(ngptlw, ncol, nlay)
    real(c_double), intent(in) :: ciwpmcl(ngptlw, ncol, nlay), clwpmcl(ngptlw, ncol, nlay)
    real(c_double), intent(in) :: reicmcl(ncol, nlay), relqmcl(ncol, nlay)
    real(c_double), intent(out) :: uflx(ncol, nlay + 1), dflx(ncol, nlay + 1), hr(ncol, nlay)
    real(c_double), intent(out) :: uflxc(ncol, nlay + 1), dflxc(ncol, nlay + 1), hrc(ncol, nlay)
    integer(kind=im) :: icld, idrv
    icld = icld_in
    idrv = 0
    call rrtmg_lw(ncol, nlay, icld, idrv, play, plev, tlay, tlev, tsfc, &
                  h2o, o3, co2, ch4, n2o, o2, cfc11, cfc12, cfc22, ccl4, emis, &
                  0, 0, 0, cldfmcl, taucmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl, tauaer, &
                  uflx, dflx, hr, uflxc, dflxc, hrc)
  end subroutine ref_lw_mcica_run

  ! lwcldpr's coefficients for this library's copy of rrlw_cld (as
  ! ref_lw_set_cld of ref_lw_driver.f90)
  subroutine ref_lw_mcica_set_cld(c1, i0, i1, i2, i3, l0, l1) bind(C, name="ref_lw_mcica_set_cld")
    real(c_double), value :: c1, l0
    real(c_double), intent(in) :: i0(2), i1(2, 5), i2(43, 16), i3(46, 16), l1(58, 16)
    abscld1 = c1
    absice0 = i0
    absice1 = i1
    absice2 = i2
    absice3 = i3
    absliq0 = l0
    absliq1 = l1
  end subroutine ref_lw_mcica_set_cld

  ! rrtmg_lw of the McICA driver with cldprmc's water-path optics
  ! (inflglw = 2, iceflglw, liqflglw) on the subcolumns' water paths
  subroutine ref_lw_mcica_run_wp(ncol, nlay, icld_in, inflag_in, iceflag_in, liqflag_in, play, plev, tlay, &
                                 tlev, tsfc, h2o, o3, co2, ch4, n2o, o2, cfc11, cfc12, cfc22, ccl4, emis, &
                                 cldfmcl, taucmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl, tauaer, &
                                 uflx, dflx, hr, uflxc, dflxc, hrc) bind(C, name="ref_lw_mcica_run_wp")
    integer(c_int), value :: ncol, nlay, icld_in, inflag_in, iceflag_in, liqflag_in
    real(c_double), intent(in) :: play(ncol, nlay), plev(ncol, nlay + 1), tlay(ncol, nlay)
    real(c_double), intent(in) :: tlev(ncol, nlay + 1), tsfc(ncol)
    real(c_double), intent(in) :: h2o(ncol, nlay), o3(ncol, nlay), co2(ncol, nlay)
    real(c_double), intent(in) :: ch4(ncol, nlay), n2o(ncol, nlay), o2(ncol, nlay)
    real(c_double), intent(in) :: cfc11(ncol, nlay), cfc12(ncol, nlay), cfc22(ncol, nlay)
    real(c_double), intent(in) :: ccl4(ncol, nlay), emis(ncol, nbndlw), tauaer(ncol, nlay, nbndlw)
    real(c_double), intent(in) :: cldfmcl(ngptlw, ncol, nlay), taucmcl(ngptlw, ncol, nlay)
    real(c_double), intent(in) :: ciwpmcl(ngptlw, ncol, nlay), clwpmcl(ngptlw, ncol, nlay)
    real(c_double), intent(in) :: reicmcl(ncol, nlay), relqmcl(ncol, nlay)
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
                  inflag, iceflag, liqflag, cldfmcl, taucmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl, tauaer, &
                  uflx, dflx, hr, uflxc, dflxc, hrc)
  end subroutine ref_lw_mcica_run_wp

end module ref_lw_mcica_driver
