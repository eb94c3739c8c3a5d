! C-callable driver around the reference RRTMG_SW McICA Fortran
! (mcica_subcol_gen_sw.f90, rrtmg_sw_cldprmc.f90, rrtmg_sw_spcvmc.f90 and the
! McICA rrtmg_sw_rad.f90, which MiMA's build leaves out), compiled from the
! sources under the MiMA reference tree by oracle/build_ref.sh into
! oracle/_ref/librrtmg_sw_mcica_ref.so.  TEST INFRASTRUCTURE ONLY: it
! produces the golden vectors the oracle is pinned against.
!
! Entry points:
!   ref_sw_mcica_init(cpdair)   -> rrtmg_sw_ini (rrtmg_sw_init.f90:28)
!   ref_sw_mcica_subcol(...)    -> mcica_subcol_sw (mcica_subcol_gen_sw.f90:47)
!   ref_sw_mcica_run(...)       -> rrtmg_sw of the McICA driver
!                                  (rrtmg_sw_rad.f90:80), iaer = 0, inflgsw = 0
!   ref_sw_mcica_run_wp(...)    -> the same with inflgsw = 2 (cldprmc_sw's
!                                  water-path optics) and given flags, radii
module ref_sw_mcica_driver
  use iso_c_binding
  use parkind, only : im => kind_im, rb => kind_rb
  use parrrsw, only : nbndsw, ngptsw, naerec
  use rrtmg_sw_init, only : rrtmg_sw_ini
  use rrtmg_sw_rad, only : rrtmg_sw
  use mcica_subcol_gen_sw, only : mcica_subcol_sw
  implicit none
  ! the generator ref_sw_mcica_subcol runs: 0 KISS, 1 Mersenne Twister
  integer(kind=im), save :: g_irng = 0
contains
  subroutine ref_sw_mcica_set_irng(irng) bind(C, name="ref_sw_mcica_set_irng")
    integer(c_int), value :: irng
    g_irng = irng
  end subroutine ref_sw_mcica_set_irng

  subroutine ref_sw_mcica_init(cpdair) bind(C, name="ref_sw_mcica_init")
    real(c_double), value :: cpdair
    call rrtmg_sw_ini(cpdair)
  end subroutine ref_sw_mcica_init

  ! the subcolumns for icld, permuteseed and the generator of
  ! ref_sw_mcica_set_irng (KISS unless set); effective radii
  ! zero (copied through, unread with inflgsw = 0)
  subroutine ref_sw_mcica_subcol(ncol, nlay, icld_in, permuteseed_in, play, cldfrac, ciwp, clwp, &
                                 tauc, ssac, asmc, fsfc, cldfmcl, ciwpmcl, clwpmcl, taucmcl, &
                                 ssacmcl, asmcmcl, fsfcmcl) bind(C, name="ref_sw_mcica_subcol")
    integer(c_int), value :: ncol, nlay, icld_in, permuteseed_in
    real(c_double), intent(in) :: play(ncol, nlay), cldfrac(ncol, nlay), ciwp(ncol, nlay), clwp(ncol, nlay)
    real(c_double), intent(in) :: tauc(nbndsw, ncol, nlay), ssac(nbndsw, ncol, nlay)
    real(c_double), intent(in) :: asmc(nbndsw, ncol, nlay), fsfc(nbndsw, ncol, nlay)
    real(c_double), intent(out) :: cldfmcl(ngptsw, ncol, nlay), ciwpmcl(ngptsw, ncol, nlay)
    real(c_double), intent(out) :: clwpmcl(ngptsw, ncol, nlay), taucmcl(ngptsw, ncol, nlay)
    real(c_double), intent(out) :: ssacmcl(ngptsw, ncol, nlay), asmcmcl(ngptsw, ncol, nlay)
    real(c_double), intent(out) :: fsfcmcl(ngptsw, ncol, nlay)
    integer(kind=im) :: icld, permuteseed, irng
    real(kind=rb), allocatable :: rei(:, :), rel(:, :), reicmcl(:, :), relqmcl(:, :)
    allocate(rei(ncol, nlay), rel(ncol, nlay), reicmcl(ncol, nlay), relqmcl(ncol, nlay))
    rei = 0.0_rb
    rel = 0.0_rb
    icld = icld_in
    permuteseed = permuteseed_in
    irng = g_irng
    call mcica_subcol_sw(1, ncol, nlay, icld, permuteseed, irng, play, cldfrac, ciwp, clwp, rei, rel, &
                         tauc, ssac, asmc, fsfc, cldfmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl, &
                         taucmcl, ssacmcl, asmcmcl, fsfcmcl)
  end subroutine ref_sw_mcica_subcol

  ! rrtmg_sw of the McICA driver on given subcolumns (inflgsw = 0); h2o and
  ! O3 are volume mixing ratios here
  subroutine ref_sw_mcica_run(ncol, nlay, icld_in, play, plev, tlay, tlev, tsfc, h2o, o3, co2, ch4, n2o, o2, &
                              asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, &
                              cldfmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl, ciwpmcl, clwpmcl, &
                              swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc) bind(C, name="ref_sw_mcica_run")
    integer(c_int), value :: ncol, nlay, dyofyr, icld_in
    real(c_double), value :: adjes, scon
    real(c_double), intent(in) :: play(ncol, nlay), plev(ncol, nlay + 1), tlay(ncol, nlay)
    real(c_double), intent(in) :: tlev(ncol, nlay + 1), tsfc(ncol)
    real(c_double), intent(in) :: h2o(ncol, nlay), o3(ncol, nlay), co2(ncol, nlay)
    real(c_double), intent(in) :: ch4(ncol, nlay), n2o(ncol, nlay), o2(ncol, nlay)
    real(c_double), intent(in) :: asdir(ncol), asdif(ncol), aldir(ncol), aldif(ncol), coszen(ncol)
    real(c_double), intent(in) :: cldfmcl(ngptsw, ncol, nlay), taucmcl(ngptsw, ncol, nlay)
    real(c_double), intent(in) :: ssacmcl(ngptsw, ncol, nlay), asmcmcl(ngptsw, ncol, nlay)
    real(c_double), intent(in) :: fsfcmcl(ngptsw, ncol, nlay), ciwpmcl(ngptsw, ncol, nlay)
    real(c_double), intent(in) :: clwpmcl(ngptsw, ncol, nlay)
    real(c_double), intent(out) :: swuflx(ncol, nlay + 1), swdflx(ncol, nlay + 1), swhr(ncol, nlay)
    real(c_double), intent(out) :: swuflxc(ncol, nlay + 1), swdflxc(ncol, nlay + 1), swhrc(ncol, nlay)
    integer(kind=im) :: icld, iaer, dy
    real(kind=rb), allocatable :: rzero(:, :), aerb(:, :, :), ecaer(:, :, :)
    allocate(rzero(ncol, nlay), aerb(ncol, nlay, nbndsw), ecaer(ncol, nlay, naerec))
    rzero = 0.0_rb
    aerb = 0.0_rb
    ecaer = 0.0_rb
    icld = icld_in
    iaer = 0
    dy = dyofyr
    call rrtmg_sw(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, &
                  h2o, o3, co2, ch4, n2o, o2, asdir, asdif, aldir, aldif, &
                  coszen, adjes, dy, scon, 0, 0, 0, cldfmcl, &
                  taucmcl, ssacmcl, asmcmcl, fsfcmcl, ciwpmcl, clwpmcl, rzero, rzero, &
                  aerb, aerb, aerb, ecaer, &
                  swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc)
  end subroutine ref_sw_mcica_run

  ! rrtmg_sw of the McICA driver with water-path optics (inflgsw = 2):
  ! cldprmc_sw forms each cloudy subcolumn's optics from ciwpmcl, clwpmcl
  ! and the layer radii reicmcl, relqmcl (ncol, nlay)
  subroutine ref_sw_mcica_run_wp(ncol, nlay, icld_in, iceflag, liqflag, play, plev, tlay, tlev, tsfc, h2o, o3, &
                                 co2, ch4, n2o, o2, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, &
                                 cldfmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl, &
                                 swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc) bind(C, name="ref_sw_mcica_run_wp")
    integer(c_int), value :: ncol, nlay, dyofyr, icld_in, iceflag, liqflag
    real(c_double), value :: adjes, scon
    real(c_double), intent(in) :: play(ncol, nlay), plev(ncol, nlay + 1), tlay(ncol, nlay)
    real(c_double), intent(in) :: tlev(ncol, nlay + 1), tsfc(ncol)
    real(c_double), intent(in) :: h2o(ncol, nlay), o3(ncol, nlay), co2(ncol, nlay)
    real(c_double), intent(in) :: ch4(ncol, nlay), n2o(ncol, nlay), o2(ncol, nlay)
    real(c_double), intent(in) :: asdir(ncol), asdif(ncol), aldir(ncol), aldif(ncol), coszen(ncol)
    real(c_double), intent(in) :: cldfmcl(ngptsw, ncol, nlay), taucmcl(ngptsw, ncol, nlay)
    real(c_double), intent(in) :: ssacmcl(ngptsw, ncol, nlay), asmcmcl(ngptsw, ncol, nlay)
    real(c_double), intent(in) :: fsfcmcl(ngptsw, ncol, nlay), ciwpmcl(ngptsw, ncol, nlay)
    real(c_double), intent(in) :: clwpmcl(ngptsw, ncol, nlay), reicmcl(ncol, nlay), relqmcl(ncol, nlay)
    real(c_double), intent(out) :: swuflx(ncol, nlay + 1), swdflx(ncol, nlay + 1), swhr(ncol, nlay)
    real(c_double), intent(out) :: swuflxc(ncol, nlay + 1), swdflxc(ncol, nlay + 1), swhrc(ncol, nlay)
    integer(kind=im) :: icld, iaer, dy, inflg, iceflg, liqflg
    real(kind=rb), allocatable :: aerb(:, :, :), ecaer(:, :, :)
    allocate(aerb(ncol, nlay, nbndsw), ecaer(ncol, nlay, naerec))
    aerb = 0.0_rb
    ecaer = 0.0_rb
    icld = icld_in
    iaer = 0
    dy = dyofyr
    inflg = 2
    iceflg = iceflag
    liqflg = liqflag
    call rrtmg_sw(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, &
                  h2o, o3, co2, ch4, n2o, o2, asdir, asdif, aldir, aldif, &
                  coszen, adjes, dy, scon, inflg, iceflg, liqflg, cldfmcl, &
                  taucmcl, ssacmcl, asmcmcl, fsfcmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl, &
                  aerb, aerb, aerb, ecaer, &
                  swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc)
  end subroutine ref_sw_mcica_run_wp
end module ref_sw_mcica_driver
