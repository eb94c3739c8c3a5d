! C-callable driver around the reference RRTMG_SW Fortran, compiled from the
! sources under the MiMA reference tree by oracle/build_ref.sh into
! oracle/_ref/librrtmg_sw_ref.so.  TEST INFRASTRUCTURE ONLY: it produces the
! golden vectors the oracle is pinned against; nothing in the product loads it.
!
! Entry points:
!   ref_sw_init(cpdair)            -> rrtmg_sw_ini (rrtmg_sw_init.f90:28)
!   ref_sw_run(...)                -> rrtmg_sw clear sky, icld = iaer = 0
!                                     (rrtmg_sw_rad.nomcica.f90:78), with the
!                                     cloud/aerosol arguments MiMA passes
!                                     (RR/rrtm_radiation.f90:701-713)
!   ref_sw_run_cloud(...)          -> rrtmg_sw with clouds (icld >= 1, inflgsw = 0)
!   ref_sw_run_cloud_wp(...)       -> the same with water-path cloud optics
!                                     (inflgsw = 2, swcldpr's coefficients)
!   ref_sw_column(...)             -> inatm_sw + setcoef_sw + taumol_sw for one
!                                     column: jp/jt/jt1, laytrop, taug, taur,
!                                     sfluxzen
!   ref_sw_reftra(...)             -> reftra_sw on given layers
module ref_sw_driver
  use iso_c_binding
  use parkind, only : im => kind_im, rb => kind_rb
  use parrrsw, only : nbndsw, ngptsw, naerec, mxmol, jpband
  use rrtmg_sw_init, only : rrtmg_sw_ini
  use rrtmg_sw_rad, only : rrtmg_sw, inatm_sw
  use rrtmg_sw_setcoef, only : setcoef_sw
  use rrtmg_sw_taumol, only : taumol_sw
  use rrtmg_sw_reftra, only : reftra_sw
  implicit none
contains

  subroutine ref_sw_init(cpdair) bind(C, name="ref_sw_init")
    real(c_double), value :: cpdair
    call rrtmg_sw_ini(cpdair)
  end subroutine ref_sw_init

  subroutine ref_sw_run(ncol, nlay, play, plev, tlay, tlev, tsfc, h2o, o3, co2, ch4, n2o, o2, &
                        asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, &
                        swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc) bind(C, name="ref_sw_run")
    integer(c_int), value :: ncol, nlay, dyofyr
    real(c_double), value :: adjes, scon
    real(c_double), intent(in) :: play(ncol, nlay), plev(ncol, nlay + 1), tlay(ncol, nlay)
    real(c_double), intent(in) :: tlev(ncol, nlay + 1), tsfc(ncol)
    real(c_double), intent(in) :: h2o(ncol, nlay), o3(ncol, nlay), co2(ncol, nlay)
    real(c_double), intent(in) :: ch4(ncol, nlay), n2o(ncol, nlay), o2(ncol, nlay)
    real(c_double), intent(in) :: asdir(ncol), asdif(ncol), aldir(ncol), aldif(ncol), coszen(ncol)
    real(c_double), intent(out) :: swuflx(ncol, nlay + 1), swdflx(ncol, nlay + 1), swhr(ncol, nlay)
    real(c_double), intent(out) :: swuflxc(ncol, nlay + 1), swdflxc(ncol, nlay + 1), swhrc(ncol, nlay)
    integer(kind=im) :: icld, iaer, dy
    real(kind=rb), allocatable :: zeros(:, :), tencl(:, :), cldb(:, :, :), aerb(:, :, :), ecaer(:, :, :)
    allocate(zeros(ncol, nlay), tencl(ncol, nlay), cldb(nbndsw, ncol, nlay))
    allocate(aerb(ncol, nlay, nbndsw), ecaer(ncol, nlay, naerec))
    zeros = 0.0_rb
    tencl = 10.0_rb
    cldb = 0.0_rb
    aerb = 0.0_rb
    ecaer = 0.0_rb
    icld = 0
    iaer = 0
    dy = dyofyr
    call rrtmg_sw(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, &
                  h2o, o3, co2, ch4, n2o, o2, asdir, asdif, aldir, aldif, &
                  coszen, adjes, dy, scon, 0, 0, 0, zeros, &
                  cldb, cldb, cldb, cldb, zeros, zeros, tencl, tencl, &
                  aerb, aerb, aerb, ecaer, &
                  swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc)
  end subroutine ref_sw_run

  ! rrtmg_sw with clouds: icld (the nomcica path takes any icld >= 1 the
  ! same way), inflgsw = 0 (cldprop_sw delta-scales the input optics), the
  ! caller's fractions (overcast or clear only: the reference stops on
  ! partial cloud) and cloud optics taucld, ssacld, asmcld, fsfcld
  ! (nbndsw, ncol, nlay); effective radii 10 um as MiMA passes them
  subroutine ref_sw_run_cloud(ncol, nlay, icld_in, play, plev, tlay, tlev, tsfc, h2o, o3, co2, ch4, n2o, o2, &
                              asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, &
                              cldfr, taucld, ssacld, asmcld, fsfcld, cicewp, cliqwp, &
                              swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc) bind(C, name="ref_sw_run_cloud")
    integer(c_int), value :: ncol, nlay, dyofyr, icld_in
    real(c_double), value :: adjes, scon
    real(c_double), intent(in) :: play(ncol, nlay), plev(ncol, nlay + 1), tlay(ncol, nlay)
    real(c_double), intent(in) :: tlev(ncol, nlay + 1), tsfc(ncol)
    real(c_double), intent(in) :: h2o(ncol, nlay), o3(ncol, nlay), co2(ncol, nlay)
    real(c_double), intent(in) :: ch4(ncol, nlay), n2o(ncol, nlay), o2(ncol, nlay)
    real(c_double), intent(in) :: asdir(ncol), asdif(ncol), aldir(ncol), aldif(ncol), coszen(ncol)
    real(c_double), intent(in) :: cldfr(ncol, nlay), cicewp(ncol, nlay), cliqwp(ncol, nlay)
    real(c_double), intent(in) :: taucld(nbndsw, ncol, nlay), ssacld(nbndsw, ncol, nlay)
    real(c_double), intent(in) :: asmcld(nbndsw, ncol, nlay), fsfcld(nbndsw, ncol, nlay)
    real(c_double), intent(out) :: swuflx(ncol, nlay + 1), swdflx(ncol, nlay + 1), swhr(ncol, nlay)
    real(c_double), intent(out) :: swuflxc(ncol, nlay + 1), swdflxc(ncol, nlay + 1), swhrc(ncol, nlay)
    integer(kind=im) :: icld, iaer, dy
    real(kind=rb), allocatable :: tencl(:, :), aerb(:, :, :), ecaer(:, :, :)
    allocate(tencl(ncol, nlay), aerb(ncol, nlay, nbndsw), ecaer(ncol, nlay, naerec))
    tencl = 10.0_rb
    aerb = 0.0_rb
    ecaer = 0.0_rb
    icld = icld_in
    iaer = 0
    dy = dyofyr
    call rrtmg_sw(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, &
                  h2o, o3, co2, ch4, n2o, o2, asdir, asdif, aldir, aldif, &
                  coszen, adjes, dy, scon, 0, 0, 0, cldfr, &
                  taucld, ssacld, asmcld, fsfcld, cicewp, cliqwp, tencl, tencl, &
                  aerb, aerb, aerb, ecaer, &
                  swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc)
  end subroutine ref_sw_run_cloud

  ! rrtmg_sw with clouds given by water paths (inflgsw, iceflgsw, liqflgsw;
  ! swcldpr's coefficients as rrtmg_sw_ini set them) and effective radii
  ! reice, reliq (ncol, nlay); taucld only enters the cloud test
  subroutine ref_sw_run_cloud_wp(ncol, nlay, icld_in, inflag_in, iceflag_in, liqflag_in, play, plev, tlay, &
                                 tlev, tsfc, h2o, o3, co2, ch4, n2o, o2, asdir, asdif, aldir, aldif, coszen, &
                                 adjes, dyofyr, scon, cldfr, taucld, cicewp, cliqwp, reice, reliq, &
                                 swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc) bind(C, name="ref_sw_run_cloud_wp")
    integer(c_int), value :: ncol, nlay, dyofyr, icld_in, inflag_in, iceflag_in, liqflag_in
    real(c_double), value :: adjes, scon
    real(c_double), intent(in) :: play(ncol, nlay), plev(ncol, nlay + 1), tlay(ncol, nlay)
    real(c_double), intent(in) :: tlev(ncol, nlay + 1), tsfc(ncol)
    real(c_double), intent(in) :: h2o(ncol, nlay), o3(ncol, nlay), co2(ncol, nlay)
    real(c_double), intent(in) :: ch4(ncol, nlay), n2o(ncol, nlay), o2(ncol, nlay)
    real(c_double), intent(in) :: asdir(ncol), asdif(ncol), aldir(ncol), aldif(ncol), coszen(ncol)
    real(c_double), intent(in) :: cldfr(ncol, nlay), cicewp(ncol, nlay), cliqwp(ncol, nlay)
    real(c_double), intent(in) :: reice(ncol, nlay), reliq(ncol, nlay), taucld(nbndsw, ncol, nlay)
    real(c_double), intent(out) :: swuflx(ncol, nlay + 1), swdflx(ncol, nlay + 1), swhr(ncol, nlay)
    real(c_double), intent(out) :: swuflxc(ncol, nlay + 1), swdflxc(ncol, nlay + 1), swhrc(ncol, nlay)
    integer(kind=im) :: icld, iaer, dy, inflag, iceflag, liqflag
    real(kind=rb), allocatable :: cldb(:, :, :), aerb(:, :, :), ecaer(:, :, :)
    allocate(cldb(nbndsw, ncol, nlay), aerb(ncol, nlay, nbndsw), ecaer(ncol, nlay, naerec))
    cldb = 0.0_rb
    aerb = 0.0_rb
    ecaer = 0.0_rb
    icld = icld_in
    iaer = 0
    dy = dyofyr
    inflag = inflag_in
    iceflag = iceflag_in
    liqflag = liqflag_in
    call rrtmg_sw(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, &
                  h2o, o3, co2, ch4, n2o, o2, asdir, asdif, aldir, aldif, &
                  coszen, adjes, dy, scon, inflag, iceflag, liqflag, cldfr, &
                  taucld, cldb, cldb, cldb, cicewp, cliqwp, reice, reliq, &
                  aerb, aerb, aerb, ecaer, &
                  swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc)
  end subroutine ref_sw_run_cloud_wp

  subroutine ref_sw_column(nlay, play, plev, tlay, tlev, h2o, o3, co2, ch4, n2o, o2, &
                           laytrop, jp, jt, jt1, fac00, fac01, fac10, fac11, colmol, &
                           taug, taur, sfluxzen) bind(C, name="ref_sw_column")
    integer(c_int), value :: nlay
    real(c_double), intent(in) :: play(1, nlay), plev(1, nlay + 1), tlay(1, nlay), tlev(1, nlay + 1)
    real(c_double), intent(in) :: h2o(1, nlay), o3(1, nlay), co2(1, nlay), ch4(1, nlay)
    real(c_double), intent(in) :: n2o(1, nlay), o2(1, nlay)
    integer(c_int), intent(out) :: laytrop, jp(nlay), jt(nlay), jt1(nlay)
    real(c_double), intent(out) :: fac00(nlay), fac01(nlay), fac10(nlay), fac11(nlay), colmol(nlay)
    real(c_double), intent(out) :: taug(nlay, ngptsw), taur(nlay, ngptsw), sfluxzen(ngptsw)
    integer(kind=im) :: nlayers, inflag, iceflag, liqflag, layswtch, laylow, lt
    integer(kind=im) :: ijp(nlay), ijt(nlay), ijt1(nlay), indself(nlay), indfor(nlay)
    real(kind=rb) :: pavel(nlay), tavel(nlay), pz(0:nlay), tz(0:nlay), tbound, pdp(nlay)
    real(kind=rb) :: coldry(nlay), wkl(mxmol, nlay), adjflux(jpband), solvar(jpband)
    real(kind=rb) :: cldfrac(nlay), tauc(nbndsw, nlay), ssac(nbndsw, nlay), asmc(nbndsw, nlay)
    real(kind=rb) :: fsfc(nbndsw, nlay), ciwp(nlay), clwp(nlay), rei(nlay), rel(nlay)
    real(kind=rb) :: taua(nlay, nbndsw), ssaa(nlay, nbndsw), asma(nlay, nbndsw)
    real(kind=rb) :: co2mult(nlay), colch4(nlay), colco2(nlay), colh2o(nlay), cmol(nlay)
    real(kind=rb) :: coln2o(nlay), colo2(nlay), colo3(nlay), f00(nlay), f01(nlay), f10(nlay), f11(nlay)
    real(kind=rb) :: selffac(nlay), selffrac(nlay), forfac(nlay), forfrac(nlay)
    real(kind=rb) :: zeros(1, nlay), tencl(1, nlay), cldb(nbndsw, 1, nlay), aerb(1, nlay, nbndsw)
    real(kind=rb) :: tsfc(1)
    zeros = 0.0_rb
    tencl = 10.0_rb
    cldb = 0.0_rb
    aerb = 0.0_rb
    tsfc = tlev(1, 1)
    call inatm_sw(1, nlay, 0, 0, play, plev, tlay, tlev, tsfc, h2o, o3, co2, ch4, n2o, o2, &
                  1.0_rb, 0, 1368.22_rb, 0, 0, 0, zeros, cldb, cldb, cldb, cldb, zeros, zeros, &
                  tencl, tencl, aerb, aerb, aerb, &
                  nlayers, pavel, pz, pdp, tavel, tz, tbound, coldry, wkl, &
                  adjflux, solvar, inflag, iceflag, liqflag, cldfrac, tauc, &
                  ssac, asmc, fsfc, ciwp, clwp, rei, rel, taua, ssaa, asma)
    call setcoef_sw(nlayers, pavel, tavel, pz, tz, tbound, coldry, wkl, &
                    lt, layswtch, laylow, ijp, ijt, ijt1, &
                    co2mult, colch4, colco2, colh2o, cmol, coln2o, &
                    colo2, colo3, f00, f01, f10, f11, &
                    selffac, selffrac, indself, forfac, forfrac, indfor)
    call taumol_sw(nlayers, colh2o, colco2, colch4, colo2, colo3, cmol, &
                   lt, ijp, ijt, ijt1, f00, f01, f10, f11, &
                   selffac, selffrac, indself, forfac, forfrac, indfor, &
                   sfluxzen, taug, taur)
    laytrop = lt
    jp = ijp
    jt = ijt
    jt1 = ijt1
    fac00 = f00
    fac01 = f01
    fac10 = f10
    fac11 = f11
    colmol = cmol
  end subroutine ref_sw_column

  subroutine ref_sw_reftra(nlay, gg, mu0, tau, w, pref, prefd, ptra, ptrad) bind(C, name="ref_sw_reftra")
    integer(c_int), value :: nlay
    real(c_double), value :: mu0
    real(c_double), intent(in) :: gg(nlay), tau(nlay), w(nlay)
    real(c_double), intent(out) :: pref(nlay), prefd(nlay), ptra(nlay), ptrad(nlay)
    logical :: chk(nlay)
    real(kind=rb) :: r(nlay + 1), rd(nlay + 1), t(nlay + 1), td(nlay + 1)
    chk = .true.
    call reftra_sw(nlay, chk, gg, mu0, tau, w, r, rd, t, td)
    pref = r(1:nlay)
    prefd = rd(1:nlay)
    ptra = t(1:nlay)
    ptrad = td(1:nlay)
  end subroutine ref_sw_reftra

end module ref_sw_driver
