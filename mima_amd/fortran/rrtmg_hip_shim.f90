! Fortran drop-in for the RRTMG modules MiMA uses, on top of the C ABI in
! include/rrtmg_hip.h.  It re-provides, with the reference's module names,
! procedure names and dummy-argument lists:
!   parkind        (rrtmg_*/gcm_model/modules/parkind.f90)       kind_im, kind_rb
!   parrrtm        (rrtmg_lw/gcm_model/modules/parrrtm.f90)       nbndlw, ngptlw
!   parrrsw        (rrtmg_sw/gcm_model/modules/parrrsw.f90)       nbndsw, ngptsw, naerec
!   rrtmg_lw_init  rrtmg_lw_ini(cpdair)   (rrtmg_lw_init.f90:28)
!   rrtmg_sw_init  rrtmg_sw_ini(cpdair)   (rrtmg_sw_init.f90:28)
!   rrtmg_lw_rad   rrtmg_lw(...)          (rrtmg_lw_rad.nomcica.f90:80-89)
!   rrtmg_sw_rad   rrtmg_sw(...)          (rrtmg_sw_rad.nomcica.f90:78-88)
! so physics_driver.f90:132-135 / :577-578 and rrtm_radiation.f90:473-474 /
! :686-749 compile and run unchanged against it.  Arrays may be any actual
! argument the reference accepts (sections, expression temporaries); the
! compiler passes contiguous copies to the C entry points where needed.
! A failing call stops the run with the library's message, as the
! reference's own `stop` statements do.
module parkind
  implicit none
  integer, parameter :: kind_im = selected_int_kind(6)
  integer, parameter :: kind_rb = selected_real_kind(12)
end module parkind

module parrrtm
  use parkind, only : im => kind_im
  implicit none
  integer(kind=im), parameter :: nbndlw = 16
  integer(kind=im), parameter :: ngptlw = 140
  integer(kind=im), parameter :: mxlay = 203
end module parrrtm

module parrrsw
  use parkind, only : im => kind_im, rb => kind_rb
  implicit none
  integer(kind=im), parameter :: nbndsw = 14
  integer(kind=im), parameter :: ngptsw = 112
  integer(kind=im), parameter :: naerec = 6
  integer(kind=im), parameter :: mxlay = 203
  integer(kind=im), parameter :: jpb1 = 16, jpb2 = 29
  real(kind=rb), parameter :: rrsw_scon = 1.36822e+03_rb
end module parrrsw

module rrtmg_hip_capi
  use iso_c_binding
  implicit none
  integer(c_int), parameter :: RRTMG_HIP_HOST = 0
  interface
    integer(c_int) function c_sw_init(cpdair, table_dir) bind(C, name="rrtmg_hip_sw_init")
      import :: c_int, c_double, c_ptr
      real(c_double), value :: cpdair
      type(c_ptr), value :: table_dir
    end function
    integer(c_int) function c_lw_init(cpdair, table_dir) bind(C, name="rrtmg_hip_lw_init")
      import :: c_int, c_double, c_ptr
      real(c_double), value :: cpdair
      type(c_ptr), value :: table_dir
    end function
    integer(c_int) function c_set_precision(mode) bind(C, name="rrtmg_hip_set_precision")
      import :: c_int
      integer(c_int), value :: mode
    end function
    type(c_ptr) function c_last_error() bind(C, name="rrtmg_hip_last_error")
      import :: c_ptr
    end function
    integer(c_int) function c_sw(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, &
                                 h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, &
                                 asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, &
                                 swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc, &
                                 ptr_kind, stream) bind(C, name="rrtmg_hip_sw")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: ncol, nlay, icld, iaer, dyofyr, ptr_kind
      real(c_double), intent(in) :: play(*), plev(*), tlay(*), tlev(*), tsfc(*)
      real(c_double), intent(in) :: h2ovmr(*), o3vmr(*), co2vmr(*), ch4vmr(*), n2ovmr(*), o2vmr(*)
      real(c_double), intent(in) :: asdir(*), asdif(*), aldir(*), aldif(*), coszen(*)
      real(c_double), value :: adjes, scon
      real(c_double), intent(out) :: swuflx(*), swdflx(*), swhr(*), swuflxc(*), swdflxc(*), swhrc(*)
      type(c_ptr), value :: stream
    end function
    integer(c_int) function c_lw(ncol, nlay, icld, idrv, play, plev, tlay, tlev, tsfc, &
                                 h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, &
                                 cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, tauaer, &
                                 uflx, dflx, hr, uflxc, dflxc, hrc, ptr_kind, stream) &
                                 bind(C, name="rrtmg_hip_lw")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: ncol, nlay, icld, idrv, ptr_kind
      real(c_double), intent(in) :: play(*), plev(*), tlay(*), tlev(*), tsfc(*)
      real(c_double), intent(in) :: h2ovmr(*), o3vmr(*), co2vmr(*), ch4vmr(*), n2ovmr(*), o2vmr(*)
      real(c_double), intent(in) :: cfc11vmr(*), cfc12vmr(*), cfc22vmr(*), ccl4vmr(*), emis(*), tauaer(*)
      real(c_double), intent(out) :: uflx(*), dflx(*), hr(*), uflxc(*), dflxc(*), hrc(*)
      type(c_ptr), value :: stream
    end function
    integer(c_int) function c_sw_cloud(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, &
                                       h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, &
                                       asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, &
                                       inflgsw, iceflgsw, liqflgsw, cldfr, taucld, ssacld, asmcld, fsfcld, &
                                       cicewp, cliqwp, reice, reliq, tauaer, ssaaer, asmaer, ecaer, &
                                       swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc, ptr_kind, stream) &
                                       bind(C, name="rrtmg_hip_sw_cloud")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: ncol, nlay, icld, iaer, dyofyr, inflgsw, iceflgsw, liqflgsw, ptr_kind
      real(c_double), intent(in) :: play(*), plev(*), tlay(*), tlev(*), tsfc(*)
      real(c_double), intent(in) :: h2ovmr(*), o3vmr(*), co2vmr(*), ch4vmr(*), n2ovmr(*), o2vmr(*)
      real(c_double), intent(in) :: asdir(*), asdif(*), aldir(*), aldif(*), coszen(*)
      real(c_double), value :: adjes, scon
      real(c_double), intent(in) :: cldfr(*), taucld(*), ssacld(*), asmcld(*), fsfcld(*)
      real(c_double), intent(in) :: cicewp(*), cliqwp(*), reice(*), reliq(*)
      real(c_double), intent(in) :: tauaer(*), ssaaer(*), asmaer(*), ecaer(*)
      real(c_double), intent(out) :: swuflx(*), swdflx(*), swhr(*), swuflxc(*), swdflxc(*), swhrc(*)
      type(c_ptr), value :: stream
    end function
    integer(c_int) function c_lw_cloud(ncol, nlay, icld, idrv, play, plev, tlay, tlev, tsfc, &
                                       h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, &
                                       cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, &
                                       inflglw, iceflglw, liqflglw, cldfr, taucld, cicewp, cliqwp, &
                                       reice, reliq, tauaer, &
                                       uflx, dflx, hr, uflxc, dflxc, hrc, ptr_kind, stream) &
                                       bind(C, name="rrtmg_hip_lw_cloud")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: ncol, nlay, icld, idrv, inflglw, iceflglw, liqflglw, ptr_kind
      real(c_double), intent(in) :: play(*), plev(*), tlay(*), tlev(*), tsfc(*)
      real(c_double), intent(in) :: h2ovmr(*), o3vmr(*), co2vmr(*), ch4vmr(*), n2ovmr(*), o2vmr(*)
      real(c_double), intent(in) :: cfc11vmr(*), cfc12vmr(*), cfc22vmr(*), ccl4vmr(*), emis(*)
      real(c_double), intent(in) :: cldfr(*), taucld(*), cicewp(*), cliqwp(*), reice(*), reliq(*), tauaer(*)
      real(c_double), intent(out) :: uflx(*), dflx(*), hr(*), uflxc(*), dflxc(*), hrc(*)
      type(c_ptr), value :: stream
    end function
    integer(c_int) function c_lw_vjp(ncol, nlay, play, plev, tlay, tlev, tsfc, &
                                     h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, &
                                     cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, tauaer, &
                                     uflx_bar, dflx_bar, hr_bar, tlay_bar, tlev_bar, h2ovmr_bar, &
                                     o3vmr_bar, tsfc_bar, ptr_kind, stream) bind(C, name="rrtmg_hip_lw_vjp")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: ncol, nlay, ptr_kind
      real(c_double), intent(in) :: play(*), plev(*), tlay(*), tlev(*), tsfc(*)
      real(c_double), intent(in) :: h2ovmr(*), o3vmr(*), co2vmr(*), ch4vmr(*), n2ovmr(*), o2vmr(*)
      real(c_double), intent(in) :: cfc11vmr(*), cfc12vmr(*), cfc22vmr(*), ccl4vmr(*), emis(*)
      type(c_ptr), value :: tauaer
      real(c_double), intent(in) :: uflx_bar(*), dflx_bar(*), hr_bar(*)
      real(c_double), intent(out) :: tlay_bar(*), tlev_bar(*), h2ovmr_bar(*), o3vmr_bar(*), tsfc_bar(*)
      type(c_ptr), value :: stream
    end function
    integer(c_int) function c_sw_vjp(ncol, nlay, play, plev, tlay, tlev, tsfc, &
                                     h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, &
                                     asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, &
                                     swuflx_bar, swdflx_bar, swhr_bar, tlay_bar, h2ovmr_bar, o3vmr_bar, &
                                     ptr_kind, stream) bind(C, name="rrtmg_hip_sw_vjp")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: ncol, nlay, dyofyr, ptr_kind
      real(c_double), intent(in) :: play(*), plev(*), tlay(*), tlev(*), tsfc(*)
      real(c_double), intent(in) :: h2ovmr(*), o3vmr(*), co2vmr(*), ch4vmr(*), n2ovmr(*), o2vmr(*)
      real(c_double), intent(in) :: asdir(*), asdif(*), aldir(*), aldif(*), coszen(*)
      real(c_double), value :: adjes, scon
      real(c_double), intent(in) :: swuflx_bar(*), swdflx_bar(*), swhr_bar(*)
      real(c_double), intent(out) :: tlay_bar(*), h2ovmr_bar(*), o3vmr_bar(*)
      type(c_ptr), value :: stream
    end function
  end interface
contains
  subroutine check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == 0) return
    call c_f_pointer(c_last_error(), msg, [512])
    text = ' '
    do i = 1, 512
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    write(*, '(a,a,a,i0,a,a)') 'FATAL from ', where, ': rrtmg_hip error ', rc, ': ', trim(text)
    error stop 1
  end subroutine check
end module rrtmg_hip_capi

module rrtmg_lw_init
  use parkind, only : rb => kind_rb
  implicit none
  public :: rrtmg_lw_ini
contains
  subroutine rrtmg_lw_ini(cpdair)
    use iso_c_binding
    use rrtmg_hip_capi
    real(kind=rb), intent(in) :: cpdair
    call check(c_lw_init(real(cpdair, c_double), c_null_ptr), 'rrtmg_lw_ini')
  end subroutine rrtmg_lw_ini
end module rrtmg_lw_init

module rrtmg_sw_init
  use parkind, only : rb => kind_rb
  implicit none
  public :: rrtmg_sw_ini
contains
  subroutine rrtmg_sw_ini(cpdair)
    use iso_c_binding
    use rrtmg_hip_capi
    real(kind=rb), intent(in) :: cpdair
    call check(c_sw_init(real(cpdair, c_double), c_null_ptr), 'rrtmg_sw_ini')
  end subroutine rrtmg_sw_ini
end module rrtmg_sw_init

module rrtmg_lw_rad
  use parkind, only : im => kind_im, rb => kind_rb
  implicit none
  public :: rrtmg_lw
contains
  subroutine rrtmg_lw(ncol, nlay, icld, idrv, &
                      play, plev, tlay, tlev, tsfc, &
                      h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, &
                      cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, &
                      inflglw, iceflglw, liqflglw, cldfr, &
                      taucld, cicewp, cliqwp, reice, reliq, &
                      tauaer, &
                      uflx, dflx, hr, uflxc, dflxc, hrc, &
                      duflx_dt, duflxc_dt)
    use iso_c_binding
    use rrtmg_hip_capi
    integer(kind=im), intent(in) :: ncol, nlay
    integer(kind=im), intent(inout) :: icld
    integer(kind=im), intent(in) :: idrv
    real(kind=rb), intent(in) :: play(:,:), plev(:,:), tlay(:,:), tlev(:,:), tsfc(:)
    real(kind=rb), intent(in) :: h2ovmr(:,:), o3vmr(:,:), co2vmr(:,:), ch4vmr(:,:), n2ovmr(:,:), o2vmr(:,:)
    real(kind=rb), intent(in) :: cfc11vmr(:,:), cfc12vmr(:,:), cfc22vmr(:,:), ccl4vmr(:,:), emis(:,:)
    integer(kind=im), intent(in) :: inflglw, iceflglw, liqflglw
    real(kind=rb), intent(in) :: cldfr(:,:), cicewp(:,:), cliqwp(:,:), reice(:,:), reliq(:,:)
    real(kind=rb), intent(in) :: taucld(:,:,:)
    real(kind=rb), intent(in) :: tauaer(:,:,:)
    real(kind=rb), intent(out) :: uflx(:,:), dflx(:,:), hr(:,:), uflxc(:,:), dflxc(:,:), hrc(:,:)
    real(kind=rb), intent(out), optional :: duflx_dt(:,:), duflxc_dt(:,:)
    ! out-of-range icld becomes 2 (rrtmg_lw_rad.nomcica.f90:437); icld = 0 is
    ! clear sky, icld = 1..3 with clouds (inflglw = 0: in-cloud optical depths)
    if (icld < 0 .or. icld > 3) icld = 2
    if (icld == 0) then
      call check(c_lw(ncol, nlay, icld, idrv, play, plev, tlay, tlev, tsfc, &
                      h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, &
                      cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, tauaer, &
                      uflx, dflx, hr, uflxc, dflxc, hrc, RRTMG_HIP_HOST, c_null_ptr), 'rrtmg_lw')
    else
      call check(c_lw_cloud(ncol, nlay, icld, idrv, play, plev, tlay, tlev, tsfc, &
                            h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, &
                            cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, &
                            inflglw, iceflglw, liqflglw, cldfr, taucld, cicewp, cliqwp, &
                            reice, reliq, tauaer, &
                            uflx, dflx, hr, uflxc, dflxc, hrc, RRTMG_HIP_HOST, c_null_ptr), 'rrtmg_lw')
    end if
    if (present(duflx_dt)) duflx_dt = 0.0_rb
    if (present(duflxc_dt)) duflxc_dt = 0.0_rb
  end subroutine rrtmg_lw
end module rrtmg_lw_rad

module rrtmg_sw_rad
  use parkind, only : im => kind_im, rb => kind_rb
  implicit none
  public :: rrtmg_sw
contains
  subroutine rrtmg_sw(ncol, nlay, icld, iaer, &
                      play, plev, tlay, tlev, tsfc, &
                      h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, &
                      asdir, asdif, aldir, aldif, &
                      coszen, adjes, dyofyr, scon, &
                      inflgsw, iceflgsw, liqflgsw, cldfr, &
                      taucld, ssacld, asmcld, fsfcld, &
                      cicewp, cliqwp, reice, reliq, &
                      tauaer, ssaaer, asmaer, ecaer, &
                      swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc)
    use iso_c_binding
    use rrtmg_hip_capi
    integer(kind=im), intent(in) :: ncol, nlay
    integer(kind=im), intent(inout) :: icld, iaer
    real(kind=rb), intent(in) :: play(:,:), plev(:,:), tlay(:,:), tlev(:,:), tsfc(:)
    real(kind=rb), intent(in) :: h2ovmr(:,:), o3vmr(:,:), co2vmr(:,:), ch4vmr(:,:), n2ovmr(:,:), o2vmr(:,:)
    real(kind=rb), intent(in) :: asdir(:), asdif(:), aldir(:), aldif(:), coszen(:)
    real(kind=rb), intent(in) :: adjes, scon
    integer(kind=im), intent(in) :: dyofyr
    integer(kind=im), intent(in) :: inflgsw, iceflgsw, liqflgsw
    real(kind=rb), intent(in) :: cldfr(:,:), cicewp(:,:), cliqwp(:,:), reice(:,:), reliq(:,:)
    real(kind=rb), intent(in) :: taucld(:,:,:), ssacld(:,:,:), asmcld(:,:,:), fsfcld(:,:,:)
    real(kind=rb), intent(in) :: tauaer(:,:,:), ssaaer(:,:,:), asmaer(:,:,:), ecaer(:,:,:)
    real(kind=rb), intent(out) :: swuflx(:,:), swdflx(:,:), swhr(:,:), swuflxc(:,:), swdflxc(:,:), swhrc(:,:)
    ! flag normalisation as in rrtmg_sw_rad.nomcica.f90:471,479; icld = 0 is
    ! clear sky, icld >= 1 with clouds (inflgsw = 0: the clouds' optics)
    if (icld < 0 .or. icld > 3) icld = 2
    if (iaer /= 0 .and. iaer /= 6 .and. iaer /= 10) iaer = 0
    if (icld == 0) then
      call check(c_sw(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, &
                      h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, &
                      asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, &
                      swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc, RRTMG_HIP_HOST, c_null_ptr), 'rrtmg_sw')
    else
      call check(c_sw_cloud(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, &
                            h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, &
                            asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, &
                            inflgsw, iceflgsw, liqflgsw, cldfr, taucld, ssacld, asmcld, fsfcld, &
                            cicewp, cliqwp, reice, reliq, tauaer, ssaaer, asmaer, ecaer, &
                            swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc, RRTMG_HIP_HOST, c_null_ptr), &
                 'rrtmg_sw')
    end if
  end subroutine rrtmg_sw
end module rrtmg_sw_rad

! Adjoint entry points (no counterpart in the reference): the backward of
! rrtmg_lw / rrtmg_sw, for Fortran callers of the same arrays.  Given the
! cotangents of the outputs (uflx, dflx, hr or swuflx, swdflx, swhr), they
! return the gradients of sum(cotangent * output) with respect to the layer
! temperature, water vapour and ozone (and, for LW, level and surface
! temperature).  Clear sky without aerosols, as MiMA calls rrtmg_lw/rrtmg_sw.
module rrtmg_adjoint
  use parkind, only : im => kind_im, rb => kind_rb
  use parrrtm, only : nbndlw
  implicit none
contains
  subroutine rrtmg_lw_adj(ncol, nlay, play, plev, tlay, tlev, tsfc, &
                          h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, &
                          cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, &
                          uflx_bar, dflx_bar, hr_bar, &
                          tlay_bar, tlev_bar, h2ovmr_bar, o3vmr_bar, tsfc_bar)
    use iso_c_binding
    use rrtmg_hip_capi
    integer(kind=im), intent(in) :: ncol, nlay
    real(kind=rb), intent(in) :: play(:,:), plev(:,:), tlay(:,:), tlev(:,:), tsfc(:)
    real(kind=rb), intent(in) :: h2ovmr(:,:), o3vmr(:,:), co2vmr(:,:), ch4vmr(:,:), n2ovmr(:,:), o2vmr(:,:)
    real(kind=rb), intent(in) :: cfc11vmr(:,:), cfc12vmr(:,:), cfc22vmr(:,:), ccl4vmr(:,:), emis(:,:)
    real(kind=rb), intent(in) :: uflx_bar(:,:), dflx_bar(:,:), hr_bar(:,:)
    real(kind=rb), intent(out) :: tlay_bar(:,:), tlev_bar(:,:), h2ovmr_bar(:,:), o3vmr_bar(:,:), tsfc_bar(:)
    call check(c_lw_vjp(ncol, nlay, play, plev, tlay, tlev, tsfc, &
                        h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, &
                        cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, c_null_ptr, &
                        uflx_bar, dflx_bar, hr_bar, tlay_bar, tlev_bar, h2ovmr_bar, o3vmr_bar, tsfc_bar, &
                        RRTMG_HIP_HOST, c_null_ptr), 'rrtmg_lw_adj')
  end subroutine rrtmg_lw_adj

  subroutine rrtmg_sw_adj(ncol, nlay, play, plev, tlay, tlev, tsfc, &
                          h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, &
                          asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, &
                          swuflx_bar, swdflx_bar, swhr_bar, &
                          tlay_bar, h2ovmr_bar, o3vmr_bar)
    use iso_c_binding
    use rrtmg_hip_capi
    integer(kind=im), intent(in) :: ncol, nlay, dyofyr
    real(kind=rb), intent(in) :: play(:,:), plev(:,:), tlay(:,:), tlev(:,:), tsfc(:)
    real(kind=rb), intent(in) :: h2ovmr(:,:), o3vmr(:,:), co2vmr(:,:), ch4vmr(:,:), n2ovmr(:,:), o2vmr(:,:)
    real(kind=rb), intent(in) :: asdir(:), asdif(:), aldir(:), aldif(:), coszen(:)
    real(kind=rb), intent(in) :: adjes, scon
    real(kind=rb), intent(in) :: swuflx_bar(:,:), swdflx_bar(:,:), swhr_bar(:,:)
    real(kind=rb), intent(out) :: tlay_bar(:,:), h2ovmr_bar(:,:), o3vmr_bar(:,:)
    call check(c_sw_vjp(ncol, nlay, play, plev, tlay, tlev, tsfc, &
                        h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, &
                        asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, &
                        swuflx_bar, swdflx_bar, swhr_bar, tlay_bar, h2ovmr_bar, o3vmr_bar, &
                        RRTMG_HIP_HOST, c_null_ptr), 'rrtmg_sw_adj')
  end subroutine rrtmg_sw_adj
end module rrtmg_adjoint
