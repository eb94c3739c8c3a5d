! Test driver of the Fortran drop-in: calls rrtmg_sw_ini/rrtmg_lw_ini and
! rrtmg_sw/rrtmg_lw the way MiMA's run_rrtmg does (RR/rrtm_radiation.f90:652-749):
! surface-first arrays built by reversing top-first GCM arrays (non-contiguous
! actual arguments), expression temporaries such as 10*ones for the cloud
! radii and MiMA's zero cloud/aerosol arguments.
!
! usage: shim_driver <input.bin> <output.bin>
!   input  (stream): int32 ncol, nlay; then float64 arrays, top-first
!          (ncol, nlay) or (ncol, nlay+1): play, plev, tlay, tlev; h2o, o3, co2;
!          (ncol): tsfc, albedo, coszen; scalars solrad(adjes), solr_cnst(scon);
!          then the cloud fields of a cloudy LW call, top-first: cldfr
!          (ncol, nlay), taucld (nbndlw, ncol, nlay); then those of a cloudy
!          SW call: cldfr (ncol, nlay), taucld, ssacld, asmcld, fsfcld
!          (nbndsw, ncol, nlay)
!   output (stream): swuflx, swdflx, swhr, uflx, dflx, hr, surface-first; then
!          the adjoint entry points' gradients for cotangents hr_bar = swhr_bar =
!          1 and uflx_bar = dflx_bar = swuflx_bar = swdflx_bar = 0:
!          LW tlay_bar, tlev_bar, h2ovmr_bar, o3vmr_bar, tsfc_bar; SW tlay_bar,
!          h2ovmr_bar, o3vmr_bar; then the cloudy LW call's (icld = 2,
!          inflglw = 0) uflx, dflx, hr, uflxc, dflxc, hrc; then the cloudy
!          SW call's (icld = 2, inflgsw = 0) swuflx, swdflx, swhr, swuflxc,
!          swdflxc, swhrc
program shim_driver
  use parkind, only : im => kind_im, rb => kind_rb
  use parrrtm, only : nbndlw
  use parrrsw, only : nbndsw
  use rrtmg_lw_init, only : rrtmg_lw_ini
  use rrtmg_sw_init, only : rrtmg_sw_ini
  use rrtmg_lw_rad, only : rrtmg_lw
  use rrtmg_sw_rad, only : rrtmg_sw
  use rrtmg_adjoint, only : rrtmg_lw_adj, rrtmg_sw_adj
  implicit none
  character(len=512) :: fin, fout
  integer :: u, ncol, nlay, sk
  integer(kind=im) :: icld, iaer, idrv, dyofyr
  real(kind=rb), allocatable :: pf(:,:), ph(:,:), tf(:,:), th(:,:), q(:,:), o3(:,:), co2(:,:)
  real(kind=rb), allocatable :: tsrf(:), albedo(:), cosz(:), zeros(:,:), ones(:,:), emis(:,:)
  real(kind=rb), allocatable :: taucld(:,:,:), sw_zro(:,:,:), tauaer(:,:,:), zro_sw(:,:,:)
  real(kind=rb), allocatable :: swuflx(:,:), swdflx(:,:), swhr(:,:), swuflxc(:,:), swdflxc(:,:), swhrc(:,:)
  real(kind=rb), allocatable :: uflx(:,:), dflx(:,:), hr(:,:), uflxc(:,:), dflxc(:,:), hrc(:,:)
  real(kind=rb), allocatable :: zl(:,:), gtl(:,:), gtv(:,:), gq(:,:), go(:,:), gts(:), sgt(:,:), sgq(:,:), sgo(:,:)
  real(kind=rb), allocatable :: sf(:,:), sh(:,:), sq(:,:), so(:,:), tcopy(:)
  real(kind=rb), allocatable :: cfr(:,:), tcl(:,:,:)
  real(kind=rb), allocatable :: scf(:,:), stau(:,:,:), sssa(:,:,:), sasm(:,:,:), sfsf(:,:,:)
  real(kind=rb) :: solrad, solr_cnst
  real(kind=rb), parameter :: cp_air = 287.04_rb / (2._rb / 7._rb)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) ncol, nlay
  sk = nlay
  allocate(pf(ncol, nlay), ph(ncol, nlay + 1), tf(ncol, nlay), th(ncol, nlay + 1))
  allocate(q(ncol, nlay), o3(ncol, nlay), co2(ncol, nlay), tsrf(ncol), albedo(ncol), cosz(ncol))
  read(u) pf, ph, tf, th, q, o3, co2, tsrf, albedo, cosz, solrad, solr_cnst
  allocate(cfr(ncol, nlay), tcl(nbndlw, ncol, nlay))
  read(u) cfr, tcl
  allocate(scf(ncol, nlay), stau(nbndsw, ncol, nlay), sssa(nbndsw, ncol, nlay))
  allocate(sasm(nbndsw, ncol, nlay), sfsf(nbndsw, ncol, nlay))
  read(u) scf, stau, sssa, sasm, sfsf
  close(u)

  allocate(zeros(ncol, nlay), ones(ncol, nlay), emis(ncol, nbndlw))
  allocate(taucld(nbndlw, ncol, nlay), sw_zro(nbndsw, ncol, nlay))
  allocate(tauaer(ncol, nlay, nbndlw), zro_sw(ncol, nlay, nbndsw))
  zeros = 0._rb
  ones = 1._rb
  emis = 1._rb
  taucld = 0._rb
  sw_zro = 0._rb
  tauaer = 0._rb
  zro_sw = 0._rb
  allocate(swuflx(ncol, nlay + 1), swdflx(ncol, nlay + 1), swhr(ncol, nlay))
  allocate(swuflxc(ncol, nlay + 1), swdflxc(ncol, nlay + 1), swhrc(ncol, nlay))
  allocate(uflx(ncol, nlay + 1), dflx(ncol, nlay + 1), hr(ncol, nlay))
  allocate(uflxc(ncol, nlay + 1), dflxc(ncol, nlay + 1), hrc(ncol, nlay))

  call rrtmg_lw_ini(cp_air)
  call rrtmg_sw_ini(cp_air)
  icld = 0
  iaer = 0
  idrv = 0
  dyofyr = 0
  ! surface-first views of the top-first state: non-contiguous actuals
  call rrtmg_sw(ncol, nlay, icld, iaer, &
                pf(:, sk:1:-1), ph(:, sk+1:1:-1), tf(:, sk:1:-1), th(:, sk+1:1:-1), tsrf, &
                q(:, sk:1:-1), o3(:, sk:1:-1), co2(:, sk:1:-1), zeros, zeros, zeros, &
                albedo, albedo, albedo, albedo, &
                cosz, solrad, dyofyr, solr_cnst, &
                0, 0, 0, &
                zeros, taucld(1:nbndsw,:,:), sw_zro, sw_zro, sw_zro, &
                zeros, zeros, 10*ones, 10*ones, &
                zro_sw, zro_sw, zro_sw, zro_sw, &
                swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc)
  call rrtmg_lw(ncol, nlay, icld, idrv, &
                pf(:, sk:1:-1), ph(:, sk+1:1:-1), tf(:, sk:1:-1), th(:, sk+1:1:-1), tsrf, &
                q(:, sk:1:-1), o3(:, sk:1:-1), co2(:, sk:1:-1), &
                zeros, zeros, zeros, &
                zeros, zeros, zeros, zeros, &
                emis, 0, 0, 0, &
                zeros, taucld, zeros, zeros, 10*ones, 10*ones, &
                tauaer, &
                uflx, dflx, hr, uflxc, dflxc, hrc)

  ! adjoint entry points on the same surface-first state (contiguous copies)
  allocate(zl(ncol, nlay + 1), gtl(ncol, nlay), gtv(ncol, nlay + 1), gq(ncol, nlay), go(ncol, nlay), gts(ncol))
  allocate(sgt(ncol, nlay), sgq(ncol, nlay), sgo(ncol, nlay))
  allocate(sf(ncol, nlay), sh(ncol, nlay + 1), sq(ncol, nlay), so(ncol, nlay), tcopy(ncol))
  zl = 0._rb
  sf = tf(:, sk:1:-1)
  sh = th(:, sk+1:1:-1)
  sq = q(:, sk:1:-1)
  so = o3(:, sk:1:-1)
  tcopy = tsrf
  call rrtmg_lw_adj(ncol, nlay, pf(:, sk:1:-1), ph(:, sk+1:1:-1), sf, sh, tcopy, &
                    sq, so, co2(:, sk:1:-1), zeros, zeros, zeros, zeros, zeros, zeros, zeros, emis, &
                    zl, zl, ones, gtl, gtv, gq, go, gts)
  call rrtmg_sw_adj(ncol, nlay, pf(:, sk:1:-1), ph(:, sk+1:1:-1), sf, sh, tcopy, &
                    sq, so, co2(:, sk:1:-1), zeros, zeros, zeros, &
                    albedo, albedo, albedo, albedo, cosz, solrad, dyofyr, solr_cnst, &
                    zl, zl, ones, sgt, sgq, sgo)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) swuflx, swdflx, swhr, uflx, dflx, hr
  write(u) gtl, gtv, gq, go, gts, sgt, sgq, sgo

  ! LW with clouds: maximum-random overlap, in-cloud optical depths
  icld = 2
  call rrtmg_lw(ncol, nlay, icld, idrv, &
                pf(:, sk:1:-1), ph(:, sk+1:1:-1), tf(:, sk:1:-1), th(:, sk+1:1:-1), tsrf, &
                q(:, sk:1:-1), o3(:, sk:1:-1), co2(:, sk:1:-1), &
                zeros, zeros, zeros, &
                zeros, zeros, zeros, zeros, &
                emis, 0, 0, 0, &
                cfr(:, sk:1:-1), tcl(:, :, sk:1:-1), zeros, zeros, 10*ones, 10*ones, &
                tauaer, &
                uflx, dflx, hr, uflxc, dflxc, hrc)
  write(u) uflx, dflx, hr, uflxc, dflxc, hrc

  ! SW with clouds: overcast layers, the clouds' optics (inflgsw = 0)
  icld = 2
  call rrtmg_sw(ncol, nlay, icld, iaer, &
                pf(:, sk:1:-1), ph(:, sk+1:1:-1), tf(:, sk:1:-1), th(:, sk+1:1:-1), tsrf, &
                q(:, sk:1:-1), o3(:, sk:1:-1), co2(:, sk:1:-1), zeros, zeros, zeros, &
                albedo, albedo, albedo, albedo, &
                cosz, solrad, dyofyr, solr_cnst, &
                0, 0, 0, &
                scf(:, sk:1:-1), stau(:, :, sk:1:-1), sssa(:, :, sk:1:-1), sasm(:, :, sk:1:-1), sfsf(:, :, sk:1:-1), &
                zeros, zeros, 10*ones, 10*ones, &
                zro_sw, zro_sw, zro_sw, zro_sw, &
                swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc)
  write(u) swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc
  close(u)
end program shim_driver
