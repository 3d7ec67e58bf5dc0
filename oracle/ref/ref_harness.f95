! Test driver for the reference CPU physics.  It is compiled against the reference's own
! Fortran sources where they lie (src/Fortran/*.f95; see oracle/ref/Makefile) and calls the
! reference subroutines in the order of the reference time loop (main_ocl.f95:264-299, with
! EXTERNAL_WIND_PROFILE and IFBF=1).  Inputs come from a raw stream file written by
! tests/golden/make_ref_vectors.py; outputs (full fields) go to a raw stream file.
!
! Modes
!   0  trajectory: init coefficients (init.f95, built without the GIS read), supplied masks,
!      the ifdata.f95:250-273 sequence with bondv1_ext_wind_profile, then nsteps time steps;
!      the state is dumped after each requested step (0 = after the init sequence).
!   1  single stage: the full state is read from the input, one reference stage runs, the
!      state is dumped.
!   2  terrain masks: feedbfm.f95 with its own GIS read (run with cwd = src/Fortran), zbm and
!      the four masks are dumped.
!   1  (stage 13) anime.f95 on the state read, with n = n0 = nmax (so it writes), into the
!      files anime22.bin and anime23.bin of the working directory; dxl and dyl are the running
!      sums of dx1 and dy1 (grid.f95:26-38).
!   3  statistics: as mode 0 (no state dumps), with aveflow.f95 called after every step
!      (n1 from the input, nmax = nsteps); the eighteen statistics arrays are dumped at the end.
!      aveu and avew are passed as sections of buffers that start with ip*jp zeros, so the
!      reads of aveflow.f95:105-106 at level k-1 = 0 (before the arrays) see zeros.

subroutine noop
  ! press.f95:116 calls noop, which the reference leaves to an absent noop.c
  ! (SConstruct:125-126); it is an empty call target after the SOR loop's exit label.
end subroutine noop

program ref_harness
  use common_sn
  use module_velnw
  use module_bondv1_ext_wind_profile
  use module_velFG
  use module_vel2
  use module_feedbf
  use module_les
  use module_boundsm
  use module_adam
  use module_press
  use module_bondFG
  use module_boundp
  use module_init
  use module_feedbfm
  use module_aveflow
  use module_anime
  implicit none

  real(kind=4), dimension(0:ip+2,0:jp+2,0:kp+1) :: p
  real(kind=4), dimension(0:ip+1,-1:jp+1,0:kp+1) :: u, v
  real(kind=4), dimension(0:ip+1,-1:jp+1,-1:kp+1) :: w
  real(kind=4), dimension(0:ip,0:jp,0:kp) :: usum, vsum, wsum, f, g, h, fx, fy, fz
  real(kind=4), dimension(ip,jp,kp) :: fold, gold, hold, fghold, cn1
  real(kind=4), dimension(0:ip+1,0:jp+1,0:kp+1) :: rhs
  real(kind=4), dimension(-1:ip+1,-1:jp+1,0:kp+1) :: sm
  real(kind=4), dimension(-1:ip+2,0:jp+2,0:kp+2) :: diu1, diu5, cov1, cov5, nou1, nou5
  real(kind=4), dimension(0:ip+2,0:jp+2,0:kp+2) :: diu2, diu3, diu4, diu6, diu7, diu8, diu9
  real(kind=4), dimension(0:ip+2,0:jp+2,0:kp+2) :: cov2, cov3, cov4, cov6, cov7, cov8, cov9
  real(kind=4), dimension(0:ip+2,0:jp+2,0:kp+2) :: nou2, nou3, nou4, nou6, nou7, nou8, nou9
  real(kind=4), dimension(0:ip,jp,kp) :: dfu1
  real(kind=4), dimension(ip,0:jp,kp) :: dfv1
  real(kind=4), dimension(ip,jp,kp) :: dfw1
  real(kind=4), dimension(0:ip+1,0:jp+1,0:kp+1) :: amask1, dmask1
  real(kind=4), dimension(-1:ip+1,0:jp+1,0:kp+1) :: bmask1
  real(kind=4), dimension(0:ip+1,-1:jp+1,0:kp+1) :: cmask1
  real(kind=4), dimension(-1:ip+1) :: dx1
  real(kind=4), dimension(0:ip) :: dxs
  real(kind=4), dimension(0:jp+1) :: dy1
  real(kind=4), dimension(0:jp) :: dys
  real(kind=4), dimension(-1:kp+2) :: dzn, dzs
  real(kind=4), dimension(kp+2) :: z2
  real(kind=4), dimension(ip) :: cn2l, cn2s
  real(kind=4), dimension(jp) :: cn3l, cn3s
  real(kind=4), dimension(kp) :: cn4l, cn4s, delx1
  real(kind=4), dimension(jp,kp) :: wind_profile
  real(kind=4), dimension(-1:ipmax+1,-1:jpmax+1) :: zbm
  ! mode 3 only; on the heap so that the large grids keep their static arrays below 2 GiB
  real(kind=4), allocatable :: aveubuf(:), avewbuf(:)
  real(kind=4), allocatable, dimension(:,:,:) :: avev, avep, avel, aveuu, avevv, aveww, avesm, &
                                                 avesmsm, uwfx
  real(kind=4), allocatable, dimension(:,:) :: avesu, avesv, avesw, avesuu, avesvv, avesww, uwfxs
  integer :: n1
  real(kind=4) :: time
  character(len=70) :: data10, data11
  real(kind=4) :: dt, ro, vn, alpha, beta
  integer :: mode, nip, njp, nkp, nsteps, ndump, stage, nstage, n, idump, im, jm, km
  integer, allocatable :: dump_steps(:)
  integer(kind=8) :: clk0, clk1, clkrate
  character(len=512) :: fin, fout
  character(len=70) :: data20, data21, data22, data23
  real(kind=4), dimension(0:ip) :: dxl
  real(kind=4), dimension(0:jp) :: dyl
  integer :: ii

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  read(10) mode, nip, njp, nkp
  if (nip /= ip .or. njp /= jp .or. nkp /= kp) then
    print *, 'ref_harness: grid mismatch', nip, njp, nkp, ip, jp, kp
    stop 2
  end if
  im = ip; jm = jp; km = kp
  data20 = ''; data21 = ''
  fghold = 0.0; fx = 0.0; fy = 0.0; fz = 0.0; rhs = 0.0; sm = 0.0
  fold = 0.0; gold = 0.0; hold = 0.0; usum = 0.0; vsum = 0.0; wsum = 0.0
  amask1 = 0.0; bmask1 = 0.0; cmask1 = 0.0; dmask1 = 0.0; zbm = 0.0

  call zero_work_arrays()
  if (mode == 2) then
    read(10) dzn, z2, zbm
    call feedbfm(km, jm, im, amask1, bmask1, cmask1, dmask1, zbm, z2, dzn)
    write(11) zbm, amask1, bmask1, cmask1, dmask1
    close(11)
    stop
  end if

  read(10) dt, ro, vn, alpha, beta
  read(10) dx1, dy1, dzn, dzs, dxs, dys, z2
  read(10) amask1, bmask1, cmask1, dmask1, wind_profile

  if (mode == 0 .or. mode == 3) then
    read(10) nsteps, ndump
    allocate(dump_steps(ndump))
    read(10) dump_steps
    if (mode == 3) then
      read(10) n1
      allocate(aveubuf(ip*jp + ip*jp*kp), avewbuf(ip*jp + ip*jp*kp))
      allocate(avev(ip,jp,kp), avep(ip,jp,kp), avel(ip,jp,kp), aveuu(ip,jp,kp), avevv(ip,jp,kp))
      allocate(aveww(ip,jp,kp), avesm(ip,jp,kp), avesmsm(ip,jp,kp), uwfx(ip,jp,kp))
      allocate(avesu(ip,kp), avesv(ip,kp), avesw(ip,kp), avesuu(ip,kp), avesvv(ip,kp))
      allocate(avesww(ip,kp), uwfxs(ip,kp))
      aveubuf = 0.0; avewbuf = 0.0; uwfxs = 0.0; time = 0.0
      data10 = ''; data11 = ''
    end if
    call init(km, jm, im, u, v, w, p, cn2s, dxs, cn2l, cn3s, dys, cn3l, dzs, cn4s, cn4l, cn1, &
              amask1, bmask1, cmask1, dmask1, zbm, z2, dzn)
    ! init.f95 is built with IFBF=0 (no GIS read): the masks come from the input file
    rewind(10)
    call skip_to_masks()
    read(10) amask1, bmask1, cmask1, dmask1, wind_profile
    ! ifdata.f95:250-256 (zero_arrays); ifdata.f95 itself is not built because it uses
    ! bondv1.f95, whose fixed inflow levels k=1..78 do not compile for kp < 77
    call zero_work_arrays()
    n = 1
    ! ifdata.f95:258-273
    call bondv1_ext_wind_profile(jm, u, z2, dzn, v, w, km, n, im, dt, dxs, wind_profile)
    call boundp1(km, jm, p, im)
    call boundp2(jm, im, p, km)
    call velfg(km, jm, im, dx1, cov1, cov2, cov3, dfu1, diu1, diu2, dy1, diu3, dzn, vn, f, cov4, &
               cov5, cov6, dfv1, diu4, diu5, diu6, g, cov7, cov8, cov9, dfw1, diu7, diu8, diu9, &
               dzs, h, nou1, u, nou5, v, nou9, w, nou2, nou3, nou4, nou6, nou7, nou8)
    call feedbf(km, jm, im, usum, u, bmask1, vsum, v, cmask1, wsum, w, dmask1, alpha, dt, beta, &
                fx, fy, fz, f, g, h)
    call les(km, delx1, dx1, dy1, dzn, jm, im, diu1, diu2, diu3, diu4, diu5, diu6, diu7, diu8, &
             diu9, sm, f, g, h)
    do idump = 1, ndump
      if (dump_steps(idump) == 0) call dump_state()
    end do
    call system_clock(clk0, clkrate)
    do n = 1, nsteps
      call one_step(n)
      if (mode == 3) then
        call aveflow(n, n1, km, jm, im, aveubuf(ip*jp+1), avev, avewbuf(ip*jp+1), avep, avel, &
                     aveuu, avevv, aveww, avesm, avesmsm, uwfx, avesu, avesv, avesw, avesuu, &
                     avesvv, avesww, u, v, w, p, sm, nsteps, uwfxs, data10, time, data11)
      end if
      do idump = 1, ndump
        if (dump_steps(idump) == n) call dump_state()
      end do
    end do
    call system_clock(clk1)
    print '(A,F12.6)', 'steps_seconds=', real(clk1 - clk0, 8) / real(clkrate, 8)
    if (mode == 3) then
      write(11) aveubuf(ip*jp+1:), avev, avewbuf(ip*jp+1:), avep, avel, aveuu, avevv, aveww, &
                avesm, avesmsm, uwfx, avesu, avesv, avesw, avesuu, avesvv, avesww, uwfxs
    end if
  else if (mode == 1) then
    read(10) stage, nstage
    call read_state()
    n = nstage
    select case (stage)
    case (1)
      call velnw(km, jm, im, p, ro, dxs, u, dt, f, dys, v, g, dzs, w, h)
    case (2)
      call bondv1_ext_wind_profile(jm, u, z2, dzn, v, w, km, n, im, dt, dxs, wind_profile)
    case (3)
      call velfg(km, jm, im, dx1, cov1, cov2, cov3, dfu1, diu1, diu2, dy1, diu3, dzn, vn, f, cov4, &
                 cov5, cov6, dfv1, diu4, diu5, diu6, g, cov7, cov8, cov9, dfw1, diu7, diu8, diu9, &
                 dzs, h, nou1, u, nou5, v, nou9, w, nou2, nou3, nou4, nou6, nou7, nou8)
    case (4)
      call feedbf(km, jm, im, usum, u, bmask1, vsum, v, cmask1, wsum, w, dmask1, alpha, dt, beta, &
                  fx, fy, fz, f, g, h)
    case (5)
      call les(km, delx1, dx1, dy1, dzn, jm, im, diu1, diu2, diu3, diu4, diu5, diu6, diu7, diu8, &
               diu9, sm, f, g, h)
    case (6)
      call adam(n, n + 1, data21, fold, im, jm, km, gold, hold, fghold, f, g, h)
    case (7)
      call press(km, jm, im, rhs, u, dx1, v, dy1, w, dzn, f, g, h, dt, cn1, cn2l, p, cn2s, cn3l, &
                 cn3s, cn4l, cn4s, n, n + 1, data20, usum, vsum, wsum)
    case (8)
      call boundp1(km, jm, p, im)
    case (9)
      call boundp2(jm, im, p, km)
    case (10)
      call bondfg(km, jm, f, im, g, h)
    case (11)
      call boundsm(km, jm, sm, im)
    case (12)
      call one_step(n)
    case (13)
      dxl(0) = 0.
      do ii = 1, ip
        dxl(ii) = dxl(ii-1) + dx1(ii)
      end do
      dyl(0) = 0.
      do ii = 1, jp
        dyl(ii) = dyl(ii-1) + dy1(ii)
      end do
      data22 = 'anime22.bin'
      data23 = 'anime23.bin'
      call anime(n, n, n, km, jm, im, dxl, dx1, dyl, dy1, z2, data22, data23, u, w, v, amask1)
    case default
      print *, 'ref_harness: unknown stage', stage
      stop 3
    end select
    call dump_state()
  end if
  close(11)

contains

  subroutine zero_work_arrays()
    cov1 = 0.0; cov2 = 0.0; cov3 = 0.0; cov4 = 0.0; cov5 = 0.0; cov6 = 0.0; cov7 = 0.0
    cov8 = 0.0; cov9 = 0.0; nou1 = 0.0; nou2 = 0.0; nou3 = 0.0; nou4 = 0.0; nou5 = 0.0
    nou6 = 0.0; nou7 = 0.0; nou8 = 0.0; nou9 = 0.0; dfu1 = 0.0; dfv1 = 0.0; dfw1 = 0.0
    diu1 = 0.0; diu2 = 0.0; diu3 = 0.0; diu4 = 0.0; diu5 = 0.0; diu6 = 0.0; diu7 = 0.0
    diu8 = 0.0; diu9 = 0.0; f = 0.0; g = 0.0; h = 0.0; p = 0.0; u = 0.0; v = 0.0; w = 0.0
  end subroutine zero_work_arrays

  subroutine skip_to_masks()
    integer :: hdr(4)
    real(kind=4) :: sc(5)
    read(10) hdr
    read(10) sc
    read(10) dx1, dy1, dzn, dzs, dxs, dys, z2
  end subroutine skip_to_masks

  ! main_ocl.f95:264-299
  subroutine one_step(nn)
    integer, intent(in) :: nn
    call velnw(km, jm, im, p, ro, dxs, u, dt, f, dys, v, g, dzs, w, h)
    call bondv1_ext_wind_profile(jm, u, z2, dzn, v, w, km, nn, im, dt, dxs, wind_profile)
    call velfg(km, jm, im, dx1, cov1, cov2, cov3, dfu1, diu1, diu2, dy1, diu3, dzn, vn, f, cov4, &
               cov5, cov6, dfv1, diu4, diu5, diu6, g, cov7, cov8, cov9, dfw1, diu7, diu8, diu9, &
               dzs, h, nou1, u, nou5, v, nou9, w, nou2, nou3, nou4, nou6, nou7, nou8)
    call feedbf(km, jm, im, usum, u, bmask1, vsum, v, cmask1, wsum, w, dmask1, alpha, dt, beta, &
                fx, fy, fz, f, g, h)
    call les(km, delx1, dx1, dy1, dzn, jm, im, diu1, diu2, diu3, diu4, diu5, diu6, diu7, diu8, &
             diu9, sm, f, g, h)
    call adam(nn, nsteps_or(nn), data21, fold, im, jm, km, gold, hold, fghold, f, g, h)
    call press(km, jm, im, rhs, u, dx1, v, dy1, w, dzn, f, g, h, dt, cn1, cn2l, p, cn2s, cn3l, &
               cn3s, cn4l, cn4s, nn, nsteps_or(nn), data20, usum, vsum, wsum)
  end subroutine one_step

  integer function nsteps_or(nn)
    integer, intent(in) :: nn
    nsteps_or = max(nn + 1, 2)
  end function nsteps_or

  subroutine read_state()
    read(10) p, u, v, w, usum, vsum, wsum, f, g, h, fold, gold, hold, rhs, sm
    read(10) diu1, diu2, diu3, diu4, diu5, diu6, diu7, diu8, diu9
    read(10) cn1, cn2l, cn2s, cn3l, cn3s, cn4l, cn4s, delx1
  end subroutine read_state

  subroutine dump_state()
    write(11) p, u, v, w, usum, vsum, wsum, f, g, h, fold, gold, hold, rhs, sm
    write(11) diu1, diu2, diu3, diu4, diu5, diu6, diu7, diu8, diu9
    write(11) cn1, cn2l, cn2s, cn3l, cn3s, cn4l, cn4s, delx1
  end subroutine dump_state

end program ref_harness
