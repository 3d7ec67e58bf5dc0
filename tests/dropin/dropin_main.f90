! Test host for the Fortran drop-in (les_halo_amd/fortran/module_les_hip.f90).  It follows the
! reference host program's GPU path, main_ocl.f95:198-253: the state after the init sequence
! (read from a stream file, in the Fortran bounds of main_ocl.f95:40-167), one call of
! initialise_LES_kernel, then run_LES_kernel(n, nmax, wind_profile) for n = 1..nmax; the state
! is then read back and written to a stream file.
!   dropin_main in.bin out.bin
! in.bin:  ip, jp, kp, nmax (int32), dt (float32), then the arrays in initialise_LES_kernel's
!          argument order.  out.bin: p, u, v, w, usum, vsum, wsum, f, g, h, fold, gold, hold,
!          rhs, sm.
program dropin_main
  use iso_c_binding
  use module_LES_hip
  implicit none
  integer(c_int) :: ip, jp, kp, nmax, n
  real(c_float) :: dt
  real(c_float), allocatable :: p(:,:,:), u(:,:,:), v(:,:,:), w(:,:,:)
  real(c_float), allocatable :: usum(:,:,:), vsum(:,:,:), wsum(:,:,:), f(:,:,:), g(:,:,:), h(:,:,:)
  real(c_float), allocatable :: fold(:,:,:), gold(:,:,:), hold(:,:,:), cn1(:,:,:)
  real(c_float), allocatable :: diu1(:,:,:), diu2(:,:,:), diu3(:,:,:), diu4(:,:,:), diu5(:,:,:)
  real(c_float), allocatable :: diu6(:,:,:), diu7(:,:,:), diu8(:,:,:), diu9(:,:,:)
  real(c_float), allocatable :: amask1(:,:,:), bmask1(:,:,:), cmask1(:,:,:), dmask1(:,:,:)
  real(c_float), allocatable :: rhs(:,:,:), sm(:,:,:)
  real(c_float), allocatable :: cn2l(:), cn2s(:), cn3l(:), cn3s(:), cn4l(:), cn4s(:)
  real(c_float), allocatable :: dxs(:), dys(:), dzs(:), dx1(:), dy1(:), dzn(:)
  real(c_float), allocatable :: wind_profile(:,:)
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) ip, jp, kp, nmax, dt
  ! main_ocl.f95:40-167
  allocate(p(0:ip+2,0:jp+2,0:kp+1))
  allocate(u(0:ip+1,-1:jp+1,0:kp+1), v(0:ip+1,-1:jp+1,0:kp+1), w(0:ip+1,-1:jp+1,-1:kp+1))
  allocate(usum(0:ip,0:jp,0:kp), vsum(0:ip,0:jp,0:kp), wsum(0:ip,0:jp,0:kp))
  allocate(f(0:ip,0:jp,0:kp), g(0:ip,0:jp,0:kp), h(0:ip,0:jp,0:kp))
  allocate(fold(ip,jp,kp), gold(ip,jp,kp), hold(ip,jp,kp), cn1(ip,jp,kp))
  allocate(diu1(-1:ip+2,0:jp+2,0:kp+2), diu5(-1:ip+2,0:jp+2,0:kp+2))
  allocate(diu2(0:ip+2,0:jp+2,0:kp+2), diu3(0:ip+2,0:jp+2,0:kp+2), diu4(0:ip+2,0:jp+2,0:kp+2))
  allocate(diu6(0:ip+2,0:jp+2,0:kp+2), diu7(0:ip+2,0:jp+2,0:kp+2), diu8(0:ip+2,0:jp+2,0:kp+2))
  allocate(diu9(0:ip+2,0:jp+2,0:kp+2))
  allocate(amask1(0:ip+1,0:jp+1,0:kp+1), bmask1(-1:ip+1,0:jp+1,0:kp+1))
  allocate(cmask1(0:ip+1,-1:jp+1,0:kp+1), dmask1(0:ip+1,0:jp+1,0:kp+1))
  allocate(rhs(0:ip+1,0:jp+1,0:kp+1), sm(-1:ip+1,-1:jp+1,0:kp+1))
  allocate(cn2l(ip), cn2s(ip), cn3l(jp), cn3s(jp), cn4l(kp), cn4s(kp))
  allocate(dxs(0:ip), dys(0:jp), dzs(-1:kp+2), dx1(-1:ip+1), dy1(0:jp+1), dzn(-1:kp+2))
  allocate(wind_profile(jp,kp))
  read(10) p, u, v, w, usum, vsum, wsum, f, g, h, fold, gold, hold
  read(10) diu1, diu2, diu3, diu4, diu5, diu6, diu7, diu8, diu9
  read(10) amask1, bmask1, cmask1, dmask1, cn1, cn2l, cn2s, cn3l, cn3s, cn4l, cn4s
  read(10) rhs, sm, dxs, dys, dzs, dx1, dy1, dzn, wind_profile
  close(10)

  call initialise_LES_kernel(p,u,v,w,usum,vsum,wsum,f,g,h,fold,gold,hold, &
      diu1,diu2,diu3,diu4,diu5,diu6,diu7,diu8,diu9,amask1,bmask1,cmask1,dmask1, &
      cn1,cn2l,cn2s,cn3l,cn3s,cn4l,cn4s,rhs,sm,dxs,dys,dzs,dx1,dy1,dzn,wind_profile,dt, &
      ip,jp,kp)
  do n = 1, nmax
    call run_LES_kernel(n, nmax, wind_profile)
  end do
  call download_LES_field(LES_P, p)
  call download_LES_field(LES_U, u)
  call download_LES_field(LES_V, v)
  call download_LES_field(LES_W, w)
  call download_LES_field(LES_USUM, usum)
  call download_LES_field(LES_VSUM, vsum)
  call download_LES_field(LES_WSUM, wsum)
  call download_LES_field(LES_F, f)
  call download_LES_field(LES_G, g)
  call download_LES_field(LES_H, h)
  call download_LES_field(LES_FOLD, fold)
  call download_LES_field(LES_GOLD, gold)
  call download_LES_field(LES_HOLD, hold)
  call download_LES_field(LES_RHS, rhs)
  call download_LES_field(LES_SM, sm)
  call finalise_LES_kernel()
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) p, u, v, w, usum, vsum, wsum, f, g, h, fold, gold, hold, rhs, sm
  close(11)
  print '(A,I0,A)', 'dropin_main: ', nmax, ' steps done'
end program dropin_main
