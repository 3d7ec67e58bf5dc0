! Fortran binding of libles_hip.so for the reference's host program.
!
! It takes the place of the reference's GPU host module (use module_LES_combined_ocl,
! src/OpenCL/Wrappers/module_LES_combined_ocl.f95, generated from
! src/OpenCL/module_LES_combined_ocl_TEMPL_V2.f95): the two subroutines the host loop calls,
! initialise_LES_kernel (TEMPL:18-30, called at main_ocl.f95:198-210) and
! run_LES_kernel(n, nmax, wind_profile) (TEMPL:219-230, called at main_ocl.f95:241-253), keep
! their names and argument lists (external wind profile form) and call the C ABI of
! include/les_hip.h.  The reference's subroutines return no status; these stop with the
! library's message when a call fails.
module module_LES_hip
  use iso_c_binding
  implicit none
  private
  public :: initialise_LES_kernel, run_LES_kernel, finalise_LES_kernel, download_LES_field
  public :: LES_P, LES_U, LES_V, LES_W, LES_USUM, LES_VSUM, LES_WSUM, LES_F, LES_G, LES_H, &
            LES_FOLD, LES_GOLD, LES_HOLD, LES_RHS, LES_SM
  ! field ids of include/les_hip.h (enum les_field)
  integer(c_int), parameter :: LES_P = 0, LES_U = 1, LES_V = 2, LES_W = 3, LES_USUM = 4, &
                               LES_VSUM = 5, LES_WSUM = 6, LES_F = 7, LES_G = 8, LES_H = 9, &
                               LES_FOLD = 10, LES_GOLD = 11, LES_HOLD = 12, LES_RHS = 13, &
                               LES_SM = 14

  interface
    integer(c_int) function c_initialise(p,u,v,w,usum,vsum,wsum,f,g,h, &
        fold,gold,hold,diu1,diu2,diu3,diu4,diu5,diu6,diu7,diu8,diu9, &
        amask1,bmask1,cmask1,dmask1,cn1,cn2l,cn2s,cn3l,cn3s,cn4l,cn4s, &
        rhs,sm,dxs,dys,dzs,dx1,dy1,dzn,wind_profile,dt,im,jm,km) &
        bind(C, name="les_initialise_LES_kernel")
      import :: c_int, c_float
      real(c_float), intent(in) :: p(*),u(*),v(*),w(*),usum(*),vsum(*),wsum(*),f(*),g(*),h(*)
      real(c_float), intent(in) :: fold(*),gold(*),hold(*)
      real(c_float), intent(in) :: diu1(*),diu2(*),diu3(*),diu4(*),diu5(*),diu6(*),diu7(*), &
                                   diu8(*),diu9(*)
      real(c_float), intent(in) :: amask1(*),bmask1(*),cmask1(*),dmask1(*)
      real(c_float), intent(in) :: cn1(*),cn2l(*),cn2s(*),cn3l(*),cn3s(*),cn4l(*),cn4s(*)
      real(c_float), intent(in) :: rhs(*),sm(*),dxs(*),dys(*),dzs(*),dx1(*),dy1(*),dzn(*)
      real(c_float), intent(in) :: wind_profile(*)
      real(c_float), value :: dt
      integer(c_int), value :: im, jm, km
    end function
    integer(c_int) function c_run(n, nmax, wind_profile) bind(C, name="les_run_LES_kernel")
      import :: c_int, c_float
      integer(c_int), value :: n, nmax
      real(c_float), intent(in) :: wind_profile(*)
    end function
    integer(c_int) function c_download(hnd, field, host) bind(C, name="les_download")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: hnd
      integer(c_int), value :: field
      real(c_float), intent(out) :: host(*)
    end function
    type(c_ptr) function c_default_handle() bind(C, name="les_default_handle")
      import :: c_ptr
    end function
    type(c_ptr) function c_last_error(hnd) bind(C, name="les_last_error")
      import :: c_ptr
      type(c_ptr), value :: hnd
    end function
    integer(c_int) function c_finalise() bind(C, name="les_finalise_LES_kernel")
      import :: c_int
    end function
  end interface

contains

  subroutine check(status, what)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: what
    if (status /= 0) then
      print '(A,A,I0)', what, ' failed with status ', status
      stop 1
    end if
  end subroutine check

  ! TEMPL:18-30 / main_ocl.f95:198-210 (EXTERNAL_WIND_PROFILE form)
  subroutine initialise_LES_kernel(p,u,v,w,usum,vsum,wsum,f,g,h,fold,gold,hold, &
      diu1,diu2,diu3,diu4,diu5,diu6,diu7,diu8,diu9,amask1,bmask1,cmask1,dmask1, &
      cn1,cn2l,cn2s,cn3l,cn3s,cn4l,cn4s,rhs,sm,dxs,dys,dzs,dx1,dy1,dzn,wind_profile,dt,im,jm,km)
    real(c_float), dimension(*), intent(in) :: p,u,v,w,usum,vsum,wsum,f,g,h,fold,gold,hold, &
        diu1,diu2,diu3,diu4,diu5,diu6,diu7,diu8,diu9,amask1,bmask1,cmask1,dmask1, &
        cn1,cn2l,cn2s,cn3l,cn3s,cn4l,cn4s,rhs,sm,dxs,dys,dzs,dx1,dy1,dzn,wind_profile
    real(c_float), intent(in) :: dt
    integer, intent(in) :: im, jm, km
    call check(c_initialise(p,u,v,w,usum,vsum,wsum,f,g,h,fold,gold,hold, &
        diu1,diu2,diu3,diu4,diu5,diu6,diu7,diu8,diu9,amask1,bmask1,cmask1,dmask1, &
        cn1,cn2l,cn2s,cn3l,cn3s,cn4l,cn4s,rhs,sm,dxs,dys,dzs,dx1,dy1,dzn,wind_profile, &
        dt,int(im,c_int),int(jm,c_int),int(km,c_int)), 'initialise_LES_kernel')
  end subroutine initialise_LES_kernel

  ! TEMPL:219-230 / main_ocl.f95:241-253: one time step n on the device
  subroutine run_LES_kernel(n, nmax, wind_profile)
    integer, intent(in) :: n, nmax
    real(c_float), dimension(*), intent(in) :: wind_profile
    call check(c_run(int(n,c_int), int(nmax,c_int), wind_profile), 'run_LES_kernel')
  end subroutine run_LES_kernel

  ! Copies one array of the device state (field id above) into the caller's Fortran array
  ! (the reference reads results back with oclRead*ArrayBuffer, TEMPL:1035-1056).
  subroutine download_LES_field(field, host)
    integer(c_int), intent(in) :: field
    real(c_float), dimension(*), intent(out) :: host
    call check(c_download(c_default_handle(), field, host), 'les_download')
  end subroutine download_LES_field

  subroutine finalise_LES_kernel()
    call check(c_finalise(), 'les_finalise_LES_kernel')
  end subroutine finalise_LES_kernel

end module module_LES_hip
