! Writes a data30/data31 pair with the record statements of the reference's restart I/O
! (press.f95:192-199 and adam.f95:24-29 write the data20/data21 arrays this way; ifdata.f95:229-244
! reads data30 = (n, time) then the same seven arrays, and data31 = the four adam arrays).
! Built with the image's Fortran compiler by make_restart_vectors.py; the files it writes are the
! fixtures of tests/test_restart.py.  Values: array q (u, v, w, p, usum, vsum, wsum, fold, gold,
! hold) holds q*1000 + i + 10*j + 100*k; fghold is zero.
program restart_records
  implicit none
  integer, parameter :: im = 6, jm = 4, km = 3
  real(kind=4) :: a(im, jm, km, 11)
  integer :: i, j, k, q, n
  real(kind=4) :: time
  do q = 1, 10
    do k = 1, km
      do j = 1, jm
        do i = 1, im
          a(i, j, k, q) = real(q * 1000 + i + 10 * j + 100 * k)
        end do
      end do
    end do
  end do
  a(:, :, :, 11) = 0.0
  n = 1234
  time = 61.5
  open(unit=30, file='data30.bin', form='unformatted', status='unknown')
  write(30) n, time
  do q = 1, 7
    write(30) (((a(i, j, k, q), i=1,im), j=1,jm), k=1,km)
  end do
  close(30)
  open(unit=31, file='data31.bin', form='unformatted', status='unknown')
  do q = 8, 11
    write(31) (((a(i, j, k, q), i=1,im), j=1,jm), k=1,km)
  end do
  close(31)
end program restart_records
