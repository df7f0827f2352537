! Fortran 2003 FFTW caller through iso_c_binding and include/fftw_wrapper.f03
! (bind(C) interfaces of the C entry points): complex and real 1-D / 2-D
! transforms in both precisions on closed-form inputs whose DFT is known
! exactly.  C planners take dimensions in row-major order, so a Fortran
! array a(nx, ny) is planned as (ny, nx).  Prints PASS or FAIL <case>.
program modern
  use, intrinsic :: iso_c_binding
  implicit none
  include 'fftw_wrapper.f03'
  integer, parameter :: n = 128, nx = 16, ny = 12
  type(C_PTR) :: plan
  complex(C_DOUBLE_COMPLEX) :: x(n), y(n), a(nx, ny), b(nx, ny)
  complex(C_FLOAT_COMPLEX) :: xs(n), ys(n)
  real(C_DOUBLE) :: r(nx, ny)
  complex(C_DOUBLE_COMPLEX) :: c(nx / 2 + 1, ny), expect
  real(C_DOUBLE) :: twopi, err, add, mul, fmas
  integer :: i, j, k, nfail
  twopi = 8.0d0 * atan(1.0d0)
  nfail = 0

  ! 1-D double forward: tone at f = 9
  do i = 1, n
    x(i) = exp(cmplx(0.0d0, twopi * 9.0d0 * (i - 1) / n, kind=C_DOUBLE))
  end do
  plan = fftw_plan_dft_1d(n, x, y, FFTW_FORWARD, FFTW_ESTIMATE)
  call fftw_execute_dft(plan, x, y)
  call fftw_flops(plan, add, mul, fmas)
  call fftw_destroy_plan(plan)
  err = 0.0d0
  do k = 1, n
    expect = (0.0d0, 0.0d0)
    if (k == 10) expect = cmplx(real(n, C_DOUBLE), 0.0d0, kind=C_DOUBLE)
    err = max(err, abs(y(k) - expect))
  end do
  if (err > 1.0d-10 .or. add + mul + fmas <= 0.0d0) then
    print *, 'FAIL 1d-double', err, add, mul, fmas
    nfail = nfail + 1
  end if

  ! 2-D double forward, a(nx, ny): tone (fx, fy) = (5, 4), planned as (ny, nx)
  do j = 1, ny
    do i = 1, nx
      a(i, j) = exp(cmplx(0.0d0, twopi * (5.0d0 * (i - 1) / nx + 4.0d0 * (j - 1) / ny), kind=C_DOUBLE))
    end do
  end do
  plan = fftw_plan_dft_2d(ny, nx, a, b, FFTW_FORWARD, FFTW_ESTIMATE)
  call fftw_execute(plan)
  call fftw_destroy_plan(plan)
  err = 0.0d0
  do j = 1, ny
    do i = 1, nx
      expect = (0.0d0, 0.0d0)
      if (i == 6 .and. j == 5) expect = cmplx(real(nx * ny, C_DOUBLE), 0.0d0, kind=C_DOUBLE)
      err = max(err, abs(b(i, j) - expect))
    end do
  end do
  if (err > 1.0d-10) then
    print *, 'FAIL 2d-double', err
    nfail = nfail + 1
  end if

  ! 1-D single backward: tone at f = 3 -> bin n - 3
  do i = 1, n
    xs(i) = cmplx(exp(cmplx(0.0d0, twopi * 3.0d0 * (i - 1) / n, kind=C_DOUBLE)), kind=C_FLOAT)
  end do
  plan = fftwf_plan_dft_1d(n, xs, ys, FFTW_BACKWARD, FFTW_ESTIMATE)
  call fftwf_execute_dft(plan, xs, ys)
  call fftwf_destroy_plan(plan)
  err = 0.0d0
  do k = 1, n
    expect = (0.0d0, 0.0d0)
    if (k == n - 2) expect = cmplx(real(n, C_DOUBLE), 0.0d0, kind=C_DOUBLE)
    err = max(err, abs(cmplx(ys(k), kind=C_DOUBLE) - expect))
  end do
  if (err > 1.0d-3) then
    print *, 'FAIL 1d-single', err
    nfail = nfail + 1
  end if

  ! 2-D R2C double, r(nx, ny) = cos(2 pi (2 i / nx + 3 j / ny)): half spectrum
  ! c(nx/2+1, ny) holds nx ny / 2 at (fx, fy) = (2, 3)
  do j = 1, ny
    do i = 1, nx
      r(i, j) = cos(twopi * (2.0d0 * (i - 1) / nx + 3.0d0 * (j - 1) / ny))
    end do
  end do
  plan = fftw_plan_dft_r2c_2d(ny, nx, r, c, FFTW_ESTIMATE)
  call fftw_execute_dft_r2c(plan, r, c)
  call fftw_destroy_plan(plan)
  err = 0.0d0
  do j = 1, ny
    do i = 1, nx / 2 + 1
      expect = (0.0d0, 0.0d0)
      if (i == 3 .and. j == 4) expect = cmplx(real(nx * ny, C_DOUBLE) / 2.0d0, 0.0d0, kind=C_DOUBLE)
      err = max(err, abs(c(i, j) - expect))
    end do
  end do
  if (err > 1.0d-10) then
    print *, 'FAIL r2c-2d-double', err
    nfail = nfail + 1
  end if

  if (nfail == 0) then
    print *, 'PASS modern'
  else
    stop 1
  end if
end program modern
