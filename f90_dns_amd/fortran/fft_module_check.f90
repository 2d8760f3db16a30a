! fft_module_check.f90 -- drives the drop-in fftw3_dns_module exactly as the reference's solver
! does (setup_fftw3_plans_dns, fftw3_forward_2d_dns, fftw3_backward_2d_dns, destroy) on the
! round-trip known answer of test_fft_roundtrip (DNS_pressure_BC_3D.f90:2002-2069) and a spectral
! x-derivative; prints the errors and PASS/FAIL with the reference's 1e-12 tolerance.
program fft_module_check
    use fftw3_dns_module
    use iso_fortran_env, only: wp => real64
    implicit none
    integer, parameter :: nx = 32, ny = 32
    real(wp), parameter :: pi = 4.0_wp * atan(1.0_wp)
    real(wp) :: xlen, ylen, alpha, beta, x, y, err_rt, err_dx
    real(wp), allocatable :: f(:,:), g(:,:), dfdx(:,:)
    complex(wp), allocatable :: fh(:,:)
    type(fftw3_dns_plans) :: plans
    integer :: i, j
    xlen = 2.0_wp * pi; ylen = pi
    alpha = 2.0_wp * pi / xlen; beta = 2.0_wp * pi / ylen
    allocate(f(nx,ny), g(nx,ny), dfdx(nx,ny), fh(nx/2+1,ny))
    call fftw3_initialize_threading()
    call setup_fftw3_plans_dns(plans, nx, ny, g, fh)
    do j = 1, ny
        y = real(j-1, wp) * ylen / real(ny, wp)
        do i = 1, nx
            x = real(i-1, wp) * xlen / real(nx, wp)
            f(i,j) = sin(2.0_wp * alpha * x) * cos(3.0_wp * beta * y)
            dfdx(i,j) = 2.0_wp * alpha * cos(2.0_wp * alpha * x) * cos(3.0_wp * beta * y)
        end do
    end do
    g = f
    call fftw3_forward_2d_dns(plans, g, fh)
    call fftw3_backward_2d_dns(plans, fh, g)
    err_rt = maxval(abs(g - f))
    g = f
    call fftw3_forward_2d_dns(plans, g, fh)
    do j = 1, ny
        do i = 1, nx/2 + 1
            fh(i,j) = fh(i,j) * cmplx(0.0_wp, real(i-1, wp) * alpha, kind=wp)
        end do
    end do
    call fftw3_backward_2d_dns(plans, fh, g)
    err_dx = maxval(abs(g - dfdx))
    call destroy_fftw3_plans_dns(plans)
    write(*,'(A,ES10.3)') ' round-trip max error: ', err_rt
    write(*,'(A,ES10.3)') ' d/dx max error:       ', err_dx
    if (err_rt < 1.0e-12_wp .and. err_dx < 1.0e-12_wp) then
        write(*,'(A)') ' FFT MODULE PASSED'
    else
        write(*,'(A)') ' FFT MODULE FAILED'
        error stop 1
    end if
end program fft_module_check
