! fftw3_dns_module.f90 -- the reference's FFT boundary module (fftw3_dns_module.f90:19-152 of the
! reference) with the same type, procedure names and argument lists, backed by the MI355X plane
! transforms of include/f90dns.h instead of FFTW3.  Forward: unnormalised r2c of real(nx,ny)
! into complex(nx/2+1,ny); backward: c2r divided by nx*ny.  Errors stop the program as the
! reference's do (fftw3_dns_module.f90:77-82, 101-103).
module fftw3_dns_module
    use iso_c_binding
    use dns_c_api, only: c_dns_fft_plan_create, c_dns_fft_plan_destroy, c_dns_fft_forward, c_dns_fft_backward, &
                         dns_error_message, DNS_PREC_FP64
    use iso_fortran_env, only: wp => real64
    implicit none
    private
    public :: fftw3_dns_plans, fftw3_initialize_threading, setup_fftw3_plans_dns, fftw3_forward_2d_dns, &
              fftw3_backward_2d_dns, destroy_fftw3_plans_dns

    type :: fftw3_dns_plans
        type(c_ptr) :: plan_forward = c_null_ptr     ! the device plan (one handle serves both directions)
        type(c_ptr) :: plan_backward = c_null_ptr
        integer :: nx = 0, ny = 0, nxhp = 0
        logical :: initialized = .false.
    end type fftw3_dns_plans

contains

    subroutine fftw3_initialize_threading()
        ! the GPU transforms need no host threads; kept for the reference's call sequence
        write(*,'(A)') ' FFT on the GPU (gfx950 plane transforms)'
    end subroutine fftw3_initialize_threading

    subroutine setup_fftw3_plans_dns(plans, nx, ny, sample_real, sample_complex)
        type(fftw3_dns_plans), intent(out) :: plans
        integer, intent(in) :: nx, ny
        real(wp), intent(inout) :: sample_real(nx, ny)
        complex(wp), intent(inout) :: sample_complex(nx/2+1, ny)
        integer(c_int) :: rc
        plans%nx = nx; plans%ny = ny; plans%nxhp = nx/2 + 1
        rc = c_dns_fft_plan_create(int(nx, c_int), int(ny, c_int), DNS_PREC_FP64, plans%plan_forward)
        if (rc /= 0) then
            write(*,'(A,A)') ' Error: Failed to create the GPU FFT plan: ', dns_error_message(c_null_ptr)
            stop
        end if
        plans%plan_backward = plans%plan_forward
        plans%initialized = .true.
        if (.false.) sample_real(1,1) = real(sample_complex(1,1), wp)
    end subroutine setup_fftw3_plans_dns

    subroutine fftw3_forward_2d_dns(plans, real_field, complex_field)
        type(fftw3_dns_plans), intent(in) :: plans
        real(wp), intent(inout), target :: real_field(plans%nx, plans%ny)
        complex(wp), intent(out), target :: complex_field(plans%nxhp, plans%ny)
        if (.not. plans%initialized) stop 'Error: FFTW3 plans not initialized'
        if (c_dns_fft_forward(plans%plan_forward, c_loc(real_field), c_loc(complex_field), 1_c_int) /= 0) &
            stop 'Error: GPU forward transform failed'
    end subroutine fftw3_forward_2d_dns

    subroutine fftw3_backward_2d_dns(plans, complex_field, real_field)
        type(fftw3_dns_plans), intent(in) :: plans
        complex(wp), intent(inout), target :: complex_field(plans%nxhp, plans%ny)
        real(wp), intent(out), target :: real_field(plans%nx, plans%ny)
        if (.not. plans%initialized) stop 'Error: FFTW3 plans not initialized'
        if (c_dns_fft_backward(plans%plan_backward, c_loc(complex_field), c_loc(real_field), 1_c_int) /= 0) &
            stop 'Error: GPU backward transform failed'
    end subroutine fftw3_backward_2d_dns

    subroutine destroy_fftw3_plans_dns(plans)
        type(fftw3_dns_plans), intent(inout) :: plans
        integer(c_int) :: rc
        if (plans%initialized) then
            rc = c_dns_fft_plan_destroy(plans%plan_forward)
            plans%plan_forward = c_null_ptr; plans%plan_backward = c_null_ptr
            plans%initialized = .false.
        end if
    end subroutine destroy_fftw3_plans_dns

end module fftw3_dns_module
