! dns_3d_gpu.f90 -- Fortran driver of the MI355X channel-DNS step.
!
! It reads the reference's input file (the eight namelists of DNS_pressure_BC_3D.f90:559-568, in
! the same order, plus an optional &gpu namelist) and runs the time loop with the reference's
! stage subroutine names and call order (DNS_pressure_BC_3D.f90:286-361).  Each stage is a thin
! wrapper over the C ABI of include/f90dns.h through ISO_C_BINDING; the fields stay on the GPU and
! come back only for diagnostics, the output files and the final restart file.  Every nwrt
! steps it prints the diagnostic line and writes output_time_series.dat and flow_statistics.dat,
! every 10 nwrt steps a velocity snapshot and the plane-mean profile, in a directory named after
! the input file and the start time, as the reference does (DNS_pressure_BC_3D.f90:328-338,
! 2557-2586, 2664-2694, 2702-2948).
!
! Usage: dns_3d_gpu [input.dat]
program dns_3d_gpu
    use iso_c_binding
    use iso_fortran_env, only: wp => real64, int32, int64
    use dns_c_api
    implicit none

    real(wp), parameter :: pi = 4.0_wp * atan(1.0_wp)

    ! &grid, &time_control, &simulation, &output, &flow_control, &perturbations,
    ! &pressure_solver, &computational (names as in the reference)
    integer :: nx_input, ny_input, nz_input
    integer :: istart, nsteps, nwrt, restart_write_freq
    real(wp) :: re, dt, alpha, beta, ta, ybar, cgstol, cs, u00, wavlen, xlen, ylen, u_centerline
    logical :: use_crank_nicolson
    integer :: iform, iles
    integer :: flow_control_method, controller_update_freq
    real(wp) :: target_pressure_gradient, target_bulk_velocity, controller_gain
    logical :: enable_perturbations, zero_w_perturbations, use_analytical_perturbations
    real(wp) :: perturbation_amplitude
    integer :: analytical_mode_x, analytical_mode_y
    logical :: use_weak_form_divergence, use_dealias_23_rule
    ! &gpu (optional): precision 0 = FP64, 1 = mixed; lgl_derivative 0 = reference D1, 1 = exact;
    ! seed/noise_kmax for the seeded noise used when use_analytical_perturbations is false;
    ! fused_step = .true. launches one fused step instead of the five stage calls.
    integer :: precision, lgl_derivative, noise_kmax
    integer(int64) :: seed
    logical :: fused_step

    namelist /grid/ nx_input, ny_input, nz_input
    namelist /time_control/ istart, dt, nsteps, nwrt, restart_write_freq
    namelist /simulation/ re, alpha, beta, ta, ybar, cgstol, cs, u00, wavlen, xlen, ylen, use_crank_nicolson, &
                          u_centerline
    namelist /output/ iform, iles
    namelist /flow_control/ flow_control_method, target_pressure_gradient, target_bulk_velocity, controller_gain, &
                            controller_update_freq
    namelist /perturbations/ enable_perturbations, perturbation_amplitude, zero_w_perturbations, &
                             use_analytical_perturbations, analytical_mode_x, analytical_mode_y
    namelist /pressure_solver/ use_weak_form_divergence
    namelist /computational/ use_dealias_23_rule
    namelist /gpu/ precision, lgl_derivative, seed, noise_kmax, fused_step

    character(len=256) :: input_filename
    type(c_ptr) :: ctx = c_null_ptr
    type(dns_params_t) :: prm
    integer :: nx, ny, nz, istep, ios
    integer(c_int) :: rc
    real(wp) :: time, max_divergence
    integer(int64) :: t0, t1, rate
    real(wp), allocatable, target :: u(:,:,:), v(:,:,:), w(:,:,:), p_total(:,:,:), un(:,:,:), vn(:,:,:), wn(:,:,:)
    real(wp), allocatable, target :: zpts(:), zwts(:)
    character(len=256) :: output_directory

    input_filename = 'input.dat'
    if (command_argument_count() >= 1) call get_command_argument(1, input_filename)

    write(*,'(A)') ' ============================================'
    write(*,'(A)') '   3D DNS SOLVER - CHANNEL FLOW (MI355X)'
    write(*,'(A)') ' ============================================'

    call read_input_3d()
    nx = nx_input; ny = ny_input; nz = nz_input

    prm%nx = nx; prm%ny = ny; prm%nz = nz
    prm%xlen = xlen; prm%ylen = ylen; prm%re = re; prm%dt = dt
    prm%use_crank_nicolson = merge(1, 0, use_crank_nicolson)
    prm%use_dealias_23_rule = merge(1, 0, use_dealias_23_rule)
    prm%use_weak_form_divergence = merge(1, 0, use_weak_form_divergence)
    prm%flow_control_method = flow_control_method
    prm%target_pressure_gradient = target_pressure_gradient
    prm%target_bulk_velocity = target_bulk_velocity
    prm%controller_gain = controller_gain
    prm%controller_update_freq = controller_update_freq
    prm%istart = istart
    prm%u_centerline = u_centerline
    prm%enable_perturbations = merge(1, 0, enable_perturbations)
    prm%perturbation_kind = merge(0, 1, use_analytical_perturbations)
    prm%analytical_mode_x = analytical_mode_x
    prm%analytical_mode_y = analytical_mode_y
    prm%perturbation_amplitude = perturbation_amplitude
    prm%seed = seed
    prm%noise_kmax = noise_kmax
    prm%precision = precision
    prm%lgl_derivative = lgl_derivative

    rc = c_dns_create(prm, ctx)
    if (rc /= DNS_OK) then
        write(*,'(A,A)') ' ERROR: dns_create: ', dns_error_message(c_null_ptr)
        error stop 2
    end if
    call check(c_dns_init_fields(ctx), 'dns_init_fields')
    write(*,'(A,I0,A,I0,A,I0)') ' Grid: ', nx, 'x', ny, 'x', nz
    allocate(zpts(nz), zwts(nz))
    call check(c_dns_grid_nodes(ctx, c_loc(zpts), c_loc(zwts)), 'dns_grid_nodes')
    call initialize_output_files()
    call check_divergence_3d(0)

    allocate(u(nx,ny,nz), v(nx,ny,nz), w(nx,ny,nz), p_total(nx,ny,nz))
    allocate(un(nx,ny,nz), vn(nx,ny,nz), wn(nx,ny,nz))
    time = 0.0_wp
    write(*,'(A)') ' Starting main time integration...'
    call system_clock(t0, rate)
    do istep = 1, nsteps
        time = time + dt
        if (istep == nsteps) then          ! un, vn, wn of the final restart file
            call check(c_dns_get_fields(ctx, c_loc(un), c_loc(vn), c_loc(wn), c_null_ptr), 'dns_get_fields')
        end if
        if (fused_step) then
            call check(c_dns_step(ctx, 1_c_int), 'dns_step')
        else
            call check(c_dns_begin_step(ctx), 'dns_begin_step')
            call compute_source_terms_3d()
            call viscous_step_3d()
            call solve_pressure_3d()
            call pressure_correction_3d()
            call apply_boundary_conditions_3d()
        end if
        if (mod(istep, nwrt) == 0) then
            call output_diagnostics_3d(istep)
            call write_time_series_data(istep)
            call write_flow_statistics(istep)
            if (mod(istep, 10 * nwrt) == 0) then
                call write_velocity_field_snapshot(istep)
                call write_spanwise_averaged_profile(istep)
            end if
        end if
        if (mod(istep, 50) == 0) call check_divergence_3d(istep)
    end do
    call check(c_dns_synchronize(ctx), 'dns_synchronize')
    call system_clock(t1)

    call write_restart_file_3d(nsteps)
    write(*,'(A)') ' ========================================================'
    write(*,'(A)') '               SIMULATION COMPLETED SUCCESSFULLY'
    write(*,'(A)') ' ========================================================'
    write(*,'(A,F12.6,A)') ' Total simulation time:   ', time, ' (non-dimensional)'
    write(*,'(A,I0,A)')     ' Total time steps:       ', nsteps, ' steps'
    write(*,'(A,F10.3,A)')  ' Total wall clock time:  ', real(t1 - t0, wp) / real(rate, wp), ' seconds'
    if (t1 > t0) write(*,'(A,ES12.4,A)') ' Performance:            ', &
        real(nsteps, wp) * real(nx, wp) * real(ny, wp) * real(nz, wp) * real(rate, wp) / real(t1 - t0, wp), &
        ' grid-point steps/second'
    write(*,'(A,ES12.4)')   ' Max divergence:         ', max_divergence
    call check(c_dns_destroy(ctx), 'dns_destroy')

contains

    subroutine check(code, what)
        integer(c_int), intent(in) :: code
        character(len=*), intent(in) :: what
        if (code /= DNS_OK) then
            write(*,'(A,A,A,A)') ' ERROR: ', what, ': ', dns_error_message(ctx)
            error stop 3
        end if
    end subroutine check

    ! read_input_3d (DNS_pressure_BC_3D.f90:554-708): defaults, then the eight namelists in order
    subroutine read_input_3d()
        nx_input = 128; ny_input = 32; nz_input = 33
        istart = 0; nsteps = 1000; nwrt = 100; restart_write_freq = 1000
        re = 180.0_wp; dt = 0.01_wp; alpha = 1.0_wp; beta = 1.0_wp
        ta = 0.0_wp; ybar = 2.0_wp; cgstol = 1e-6_wp; cs = 0.1_wp; u00 = 0.0_wp; wavlen = 1.0_wp
        xlen = 2.0_wp * pi; ylen = 2.0_wp * pi; use_crank_nicolson = .true.; u_centerline = 1.5_wp
        iform = 0; iles = 0
        flow_control_method = 1; target_pressure_gradient = 0.03_wp
        target_bulk_velocity = 1.0_wp; controller_gain = 0.15_wp; controller_update_freq = 7
        enable_perturbations = .false.; perturbation_amplitude = 0.01_wp; zero_w_perturbations = .false.
        use_analytical_perturbations = .false.; analytical_mode_x = 1; analytical_mode_y = 1
        use_weak_form_divergence = .false.; use_dealias_23_rule = .false.
        precision = 0; lgl_derivative = 0; seed = 0_int64; noise_kmax = 8; fused_step = .false.
        open(7, file=trim(input_filename), status='old', iostat=ios)
        if (ios == 0) then
            read(7, nml=grid)
            read(7, nml=time_control)
            read(7, nml=simulation)
            read(7, nml=output)
            read(7, nml=flow_control)
            read(7, nml=perturbations)
            read(7, nml=pressure_solver)
            read(7, nml=computational)
            read(7, nml=gpu, iostat=ios)
            close(7)
            write(*,'(A,A)') ' Input read from ', trim(input_filename)
        else
            write(*,'(A,A)') ' WARNING: Could not open input file: ', trim(input_filename)
            write(*,'(A)') ' Using default parameters'
        end if
        if (flow_control_method /= 1 .and. flow_control_method /= 2) then
            write(*,'(A,I0)') ' ERROR: Invalid flow_control_method, got ', flow_control_method
            stop
        end if
        write(*,'(A,F10.1)') ' Reynolds number: ', re
        write(*,'(A,F8.5)') ' Time step: ', dt
        write(*,'(A,I0)') ' Number of steps: ', nsteps
    end subroutine read_input_3d

    ! the stage subroutines (DNS_pressure_BC_3D.f90:314-326)
    subroutine compute_source_terms_3d()
        call check(c_dns_compute_source_terms(ctx), 'compute_source_terms_3d')
    end subroutine compute_source_terms_3d

    subroutine viscous_step_3d()
        call check(c_dns_viscous_step(ctx), 'viscous_step_3d')
    end subroutine viscous_step_3d

    subroutine solve_pressure_3d()
        call check(c_dns_solve_pressure(ctx), 'solve_pressure_3d')
    end subroutine solve_pressure_3d

    subroutine pressure_correction_3d()
        call check(c_dns_pressure_correction(ctx), 'pressure_correction_3d')
    end subroutine pressure_correction_3d

    subroutine apply_boundary_conditions_3d()
        call check(c_dns_apply_boundary_conditions(ctx), 'apply_boundary_conditions_3d')
    end subroutine apply_boundary_conditions_3d

    ! output_diagnostics_3d (DNS_pressure_BC_3D.f90:2094-2146): same line format
    subroutine output_diagnostics_3d(step)
        integer, intent(in) :: step
        type(dns_diag_t) :: d
        call check(c_dns_diagnostics(ctx, d), 'output_diagnostics_3d')
        if (flow_control_method == 2) then
            write(*,'(A,I6,A,F12.6,A,5(ES12.4,A),F8.5,A,ES12.4)') ' Step ', step, ', t=', time, &
                ', |u|max=', d%umax, ', |v|max=', d%vmax, ', |w|max=', d%wmax, ', E=', d%energy, &
                ', dP/dx=', d%pressure_gradient, ', U_bulk=', d%bulk_velocity, ', dP/dx_avg=', d%avg_pressure_gradient
        else
            write(*,'(A,I6,A,F12.6,A,5(ES12.4,A))') ' Step ', step, ', t=', time, &
                ', |u|max=', d%umax, ', |v|max=', d%vmax, ', |w|max=', d%wmax, ', E=', d%energy, &
                ', dP/dx=', d%pressure_gradient
        end if
    end subroutine output_diagnostics_3d

    ! create_output_directory + initialize_output_files (DNS_pressure_BC_3D.f90:2557-2586,
    ! 2664-2694): <input file without .dat>_<YYYYMMDD_HHMMSS>, and the two file headers
    subroutine initialize_output_files()
        character(len=256) :: base_name
        character(len=10) :: date_str, time_str
        integer :: un_
        call date_and_time(date_str, time_str)
        base_name = trim(input_filename)
        if (index(base_name, '.dat') > 0) base_name = base_name(1:index(base_name, '.dat') - 1)
        output_directory = trim(base_name) // '_' // date_str(1:4) // date_str(5:6) // date_str(7:8) // '_' // &
                           time_str(1:2) // time_str(3:4) // time_str(5:6)
        call execute_command_line('mkdir -p ' // trim(output_directory))
        write(*,'(A)') ' Output directory created:'
        write(*,'(A,A)') '   ', trim(output_directory)
        open(newunit=un_, file=trim(output_directory) // '/output_time_series.dat', status='replace')
        write(un_,'(A)') '# DNS 3D Time Series Data'
        write(un_,'(A)') '# Columns: step, time, u_max, v_max, w_max, energy, pressure_gradient, bulk_velocity'
        close(un_)
        open(newunit=un_, file=trim(output_directory) // '/flow_statistics.dat', status='replace')
        write(un_,'(A)') '# DNS 3D Flow Statistics Summary'
        write(un_,'(A)') '# Updated periodically during simulation'
        close(un_)
    end subroutine initialize_output_files

    ! write_time_series_data (DNS_pressure_BC_3D.f90:2702-2748): one line per call, the averaged
    ! G and U_b under method 2, the current G under method 1
    subroutine write_time_series_data(step)
        integer, intent(in) :: step
        type(dns_diag_t) :: d
        integer :: un_
        call check(c_dns_diagnostics(ctx, d), 'write_time_series_data')
        open(newunit=un_, file=trim(output_directory) // '/output_time_series.dat', status='old', position='append')
        if (flow_control_method == 2) then
            write(un_,'(I8,7(ES15.6))') step, time, d%umax, d%vmax, d%wmax, d%energy, d%avg_pressure_gradient, &
                                        d%bulk_velocity
        else
            write(un_,'(I8,6(ES15.6))') step, time, d%umax, d%vmax, d%wmax, d%energy, d%pressure_gradient
        end if
        close(un_)
    end subroutine write_time_series_data

    ! write_flow_statistics (DNS_pressure_BC_3D.f90:2849-2948): the file is rewritten each call.
    ! The wall shear comes from the plane means of u next to each wall (one-sided differences over
    ! the LGL nodes, as the reference); the bulk velocity is the LGL-weighted plane-mean profile
    ! integral of the library (the reference's method-1 branch divides the same sum by 2 instead
    ! of by the sum of the weights, which is 2).
    subroutine write_flow_statistics(step)
        integer, intent(in) :: step
        type(dns_diag_t) :: d
        real(wp) :: nu, du_dz_lower, du_dz_upper, tau_lower, tau_upper, ut
        integer :: un_
        call check(c_dns_diagnostics(ctx, d), 'write_flow_statistics')
        call check(c_dns_get_fields(ctx, c_loc(u), c_null_ptr, c_null_ptr, c_null_ptr), 'dns_get_fields')
        nu = 1.0_wp / re
        du_dz_lower = sum((u(:,:,2) - u(:,:,1)) / (zpts(2) - zpts(1))) / real(nx * ny, wp)
        du_dz_upper = sum((u(:,:,nz) - u(:,:,nz-1)) / (zpts(nz) - zpts(nz-1))) / real(nx * ny, wp)
        tau_lower = nu * du_dz_lower
        tau_upper = nu * du_dz_upper
        ut = sqrt((abs(tau_lower) + abs(tau_upper)) / 2.0_wp)
        open(newunit=un_, file=trim(output_directory) // '/flow_statistics.dat', status='replace')
        write(un_,'(A)') '# DNS 3D Flow Statistics Summary'
        write(un_,'(A,I0)') '# Time step: ', step
        write(un_,'(A,F12.6)') '# Time: ', time
        write(un_,'(A)')
        write(un_,'(A)') '# Maximum velocities:'
        write(un_,'(A,ES15.6)') 'u_max = ', d%umax
        write(un_,'(A,ES15.6)') 'v_max = ', d%vmax
        write(un_,'(A,ES15.6)') 'w_max = ', d%wmax
        write(un_,'(A)')
        write(un_,'(A)') '# Energy and bulk flow:'
        write(un_,'(A,ES15.6)') 'kinetic_energy = ', d%energy
        write(un_,'(A,ES15.6)') 'bulk_velocity = ', d%bulk_velocity
        if (flow_control_method == 2) then
            write(un_,'(A,ES15.6)') 'pressure_gradient = ', d%avg_pressure_gradient
        else
            write(un_,'(A,ES15.6)') 'pressure_gradient = ', d%pressure_gradient
        end if
        write(un_,'(A)')
        write(un_,'(A)') '# Wall shear stress:'
        write(un_,'(A,ES15.6)') 'tau_wall_lower = ', tau_lower
        write(un_,'(A,ES15.6)') 'tau_wall_upper = ', tau_upper
        write(un_,'(A,ES15.6)') 'friction_velocity_avg = ', ut
        write(un_,'(A,ES15.6)') 're_tau = ', ut * re
        close(un_)
    end subroutine write_flow_statistics

    ! write_velocity_field_snapshot (DNS_pressure_BC_3D.f90:2756-2791): every point as text
    subroutine write_velocity_field_snapshot(step)
        integer, intent(in) :: step
        character(len=32) :: filename
        integer :: i, j, k, un_
        call check(c_dns_get_fields(ctx, c_loc(u), c_loc(v), c_loc(w), c_null_ptr), 'dns_get_fields')
        write(filename, '(A,I5.5,A)') 'velocity_field_', step, '.dat'
        open(newunit=un_, file=trim(output_directory) // '/' // trim(filename), status='replace')
        write(un_,'(A)') '# DNS 3D Velocity Field Snapshot'
        write(un_,'(A,I0)') '# Time step: ', step
        write(un_,'(A,F12.6)') '# Time: ', time
        write(un_,'(A)') '# Columns: i, j, k, x, y, z, u, v, w'
        do k = 1, nz
            do j = 1, ny
                do i = 1, nx
                    write(un_,'(3I6,6ES15.6)') i, j, k, real(i - 1, wp) * xlen / real(nx, wp), &
                        real(j - 1, wp) * ylen / real(ny, wp), zpts(k), u(i,j,k), v(i,j,k), w(i,j,k)
                end do
            end do
        end do
        close(un_)
        write(*,'(A,A)') ' Velocity field snapshot written: ', trim(filename)
    end subroutine write_velocity_field_snapshot

    ! write_spanwise_averaged_profile (DNS_pressure_BC_3D.f90:2799-2841): plane means per node
    ! (u, v, w as fetched by the snapshot written just before)
    subroutine write_spanwise_averaged_profile(step)
        integer, intent(in) :: step
        character(len=32) :: filename
        integer :: k, un_
        write(filename, '(A,I5.5,A)') 'spanwise_profile_', step, '.dat'
        open(newunit=un_, file=trim(output_directory) // '/' // trim(filename), status='replace')
        write(un_,'(A)') '# DNS 3D Spanwise-Averaged Velocity Profile'
        write(un_,'(A,I0)') '# Time step: ', step
        write(un_,'(A,F12.6)') '# Time: ', time
        write(un_,'(A)') '# Columns: k, z, u_avg, v_avg, w_avg'
        do k = 1, nz
            write(un_,'(I6,4ES15.6)') k, zpts(k), sum(u(:,:,k)) / real(nx * ny, wp), &
                sum(v(:,:,k)) / real(nx * ny, wp), sum(w(:,:,k)) / real(nx * ny, wp)
        end do
        close(un_)
        write(*,'(A,A)') ' Spanwise profile written: ', trim(filename)
    end subroutine write_spanwise_averaged_profile

    ! check_divergence_3d (DNS_pressure_BC_3D.f90:2171-2190)
    subroutine check_divergence_3d(step)
        integer, intent(in) :: step
        call check(c_dns_check_divergence(ctx, max_divergence), 'check_divergence_3d')
        write(*,'(A,I6,A,E12.4)') ' Step ', step, ', Max divergence: ', max_divergence
    end subroutine check_divergence_3d

    ! write_restart_file_3d layout (DNS_pressure_BC_3D.f90:2999-3027), written to ./restart.dat
    subroutine write_restart_file_3d(step)
        integer, intent(in) :: step
        type(dns_diag_t) :: d
        integer :: un_
        call check(c_dns_get_fields(ctx, c_loc(u), c_loc(v), c_loc(w), c_loc(p_total)), 'dns_get_fields')
        call check(c_dns_diagnostics(ctx, d), 'dns_diagnostics')
        open(newunit=un_, file='restart.dat', form='unformatted', access='stream', status='replace', iostat=ios)
        if (ios /= 0) then
            write(*,'(A)') ' ERROR: Cannot create restart file: restart.dat'
            return
        end if
        write(un_) 'DNS3D_RESTART_v1.0'
        write(un_) int(step, int32), time, dt
        write(un_) int(nx, int32), int(ny, int32), int(nz, int32)
        write(un_) re, 2.0_wp * pi / xlen, 2.0_wp * pi / ylen
        write(un_) u
        write(un_) v
        write(un_) w
        write(un_) un
        write(un_) vn
        write(un_) wn
        write(un_) p_total
        write(un_) d%pressure_gradient, d%bulk_velocity, d%bulk_error_integral, d%bulk_error_prev
        close(un_)
        write(*,'(A,I0,A,F12.6,A)') ' Restart file written at step ', step, ', time = ', time, ' to restart.dat'
    end subroutine write_restart_file_3d

end program dns_3d_gpu
