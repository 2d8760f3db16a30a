! dns_3d_gpu.f90 -- Fortran driver of the MI355X channel-DNS step.
!
! It takes the reference's command line (process_command_line_args, DNS_pressure_BC_3D.f90:445-528:
! [--restart [file.dat]] [input.dat], --help, default input_3d.dat), reads the reference's input
! file (the eight namelists of :559-568, in the same order, plus an optional &gpu namelist) and
! runs the time loop with the reference's stage subroutine names and call order (:286-361).  Each
! stage is a thin wrapper over the C ABI of include/f90dns.h through ISO_C_BINDING; the fields stay
! on the GPU and come back only for diagnostics, the output files and the restart files.  In a
! directory named after the input file and the start time (:2557-2586) it writes
! simulation_summary.txt (:2594-2659), output_time_series.dat and flow_statistics.dat every nwrt
! steps, a velocity snapshot and the plane-mean profile every 10 nwrt steps (:328-338, 2664-2948),
! restart.dat every restart_write_freq steps before the step body and after the last step
! (:307-309, 367, 2976-3038); it prints the timing summary every 100 steps (:2207-2230) and the
! final summary (:373-421).
!
! Restart reading keeps the reference's behaviour: read_input_3d resets the restart request of
! the command line (:626-627) and the file is looked for in the new output directory (:3075),
! so a run always starts new.  &gpu restart_file = 'path' (not in the reference) reads a restart
! file of the same layout and resumes from it.
!
! Usage: dns_3d_gpu [--restart [file.dat]] [input.dat] | --help
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
    ! &gpu (optional): precision 0 = FP64, 1 = mixed, 2 = mixed with the D1 contraction in FP32;
    ! lgl_derivative 0 = reference D1, 1 = exact;
    ! seed/noise_kmax for the seeded noise used when use_analytical_perturbations is false;
    ! fused_step = .true. launches one fused step instead of the five stage calls; restart_file
    ! names a restart file to resume from.
    integer :: precision, lgl_derivative, noise_kmax
    integer(int64) :: seed
    logical :: fused_step
    character(len=256) :: restart_file
    ! the 2D solver's numerics on the 3D step (no 3D counterpart): rk4_convection = the 4-stage
    ! 1/irk convection (DNS_pressure_BC_2D.f90:931-1014), jacgs_pressure = the Jacobi-CG
    ! pressure solve (:2361-2476) with &simulation cgstol
    logical :: rk4_convection, jacgs_pressure

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
    namelist /gpu/ precision, lgl_derivative, seed, noise_kmax, fused_step, restart_file, rk4_convection, &
                   jacgs_pressure

    character(len=256) :: input_filename, restart_filename
    logical :: restart_from_file
    integer :: restart_step, current_step
    real(wp) :: restart_time
    type(c_ptr) :: ctx = c_null_ptr
    type(dns_params_t) :: prm
    integer :: nx, ny, nz, istep, ios, istart_loop, iend_loop
    integer(c_int) :: rc
    real(wp) :: time, max_divergence, g_restart, ub_restart, int_restart, prev_restart
    ! step timing (print_step_timing, :2207-2230), wall clock of each synchronised step
    integer, parameter :: timing_interval = 100
    integer(int64) :: t0, t1, ts0, ts1, rate
    real(wp) :: step_duration, total_sim_time, min_step_time, max_step_time, avg_step_time, total_wall_time
    real(wp), allocatable, target :: u(:,:,:), v(:,:,:), w(:,:,:), p_total(:,:,:), un(:,:,:), vn(:,:,:), wn(:,:,:)
    real(wp), allocatable, target :: zpts(:), zwts(:)
    character(len=256) :: output_directory

    write(*,'(A)') ' ============================================'
    write(*,'(A)') '   3D DNS SOLVER - CHANNEL FLOW (MI355X)'
    write(*,'(A)') ' ============================================'

    call process_command_line_args()
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
    prm%convection_scheme = merge(1, 0, rk4_convection)
    prm%pressure_solver = merge(1, 0, jacgs_pressure)
    prm%cgstol = cgstol

    rc = c_dns_create(prm, ctx)
    if (rc /= DNS_OK) then
        write(*,'(A,A)') ' ERROR: dns_create: ', dns_error_message(c_null_ptr)
        error stop 2
    end if
    call check(c_dns_init_fields(ctx), 'dns_init_fields')
    write(*,'(A,I0,A,I0,A,I0)') ' Grid: ', nx, 'x', ny, 'x', nz
    allocate(zpts(nz), zwts(nz))
    call check(c_dns_grid_nodes(ctx, c_loc(zpts), c_loc(zwts)), 'dns_grid_nodes')
    allocate(u(nx,ny,nz), v(nx,ny,nz), w(nx,ny,nz), p_total(nx,ny,nz))
    allocate(un(nx,ny,nz), vn(nx,ny,nz), wn(nx,ny,nz))
    call check_divergence_3d(0)
    call initialize_output_files()
    call write_simulation_summary()

    ! restart detection (:220-276): the command-line request was reset by read_input_3d, and the
    ! file is looked for in the new output directory, so this finds nothing, as in the reference
    current_step = 0
    if (restart_from_file) then
        write(*,'(A,A)') ' Command line restart requested from: ', trim(restart_filename)
    else
        restart_filename = 'restart.dat'
    end if
    call read_restart_file_3d(trim(output_directory) // '/' // trim(restart_filename), restart_step, restart_time, &
                              restart_from_file)
    if (.not. restart_from_file .and. len_trim(restart_file) > 0) then
        restart_filename = restart_file
        call read_restart_file_3d(trim(restart_file), restart_step, restart_time, restart_from_file)
    end if
    if (restart_from_file) then
        current_step = restart_step
        time = restart_time
        write(*,'(A)') ' ============================================'
        write(*,'(A)') '   RESTARTING FROM CHECKPOINT'
        write(*,'(A,A)') '   Restart file:  ', trim(restart_filename)
        write(*,'(A,I0)') '   Restart step:  ', restart_step
        write(*,'(A,F12.6)') '   Restart time:  ', restart_time
        write(*,'(A,I0)') '   Remaining steps: ', nsteps - restart_step
        if (restart_step >= nsteps) then
            write(*,'(A)') ' ============================================'
            write(*,'(A)') '   SIMULATION ALREADY COMPLETE'
            write(*,'(A,I0)') '   Restart step: ', restart_step
            write(*,'(A,I0)') '   Target steps: ', nsteps
            write(*,'(A)') '   Use larger nsteps to continue simulation.'
            write(*,'(A)') ' ============================================'
            stop
        end if
        write(*,'(A)') ' ============================================'
    else
        write(*,'(A)') ' ============================================'
        write(*,'(A)') '   STARTING NEW SIMULATION'
        write(*,'(A)') '   No restart requested'
        write(*,'(A)') ' ============================================'
        time = 0.0_wp
    end if

    total_sim_time = 0.0_wp; min_step_time = huge(1.0_wp); max_step_time = 0.0_wp; avg_step_time = 0.0_wp
    write(*,'(A)') ' Starting main time integration...'
    call system_clock(t0, rate)
    istart_loop = current_step + 1
    iend_loop = nsteps
    do istep = istart_loop, iend_loop
        call system_clock(ts0)
        current_step = istep
        if (restart_from_file .and. istep == istart_loop) then
            time = restart_time + dt
        else
            time = time + dt
        end if
        ! un = u (:300): fetched where a restart file needs them
        if (istep == nsteps .or. mod(istep, restart_write_freq) == 0) then
            call check(c_dns_get_fields(ctx, c_loc(un), c_loc(vn), c_loc(wn), c_null_ptr), 'dns_get_fields')
        end if
        if (mod(istep, restart_write_freq) == 0) call write_restart_file_3d(istep, time, .true.)
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
        call check(c_dns_synchronize(ctx), 'dns_synchronize')
        call system_clock(ts1)
        step_duration = real(ts1 - ts0, wp) / real(rate, wp)
        total_sim_time = total_sim_time + step_duration
        min_step_time = min(min_step_time, step_duration)
        max_step_time = max(max_step_time, step_duration)
        avg_step_time = total_sim_time / real(istep, wp)
        if (mod(istep, timing_interval) == 0) call print_step_timing(istep, step_duration)
    end do
    call check(c_dns_synchronize(ctx), 'dns_synchronize')
    call system_clock(t1)
    total_wall_time = real(t1 - t0, wp) / real(rate, wp)

    call write_restart_file_3d(nsteps, time, .false.)
    call final_summary()
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
        restart_file = ''; rk4_convection = .false.; jacgs_pressure = .false.
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
        ! the restart state is reset here, after the command line was read (:626-630)
        restart_filename = 'restart.dat'
        restart_from_file = .false.
        current_step = 0
        restart_step = 0
        restart_time = 0.0_wp
    end subroutine read_input_3d

    ! process_command_line_args (DNS_pressure_BC_3D.f90:445-528)
    subroutine process_command_line_args()
        integer :: num_args, i
        character(len=256) :: arg
        logical :: restart_requested
        restart_requested = .false.
        input_filename = 'input_3d.dat'
        restart_filename = 'restart.dat'
        restart_from_file = .false.
        num_args = command_argument_count()
        if (num_args == 0) then
            write(*,'(A,A)') ' Using default input file: ', trim(input_filename)
            return
        end if
        i = 1
        do while (i <= num_args)
            call get_command_argument(i, arg)
            if (trim(arg) == '--help' .or. trim(arg) == '-h') then
                write(*,'(A)') ' '
                write(*,'(A)') ' USAGE:'
                write(*,'(A)') '   ./dns_3d_gpu [options] [input_file.dat]'
                write(*,'(A)') ' '
                write(*,'(A)') ' OPTIONS:'
                write(*,'(A)') '   --restart [file]    : Restart from checkpoint file (default: restart.dat)'
                write(*,'(A)') '   --help, -h          : Show this help message'
                write(*,'(A)') ' '
                write(*,'(A)') ' ARGUMENTS:'
                write(*,'(A)') '   input_file.dat      : Specify input file (default: input_3d.dat)'
                write(*,'(A)') ' '
                write(*,'(A)') ' EXAMPLES:'
                write(*,'(A)') '   ./dns_3d_gpu                      # New simulation, input_3d.dat'
                write(*,'(A)') '   ./dns_3d_gpu --restart            # Restart from restart.dat'
                write(*,'(A)') '   ./dns_3d_gpu --restart custom.dat # Restart from custom.dat'
                write(*,'(A)') '   ./dns_3d_gpu input_method2.dat    # New simulation, custom input'
                write(*,'(A)') '   ./dns_3d_gpu --restart input_method2.dat # Restart, custom input'
                write(*,'(A)') ' '
                stop
            else if (trim(arg) == '--restart') then
                restart_requested = .true.
                if (i + 1 <= num_args) then
                    call get_command_argument(i + 1, arg)
                    if (arg(1:2) /= '--' .and. index(arg, '.dat') > 0) then
                        restart_filename = trim(arg)
                        i = i + 1
                        write(*,'(A,A)') ' Restart requested from file: ', trim(restart_filename)
                    else
                        write(*,'(A,A)') ' Restart requested from default file: ', trim(restart_filename)
                    end if
                else
                    write(*,'(A,A)') ' Restart requested from default file: ', trim(restart_filename)
                end if
            else if (index(arg, '.dat') > 0) then
                input_filename = trim(arg)
                write(*,'(A,A)') ' Using specified input file: ', trim(input_filename)
            else
                write(*,'(A,A)') ' ERROR: Unknown argument: ', trim(arg)
                write(*,'(A)') ' Use --help for usage information'
                stop
            end if
            i = i + 1
        end do
        if (restart_requested) then
            restart_from_file = .true.
            write(*,'(A)') ' Command line restart mode enabled'
        end if
    end subroutine process_command_line_args

    ! write_simulation_summary (DNS_pressure_BC_3D.f90:2594-2659): run parameters, written once
    ! after the output directory is created
    subroutine write_simulation_summary()
        character(len=10) :: date_str, time_str
        integer :: un_
        call date_and_time(date_str, time_str)
        open(newunit=un_, file=trim(output_directory) // '/simulation_summary.txt', status='replace')
        write(un_,'(A)') '============================================'
        write(un_,'(A)') '   3D DNS SIMULATION SUMMARY'
        write(un_,'(A)') '============================================'
        write(un_,'(A)') ' '
        write(un_,'(A,A)') ' Input file: ', trim(input_filename)
        write(un_,'(A,A)') ' Output directory: ', trim(output_directory)
        write(un_,'(A,A,A,A,A,A)') ' Start time: ', date_str(1:4), '-', date_str(5:6), '-', date_str(7:8)
        write(un_,'(A,A,A,A,A,A)') '             ', time_str(1:2), ':', time_str(3:4), ':', time_str(5:6)
        write(un_,'(A)') ' '
        write(un_,'(A)') ' GRID PARAMETERS:'
        write(un_,'(A,I0,A,I0,A,I0)') '   Grid size: ', nx, ' ' // char(195) // char(151) // ' ', ny, &
                                      ' ' // char(195) // char(151) // ' ', nz
        write(un_,'(A,F8.4)') '   Domain Lx: ', xlen
        write(un_,'(A,F8.4)') '   Domain Ly: ', ylen
        write(un_,'(A,F8.4)') '   Domain Lz: ', 2.0_wp
        write(un_,'(A)') ' '
        write(un_,'(A)') ' SIMULATION PARAMETERS:'
        write(un_,'(A,F10.4)') '   Reynolds number: ', re
        write(un_,'(A,F10.6)') '   Time step: ', dt
        write(un_,'(A,I0)') '   Total steps: ', nsteps
        write(un_,'(A,I0)') '   Output frequency: ', nwrt
        write(un_,'(A,F10.4)') '   Total time: ', real(nsteps, wp) * dt
        write(un_,'(A)') ' '
        write(un_,'(A)') ' FLOW CONTROL:'
        write(un_,'(A,I0)') '   Method: ', flow_control_method
        if (flow_control_method == 2) then
            write(un_,'(A,F10.6)') '   Target bulk velocity: ', target_bulk_velocity
            write(un_,'(A,F10.6)') '   Controller gain: ', controller_gain
            write(un_,'(A,I0)') '   Update frequency: ', controller_update_freq
        end if
        write(un_,'(A)') ' '
        write(un_,'(A)') ' OUTPUT FILES:'
        write(un_,'(A)') '   - simulation_summary.txt (this file)'
        write(un_,'(A)') '   - output_time_series.dat (time evolution)'
        write(un_,'(A)') '   - flow_statistics.dat (statistical summaries)'
        write(un_,'(A)') '   - velocity_field_XXXXX.dat (field snapshots)'
        write(un_,'(A)') '   - spanwise_profile_XXXXX.dat (averaged profiles)'
        write(un_,'(A)') ' '
        write(un_,'(A)') ' RESTART CONFIGURATION (Phase 5):'
        write(un_,'(A,I0)') '   Restart write frequency: ', restart_write_freq
        write(un_,'(A,A)') '   Restart filename: ', trim(restart_filename)
        if (restart_from_file) then
            write(un_,'(A)') '   Restart mode: ENABLED (restarting from checkpoint)'
            write(un_,'(A,I0)') '   Restart step: ', restart_step
            write(un_,'(A,F12.6)') '   Restart time: ', restart_time
        else
            write(un_,'(A)') '   Restart mode: DISABLED (new simulation)'
        end if
        write(un_,'(A)') ' '
        write(un_,'(A)') '============================================'
        close(un_)
        write(*,'(A)') ' Simulation summary written: simulation_summary.txt'
    end subroutine write_simulation_summary

    ! print_step_timing (DNS_pressure_BC_3D.f90:2207-2230), on wall-clock step times
    subroutine print_step_timing(step, current_time)
        integer, intent(in) :: step
        real(wp), intent(in) :: current_time
        write(*,'(A)') ' ================================='
        write(*,'(A,I6)') ' TIMING SUMMARY - Step: ', step
        write(*,'(A)') ' ================================='
        write(*,'(A,F8.4,A)') ' Current step time: ', current_time, ' seconds'
        write(*,'(A,F8.4,A)') ' Average step time: ', avg_step_time, ' seconds'
        write(*,'(A,F8.4,A)') ' Min step time:     ', min_step_time, ' seconds'
        write(*,'(A,F8.4,A)') ' Max step time:     ', max_step_time, ' seconds'
        write(*,'(A,F6.2)')   ' Performance ratio: ', current_time / avg_step_time
        write(*,'(A,E12.4)')  ' Max divergence:    ', max_divergence
        write(*,'(A,F8.1,A)') ' Est. time remaining: ', avg_step_time * real(nsteps - step, wp), ' seconds'
        write(*,'(A)') ' ================================='
        write(*,*)
    end subroutine print_step_timing

    ! the end-of-run block (DNS_pressure_BC_3D.f90:373-421), wall clock in place of cpu_time
    subroutine final_summary()
        type(dns_diag_t) :: d
        call check(c_dns_diagnostics(ctx, d), 'dns_diagnostics')
        write(*,'(A)') ' ========================================================'
        write(*,'(A)') '               SIMULATION COMPLETED SUCCESSFULLY'
        write(*,'(A)') ' ========================================================'
        write(*,'(A)')
        write(*,'(A)') ' FINAL PERFORMANCE SUMMARY:'
        write(*,'(A)') ' ----------------------------------------'
        write(*,'(A,F12.6,A)') ' Total simulation time:   ', time, ' (non-dimensional)'
        write(*,'(A,I0,A)')     ' Total time steps:       ', nsteps, ' steps'
        write(*,'(A,F10.3,A)')  ' Total wall clock time:  ', total_wall_time, ' seconds'
        write(*,'(A,F10.6,A)')  ' Average time per step:  ', avg_step_time, ' seconds'
        write(*,'(A,F10.6,A)')  ' Fastest step time:      ', min_step_time, ' seconds'
        write(*,'(A,F10.6,A)')  ' Slowest step time:      ', max_step_time, ' seconds'
        if (total_wall_time > 0.0_wp) then
            write(*,'(A,F8.1,A)') ' Performance:            ', real(nsteps - istart_loop + 1, wp) / total_wall_time, &
                ' steps/second'
            write(*,'(A,ES12.4,A)') ' Performance:            ', real(nsteps - istart_loop + 1, wp) * real(nx, wp) * &
                real(ny, wp) * real(nz, wp) / total_wall_time, ' grid-point steps/second'
        end if
        write(*,'(A)')
        write(*,'(A)') ' FINAL SIMULATION STATE:'
        write(*,'(A)') ' ----------------------------------------'
        write(*,'(A,F12.6)')    ' Final time:             ', time
        write(*,'(A,I0)')       ' Time steps completed:   ', nsteps
        write(*,'(A,ES12.4)')   ' Max divergence:         ', max_divergence
        if (flow_control_method == 2) then
            write(*,'(A,F10.6)') ' Final bulk velocity:    ', d%bulk_velocity
            write(*,'(A,ES12.4)') ' Final pressure gradient:', d%avg_pressure_gradient
        end if
        write(*,'(A)')
        write(*,'(A)') ' OUTPUT FILES GENERATED:'
        write(*,'(A)') ' ----------------------------------------'
        write(*,'(A,A)')       ' Output directory:       ', trim(output_directory)
        write(*,'(A)')         ' Files created:'
        write(*,'(A)')         '   ' // char(226) // char(156) // char(147) // ' output_time_series.dat    (time evolution data)'
        write(*,'(A)')         '   ' // char(226) // char(156) // char(147) // ' flow_statistics.dat       (final flow statistics)'
        write(*,'(A)')         '   ' // char(226) // char(156) // char(147) // ' simulation_summary.txt    (run parameters)'
        write(*,'(A)')         '   ' // char(226) // char(156) // char(147) // ' velocity_field snapshots  (periodic velocity data)'
        write(*,'(A)')         '   ' // char(226) // char(156) // char(147) // ' spanwise_profile data     (wall-normal profiles)'
        write(*,'(A)')
        write(*,'(A)') ' ========================================================'
        write(*,'(A)') '      3D DNS SIMULATION ANALYSIS READY FOR PROCESSING'
        write(*,'(A)') ' ========================================================'
        write(*,'(A)')
        write(*,'(A)') ' Time integration completed successfully!'
        write(*,'(A)') ' ============================================'
    end subroutine final_summary

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

    ! write_restart_file_3d (DNS_pressure_BC_3D.f90:2976-3038): <output directory>/restart.dat,
    ! the layout of :2999-3027.  periodic: written at the top of step `step` (:307-309), so u = un
    ! are the fields before the step and `t` is already advanced; otherwise after the last step
    ! (:367), with un the fields before it.  The flow-control record holds G, U_b, the error
    ! integral and the previous error: U_b is that of the fields written under method 2 and 0 under
    ! method 1 (where the reference never computes it).
    subroutine write_restart_file_3d(step, t, periodic)
        integer, intent(in) :: step
        real(wp), intent(in) :: t
        logical, intent(in) :: periodic
        type(dns_diag_t) :: d
        integer :: un_
        character(len=512) :: filepath
        real(wp) :: ub
        call check(c_dns_get_fields(ctx, c_loc(u), c_loc(v), c_loc(w), c_loc(p_total)), 'dns_get_fields')
        call check(c_dns_diagnostics(ctx, d), 'dns_diagnostics')
        ub = merge(d%bulk_velocity, 0.0_wp, flow_control_method == 2)
        filepath = trim(output_directory) // '/restart.dat'
        open(newunit=un_, file=trim(filepath), form='unformatted', access='stream', status='replace', iostat=ios)
        if (ios /= 0) then
            write(*,'(A,A)') ' ERROR: Cannot create restart file: ', trim(filepath)
            write(*,'(A,I0)') ' I/O error code: ', ios
            return
        end if
        write(un_) 'DNS3D_RESTART_v1.0'
        write(un_) int(step, int32), t, dt
        write(un_) int(nx, int32), int(ny, int32), int(nz, int32)
        write(un_) re, 2.0_wp * pi / xlen, 2.0_wp * pi / ylen
        write(un_) u
        write(un_) v
        write(un_) w
        if (periodic) then
            write(un_) u
            write(un_) v
            write(un_) w
        else
            write(un_) un
            write(un_) vn
            write(un_) wn
        end if
        write(un_) p_total
        write(un_) d%pressure_gradient
        write(un_) ub
        write(un_) d%bulk_error_integral
        write(un_) d%bulk_error_prev
        close(un_)
        write(*,'(A,I0,A,F12.6,A)') ' ' // char(226) // char(156) // char(147) // ' Restart file written at step ', &
            step, ', time = ', t, ' to ' // trim(filepath)
    end subroutine write_restart_file_3d

    ! read_restart_file_3d (DNS_pressure_BC_3D.f90:3052-3194): header, step and time, grid check,
    ! Re warning, u, v, w, un, vn, wn, p_total, flow-control state; the fields go to the device
    ! (dns_set_fields with p: the wall pressure gradient of the first step stays zero, as the
    ! reference's grad_p_prev does), the controller state with dns_set_controller.
    subroutine read_restart_file_3d(filepath, step_out, time_out, success)
        character(len=*), intent(in) :: filepath
        integer, intent(out) :: step_out
        real(wp), intent(out) :: time_out
        logical, intent(out) :: success
        character(len=18) :: header_check
        integer :: un_, nx_file, ny_file, nz_file
        integer(int32) :: i4, n4(3)
        real(wp) :: re_file, alpha_file, beta_file, dt_file
        success = .false.
        step_out = 0
        time_out = 0.0_wp
        open(newunit=un_, file=trim(filepath), form='unformatted', access='stream', status='old', iostat=ios)
        if (ios /= 0) then
            write(*,'(A,A)') ' WARNING: Restart file not found: ', trim(filepath)
            return
        end if
        read(un_, iostat=ios) header_check
        if (ios /= 0 .or. trim(header_check) /= 'DNS3D_RESTART_v1.0') then
            write(*,'(A)') ' ERROR: Invalid restart file header'
            close(un_)
            return
        end if
        read(un_, iostat=ios) i4, time_out, dt_file
        if (ios /= 0) then
            write(*,'(A)') ' ERROR: Cannot read time information from restart file'
            close(un_)
            return
        end if
        step_out = int(i4)
        read(un_, iostat=ios) n4
        if (ios /= 0) then
            write(*,'(A)') ' ERROR: Cannot read grid dimensions from restart file'
            close(un_)
            return
        end if
        nx_file = int(n4(1)); ny_file = int(n4(2)); nz_file = int(n4(3))
        if (nx_file /= nx .or. ny_file /= ny .or. nz_file /= nz) then
            write(*,'(A)') ' ERROR: Grid dimensions in restart file do not match current setup'
            write(*,'(A,3I0)') ' Current grid: ', nx, ny, nz
            write(*,'(A,3I0)') ' File grid:    ', nx_file, ny_file, nz_file
            close(un_)
            step_out = 0
            return
        end if
        read(un_, iostat=ios) re_file, alpha_file, beta_file
        if (ios /= 0) then
            write(*,'(A)') ' ERROR: Cannot read physical parameters from restart file'
            close(un_)
            return
        end if
        if (abs(re_file - re) > 1.0e-12_wp) &
            write(*,'(A,F12.6,A,F12.6)') ' WARNING: Reynolds number mismatch: file=', re_file, ', current=', re
        read(un_, iostat=ios) u, v, w, un, vn, wn, p_total
        if (ios /= 0) then
            write(*,'(A)') ' ERROR: Cannot read the fields from restart file'
            close(un_)
            return
        end if
        read(un_, iostat=ios) g_restart, ub_restart, int_restart, prev_restart
        if (ios /= 0) then
            write(*,'(A)') ' WARNING: Cannot read flow control state (using defaults)'
            g_restart = target_pressure_gradient
            ub_restart = 0.0_wp
            int_restart = 0.0_wp
            prev_restart = 0.0_wp
        end if
        close(un_)
        call check(c_dns_set_fields(ctx, c_loc(u), c_loc(v), c_loc(w), c_loc(p_total)), 'dns_set_fields')
        call check(c_dns_set_controller(ctx, g_restart, int_restart, prev_restart, 0.0_wp, 0_c_int), &
                   'dns_set_controller')
        restart_step = step_out
        restart_time = time_out
        success = .true.
        write(*,'(A,I0,A,F12.6,A)') ' ' // char(226) // char(156) // char(147) // &
            ' Restart file read successfully: step ', step_out, ', time = ', time_out, ' from ' // trim(filepath)
    end subroutine read_restart_file_3d

end program dns_3d_gpu
