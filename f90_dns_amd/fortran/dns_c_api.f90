! dns_c_api.f90 -- ISO_C_BINDING interfaces of include/f90dns.h (the MI355X channel-DNS C ABI).
! The derived types mirror the C structs field by field (bind(C) gives them C layout).
module dns_c_api
    use iso_c_binding
    implicit none
    private
    public :: dns_params_t, dns_diag_t, DNS_OK, DNS_PREC_FP64, DNS_PREC_MIXED, DNS_PREC_MIXED_F32D
    public :: c_dns_create, c_dns_destroy, c_dns_last_error, c_dns_init_fields, c_dns_set_fields, &
              c_dns_get_fields, c_dns_begin_step, c_dns_compute_source_terms, c_dns_viscous_step, &
              c_dns_solve_pressure, c_dns_pressure_correction, c_dns_apply_boundary_conditions, &
              c_dns_step, c_dns_step_graph, c_dns_synchronize, c_dns_diagnostics, c_dns_check_divergence, &
              c_dns_grid_nodes, c_dns_set_controller, &
              c_dns_fft_plan_create, c_dns_fft_plan_destroy, c_dns_fft_forward, c_dns_fft_backward, &
              dns_error_message

    integer(c_int), parameter :: DNS_OK = 0, DNS_PREC_FP64 = 0, DNS_PREC_MIXED = 1, &
                                      DNS_PREC_MIXED_F32D = 2

    type, bind(C) :: dns_params_t
        integer(c_int) :: nx, ny, nz
        real(c_double) :: xlen, ylen, re, dt
        integer(c_int) :: use_crank_nicolson, use_dealias_23_rule, use_weak_form_divergence, flow_control_method
        real(c_double) :: target_pressure_gradient, target_bulk_velocity, controller_gain
        integer(c_int) :: controller_update_freq, istart
        real(c_double) :: u_centerline
        integer(c_int) :: enable_perturbations, perturbation_kind, analytical_mode_x, analytical_mode_y
        real(c_double) :: perturbation_amplitude
        integer(c_long_long) :: seed
        integer(c_int) :: noise_kmax, precision, lgl_derivative
        integer(c_int) :: convection_scheme, pressure_solver
        real(c_double) :: cgstol
    end type dns_params_t

    type, bind(C) :: dns_diag_t
        real(c_double) :: umax, vmax, wmax, energy, pressure_gradient, bulk_velocity, avg_pressure_gradient
        real(c_double) :: bulk_error_integral, bulk_error_prev, time
        integer(c_int) :: step
    end type dns_diag_t

    interface
        integer(c_int) function c_dns_create(p, ctx) bind(C, name="dns_create")
            import :: c_int, c_ptr, dns_params_t
            type(dns_params_t), intent(in) :: p
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function c_dns_destroy(ctx) bind(C, name="dns_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        type(c_ptr) function c_dns_last_error(ctx) bind(C, name="dns_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function c_dns_init_fields(ctx) bind(C, name="dns_init_fields")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function c_dns_set_fields(ctx, u, v, w, p) bind(C, name="dns_set_fields")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, u, v, w, p
        end function
        integer(c_int) function c_dns_get_fields(ctx, u, v, w, p) bind(C, name="dns_get_fields")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, u, v, w, p
        end function
        integer(c_int) function c_dns_set_controller(ctx, g, integral, prev, gsum, gcount) &
                bind(C, name="dns_set_controller")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: g, integral, prev, gsum
            integer(c_int), value :: gcount
        end function
        integer(c_int) function c_dns_begin_step(ctx) bind(C, name="dns_begin_step")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function c_dns_compute_source_terms(ctx) bind(C, name="dns_compute_source_terms")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function c_dns_viscous_step(ctx) bind(C, name="dns_viscous_step")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function c_dns_solve_pressure(ctx) bind(C, name="dns_solve_pressure")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function c_dns_pressure_correction(ctx) bind(C, name="dns_pressure_correction")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function c_dns_apply_boundary_conditions(ctx) bind(C, name="dns_apply_boundary_conditions")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function c_dns_step(ctx, n) bind(C, name="dns_step")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: n
        end function
        integer(c_int) function c_dns_step_graph(ctx, n) bind(C, name="dns_step_graph")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: n
        end function
        integer(c_int) function c_dns_synchronize(ctx) bind(C, name="dns_synchronize")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function c_dns_diagnostics(ctx, d) bind(C, name="dns_diagnostics")
            import :: c_int, c_ptr, dns_diag_t
            type(c_ptr), value :: ctx
            type(dns_diag_t), intent(out) :: d
        end function
        integer(c_int) function c_dns_check_divergence(ctx, mx) bind(C, name="dns_check_divergence")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: mx
        end function
        integer(c_int) function c_dns_grid_nodes(ctx, z, w) bind(C, name="dns_grid_nodes")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, z, w
        end function
        integer(c_int) function c_dns_fft_plan_create(nx, ny, prec, plan) bind(C, name="dns_fft_plan_create")
            import :: c_int, c_ptr
            integer(c_int), value :: nx, ny, prec
            type(c_ptr), intent(out) :: plan
        end function
        integer(c_int) function c_dns_fft_plan_destroy(plan) bind(C, name="dns_fft_plan_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: plan
        end function
        integer(c_int) function c_dns_fft_forward(plan, rin, cout, nplanes) bind(C, name="dns_fft_forward")
            import :: c_int, c_ptr
            type(c_ptr), value :: plan, rin, cout
            integer(c_int), value :: nplanes
        end function
        integer(c_int) function c_dns_fft_backward(plan, cin, rout, nplanes) bind(C, name="dns_fft_backward")
            import :: c_int, c_ptr
            type(c_ptr), value :: plan, cin, rout
            integer(c_int), value :: nplanes
        end function
        integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function
    end interface

contains

    ! message of the last failure on ctx (or of the last creation failure when ctx is null)
    function dns_error_message(ctx) result(msg)
        type(c_ptr), intent(in) :: ctx
        character(len=:), allocatable :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: chars(:)
        integer :: n, i
        p = c_dns_last_error(ctx)
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: msg)
        do i = 1, n
            msg(i:i) = chars(i)
        end do
    end function dns_error_message

end module dns_c_api
