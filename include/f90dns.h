/* f90dns.h -- C ABI of the MI355X (gfx950) channel-DNS time step.
 *
 * Drop-in boundary for the 3D fractional step of DNS_pressure_BC_3D.f90.  The
 * reference's stage subroutines are argument-less internal procedures of
 * `program dns_3d` working on host-associated globals, and its FFT boundary is
 * the Fortran module fftw3_dns_module.f90.  This header replaces both with plain
 * C entry points (no C++ or torch types): a Fortran driver binds them with
 * ISO_C_BINDING, any other host language with its own FFI.
 *
 * Conventions
 *  - Physical fields are double, Fortran order f(i,j,k) at [i + nx*(j + ny*k)],
 *    i = x (periodic, fastest), j = y (periodic), k = z (LGL, k = 0 at z = -1).
 *  - Spectral planes are complex(nx/2+1, ny) stored as interleaved (re, im) double
 *    pairs, kx fastest, FFTW's r2c layout.
 *  - Every function returns 0 on success or a DNS_E* code; dns_last_error() gives
 *    the message of the last failure on that context (the reference stops the
 *    program instead: DNS_pressure_BC_3D.f90:837-875,1506,1715-1718).
 *  - All device work of a context runs in order on the context's HIP stream;
 *    functions that return host data synchronise that stream.
 */
#ifndef F90DNS_H
#define F90DNS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    DNS_OK = 0,
    DNS_EINVAL = 1,       /* bad parameter or shape */
    DNS_ENOMEM = 2,       /* device or host allocation failed */
    DNS_EDEVICE = 3,      /* HIP runtime error or no gfx950 device */
    DNS_ENONFINITE = 4,   /* NaN/Inf found in a field */
    DNS_ESOLVER = 5,      /* per-mode operator could not be factorised */
    DNS_EUNSUPPORTED = 6  /* option not implemented on this path */
};

/* DNS_PREC_MIXED: FP32 plane transforms, nonlinear product and S / omega storage, FP64 solves
 * and D1 arithmetic.  DNS_PREC_MIXED_F32D: the same, and the vorticity's wall-normal D1
 * contraction on FP32 MFMA (v_mfma_f32_16x16x4f32) with FP32 accumulation. */
enum { DNS_PREC_FP64 = 0, DNS_PREC_MIXED = 1, DNS_PREC_MIXED_F32D = 2 };

/* Run parameters: the namelist variables of read_input_3d that the 3D step reads
 * (DNS_pressure_BC_3D.f90:559-568), plus the precision of the device path. */
typedef struct dns_params {
    int nx, ny, nz;                     /* &grid nx_input, ny_input, nz_input */
    double xlen, ylen, re, dt;          /* &simulation xlen, ylen, re; &time_control dt */
    int use_crank_nicolson;             /* &simulation: theta = 0.5 if nonzero, else 1 */
    int use_dealias_23_rule;            /* &computational */
    int use_weak_form_divergence;       /* &pressure_solver: 1 = weak-form pressure RHS (:1749-1835) */
    int flow_control_method;            /* &flow_control: 1 = constant G, 2 = PI on U_bulk */
    double target_pressure_gradient;
    double target_bulk_velocity;
    double controller_gain;
    int controller_update_freq;
    int istart;                         /* &time_control istart: the PI controller's step argument */
    double u_centerline;                /* &simulation */
    int enable_perturbations;           /* &perturbations */
    int perturbation_kind;              /* 0 = analytical (mode_x, mode_y), 1 = seeded noise */
    int analytical_mode_x, analytical_mode_y;
    double perturbation_amplitude;
    unsigned long long seed;            /* seeded noise: splitmix64 seed */
    int noise_kmax;                     /* seeded noise: modes 0 <= mx, my <= noise_kmax */
    int precision;                      /* DNS_PREC_FP64, DNS_PREC_MIXED or DNS_PREC_MIXED_F32D */
    int lgl_derivative;                 /* 0: D1 as lgl_module.f90:211-263 (entries whose Lagrange
                                           denominator is <= 1e-15 set to 0: D1 is diagonal for
                                           nz >= 65); 1: exact LGL D1 (not in the reference) */
    /* Optional numerics variants of the 2D solver applied to the 3D step (no 3D counterpart):
     * convection_scheme 0 = the 3D step's single product u x omega; 1 = the 4-stage 1/irk
     * scheme of runge_kutta_convection (DNS_pressure_BC_2D.f90:931-1014): S is re-evaluated
     * at un + (dt/irk) S for irk = 4..1 (walls zero after each stage) and the last S is the
     * source term.  pressure_solver 0 = direct per-mode solve (zgesv, DNS_pressure_BC_3D.f90:1714);
     * 1 = jacgs, Jacobi-preconditioned CG (DNS_pressure_BC_2D.f90:2361-2476) on the real and
     * imaginary parts, x0 = 0, at most 500 iterations, stop at ||r|| <= cgstol (&simulation
     * cgstol); FP64 only, not with the weak form, nz <= 192.  Single contexts and batches only. */
    int convection_scheme;
    int pressure_solver;
    double cgstol;
} dns_params;

/* Diagnostics of output_diagnostics_3d / check_divergence_3d
 * (DNS_pressure_BC_3D.f90:2094-2146, 2171-2190). */
typedef struct dns_diag {
    double umax, vmax, wmax;            /* max |u|, |v|, |w| */
    double energy;                      /* LGL-weighted kinetic energy per unit volume */
    double pressure_gradient;           /* current G (current_pressure_gradient) */
    double bulk_velocity;               /* U_bulk of the current u */
    double avg_pressure_gradient;       /* running average of G (method 2) */
    double bulk_error_integral;         /* PI state */
    double bulk_error_prev;
    double time;
    int step;
} dns_diag;

typedef struct dns_ctx dns_ctx;

/* Context: grid constants, LGL operators, per-mode solve bases and the resident
 * fields.  Replaces setup_grid_3d, setup_spectral_3d, allocate_arrays_3d
 * (DNS_pressure_BC_3D.f90:728-888). */
int dns_create(const dns_params* p, dns_ctx** out);
int dns_destroy(dns_ctx* ctx);
const char* dns_last_error(const dns_ctx* ctx);   /* ctx may be NULL: last creation error */

/* initialize_fields_3d + apply_solenoidal_perturbations (DNS_pressure_BC_3D.f90:909-962,
 * 3219-3265): Poiseuille profile, perturbation, wall values; p = 0, time = 0. */
int dns_init_fields(dns_ctx* ctx);
/* Host physical fields in/out (any pointer may be NULL to skip that field). */
int dns_set_fields(dns_ctx* ctx, const double* u, const double* v, const double* w, const double* p);
int dns_get_fields(dns_ctx* ctx, double* u, double* v, double* w, double* p);
/* Device-pointer variants (physical fields already in HBM). */
int dns_set_fields_device(dns_ctx* ctx, const double* u, const double* v, const double* w, const double* p);
int dns_get_fields_device(dns_ctx* ctx, double* u, double* v, double* w, double* p);
/* PI controller / forcing state (current_pressure_gradient, bulk_velocity_error_integral,
 * bulk_velocity_error_prev, pressure_gradient_sum, pressure_gradient_count). */
int dns_set_controller(dns_ctx* ctx, double G, double integral, double prev, double gsum, int gcount);

/* The five stage subroutines of the time loop, same names and order as
 * DNS_pressure_BC_3D.f90:314-326.  dns_begin_step does `time = time + dt` (:296). */
int dns_begin_step(dns_ctx* ctx);
int dns_compute_source_terms(dns_ctx* ctx);       /* :1032-1148 */
int dns_viscous_step(dns_ctx* ctx);               /* :1285-1332 */
int dns_solve_pressure(dns_ctx* ctx);             /* :1631-1733 */
int dns_pressure_correction(dns_ctx* ctx);        /* :1903-1929 */
int dns_apply_boundary_conditions(dns_ctx* ctx);  /* :1953-1972 */
/* nsteps whole steps (begin + five stages), pressure stages fused; asynchronous. */
int dns_step(dns_ctx* ctx, int nsteps);
/* Same as dns_step, the step replayed from a captured HIP graph. */
int dns_step_graph(dns_ctx* ctx, int nsteps);
int dns_synchronize(dns_ctx* ctx);

int dns_diagnostics(dns_ctx* ctx, dns_diag* d);
int dns_check_divergence(dns_ctx* ctx, double* max_div);
/* The wall-normal grid of the context: LGL nodes zpts(nz) on [-1, 1] and weights zwts(nz)
 * (lgl_nodes_weights, lgl_module.f90:30-80, as setup_spectral_3d stores them,
 * DNS_pressure_BC_3D.f90:799); either pointer may be NULL.  The driver's flow statistics
 * (write_flow_statistics, :2903-2922) take wall differences over them. */
int dns_grid_nodes(const dns_ctx* ctx, double* z, double* w);

/* Per-step device time of the last dns_step / dns_step_graph call, from HIP events
 * recorded on the context stream: min, p50, p99, mean in milliseconds. */
int dns_step_timing(dns_ctx* ctx, double* out4);
/* Name of the HIP stream handle (hipStream_t) the context runs on. */
void* dns_stream(dns_ctx* ctx);
/* One step (begin + five stages, pressure fused) enqueued on the context stream, eager or
 * replayed from the captured graph; returns without waiting. */
int dns_enqueue_step(dns_ctx* ctx, int use_graph);
/* 0: all work of the context has finished; 1: work still running. */
int dns_query(dns_ctx* ctx);

/* ---- Batched members (no counterpart in the reference, which runs one flow) ---------
 * One context holds nmembers independent flows of the same grid, dt, domain, scheme flags
 * and precision; Re, the flow-control targets and the initial perturbation may differ per
 * member.  Every stage launch advances all members, so a step of B members costs far less
 * than B steps.  A member's results do not depend on the other members: each member is
 * computed by the same arithmetic as a context of one.  The unsuffixed field and
 * diagnostics functions act on member 0. */
int dns_create_batch(const dns_params* members, int nmembers, dns_ctx** out);
int dns_members(const dns_ctx* ctx);
size_t dns_device_bytes(const dns_ctx* ctx);     /* HBM held by the context */
int dns_set_member_fields(dns_ctx* ctx, int member, const double* u, const double* v, const double* w,
                          const double* p);
int dns_get_member_fields(dns_ctx* ctx, int member, double* u, double* v, double* w, double* p);
int dns_set_member_controller(dns_ctx* ctx, int member, double G, double integral, double prev, double gsum,
                              int gcount);
/* Reassign member `member` (same shape as the context): its Re, controller and perturbation
 * parameters; resets its controller state.  dns_init_member_fields then gives it the initial
 * field of those parameters (initialize_fields_3d for one member). */
int dns_set_member_params(dns_ctx* ctx, int member, const dns_params* p);
int dns_init_member_fields(dns_ctx* ctx, int member);
int dns_member_diagnostics(dns_ctx* ctx, int member, dns_diag* d);
int dns_member_divergence(dns_ctx* ctx, int member, double* max_div);

/* ---- FFT module (fftw3_dns_module.f90:19-152) --------------------------------
 * Plane transforms real(nx,ny) <-> complex(nx/2+1,ny): forward unnormalised, e^{-i};
 * backward e^{+i} then divided by nx*ny, with FFTW's c2r handling of non-Hermitian
 * input.  Host-array variants copy in and out; _device variants take HBM pointers. */
typedef struct dns_fft_plan dns_fft_plan;
int dns_fft_plan_create(int nx, int ny, int precision, dns_fft_plan** out);  /* setup_fftw3_plans_dns */
int dns_fft_plan_destroy(dns_fft_plan* plan);                                /* destroy_fftw3_plans_dns */
int dns_fft_forward(dns_fft_plan* plan, const double* real_in, double* complex_out, int nplanes);
int dns_fft_backward(dns_fft_plan* plan, const double* complex_in, double* real_out, int nplanes);
int dns_fft_forward_device(dns_fft_plan* plan, const double* real_in, double* complex_out, int nplanes, void* stream);
int dns_fft_backward_device(dns_fft_plan* plan, const double* complex_in, double* real_out, int nplanes, void* stream);

/* Allocation-free plane transforms on HBM pointers (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  direction DNS_FFT_FORWARD is dns_fft_forward_device,
 * DNS_FFT_BACKWARD dns_fft_backward_device, with the same results; the plan's device must be
 * current.  Where the plane does not fit the LDS in FP64 (e.g. 192 x 192) the transform takes two
 * passes through a scratch array that the library holds in device memory for the life of the
 * process (16 MiB per device, set up by dns_fft_plan_create); batches larger than it holds run
 * in chunks.  The call allocates nothing.  Every call on one device shares that scratch: calls
 * that overlap on one device (other streams, other threads) race and must be serialised by the
 * caller; calls on one stream are ordered. */
enum { DNS_FFT_FORWARD = -1, DNS_FFT_BACKWARD = 1 };
int dns_fft_execute_device_noalloc(dns_fft_plan* plan, int direction, const double* in, double* out, int nplanes,
                                   void* stream);

/* Library version string. */
const char* dns_version(void);

#ifdef __cplusplus
}
#endif
#endif
