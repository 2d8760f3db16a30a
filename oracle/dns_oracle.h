/* dns_oracle.h -- CPU FP64 restatement of the 3D channel-DNS fractional time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product in
 * f90_dns_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product never calls it.
 *
 * It restates, in plain C, the algorithm of DNS_pressure_BC_3D.f90 as written
 * (physical-space fields, one 2D r2c/c2r per z-plane per derivative, per-mode
 * dense LU solves), with every quirk of the reference kept: double lifting of
 * the Helmholtz boundary values, the PI controller running with the input
 * `istart`, post-filter 2/3 dealiasing, collocation D2 = D1*D1 for the explicit
 * diffusion vs Galerkin D1^T W D1 for the implicit operator, and the 1e-15
 * threshold on the Lagrange denominators in the differentiation matrix.
 */
#ifndef DNS_ORACLE_H
#define DNS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Run parameters: the fields of the reference's eight namelists that the 3D
 * step reads (DNS_pressure_BC_3D.f90:559-568). */
typedef struct orc_params {
    int nx, ny, nz;
    double xlen, ylen, re, dt;
    int use_crank_nicolson;        /* theta = 0.5 if set, else 1 (backward Euler) */
    int use_dealias_23_rule;
    int flow_control_method;       /* 1 = constant G, 2 = PI on bulk velocity */
    double target_pressure_gradient;
    double target_bulk_velocity;
    double controller_gain;
    int controller_update_freq;
    int istart;                    /* the PI controller is called with this value */
    double u_centerline;
    /* initial condition */
    int enable_perturbations;
    int perturbation_kind;         /* 0 = analytical (mode_x, mode_y), 1 = seeded noise */
    int analytical_mode_x, analytical_mode_y;
    double perturbation_amplitude;
    unsigned long long seed;       /* seeded noise generator */
    int noise_kmax;                /* seeded noise: 0 <= mx, my <= noise_kmax */
    /* oracle execution mode */
    int refactor_every_step;       /* 1: rebuild + LU every mode every step, as written */
    int lgl_derivative;            /* 0: reference D1 (1e-15 denominator threshold), 1: exact LGL D1 */
    int use_weak_form_divergence;  /* 1: solve_pressure_3d_weak_form (DNS_pressure_BC_3D.f90:1749-1835) */
    int fft_variant;               /* 0: Stockham factors 4s, 2s, odd primes; 1: radix 2 in place of 4,
                                      odd factors first (an equivalent transform, other rounding) */
    /* optional numerics variants of the 2D solver, applied to the 3D step (parity unpinned) */
    int convection_scheme;         /* 0: the 3D step's single evaluation of u x omega; 1: the 4-stage
                                      1/irk scheme of runge_kutta_convection (DNS_pressure_BC_2D.f90:931-1014) */
    int pressure_solver;           /* 0: direct per-mode solve (zgesv, :1714); 1: jacgs, Jacobi-
                                      preconditioned CG (DNS_pressure_BC_2D.f90:2361-2476), real and
                                      imaginary parts separately, x0 = 0, 500 iterations, tol = cgstol */
    double cgstol;                 /* &simulation cgstol: absolute residual tolerance of jacgs */
} orc_params;

/* PI controller state (DNS_pressure_BC_3D.f90:145-152): current_pressure_gradient,
 * current_bulk_velocity, bulk_velocity_error, _integral, _prev, pressure_gradient_sum,
 * pressure_gradient_count, avg_pressure_gradient. */
typedef struct orc_pi_state {
    double G, Ub, err, err_int, err_prev, gsum, gavg;
    int gcount;
} orc_pi_state;

typedef struct orc_diag {
    double umax, vmax, wmax, energy;
    double pressure_gradient;      /* current G */
    double bulk_velocity;          /* U_b of the current u */
    double avg_pressure_gradient;
    double bulk_error_integral, bulk_error_prev;
    double time;
    int step;
} orc_diag;

typedef struct orc_ctx orc_ctx;

orc_ctx* orc_create(const orc_params* p);
void orc_destroy(orc_ctx* c);
int orc_init_fields(orc_ctx* c);       /* Poiseuille + perturbation + BCs */
int orc_set_fields(orc_ctx* c, const double* u, const double* v, const double* w, const double* p);
int orc_get_fields(const orc_ctx* c, double* u, double* v, double* w, double* p);
int orc_set_wall_gradients(orc_ctx* c, const double* gpx, const double* gpy);
int orc_get_wall_gradients(const orc_ctx* c, double* gpx, double* gpy); /* (nx,ny,2) walls k=1,nz */
int orc_set_controller(orc_ctx* c, double G, double integral, double prev, double gsum, int gcount);

/* the five stage subroutines, in the reference's order (DNS_pressure_BC_3D.f90:314-326) */
void orc_compute_source_terms(orc_ctx* c);
void orc_viscous_step(orc_ctx* c);
void orc_solve_pressure(orc_ctx* c);
void orc_pressure_correction(orc_ctx* c);
void orc_apply_boundary_conditions(orc_ctx* c);
/* nsteps full steps: time += dt, the five stages */
void orc_step(orc_ctx* c, int nsteps);

void orc_diagnostics(orc_ctx* c, orc_diag* d);
double orc_max_divergence(orc_ctx* c);
const double* orc_source_terms(const orc_ctx* c, int comp);   /* su, sv, sw after stage 1 */
const double* orc_star_fields(const orc_ctx* c, int comp);    /* u*, v*, w* after stage 2 */
const double* orc_phi(const orc_ctx* c);                      /* phi after stage 3 */

/* update_pressure_gradient_3d (DNS_pressure_BC_3D.f90:2488-2549) for a given bulk velocity ub
 * (the value calculate_bulk_velocity_3d would return): returns 1 when the controller ran, 0 when
 * mod(istep, controller_update_freq) /= 0 left the state untouched. */
int orc_pi_update(const orc_params* p, orc_pi_state* s, int istep, double ub);
/* calculate_bulk_velocity_3d (DNS_pressure_BC_3D.f90:2434-2464) of a field u(nx,ny,nz) over the
 * LGL weights of nz nodes. */
double orc_bulk_velocity(int nx, int ny, int nz, const double* u);

/* component-level entry points (unit parity) */
void orc_lgl(int n, double* nodes, double* weights);
void orc_diff_matrix(int n, const double* nodes, double* d1 /* column-major n x n */);
void orc_diff_matrix_exact(int n, const double* nodes, double* d1 /* column-major n x n */);
int orc_status(const orc_ctx* c);   /* 0, or k+1 when a per-mode LU met a zero pivot at column k
                                       (the reference's zgesv info /= 0 stop) */
/* jacgs (DNS_pressure_BC_2D.f90:2361-2476) on one real system a(n,n) column-major, x in/out;
 * returns 0, or -1 when niter iterations did not reach ||r|| <= tol; *iters gets the count. */
int orc_jacgs(int n, const double* a, double* x, const double* b, int niter, double tol, int* iters);
/* solves of the last jacgs pressure stage: total iterations and non-converged systems */
void orc_jacgs_stats(const orc_ctx* c, long long* iterations, int* nonconverged);
void orc_wavenumbers(int nx, int ny, double xlen, double ylen, double* xw, double* yw);
void orc_dealias_mask(int nx, int ny, unsigned char* mask /* (nx/2+1, ny) */);
int  orc_fft_r2c(int nx, int ny, int nplanes, const double* in, double* out_interleaved);
int  orc_fft_c2r(int nx, int ny, int nplanes, const double* in_interleaved, double* out);
int  orc_make_initial_field(const orc_params* p, double* u, double* v, double* w);

#ifdef __cplusplus
}
#endif
#endif
